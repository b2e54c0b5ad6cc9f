// The C-ABI of tilestream_c.h over the C++ API.  Record types are
// layout-identical to their C++ counterparts (checked below), so arrays are
// passed through without copying.
#include "plan.h"

#include <tilestream/batch.h>
#include <tilestream/cpu.h>
#include <tilestream/hip.h>
#include <tilestream/serve.h>
#include <tilestream/tiff.h>
#include <tilestream_c.h>

#include <cstddef>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

using namespace tilestream;

#define TS_SAME_LAYOUT(C_TYPE, CXX_TYPE) \
    static_assert(sizeof(C_TYPE) == sizeof(CXX_TYPE) && alignof(C_TYPE) == alignof(CXX_TYPE), #C_TYPE)

TS_SAME_LAYOUT(ts_image_spec, ImageSpec);
TS_SAME_LAYOUT(ts_tile_shape, TileShape);
TS_SAME_LAYOUT(ts_roi, Roi);
TS_SAME_LAYOUT(ts_const_bytes, ConstBytes);
TS_SAME_LAYOUT(ts_mutable_bytes, MutableBytes);
TS_SAME_LAYOUT(ts_const_image_view, ConstImageView);
TS_SAME_LAYOUT(ts_image_view, ImageView);
TS_SAME_LAYOUT(ts_device_result, DeviceResult);
TS_SAME_LAYOUT(ts_workspace_spec, WorkspaceSpec);
TS_SAME_LAYOUT(ts_parse_limits, ParseLimits);
TS_SAME_LAYOUT(ts_parameter, hip::Parameter);
TS_SAME_LAYOUT(ts_calibration, hip::Calibration);
TS_SAME_LAYOUT(ts_tile_mapping, hip::TileMapping);
TS_SAME_LAYOUT(ts_compressed_input, hip::CompressedInput);
TS_SAME_LAYOUT(ts_batch_item, hip::BatchItem);
TS_SAME_LAYOUT(ts_batch_request, hip::BatchRequest);
TS_SAME_LAYOUT(ts_tile_metadata, TileMetadata);
TS_SAME_LAYOUT(ts_segment_metadata, SegmentMetadata);
static_assert(offsetof(ts_batch_item, calibration) == offsetof(hip::BatchItem, calibration));
static_assert(offsetof(ts_batch_item, roi) == offsetof(hip::BatchItem, roi));
static_assert(offsetof(ts_compressed_input, trusted_tile) == offsetof(hip::CompressedInput, trusted_tile));
static_assert(offsetof(ts_parameter, map) == offsetof(hip::Parameter, map));
static_assert(offsetof(ts_batch_request, tile) == offsetof(hip::BatchRequest, tile));

struct ts_plan {
    std::unique_ptr<hip::BatchPlan> plan;
};

struct ts_queue {
    std::unique_ptr<hip::DecodeQueue> queue;
};

TS_SAME_LAYOUT(ts_launch_shape, hip::LaunchShape);
TS_SAME_LAYOUT(ts_queue_stats, hip::QueueStats);

struct ts_archive_set {
    std::unique_ptr<hip::ArchiveSet> set;
};

struct ts_server {
    std::unique_ptr<hip::Server> server;
};

TS_SAME_LAYOUT(ts_archive_info, hip::ArchiveInfo);
TS_SAME_LAYOUT(ts_serve_request, hip::ServeRequest);
TS_SAME_LAYOUT(ts_server_options, hip::ServerOptions);
TS_SAME_LAYOUT(ts_server_stats, hip::ServerStats);
static_assert(offsetof(ts_serve_request, tile) == offsetof(hip::ServeRequest, tile));
static_assert(offsetof(ts_serve_request, roi) == offsetof(hip::ServeRequest, roi));
static_assert(offsetof(ts_serve_request, output) == offsetof(hip::ServeRequest, output));
static_assert(offsetof(ts_serve_request, input) == offsetof(hip::ServeRequest, input));
static_assert(offsetof(ts_serve_request, compressed) == offsetof(hip::ServeRequest, compressed));
static_assert(offsetof(ts_serve_request, offset) == offsetof(hip::ServeRequest, offset));
static_assert(offsetof(ts_server_options, encode_window) == offsetof(hip::ServerOptions, encode_window));
static_assert(offsetof(ts_server_options, admission_bytes) == offsetof(hip::ServerOptions, admission_bytes));
static_assert(offsetof(ts_server_options, staging_bytes) == offsetof(hip::ServerOptions, staging_bytes));
static_assert(offsetof(ts_archive_info, host) == offsetof(hip::ArchiveInfo, host));
static_assert(offsetof(ts_server_options, shape) == offsetof(hip::ServerOptions, shape));
static_assert(offsetof(ts_server_options, large_bytes) == offsetof(hip::ServerOptions, large_bytes));
static_assert(offsetof(ts_server_options, block_tasks) == offsetof(hip::ServerOptions, block_tasks));
static_assert(offsetof(ts_archive_info, tile_count) == offsetof(hip::ArchiveInfo, tile_count));

namespace {

template <typename To, typename From> To &as(From &v) { return *reinterpret_cast<To *>(&v); }
template <typename To, typename From> const To &as(const From &v) { return *reinterpret_cast<const To *>(&v); }

ImageMetadata *meta(ts_metadata *m) { return reinterpret_cast<ImageMetadata *>(m); }
const ImageMetadata *meta(const ts_metadata *m) { return reinterpret_cast<const ImageMetadata *>(m); }

uint32_t code(Status s) { return static_cast<uint32_t>(s); }

hipStream_t stream_of(void *s) { return static_cast<hipStream_t>(s); }

using Exec = Status (*)(const hip::BatchPlan &, MutableBytes, hip::ResultBuffer, hipStream_t) noexcept;

uint32_t exec(Exec fn, const ts_plan *plan, ts_mutable_bytes ws, ts_device_result *results, uint64_t n, void *s) {
    if (plan == nullptr || plan->plan == nullptr) return code(Status::invalid_argument);
    return code(fn(*plan->plan, as<MutableBytes>(ws), {reinterpret_cast<DeviceResult *>(results), n}, stream_of(s)));
}

} // namespace

extern "C" {

uint32_t ts_image_span(const ts_image_spec *image, uint64_t pitch, uint64_t *bytes) {
    return code(image_span(as<ImageSpec>(*image), pitch, *bytes));
}

uint32_t ts_compress_bound(const ts_image_spec *image, ts_tile_shape tile, uint64_t *bytes) {
    return code(compress_bound(as<ImageSpec>(*image), as<TileShape>(tile), *bytes));
}

uint32_t ts_workspace_size(const ts_workspace_spec *spec, uint64_t *bytes) {
    return code(workspace_size(as<WorkspaceSpec>(*spec), *bytes));
}

uint32_t ts_parse_metadata(ts_const_bytes source, const ts_parse_limits *limits, ts_metadata **out) {
    try {
        auto parsed = std::make_unique<ImageMetadata>();
        const ParseLimits lim = limits != nullptr ? as<ParseLimits>(*limits) : ParseLimits{};
        const Status st = parse_metadata(as<ConstBytes>(source), lim, *parsed);
        if (st == Status::success) *out = reinterpret_cast<ts_metadata *>(parsed.release());
        return code(st);
    } catch (const std::bad_alloc &) {
        return code(Status::resource_limit);
    }
}

void ts_metadata_free(ts_metadata *m) { delete meta(m); }

void ts_tiff_default_options(ts_tiff_read_limits *limits, ts_tiff_write_options *options) {
    if (limits != nullptr) {
        const io::TiffReadLimits l;
        *limits = {l.max_width, l.max_height, l.max_decoded_bytes};
    }
    if (options != nullptr) {
        const io::TiffWriteOptions o;
        *options = {o.bigtiff ? 1u : 0u, static_cast<uint32_t>(o.compression), o.horizontal_predictor ? 1u : 0u,
                    o.tile_width, o.tile_height, o.deflate_level};
    }
}

uint32_t ts_read_tiff(const char *path, const ts_tiff_read_limits *limits, ts_tiff_image **out) {
    if (out == nullptr) return code(Status::invalid_argument);
    try {
        auto image = std::make_unique<io::TiffImage>();
        io::TiffReadLimits l;
        if (limits != nullptr) l = {limits->max_width, limits->max_height, limits->max_decoded_bytes};
        const Status st = io::read_tiff(path, *image, l);
        if (st == Status::success) *out = reinterpret_cast<ts_tiff_image *>(image.release());
        return code(st);
    } catch (const std::bad_alloc &) {
        return code(Status::resource_limit);
    }
}

void ts_tiff_image_get(const ts_tiff_image *image, ts_image_spec *spec, uint64_t *pitch_bytes, const void **pixels) {
    const auto *t = reinterpret_cast<const io::TiffImage *>(image);
    if (spec != nullptr) *spec = as<ts_image_spec>(t->image);
    if (pitch_bytes != nullptr) *pitch_bytes = t->pitch_bytes;
    if (pixels != nullptr) *pixels = t->pixels.data();
}

void ts_tiff_image_free(ts_tiff_image *image) { delete reinterpret_cast<io::TiffImage *>(image); }

uint32_t ts_write_tiff(const char *path, ts_const_image_view image, const ts_tiff_write_options *options) {
    io::TiffWriteOptions o;
    if (options != nullptr) {
        if (options->compression > 1) return code(Status::invalid_argument);
        o.bigtiff = options->bigtiff != 0;
        o.compression = static_cast<io::TiffCompression>(options->compression);
        o.horizontal_predictor = options->horizontal_predictor != 0;
        o.tile_width = options->tile_width;
        o.tile_height = options->tile_height;
        o.deflate_level = options->deflate_level;
    }
    return code(io::write_tiff(path, as<ConstImageView>(image), o));
}

void ts_metadata_get_summary(const ts_metadata *m, ts_metadata_summary *out) {
    const ImageMetadata &x = *meta(m);
    out->image = as<ts_image_spec>(x.image);
    out->tile = as<ts_tile_shape>(x.tile);
    out->file_bytes = x.file_bytes;
    out->tile_count = x.tiles.size();
    out->segment_count = x.segments.size();
}

uint32_t ts_metadata_get_tile(const ts_metadata *m, uint64_t i, ts_tile_metadata *out) {
    if (i >= meta(m)->tiles.size()) return code(Status::invalid_argument);
    std::memcpy(out, &meta(m)->tiles[i], sizeof *out);
    return 0;
}

uint32_t ts_metadata_set_tile(ts_metadata *m, uint64_t i, const ts_tile_metadata *in) {
    if (i >= meta(m)->tiles.size()) return code(Status::invalid_argument);
    std::memcpy(&meta(m)->tiles[i], in, sizeof *in);
    return 0;
}

uint32_t ts_metadata_get_segment(const ts_metadata *m, uint64_t i, ts_segment_metadata *out) {
    if (i >= meta(m)->segments.size()) return code(Status::invalid_argument);
    std::memcpy(out, &meta(m)->segments[i], sizeof *out);
    return 0;
}

uint32_t ts_metadata_set_segment(ts_metadata *m, uint64_t i, const ts_segment_metadata *in) {
    if (i >= meta(m)->segments.size()) return code(Status::invalid_argument);
    std::memcpy(&meta(m)->segments[i], in, sizeof *in);
    return 0;
}

uint32_t ts_metadata_truncate(ts_metadata *m, uint64_t tiles, uint64_t segments) {
    ImageMetadata &x = *meta(m);
    if (tiles > x.tiles.size() || segments > x.segments.size()) return code(Status::invalid_argument);
    x.tiles.resize(static_cast<std::size_t>(tiles));
    x.segments.resize(static_cast<std::size_t>(segments));
    return 0;
}

void ts_metadata_set_file_bytes(ts_metadata *m, uint64_t file_bytes) { meta(m)->file_bytes = file_bytes; }

uint32_t ts_cpu_compress(ts_const_image_view input, ts_tile_shape tile, ts_mutable_bytes output, uint64_t *written) {
    return code(cpu::compress(as<ConstImageView>(input), as<TileShape>(tile), as<MutableBytes>(output), *written));
}

uint32_t ts_cpu_decompress(ts_const_bytes input, ts_image_view output, const ts_parse_limits *limits) {
    const ParseLimits lim = limits != nullptr ? as<ParseLimits>(*limits) : ParseLimits{};
    return code(cpu::decompress(as<ConstBytes>(input), as<ImageView>(output), lim));
}

uint32_t ts_prepare_batch(const ts_batch_request *request, ts_plan **plan, void *stream) {
    if (request == nullptr || plan == nullptr) return code(Status::invalid_argument);
    try {
        std::unique_ptr<hip::BatchPlan> fresh;
        const Status st = hip::prepare_batch(as<hip::BatchRequest>(*request), fresh, stream_of(stream));
        if (st != Status::success) return code(st);
        auto *handle = new ts_plan{std::move(fresh)};
        delete *plan;
        *plan = handle;
        return 0;
    } catch (const std::bad_alloc &) {
        return code(Status::resource_limit);
    }
}

void ts_plan_destroy(ts_plan *plan) { delete plan; }

void ts_plan_workspace_spec(const ts_plan *plan, ts_workspace_spec *out) {
    const WorkspaceSpec spec = plan != nullptr && plan->plan ? plan->plan->workspace_spec() : WorkspaceSpec{};
    std::memcpy(out, &spec, sizeof spec);
}

uint32_t ts_compress_async(const ts_plan *p, ts_mutable_bytes ws, ts_device_result *r, uint64_t n, void *s) {
    return exec(hip::compress_async, p, ws, r, n, s);
}
uint32_t ts_compress_async_static(const ts_plan *p, ts_device_result *r, uint64_t n, void *s) {
    if (p == nullptr || p->plan == nullptr) return code(Status::invalid_argument);
    return code(hip::compress_async_static(*p->plan, {reinterpret_cast<DeviceResult *>(r), n}, stream_of(s)));
}
uint32_t ts_decompress_async(const ts_plan *p, ts_mutable_bytes ws, ts_device_result *r, uint64_t n, void *s) {
    return exec(hip::decompress_async, p, ws, r, n, s);
}
uint32_t ts_decode_roi_async(const ts_plan *p, ts_mutable_bytes ws, ts_device_result *r, uint64_t n, void *s) {
    return exec(hip::decode_roi_async, p, ws, r, n, s);
}
uint32_t ts_calibrate_async(const ts_plan *p, ts_mutable_bytes ws, ts_device_result *r, uint64_t n, void *s) {
    return exec(hip::calibrate_async, p, ws, r, n, s);
}
uint32_t ts_decode_calibrate_async(const ts_plan *p, ts_mutable_bytes ws, ts_device_result *r, uint64_t n, void *s) {
    return exec(hip::decode_calibrate_async, p, ws, r, n, s);
}

uint32_t ts_buffer_allocate(uint64_t bytes, void *stream, void **data) {
    // DeviceBuffer::allocate rules, with ownership handed to the caller.
    if (bytes > max_bytes) return code(Status::size_overflow);
    if (bytes == 0) {
        *data = nullptr;
        return 0;
    }
    void *p = nullptr;
    if (hipMallocAsync(&p, static_cast<std::size_t>(bytes), stream_of(stream)) != hipSuccess)
        return code(hip::detail::hip_failure());
    if (reinterpret_cast<std::uintptr_t>(p) % workspace_alignment != 0) {
        static_cast<void>(hipFreeAsync(p, stream_of(stream)));
        return code(Status::hip_error);
    }
    *data = p;
    return 0;
}

uint32_t ts_buffer_release(void *data, void *stream) {
    if (data == nullptr) return 0;
    return code(hip::detail::hip_status(hipFreeAsync(data, stream_of(stream))));
}

uint32_t ts_synchronize(void *stream) { return code(hip::synchronize(stream_of(stream))); }

uint32_t ts_plan_set_decode_shape(ts_plan *plan, uint32_t group, uint32_t block, uint32_t grid, uint32_t depth) {
    if (plan == nullptr || plan->plan == nullptr) return code(Status::invalid_argument);
    // Plans keep tile tables only: the block decoder (group 0) is the queue's.
    if (group == 0 || !tsk_decode_supported(group, block, depth)) return code(Status::invalid_argument);
    auto *state = const_cast<hip::BatchPlan::Impl *>(plan->plan->impl());
    state->shape = {group, block, grid, depth};
    return 0;
}

uint32_t ts_queue_prepare(const ts_batch_item *items, const uint32_t *operations, uint64_t count, ts_queue **queue,
                          void *stream) {
    if (queue == nullptr) return code(Status::invalid_argument);
    try {
        std::unique_ptr<hip::DecodeQueue> fresh;
        const Status st = hip::DecodeQueue::prepare(reinterpret_cast<const hip::BatchItem *>(items),
                                                    reinterpret_cast<const hip::Operation *>(operations), count,
                                                    fresh, stream_of(stream));
        if (st != Status::success) return code(st);
        auto *handle = new ts_queue{std::move(fresh)};
        delete *queue;
        *queue = handle;
        return 0;
    } catch (const std::bad_alloc &) {
        return code(Status::resource_limit);
    }
}

void ts_queue_destroy(ts_queue *queue) { delete queue; }
uint64_t ts_queue_workspace_bytes(const ts_queue *q) { return q->queue->workspace_bytes(); }
void ts_queue_get_stats(const ts_queue *q, ts_queue_stats *out) {
    const hip::QueueStats s = q->queue->stats();
    std::memcpy(out, &s, sizeof s);
}
void ts_queue_get_shape(const ts_queue *q, ts_launch_shape *out) {
    const hip::LaunchShape s = q->queue->shape();
    std::memcpy(out, &s, sizeof s);
}
uint32_t ts_queue_set_shape(ts_queue *q, const ts_launch_shape *shape) {
    if (!tsk_decode_supported(shape->group, shape->block, shape->depth)) return code(Status::invalid_argument);
    q->queue->set_shape(as<hip::LaunchShape>(*shape));
    return 0;
}
uint32_t ts_queue_run(ts_queue *q, ts_mutable_bytes ws, ts_device_result *r, uint64_t n, void *s) {
    return code(q->queue->run(as<MutableBytes>(ws), {reinterpret_cast<DeviceResult *>(r), n}, stream_of(s)));
}
uint32_t ts_queue_capture(ts_queue *q, ts_mutable_bytes ws, ts_device_result *r, uint64_t n, void *s) {
    return code(q->queue->capture(as<MutableBytes>(ws), {reinterpret_cast<DeviceResult *>(r), n}, stream_of(s)));
}
uint32_t ts_queue_launch(ts_queue *q, void *s) { return code(q->queue->launch(stream_of(s))); }
uint32_t ts_queue_autotune(ts_queue *q, ts_mutable_bytes ws, ts_device_result *r, uint64_t n, void *s,
                           const ts_launch_shape *candidates, uint64_t count, uint32_t repeats, float *ms) {
    try {
        std::vector<hip::LaunchShape> c(count);
        for (uint64_t i = 0; i < count; ++i) {
            c[i] = as<hip::LaunchShape>(candidates[i]);
            if (!tsk_decode_supported(c[i].group, c[i].block, c[i].depth)) return code(Status::invalid_argument);
        }
        std::vector<float> times;
        const Status st = q->queue->autotune(as<MutableBytes>(ws), {reinterpret_cast<DeviceResult *>(r), n},
                                             stream_of(s), c, repeats, times);
        for (uint64_t i = 0; i < count && i < times.size(); ++i) ms[i] = times[i];
        return code(st);
    } catch (const std::bad_alloc &) {
        return code(Status::resource_limit);
    }
}

uint32_t ts_archive_set_create(ts_archive_set **set) {
    if (set == nullptr) return code(Status::invalid_argument);
    try {
        std::unique_ptr<hip::ArchiveSet> fresh;
        if (const Status st = hip::ArchiveSet::create(fresh); st != Status::success) return code(st);
        auto *handle = new ts_archive_set{std::move(fresh)};
        delete *set;
        *set = handle;
        return 0;
    } catch (const std::bad_alloc &) {
        return code(Status::resource_limit);
    }
}
void ts_archive_set_destroy(ts_archive_set *set) { delete set; }
uint32_t ts_archive_set_add(ts_archive_set *set, ts_const_bytes file, const ts_metadata *metadata, uint32_t *id) {
    if (set == nullptr || metadata == nullptr || id == nullptr) return code(Status::invalid_argument);
    return code(set->set->add(as<ConstBytes>(file), *meta(metadata), *id));
}
uint32_t ts_archive_set_add_device(ts_archive_set *set, ts_const_bytes file, ts_image_spec image, ts_tile_shape tile,
                                   const ts_device_result *producer, void *stream, uint32_t *id) {
    if (set == nullptr || id == nullptr) return code(Status::invalid_argument);
    return code(set->set->add_device(as<ConstBytes>(file), as<ImageSpec>(image), as<TileShape>(tile),
                                     reinterpret_cast<const DeviceResult *>(producer), stream_of(stream), *id));
}
uint32_t ts_archive_set_commit(ts_archive_set *set, void *stream) {
    if (set == nullptr) return code(Status::invalid_argument);
    return code(set->set->commit(stream_of(stream)));
}
uint32_t ts_archive_set_size(const ts_archive_set *set) { return set->set->size(); }
uint32_t ts_archive_set_info(const ts_archive_set *set, uint32_t id, ts_archive_info *out) {
    if (set == nullptr || out == nullptr) return code(Status::invalid_argument);
    return code(set->set->info(id, as<hip::ArchiveInfo>(*out)));
}
void ts_server_default_options(ts_server_options *out) {
    const hip::ServerOptions o;
    std::memcpy(out, &o, sizeof o);
}
uint32_t ts_server_create(const ts_archive_set *set, const ts_server_options *options, ts_server **server) {
    if (set == nullptr || options == nullptr || server == nullptr) return code(Status::invalid_argument);
    try {
        std::unique_ptr<hip::Server> fresh;
        if (const Status st = hip::Server::create(*set->set, as<hip::ServerOptions>(*options), fresh);
            st != Status::success)
            return code(st);
        auto *handle = new ts_server{std::move(fresh)};
        delete *server;
        *server = handle;
        return 0;
    } catch (const std::bad_alloc &) {
        return code(Status::resource_limit);
    }
}
void ts_server_destroy(ts_server *server) { delete server; }
uint32_t ts_server_check(const ts_server *server, const ts_serve_request *request) {
    if (server == nullptr || request == nullptr) return code(Status::invalid_argument);
    return code(server->server->check(as<hip::ServeRequest>(*request)));
}
uint32_t ts_server_replay(ts_server *server, const ts_serve_request *requests, const uint64_t *arrival_ns,
                          uint64_t count, uint64_t *latency_ns, ts_device_result *results, ts_server_stats *stats) {
    if (server == nullptr || stats == nullptr) return code(Status::invalid_argument);
    return code(server->server->replay(reinterpret_cast<const hip::ServeRequest *>(requests), arrival_ns, count,
                                       latency_ns, reinterpret_cast<DeviceResult *>(results),
                                       as<hip::ServerStats>(*stats)));
}
uint32_t ts_server_run_batch(ts_server *server, const ts_serve_request *requests, uint64_t count,
                             ts_device_result *results) {
    if (server == nullptr) return code(Status::invalid_argument);
    return code(server->server->run_batch(reinterpret_cast<const hip::ServeRequest *>(requests), count,
                                          reinterpret_cast<DeviceResult *>(results)));
}

int ts_hip_malloc(uint64_t bytes, void **data) { return hipMalloc(data, bytes); }
int ts_hip_free(void *data) { return hipFree(data); }
int ts_hip_host_malloc(uint64_t bytes, void **data) { return hipHostMalloc(data, bytes, hipHostMallocMapped); }
int ts_hip_host_free(void *data) { return hipHostFree(data); }
// Both helpers complete before returning.  hipMemset and device-to-device
// hipMemcpy may still run on the null stream after they return, and the null
// stream does not order work on streams created non-blocking.
int ts_hip_memcpy(void *dst, const void *src, uint64_t bytes) {
    hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyDefault);
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    return e;
}
int ts_hip_memcpy_async(void *dst, const void *src, uint64_t bytes, void *stream) {
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, stream_of(stream));
}
int ts_hip_memset(void *dst, int value, uint64_t bytes) {
    hipError_t e = hipMemset(dst, value, bytes);
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    return e;
}
int ts_hip_stream_create(void **stream) {
    hipStream_t s = nullptr;
    const hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    *stream = s;
    return e;
}
int ts_hip_stream_destroy(void *stream) { return hipStreamDestroy(stream_of(stream)); }
int ts_hip_stream_synchronize(void *stream) { return hipStreamSynchronize(stream_of(stream)); }
int ts_hip_device_synchronize(void) { return hipDeviceSynchronize(); }
int ts_hip_get_last_error(void) { return hipGetLastError(); }
int ts_hip_peek_last_error(void) { return hipPeekAtLastError(); }
int ts_hip_device_count(int *count) { return hipGetDeviceCount(count); }
int ts_hip_event_create(void **event) {
    hipEvent_t e = nullptr;
    const hipError_t r = hipEventCreate(&e);
    *event = e;
    return r;
}
int ts_hip_event_destroy(void *event) { return hipEventDestroy(static_cast<hipEvent_t>(event)); }
int ts_hip_event_record(void *event, void *stream) {
    return hipEventRecord(static_cast<hipEvent_t>(event), stream_of(stream));
}
int ts_hip_event_synchronize(void *event) { return hipEventSynchronize(static_cast<hipEvent_t>(event)); }
int ts_hip_stream_wait_event(void *stream, void *event) {
    return hipStreamWaitEvent(stream_of(stream), static_cast<hipEvent_t>(event), 0);
}
int ts_hip_event_elapsed_ms(void *start, void *stop, float *ms) {
    return hipEventElapsedTime(ms, static_cast<hipEvent_t>(start), static_cast<hipEvent_t>(stop));
}

uint64_t ts_abi_sizeof(const char *name) {
    const std::string n = name;
#define TS_SIZE(T) \
    if (n == #T) return sizeof(T);
    TS_SIZE(ts_image_spec)
    TS_SIZE(ts_tile_shape)
    TS_SIZE(ts_roi)
    TS_SIZE(ts_const_bytes)
    TS_SIZE(ts_mutable_bytes)
    TS_SIZE(ts_const_image_view)
    TS_SIZE(ts_image_view)
    TS_SIZE(ts_device_result)
    TS_SIZE(ts_workspace_spec)
    TS_SIZE(ts_parse_limits)
    TS_SIZE(ts_metadata_summary)
    TS_SIZE(ts_tile_metadata)
    TS_SIZE(ts_segment_metadata)
    TS_SIZE(ts_parameter)
    TS_SIZE(ts_calibration)
    TS_SIZE(ts_tile_mapping)
    TS_SIZE(ts_compressed_input)
    TS_SIZE(ts_batch_item)
    TS_SIZE(ts_batch_request)
    TS_SIZE(ts_launch_shape)
    TS_SIZE(ts_queue_stats)
    TS_SIZE(ts_archive_info)
    TS_SIZE(ts_serve_request)
    TS_SIZE(ts_server_options)
    TS_SIZE(ts_server_stats)
    TS_SIZE(ts_tiff_read_limits)
    TS_SIZE(ts_tiff_write_options)
#undef TS_SIZE
    return 0;
}

} // extern "C"
