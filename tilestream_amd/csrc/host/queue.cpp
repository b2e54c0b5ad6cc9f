// DecodeQueue: one persistent launch over a mixed batch (tilestream/batch.h).
#include "plan.h"

#include <tilestream/batch.h>

#include <algorithm>
#include <new>
#include <stdexcept>


namespace tilestream::hip {

struct DecodeQueue::Impl {
    int device = 0;
    bool pageable = false;
    std::uint64_t count = 0;
    bool has_calibrated = false, has_plain = false;
    bool has_u16 = false, has_f32 = false;
    detail::Footprint footprint;
    void *storage = nullptr;
    const tsk_decode_image *images = nullptr;
    const tsk_decode_task *tasks = nullptr;
    const std::uint16_t *rows = nullptr;
    std::uint64_t task_count = 0;
    std::uint64_t row_count = 0;
    // Block decoder tables (block_count 0: every shape uses the tile decoder).
    const tsk_block *blocks = nullptr;
    std::uint64_t block_count = 0;
    QueueStats stats;
    LaunchShape shape;
    hipGraphExec_t graph = nullptr;

    ~Impl() {
        if (graph != nullptr) static_cast<void>(hipGraphExecDestroy(graph));
        if (storage != nullptr) static_cast<void>(hipFreeAsync(storage, hipStreamPerThread));
    }

    // One pixel type and one output kind use a specialised kernel; anything
    // else the kernel that dispatches per image.
    std::uint32_t kind() const noexcept {
        if (has_u16 != has_f32 && has_calibrated != has_plain) {
            if (has_u16) return has_calibrated ? TSK_KIND_U16_CAL : TSK_KIND_U16;
            return has_calibrated ? TSK_KIND_F32_CAL : TSK_KIND_F32;
        }
        return TSK_KIND_MIXED;
    }
};

namespace {

constexpr std::uint64_t kCounterBytes = 256;

template <typename T> bool upload(void *storage, std::uint64_t at, const std::vector<T> &v, hipStream_t stream) {
    return v.empty() || hipMemcpyAsync(static_cast<std::uint8_t *>(storage) + at, v.data(), v.size() * sizeof(T),
                                       hipMemcpyHostToDevice, stream) == hipSuccess;
}

} // namespace

DecodeQueue::DecodeQueue(std::unique_ptr<Impl> impl) noexcept : impl_(std::move(impl)) {}
DecodeQueue::~DecodeQueue() = default;

Status DecodeQueue::prepare(const BatchItem *items, const Operation *operations, std::uint64_t count,
                            std::unique_ptr<DecodeQueue> &queue, hipStream_t stream) noexcept {
    if (count > UINT32_MAX) return Status::resource_limit;
    if (count != 0 && (items == nullptr || operations == nullptr)) return Status::invalid_argument;
    try {
        auto impl = std::make_unique<Impl>();
        impl->count = count;
        detail::DecodeDraft d;
        for (std::uint64_t i = 0; i < count; ++i) {
            const Operation op = operations[i];
            if (op != Operation::decompress && op != Operation::decode_roi && op != Operation::decode_calibrate)
                return Status::invalid_argument;
            if (const Status st = detail::draft_decode_item(items[i], op, false, d); st != Status::success) return st;
        }
        if (!impl->footprint.admit(std::move(d.writable), std::move(d.readable))) return Status::invalid_argument;
        QueueStats &s = impl->stats;
        s.images = count;
        s.tasks = d.tasks.size();
        for (const tsk_decode_image &im : d.images) {
            const std::uint64_t area = std::uint64_t{im.roi_w} * im.roi_h;
            const std::uint64_t stored = im.type == TSK_UINT16 ? 2 : 4;
            s.decoded_bytes += area * stored;
            s.output_bytes += area * (im.calibrated ? 4 : stored);
            s.map_bytes += area * 4 * ((im.offset_map != nullptr) + (im.gain_map != nullptr));
            impl->has_calibrated = impl->has_calibrated || im.calibrated;
            impl->has_plain = impl->has_plain || !im.calibrated;
            impl->has_u16 = impl->has_u16 || im.type == TSK_UINT16;
            impl->has_f32 = impl->has_f32 || im.type == TSK_FLOAT32;
        }
        for (const tsk_decode_task &t : d.tasks) s.payload_bytes += std::uint64_t{t.words} * 4;
        if (count == 0) {
            queue.reset(new DecodeQueue(std::move(impl)));
            return Status::success;
        }

        int pageable = 0;
        if (hipGetDevice(&impl->device) != hipSuccess ||
            hipDeviceGetAttribute(&pageable, hipDeviceAttributePageableMemoryAccess, impl->device) != hipSuccess)
            return detail::hip_failure();
        impl->pageable = pageable != 0;
        if (const Status st = detail::stream_device(stream, impl->device); st != Status::success) return st;
        for (const void *p : d.pointers)
            if (const Status st = detail::accessible(p, impl->device, impl->pageable); st != Status::success) return st;

        std::vector<tsk_block> blocks;
        if (!detail::plan_blocks(d, blocks)) blocks.clear();
        const auto next = [](std::uint64_t at, std::uint64_t bytes) { return (at + bytes + 15) / 16 * 16; };
        const std::uint64_t images_at = 0;
        const std::uint64_t tasks_at = next(images_at, d.images.size() * sizeof(tsk_decode_image));
        const std::uint64_t rows_at = next(tasks_at, d.tasks.size() * sizeof(tsk_decode_task));
        const std::uint64_t blocks_at = next(rows_at, d.rows.size() * sizeof(std::uint16_t));
        const std::uint64_t total = next(blocks_at, blocks.size() * sizeof(tsk_block)) + 16;
        if (hipMallocAsync(&impl->storage, total, stream) != hipSuccess) {
            impl->storage = nullptr;
            return detail::hip_failure();
        }
        auto *base = static_cast<std::uint8_t *>(impl->storage);
        impl->images = reinterpret_cast<const tsk_decode_image *>(base + images_at);
        impl->tasks = reinterpret_cast<const tsk_decode_task *>(base + tasks_at);
        impl->rows = reinterpret_cast<const std::uint16_t *>(base + rows_at);
        impl->task_count = d.tasks.size();
        impl->row_count = d.rows.size();
        impl->blocks = reinterpret_cast<const tsk_block *>(base + blocks_at);
        impl->block_count = blocks.size();
        if (tsk_preload(1) != hipSuccess || tsk_preload(3) != hipSuccess) return detail::hip_failure();
        if (!upload(impl->storage, images_at, d.images, stream) || !upload(impl->storage, tasks_at, d.tasks, stream) ||
            !upload(impl->storage, rows_at, d.rows, stream) || !upload(impl->storage, blocks_at, blocks, stream) ||
            hipStreamSynchronize(stream) != hipSuccess)
            return detail::hip_failure();
        queue.reset(new DecodeQueue(std::move(impl)));
        return Status::success;
    } catch (const std::bad_alloc &) {
        return Status::resource_limit;
    } catch (const std::length_error &) {
        return Status::resource_limit;
    }
}

std::uint64_t DecodeQueue::workspace_bytes() const noexcept { return impl_->count ? kCounterBytes : 0; }
std::uint64_t DecodeQueue::count() const noexcept { return impl_->count; }
QueueStats DecodeQueue::stats() const noexcept { return impl_->stats; }
LaunchShape DecodeQueue::shape() const noexcept { return impl_->shape; }
void DecodeQueue::set_shape(const LaunchShape &shape) noexcept { impl_->shape = shape; }

Status DecodeQueue::run(MutableBytes workspace, ResultBuffer results, hipStream_t stream) noexcept {
    Impl &q = *impl_;
    if (q.count == 0) return Status::success;
    // The same host checks as an execution of a plan.
    const detail::CallNeeds needs{&q.footprint, q.device, q.pageable, q.count, kCounterBytes};
    if (const Status st = detail::check_call(needs, workspace, results, stream); st != Status::success) return st;
    const LaunchShape &sh = q.shape;
    // Group 0 selects the block decoder, which needs every payload address on
    // the host (not the case for encoder output decoded from its device index).
    const bool block_mode = sh.group == 0;
    if (block_mode && q.block_count == 0) return Status::invalid_argument;
    std::uint32_t grid = sh.grid;
    if (grid == 0) {
        int cus = 0, per_cu = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, q.device) != hipSuccess ||
            tsk_decode_occupancy(q.kind(), sh.group, sh.block, sh.depth, &per_cu) != hipSuccess)
            return detail::hip_failure();
        grid = static_cast<std::uint32_t>(std::max(1, cus * std::max(1, per_cu)));
    }
    if (!tsk_decode_supported(sh.group, sh.block, sh.depth)) return Status::invalid_argument;
    auto *counter = static_cast<std::uint64_t *>(workspace.data);
    // One kernel zeroes the results and the work counter.
    if (tsk_results_begin(results.data, static_cast<std::uint32_t>(q.count),
                          sh.chunk != 0 ? static_cast<std::uint32_t *>(workspace.data) : nullptr, 2, stream) != hipSuccess)
        return detail::hip_failure();
    tsk_decode_args a{};
    a.images = q.images;
    a.tasks = q.tasks;
    a.rows = q.rows;
    a.results = results.data;
    a.queue = sh.chunk != 0 ? counter : nullptr;
    a.task_count = block_mode ? q.block_count : q.task_count;
    if (block_mode) {
        a.blocks = q.blocks;
    }
    a.row_count = q.row_count;
    a.image_count = static_cast<std::uint32_t>(q.count);
    a.kind = q.kind();
    a.group = sh.group;
    a.block = sh.block;
    a.grid = grid;
    a.chunk = sh.chunk;
    a.depth = sh.depth;
    a.run = sh.run;
    if (tsk_decode(&a, stream) != hipSuccess ||
        tsk_results_complete(q.images, static_cast<std::uint32_t>(q.count), results.data, stream) != hipSuccess)
        return detail::hip_failure();
    return Status::success;
}

Status DecodeQueue::capture(MutableBytes workspace, ResultBuffer results, hipStream_t stream) noexcept {
    Impl &q = *impl_;
    if (q.graph != nullptr) {
        static_cast<void>(hipGraphExecDestroy(q.graph));
        q.graph = nullptr;
    }
    if (hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return detail::hip_failure();
    const Status st = run(workspace, results, stream);
    hipGraph_t graph = nullptr;
    const hipError_t end = hipStreamEndCapture(stream, &graph);
    if (st != Status::success) {
        if (graph != nullptr) static_cast<void>(hipGraphDestroy(graph));
        return st;
    }
    if (end != hipSuccess) return detail::hip_failure();
    const hipError_t inst = hipGraphInstantiate(&q.graph, graph, nullptr, nullptr, 0);
    static_cast<void>(hipGraphDestroy(graph));
    if (inst != hipSuccess) {
        q.graph = nullptr;
        return detail::hip_failure();
    }
    return Status::success;
}

Status DecodeQueue::launch(hipStream_t stream) noexcept {
    if (impl_->count == 0) return Status::success;
    if (impl_->graph == nullptr) return Status::invalid_argument;
    return detail::hip_status(hipGraphLaunch(impl_->graph, stream));
}

Status DecodeQueue::autotune(MutableBytes workspace, ResultBuffer results, hipStream_t stream,
                             const std::vector<LaunchShape> &candidates, std::uint32_t repeats,
                             std::vector<float> &milliseconds) noexcept {
    try {
        milliseconds.assign(candidates.size(), 0.0F);
        if (candidates.empty() || impl_->count == 0) return Status::success;
        hipEvent_t start = nullptr, stop = nullptr;
        if (hipEventCreate(&start) != hipSuccess || hipEventCreate(&stop) != hipSuccess) return detail::hip_failure();
        float best = 0.0F;
        LaunchShape winner = impl_->shape;
        Status result = Status::success;
        std::vector<float> samples;
        for (std::size_t c = 0; c < candidates.size() && result == Status::success; ++c) {
            impl_->shape = candidates[c];
            result = run(workspace, results, stream); // warm-up
            samples.clear();
            for (std::uint32_t r = 0; r < std::max(1u, repeats) && result == Status::success; ++r) {
                float ms = 0.0F;
                if (hipEventRecord(start, stream) != hipSuccess) result = detail::hip_failure();
                if (result == Status::success) result = run(workspace, results, stream);
                if (result == Status::success &&
                    (hipEventRecord(stop, stream) != hipSuccess || hipEventSynchronize(stop) != hipSuccess ||
                     hipEventElapsedTime(&ms, start, stop) != hipSuccess))
                    result = detail::hip_failure();
                samples.push_back(ms);
            }
            if (result != Status::success) break;
            std::sort(samples.begin(), samples.end());
            milliseconds[c] = samples[samples.size() / 2];
            if (c == 0 || milliseconds[c] < best) {
                best = milliseconds[c];
                winner = candidates[c];
            }
        }
        impl_->shape = winner;
        static_cast<void>(hipEventDestroy(start));
        static_cast<void>(hipEventDestroy(stop));
        return result;
    } catch (const std::bad_alloc &) {
        return Status::resource_limit;
    }
}

} // namespace tilestream::hip
