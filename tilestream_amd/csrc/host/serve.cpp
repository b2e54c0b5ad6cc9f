// Serving resident archives (tilestream/serve.h): archive tile tables planned
// once, region and encode requests described by a few bytes each, continuous
// batching over several streams.
#include "plan.h"

#include <tilestream/format.h>
#include <tilestream/serve.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <new>
#include <stdexcept>
#include <thread>

namespace tilestream::hip {

namespace {

struct Archive {
    ArchiveInfo info;
    const std::uint8_t *bytes = nullptr; // device-visible address (device or mapped pinned host memory)
    std::uint64_t tile_base = 0;         // first entry in the set's tile table
    std::uint32_t tile_columns = 0;
};

// 16-byte blocks holding a payload: what staging copies and the decoder reads.
inline std::uint64_t staged_blocks(const tsk_decode_task &t) {
    const auto p = reinterpret_cast<std::uintptr_t>(t.payload);
    return ((p + std::uint64_t{t.words} * 4 + 15) / 16) - p / 16;
}

constexpr std::uint64_t align_up(std::uint64_t v, std::uint64_t a) { return (v + a - 1) / a * a; }

} // namespace

// Tile table of an archive added from device memory, built there; copied
// into the committed tables at every commit.
struct DeviceTable {
    void *storage = nullptr;           // tasks, then row offsets
    std::uint64_t tile_base = 0, tiles = 0;
    std::uint64_t rows_base = 0, rows = 0;
    std::uint64_t rows_at = 0;         // byte offset of the row offsets in storage
};

struct ArchiveSet::Impl {
    int device = 0;
    bool pageable = false;
    std::vector<Archive> archives;
    std::vector<tsk_decode_task> tiles;
    std::vector<std::uint16_t> rows;
    std::vector<DeviceTable> device_tables;
    // committed tables
    void *storage = nullptr;
    const tsk_decode_task *device_tiles = nullptr;
    const std::uint16_t *device_rows = nullptr;
    std::uint64_t committed = 0;
    std::uint64_t row_count = 0;
    bool any_host = false;

    ~Impl() {
        if (storage != nullptr) static_cast<void>(hipFree(storage));
        for (const DeviceTable &t : device_tables) static_cast<void>(hipFree(t.storage));
    }
};

ArchiveSet::ArchiveSet(std::unique_ptr<Impl> impl) noexcept : impl_(std::move(impl)) {}
ArchiveSet::~ArchiveSet() = default;

Status ArchiveSet::create(std::unique_ptr<ArchiveSet> &set) noexcept {
    try {
        auto impl = std::make_unique<Impl>();
        int pageable = 0;
        if (hipGetDevice(&impl->device) != hipSuccess ||
            hipDeviceGetAttribute(&pageable, hipDeviceAttributePageableMemoryAccess, impl->device) != hipSuccess)
            return detail::hip_failure();
        impl->pageable = pageable != 0;
        set.reset(new ArchiveSet(std::move(impl)));
        return Status::success;
    } catch (const std::bad_alloc &) {
        return Status::resource_limit;
    }
}

// Every tile of the archive is checked as an ROI decode checks the tiles it
// selects (reference plan.cpp:238-279), then entered into the tile table with
// its absolute payload address and, for predictive tiles, its row offsets.
Status ArchiveSet::add(ConstBytes device_file, const ImageMetadata &m, std::uint32_t &id) noexcept {
    Impl &s = *impl_;
    if (s.archives.size() >= UINT32_MAX) return Status::resource_limit;
    std::uint64_t bound = 0;
    if (const Status st = compress_bound(m.image, m.tile, bound); st != Status::success) return st;
    if (const Status st = host::byte_range(device_file.data, device_file.size); st != Status::success) return st;
    if (device_file.data == nullptr || reinterpret_cast<std::uintptr_t>(device_file.data) % 4 != 0)
        return Status::invalid_argument;
    if (m.file_bytes > device_file.size) return Status::buffer_too_small;
    const std::uint64_t across = host::ceil_div(m.image.width, m.tile.width);
    const std::uint64_t down = host::ceil_div(m.image.height, m.tile.height);
    if (m.tiles.size() != across * down) return Status::malformed_data;
    if (const Status st = detail::accessible(device_file.data, s.device, s.pageable); st != Status::success) return st;
    hipPointerAttribute_t attr{};
    bool host = false;
    if (hipPointerGetAttributes(&attr, device_file.data) == hipSuccess) host = attr.type == hipMemoryTypeHost;
    else static_cast<void>(hipGetLastError());
    try {
        const std::size_t tiles_before = s.tiles.size(), rows_before = s.rows.size();
        auto rollback = [&](Status st) {
            s.tiles.resize(tiles_before);
            s.rows.resize(rows_before);
            return st;
        };
        if (tiles_before + m.tiles.size() > UINT32_MAX) return Status::resource_limit;
        const auto *base = static_cast<const std::uint8_t *>(device_file.data);
        for (std::uint64_t index = 0; index < m.tiles.size(); ++index) {
            if (const Status st = detail::check_tile(m, index); st != Status::success) return rollback(st);
            const TileMetadata &t = m.tiles[static_cast<std::size_t>(index)];
            tsk_decode_task task{};
            task.payload = base + t.offset;
            task.column = static_cast<std::uint32_t>(index % across);
            task.row = static_cast<std::uint32_t>(index / across);
            task.image = 0; // request mode: the request is the image
            task.rows = TSK_NO_ROWS;
            task.words = static_cast<std::uint32_t>(t.length / 4);
            task.shape = TSK_TASK_SHAPE(static_cast<std::uint32_t>(t.mode), t.height, t.width);
            if (t.mode == TileMode::predict_bitpack) {
                if (s.rows.size() + t.height > UINT32_MAX) return rollback(Status::resource_limit);
                task.rows = static_cast<std::uint32_t>(s.rows.size());
                for (std::uint64_t r = 0; r < t.height; ++r) {
                    const SegmentMetadata &g = m.segments[static_cast<std::size_t>(t.first_segment + r)];
                    s.rows.push_back(static_cast<std::uint16_t>((g.payload_offset - format::segment_header_bytes) / 4));
                }
            }
            s.tiles.push_back(task);
        }
        Archive a;
        a.info.image = m.image;
        a.info.tile = m.tile;
        a.info.tile_count = m.tiles.size();
        a.info.file_bytes = m.file_bytes;
        a.info.host = host ? 1u : 0u;
        a.bytes = base;
        a.tile_base = tiles_before;
        a.tile_columns = static_cast<std::uint32_t>(across);
        s.archives.push_back(a);
        s.any_host = s.any_host || host;
        id = static_cast<std::uint32_t>(s.archives.size() - 1);
        return Status::success;
    } catch (const std::bad_alloc &) {
        return Status::resource_limit;
    } catch (const std::length_error &) {
        return Status::resource_limit;
    }
}

Status ArchiveSet::add_device(ConstBytes device_file, const ImageSpec &image, TileShape tile,
                              const DeviceResult *producer, hipStream_t stream, std::uint32_t &id) noexcept {
    Impl &s = *impl_;
    if (s.archives.size() >= UINT32_MAX) return Status::resource_limit;
    std::uint64_t bound = 0;
    if (const Status st = compress_bound(image, tile, bound); st != Status::success) return st;
    if (const Status st = host::byte_range(device_file.data, device_file.size); st != Status::success) return st;
    if (device_file.data == nullptr || reinterpret_cast<std::uintptr_t>(device_file.data) % 4 != 0)
        return Status::invalid_argument;
    if (const Status st = detail::stream_device(stream, s.device); st != Status::success) return st;
    // Device memory only: tables of pinned host archives are built on the host.
    if (const Status st = detail::on_device(device_file.data, s.device); st != Status::success) return st;
    if (producer != nullptr &&
        (reinterpret_cast<std::uintptr_t>(producer) % 8 != 0 ||
         detail::accessible(producer, s.device, s.pageable) != Status::success))
        return Status::invalid_argument;
    const std::uint64_t across = host::ceil_div(image.width, tile.width);
    const std::uint64_t count = across * host::ceil_div(image.height, tile.height);
    const std::uint64_t rows = count * tile.height;
    tsk_index_args a{};
    a.src = static_cast<const std::uint8_t *>(device_file.data);
    a.producer = producer;
    a.capacity = device_file.size;
    a.bound = bound;
    a.tile_count = count;
    a.width = image.width;
    a.height = image.height;
    a.tile_columns = static_cast<std::uint32_t>(across);
    a.tile_w = tile.width;
    a.tile_h = tile.height;
    a.type = static_cast<std::uint32_t>(image.type);
    // The header first: only a valid header (whose index lies inside the
    // file) sizes the tables.
    void *record_storage = nullptr;
    if (hipMalloc(&record_storage, sizeof(tsk_index_record)) != hipSuccess) return detail::hip_failure();
    a.record = static_cast<tsk_index_record *>(record_storage);
    tsk_index_record record{};
    const auto read_record = [&] {
        return hipMemcpyAsync(&record, a.record, sizeof record, hipMemcpyDeviceToHost, stream) == hipSuccess &&
               hipStreamSynchronize(stream) == hipSuccess;
    };
    DeviceTable table;
    const auto release = [&](Status st) {
        static_cast<void>(hipFree(record_storage));
        if (table.storage != nullptr) static_cast<void>(hipFree(table.storage));
        return st;
    };
    if (tsk_index_header(&a, stream) != hipSuccess || !read_record()) return release(detail::hip_failure());
    if (record.header_status == TSK_INDEX_OTHER_IMAGE) return release(Status::invalid_argument);
    if (record.header_status != 0) return release(static_cast<Status>(record.header_status));
    if (s.tiles.size() + count > UINT32_MAX || s.rows.size() + rows > UINT32_MAX)
        return release(Status::resource_limit);
    table.tile_base = s.tiles.size();
    table.tiles = count;
    table.rows_base = s.rows.size();
    table.rows = rows;
    table.rows_at = align_up(count * sizeof(tsk_decode_task), 256);
    if (hipMalloc(&table.storage, table.rows_at + rows * sizeof(std::uint16_t)) != hipSuccess)
        return release(detail::hip_failure());
    auto *base = static_cast<std::uint8_t *>(table.storage);
    a.tasks = reinterpret_cast<tsk_decode_task *>(base);
    a.rows = reinterpret_cast<std::uint16_t *>(base + table.rows_at);
    a.rows_base = static_cast<std::uint32_t>(table.rows_base);
    if (tsk_index_tiles(&a, stream) != hipSuccess || !read_record()) return release(detail::hip_failure());
    if (record.tile_failures != 0) return release(Status::malformed_data);
    static_cast<void>(hipFree(record_storage));
    record_storage = nullptr;
    try {
        s.device_tables.push_back(table);
        try {
            s.tiles.resize(s.tiles.size() + count);
            s.rows.resize(s.rows.size() + rows);
            Archive ar;
            ar.info.image = image;
            ar.info.tile = tile;
            ar.info.tile_count = count;
            ar.info.file_bytes = record.file_bytes;
            ar.info.host = 0;
            ar.bytes = static_cast<const std::uint8_t *>(device_file.data);
            ar.tile_base = table.tile_base;
            ar.tile_columns = static_cast<std::uint32_t>(across);
            s.archives.push_back(ar);
        } catch (...) {
            s.tiles.resize(table.tile_base);
            s.rows.resize(table.rows_base);
            s.device_tables.pop_back();
            throw;
        }
    } catch (const std::bad_alloc &) {
        return release(Status::resource_limit);
    } catch (const std::length_error &) {
        return release(Status::resource_limit);
    }
    id = static_cast<std::uint32_t>(s.archives.size() - 1);
    return Status::success;
}

Status ArchiveSet::commit(hipStream_t stream) noexcept {
    Impl &s = *impl_;
    if (const Status st = detail::stream_device(stream, s.device); st != Status::success) return st;
    const std::uint64_t rows_at = align_up(s.tiles.size() * sizeof(tsk_decode_task), 256);
    const std::uint64_t total = rows_at + s.rows.size() * sizeof(std::uint16_t) + 256;
    void *storage = nullptr;
    if (hipMalloc(&storage, total) != hipSuccess) return detail::hip_failure();
    auto *base = static_cast<std::uint8_t *>(storage);
    bool copied = (s.tiles.empty() || hipMemcpyAsync(base, s.tiles.data(), s.tiles.size() * sizeof(tsk_decode_task),
                                                     hipMemcpyHostToDevice, stream) == hipSuccess) &&
                  (s.rows.empty() || hipMemcpyAsync(base + rows_at, s.rows.data(), s.rows.size() * sizeof(std::uint16_t),
                                                    hipMemcpyHostToDevice, stream) == hipSuccess);
    // Tables built on the device replace their placeholders (same stream, after the upload).
    for (const DeviceTable &t : s.device_tables) {
        if (!copied) break;
        const auto *from = static_cast<const std::uint8_t *>(t.storage);
        copied = hipMemcpyAsync(base + t.tile_base * sizeof(tsk_decode_task), from, t.tiles * sizeof(tsk_decode_task),
                                hipMemcpyDeviceToDevice, stream) == hipSuccess &&
                 (t.rows == 0 || hipMemcpyAsync(base + rows_at + t.rows_base * sizeof(std::uint16_t), from + t.rows_at,
                                                t.rows * sizeof(std::uint16_t), hipMemcpyDeviceToDevice,
                                                stream) == hipSuccess);
    }
    if (!copied || hipStreamSynchronize(stream) != hipSuccess) {
        static_cast<void>(hipFree(storage));
        return detail::hip_failure();
    }
    if (s.storage != nullptr) static_cast<void>(hipFree(s.storage));
    s.storage = storage;
    s.device_tiles = reinterpret_cast<const tsk_decode_task *>(base);
    s.device_rows = reinterpret_cast<const std::uint16_t *>(base + rows_at);
    s.committed = s.archives.size();
    s.row_count = s.rows.size();
    return Status::success;
}

std::uint32_t ArchiveSet::size() const noexcept { return static_cast<std::uint32_t>(impl_->archives.size()); }

Status ArchiveSet::info(std::uint32_t id, ArchiveInfo &out) const noexcept {
    if (id >= impl_->archives.size()) return Status::invalid_argument;
    out = impl_->archives[id].info;
    return Status::success;
}

// ---------------------------------------------------------------------------

namespace {

// One batch in flight: a stream, the batch's tables staged in pinned host
// memory and copied to the device by a kernel, the final results written back
// into pinned host memory by kernels, and the event that marks its end.
//
// Table layout (host and device alike): work counter | DeviceResult[n]
// (region requests first, then encode requests) | tsk_decode_image[regions] |
// tsk_request[regions] | request of every decode task (batches decoded by
// blocks: of every block) | staging block of every decode task |
// tsk_encode_image[encodes] (uint16 images, then float32) | encode window
// entries | per window, the entry of each block of 64 tiles.  Block records,
// built on the device, have a device buffer of their own.
struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    std::uint8_t *host = nullptr;      // tables, then the published results
    std::uint8_t *host_view = nullptr; // the same memory as the device addresses it
    std::uint8_t *device = nullptr;
    std::uint8_t *scratch = nullptr;   // encoder workspace (look-back buffers, carried cursor)
    std::uint8_t *staging = nullptr;   // payloads of tiles of host archives
    tsk_block *records = nullptr;      // block records of a batch decoded by blocks
    std::uint64_t published = 0;       // offset of the published results in `host`
    std::vector<std::uint64_t> order;  // request index of each result of the current batch
    std::uint64_t bytes = 0;           // output bytes admitted
    bool busy = false;
    bool large = false;                // holds a batch of large requests
};

constexpr std::uint64_t kCounterBytes = 256;

// Block records per slot (416 bytes each, 27 MB): a batch that needs more is
// decoded tile by tile.
constexpr std::uint64_t kBlockRecords = 65536;

// The block geometry of a region request, as the host block planner makes it
// (plan_blocks): across x down tiles whose output is at most 64 rows of at
// most 1 KiB.  False when the block decoder cannot take the request: a host
// archive (its payloads are staged tile by tile), or a geometry whose blocks
// could exceed the block stage with every tile RAW and 16-byte misaligned.
bool block_geometry(const ArchiveInfo &info, bool calibrated, std::uint32_t &across, std::uint32_t &down) {
    if (info.host) return false;
    const std::uint32_t stored = info.image.type == PixelType::uint16 ? 2u : 4u;
    const std::uint32_t out = calibrated ? 4u : stored;
    const std::uint32_t tw = info.tile.width, th = info.tile.height;
    if (tw == 0 || th == 0) return false;
    across = std::max(1u, 1024u / (tw * out));
    down = std::max(1u, 64u / th);
    if (across * down > 8) return false;
    const std::uint64_t per_tile = (std::uint64_t{tw} * th * stored + 3) / 4 * 4 / 16 + 2;
    return across * down * per_tile <= TSK_BLOCK_STAGE_USED;
}

struct Layout {
    std::uint64_t results, images, requests, index, staged, encode, entries, blocks, total;
};

Layout layout_for(std::uint64_t regions, std::uint64_t encodes, std::uint64_t tasks, std::uint64_t entries,
                  std::uint64_t blocks) {
    Layout l{};
    l.results = kCounterBytes;
    l.images = align_up(l.results + (regions + encodes) * sizeof(DeviceResult), 256);
    l.requests = align_up(l.images + regions * sizeof(tsk_decode_image), 256);
    l.index = align_up(l.requests + regions * sizeof(tsk_request), 256);
    l.staged = align_up(l.index + tasks * sizeof(std::uint32_t), 256);
    l.encode = align_up(l.staged + tasks * sizeof(std::uint32_t), 256);
    l.entries = align_up(l.encode + encodes * sizeof(tsk_encode_image), 256);
    l.blocks = align_up(l.entries + entries * sizeof(tsk_window_entry), 16);
    l.total = align_up(l.blocks + blocks * sizeof(std::uint32_t), 16);
    return l;
}

// Encoder workspace per slot: the carried cursor, then two look-back buffers
// of one window of tiles (alternating by window parity).
struct Scratch {
    static constexpr std::uint64_t carry = 0;
    static constexpr std::uint64_t state = 256;
    static std::uint64_t total(std::uint32_t window) { return state + 2 * std::uint64_t{window} * 8; }
};

// What a request costs a batch.
struct Extent {
    std::uint32_t col0 = 0, row0 = 0, cols = 0, rows = 0; // region: its tiles
    std::uint64_t tasks = 0;   // tiles decoded or encoded
    std::uint64_t bytes = 0;   // output bytes (encode: the compress bound)
    std::uint64_t decoded = 0; // region pixels x stored sample bytes, or encoded input bytes
    std::uint64_t staged = 0;  // region of a host archive: staging bytes (16-byte blocks)
};

} // namespace

struct Server::Impl {
    const ArchiveSet::Impl *set = nullptr;
    ServerOptions options;
    int device = 0;
    int cus = 0;
    int per_cu[5] = {};
    int per_cu_blocks[5] = {}; // the block decoder's resident workgroups per CU, per kind
    bool pageable = false;
    std::vector<Slot> slots;
    std::vector<Extent> extents; // per request of the current replay

    ~Impl() {
        for (Slot &s : slots) {
            if (s.stream != nullptr) static_cast<void>(hipStreamSynchronize(s.stream));
            if (s.done != nullptr) static_cast<void>(hipEventDestroy(s.done));
            if (s.stream != nullptr) static_cast<void>(hipStreamDestroy(s.stream));
            if (s.host != nullptr) static_cast<void>(hipHostFree(s.host));
            if (s.device != nullptr) static_cast<void>(hipFree(s.device));
            if (s.scratch != nullptr) static_cast<void>(hipFree(s.scratch));
            if (s.staging != nullptr) static_cast<void>(hipFree(s.staging));
            if (s.records != nullptr) static_cast<void>(hipFree(s.records));
        }
    }

    [[nodiscard]] Layout capacity() const noexcept {
        const std::uint64_t n = options.max_batch;
        const std::uint64_t windows = 2 * (options.max_tasks / options.encode_window + 2);
        return layout_for(n, n, options.max_tasks, n + windows, options.max_tasks / 64 + windows);
    }
    Status extent(const ServeRequest &r, Extent &e) const noexcept;
    // Waits for every batch in flight and frees the slots.
    void drain() noexcept {
        for (Slot &slot : slots)
            if (slot.busy) static_cast<void>(hipStreamSynchronize(slot.stream));
        for (Slot &slot : slots) slot.busy = false;
    }
    // Allocates its window tables: may throw std::bad_alloc (replay turns it into resource_limit).
    Status launch(Slot &slot, const ServeRequest *requests, const std::uint64_t *ids, std::uint64_t count, bool large);
};

Server::Server(std::unique_ptr<Impl> impl) noexcept : impl_(std::move(impl)) {}
Server::~Server() = default;

Status Server::create(const ArchiveSet &archives, const ServerOptions &options, std::unique_ptr<Server> &server) noexcept {
    const ArchiveSet::Impl *set = archives.impl();
    if (set == nullptr || set->committed == 0 || set->committed != set->archives.size()) return Status::invalid_argument;
    if (options.streams == 0 || options.streams > 16 || options.max_batch == 0 || options.max_tasks == 0 ||
        options.admission_bytes == 0 || options.encode_window < 64 || options.encode_window > (1u << 20) ||
        options.staging_bytes < 16 || options.staging_bytes / 16 > 0xfffffffeull)
        return Status::invalid_argument;
    const LaunchShape &sh = options.shape;
    if (!tsk_decode_supported(sh.group, sh.block, sh.depth)) return Status::invalid_argument;
    try {
        auto impl = std::make_unique<Impl>();
        impl->set = set;
        impl->options = options;
        impl->pageable = set->pageable;
        if (hipGetDevice(&impl->device) != hipSuccess || impl->device != set->device ||
            hipDeviceGetAttribute(&impl->cus, hipDeviceAttributeMultiprocessorCount, impl->device) != hipSuccess)
            return detail::hip_failure();
        for (std::uint32_t kind = 0; kind < 5; ++kind)
            if (tsk_decode_occupancy(kind, sh.group, sh.block, sh.depth, &impl->per_cu[kind]) != hipSuccess ||
                tsk_decode_occupancy(kind, 0, 512, 2, &impl->per_cu_blocks[kind]) != hipSuccess)
                return detail::hip_failure();
        if (tsk_preload(0) != hipSuccess || tsk_preload(1) != hipSuccess || tsk_preload(3) != hipSuccess)
            return detail::hip_failure();
        const Layout cap = impl->capacity();
        impl->slots.resize(options.streams);
        for (Slot &s : impl->slots) {
            if (hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking) != hipSuccess) {
                s.stream = nullptr;
                return detail::hip_failure();
            }
            if (hipEventCreateWithFlags(&s.done, hipEventDisableTiming) != hipSuccess) {
                s.done = nullptr;
                return detail::hip_failure();
            }
            void *h = nullptr, *hv = nullptr, *d = nullptr, *w = nullptr;
            s.published = align_up(cap.total, 256);
            const std::uint64_t host_bytes = s.published + 2 * std::uint64_t{options.max_batch} * sizeof(DeviceResult);
            // Coherent: the device reads the tables and writes the results over
            // the bus while the host polls.
            if (hipHostMalloc(&h, host_bytes, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess)
                return detail::hip_failure();
            s.host = static_cast<std::uint8_t *>(h);
            if (hipHostGetDevicePointer(&hv, h, 0) != hipSuccess) return detail::hip_failure();
            s.host_view = static_cast<std::uint8_t *>(hv);
            if (hipMalloc(&d, align_up(cap.total, 256)) != hipSuccess) return detail::hip_failure();
            s.device = static_cast<std::uint8_t *>(d);
            if (hipMalloc(&w, Scratch::total(options.encode_window)) != hipSuccess) return detail::hip_failure();
            s.scratch = static_cast<std::uint8_t *>(w);
            if (options.block_tasks != 0) {
                void *b = nullptr;
                if (hipMalloc(&b, kBlockRecords * sizeof(tsk_block)) != hipSuccess) return detail::hip_failure();
                s.records = static_cast<tsk_block *>(b);
            }
            if (set->any_host) {
                void *g = nullptr;
                if (hipMalloc(&g, options.staging_bytes) != hipSuccess) return detail::hip_failure();
                s.staging = static_cast<std::uint8_t *>(g);
            }
            s.order.reserve(options.max_batch);
        }
        server.reset(new Server(std::move(impl)));
        return Status::success;
    } catch (const std::bad_alloc &) {
        return Status::resource_limit;
    } catch (const std::length_error &) {
        return Status::resource_limit;
    }
}

Status Server::Impl::extent(const ServeRequest &r, Extent &e) const noexcept {
    if (r.kind == RequestKind::encode) {
        // The checks of a compress item (reference plan.cpp:351-465, draft_compress here).
        std::uint64_t bound = 0;
        if (const Status st = host::check_view(r.input); st != Status::success) return st;
        if (const Status st = compress_bound(r.input.image, r.tile, bound); st != Status::success) return st;
        if (const Status st = host::byte_range(r.compressed.data, r.compressed.size); st != Status::success) return st;
        if (r.compressed.data == nullptr || reinterpret_cast<std::uintptr_t>(r.compressed.data) % 4 != 0)
            return Status::invalid_argument;
        if (r.compressed.size < bound) return Status::buffer_too_small;
        e.tasks = host::ceil_div(r.input.image.width, r.tile.width) * host::ceil_div(r.input.image.height, r.tile.height);
        if (e.tasks > options.max_tasks) return Status::resource_limit;
        const std::uint64_t s = r.input.image.type == PixelType::uint16 ? 2 : 4;
        e.bytes = bound;
        e.decoded = std::uint64_t{r.input.image.width} * r.input.image.height * s;
        return Status::success;
    }
    if (r.kind != RequestKind::region) return Status::invalid_argument;
    if (r.archive >= set->committed || r.calibrated > 1) return Status::invalid_argument;
    const Archive &a = set->archives[r.archive];
    const ImageSpec &image = a.info.image;
    const Roi &roi = r.roi;
    if (roi.width == 0 || roi.height == 0 || std::uint64_t{roi.x} + roi.width > image.width ||
        std::uint64_t{roi.y} + roi.height > image.height)
        return Status::invalid_roi;
    if (const Status st = host::check_view(host::as_const(r.output)); st != Status::success) return st;
    if (r.output.image.type != (r.calibrated ? PixelType::float32 : image.type) || r.output.image.width != roi.width ||
        r.output.image.height != roi.height)
        return Status::invalid_argument;
    const TileShape &t = a.info.tile;
    e.col0 = roi.x / t.width;
    e.row0 = roi.y / t.height;
    e.cols = (roi.x + roi.width - 1) / t.width - e.col0 + 1;
    e.rows = (roi.y + roi.height - 1) / t.height - e.row0 + 1;
    e.tasks = std::uint64_t{e.cols} * e.rows;
    if (e.tasks > options.max_tasks) return Status::resource_limit;
    const std::uint64_t area = std::uint64_t{roi.width} * roi.height;
    const std::uint64_t stored = image.type == PixelType::uint16 ? 2 : 4;
    e.decoded = area * stored;
    e.bytes = area * (r.calibrated ? 4 : stored);
    if (a.info.host) {
        for (std::uint32_t dr = 0; dr < e.rows; ++dr)
            for (std::uint32_t dc = 0; dc < e.cols; ++dc)
                e.staged += 16 * staged_blocks(set->tiles[a.tile_base + std::uint64_t{e.row0 + dr} * a.tile_columns +
                                                          e.col0 + dc]);
        if (e.staged > options.staging_bytes) return Status::resource_limit;
    }
    return Status::success;
}

Status Server::check(const ServeRequest &request) const noexcept {
    Extent e;
    return impl_->extent(request, e);
}

// Stages the batch's tables, copies them to the device with one kernel, runs
// the region decode in request mode and the encode windows, and copies the
// final results into pinned memory; then marks the end.
Status Server::Impl::launch(Slot &slot, const ServeRequest *requests, const std::uint64_t *ids, std::uint64_t count,
                            bool large) {
    std::uint64_t regions = 0, tasks = 0;
    std::uint64_t encodes[2] = {0, 0}; // uint16, float32
    std::vector<std::uint64_t> enc_tiles[2];
    for (std::uint64_t k = 0; k < count; ++k) {
        const ServeRequest &r = requests[ids[k]];
        if (r.kind == RequestKind::region) {
            ++regions;
            tasks += extents[ids[k]].tasks;
        } else {
            const int g = r.input.image.type == PixelType::uint16 ? 0 : 1;
            ++encodes[g];
            enc_tiles[g].push_back(extents[ids[k]].tasks);
        }
    }
    // Encode windows per pixel type; the entries of both groups in one table.
    detail::WindowSplit split[2] = {detail::split_into_windows(enc_tiles[0], options.encode_window),
                                    detail::split_into_windows(enc_tiles[1], options.encode_window)};
    std::vector<tsk_window_entry> entries = std::move(split[0].entries);
    const std::size_t group1_entries = entries.size();
    entries.insert(entries.end(), split[1].entries.begin(), split[1].entries.end());
    for (detail::WindowSpan &w : split[1].windows) w.first_entry += group1_entries;
    std::vector<std::uint32_t> blocks;
    std::vector<std::size_t> blocks_first[2];
    for (int g = 0; g < 2; ++g) {
        for (const detail::WindowSpan &w : split[g].windows) {
            blocks_first[g].push_back(blocks.size());
            detail::append_block_owners(entries.data() + w.first_entry, w.entry_count, w.tile_count, blocks);
        }
    }
    const Layout l = layout_for(regions, encodes[0] + encodes[1], tasks, entries.size(), blocks.size());

    std::uint8_t *h = slot.host;
    std::memset(h, 0, l.images); // work counter and results start at zero
    auto *staged = reinterpret_cast<std::uint32_t *>(h + l.staged);
    bool staging = false;
    for (std::uint64_t k = 0; k < count && !staging; ++k) staging = extents[ids[k]].staged != 0;
    // A batch of at least block_tasks tiles, all of device archives whose
    // blocks fit the block stage, is decoded by blocks: fewer, larger units
    // (up to 8 tiles), each writing whole 1 KiB runs of its output rows.
    bool by_blocks = slot.records != nullptr && regions != 0 && !staging && tasks >= options.block_tasks;
    std::uint64_t block_total = 0;
    for (std::uint64_t k = 0; k < count && by_blocks; ++k) {
        const ServeRequest &r = requests[ids[k]];
        if (r.kind != RequestKind::region) continue;
        const Extent &e = extents[ids[k]];
        std::uint32_t across = 0, down = 0;
        by_blocks = block_geometry(set->archives[r.archive].info, r.calibrated != 0, across, down);
        block_total += host::ceil_div(e.cols, across) * host::ceil_div(e.rows, down);
    }
    by_blocks = by_blocks && block_total <= kBlockRecords;
    std::uint64_t block_next = 0;
    std::uint64_t staged_block = 0;
    auto *images = reinterpret_cast<tsk_decode_image *>(h + l.images);
    auto *reqs = reinterpret_cast<tsk_request *>(h + l.requests);
    auto *index = reinterpret_cast<std::uint32_t *>(h + l.index);
    auto *enc = reinterpret_cast<tsk_encode_image *>(h + l.encode);
    if (!entries.empty())
        std::memcpy(h + l.entries, entries.data(), entries.size() * sizeof(tsk_window_entry));
    if (!blocks.empty()) std::memcpy(h + l.blocks, blocks.data(), blocks.size() * sizeof(std::uint32_t));
    slot.order.assign(regions + encodes[0] + encodes[1], 0);
    bool u16 = false, f32 = false, cal = false, plain = false;
    std::uint64_t task = 0, j = 0, e_next[2] = {0, encodes[0]};
    for (std::uint64_t k = 0; k < count; ++k) {
        const ServeRequest &r = requests[ids[k]];
        const Extent &e = extents[ids[k]];
        if (r.kind == RequestKind::encode) {
            const int g = r.input.image.type == PixelType::uint16 ? 0 : 1;
            const std::uint64_t at = e_next[g]++;
            tsk_encode_image &im = enc[at];
            im = tsk_encode_image{};
            im.in = static_cast<const std::uint8_t *>(r.input.data);
            im.out = static_cast<std::uint8_t *>(r.compressed.data);
            im.in_pitch = r.input.pitch_bytes;
            im.tile_count = e.tasks;
            im.width = r.input.image.width;
            im.height = r.input.image.height;
            im.tile_columns = static_cast<std::uint32_t>(host::ceil_div(r.input.image.width, r.tile.width));
            im.tile_w = r.tile.width;
            im.tile_h = r.tile.height;
            im.type = static_cast<std::uint32_t>(r.input.image.type);
            slot.order[regions + at] = ids[k];
            continue;
        }
        const Archive &a = set->archives[r.archive];
        tsk_decode_image &im = images[j];
        im = tsk_decode_image{};
        im.src = a.bytes;
        im.out = static_cast<std::uint8_t *>(r.output.data);
        im.src_size = a.info.file_bytes;
        im.out_pitch = r.output.pitch_bytes;
        im.width = a.info.image.width;
        im.height = a.info.image.height;
        im.roi_x = r.roi.x;
        im.roi_y = r.roi.y;
        im.roi_w = r.roi.width;
        im.roi_h = r.roi.height;
        im.tile_columns = a.tile_columns;
        im.tile_w = a.info.tile.width;
        im.tile_h = a.info.tile.height;
        im.type = static_cast<std::uint8_t>(a.info.image.type);
        im.calibrated = static_cast<std::uint8_t>(r.calibrated);
        im.offset = r.calibrated ? r.offset : 0.0F;
        im.gain = r.calibrated ? r.gain : 1.0F;
        tsk_request &q = reqs[j];
        q.first_task = task;
        q.tile_base = static_cast<std::uint32_t>(a.tile_base);
        q.tile_columns = a.tile_columns;
        q.col0 = e.col0;
        q.row0 = e.row0;
        q.cols = e.cols;
        q.rows = e.rows;
        q.first_block = static_cast<std::uint32_t>(block_next);
        q.across = 0;
        q.down = 0;
        if (by_blocks) {
            std::uint32_t across = 0, down = 0;
            static_cast<void>(block_geometry(a.info, r.calibrated != 0, across, down));
            q.across = static_cast<std::uint16_t>(across);
            q.down = static_cast<std::uint16_t>(down);
            const std::uint64_t nb = host::ceil_div(e.cols, across) * host::ceil_div(e.rows, down);
            std::fill_n(index + block_next, nb, static_cast<std::uint32_t>(j));
            block_next += nb;
        } else {
            std::fill_n(index + task, e.tasks, static_cast<std::uint32_t>(j));
        }
        if (staging) {
            // Tiles of a host archive get consecutive 16-byte blocks of staging.
            for (std::uint64_t k = 0; k < e.tasks; ++k) {
                if (!a.info.host) {
                    staged[task + k] = TSK_NOT_STAGED;
                    continue;
                }
                const std::uint64_t dr = k / e.cols, dc = k % e.cols;
                staged[task + k] = static_cast<std::uint32_t>(staged_block);
                staged_block += staged_blocks(set->tiles[a.tile_base + (e.row0 + dr) * a.tile_columns + e.col0 + dc]);
            }
        }
        task += e.tasks;
        slot.order[j] = ids[k];
        ++j;
        u16 = u16 || a.info.image.type == PixelType::uint16;
        f32 = f32 || a.info.image.type == PixelType::float32;
        cal = cal || r.calibrated != 0;
        plain = plain || r.calibrated == 0;
    }

    // Kernels read the staged tables from pinned memory and write the final
    // results back into it: they start sooner than runtime copies.
    if (tsk_copy_blocks(slot.device, slot.host_view, l.total / 16, slot.stream) != hipSuccess)
        return detail::hip_failure();
    if (regions != 0) {
        std::uint32_t kind = TSK_KIND_MIXED;
        if (u16 != f32 && cal != plain)
            kind = u16 ? (cal ? TSK_KIND_U16_CAL : TSK_KIND_U16) : (cal ? TSK_KIND_F32_CAL : TSK_KIND_F32);
        const LaunchShape &sh = options.shape;
        std::uint64_t resident = static_cast<std::uint64_t>(cus) *
                                 static_cast<std::uint64_t>(std::max(1, by_blocks ? per_cu_blocks[kind] : per_cu[kind]));
        // A batch of large requests is a persistent grid that holds its CUs to
        // the end: it takes (streams - 1) / streams of them, so batches of small
        // requests on the other streams find CUs free.
        if (large && slots.size() > 1) resident = std::max<std::uint64_t>(1, resident * (slots.size() - 1) / slots.size());
        const std::uint64_t units = by_blocks ? block_total : tasks;
        const std::uint64_t grid = sh.grid != 0 && !by_blocks ? sh.grid : std::min<std::uint64_t>(resident, units);
        const std::uint32_t chunk = by_blocks ? 1u : sh.chunk;
        const bool queued = chunk != 0 && units > grid;
        tsk_decode_args args{};
        args.images = reinterpret_cast<const tsk_decode_image *>(slot.device + l.images);
        args.tasks = set->device_tiles;
        args.rows = set->device_rows;
        args.results = slot.device + l.results;
        args.queue = queued ? reinterpret_cast<std::uint64_t *>(slot.device) : nullptr;
        args.task_count = units;
        args.row_count = set->row_count;
        args.image_count = static_cast<std::uint32_t>(regions);
        args.kind = kind;
        args.group = by_blocks ? 0u : sh.group;
        args.block = by_blocks ? 512u : sh.block;
        args.grid = static_cast<std::uint32_t>(std::max<std::uint64_t>(1, grid));
        args.chunk = queued ? chunk : 0;
        args.depth = by_blocks ? 2u : sh.depth;
        args.run = by_blocks ? 0u : sh.run;
        if (by_blocks) {
            tsk_block_build_args build{};
            build.requests = reinterpret_cast<const tsk_request *>(slot.device + l.requests);
            build.block_index = reinterpret_cast<const std::uint32_t *>(slot.device + l.index);
            build.images = args.images;
            build.tiles = set->device_tiles;
            build.blocks = slot.records;
            build.block_count = block_total;
            if (tsk_build_blocks(&build, slot.stream) != hipSuccess) return detail::hip_failure();
            args.blocks = slot.records;
        } else {
            args.request_count = static_cast<std::uint32_t>(regions);
            args.requests = reinterpret_cast<const tsk_request *>(slot.device + l.requests);
            args.request_index = reinterpret_cast<const std::uint32_t *>(slot.device + l.index);
        }
        if (staging) {
            args.staged = reinterpret_cast<const std::uint32_t *>(slot.device + l.staged);
            args.staging = slot.staging;
            if (tsk_stage_payloads(&args, slot.stream) != hipSuccess) return detail::hip_failure();
        }
        if (tsk_decode(&args, slot.stream) != hipSuccess ||
            tsk_results_publish(args.images, static_cast<std::uint32_t>(regions), args.results,
                                slot.host_view + slot.published, slot.stream) != hipSuccess)
            return detail::hip_failure();
    }
    if (encodes[0] + encodes[1] != 0) {
        tsk_encode_args a{};
        a.run = detail::encode_run();
        a.state = reinterpret_cast<std::uint64_t *>(slot.scratch + Scratch::state);
        a.carry = reinterpret_cast<std::uint64_t *>(slot.scratch + Scratch::carry);
        a.capacity = options.encode_window;
        if (tsk_encode_reset(&a, slot.stream) != hipSuccess) return detail::hip_failure();
        // Window parity alternates over the whole sequence (both pixel types):
        // each window's look-back buffer was cleared by the window before it,
        // or by the reset for the first.
        std::uint32_t sequence = 0;
        for (int g = 0; g < 2; ++g) {
            if (encodes[g] == 0) continue;
            a.images = reinterpret_cast<const tsk_encode_image *>(slot.device + l.encode) + (g == 0 ? 0 : encodes[0]);
            a.results = slot.device + l.results + (regions + (g == 0 ? 0 : encodes[0])) * sizeof(DeviceResult);
            a.type = g == 0 ? TSK_UINT16 : TSK_FLOAT32;
            for (std::size_t w = 0; w < split[g].windows.size(); ++w) {
                const detail::WindowSpan &span = split[g].windows[w];
                a.entries = reinterpret_cast<const tsk_window_entry *>(slot.device + l.entries) + span.first_entry;
                a.blocks = reinterpret_cast<const std::uint32_t *>(slot.device + l.blocks) + blocks_first[g][w];
                a.entry_count = span.entry_count;
                a.tile_count = span.tile_count;
                a.parity = sequence++ & 1u;
                if (tsk_encode_window(&a, slot.stream) != hipSuccess) return detail::hip_failure();
            }
        }
    }
    if (encodes[0] + encodes[1] != 0 &&
        tsk_copy_blocks(slot.host_view + slot.published + regions * sizeof(DeviceResult),
                        slot.device + l.results + regions * sizeof(DeviceResult), encodes[0] + encodes[1],
                        slot.stream) != hipSuccess)
        return detail::hip_failure();
    if (hipEventRecord(slot.done, slot.stream) != hipSuccess) return detail::hip_failure();
    slot.busy = true;
    return Status::success;
}

Status Server::replay(const ServeRequest *requests, const std::uint64_t *arrival_ns, std::uint64_t count,
                      std::uint64_t *latency_ns, DeviceResult *results, ServerStats &stats) noexcept {
    Impl &s = *impl_;
    stats = ServerStats{};
    if (count == 0) return Status::success;
    if (requests == nullptr || arrival_ns == nullptr) return Status::invalid_argument;
    int current = 0;
    if (hipGetDevice(&current) != hipSuccess) return detail::hip_failure();
    if (current != s.device) return Status::invalid_argument;
    try {
        // Validate everything before the clock starts.
        s.extents.assign(count, Extent{});
        std::vector<detail::Span> writable, readable;
        writable.reserve(count);
        for (std::uint64_t i = 0; i < count; ++i) {
            if (i > 0 && arrival_ns[i] < arrival_ns[i - 1]) return Status::invalid_argument;
            const ServeRequest &r = requests[i];
            if (const Status st = s.extent(r, s.extents[i]); st != Status::success) return st;
            if (r.kind == RequestKind::encode) {
                if (const Status st = detail::accessible(r.input.data, s.device, s.pageable); st != Status::success)
                    return st;
                if (const Status st = detail::accessible(r.compressed.data, s.device, s.pageable); st != Status::success)
                    return st;
                const auto in = reinterpret_cast<std::uintptr_t>(r.input.data);
                const auto out = reinterpret_cast<std::uintptr_t>(r.compressed.data);
                readable.push_back({in, in + r.input.capacity_bytes});
                writable.push_back({out, out + r.compressed.size});
                continue;
            }
            std::uint64_t span = 0;
            if (const Status st = image_span(r.output.image, r.output.pitch_bytes, span); st != Status::success)
                return st;
            if (const Status st = detail::accessible(r.output.data, s.device, s.pageable); st != Status::success)
                return st;
            const auto at = reinterpret_cast<std::uintptr_t>(r.output.data);
            writable.push_back({at, at + span});
        }
        for (const Archive &a : s.set->archives) {
            const auto at = reinterpret_cast<std::uintptr_t>(a.bytes);
            readable.push_back({at, at + a.info.file_bytes});
        }
        detail::Footprint footprint;
        if (!footprint.admit(std::move(writable), std::move(readable))) return Status::invalid_argument;

        using clock = std::chrono::steady_clock;
        const auto start = clock::now();
        auto elapsed = [&] {
            return static_cast<std::uint64_t>(std::chrono::duration_cast<std::chrono::nanoseconds>(clock::now() - start).count());
        };
        // Arrived requests wait in two queues by size class, in arrival order.
        std::vector<std::uint64_t> queue[2]; // 0: small, 1: large (output bytes >= large_bytes)
        std::size_t head[2] = {0, 0};
        std::uint64_t arrived = 0, dispatched = 0, finished = 0, in_flight = 0;
        std::uint64_t last_done = 0;
        std::size_t large_busy = 0;
        Status failure = Status::success;
        while (finished < count) {
            const std::uint64_t now = elapsed();
            for (; arrived < count && arrival_ns[arrived] <= now; ++arrived)
                queue[s.extents[arrived].bytes >= s.options.large_bytes ? 1 : 0].push_back(arrived);
            bool progress = false;
            for (Slot &slot : s.slots) {
                if (!slot.busy) continue;
                const hipError_t q = hipEventQuery(slot.done);
                if (q == hipErrorNotReady) continue;
                if (q != hipSuccess) {
                    failure = detail::hip_failure();
                    break;
                }
                const std::uint64_t t = elapsed();
                const auto *res = reinterpret_cast<const volatile DeviceResult *>(slot.host + slot.published);
                for (std::size_t k = 0; k < slot.order.size(); ++k) {
                    const std::uint64_t i = slot.order[k];
                    const Status st = res[k].status;
                    if (latency_ns != nullptr) latency_ns[i] = t - arrival_ns[i];
                    if (results != nullptr) {
                        results[i].status = st;
                        results[i].reserved = 0;
                        results[i].bytes_written = res[k].bytes_written;
                    }
                    if (st != Status::success) ++stats.failed;
                    stats.decoded_bytes += s.extents[i].decoded;
                }
                finished += slot.order.size();
                in_flight -= slot.bytes;
                last_done = t;
                slot.busy = false;
                if (slot.large) --large_busy;
                slot.large = false;
                progress = true;
            }
            if (failure != Status::success) break;
            for (Slot &slot : s.slots) {
                if (slot.busy || dispatched == arrived) continue;
                // One class per batch: the older head, but large batches
                // leave at least one stream to small requests.
                const bool small_waiting = head[0] < queue[0].size(), large_waiting = head[1] < queue[1].size();
                const bool large_room = s.slots.size() == 1 || large_busy + 1 < s.slots.size();
                int q = 0;
                if (small_waiting && large_waiting && large_room) q = queue[1][head[1]] < queue[0][head[0]] ? 1 : 0;
                else if (!small_waiting && large_waiting && large_room) q = 1;
                else if (!small_waiting) break;
                // Admit that class's requests in order while the batch and the
                // device have room; one request is always admitted alone.
                const std::uint64_t *ids = queue[q].data() + head[q];
                const std::uint64_t waiting = queue[q].size() - head[q];
                std::uint64_t n = 0, tasks = 0, bytes = 0, staged = 0;
                bool held = false;
                while (n < waiting && n < s.options.max_batch) {
                    const Extent &e = s.extents[ids[n]];
                    if (tasks + e.tasks > s.options.max_tasks) break;
                    if (staged + e.staged > s.options.staging_bytes) break;
                    if (in_flight + bytes + e.bytes > s.options.admission_bytes && (n != 0 || in_flight != 0)) {
                        held = true;
                        break;
                    }
                    tasks += e.tasks;
                    bytes += e.bytes;
                    staged += e.staged;
                    ++n;
                }
                if (n == 0) break;
                if (const Status st = s.launch(slot, requests, ids, n, q == 1); st != Status::success) {
                    failure = st;
                    break;
                }
                slot.bytes = bytes;
                slot.large = q == 1;
                if (slot.large) ++large_busy;
                in_flight += bytes;
                stats.peak_in_flight = std::max(stats.peak_in_flight, in_flight);
                if (held) ++stats.admission_holds;
                head[q] += n;
                dispatched += n;
                ++stats.batches;
                progress = true;
            }
            if (failure != Status::success) break;
            if (!progress && dispatched == arrived && arrived < count && in_flight == 0) {
                // Idle until the next arrival.
                const std::uint64_t wait = arrival_ns[arrived] - std::min(arrival_ns[arrived], elapsed());
                if (wait > 200000) std::this_thread::sleep_for(std::chrono::nanoseconds(wait - 100000));
            }
        }
        if (failure != Status::success) {
            s.drain();
            return failure;
        }
        stats.requests = count;
        stats.seconds = static_cast<double>(last_done - std::min(last_done, arrival_ns[0])) * 1e-9;
        return Status::success;
    } catch (const std::bad_alloc &) {
        s.drain();
        return Status::resource_limit;
    } catch (const std::length_error &) {
        s.drain();
        return Status::resource_limit;
    }
}

Status Server::run_batch(const ServeRequest *requests, std::uint64_t count, DeviceResult *results) noexcept {
    try {
        std::vector<std::uint64_t> arrivals(count, 0);
        ServerStats stats;
        return replay(requests, arrivals.data(), count, nullptr, results, stats);
    } catch (const std::bad_alloc &) {
        return Status::resource_limit;
    }
}

} // namespace tilestream::hip
