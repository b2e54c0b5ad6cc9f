// A CPU stand-in for the HIP runtime calls and kernel launchers the host
// planner makes, so that plan.cpp, queue.cpp and serve.cpp run whole (under
// AddressSanitizer + UBSan, built by test_host_sanitizers.py) on a machine
// without a GPU.  TEST INFRASTRUCTURE ONLY.
//
// "Device" memory is host memory tracked in a registry, so pointer queries
// answer as the runtime would.  The launchers do no decoding: they check that
// every table entry a kernel would read lies inside the tables the planner
// built (task, row, block and window indices, payload spans inside their
// source) and write the result records a successful launch would, so the
// callers' flows go on.  A violated check aborts the process.
#include <hip/hip_runtime_api.h>
#include <tilestream_kernels.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iterator>
#include <map>
#include <mutex>

namespace {

enum class Kind { device, pinned };
struct Allocation {
    std::size_t size;
    Kind kind;
};

std::mutex registry_lock;
std::map<std::uintptr_t, Allocation> &registry() {
    static std::map<std::uintptr_t, Allocation> r;
    return r;
}
thread_local hipError_t last_error = hipSuccess;

hipError_t fail(hipError_t e) {
    last_error = e;
    return e;
}

void check(bool ok, const char *what) {
    if (ok) return;
    std::fprintf(stderr, "host stub: kernel would read outside its tables: %s\n", what);
    std::abort();
}

// The allocation holding [p, p + bytes), if any.
bool inside_allocation(const void *p, std::uint64_t bytes) {
    const std::lock_guard<std::mutex> hold(registry_lock);
    const auto at = reinterpret_cast<std::uintptr_t>(p);
    auto it = registry().upper_bound(at);
    if (it == registry().begin()) return false;
    --it;
    return at >= it->first && at + bytes <= it->first + it->second.size;
}

void *allocate(std::size_t size, Kind kind) {
    // 256-byte alignment, as the runtime gives.
    void *p = std::aligned_alloc(256, (size + 255) / 256 * 256 + 256);
    if (p == nullptr) return nullptr;
    std::memset(p, 0, size);
    const std::lock_guard<std::mutex> hold(registry_lock);
    registry()[reinterpret_cast<std::uintptr_t>(p)] = {size, kind};
    return p;
}

hipError_t release(void *p) {
    if (p == nullptr) return hipSuccess;
    {
        const std::lock_guard<std::mutex> hold(registry_lock);
        if (registry().erase(reinterpret_cast<std::uintptr_t>(p)) == 0) return fail(hipErrorInvalidValue);
    }
    std::free(p);
    return hipSuccess;
}

struct Object {
    int tag;
};

std::uint32_t sample_of(const tsk_decode_image &img) { return img.calibrated ? 4u : (img.type == TSK_UINT16 ? 2u : 4u); }

void complete(const tsk_decode_image *images, std::uint32_t count, void *out) {
    auto *r = static_cast<std::uint32_t *>(out);
    for (std::uint32_t i = 0; i < count; ++i) {
        const std::uint64_t bytes = std::uint64_t{images[i].roi_w} * images[i].roi_h * sample_of(images[i]);
        r[4 * i] = 0;
        r[4 * i + 1] = 0;
        std::memcpy(&r[4 * i + 2], &bytes, 8);
    }
}

void check_tile_task(const tsk_decode_args &a, const tsk_decode_task &t, const tsk_decode_image &img) {
    const std::uint32_t mode = t.shape & 0xffu, h = (t.shape >> 8) & 0xffu, w = t.shape >> 16;
    check(w >= 1 && w <= img.tile_w && h >= 1 && h <= img.tile_h, "tile size");
    check(t.column < img.tile_columns, "tile column");
    if (t.payload != nullptr) {
        check(t.payload >= img.src && t.payload + std::uint64_t{t.words} * 4 <= img.src + img.src_size, "payload span");
        if (mode == 2 && t.rows != TSK_NO_ROWS) {
            check(std::uint64_t{t.rows} + h <= a.row_count, "row table");
            // Each row's descriptor (seed and width words) inside the payload.
            for (std::uint32_t r = 0; r < h; ++r) check(a.rows[t.rows + r] + 2u <= t.words, "row descriptor");
        }
    }
}

} // namespace

extern "C" {

hipError_t hipGetDevice(int *device) {
    *device = 0;
    return hipSuccess;
}

hipError_t hipDeviceGetAttribute(int *value, hipDeviceAttribute_t attr, int) {
    *value = attr == hipDeviceAttributeMultiprocessorCount ? 256 : 0;
    return hipSuccess;
}

hipError_t hipGetLastError(void) {
    const hipError_t e = last_error;
    last_error = hipSuccess;
    return e;
}

hipError_t hipPeekAtLastError(void) { return last_error; }

hipError_t hipMalloc(void **p, size_t size) {
    *p = allocate(size, Kind::device);
    return *p != nullptr ? hipSuccess : fail(hipErrorOutOfMemory);
}

hipError_t hipMallocAsync(void **p, size_t size, hipStream_t) { return hipMalloc(p, size); }
hipError_t hipFree(void *p) { return release(p); }
hipError_t hipFreeAsync(void *p, hipStream_t) { return release(p); }

hipError_t hipHostMalloc(void **p, size_t size, unsigned int) {
    *p = allocate(size, Kind::pinned);
    return *p != nullptr ? hipSuccess : fail(hipErrorOutOfMemory);
}

hipError_t hipHostFree(void *p) { return release(p); }

hipError_t hipHostGetDevicePointer(void **device, void *host, unsigned int) {
    *device = host;
    return hipSuccess;
}

hipError_t hipPointerGetAttributes(hipPointerAttribute_t *attr, const void *p) {
    const std::lock_guard<std::mutex> hold(registry_lock);
    const auto at = reinterpret_cast<std::uintptr_t>(p);
    auto it = registry().upper_bound(at);
    if (it != registry().begin()) {
        --it;
        if (at < it->first + it->second.size) {
            *attr = hipPointerAttribute_t{};
            attr->type = it->second.kind == Kind::device ? hipMemoryTypeDevice : hipMemoryTypeHost;
            attr->device = 0;
            attr->devicePointer = const_cast<void *>(p);
            attr->hostPointer = it->second.kind == Kind::pinned ? const_cast<void *>(p) : nullptr;
            return hipSuccess;
        }
    }
    return fail(hipErrorInvalidValue);
}

hipError_t hipMemcpyAsync(void *dst, const void *src, size_t bytes, hipMemcpyKind, hipStream_t) {
    if (bytes != 0) std::memcpy(dst, src, bytes);
    return hipSuccess;
}

hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned int) {
    *s = reinterpret_cast<hipStream_t>(new Object{1});
    return hipSuccess;
}

hipError_t hipStreamDestroy(hipStream_t s) {
    delete reinterpret_cast<Object *>(s);
    return hipSuccess;
}

hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }

hipError_t hipStreamGetDevice(hipStream_t, hipDevice_t *device) {
    *device = 0;
    return hipSuccess;
}

hipError_t hipEventCreate(hipEvent_t *e) {
    *e = reinterpret_cast<hipEvent_t>(new Object{2});
    return hipSuccess;
}

hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned) { return hipEventCreate(e); }

hipError_t hipEventDestroy(hipEvent_t e) {
    delete reinterpret_cast<Object *>(e);
    return hipSuccess;
}

hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipEventQuery(hipEvent_t) { return hipSuccess; }

hipError_t hipEventElapsedTime(float *ms, hipEvent_t, hipEvent_t) {
    *ms = 0.001F;
    return hipSuccess;
}

hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) { return hipSuccess; }

hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t *g) {
    *g = reinterpret_cast<hipGraph_t>(new Object{3});
    return hipSuccess;
}

hipError_t hipGraphDestroy(hipGraph_t g) {
    delete reinterpret_cast<Object *>(g);
    return hipSuccess;
}

hipError_t hipGraphInstantiate(hipGraphExec_t *x, hipGraph_t, hipGraphNode_t *, char *, size_t) {
    *x = reinterpret_cast<hipGraphExec_t>(new Object{4});
    return hipSuccess;
}

hipError_t hipGraphExecDestroy(hipGraphExec_t x) {
    delete reinterpret_cast<Object *>(x);
    return hipSuccess;
}

hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t) { return hipSuccess; }

// ---- kernel launchers --------------------------------------------------------

int tsk_preload(uint32_t) { return hipSuccess; }
int tsk_decode_supported(uint32_t group, uint32_t block, uint32_t depth) {
    return group == 0 ? (block == 512 && depth == 2) : (block >= 64 && block <= 512 && depth >= 1 && depth <= 3);
}
int tsk_decode_occupancy(uint32_t, uint32_t, uint32_t, uint32_t, int *blocks) {
    *blocks = 2;
    return hipSuccess;
}

int tsk_results_begin(void *results, uint32_t count, uint32_t *counter, uint32_t words, void *) {
    if (count != 0) check(inside_allocation(results, 16ull * count), "result records");
    if (counter != nullptr) std::memset(counter, 0, 4ull * words);
    if (count != 0) std::memset(results, 0, 16ull * count);
    return hipSuccess;
}

int tsk_results_complete(const tsk_decode_image *images, uint32_t count, void *results, void *) {
    complete(images, count, results);
    return hipSuccess;
}

int tsk_results_publish(const tsk_decode_image *images, uint32_t count, const void *, void *out, void *) {
    complete(images, count, out);
    return hipSuccess;
}

int tsk_decode(const tsk_decode_args *args, void *) {
    const tsk_decode_args &a = *args;
    check(a.grid != 0, "grid");
    if (a.group == 0) {
        check(a.blocks != nullptr && a.request_count == 0, "block tables");
        check(inside_allocation(a.blocks, a.task_count * sizeof(tsk_block)), "block table");
        for (std::uint64_t i = 0; i < a.task_count; ++i) {
            const tsk_block &b = a.blocks[i];
            check(b.image < a.image_count && b.count == std::uint32_t{b.across} * b.down && b.count <= 8, "block shape");
            check(b.blocks <= TSK_BLOCK_STAGE_USED, "block stage");
            std::uint32_t staged = 0;
            for (std::uint32_t t = 0; t < b.count; ++t) {
                const tsk_block_tile &bt = b.tiles[t];
                const auto p = reinterpret_cast<std::uintptr_t>(bt.payload);
                check(bt.stage == staged, "stage prefix");
                staged += static_cast<std::uint32_t>(((p + std::uint64_t{bt.words} * 4 + 15) >> 4) - (p >> 4));
                tsk_decode_task task{};
                task.payload = bt.payload;
                task.column = b.col0 + t % b.across;
                task.rows = bt.rows;
                task.words = bt.words;
                task.shape = bt.shape;
                check_tile_task(a, task, a.images[b.image]);
            }
            check(staged == b.blocks, "staged blocks");
        }
        return hipSuccess;
    }
    if (a.request_count != 0) {
        for (std::uint64_t i = 0; i < a.task_count; ++i) {
            const std::uint32_t r = a.request_index[i];
            check(r < a.request_count, "request index");
            const tsk_request &q = a.requests[r];
            check(i >= q.first_task && q.cols != 0, "request task");
        }
        return hipSuccess;
    }
    check(inside_allocation(a.tasks, a.task_count * sizeof(tsk_decode_task)), "task table");
    for (std::uint64_t i = 0; i < a.task_count; ++i) {
        const tsk_decode_task &t = a.tasks[i];
        check(t.image < a.image_count, "task image");
        check_tile_task(a, t, a.images[t.image]);
    }
    return hipSuccess;
}

// The block records of a request batch, built as the device kernel builds
// them, with every table index checked.
int tsk_build_blocks(const tsk_block_build_args *args, void *) {
    const tsk_block_build_args &a = *args;
    check(inside_allocation(a.blocks, a.block_count * sizeof(tsk_block)), "block records");
    check(inside_allocation(a.block_index, a.block_count * sizeof(std::uint32_t)), "block index");
    for (std::uint64_t b = 0; b < a.block_count; ++b) {
        const std::uint32_t r = a.block_index[b];
        check(inside_allocation(a.requests + r, sizeof(tsk_request)), "block request");
        const tsk_request &q = a.requests[r];
        check(q.across >= 1 && q.down >= 1 && std::uint32_t{q.across} * q.down <= 8 && b >= q.first_block, "block grid");
        const std::uint32_t local = static_cast<std::uint32_t>(b - q.first_block);
        const std::uint32_t nbx = (q.cols + q.across - 1) / q.across;
        const std::uint32_t by = local / nbx, bx = local - by * nbx;
        check(by * q.down < q.rows, "block row");
        const std::uint32_t across = std::min<std::uint32_t>(q.across, q.cols - bx * q.across);
        const std::uint32_t down = std::min<std::uint32_t>(q.down, q.rows - by * q.down);
        const tsk_decode_image &img = a.images[r];
        tsk_block rec{};
        rec.image = r;
        rec.across = static_cast<std::uint16_t>(across);
        rec.down = static_cast<std::uint16_t>(down);
        rec.col0 = q.col0 + bx * q.across;
        rec.row0 = q.row0 + by * q.down;
        rec.count = across * down;
        bool full = across == q.across && down == q.down;
        for (std::uint32_t k = 0; k < rec.count; ++k) {
            const std::uint64_t at = std::uint64_t{q.tile_base} + std::uint64_t{rec.row0 + k / across} * q.tile_columns +
                                     rec.col0 + k % across;
            check(inside_allocation(a.tiles + at, sizeof(tsk_decode_task)), "block tile");
            const tsk_decode_task &t = a.tiles[at];
            const auto p = reinterpret_cast<std::uintptr_t>(t.payload);
            tsk_block_tile &bt = rec.tiles[k];
            bt.payload = t.payload;
            bt.words = t.words;
            bt.shape = t.shape;
            bt.rows = t.rows;
            bt.stage = rec.blocks;
            rec.blocks += static_cast<std::uint32_t>(((p + std::uint64_t{t.words} * 4 + 15) >> 4) - (p >> 4));
            full = full && t.shape == TSK_TASK_SHAPE(2u, img.tile_h, img.tile_w) && t.rows != TSK_NO_ROWS;
        }
        rec.flags = full ? TSK_BLOCK_FULL : 0u;
        rec.image_copy = img;
        a.blocks[b] = rec;
    }
    return hipSuccess;
}

int tsk_stage_payloads(const tsk_decode_args *args, void *) {
    check(args->staged != nullptr && args->staging != nullptr, "staging tables");
    return hipSuccess;
}

int tsk_encode_reset(const tsk_encode_args *a, void *) {
    check(inside_allocation(a->state, 16ull * a->capacity), "look-back buffers");
    std::memset(a->state, 0, 16ull * a->capacity);
    return hipSuccess;
}

int tsk_encode_window(const tsk_encode_args *a, void *) {
    check(a->tile_count <= a->capacity, "window capacity");
    std::uint64_t covered = 0;
    for (std::uint32_t e = 0; e < a->entry_count; ++e) {
        const tsk_window_entry &x = a->entries[e];
        check(x.first == covered && x.count != 0, "window entries");
        covered += x.count;
        const tsk_encode_image &img = a->images[x.image];
        check(x.start + x.count <= img.tile_count, "entry tiles");
        if (x.flags & 2u) {
            auto *r = static_cast<std::uint32_t *>(a->results) + 4 * x.image;
            r[0] = 0;
            r[1] = 0;
            const std::uint64_t bytes = 64 + 32 * img.tile_count;
            std::memcpy(&r[2], &bytes, 8);
        }
    }
    check(covered == a->tile_count, "window tiles");
    if (a->blocks != nullptr)
        for (std::uint32_t b = 0; b * 64u < a->tile_count; ++b) check(a->blocks[b] < a->entry_count, "block owners");
    return hipSuccess;
}

int tsk_index_header(const tsk_index_args *a, void *) {
    check(inside_allocation(a->record, sizeof(tsk_index_record)), "index record");
    check(inside_allocation(a->src, a->capacity), "index source");
    std::uint64_t size = a->capacity;
    if (a->producer != nullptr) std::memcpy(&size, static_cast<const std::uint8_t *>(a->producer) + 8, 8);
    *a->record = tsk_index_record{0u, 0u, size};
    return hipSuccess;
}

int tsk_index_tiles(const tsk_index_args *a, void *) {
    check(inside_allocation(a->tasks, a->tile_count * sizeof(tsk_decode_task)), "index tasks");
    check(inside_allocation(a->rows, a->tile_count * a->tile_h * sizeof(std::uint16_t)), "index rows");
    std::memset(a->tasks, 0, a->tile_count * sizeof(tsk_decode_task));
    return hipSuccess;
}

int tsk_scratch_region(void **address, uint64_t *bytes) {
    static void *scratch = allocate(TSK_SCRATCH_BYTES, Kind::device);
    *address = scratch;
    *bytes = TSK_SCRATCH_BYTES;
    return hipSuccess;
}

int tsk_calibrate(const tsk_calibrate_args *a, void *) {
    std::uint64_t rows = 0;
    for (std::uint32_t i = 0; i < a->image_count; ++i) {
        check(a->first_rows[i] == rows, "calibration rows");
        rows += a->images[i].height;
        auto *r = static_cast<std::uint32_t *>(a->results) + 4 * i;
        r[0] = 0;
        const std::uint64_t bytes = std::uint64_t{a->images[i].width} * a->images[i].height * 4;
        std::memcpy(&r[2], &bytes, 8);
    }
    check(rows == a->total_rows, "total rows");
    return hipSuccess;
}

int tsk_copy_blocks(void *dst, const void *src, uint64_t blocks, void *) {
    check(inside_allocation(dst, blocks * 16) && inside_allocation(src, blocks * 16), "copy span");
    std::memcpy(dst, src, blocks * 16);
    return hipSuccess;
}

} // extern "C"
