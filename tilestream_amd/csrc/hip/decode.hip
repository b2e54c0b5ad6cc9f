// Decode launchers, result-record kernels and small utility kernels.  The
// decode kernels themselves are instantiated per kind in decode_<kind>.hip
// from the templates of decode_core.h.
#include "decode_core.h"

using tsk::DecodeFn;
using tsk::ResultBits;

// The kernel lookups of each kind, one per lane-group width 1, 2, 4, 8, 16.
#define TSK_DECLARE_TABLE(KIND_NAME, G) DecodeFn TSK_TABLE_NAME(KIND_NAME, G)(std::uint32_t block, std::uint32_t depth);
#define TSK_DECLARE_TABLES(KIND_NAME)                                                                  \
    TSK_DECLARE_TABLE(KIND_NAME, 1)                                                                   \
    TSK_DECLARE_TABLE(KIND_NAME, 2)                                                                   \
    TSK_DECLARE_TABLE(KIND_NAME, 4) TSK_DECLARE_TABLE(KIND_NAME, 8) TSK_DECLARE_TABLE(KIND_NAME, 16)
#define TSK_TABLE_ROW(KIND_NAME)                                                                       \
    {TSK_TABLE_NAME(KIND_NAME, 1), TSK_TABLE_NAME(KIND_NAME, 2), TSK_TABLE_NAME(KIND_NAME, 4),            \
     TSK_TABLE_NAME(KIND_NAME, 8), TSK_TABLE_NAME(KIND_NAME, 16)}

extern "C" {
TSK_DECLARE_TABLES(u16)
TSK_DECLARE_TABLES(f32)
TSK_DECLARE_TABLES(u16_cal)
TSK_DECLARE_TABLES(f32_cal)
TSK_DECLARE_TABLES(mixed)
DecodeFn tsk_block_table(std::uint32_t kind, std::uint32_t block);
}

namespace {

// Copies the 16-byte blocks that hold each staged task's payload from the
// resident address (mapped host memory, read over the host link) into the
// task's place in device staging.  Workgroups stride over the tasks.
__global__ void __launch_bounds__(256) stage_kernel(tsk_decode_args a) {
    using tsk::gptr;
    using tsk::u32x4;
    for (std::uint64_t i = blockIdx.x; i < a.task_count; i += gridDim.x) {
        const std::uint32_t block = tsk::load_constant(a.staged, i);
        if (block == TSK_NOT_STAGED) continue;
        const std::uint32_t r = tsk::load_constant(a.request_index, i);
        const tsk_request q = tsk::load_constant(a.requests, r);
        const std::uint32_t local = static_cast<std::uint32_t>(i - q.first_task);
        const std::uint32_t dr = local / q.cols;
        const std::uint32_t col = q.col0 + (local - dr * q.cols);
        const std::uint32_t row = q.row0 + dr;
        const tsk_decode_task t =
            tsk::load_constant(a.tasks, std::uint64_t{q.tile_base} + std::uint64_t{row} * q.tile_columns + col);
        const std::uintptr_t first = reinterpret_cast<std::uintptr_t>(t.payload) & ~std::uintptr_t{15};
        const std::uintptr_t end = (reinterpret_cast<std::uintptr_t>(t.payload) + std::uint64_t{t.words} * 4 + 15) &
                                   ~std::uintptr_t{15};
        const std::uint32_t blocks = static_cast<std::uint32_t>((end - first) / 16);
        const gptr<const u32x4> src = tsk::global(reinterpret_cast<const u32x4 *>(first));
        const gptr<u32x4> dst = tsk::global(reinterpret_cast<u32x4 *>(const_cast<std::uint8_t *>(a.staging)) +
                                            std::uint64_t{block});
        for (std::uint32_t k = threadIdx.x; k < blocks; k += blockDim.x) dst[k] = __builtin_nontemporal_load(src + k);
    }
}

// Zeroes the result records and, when given, the work counter words.
__global__ void results_begin_kernel(ResultBits *results, std::uint32_t count, std::uint32_t *counter,
                                     std::uint32_t words) {
    const std::uint64_t stride = std::uint64_t{gridDim.x} * blockDim.x;
    for (std::uint64_t i = std::uint64_t{blockIdx.x} * blockDim.x + threadIdx.x; i < count; i += stride)
        results[i] = {0u, 0u, 0u};
    if (counter != nullptr && blockIdx.x == 0 && threadIdx.x < words) counter[threadIdx.x] = 0;
}

__global__ void results_complete_kernel(const tsk_decode_image *images, std::uint32_t count, const ResultBits *results,
                                        ResultBits *out) {
    const std::uint64_t stride = std::uint64_t{gridDim.x} * blockDim.x;
    for (std::uint64_t i = std::uint64_t{blockIdx.x} * blockDim.x + threadIdx.x; i < count; i += stride) {
        const tsk_decode_image &img = images[i];
        std::uint32_t status = results[i].status;
        if (img.producer != nullptr) {
            const std::uint32_t produced = static_cast<const ResultBits *>(img.producer)->status;
            if (produced != 0) status = produced;
        }
        const std::uint64_t s = img.calibrated ? 4 : (img.type == TSK_UINT16 ? 2 : 4);
        out[i] = {status, 0u, status == 0 ? std::uint64_t{img.roi_w} * img.roi_h * s : 0u};
    }
}

// Streaming copy used as the known-size reference when HBM counters are scaled.
__global__ void copy_kernel(uint4 *__restrict__ dst, const uint4 *__restrict__ src, std::uint64_t blocks) {
    const std::uint64_t stride = std::uint64_t{gridDim.x} * blockDim.x;
    for (std::uint64_t i = std::uint64_t{blockIdx.x} * blockDim.x + threadIdx.x; i < blocks; i += stride)
        dst[i] = src[i];
}

// The fastest plain copy shape measured on this hardware (four 16-byte blocks
// in flight per thread, non-temporal, 32768 workgroups): the practical
// ceiling that a kernel reading and writing about as much is compared with.
__global__ void __launch_bounds__(256) stream_copy_kernel(tsk::u32x4 *__restrict__ dst,
                                                          const tsk::u32x4 *__restrict__ src, std::uint64_t blocks) {
    constexpr int kUnroll = 4;
    const std::uint64_t stride = std::uint64_t{gridDim.x} * blockDim.x * kUnroll;
    for (std::uint64_t base = std::uint64_t{blockIdx.x} * blockDim.x * kUnroll + threadIdx.x; base < blocks;
         base += stride) {
        tsk::u32x4 v[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            const std::uint64_t i = base + std::uint64_t{static_cast<unsigned>(u)} * blockDim.x;
            if (i < blocks) v[u] = __builtin_nontemporal_load(src + i);
        }
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            const std::uint64_t i = base + std::uint64_t{static_cast<unsigned>(u)} * blockDim.x;
            if (i < blocks) __builtin_nontemporal_store(v[u], dst + i);
        }
    }
}

// Read-only and write-only streams in their fastest measured shapes (reads:
// four non-temporal 16-byte loads in flight per thread; writes: one
// non-temporal 16-byte store per thread and step; 32768 workgroups each):
// the two one-sided ceilings that bound a kernel's mix of reads and writes.
__global__ void __launch_bounds__(256) stream_read_kernel(const tsk::u32x4 *__restrict__ src, std::uint64_t blocks,
                                                          std::uint32_t *sink) {
    constexpr int kUnroll = 4;
    const std::uint64_t stride = std::uint64_t{gridDim.x} * blockDim.x * kUnroll;
    std::uint32_t acc = 0;
    for (std::uint64_t base = std::uint64_t{blockIdx.x} * blockDim.x * kUnroll + threadIdx.x; base < blocks;
         base += stride) {
        tsk::u32x4 v[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            const std::uint64_t i = base + std::uint64_t{static_cast<unsigned>(u)} * blockDim.x;
            v[u] = i < blocks ? __builtin_nontemporal_load(src + i) : tsk::u32x4{0, 0, 0, 0};
        }
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) acc ^= v[u].x ^ v[u].y ^ v[u].z ^ v[u].w;
    }
    // One word that is never all-ones in practice keeps the loads live.
    if (acc == 0xffffffffu) sink[0] = acc;
}

__global__ void __launch_bounds__(256) stream_write_kernel(tsk::u32x4 *__restrict__ dst, std::uint64_t blocks,
                                                           std::uint32_t value) {
    const std::uint64_t stride = std::uint64_t{gridDim.x} * blockDim.x;
    for (std::uint64_t i = std::uint64_t{blockIdx.x} * blockDim.x + threadIdx.x; i < blocks; i += stride)
        __builtin_nontemporal_store(tsk::u32x4{value, value, value, value}, dst + i);
}

__global__ void fill_kernel(std::uint32_t *data, std::uint64_t words, std::uint32_t value) {
    const std::uint64_t stride = std::uint64_t{gridDim.x} * blockDim.x;
    for (std::uint64_t i = std::uint64_t{blockIdx.x} * blockDim.x + threadIdx.x; i < words; i += stride)
        data[i] = value;
}

DecodeFn decode_function(std::uint32_t kind, std::uint32_t group, std::uint32_t block, std::uint32_t depth) {
    if (group == 0) return depth == 2 ? tsk_block_table(kind, block) : nullptr;
    using Table = DecodeFn (*)(std::uint32_t, std::uint32_t);
    // Rows in TSK_KIND_* order.
    static constexpr Table tables[5][5] = {TSK_TABLE_ROW(u16), TSK_TABLE_ROW(f32), TSK_TABLE_ROW(u16_cal),
                                           TSK_TABLE_ROW(f32_cal), TSK_TABLE_ROW(mixed)};
    static_assert(TSK_KIND_U16 == 0 && TSK_KIND_F32 == 1 && TSK_KIND_U16_CAL == 2 && TSK_KIND_F32_CAL == 3 &&
                  TSK_KIND_MIXED == 4);
    if (!tsk::compiled(group, block, depth) || kind > TSK_KIND_MIXED) return nullptr;
    const int width = group == 1 ? 0 : group == 2 ? 1 : group == 4 ? 2 : group == 8 ? 3 : 4;
    return tables[kind][width](block, depth);
}

std::uint32_t small_grid(std::uint64_t n) {
    const std::uint64_t blocks = n / 256 + 1;
    return static_cast<std::uint32_t>(blocks < 1024 ? blocks : 1024);
}

} // namespace

extern "C" {

int tsk_results_begin(void *results, std::uint32_t count, std::uint32_t *counter, std::uint32_t words,
                      void *stream) {
    if (count == 0 && counter == nullptr) return hipSuccess;
    if (words > 256) return hipErrorInvalidValue;
    hipLaunchKernelGGL(results_begin_kernel, dim3(small_grid(count)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), static_cast<ResultBits *>(results), count, counter, words);
    return hipGetLastError();
}

int tsk_results_complete(const tsk_decode_image *images, std::uint32_t count, void *results, void *stream) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(results_complete_kernel, dim3(small_grid(count)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), images, count, static_cast<const ResultBits *>(results),
                       static_cast<ResultBits *>(results));
    return hipGetLastError();
}

int tsk_stage_payloads(const tsk_decode_args *a, void *stream) {
    if (a->task_count == 0 || a->staged == nullptr) return hipSuccess;
    const std::uint64_t grid = a->task_count < 8192 ? a->task_count : 8192;
    hipLaunchKernelGGL(stage_kernel, dim3(static_cast<std::uint32_t>(grid)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), *a);
    return hipGetLastError();
}

int tsk_results_publish(const tsk_decode_image *images, std::uint32_t count, const void *results, void *out,
                        void *stream) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(results_complete_kernel, dim3(small_grid(count)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), images, count, static_cast<const ResultBits *>(results),
                       static_cast<ResultBits *>(out));
    return hipGetLastError();
}

int tsk_decode_supported(std::uint32_t group, std::uint32_t block, std::uint32_t depth) {
    if (group == 0) return depth == 2 && block == 512 ? 1 : 0;
    return tsk::compiled(group, block, depth) ? 1 : 0;
}

int tsk_decode(const tsk_decode_args *args, void *stream) {
    if (args->task_count == 0) return hipSuccess;
    const DecodeFn fn = decode_function(args->kind, args->group, args->block, args->depth);
    if (fn == nullptr || args->grid == 0 || (args->queue != nullptr && args->chunk == 0)) return hipErrorInvalidValue;
    if (args->group == 0 && (args->blocks == nullptr || args->request_count != 0))
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(fn, dim3(args->grid), dim3(args->block), 0, static_cast<hipStream_t>(stream), *args);
    return hipGetLastError();
}

int tsk_copy_blocks(void *dst, const void *src, std::uint64_t blocks, void *stream) {
    if (blocks == 0) return hipSuccess;
    const std::uint64_t groups = (blocks + 255) / 256;
    hipLaunchKernelGGL(copy_kernel, dim3(groups < 4096 ? static_cast<unsigned>(groups) : 4096u), dim3(256), 0,
                       static_cast<hipStream_t>(stream),
                       static_cast<uint4 *>(dst), static_cast<const uint4 *>(src), blocks);
    return hipGetLastError();
}

int tsk_stream_copy(void *dst, const void *src, std::uint64_t blocks, void *stream) {
    if (blocks == 0) return hipSuccess;
    const std::uint64_t groups = (blocks + 1023) / 1024;
    hipLaunchKernelGGL(stream_copy_kernel, dim3(groups < 32768 ? static_cast<unsigned>(groups) : 32768u), dim3(256),
                       0, static_cast<hipStream_t>(stream), static_cast<tsk::u32x4 *>(dst),
                       static_cast<const tsk::u32x4 *>(src), blocks);
    return hipGetLastError();
}

int tsk_stream_read(const void *src, std::uint64_t blocks, void *sink, void *stream) {
    if (blocks == 0) return hipSuccess;
    const std::uint64_t groups = (blocks + 1023) / 1024;
    hipLaunchKernelGGL(stream_read_kernel, dim3(groups < 32768 ? static_cast<unsigned>(groups) : 32768u), dim3(256),
                       0, static_cast<hipStream_t>(stream), static_cast<const tsk::u32x4 *>(src), blocks,
                       static_cast<std::uint32_t *>(sink));
    return hipGetLastError();
}

int tsk_stream_write(void *dst, std::uint64_t blocks, std::uint32_t value, void *stream) {
    if (blocks == 0) return hipSuccess;
    const std::uint64_t groups = (blocks + 255) / 256;
    hipLaunchKernelGGL(stream_write_kernel, dim3(groups < 32768 ? static_cast<unsigned>(groups) : 32768u), dim3(256),
                       0, static_cast<hipStream_t>(stream), static_cast<tsk::u32x4 *>(dst), blocks, value);
    return hipGetLastError();
}

int tsk_decode_occupancy(std::uint32_t kind, std::uint32_t group, std::uint32_t block, std::uint32_t depth,
                         int *blocks) {
    const DecodeFn fn = decode_function(kind, group, block, depth);
    if (fn == nullptr) return hipErrorInvalidValue;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, reinterpret_cast<const void *>(fn),
                                                        static_cast<int>(block), 0);
}

int tsk_fill_u32(std::uint32_t *data, std::uint64_t words, std::uint32_t value, void *stream) {
    if (words == 0) return hipSuccess;
    hipLaunchKernelGGL(fill_kernel, dim3(small_grid(words)), dim3(256), 0, static_cast<hipStream_t>(stream), data,
                       words, value);
    return hipGetLastError();
}

// Loads every decode kernel the kinds of an operation can launch.
int tsk_preload_decode(std::uint32_t calibrated) {
    hipFuncAttributes attr;
    const std::uint32_t kinds[3] = {calibrated ? std::uint32_t{TSK_KIND_U16_CAL} : std::uint32_t{TSK_KIND_U16},
                                    calibrated ? std::uint32_t{TSK_KIND_F32_CAL} : std::uint32_t{TSK_KIND_F32},
                                    TSK_KIND_MIXED};
    for (std::uint32_t kind : kinds)
        for (std::uint32_t g : {0u, 1u, 2u, 4u, 8u, 16u})
            for (std::uint32_t b : {64u, 128u, 256u, 512u})
                for (std::uint32_t d : {1u, 2u, 3u}) {
                    const DecodeFn fn = decode_function(kind, g, b, d);
                    if (fn == nullptr) continue;
                    const hipError_t e = hipFuncGetAttributes(&attr, reinterpret_cast<const void *>(fn));
                    if (e != hipSuccess) return e;
                }
    hipError_t e = hipFuncGetAttributes(&attr, reinterpret_cast<const void *>(results_begin_kernel));
    if (e == hipSuccess) e = hipFuncGetAttributes(&attr, reinterpret_cast<const void *>(results_complete_kernel));
    if (e == hipSuccess) e = hipFuncGetAttributes(&attr, reinterpret_cast<const void *>(fill_kernel));
    return e;
}

} // extern "C"
