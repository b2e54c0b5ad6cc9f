// Tile decoder for gfx950 (CDNA4, wave64).
//
// Behaviour follows the reference decoder (reference src/cuda/decoder.cu:176-305):
// every tile intersecting the region is reconstructed from column 0, only
// region pixels are written, the trusted path validates the device-resident
// index and row descriptors first.  The execution scheme is this library's:
//
//  * A workgroup decodes one tile at a time and loops over tiles, either with a
//    static stride or by drawing chunks from a device work counter, so one
//    launch can drain a whole mixed batch.
//  * The tile payload is staged into LDS with 16-byte loads from the 16-byte
//    block that contains its first word (the block never crosses a page).
//  * A row of tile width W is split into segments of S = min(64, W / G) lanes;
//    each lane decodes G consecutive columns sequentially in registers (a
//    funnel shift per residual), the S lane totals are scanned with DPP
//    row_shr 1/2/4/8 and row_bcast 15/31, and rows wider than 64 G columns
//    carry the running value across passes.
//  * uint16 reconstruction is a 32-bit add of inverse-zigzag terms truncated to
//    16 bits at the store (exact modulo 2^16); float32 is a running XOR of bit
//    words.  Groups fully inside the region are written with 16-byte stores when
//    aligned.
#include <hip/hip_runtime.h>

#include <tilestream_kernels.h>

#include <cstdint>

namespace {

// Largest payload (64 x 64 float32 RAW) in words, plus up to 3 words of 16-byte
// misalignment ahead of it and the farthest read past a row end: a lane of the
// last segment reads (G S - 1) b / 32 + 1 <= 256 words beyond the row start.
constexpr std::uint32_t kPayloadWords = 4096;
constexpr std::uint32_t kStageWords = kPayloadWords + 8 + 280;

struct DeviceResultBits {
    std::uint32_t status;
    std::uint32_t reserved;
    std::uint64_t bytes;
};

constexpr std::uint32_t kMalformed = 5;

__device__ inline void fail(void *results, std::uint32_t image, std::uint32_t status) {
    auto *r = static_cast<DeviceResultBits *>(results) + image;
    atomicCAS(&r->status, 0u, status);
}

template <bool U16> struct Ring {
    // uint16: additive group of integers modulo 2^32 (truncated to 16 bits at the
    // end); float32: XOR group of bit words.
    static __device__ inline std::uint32_t add(std::uint32_t a, std::uint32_t b) { return U16 ? a + b : a ^ b; }
    static __device__ inline std::uint32_t sub(std::uint32_t a, std::uint32_t b) { return U16 ? a - b : a ^ b; }
    // Inverse zigzag as an unsigned delta: even z -> z / 2, odd z -> ~(z / 2),
    // which is 65535 - z / 2 modulo 2^16.
    static __device__ inline std::uint32_t term(std::uint32_t z) { return U16 ? (z >> 1) ^ (0u - (z & 1u)) : z; }
};

template <bool U16, std::uint32_t Ctrl, std::uint32_t RowMask>
__device__ inline std::uint32_t dpp_combine(std::uint32_t v) {
    // Lanes whose source is outside the DPP row or masked out read 0, the identity.
    const std::uint32_t t = static_cast<std::uint32_t>(
        __builtin_amdgcn_update_dpp(0, static_cast<int>(v), Ctrl, RowMask, 0xf, true));
    return Ring<U16>::add(v, t);
}

template <bool U16, std::uint32_t Ctrl>
__device__ inline std::uint32_t dpp_combine_if(std::uint32_t v, bool take) {
    const std::uint32_t t = static_cast<std::uint32_t>(
        __builtin_amdgcn_update_dpp(0, static_cast<int>(v), Ctrl, 0xf, 0xf, true));
    return take ? Ring<U16>::add(v, t) : v;
}

// Inclusive scan of v over segments of S consecutive lanes (S a power of two,
// 4 <= S <= 64, uniform).  p is the lane's position in its segment.
template <bool U16> __device__ inline std::uint32_t segment_scan(std::uint32_t v, std::uint32_t p, std::uint32_t S) {
    if (S >= 16) {
        v = dpp_combine<U16, 0x111, 0xf>(v); // row_shr:1
        v = dpp_combine<U16, 0x112, 0xf>(v); // row_shr:2
        v = dpp_combine<U16, 0x114, 0xf>(v); // row_shr:4
        v = dpp_combine<U16, 0x118, 0xf>(v); // row_shr:8
        if (S >= 32) v = dpp_combine<U16, 0x142, 0xa>(v); // row_bcast:15 into rows 1 and 3
        if (S == 64) v = dpp_combine<U16, 0x143, 0xc>(v); // row_bcast:31 into rows 2 and 3
    } else {
        // Segments narrower than a DPP row: keep lanes from pulling across.
        v = dpp_combine_if<U16, 0x111>(v, p >= 1);
        v = dpp_combine_if<U16, 0x112>(v, p >= 2);
        if (S == 8) v = dpp_combine_if<U16, 0x114>(v, p >= 4);
    }
    return v;
}

template <int G> struct Group {
    std::uint32_t v[G];
};

struct Region {
    std::uint8_t *line;    // output row of this tile row
    std::int64_t ex_tile;  // output column of tile column 0 (may be negative)
    std::uint64_t ey;      // output row index
    std::uint32_t left, right;
};

__device__ inline float sample_float(std::uint32_t bits, bool u16) {
    return u16 ? static_cast<float>(bits & 0xffffu) : __uint_as_float(bits);
}

// Stores G samples (or their calibrated values) for tile columns col0..col0+G-1.
template <int G, bool U16, bool CAL>
__device__ inline void emit(const tsk_decode_image &img, const Region &r, std::uint32_t col0, const Group<G> &g) {
    const bool inside = col0 >= r.left && col0 + G <= r.right;
    const std::int64_t ex0 = r.ex_tile + col0;
    if constexpr (CAL) {
        float out[G];
        const std::uint8_t *orow = img.offset_map ? img.offset_map + r.ey * img.offset_pitch : nullptr;
        const std::uint8_t *grow = img.gain_map ? img.gain_map + r.ey * img.gain_pitch : nullptr;
#pragma unroll
        for (int k = 0; k < G; ++k) {
            const std::uint32_t c = col0 + k;
            float o = img.offset, gn = img.gain;
            if (c >= r.left && c < r.right) {
                if (orow) o = reinterpret_cast<const float *>(orow)[ex0 + k];
                if (grow) gn = reinterpret_cast<const float *>(grow)[ex0 + k];
            }
            // Two roundings to nearest even, never contracted into an FMA.
            out[k] = __fmul_rn(__fsub_rn(sample_float(g.v[k], U16), o), gn);
        }
        float *dst = reinterpret_cast<float *>(r.line) + ex0;
        if (inside && G >= 4 && (reinterpret_cast<std::uintptr_t>(dst) & 15u) == 0) {
#pragma unroll
            for (int k = 0; k + 4 <= G; k += 4)
                *reinterpret_cast<float4 *>(dst + k) = make_float4(out[k], out[k + 1], out[k + 2], out[k + 3]);
        } else {
#pragma unroll
            for (int k = 0; k < G; ++k)
                if (col0 + k >= r.left && col0 + k < r.right) dst[k] = out[k];
        }
    } else if constexpr (U16) {
        std::uint16_t *dst = reinterpret_cast<std::uint16_t *>(r.line) + ex0;
        constexpr int kBytes = 2 * G;
        if (G >= 2 && inside && (reinterpret_cast<std::uintptr_t>(dst) & ((kBytes >= 16 ? 16 : kBytes) - 1)) == 0) {
            std::uint32_t w[G / 2 > 0 ? G / 2 : 1];
#pragma unroll
            for (int i = 0; i < G / 2; ++i) w[i] = (g.v[2 * i] & 0xffffu) | (g.v[2 * i + 1] << 16);
            if constexpr (G >= 8) {
#pragma unroll
                for (int i = 0; i + 4 <= G / 2; i += 4)
                    *reinterpret_cast<uint4 *>(dst + 2 * i) = make_uint4(w[i], w[i + 1], w[i + 2], w[i + 3]);
            } else if constexpr (G == 4) {
                *reinterpret_cast<uint2 *>(dst) = make_uint2(w[0], w[1]);
            } else if constexpr (G == 2) {
                *reinterpret_cast<std::uint32_t *>(dst) = w[0];
            }
        } else {
#pragma unroll
            for (int k = 0; k < G; ++k)
                if (col0 + k >= r.left && col0 + k < r.right) dst[k] = static_cast<std::uint16_t>(g.v[k]);
        }
    } else {
        std::uint32_t *dst = reinterpret_cast<std::uint32_t *>(r.line) + ex0;
        constexpr int kBytes = 4 * G;
        if (inside && (reinterpret_cast<std::uintptr_t>(dst) & ((kBytes >= 16 ? 16 : kBytes) - 1)) == 0) {
            if constexpr (G >= 4) {
#pragma unroll
                for (int k = 0; k + 4 <= G; k += 4)
                    *reinterpret_cast<uint4 *>(dst + k) = make_uint4(g.v[k], g.v[k + 1], g.v[k + 2], g.v[k + 3]);
            } else if constexpr (G == 2) {
                *reinterpret_cast<uint2 *>(dst) = make_uint2(g.v[0], g.v[1]);
            } else {
                dst[0] = g.v[0];
            }
        } else {
#pragma unroll
            for (int k = 0; k < G; ++k)
                if (col0 + k >= r.left && col0 + k < r.right) dst[k] = g.v[k];
        }
    }
}

struct TileView {
    std::uint64_t tx, ty;
    std::uint32_t w, h;
    std::uint32_t top, bottom, left, right;
    std::uint32_t mode;
    std::uint32_t base;  // LDS word of payload word 0
};

// Decodes the staged tile and writes its region pixels.
template <int G, bool U16, bool CAL>
__device__ void decode_tile(const tsk_decode_image &img, const TileView &t, const std::uint32_t *stage,
                            const std::uint16_t *row_start) {
    const std::uint32_t S = (img.tile_w / G) < 64u ? img.tile_w / G : 64u;
    const std::uint32_t passes = img.tile_w / (S * G);
    const std::uint32_t p = threadIdx.x % S;
    const std::uint32_t rows_per_step = blockDim.x / S;
    const std::uint32_t *payload = stage + t.base;

    for (std::uint32_t row = t.top + threadIdx.x / S; row < t.bottom; row += rows_per_step) {
        Region r;
        r.ey = t.ty + row - img.roi_y;
        r.line = img.out + r.ey * img.out_pitch;
        r.ex_tile = static_cast<std::int64_t>(t.tx) - static_cast<std::int64_t>(img.roi_x);
        r.left = t.left;
        r.right = t.right;
        if (t.mode == 2) {
            const std::uint32_t start = row_start[row];
            const std::uint32_t seed = payload[start];
            const std::uint32_t b = payload[start + 1];
            const std::uint32_t mask = b >= 32 ? 0xffffffffu : ((1u << b) - 1u);
            const std::uint32_t *words = payload + start + 2;
            std::uint32_t carry = seed;
            for (std::uint32_t pass = 0; pass < passes; ++pass) {
                const std::uint32_t col0 = (pass * S + p) * G;
                // Residual j of this lane is residual col0 + j of the row: it moves
                // from column col0 + j to col0 + j + 1.
                std::uint32_t bit = col0 * b;
                std::uint32_t acc = 0;
                Group<G> pre;
#pragma unroll
                for (int j = 0; j < G; ++j) {
                    const std::uint32_t w = bit >> 5;
                    const std::uint32_t z = __builtin_amdgcn_alignbit(words[w + 1], words[w], bit) & mask;
                    acc = Ring<U16>::add(acc, Ring<U16>::term(z));
                    pre.v[j] = acc;
                    bit += b;
                }
                const std::uint32_t incl = segment_scan<U16>(acc, p, S);
                const std::uint32_t first = Ring<U16>::add(carry, Ring<U16>::sub(incl, acc));
                if (passes > 1)
                    carry = Ring<U16>::add(carry, static_cast<std::uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(incl), 63)));
                Group<G> v;
                v.v[0] = first;
#pragma unroll
                for (int k = 1; k < G; ++k) v.v[k] = Ring<U16>::add(first, pre.v[k - 1]);
                emit<G, U16, CAL>(img, r, col0, v);
            }
        } else {
            for (std::uint32_t pass = 0; pass < passes; ++pass) {
                const std::uint32_t col0 = (pass * S + p) * G;
                Group<G> v;
                if (t.mode == 0) {
                    const std::uint32_t idx0 = row * t.w + col0;
#pragma unroll
                    for (int k = 0; k < G; ++k) {
                        const std::uint32_t idx = idx0 + k;
                        v.v[k] = U16 ? (payload[idx >> 1] >> ((idx & 1u) * 16)) : payload[idx];
                    }
                } else {
#pragma unroll
                    for (int k = 0; k < G; ++k) v.v[k] = payload[0];
                }
                emit<G, U16, CAL>(img, r, col0, v);
            }
        }
    }
}

template <int G, int KIND>
__global__ void __launch_bounds__(512) decode_kernel(tsk_decode_args a) {
    __shared__ __attribute__((aligned(16))) std::uint32_t stage[kStageWords];
    __shared__ std::uint16_t row_start[64];
    __shared__ std::uint32_t shared_flag;
    __shared__ std::uint64_t shared_next;

    const std::uint32_t tid = threadIdx.x;
    const std::uint32_t chunk = a.queue ? a.chunk : 1u;
    std::uint64_t next = a.queue ? 0 : blockIdx.x;
    std::uint64_t end = 0;

    for (;;) {
        if (a.queue) {
            if (next >= end) {
                __syncthreads();
                if (tid == 0) shared_next = atomicAdd(a.queue, chunk);
                __syncthreads();
                next = shared_next;
                end = next + chunk;
            }
            if (next >= a.task_count) break;
        } else if (next >= a.task_count) {
            break;
        }
        const std::uint64_t i = next;
        next += a.queue ? 1 : gridDim.x;

        const tsk_decode_task task = a.tasks[i];
        const tsk_decode_image &img = a.images[task.image];
        const bool u16 = img.type == TSK_UINT16;
        const std::uint64_t col = task.tile_index % img.tile_columns;
        const std::uint64_t trow = task.tile_index / img.tile_columns;
        TileView t;
        t.tx = col * img.tile_w;
        t.ty = trow * img.tile_h;
        t.w = static_cast<std::uint32_t>(img.width - t.tx < img.tile_w ? img.width - t.tx : img.tile_w);
        t.h = static_cast<std::uint32_t>(img.height - t.ty < img.tile_h ? img.height - t.ty : img.tile_h);
        t.mode = task.mode;
        std::uint64_t offset = task.offset;
        std::uint32_t words = task.words;

        if (img.producer != nullptr) {
            // Trusted encoder output: read the result and the index entry only now,
            // after the producer completed.  All threads see the same values.
            const auto *prod = static_cast<const DeviceResultBits *>(img.producer);
            if (prod->status != 0) continue;
            const std::uint8_t *slot = img.src + 64 + task.tile_index * 32;
            const std::uint32_t *s32 = reinterpret_cast<const std::uint32_t *>(slot);
            const std::uint32_t mode = s32[0] & 0xffu;
            const std::uint32_t dims = s32[1];
            const std::uint64_t off = s32[2] | (std::uint64_t{s32[3]} << 32);
            const std::uint64_t len = s32[4] | (std::uint64_t{s32[5]} << 32);
            const std::uint64_t end_bytes = prod->bytes < img.src_size ? prod->bytes : img.src_size;
            const std::uint64_t raw4 = (std::uint64_t{t.w} * t.h * (u16 ? 2 : 4) + 3) / 4 * 4;
            if (mode > 2 || (dims & 0xffffu) != t.w || (dims >> 16) != t.h || off % 4 != 0 || len % 4 != 0 ||
                len == 0 || len > raw4 || off > end_bytes || len > end_bytes - off) {
                if (tid == 0) fail(a.results, task.image, kMalformed);
                continue;
            }
            t.mode = mode;
            offset = off;
            words = static_cast<std::uint32_t>(len / 4);
        }

        // Stage the payload: 16-byte blocks starting at the block holding word 0.
        const std::uintptr_t addr = reinterpret_cast<std::uintptr_t>(img.src + offset);
        const std::uint32_t mis = static_cast<std::uint32_t>((addr >> 2) & 3u);
        const uint4 *blocks = reinterpret_cast<const uint4 *>(addr - 4u * mis);
        const std::uint32_t nblocks = (words + mis + 3) >> 2;
        __syncthreads(); // the previous tile is done with the LDS
        for (std::uint32_t k = tid; k < nblocks; k += blockDim.x)
            *reinterpret_cast<uint4 *>(stage + 4 * k) = blocks[k];
        if (t.mode == 2 && img.producer == nullptr && tid < t.h) row_start[tid] = a.rows[task.rows + tid];
        __syncthreads();
        if (t.mode == 2 && img.producer != nullptr) {
            if (tid == 0) {
                // Serial walk of the row descriptors (reference decoder.cu:86-110).
                const std::uint32_t widest = u16 ? 16u : 32u;
                std::uint32_t cur = 0;
                bool good = true;
                for (std::uint32_t r = 0; r < t.h && good; ++r) {
                    if (words - cur < 2) {
                        good = false;
                        break;
                    }
                    const std::uint32_t seed = stage[mis + cur];
                    const std::uint32_t bits = stage[mis + cur + 1];
                    if (bits > widest || (u16 && seed > 65535u)) {
                        good = false;
                        break;
                    }
                    const std::uint32_t packed = ((t.w - 1) * bits + 31) / 32;
                    if (words - cur - 2 < packed) {
                        good = false;
                        break;
                    }
                    row_start[r] = static_cast<std::uint16_t>(cur);
                    cur += 2 + packed;
                }
                shared_flag = good && cur == words;
            }
            __syncthreads();
            if (!shared_flag) {
                if (tid == 0) fail(a.results, task.image, kMalformed);
                continue;
            }
        }

        const std::uint64_t rx = img.roi_x, ry = img.roi_y;
        t.top = ry > t.ty ? static_cast<std::uint32_t>(ry - t.ty) : 0u;
        t.bottom = ry + img.roi_h < t.ty + t.h ? static_cast<std::uint32_t>(ry + img.roi_h - t.ty) : t.h;
        t.left = rx > t.tx ? static_cast<std::uint32_t>(rx - t.tx) : 0u;
        t.right = rx + img.roi_w < t.tx + t.w ? static_cast<std::uint32_t>(rx + img.roi_w - t.tx) : t.w;
        t.base = mis;

        if constexpr (KIND == TSK_KIND_U16) {
            decode_tile<G, true, false>(img, t, stage, row_start);
        } else if constexpr (KIND == TSK_KIND_F32) {
            decode_tile<G, false, false>(img, t, stage, row_start);
        } else if constexpr (KIND == TSK_KIND_U16_CAL) {
            decode_tile<G, true, true>(img, t, stage, row_start);
        } else if constexpr (KIND == TSK_KIND_F32_CAL) {
            decode_tile<G, false, true>(img, t, stage, row_start);
        } else {
            if (u16) {
                if (img.calibrated) decode_tile<G, true, true>(img, t, stage, row_start);
                else decode_tile<G, true, false>(img, t, stage, row_start);
            } else {
                if (img.calibrated) decode_tile<G, false, true>(img, t, stage, row_start);
                else decode_tile<G, false, false>(img, t, stage, row_start);
            }
        }
    }
}

__global__ void results_begin_kernel(DeviceResultBits *results, std::uint32_t count) {
    const std::uint64_t stride = std::uint64_t{gridDim.x} * blockDim.x;
    for (std::uint64_t i = std::uint64_t{blockIdx.x} * blockDim.x + threadIdx.x; i < count; i += stride)
        results[i] = {0u, 0u, 0u};
}

__global__ void results_complete_kernel(const tsk_decode_image *images, std::uint32_t count, DeviceResultBits *results) {
    const std::uint64_t stride = std::uint64_t{gridDim.x} * blockDim.x;
    for (std::uint64_t i = std::uint64_t{blockIdx.x} * blockDim.x + threadIdx.x; i < count; i += stride) {
        const tsk_decode_image &img = images[i];
        std::uint32_t status = results[i].status;
        if (img.producer != nullptr) {
            const std::uint32_t produced = static_cast<const DeviceResultBits *>(img.producer)->status;
            if (produced != 0) status = produced;
        }
        const std::uint64_t s = img.calibrated ? 4 : (img.type == TSK_UINT16 ? 2 : 4);
        results[i] = {status, 0u, status == 0 ? std::uint64_t{img.roi_w} * img.roi_h * s : 0u};
    }
}

__global__ void fill_kernel(std::uint32_t *data, std::uint64_t words, std::uint32_t value) {
    const std::uint64_t stride = std::uint64_t{gridDim.x} * blockDim.x;
    for (std::uint64_t i = std::uint64_t{blockIdx.x} * blockDim.x + threadIdx.x; i < words; i += stride)
        data[i] = value;
}

using DecodeFn = void (*)(tsk_decode_args);

template <int KIND> DecodeFn pick(std::uint32_t group) {
    switch (group) {
    case 1: return decode_kernel<1, KIND>;
    case 2: return decode_kernel<2, KIND>;
    case 4: return decode_kernel<4, KIND>;
    case 8: return decode_kernel<8, KIND>;
    case 16: return decode_kernel<16, KIND>;
    default: return nullptr;
    }
}

DecodeFn decode_function(std::uint32_t kind, std::uint32_t group) {
    switch (kind) {
    case TSK_KIND_U16: return pick<TSK_KIND_U16>(group);
    case TSK_KIND_F32: return pick<TSK_KIND_F32>(group);
    case TSK_KIND_U16_CAL: return pick<TSK_KIND_U16_CAL>(group);
    case TSK_KIND_F32_CAL: return pick<TSK_KIND_F32_CAL>(group);
    case TSK_KIND_MIXED: return pick<TSK_KIND_MIXED>(group);
    default: return nullptr;
    }
}

std::uint32_t small_grid(std::uint64_t n) {
    const std::uint64_t blocks = n / 256 + 1;
    return static_cast<std::uint32_t>(blocks < 1024 ? blocks : 1024);
}

} // namespace

extern "C" {

int tsk_results_begin(void *results, std::uint32_t count, void *stream) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(results_begin_kernel, dim3(small_grid(count)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), static_cast<DeviceResultBits *>(results), count);
    return hipGetLastError();
}

int tsk_results_complete(const tsk_decode_image *images, std::uint32_t count, void *results, void *stream) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(results_complete_kernel, dim3(small_grid(count)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), images, count, static_cast<DeviceResultBits *>(results));
    return hipGetLastError();
}

int tsk_decode(const tsk_decode_args *args, void *stream) {
    if (args->task_count == 0) return hipSuccess;
    const DecodeFn fn = decode_function(args->kind, args->group);
    const std::uint32_t block = args->block;
    if (fn == nullptr || (block != 64 && block != 128 && block != 256 && block != 512) || args->grid == 0 ||
        (args->queue != nullptr && args->chunk == 0))
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(fn, dim3(args->grid), dim3(block), 0, static_cast<hipStream_t>(stream), *args);
    return hipGetLastError();
}

int tsk_decode_occupancy(std::uint32_t kind, std::uint32_t group, std::uint32_t block, int *blocks) {
    const DecodeFn fn = decode_function(kind, group);
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

// Loads the decode kernels (all lane-group widths of the kinds `op` uses).
int tsk_preload_decode(std::uint32_t calibrated) {
    hipFuncAttributes attr;
    const std::uint32_t kinds[3] = {calibrated ? std::uint32_t{TSK_KIND_U16_CAL} : std::uint32_t{TSK_KIND_U16},
                                    calibrated ? std::uint32_t{TSK_KIND_F32_CAL} : std::uint32_t{TSK_KIND_F32},
                                    TSK_KIND_MIXED};
    for (std::uint32_t kind : kinds)
        for (std::uint32_t g : {1u, 2u, 4u, 8u, 16u}) {
            const hipError_t e = hipFuncGetAttributes(&attr, reinterpret_cast<const void *>(decode_function(kind, g)));
            if (e != hipSuccess) return e;
        }
    hipError_t e = hipFuncGetAttributes(&attr, reinterpret_cast<const void *>(results_begin_kernel));
    if (e == hipSuccess) e = hipFuncGetAttributes(&attr, reinterpret_cast<const void *>(results_complete_kernel));
    if (e == hipSuccess) e = hipFuncGetAttributes(&attr, reinterpret_cast<const void *>(fill_kernel));
    return e;
}

} // extern "C"
