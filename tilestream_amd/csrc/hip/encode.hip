// Canonical GPU encoder for gfx950: the same bytes as the CPU encoder.
//
// Per window of at most 4096 tiles (reference src/cuda/encoder.cu:290-327):
//   analyze  one workgroup per tile: residuals, per-row minimal widths, the
//            three candidate costs and the mode with tie order RAW < CONSTANT <
//            PREDICT_BITPACK (docs/format.md "Canonical encoder").
//   scan     one workgroup: prefix of payload lengths over the window.
//   pack     one workgroup per tile: index entry and payload; every output word
//            has a single writer that gathers the residuals overlapping it.
//   finish   per image: the 64-byte header and the result record, or the output
//            cursor carried into the next window.
#include <hip/hip_runtime.h>

#include <tilestream_kernels.h>

#include <cstdint>

namespace {

constexpr std::uint32_t kThreads = 256;

struct DeviceResultBits {
    std::uint32_t status;
    std::uint32_t reserved;
    std::uint64_t bytes;
};

struct Geometry {
    std::uint64_t x, y;
    std::uint32_t w, h;
};

__device__ inline Geometry geometry(const tsk_encode_image &img, std::uint64_t tile) {
    Geometry g;
    g.x = tile % img.tile_columns * img.tile_w;
    g.y = tile / img.tile_columns * img.tile_h;
    g.w = static_cast<std::uint32_t>(img.width - g.x < img.tile_w ? img.width - g.x : img.tile_w);
    g.h = static_cast<std::uint32_t>(img.height - g.y < img.tile_h ? img.height - g.y : img.tile_h);
    return g;
}

// Window entries are ordered by their first tile.
__device__ inline const tsk_window_entry &owner(const tsk_window_entry *e, std::uint32_t n, std::uint32_t tile) {
    std::uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const std::uint32_t mid = lo + (hi - lo) / 2;
        if (e[mid].first <= tile) lo = mid;
        else hi = mid;
    }
    return e[lo];
}

__device__ inline std::uint64_t payload_start(const tsk_encode_image &img) { return 64 + 32 * img.tile_count; }

template <bool U16> __device__ inline std::uint32_t residual(std::uint32_t cur, std::uint32_t prev) {
    if constexpr (U16) {
        const std::uint32_t d = (cur - prev) & 0xffffu;
        return d < 32768u ? 2u * d : 2u * (65536u - d) - 1u;
    } else {
        return cur ^ prev;
    }
}

// Loads the tile's samples row-major (w wide) into LDS.
template <bool U16>
__device__ inline void load_tile(const tsk_encode_image &img, const Geometry &g, std::uint32_t *px) {
    const std::uint32_t n = g.w * g.h;
    const std::uint8_t *origin = img.in + g.y * img.in_pitch + g.x * (U16 ? 2 : 4);
    for (std::uint32_t k = threadIdx.x; k < n; k += kThreads) {
        const std::uint32_t r = k / g.w, c = k - r * g.w;
        const std::uint8_t *p = origin + std::uint64_t{r} * img.in_pitch;
        px[k] = U16 ? reinterpret_cast<const std::uint16_t *>(p)[c] : reinterpret_cast<const std::uint32_t *>(p)[c];
    }
}

__device__ inline std::uint32_t row_cost(std::uint32_t w, std::uint32_t b) { return 8u + (((w - 1u) * b + 31u) / 32u) * 4u; }

template <bool U16> __global__ void __launch_bounds__(kThreads) analyze_kernel(tsk_encode_args a) {
    __shared__ std::uint32_t px[4096];
    __shared__ std::uint32_t row_or[64];

    const std::uint32_t t = blockIdx.x;
    const tsk_window_entry &e = owner(a.entries, a.entry_count, t);
    const tsk_encode_image &img = a.images[e.image];
    const Geometry g = geometry(img, e.start + (t - e.first));
    if (threadIdx.x < 64) row_or[threadIdx.x] = 0;
    load_tile<U16>(img, g, px);
    __syncthreads();

    const std::uint32_t n = g.w * g.h;
    const std::uint32_t first = px[0];
    bool same = true;
    for (std::uint32_t k = threadIdx.x; k < n; k += kThreads) {
        const std::uint32_t r = k / g.w, c = k - r * g.w;
        same = same && px[k] == first;
        if (c != 0) {
            const std::uint32_t z = residual<U16>(px[k], px[k - 1]);
            if (z != 0) atomicOr(&row_or[r], z);
        }
    }
    const bool uniform = __syncthreads_and(same) != 0;

    // Row widths and the predictive cost, reduced over the first wave (h <= 64).
    std::uint32_t cost = 0;
    if (threadIdx.x < g.h) {
        const std::uint32_t any = row_or[threadIdx.x];
        const std::uint32_t b = any == 0 ? 0u : 32u - static_cast<std::uint32_t>(__builtin_clz(any));
        a.rows[std::uint64_t{t} * a.max_rows + threadIdx.x] = static_cast<std::uint8_t>(b);
        cost = row_cost(g.w, b);
    }
    if (threadIdx.x < 64) {
        for (int off = 32; off > 0; off >>= 1) cost += __shfl_xor(cost, off);
        if (threadIdx.x == 0) {
            const std::uint32_t raw4 = (g.w * g.h * (U16 ? 2u : 4u) + 3u) / 4u * 4u;
            std::uint32_t mode = 0, length = raw4;
            if (uniform && length > 4u) {
                mode = 1;
                length = 4;
            }
            if (cost < length) {
                mode = 2;
                length = cost;
            }
            a.tiles[2 * t] = mode;
            a.tiles[2 * t + 1] = length;
        }
    }
}

// offsets[0] = 0, offsets[t + 1] = sum of lengths of window tiles 0..t.
__global__ void __launch_bounds__(1024) scan_kernel(tsk_encode_args a) {
    __shared__ std::uint64_t wave_sums[16];
    const std::uint32_t n = a.tile_count;
    const std::uint32_t per = (n + 1023) / 1024;
    const std::uint32_t begin = threadIdx.x * per;
    std::uint64_t local = 0;
    for (std::uint32_t i = begin; i < begin + per && i < n; ++i) local += a.tiles[2 * i + 1];
    // Inclusive scan of the thread totals: within each wave, then over waves.
    const std::uint32_t lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    std::uint64_t incl = local;
    for (std::uint32_t off = 1; off < 64; off <<= 1) {
        const std::uint64_t other = __shfl_up(incl, off);
        if (lane >= off) incl += other;
    }
    if (lane == 63) wave_sums[wave] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        std::uint64_t run = 0;
        for (std::uint32_t w = 0; w < 16; ++w) {
            const std::uint64_t s = wave_sums[w];
            wave_sums[w] = run;
            run += s;
        }
    }
    __syncthreads();
    std::uint64_t running = wave_sums[wave] + incl - local;
    if (threadIdx.x == 0) a.offsets[0] = 0;
    for (std::uint32_t i = begin; i < begin + per && i < n; ++i) {
        running += a.tiles[2 * i + 1];
        a.offsets[i + 1] = running;
    }
}

// Word `m` of a row's packed residuals: the bits [32 m, 32 m + 32) of the
// LSB-first stream of w - 1 residuals of b bits.
template <bool U16>
__device__ inline std::uint32_t packed_word(const std::uint32_t *row, std::uint32_t w, std::uint32_t b, std::uint32_t m) {
    const std::uint32_t lo_bit = 32u * m;
    const std::uint32_t last = w - 2u; // last residual index
    std::uint32_t j = lo_bit / b;
    const std::uint32_t jmax = (lo_bit + 31u) / b < last ? (lo_bit + 31u) / b : last;
    std::uint32_t value = 0;
    for (; j <= jmax; ++j) {
        const std::uint32_t z = residual<U16>(row[j + 1], row[j]);
        const std::uint32_t pos = j * b;
        value |= pos >= lo_bit ? z << (pos - lo_bit) : z >> (lo_bit - pos);
    }
    return value;
}

template <bool U16> __global__ void __launch_bounds__(kThreads) pack_kernel(tsk_encode_args a) {
    __shared__ std::uint32_t px[4096];
    __shared__ std::uint32_t row_word[65];
    __shared__ std::uint32_t row_bits[64];

    const std::uint32_t t = blockIdx.x;
    const tsk_window_entry &e = owner(a.entries, a.entry_count, t);
    const tsk_encode_image &img = a.images[e.image];
    const std::uint64_t index = e.start + (t - e.first);
    const Geometry g = geometry(img, index);
    const std::uint32_t mode = a.tiles[2 * t], length = a.tiles[2 * t + 1];
    const std::uint64_t base = (e.flags & 1u) ? a.carry[a.parity] : payload_start(img);
    const std::uint64_t offset = base + a.offsets[t] - a.offsets[e.first];

    if (threadIdx.x < 8) {
        // 32-byte index entry: mode, reserved, width, height, offset, length, reserved.
        std::uint32_t v = 0;
        switch (threadIdx.x) {
        case 0: v = mode; break;
        case 1: v = g.w | (g.h << 16); break;
        case 2: v = static_cast<std::uint32_t>(offset); break;
        case 3: v = static_cast<std::uint32_t>(offset >> 32); break;
        case 4: v = length; break;
        default: v = 0; break;
        }
        reinterpret_cast<std::uint32_t *>(img.out + 64 + index * 32)[threadIdx.x] = v;
    }
    load_tile<U16>(img, g, px);
    if (mode == 2 && threadIdx.x < g.h) row_bits[threadIdx.x] = a.rows[std::uint64_t{t} * a.max_rows + threadIdx.x];
    __syncthreads();
    if (mode == 2 && threadIdx.x == 0) {
        std::uint32_t cur = 0;
        for (std::uint32_t r = 0; r < g.h; ++r) {
            row_word[r] = cur;
            cur += 2u + ((g.w - 1u) * row_bits[r] + 31u) / 32u;
        }
        row_word[g.h] = cur;
    }
    __syncthreads();

    std::uint32_t *dst = reinterpret_cast<std::uint32_t *>(img.out + offset);
    const std::uint32_t words = length / 4u;
    const std::uint32_t n = g.w * g.h;
    for (std::uint32_t k = threadIdx.x; k < words; k += kThreads) {
        std::uint32_t v = 0;
        if (mode == 0) {
            if (U16) {
                v = px[2 * k] & 0xffffu;
                if (2 * k + 1 < n) v |= px[2 * k + 1] << 16;
            } else {
                v = px[k];
            }
        } else if (mode == 1) {
            v = px[0];
        } else {
            std::uint32_t lo = 0, hi = g.h;
            while (hi - lo > 1) {
                const std::uint32_t mid = (lo + hi) / 2;
                if (row_word[mid] <= k) lo = mid;
                else hi = mid;
            }
            const std::uint32_t local = k - row_word[lo];
            if (local == 0) v = px[lo * g.w];
            else if (local == 1) v = row_bits[lo];
            else v = packed_word<U16>(px + lo * g.w, g.w, row_bits[lo], local - 2);
        }
        dst[k] = v;
    }
}

__global__ void finish_kernel(tsk_encode_args a) {
    const std::uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= a.entry_count) return;
    const tsk_window_entry e = a.entries[j];
    const tsk_encode_image &img = a.images[e.image];
    const std::uint64_t base = (e.flags & 1u) ? a.carry[a.parity] : payload_start(img);
    const std::uint64_t end = base + a.offsets[e.first + e.count] - a.offsets[e.first];
    if ((e.flags & 2u) == 0) {
        a.carry[a.parity ^ 1u] = end;
        return;
    }
    // 64-byte file header (docs/format.md "File header").
    const std::uint64_t count = img.tile_count;
    const std::uint64_t payload = payload_start(img);
    std::uint32_t h[16];
    h[0] = 0x00435354u;  // 'T' 'S' 'C' 0
    h[1] = 0x0a1a0a0du;  // 13 10 26 10
    h[2] = 1u | (0u << 16);             // major 1, minor 0
    h[3] = 64u | (img.type << 16);      // header size, pixel type, flags 0
    h[4] = img.width;
    h[5] = img.height;
    h[6] = img.tile_w | (std::uint32_t{img.tile_h} << 16);
    h[7] = 32u;                          // index entry size, reserved 0
    h[8] = static_cast<std::uint32_t>(count);
    h[9] = static_cast<std::uint32_t>(count >> 32);
    h[10] = 64u;
    h[11] = 0u;
    h[12] = static_cast<std::uint32_t>(payload);
    h[13] = static_cast<std::uint32_t>(payload >> 32);
    h[14] = static_cast<std::uint32_t>(end);
    h[15] = static_cast<std::uint32_t>(end >> 32);
    std::uint32_t *out = reinterpret_cast<std::uint32_t *>(img.out);
    for (int i = 0; i < 16; ++i) out[i] = h[i];
    static_cast<DeviceResultBits *>(a.results)[e.image] = {0u, 0u, end};
}

} // namespace

extern "C" {

int tsk_encode_analyze(const tsk_encode_args *a, void *stream) {
    if (a->tile_count == 0) return hipSuccess;
    if (a->type == TSK_UINT16)
        hipLaunchKernelGGL(analyze_kernel<true>, dim3(a->tile_count), dim3(kThreads), 0, static_cast<hipStream_t>(stream), *a);
    else
        hipLaunchKernelGGL(analyze_kernel<false>, dim3(a->tile_count), dim3(kThreads), 0, static_cast<hipStream_t>(stream), *a);
    return hipGetLastError();
}

int tsk_encode_scan(const tsk_encode_args *a, void *stream) {
    hipLaunchKernelGGL(scan_kernel, dim3(1), dim3(1024), 0, static_cast<hipStream_t>(stream), *a);
    return hipGetLastError();
}

int tsk_encode_pack(const tsk_encode_args *a, void *stream) {
    if (a->tile_count == 0) return hipSuccess;
    if (a->type == TSK_UINT16)
        hipLaunchKernelGGL(pack_kernel<true>, dim3(a->tile_count), dim3(kThreads), 0, static_cast<hipStream_t>(stream), *a);
    else
        hipLaunchKernelGGL(pack_kernel<false>, dim3(a->tile_count), dim3(kThreads), 0, static_cast<hipStream_t>(stream), *a);
    return hipGetLastError();
}

int tsk_encode_finish(const tsk_encode_args *a, void *stream) {
    if (a->entry_count == 0) return hipSuccess;
    hipLaunchKernelGGL(finish_kernel, dim3((a->entry_count + 255) / 256), dim3(256), 0, static_cast<hipStream_t>(stream), *a);
    return hipGetLastError();
}

int tsk_preload_encode(void) {
    hipFuncAttributes attr;
    const void *fns[] = {reinterpret_cast<const void *>(analyze_kernel<true>), reinterpret_cast<const void *>(analyze_kernel<false>),
                         reinterpret_cast<const void *>(scan_kernel),          reinterpret_cast<const void *>(pack_kernel<true>),
                         reinterpret_cast<const void *>(pack_kernel<false>),   reinterpret_cast<const void *>(finish_kernel)};
    for (const void *f : fns) {
        const hipError_t e = hipFuncGetAttributes(&attr, f);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

} // extern "C"
