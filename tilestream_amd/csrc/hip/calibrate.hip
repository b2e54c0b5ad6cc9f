// Standalone calibration y = RN(RN(x - offset) * gain) for gfx950
// (reference src/cuda/calibrate.cu:13-62; docs/numerical-analysis.md).
// Rows of all images are concatenated; a workgroup takes rows with a stride,
// finds the row's image by binary search over first_rows, and its lanes
// stream the row four pixels at a time: 16-byte accesses to the float32
// input, the maps and the output (word alignment suffices for multi-word
// accesses), 8 bytes of uint16 input when the input row starts on a word
// (four 2-byte loads otherwise); the last width % 4 pixels one per lane.
#include <hip/hip_runtime.h>

#include <tilestream_kernels.h>

#include <cstdint>

namespace {

struct DeviceResultBits {
    std::uint32_t status;
    std::uint32_t reserved;
    std::uint64_t bytes;
};

template <typename T> using gptr = T __attribute__((address_space(1))) *;
// Vectors of words that need only word alignment.
typedef float f32x4w __attribute__((ext_vector_type(4), aligned(4)));
typedef std::uint32_t u32x2w __attribute__((ext_vector_type(2), aligned(4)));

template <bool U16> __device__ inline float sample(gptr<const std::uint8_t> row, std::uint32_t x) {
    if constexpr (U16) return static_cast<float>(reinterpret_cast<gptr<const std::uint16_t>>(row)[x]);
    else return reinterpret_cast<gptr<const float>>(row)[x];
}

// Two roundings to nearest even, never contracted into an FMA.
__device__ inline float calibrate(float x, float o, float g) { return __fmul_rn(__fsub_rn(x, o), g); }

template <bool U16> __global__ void __launch_bounds__(256) calibrate_kernel(tsk_calibrate_args a) {
    // The result records: the host checked every image before the launch, so
    // each succeeds with its FP32 output bytes; they are visible to the caller
    // when the kernel has completed, like the outputs.
    if (blockIdx.x == 0) {
        auto *results = static_cast<DeviceResultBits *>(a.results);
        for (std::uint32_t i = threadIdx.x; i < a.image_count; i += blockDim.x)
            results[i] = {0u, 0u, std::uint64_t{a.images[i].width} * a.images[i].height * 4u};
    }
    for (std::uint64_t row = blockIdx.x; row < a.total_rows; row += gridDim.x) {
        std::uint32_t lo = 0, hi = a.image_count;
        while (hi - lo > 1) {
            const std::uint32_t mid = lo + (hi - lo) / 2;
            if (a.first_rows[mid] <= row) lo = mid;
            else hi = mid;
        }
        const tsk_calibrate_image &img = a.images[lo];
        const std::uint64_t y = row - a.first_rows[lo];
        const auto in = reinterpret_cast<gptr<const std::uint8_t>>(reinterpret_cast<std::uintptr_t>(img.in + y * img.in_pitch));
        const auto out = reinterpret_cast<gptr<float>>(reinterpret_cast<std::uintptr_t>(img.out + y * img.out_pitch));
        const auto om = img.offset_map ? reinterpret_cast<gptr<const float>>(reinterpret_cast<std::uintptr_t>(
                                              img.offset_map + y * img.offset_pitch))
                                        : nullptr;
        const auto gm = img.gain_map ? reinterpret_cast<gptr<const float>>(reinterpret_cast<std::uintptr_t>(
                                            img.gain_map + y * img.gain_pitch))
                                      : nullptr;
        const std::uint32_t width = img.width, groups = width / 4;
        const bool in_words = !U16 || reinterpret_cast<std::uintptr_t>(in) % 4 == 0;
        // Each lane loads all its groups of a pass (up to kPass) before it
        // stores any: a wait for a load also waits for every store issued
        // before it, so loads placed after stores would wait once per group.
        constexpr std::uint32_t kPass = 4;
        for (std::uint32_t q0 = threadIdx.x; q0 < groups; q0 += kPass * blockDim.x) {
            float v[kPass][4];
            f32x4w o[kPass], g[kPass];
#pragma unroll
            for (std::uint32_t k = 0; k < kPass; ++k) {
                const std::uint32_t q = q0 + k * blockDim.x;
                if (q >= groups) break;
                const std::uint32_t x = 4 * q;
                if constexpr (U16) {
                    if (in_words) {
                        const u32x2w w = __builtin_nontemporal_load(reinterpret_cast<gptr<const u32x2w>>(in + 2 * x));
                        v[k][0] = static_cast<float>(w.x & 0xffffu);
                        v[k][1] = static_cast<float>(w.x >> 16);
                        v[k][2] = static_cast<float>(w.y & 0xffffu);
                        v[k][3] = static_cast<float>(w.y >> 16);
                    } else {
#pragma unroll
                        for (int c = 0; c < 4; ++c) v[k][c] = sample<true>(in, x + c);
                    }
                } else {
                    const f32x4w w = __builtin_nontemporal_load(reinterpret_cast<gptr<const f32x4w>>(in + 4 * x));
                    v[k][0] = w.x, v[k][1] = w.y, v[k][2] = w.z, v[k][3] = w.w;
                }
                o[k] = om ? __builtin_nontemporal_load(reinterpret_cast<gptr<const f32x4w>>(om + x))
                          : f32x4w{img.offset, img.offset, img.offset, img.offset};
                g[k] = gm ? __builtin_nontemporal_load(reinterpret_cast<gptr<const f32x4w>>(gm + x))
                          : f32x4w{img.gain, img.gain, img.gain, img.gain};
            }
#pragma unroll
            for (std::uint32_t k = 0; k < kPass; ++k) {
                const std::uint32_t q = q0 + k * blockDim.x;
                if (q >= groups) break;
                const f32x4w r = {calibrate(v[k][0], o[k].x, g[k].x), calibrate(v[k][1], o[k].y, g[k].y),
                                  calibrate(v[k][2], o[k].z, g[k].z), calibrate(v[k][3], o[k].w, g[k].w)};
                __builtin_nontemporal_store(r, reinterpret_cast<gptr<f32x4w>>(out + 4 * q));
            }
        }
        const std::uint32_t x = 4 * groups + threadIdx.x;
        if (x < width) out[x] = calibrate(sample<U16>(in, x), om ? om[x] : img.offset, gm ? gm[x] : img.gain);
    }
}

} // namespace

extern "C" {

int tsk_calibrate(const tsk_calibrate_args *a, void *stream) {
    if (a->image_count == 0) return hipSuccess;
    const hipStream_t s = static_cast<hipStream_t>(stream);
    const std::uint32_t grid = static_cast<std::uint32_t>(a->total_rows < 65536 ? a->total_rows : 65536);
    if (a->type == TSK_UINT16) hipLaunchKernelGGL(calibrate_kernel<true>, dim3(grid), dim3(256), 0, s, *a);
    else hipLaunchKernelGGL(calibrate_kernel<false>, dim3(grid), dim3(256), 0, s, *a);
    return hipGetLastError();
}

int tsk_preload_calibrate(void) {
    hipFuncAttributes attr;
    hipError_t e = hipFuncGetAttributes(&attr, reinterpret_cast<const void *>(calibrate_kernel<true>));
    if (e == hipSuccess) e = hipFuncGetAttributes(&attr, reinterpret_cast<const void *>(calibrate_kernel<false>));
    return e;
}

int tsk_preload_decode(std::uint32_t calibrated);
int tsk_preload_encode(void);

int tsk_preload(std::uint32_t op) {
    switch (op) {
    case 0: return tsk_preload_encode();
    case 1: return tsk_preload_decode(0);
    case 2: return tsk_preload_calibrate();
    case 3: return tsk_preload_decode(1);
    default: return hipErrorInvalidValue;
    }
}

} // extern "C"
