// Standalone calibration y = RN(RN(x - offset) * gain) for gfx950
// (reference src/cuda/calibrate.cu:13-62; docs/numerical-analysis.md).
// Rows of all images are concatenated; a workgroup takes rows with a stride,
// finds the row's image by binary search over first_rows, and its lanes stream
// the row one pixel per lane per step (consecutive lanes, consecutive pixels).
#include <hip/hip_runtime.h>

#include <tilestream_kernels.h>

#include <cstdint>

namespace {

struct DeviceResultBits {
    std::uint32_t status;
    std::uint32_t reserved;
    std::uint64_t bytes;
};

template <bool U16> __device__ inline float sample(const std::uint8_t *row, std::uint32_t x) {
    if constexpr (U16) return static_cast<float>(reinterpret_cast<const std::uint16_t *>(row)[x]);
    else return reinterpret_cast<const float *>(row)[x];
}

template <bool U16> __global__ void __launch_bounds__(256) calibrate_kernel(tsk_calibrate_args a) {
    for (std::uint64_t row = blockIdx.x; row < a.total_rows; row += gridDim.x) {
        std::uint32_t lo = 0, hi = a.image_count;
        while (hi - lo > 1) {
            const std::uint32_t mid = lo + (hi - lo) / 2;
            if (a.first_rows[mid] <= row) lo = mid;
            else hi = mid;
        }
        const tsk_calibrate_image &img = a.images[lo];
        const std::uint64_t y = row - a.first_rows[lo];
        const std::uint8_t *in = img.in + y * img.in_pitch;
        float *out = reinterpret_cast<float *>(img.out + y * img.out_pitch);
        const float *om = img.offset_map ? reinterpret_cast<const float *>(img.offset_map + y * img.offset_pitch) : nullptr;
        const float *gm = img.gain_map ? reinterpret_cast<const float *>(img.gain_map + y * img.gain_pitch) : nullptr;
        for (std::uint32_t x = threadIdx.x; x < img.width; x += blockDim.x) {
            const float o = om ? om[x] : img.offset;
            const float g = gm ? gm[x] : img.gain;
            out[x] = __fmul_rn(__fsub_rn(sample<U16>(in, x), o), g);
        }
    }
}

__global__ void report_kernel(const tsk_calibrate_image *images, std::uint32_t count, DeviceResultBits *results) {
    const std::uint64_t stride = std::uint64_t{gridDim.x} * blockDim.x;
    for (std::uint64_t i = std::uint64_t{blockIdx.x} * blockDim.x + threadIdx.x; i < count; i += stride)
        results[i] = {0u, 0u, std::uint64_t{images[i].width} * images[i].height * 4u};
}

} // namespace

extern "C" {

int tsk_calibrate(const tsk_calibrate_args *a, void *stream) {
    if (a->image_count == 0) return hipSuccess;
    const hipStream_t s = static_cast<hipStream_t>(stream);
    const std::uint32_t grid = static_cast<std::uint32_t>(a->total_rows < 65536 ? a->total_rows : 65536);
    if (a->type == TSK_UINT16) hipLaunchKernelGGL(calibrate_kernel<true>, dim3(grid), dim3(256), 0, s, *a);
    else hipLaunchKernelGGL(calibrate_kernel<false>, dim3(grid), dim3(256), 0, s, *a);
    const std::uint32_t reports = a->image_count / 256 + 1 < 1024 ? a->image_count / 256 + 1 : 1024;
    hipLaunchKernelGGL(report_kernel, dim3(reports), dim3(256), 0, s, a->images, a->image_count,
                       static_cast<DeviceResultBits *>(a->results));
    return hipGetLastError();
}

int tsk_preload_calibrate(void) {
    hipFuncAttributes attr;
    hipError_t e = hipFuncGetAttributes(&attr, reinterpret_cast<const void *>(calibrate_kernel<true>));
    if (e == hipSuccess) e = hipFuncGetAttributes(&attr, reinterpret_cast<const void *>(calibrate_kernel<false>));
    if (e == hipSuccess) e = hipFuncGetAttributes(&attr, reinterpret_cast<const void *>(report_kernel));
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
