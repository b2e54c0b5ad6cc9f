// TileStream GPU API for AMD Instinct (gfx950), HIP runtime.
//
// The types and entry points keep the reference's names, fields and semantics
// (reference include/tilestream/cuda.h:16-187) with hipStream_t in place of the
// CUDA stream type:
//   prepare_batch            <- reference cuda.h:176-177
//   compress_async           <- reference cuda.h:178-179
//   decompress_async         <- reference cuda.h:180-181
//   decode_roi_async         <- reference cuda.h:182-183
//   calibrate_async          <- reference cuda.h:184-185
//   decode_calibrate_async   <- reference cuda.h:186-187
//   DeviceBuffer, synchronize <- reference cuda.h:113-145
// Behind them the work runs in hand-written HIP kernels reached through the
// C-ABI declared in tilestream_c.h.
#pragma once

#include <tilestream/api.h>

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <memory>

namespace tilestream::hip {

enum class Operation : std::uint32_t {
    compress,
    decompress,
    decode_roi,
    calibrate,
    decode_calibrate,
};

enum class ParameterKind : std::uint32_t { scalar, map };

struct Parameter {
    ParameterKind kind = ParameterKind::scalar;
    float scalar = 0.0F;
    ConstImageView map; // FP32, exactly the output region's size, output-local coordinates.
};

// y = RN(RN(x - offset) * gain) in binary32.
struct Calibration {
    Parameter offset;
    Parameter gain{ParameterKind::scalar, 1.0F, {}};
};

// A staged copy of one validated tile payload at `input_offset` of the bytes.
struct TileMapping {
    std::uint64_t tile_index = 0;
    std::uint64_t input_offset = 0;
};

struct CompressedInput {
    ConstBytes bytes;
    const ImageMetadata *metadata = nullptr;
    const TileMapping *mappings = nullptr;
    std::uint64_t mapping_count = 0;
    // Trusted path: the output of a successful compress_async of this library.
    const DeviceResult *encoder_result = nullptr;
    ImageSpec trusted_image;
    TileShape trusted_tile;
};

struct BatchItem {
    ConstImageView input;
    ImageView output;
    MutableBytes compressed_output;
    CompressedInput compressed_input;
    Roi roi;
    Calibration calibration;
};

struct BatchRequest {
    Operation operation = Operation::compress;
    const BatchItem *items = nullptr;
    std::uint64_t count = 0;
    TileShape tile;
};

// Public native records of the reference API.  The plans of this library keep
// their own compact device tables; these types are provided for source
// compatibility (16, 40 bytes as in the reference).
struct DeviceSegment {
    std::uint32_t seed = 0;
    std::uint32_t bit_width = 0;
    std::uint64_t input_offset = 0;
};

struct DeviceTile {
    std::uint64_t tile_index = 0;
    std::uint64_t input_offset = 0;
    std::uint64_t length = 0;
    std::uint64_t first_segment = 0;
    TileMode mode = TileMode::raw;
    std::uint16_t width = 0;
    std::uint16_t height = 0;
};

struct DeviceImage {
    ConstImageView input;
    ImageView output;
    ConstBytes compressed_input;
    MutableBytes compressed_output;
    const DeviceTile *tiles = nullptr;
    const DeviceSegment *segments = nullptr;
    std::uint64_t tile_count = 0;
    const DeviceResult *encoder_result = nullptr;
    ImageSpec archive_image;
    TileShape archive_tile;
    Roi roi;
    Calibration calibration;
};

struct DeviceBatch {
    const DeviceImage *images = nullptr;
    std::uint32_t count = 0;
};

struct ResultBuffer {
    DeviceResult *data = nullptr;
    std::uint64_t count = 0;
};

static_assert(sizeof(DeviceSegment) == 16);
static_assert(sizeof(DeviceTile) == 40);

// Owning device memory from the stream-ordered allocator, aligned to
// workspace_alignment.  Execution functions never allocate.
class DeviceBuffer {
  public:
    DeviceBuffer() noexcept = default;
    ~DeviceBuffer(); // frees on hipStreamPerThread without waiting
    DeviceBuffer(DeviceBuffer &&other) noexcept;
    DeviceBuffer &operator=(DeviceBuffer &&other) noexcept;
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;

    [[nodiscard]] static Status allocate(std::uint64_t bytes, hipStream_t stream,
                                         DeviceBuffer &buffer) noexcept;
    [[nodiscard]] Status release(hipStream_t stream) noexcept;

    [[nodiscard]] void *data() const noexcept { return data_; }
    [[nodiscard]] std::uint64_t size() const noexcept { return size_; }
    [[nodiscard]] MutableBytes bytes() const noexcept { return {data_, size_}; }

  private:
    void *data_ = nullptr;
    std::uint64_t size_ = 0;
};

// Waits for `stream`; any launch or execution failure is reported as hip_error.
[[nodiscard]] Status synchronize(hipStream_t stream) noexcept;

// A validated, uploaded batch.  One execution at a time per plan.
class BatchPlan {
  public:
    ~BatchPlan();
    BatchPlan(const BatchPlan &) = delete;
    BatchPlan &operator=(const BatchPlan &) = delete;
    BatchPlan(BatchPlan &&) noexcept;
    BatchPlan &operator=(BatchPlan &&) noexcept;

    [[nodiscard]] WorkspaceSpec workspace_spec() const noexcept;

    struct Impl;
    [[nodiscard]] const Impl *impl() const noexcept { return impl_.get(); }

  private:
    std::unique_ptr<Impl> impl_;
    explicit BatchPlan(std::unique_ptr<Impl> impl) noexcept;
    friend Status prepare_batch(const BatchRequest &, std::unique_ptr<BatchPlan> &,
                                hipStream_t) noexcept;
};

[[nodiscard]] Status prepare_batch(const BatchRequest &request, std::unique_ptr<BatchPlan> &plan,
                                   hipStream_t stream) noexcept;
[[nodiscard]] Status compress_async(const BatchPlan &plan, MutableBytes workspace,
                                    ResultBuffer results, hipStream_t stream) noexcept;
[[nodiscard]] Status decompress_async(const BatchPlan &plan, MutableBytes workspace,
                                      ResultBuffer results, hipStream_t stream) noexcept;
[[nodiscard]] Status decode_roi_async(const BatchPlan &plan, MutableBytes workspace,
                                      ResultBuffer results, hipStream_t stream) noexcept;
[[nodiscard]] Status calibrate_async(const BatchPlan &plan, MutableBytes workspace,
                                     ResultBuffer results, hipStream_t stream) noexcept;
[[nodiscard]] Status decode_calibrate_async(const BatchPlan &plan, MutableBytes workspace,
                                            ResultBuffer results, hipStream_t stream) noexcept;

// compress_async without a workspace (no counterpart in the reference): for
// callers that cannot allocate device memory (code under graph capture, tight
// serving loops).  The encoder's look-back state lives in a static 1 MiB
// device array of the library, one per GPU, held until the process ends; the
// whole batch runs as one launch when its state fits (65,535 tiles), else by
// windows of at most 4096 tiles, and a plan whose state cannot fit returns
// resource_limit.  Calls share that array: calls that overlap on one GPU (two
// streams, two host threads) race and must be ordered by the caller.  It
// allocates and frees nothing.
[[nodiscard]] Status compress_async_static(const BatchPlan &plan, ResultBuffer results,
                                           hipStream_t stream) noexcept;

} // namespace tilestream::hip
