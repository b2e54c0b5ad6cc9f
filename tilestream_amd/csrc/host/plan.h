// Internal state of a prepared batch and the host checks shared by the entry
// points.
#pragma once

#include "checked.h"

#include <tilestream/hip.h>
#include <tilestream_kernels.h>

#include <cstdint>
#include <vector>

namespace tilestream::hip::detail {

// Tiles per encode window: bounds the encoder's workspace for any batch size.
inline constexpr std::uint32_t window_tiles = 4096;

struct Range {
    std::uintptr_t begin = 0;
    std::uintptr_t end = 0;
};

// Writable ranges are pairwise disjoint and disjoint from every readable range;
// readable ranges may overlap one another (reference src/cuda/plan.cpp:21-61).
class RangeSet {
  public:
    [[nodiscard]] bool assign(std::vector<Range> writable, std::vector<Range> readable);
    [[nodiscard]] bool disjoint(Range range) const noexcept;

  private:
    std::vector<Range> writable_;
    std::vector<Range> readable_;
    std::vector<std::uintptr_t> reach_; // running maximum end of the sorted readable ranges
};

struct EncodeWindow {
    const tsk_window_entry *entries = nullptr; // device
    const std::uint32_t *blocks = nullptr;     // device: entry of window tile 64 b
    std::uint32_t entry_count = 0;
    std::uint32_t tile_count = 0;
};

// Appends, for each block of 64 tiles of a window, the index of the entry
// (among the window's `count` entries, ordered by first tile) holding the
// block's first tile: the encoder's lookup from tile to entry.
inline void append_block_owners(const tsk_window_entry *entries, std::size_t count, std::uint32_t tiles,
                                std::vector<std::uint32_t> &out) {
    std::size_t e = 0;
    for (std::uint64_t tile = 0; tile < tiles; tile += 64) {
        while (e + 1 < count && entries[e + 1].first <= tile) ++e;
        out.push_back(static_cast<std::uint32_t>(e));
    }
}

// Launch shape of the decode kernel.
struct DecodeShape {
    std::uint32_t group = 8;
    std::uint32_t block = 256;
    std::uint32_t grid = 0; // 0: one workgroup per tile
    std::uint32_t depth = 1;
};

struct PlanState {
    Operation operation = Operation::compress;
    PixelType type = PixelType::uint16;
    int device = 0;
    bool pageable = false;
    std::uint64_t count = 0;
    WorkspaceSpec spec;
    host::WorkspaceLayout layout;
    RangeSet ranges;
    void *storage = nullptr; // one device allocation holding every table

    // decode / fused decode
    const tsk_decode_image *decode_images = nullptr;
    const tsk_decode_task *tasks = nullptr;
    const std::uint16_t *rows = nullptr;
    std::uint64_t task_count = 0;
    DecodeShape shape;
    // host copies, kept for the batch queue and the serving loop
    std::vector<tsk_decode_image> host_images;
    std::vector<tsk_decode_task> host_tasks;
    std::vector<std::uint16_t> host_rows;

    // encode
    const tsk_encode_image *encode_images = nullptr;
    std::vector<EncodeWindow> windows;
    // All images as one window, used when the workspace's residual region
    // holds both look-back buffers for every tile of the batch.
    EncodeWindow whole;

    // standalone calibration
    const tsk_calibrate_image *calibrate_images = nullptr;
    const std::uint64_t *first_rows = nullptr;
    std::uint64_t total_rows = 0;

    PlanState() = default;
    PlanState(const PlanState &) = delete;
    PlanState &operator=(const PlanState &) = delete;
    ~PlanState();
};

// A HIP runtime failure is reported through Status, so the runtime's last
// error is cleared and cannot be blamed on a later launch.
inline Status hip_failure() noexcept {
    static_cast<void>(hipGetLastError());
    return Status::hip_error;
}

inline Status hip_status(int error) noexcept { return error == hipSuccess ? Status::success : hip_failure(); }

[[nodiscard]] Status accessible(const void *data, int device, bool pageable) noexcept;
[[nodiscard]] Status check_tile(const ImageMetadata &metadata, std::uint64_t index) noexcept;
[[nodiscard]] Status stream_device(hipStream_t stream, int device) noexcept;
[[nodiscard]] Status check_execution(const PlanState *plan, Operation operation, MutableBytes workspace,
                                     ResultBuffer results, hipStream_t stream) noexcept;
[[nodiscard]] Status run_decode(const PlanState &plan, DeviceResult *results, hipStream_t stream,
                                const DecodeShape &shape) noexcept;

// Validates one decode item and appends its image, tasks and row offsets.  Used
// by prepare_batch and by the batch queue, which lifts the one-type rule.
struct DecodeDraft {
    std::vector<tsk_decode_image> images;
    std::vector<tsk_decode_task> tasks;
    std::vector<std::uint16_t> rows;
    std::vector<std::uint64_t> tiles_per_image;
    std::vector<Range> writable;
    std::vector<Range> readable;
    std::vector<const void *> pointers;
    bool mixed_shapes = false;
    bool first = true;
    PixelType type = PixelType::uint16;
    TileShape shape;
};

[[nodiscard]] Status draft_decode_item(const BatchItem &item, Operation operation, bool require_one_type,
                                       DecodeDraft &draft);

// Groups the drafted tiles of every item into blocks for the block decoder:
// per image, `across` x `down` neighbouring tiles of its region covering at
// most 64 output rows of at most 1 KiB, with each tile's place in the
// block's LDS stage.  False when some tile's payload is only known on the
// device (encoder output decoded without host metadata).
[[nodiscard]] bool plan_blocks(const DecodeDraft &draft, std::vector<tsk_block> &blocks);

} // namespace tilestream::hip::detail

namespace tilestream::hip {
struct BatchPlan::Impl final : detail::PlanState {};
} // namespace tilestream::hip
