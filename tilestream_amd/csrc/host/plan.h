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

// Consecutive encode tiles per XCD (tsk_encode_args.run).  On the 256-image
// batch, runs of 1, 2, 4, 8, 16, 32 tiles took 2.054, 2.041, 2.023, 2.077,
// 2.284, 2.481 ms per step; runs of 8 cut the pixel bytes fetched from 1.172 x
// to 1.053 x the pixels.
inline constexpr std::uint32_t encode_run() noexcept { return 4u; }

// A half-open address interval [lo, hi).
struct Span {
    std::uintptr_t lo = 0;
    std::uintptr_t hi = 0;
    [[nodiscard]] bool empty() const noexcept { return lo == hi; }
    [[nodiscard]] static Span of(const void *data, std::uint64_t bytes) noexcept {
        const auto at = reinterpret_cast<std::uintptr_t>(data);
        return {at, at + static_cast<std::uintptr_t>(bytes)};
    }
};

// The memory a prepared batch touches, checked against the API's aliasing
// rule: a buffer the batch writes overlaps no other buffer of the batch,
// while buffers it only reads may overlap each other.  Kept as the sorted
// union of all of them, so a buffer an execution call adds (workspace,
// results) is checked with one binary search.
class Footprint {
  public:
    // Records the batch's buffers; false if the rule is broken.
    [[nodiscard]] bool admit(std::vector<Span> written, std::vector<Span> read);
    // Whether `s` overlaps none of the recorded buffers.
    [[nodiscard]] bool avoids(Span s) const noexcept;

  private:
    std::vector<Span> union_; // sorted, coalesced
};

// Splits the concatenated tiles of consecutive images (tiles[i] of image i)
// into windows of at most `window` tiles: window k covers tiles
// [k window, (k + 1) window) of the concatenation, and holds one entry per
// image it intersects (window-relative first tile, count, image-relative
// start, flags 1: continues an earlier window, 2: finishes the image).
struct WindowSpan {
    std::size_t first_entry = 0;
    std::uint32_t entry_count = 0;
    std::uint32_t tile_count = 0;
};
struct WindowSplit {
    std::vector<tsk_window_entry> entries;
    std::vector<WindowSpan> windows;
};
[[nodiscard]] WindowSplit split_into_windows(const std::vector<std::uint64_t> &tiles, std::uint32_t window);

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
    Footprint footprint;
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
[[nodiscard]] Status on_device(const void *data, int device) noexcept;
[[nodiscard]] Status check_tile(const ImageMetadata &metadata, std::uint64_t index) noexcept;
[[nodiscard]] Status stream_device(hipStream_t stream, int device) noexcept;

// What a prepared batch asks of the buffers and stream of an execution call.
struct CallNeeds {
    const Footprint *footprint = nullptr;
    int device = 0;
    bool pageable = false;
    std::uint64_t results = 0;   // result records
    std::uint64_t workspace = 0; // workspace bytes
};
// The host checks of an execution call, before anything is enqueued, in the
// order the reference API reports them: representable ranges, capacities,
// alignment, aliasing with the batch and each other, the stream's device,
// where the buffers live, and a runtime error left pending by the caller.
[[nodiscard]] Status check_call(const CallNeeds &needs, MutableBytes workspace, ResultBuffer results,
                                hipStream_t stream) noexcept;
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
    std::vector<Span> writable;
    std::vector<Span> readable;
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
