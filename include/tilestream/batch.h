// Batched persistent decode: an extension of the reference API (no counterpart
// in the reference, which allows one pixel type per plan and launches one grid
// of one block per tile per window).
//
// A DecodeQueue holds a device work queue of (image, tile, region) tasks from
// items of any mix of pixel types, tile shapes, regions and calibrations.  One
// persistent launch drains it: workgroups draw chunks of tasks from a device
// counter that lives in the caller's workspace.  The launch (results init,
// counter reset, decode, results completion) can be captured once into a HIP
// graph and replayed.
#pragma once

#include <tilestream/hip.h>

#include <cstdint>
#include <memory>
#include <vector>

namespace tilestream::hip {

// Launch shape of the decode kernel.
struct LaunchShape {
    std::uint32_t group = 8;   // columns decoded per lane: 1, 2, 4, 8, 16
    std::uint32_t block = 256; // threads per workgroup: 64, 128, 256, 512
    std::uint32_t grid = 0;    // workgroups; 0 = resident maximum of the device
    std::uint32_t chunk = 8;   // tasks per draw from the work counter (last eighth); 0 = static stride only
    std::uint32_t depth = 2;   // tiles a workgroup holds: 1 = staged only, 2 or 3 = 1 or 2 more in flight
    std::uint32_t run = 1;     // neighbouring tasks a workgroup takes in a row from the static stride
};

struct QueueStats {
    std::uint64_t images = 0;
    std::uint64_t tasks = 0;
    std::uint64_t decoded_bytes = 0;   // region pixels x stored sample bytes
    std::uint64_t output_bytes = 0;    // region pixels x output sample bytes
    std::uint64_t payload_bytes = 0;   // padded payloads of the tiles read
    std::uint64_t map_bytes = 0;       // calibration map bytes read
};

class DecodeQueue {
  public:
    ~DecodeQueue();
    DecodeQueue(const DecodeQueue &) = delete;
    DecodeQueue &operator=(const DecodeQueue &) = delete;

    // Validates every item like prepare_batch does for `operations[i]`
    // (decompress, decode_roi or decode_calibrate) and uploads the tables.
    [[nodiscard]] static Status prepare(const BatchItem *items, const Operation *operations,
                                        std::uint64_t count, std::unique_ptr<DecodeQueue> &queue,
                                        hipStream_t stream) noexcept;

    [[nodiscard]] std::uint64_t workspace_bytes() const noexcept;
    [[nodiscard]] std::uint64_t count() const noexcept;
    [[nodiscard]] QueueStats stats() const noexcept;
    [[nodiscard]] LaunchShape shape() const noexcept;
    void set_shape(const LaunchShape &shape) noexcept;

    // Enqueues one persistent decode of the whole queue.
    [[nodiscard]] Status run(MutableBytes workspace, ResultBuffer results, hipStream_t stream) noexcept;
    // Captures run() into a HIP graph bound to these buffers; launch() replays it.
    [[nodiscard]] Status capture(MutableBytes workspace, ResultBuffer results, hipStream_t stream) noexcept;
    [[nodiscard]] Status launch(hipStream_t stream) noexcept;

    // Times every candidate shape (median of `repeats` runs after a warm-up)
    // and keeps the fastest; returns the per-candidate times in milliseconds.
    [[nodiscard]] Status autotune(MutableBytes workspace, ResultBuffer results, hipStream_t stream,
                                  const std::vector<LaunchShape> &candidates, std::uint32_t repeats,
                                  std::vector<float> &milliseconds) noexcept;

    struct Impl;

  private:
    explicit DecodeQueue(std::unique_ptr<Impl> impl) noexcept;
    std::unique_ptr<Impl> impl_;
};

} // namespace tilestream::hip
