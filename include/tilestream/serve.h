// Serving resident archives: an extension of the reference API (no counterpart
// in the reference, which plans every request on the host with prepare_batch).
// Requests decode regions of resident archives or encode device images.
//
// ArchiveSet keeps archives resident in device memory, or in pinned host
// memory, together with their tile and row tables, planned once when an
// archive is added.  Tiles of archives in host memory are staged: each batch
// copies the payloads its requests touch into device memory with a kernel
// reading over the host link, while other batches decode.  A batch of region
// requests then uploads only small per-request records; the decode kernel
// derives each request's tiles on the device.  Server runs continuous
// batching over several streams: requests that arrive while batches are on
// the GPU are admitted as stream slots and output memory free up, and
// batches on different streams overlap on the device.
#pragma once

#include <tilestream/batch.h>
#include <tilestream/hip.h>

#include <cstdint>
#include <memory>

namespace tilestream::hip {

struct ArchiveInfo {
    ImageSpec image;
    TileShape tile;
    std::uint64_t tile_count = 0;
    std::uint64_t file_bytes = 0;
    std::uint32_t host = 0; // 1: in pinned host memory, staged per batch
    std::uint32_t reserved = 0;
};

class ArchiveSet {
  public:
    ~ArchiveSet();
    ArchiveSet(const ArchiveSet &) = delete;
    ArchiveSet &operator=(const ArchiveSet &) = delete;

    [[nodiscard]] static Status create(std::unique_ptr<ArchiveSet> &set) noexcept;

    // Adds a complete archive already in device memory of the current device,
    // or in mapped pinned host memory (hipHostMalloc), validated on the host
    // into `metadata` (parse_metadata of the same bytes).  The bytes must stay
    // allocated and unchanged while the set is in use.
    [[nodiscard]] Status add(ConstBytes device_file, const ImageMetadata &metadata, std::uint32_t &id) noexcept;
    // Adds an archive in device memory of the current device without host
    // metadata: its header, index and payloads are validated on the device,
    // with the host parser's rules and header statuses, and its tile table is
    // built there (no copy of the archive to the host).  `image` and `tile`
    // are what the caller expects the file to hold (a valid file of another
    // image is invalid_argument).  With `producer` (the DeviceResult of the
    // encode that wrote the file, in device memory), the file size is its
    // bytes_written and a failed encode's status is returned; without it the
    // file is exactly device_file.size bytes.  Work is ordered on `stream`
    // after whatever wrote the file, and the call waits for it.  No reference
    // counterpart: the reference's trusted decode checks encoder output per
    // decode (src/cuda/decoder.cu:86-110,191-220); here it is checked once
    // and served many times.
    [[nodiscard]] Status add_device(ConstBytes device_file, const ImageSpec &image, TileShape tile,
                                    const DeviceResult *producer, hipStream_t stream, std::uint32_t &id) noexcept;
    // Uploads the tables of every archive added so far (synchronizes `stream`).
    [[nodiscard]] Status commit(hipStream_t stream) noexcept;

    [[nodiscard]] std::uint32_t size() const noexcept;
    [[nodiscard]] Status info(std::uint32_t id, ArchiveInfo &out) const noexcept;

    struct Impl;
    [[nodiscard]] const Impl *impl() const noexcept { return impl_.get(); }

  private:
    explicit ArchiveSet(std::unique_ptr<Impl> impl) noexcept;
    std::unique_ptr<Impl> impl_;
};

enum class RequestKind : std::uint32_t {
    region = 0, // decode a region (or all) of a resident archive
    encode = 1, // compress a device image into a canonical .tsc file
};

// One request.  region: a region of a resident archive, decoded bit-exact
// (calibrated = 0) or fused with scalar calibration into FP32 (calibrated = 1).
// encode: the device image `input` compressed with tile shape `tile` into
// `compressed` (device, at least compress_bound bytes); the result's
// bytes_written is the file size.
struct ServeRequest {
    RequestKind kind = RequestKind::region;
    std::uint32_t archive = 0;    // region
    std::uint32_t calibrated = 0; // region
    TileShape tile;               // encode
    Roi roi;                      // region
    ImageView output;             // region: device; archive pixel type, or float32 when calibrated
    ConstImageView input;         // encode: device pixels
    MutableBytes compressed;      // encode: device output
    float offset = 0.0F;          // region, calibrated
    float gain = 1.0F;
};

struct ServerOptions {
    std::uint32_t streams = 3;           // batches in flight at once
    std::uint32_t max_batch = 1024;      // requests per batch
    std::uint32_t max_tasks = 1u << 22;  // tiles per batch (a larger request is rejected)
    std::uint32_t encode_window = 1u << 20; // tiles per encoder launch, 64 to 2^20 (look-back state: 16 B per tile)
    std::uint64_t admission_bytes = 200ull << 30; // output bytes in flight (encode: its compress bound)
    std::uint64_t staging_bytes = 256ull << 20;   // per stream: device staging for tiles of host archives
    LaunchShape shape;                   // decode kernel launch shape (grid 0: sized per batch)
    // A request whose output bytes (encode: its compress bound) reach this is
    // large: it is batched only with large requests, and large batches hold at
    // most streams - 1 streams, so small requests never wait behind a large one.
    std::uint64_t large_bytes = 32ull << 20;
    // A batch of region requests with at least this many tiles, every request
    // on a device-resident archive, is decoded by the block decoder (up to 8
    // neighbouring tiles per unit, block records built on the device); 0:
    // always tile by tile.  Smaller batches keep the tile decoder, whose
    // single launch has the lower latency.
    std::uint64_t block_tasks = 16384;
};

struct ServerStats {
    std::uint64_t requests = 0;
    std::uint64_t batches = 0;
    std::uint64_t decoded_bytes = 0;     // region pixels x stored sample bytes, plus encoded input bytes
    std::uint64_t failed = 0;            // requests whose device result was not success
    double seconds = 0.0;                // first arrival to last completion
    std::uint64_t admission_holds = 0;   // batches closed early because admission_bytes was reached
    std::uint64_t peak_in_flight = 0;    // most output bytes admitted and not yet finished
};

class Server {
  public:
    ~Server();
    Server(const Server &) = delete;
    Server &operator=(const Server &) = delete;

    [[nodiscard]] static Status create(const ArchiveSet &archives, const ServerOptions &options,
                                       std::unique_ptr<Server> &server) noexcept;

    // Validates a request against the archive set (no device work).
    [[nodiscard]] Status check(const ServeRequest &request) const noexcept;

    // Open-loop replay: request i arrives arrival_ns[i] nanoseconds after the
    // start (non-decreasing); requests of one size class (large_bytes) are
    // admitted in arrival order, the older head of the two classes first.  latency_ns[i] (optional) receives completion -
    // arrival as seen by the host, results[i] (optional, host memory) the
    // request's device result (status; bytes_written: output bytes, the file
    // size of an encode).  Every request is validated before the replay starts;
    // outputs must be pairwise disjoint and disjoint from the archives and the
    // encode inputs.
    [[nodiscard]] Status replay(const ServeRequest *requests, const std::uint64_t *arrival_ns,
                                std::uint64_t count, std::uint64_t *latency_ns, DeviceResult *results,
                                ServerStats &stats) noexcept;

    // Runs `count` requests, all arriving at once, and waits for them.
    [[nodiscard]] Status run_batch(const ServeRequest *requests, std::uint64_t count,
                                   DeviceResult *results) noexcept;

    struct Impl;

  private:
    explicit Server(std::unique_ptr<Impl> impl) noexcept;
    std::unique_ptr<Impl> impl_;
};

} // namespace tilestream::hip
