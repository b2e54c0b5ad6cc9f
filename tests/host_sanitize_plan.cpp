// The host planner (plan.cpp, queue.cpp, serve.cpp) under AddressSanitizer +
// UBSan, on the CPU stand-in for the HIP runtime of host_hip_stub.cpp (built
// and run by test_host_sanitizers.py): plans of every operation, parsed,
// staged and trusted (encoder output) decode items, the batched queue in both
// decoders with graph capture, the allocation-free encode, the validation
// rejections, and the serving path.  The stand-in's launchers abort on any
// table index a kernel would read outside the planner's tables.
#include <tilestream/api.h>
#include <tilestream/batch.h>
#include <tilestream/cpu.h>
#include <tilestream/hip.h>
#include <tilestream/serve.h>
#include <tilestream/sizing.h>

#include <cstdio>
#include <cstring>
#include <memory>
#include <random>
#include <vector>

using namespace tilestream;
using namespace tilestream::hip;

namespace {

int failures = 0;

void expect(bool ok, const char *what, int line) {
    if (ok) return;
    std::printf("FAILED line %d: %s\n", line, what);
    ++failures;
}
#define EXPECT(cond) expect((cond), #cond, __LINE__)

// Buffers of the harness itself, freed at the end so that the leak check
// sees only what the planner keeps.
std::vector<void *> &owned() {
    static std::vector<void *> all;
    return all;
}

void *device_alloc(std::size_t bytes) {
    void *p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
    owned().push_back(p);
    return p;
}

// One archive on the host and in "device" memory with its metadata.
struct Archive {
    PixelType type;
    std::uint32_t w, h;
    TileShape tile;
    std::vector<std::uint8_t> pixels, file;
    ImageMetadata meta;
    void *device = nullptr;
};

Archive make_archive(PixelType type, std::uint32_t w, std::uint32_t h, TileShape tile, std::mt19937 &rng) {
    Archive a{type, w, h, tile, {}, {}, {}, nullptr};
    const std::uint32_t s = type == PixelType::uint16 ? 2 : 4;
    a.pixels.resize(std::size_t{w} * h * s);
    const int style = static_cast<int>(rng() % 3); // smooth, constant, noise: all three tile modes
    for (std::size_t i = 0; i < a.pixels.size(); ++i)
        a.pixels[i] = static_cast<std::uint8_t>(style == 0 ? (i / s) % w / 8 + (rng() & 1) : style == 1 ? 7 : rng());
    std::uint64_t bound = 0;
    EXPECT(compress_bound(ImageSpec{type, w, h}, tile, bound) == Status::success);
    a.file.resize(bound);
    std::uint64_t written = 0;
    const ConstImageView in{a.pixels.data(), ImageSpec{type, w, h}, std::uint64_t{w} * s, a.pixels.size()};
    EXPECT(cpu::compress(in, tile, MutableBytes{a.file.data(), a.file.size()}, written) == Status::success);
    a.file.resize(written);
    EXPECT(parse_metadata(ConstBytes{a.file.data(), a.file.size()}, ParseLimits{}, a.meta) == Status::success);
    a.device = device_alloc(a.file.size());
    std::memcpy(a.device, a.file.data(), a.file.size());
    return a;
}

ImageView device_view(PixelType type, std::uint32_t w, std::uint32_t h, std::uint64_t pad = 0) {
    const std::uint32_t s = type == PixelType::uint16 ? 2 : 4;
    const std::uint64_t pitch = std::uint64_t{w} * s + pad;
    return ImageView{device_alloc(pitch * h), ImageSpec{type, w, h}, pitch, pitch * h};
}

BatchItem decode_item(const Archive &a, Roi roi, bool calibrated, std::uint64_t pad) {
    BatchItem it;
    it.compressed_input.bytes = ConstBytes{a.device, a.file.size()};
    it.compressed_input.metadata = &a.meta;
    it.roi = roi;
    it.output = device_view(calibrated ? PixelType::float32 : a.type, roi.width, roi.height, pad);
    if (calibrated) {
        it.calibration.offset.scalar = 100.0F;
        it.calibration.gain.scalar = 1.05F;
    }
    return it;
}

struct Buffers {
    MutableBytes workspace;
    ResultBuffer results;
};

Buffers buffers_for(const BatchPlan &plan, std::uint64_t count) {
    std::uint64_t bytes = 0;
    EXPECT(workspace_size(plan.workspace_spec(), bytes) == Status::success);
    Buffers b;
    b.workspace = MutableBytes{bytes != 0 ? device_alloc(bytes) : nullptr, bytes};
    b.results = ResultBuffer{static_cast<DeviceResult *>(device_alloc(16 * count)), count};
    return b;
}

void plans_and_rejections(std::mt19937 &rng) {
    const TileShape shapes[3] = {{128, 32}, {64, 64}, {256, 16}};
    std::vector<Archive> archives;
    for (int i = 0; i < 12; ++i) {
        const PixelType type = i % 2 == 0 ? PixelType::uint16 : PixelType::float32;
        archives.push_back(make_archive(type, 1 + rng() % 700, 1 + rng() % 300, shapes[i % 3], rng));
    }
    for (PixelType type : {PixelType::uint16, PixelType::float32}) {
        for (Operation op : {Operation::decompress, Operation::decode_roi, Operation::decode_calibrate}) {
            std::vector<BatchItem> items;
            for (const Archive &a : archives) {
                if (a.type != type) continue;
                Roi roi{0, 0, a.w, a.h};
                if (op != Operation::decompress) {
                    roi.width = 1 + rng() % a.w;
                    roi.height = 1 + rng() % a.h;
                    roi.x = rng() % (a.w - roi.width + 1);
                    roi.y = rng() % (a.h - roi.height + 1);
                }
                items.push_back(decode_item(a, roi, op == Operation::decode_calibrate, rng() % 3 * 4));
            }
            std::unique_ptr<BatchPlan> plan;
            EXPECT(prepare_batch(BatchRequest{op, items.data(), items.size(), {}}, plan, nullptr) == Status::success);
            const Buffers b = buffers_for(*plan, items.size());
            const Status run = op == Operation::decompress   ? decompress_async(*plan, b.workspace, b.results, nullptr)
                               : op == Operation::decode_roi ? decode_roi_async(*plan, b.workspace, b.results, nullptr)
                                                             : decode_calibrate_async(*plan, b.workspace, b.results, nullptr);
            EXPECT(run == Status::success);
            // Rejections before any enqueue: workspace too small or misaligned,
            // results misaligned or overlapping an output, wrong operation.
            if (b.workspace.size >= 512) {
                EXPECT(decode_roi_async(*plan, MutableBytes{b.workspace.data, b.workspace.size - 1}, b.results, nullptr) ==
                       (op == Operation::decode_roi ? Status::buffer_too_small : Status::invalid_argument));
                auto *shifted = static_cast<std::uint8_t *>(b.workspace.data) + 16;
                EXPECT(compress_async(*plan, MutableBytes{shifted, b.workspace.size - 256}, b.results, nullptr) ==
                       Status::invalid_argument);
            }
            const ResultBuffer odd{reinterpret_cast<DeviceResult *>(reinterpret_cast<std::uint8_t *>(b.results.data) + 4),
                                   b.results.count - 1};
            EXPECT(calibrate_async(*plan, b.workspace, odd, nullptr) == Status::invalid_argument);
            const ResultBuffer over{static_cast<DeviceResult *>(items[0].output.data), items.size()};
            if (items[0].output.capacity_bytes >= 16 * items.size()) {
                const Status st = op == Operation::decompress ? decompress_async(*plan, b.workspace, over, nullptr)
                                  : op == Operation::decode_roi ? decode_roi_async(*plan, b.workspace, over, nullptr)
                                                                : decode_calibrate_async(*plan, b.workspace, over, nullptr);
                EXPECT(st == Status::invalid_argument);
            }
        }
    }
    // Staged payloads: the tiles of a region copied to the front of a buffer.
    for (const Archive &a : archives) {
        const Roi roi{a.w / 3, a.h / 3, a.w - a.w / 3, a.h - a.h / 3};
        const std::uint64_t across = (a.w + a.tile.width - 1) / a.tile.width;
        std::vector<TileMapping> maps;
        std::uint64_t at = 0;
        for (std::uint64_t r = roi.y / a.tile.height; r <= (roi.y + roi.height - 1) / a.tile.height; ++r)
            for (std::uint64_t c = roi.x / a.tile.width; c <= (roi.x + roi.width - 1) / a.tile.width; ++c) {
                const TileMetadata &t = a.meta.tiles[r * across + c];
                maps.push_back({r * across + c, at});
                at += t.length;
            }
        void *staged = device_alloc(at + 16);
        BatchItem it = decode_item(a, roi, false, 0);
        it.compressed_input.bytes = ConstBytes{staged, at};
        it.compressed_input.mappings = maps.data();
        it.compressed_input.mapping_count = maps.size();
        std::unique_ptr<BatchPlan> plan;
        EXPECT(prepare_batch(BatchRequest{Operation::decode_roi, &it, 1, {}}, plan, nullptr) == Status::success);
        // Two mappings at one offset overlap.
        if (maps.size() > 1) {
            maps[1].input_offset = maps[0].input_offset;
            std::unique_ptr<BatchPlan> bad;
            EXPECT(prepare_batch(BatchRequest{Operation::decode_roi, &it, 1, {}}, bad, nullptr) == Status::invalid_argument);
        }
    }
}

// Encoder output decoded without host metadata (the trusted path), in plans
// and in a queue next to parsed archives; the allocation-free encode.
void trusted_and_queue(std::mt19937 &rng) {
    const TileShape shapes[3] = {{128, 32}, {64, 64}, {256, 16}};
    for (int round = 0; round < 6; ++round) {
        const TileShape tile = shapes[round % 3];
        const PixelType type = round % 2 == 0 ? PixelType::uint16 : PixelType::float32;
        const std::uint32_t s = type == PixelType::uint16 ? 2 : 4;
        std::vector<BatchItem> enc;
        std::vector<std::uint64_t> bounds;
        for (int i = 0; i < 5; ++i) {
            const std::uint32_t w = 1 + rng() % 900, h = 1 + rng() % 400;
            std::uint64_t bound = 0;
            EXPECT(compress_bound(ImageSpec{type, w, h}, tile, bound) == Status::success);
            BatchItem it;
            const ImageView px = device_view(type, w, h, 8);
            it.input = ConstImageView{px.data, px.image, px.pitch_bytes, px.capacity_bytes};
            it.compressed_output = MutableBytes{device_alloc(bound), bound};
            enc.push_back(it);
            bounds.push_back(bound);
        }
        std::unique_ptr<BatchPlan> eplan;
        EXPECT(prepare_batch(BatchRequest{Operation::compress, enc.data(), enc.size(), tile}, eplan, nullptr) ==
               Status::success);
        const Buffers eb = buffers_for(*eplan, enc.size());
        EXPECT(compress_async(*eplan, eb.workspace, eb.results, nullptr) == Status::success);
        EXPECT(compress_async_static(*eplan, eb.results, nullptr) == Status::success);
        std::vector<BatchItem> dec;
        std::vector<Operation> ops;
        for (std::size_t i = 0; i < enc.size(); ++i) {
            BatchItem it;
            it.compressed_input.bytes = ConstBytes{enc[i].compressed_output.data, bounds[i]};
            it.compressed_input.encoder_result = eb.results.data + i;
            it.compressed_input.trusted_image = enc[i].input.image;
            it.compressed_input.trusted_tile = tile;
            const std::uint32_t w = enc[i].input.image.width, h = enc[i].input.image.height;
            it.roi = Roi{w / 4, h / 4, w - w / 4, h - h / 4};
            it.output = device_view(type, it.roi.width, it.roi.height);
            dec.push_back(it);
            ops.push_back(Operation::decode_roi);
        }
        std::unique_ptr<BatchPlan> dplan;
        EXPECT(prepare_batch(BatchRequest{Operation::decode_roi, dec.data(), dec.size(), {}}, dplan, nullptr) ==
               Status::success);
        const Buffers db = buffers_for(*dplan, dec.size());
        EXPECT(decode_roi_async(*dplan, db.workspace, db.results, nullptr) == Status::success);
        // The trusted items in a queue with parsed archives of the same shape.
        Archive parsed = make_archive(type, 300 + rng() % 200, 100 + rng() % 100, tile, rng);
        dec.push_back(decode_item(parsed, Roi{0, 0, parsed.w, parsed.h}, true, 4));
        ops.push_back(Operation::decode_calibrate);
        std::unique_ptr<DecodeQueue> queue;
        EXPECT(DecodeQueue::prepare(dec.data(), ops.data(), dec.size(), queue, nullptr) == Status::success);
        MutableBytes qws{device_alloc(256), 256};
        ResultBuffer qres{static_cast<DeviceResult *>(device_alloc(16 * dec.size())), dec.size()};
        // The block decoder needs host-known payloads: a queue with trusted
        // items refuses it and runs the tile decoder.
        queue->set_shape(LaunchShape{0, 512, 0, 2, 2, 1});
        EXPECT(queue->run(qws, qres, nullptr) == Status::invalid_argument);
        queue->set_shape(LaunchShape{4, 256, 0, 4, 3, 1});
        EXPECT(queue->run(qws, qres, nullptr) == Status::success);
        EXPECT(queue->capture(qws, qres, nullptr) == Status::success);
        EXPECT(queue->launch(nullptr) == Status::success);
        // A queue of parsed archives only, in the block decoder.
        std::vector<BatchItem> parsed_only(dec.end() - 1, dec.end());
        std::vector<Operation> parsed_ops(1, Operation::decode_calibrate);
        std::unique_ptr<DecodeQueue> q2;
        EXPECT(DecodeQueue::prepare(parsed_only.data(), parsed_ops.data(), 1, q2, nullptr) == Status::success);
        q2->set_shape(LaunchShape{0, 512, 0, 2, 2, 1});
        EXPECT(q2->run(qws, ResultBuffer{qres.data, 1}, nullptr) == Status::success);
        // Results in plain host memory are rejected before anything is enqueued.
        std::vector<DeviceResult> host_results(dec.size());
        EXPECT(queue->run(qws, ResultBuffer{host_results.data(), dec.size()}, nullptr) == Status::invalid_argument);
    }
}

// One field of parsed metadata changed (the mutation classes of reference
// tests/cuda_robustness_test.cpp:196-240): preparation rejects the result
// with a format status, or every table entry of the decode it plans is in
// bounds (the stand-in's launcher checks).
std::uint64_t nudge(std::uint64_t v, std::mt19937_64 &rng, std::uint64_t step = 1) {
    if (rng() % 3 == 0) return rng() % 3 == 0 ? rng() : rng() % 65536;
    return v + (rng() % 17 - 8) * step;
}

void fabricated(std::mt19937 &seed_rng) {
    std::mt19937_64 rng(seed_rng());
    const TileShape shapes[3] = {{128, 32}, {64, 64}, {256, 16}};
    std::uint64_t accepted = 0, malformed = 0;
    for (PixelType type : {PixelType::uint16, PixelType::float32}) {
        for (const TileShape shape : shapes) {
            // Constant, smooth and noisy bands: all three tile modes.
            const std::uint32_t s = type == PixelType::uint16 ? 2 : 4;
            std::vector<std::uint8_t> px(300 * 192 * s);
            for (std::size_t i = 0; i < px.size(); ++i) {
                const std::size_t row = i / s / 300;
                px[i] = static_cast<std::uint8_t>(row < 64 ? 7 : row < 128 ? (i / s) % 300 / 8 : rng());
            }
            std::uint64_t bound = 0;
            EXPECT(compress_bound(ImageSpec{type, 300, 192}, shape, bound) == Status::success);
            std::vector<std::uint8_t> file(bound);
            std::uint64_t written = 0;
            EXPECT(cpu::compress(ConstImageView{px.data(), ImageSpec{type, 300, 192}, 300ull * s, px.size()}, shape,
                                 MutableBytes{file.data(), file.size()}, written) == Status::success);
            file.resize(written);
            ImageMetadata parsed;
            EXPECT(parse_metadata(ConstBytes{file.data(), file.size()}, ParseLimits{}, parsed) == Status::success);
            void *src = device_alloc(file.size());
            std::memcpy(src, file.data(), file.size());
            void *out = device_alloc(300 * 192 * 4);
            auto *results = static_cast<DeviceResult *>(device_alloc(16));
            constexpr std::uint64_t kWorkspace = 1 << 20;
            void *workspace = device_alloc(kWorkspace);
            for (int trial = 0; trial < 500; ++trial) {
                ImageMetadata m = parsed;
                TileMetadata &t = m.tiles[rng() % m.tiles.size()];
                SegmentMetadata &g = m.segments.empty() ? m.segments.emplace_back() : m.segments[rng() % m.segments.size()];
                switch (rng() % 12) {
                case 0: t.mode = static_cast<TileMode>(rng() % 4); break;
                case 1: t.width = static_cast<std::uint16_t>(nudge(t.width, rng)); break;
                case 2: t.height = static_cast<std::uint16_t>(nudge(t.height, rng)); break;
                case 3: t.offset = nudge(t.offset, rng, 4); break;
                case 4: t.length = nudge(t.length, rng, 4); break;
                case 5: t.first_segment = nudge(t.first_segment, rng); break;
                case 6: g.seed = static_cast<std::uint32_t>(nudge(g.seed, rng)); break;
                case 7: g.bit_width = static_cast<std::uint32_t>(rng() % 40); break;
                case 8: g.payload_offset = nudge(g.payload_offset, rng, 4); break;
                case 9: g.payload_bytes = nudge(g.payload_bytes, rng, 4); break;
                case 10: m.file_bytes = nudge(m.file_bytes, rng, 4); break;
                default:
                    if (rng() % 2 == 0) m.tiles.pop_back();
                    else if (!m.segments.empty()) m.segments.pop_back();
                    break;
                }
                BatchItem it;
                it.compressed_input.bytes = ConstBytes{src, file.size()};
                it.compressed_input.metadata = &m;
                const auto x = static_cast<std::uint32_t>(rng() % 300), y = static_cast<std::uint32_t>(rng() % 192);
                it.roi = Roi{x, y, 1 + static_cast<std::uint32_t>(rng() % (300 - x)),
                             1 + static_cast<std::uint32_t>(rng() % (192 - y))};
                it.output = ImageView{out, ImageSpec{type, it.roi.width, it.roi.height}, std::uint64_t{it.roi.width} * s,
                                      300 * 192 * 4};
                std::unique_ptr<BatchPlan> plan;
                const Status st = prepare_batch(BatchRequest{Operation::decode_roi, &it, 1, {}}, plan, nullptr);
                if (st != Status::success) {
                    EXPECT(st == Status::malformed_data || st == Status::unsupported_format ||
                           st == Status::buffer_too_small);
                    malformed += st == Status::malformed_data;
                    continue;
                }
                ++accepted;
                std::uint64_t ws = 0;
                EXPECT(workspace_size(plan->workspace_spec(), ws) == Status::success && ws <= kWorkspace);
                EXPECT(decode_roi_async(*plan, MutableBytes{workspace, ws}, ResultBuffer{results, 1}, nullptr) ==
                       Status::success);
            }
        }
    }
    EXPECT(accepted > 0 && malformed > 0);
}

void serving(std::mt19937 &rng) {
    std::unique_ptr<ArchiveSet> set;
    EXPECT(ArchiveSet::create(set) == Status::success);
    std::vector<Archive> archives;
    for (int i = 0; i < 6; ++i) {
        archives.push_back(make_archive(i % 2 ? PixelType::float32 : PixelType::uint16, 200 + rng() % 600,
                                        200 + rng() % 400, TileShape{64, 64}, rng));
        std::uint32_t id = 0;
        EXPECT(set->add(ConstBytes{archives.back().device, archives.back().file.size()}, archives.back().meta, id) ==
               Status::success);
    }
    // The same archives again, tabled on the device (the stand-in launchers
    // check the table spans), before and after a commit.
    for (std::size_t i = 0; i < archives.size(); ++i) {
        const Archive &a = archives[i];
        std::uint32_t id = 0;
        EXPECT(set->add_device(ConstBytes{a.device, a.file.size()}, ImageSpec{a.type, a.w, a.h}, TileShape{64, 64},
                               nullptr, nullptr, id) == Status::success);
        if (i == 2) EXPECT(set->commit(nullptr) == Status::success);
    }
    // Rejected before any device work: host memory, a misaligned start.
    {
        std::uint32_t id = 0;
        const Archive &a = archives[0];
        EXPECT(set->add_device(ConstBytes{a.file.data(), a.file.size()}, ImageSpec{a.type, a.w, a.h},
                               TileShape{64, 64}, nullptr, nullptr, id) == Status::invalid_argument);
        EXPECT(set->add_device(ConstBytes{static_cast<const std::uint8_t *>(a.device) + 2, a.file.size() - 2},
                               ImageSpec{a.type, a.w, a.h}, TileShape{64, 64}, nullptr, nullptr, id) ==
               Status::invalid_argument);
    }
    EXPECT(set->commit(nullptr) == Status::success);
    ServerOptions options;
    options.streams = 2;
    options.encode_window = 64;
    std::unique_ptr<Server> server;
    EXPECT(Server::create(*set, options, server) == Status::success);
    std::vector<ServeRequest> requests;
    for (int i = 0; i < 40; ++i) {
        ServeRequest r;
        const Archive &a = archives[static_cast<std::size_t>(i) % archives.size()];
        if (i % 7 == 6) {
            r.kind = RequestKind::encode;
            r.tile = TileShape{128, 32};
            const ImageView px = device_view(a.type, a.w, a.h);
            r.input = ConstImageView{px.data, px.image, px.pitch_bytes, px.capacity_bytes};
            std::uint64_t bound = 0;
            EXPECT(compress_bound(px.image, r.tile, bound) == Status::success);
            r.compressed = MutableBytes{device_alloc(bound), bound};
        } else {
            r.archive = static_cast<std::uint32_t>(static_cast<std::size_t>(i) % archives.size());
            r.calibrated = i % 2;
            r.roi = Roi{static_cast<std::uint32_t>(rng() % (a.w / 2)), static_cast<std::uint32_t>(rng() % (a.h / 2)),
                        a.w / 2, a.h / 2};
            r.output = device_view(r.calibrated ? PixelType::float32 : a.type, r.roi.width, r.roi.height);
            r.offset = 100.0F;
            r.gain = 1.05F;
        }
        EXPECT(server->check(r) == Status::success);
        requests.push_back(r);
    }
    std::vector<DeviceResult> results(requests.size());
    EXPECT(server->run_batch(requests.data(), requests.size(), results.data()) == Status::success);
    std::vector<std::uint64_t> arrival(requests.size()), latency(requests.size());
    for (std::size_t i = 0; i < arrival.size(); ++i) arrival[i] = i * 1000;
    ServerStats stats;
    EXPECT(server->replay(requests.data(), arrival.data(), requests.size(), latency.data(), results.data(), stats) ==
           Status::success);
    // The same requests with every region batch decoded by blocks (records
    // built by the stand-in's block builder, which checks every table index).
    options.block_tasks = 1;
    std::unique_ptr<Server> by_blocks;
    EXPECT(Server::create(*set, options, by_blocks) == Status::success);
    EXPECT(by_blocks->run_batch(requests.data(), requests.size(), results.data()) == Status::success);
    EXPECT(by_blocks->replay(requests.data(), arrival.data(), requests.size(), latency.data(), results.data(),
                             stats) == Status::success);
    // A request past the archive's edge is rejected.
    ServeRequest bad = requests[0];
    bad.roi.x = archives[0].w;
    EXPECT(server->check(bad) != Status::success);
}

} // namespace

int main() {
    std::mt19937 rng(20261204);
    plans_and_rejections(rng);
    trusted_and_queue(rng);
    fabricated(rng);
    serving(rng);
    for (void *p : owned()) static_cast<void>(hipFree(p));
    std::printf("failures %d\nrc %d\n", failures, failures == 0 ? 0 : 1);
    return failures == 0 ? 0 : 1;
}
