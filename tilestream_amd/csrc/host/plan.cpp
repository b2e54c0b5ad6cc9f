// Batch preparation: host validation, compact device tables, upload.
//
// Validation order and the Status of every failure follow the reference
// planner (reference src/cuda/plan.cpp:128-465, 501-711) so that a caller sees
// the same result for the same request.  The device tables are this library's:
// one 32-byte task per tile and a 16-bit word offset per predictive row; the
// row seeds and widths are read from the staged payload itself.
#include "plan.h"

#include <tilestream/format.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <stdexcept>
#include <utility>

namespace tilestream::hip {
namespace detail {

using host::byte_range;

bool Footprint::admit(std::vector<Span> written, std::vector<Span> read) {
    // One sweep in order of start address.  A buffer overlaps an earlier one
    // exactly when it starts before that one ends, so it is enough to track
    // how far the written buffers seen so far reach, and how far all of them
    // reach: a written buffer may start at neither, a read one not before the
    // written reach.
    struct Tagged {
        Span span;
        bool written;
    };
    std::vector<Tagged> all;
    all.reserve(written.size() + read.size());
    for (const Span &s : written)
        if (!s.empty()) all.push_back({s, true});
    for (const Span &s : read)
        if (!s.empty()) all.push_back({s, false});
    std::sort(all.begin(), all.end(), [](const Tagged &a, const Tagged &b) { return a.span.lo < b.span.lo; });
    std::uintptr_t written_reach = 0, any_reach = 0;
    std::vector<Span> merged;
    for (const Tagged &t : all) {
        if (t.span.lo < (t.written ? any_reach : written_reach)) return false;
        any_reach = std::max(any_reach, t.span.hi);
        if (t.written) written_reach = std::max(written_reach, t.span.hi);
        if (!merged.empty() && t.span.lo <= merged.back().hi) merged.back().hi = std::max(merged.back().hi, t.span.hi);
        else merged.push_back(t.span);
    }
    union_ = std::move(merged);
    return true;
}

bool Footprint::avoids(Span s) const noexcept {
    if (s.empty()) return true;
    // The first piece of the union that ends after s starts is the only one
    // that can overlap it.
    const auto piece = std::partition_point(union_.begin(), union_.end(), [&](const Span &u) { return u.hi <= s.lo; });
    return piece == union_.end() || piece->lo >= s.hi;
}

WindowSplit split_into_windows(const std::vector<std::uint64_t> &tiles, std::uint32_t window) {
    WindowSplit out;
    std::uint64_t total = 0;
    for (std::uint64_t n : tiles) total += n;
    std::size_t image = 0;        // first image that may still have tiles at or after w0
    std::uint64_t image_begin = 0; // its first tile in the concatenation
    for (std::uint64_t w0 = 0; w0 < total; w0 += window) {
        const std::uint64_t w1 = std::min<std::uint64_t>(total, w0 + window);
        WindowSpan span;
        span.first_entry = out.entries.size();
        span.tile_count = static_cast<std::uint32_t>(w1 - w0);
        while (image < tiles.size() && image_begin + tiles[image] <= w0) image_begin += tiles[image++];
        std::uint64_t begin = image_begin;
        for (std::size_t i = image; i < tiles.size() && begin < w1; begin += tiles[i++]) {
            const std::uint64_t end = begin + tiles[i];
            const std::uint64_t lo = std::max(begin, w0), hi = std::min(end, w1);
            if (lo >= hi) continue; // an image with no tiles
            const std::uint32_t flags = (lo > begin ? 1u : 0u) | (hi == end ? 2u : 0u);
            out.entries.push_back({static_cast<std::uint32_t>(i), static_cast<std::uint32_t>(lo - w0),
                                   static_cast<std::uint32_t>(hi - lo), flags, lo - begin});
            ++span.entry_count;
        }
        out.windows.push_back(span);
    }
    return out;
}

PlanState::~PlanState() {
    // Callers finish every execution before destroying a plan, so the tables
    // can go back to the pool without waiting or ordering.
    if (storage == nullptr) return;
    static_cast<void>(hip_status(hipFreeAsync(storage, hipStreamPerThread)));
}

namespace {

// Where a pointer's memory lives, as far as the HIP runtime knows.
enum class Residence { this_device, other_device, managed, mapped_host, unmapped_host, not_registered };

Status residence_of(const void *data, int device, Residence &where) noexcept {
    hipPointerAttribute_t attr{};
    const hipError_t e = hipPointerGetAttributes(&attr, data);
    if (e == hipErrorInvalidValue) {
        // Ordinary host memory the runtime has never seen.
        static_cast<void>(hipGetLastError());
        where = Residence::not_registered;
        return Status::success;
    }
    if (e != hipSuccess) return hip_failure();
    if (attr.isManaged || attr.type == hipMemoryTypeManaged || attr.type == hipMemoryTypeUnified)
        where = Residence::managed;
    else if (attr.type == hipMemoryTypeDevice)
        where = attr.device == device ? Residence::this_device : Residence::other_device;
    else if (attr.type == hipMemoryTypeHost)
        where = attr.devicePointer == data ? Residence::mapped_host : Residence::unmapped_host;
    else
        where = Residence::not_registered;
    return Status::success;
}

} // namespace

// The plan's device may address the memory: its own device memory, managed
// memory, pinned host memory mapped at the same address, or any host memory
// when the device reports pageable memory access (docs/api.md "Alignment and
// capacity", reference interface plan.cpp:70-85).
Status accessible(const void *data, int device, bool pageable) noexcept {
    Residence where = Residence::not_registered;
    if (const Status st = residence_of(data, device, where); st != Status::success) return st;
    switch (where) {
    case Residence::this_device:
    case Residence::managed:
    case Residence::mapped_host:
        return Status::success;
    case Residence::not_registered:
        return pageable ? Status::success : Status::invalid_argument;
    default:
        return Status::invalid_argument;
    }
}

// Memory of this device itself (not host, managed or another device's).
Status on_device(const void *data, int device) noexcept {
    Residence where = Residence::not_registered;
    if (const Status st = residence_of(data, device, where); st != Status::success) return st;
    return where == Residence::this_device ? Status::success : Status::invalid_argument;
}

// A call runs on the plan's device: it must be the calling thread's current
// device and the stream's device.
Status stream_device(hipStream_t stream, int device) noexcept {
    int caller = -1;
    hipDevice_t owner = -1;
    const bool known = hipGetDevice(&caller) == hipSuccess && hipStreamGetDevice(stream, &owner) == hipSuccess;
    if (!known) return hip_failure();
    const bool same = caller == device && owner == device;
    return same ? Status::success : Status::invalid_argument;
}

namespace {

std::uint64_t columns(const ImageSpec &image, TileShape tile) noexcept { return host::ceil_div(image.width, tile.width); }
std::uint64_t tile_rows(const ImageSpec &image, TileShape tile) noexcept { return host::ceil_div(image.height, tile.height); }

// A map matches the output region exactly (reference plan.cpp:174-193).
Status check_parameter(const Parameter &p, const ImageSpec &region, std::vector<Span> &readable,
                       std::vector<const void *> &pointers) {
    if (p.kind == ParameterKind::scalar) return Status::success;
    if (p.kind != ParameterKind::map) return Status::invalid_argument;
    if (const Status st = host::check_view(p.map); st != Status::success) return st;
    if (p.map.image.type != PixelType::float32 || p.map.image.width != region.width ||
        p.map.image.height != region.height)
        return Status::invalid_argument;
    readable.push_back(Span::of(p.map.data, p.map.capacity_bytes));
    pointers.push_back(p.map.data);
    return Status::success;
}

Status check_calibration(const Calibration &c, const ImageSpec &region, std::vector<Span> &readable,
                         std::vector<const void *> &pointers) {
    const Status st = check_parameter(c.offset, region, readable, pointers);
    return st != Status::success ? st : check_parameter(c.gain, region, readable, pointers);
}

// What docs/format.md fixes for the index entry of tile `index`: the actual
// size its grid position implies, its RAW length A4(w h s), which bounds
// every payload, and the widest residual of the pixel type.
struct TileFacts {
    std::uint64_t width = 0, height = 0;
    std::uint64_t raw_length = 0;
    std::uint32_t widest = 0;
    bool u16 = true;
};

TileFacts facts_of(const ImageMetadata &m, std::uint64_t index) noexcept {
    TileFacts f;
    const std::uint64_t across = columns(m.image, m.tile);
    const std::uint64_t x = index % across * m.tile.width, y = index / across * m.tile.height;
    f.width = std::min<std::uint64_t>(m.tile.width, m.image.width - x);
    f.height = std::min<std::uint64_t>(m.tile.height, m.image.height - y);
    f.u16 = m.image.type == PixelType::uint16;
    f.raw_length = (f.width * f.height * (f.u16 ? 2 : 4) + 3) & ~std::uint64_t{3};
    f.widest = f.u16 ? 16u : 32u;
    return f;
}

} // namespace

// The parser's structural rules that keep device reads inside a tile's
// payload, checked again for each tile a plan selects (reference interface
// plan.cpp:238-279): every violation is malformed_data.
Status check_tile(const ImageMetadata &m, std::uint64_t index) noexcept {
    const TileMetadata &t = m.tiles[static_cast<std::size_t>(index)];
    const TileFacts f = facts_of(m, index);
    // The entry's size agrees with its position; the payload is word aligned
    // and inside the file.
    const bool placed = t.width == f.width && t.height == f.height && t.offset % 4 == 0 && t.offset <= m.file_bytes &&
                        t.length <= m.file_bytes - t.offset;
    if (!placed) return Status::malformed_data;
    // RAW and CONSTANT payloads have one exact length each.
    if (t.mode == TileMode::raw || t.mode == TileMode::constant) {
        const std::uint64_t exact = t.mode == TileMode::raw ? f.raw_length : 4;
        return t.length == exact ? Status::success : Status::malformed_data;
    }
    if (t.mode != TileMode::predict_bitpack || t.length > f.raw_length) return Status::malformed_data;
    // A predictive tile owns `height` consecutive row records, each inside the
    // payload with exactly the packed length its width gives.
    const std::size_t segments = m.segments.size();
    if (t.first_segment > segments || f.height > segments - t.first_segment) return Status::malformed_data;
    const auto row_fits = [&](const SegmentMetadata &g) {
        const std::uint64_t packed = ((f.width - 1) * g.bit_width + 31) / 32 * 4;
        const bool header = g.payload_offset >= format::segment_header_bytes && g.payload_offset % 4 == 0;
        const bool inside = g.payload_offset <= t.length && g.payload_bytes <= t.length - g.payload_offset;
        return g.bit_width <= f.widest && (!f.u16 || g.seed <= 0xffffu) && header && inside && g.payload_bytes == packed;
    };
    const auto first = m.segments.begin() + static_cast<std::ptrdiff_t>(t.first_segment);
    return std::all_of(first, first + static_cast<std::ptrdiff_t>(f.height), row_fits) ? Status::success
                                                                                        : Status::malformed_data;
}

namespace {

// The tiles a region intersects: columns [left, right], rows [top, bottom].
struct TileWindow {
    std::uint64_t left = 0, right = 0, top = 0, bottom = 0;
    [[nodiscard]] std::uint64_t count() const noexcept { return (right - left + 1) * (bottom - top + 1); }
};

TileWindow window_of(const Roi &roi, TileShape tile) noexcept {
    TileWindow w;
    w.left = roi.x / tile.width;
    w.right = (std::uint64_t{roi.x} + roi.width - 1) / tile.width;
    w.top = roi.y / tile.height;
    w.bottom = (std::uint64_t{roi.y} + roi.height - 1) / tile.height;
    return w;
}

// Selects the tiles intersecting the region, in tile order, each at its
// file offset or, for staged payloads, at the offset its mapping gives
// (docs/api.md "Selected-tile uploads": exactly the intersecting tiles in
// increasing order, 4-byte aligned, inside the supplied bytes, not
// overlapping one another).
Status stage_tiles(const CompressedInput &source, const Roi &roi, std::uint32_t image, DecodeDraft &d) {
    const ImageMetadata &m = *source.metadata;
    const std::uint64_t across = columns(m.image, m.tile);
    if (m.tiles.size() != across * tile_rows(m.image, m.tile)) return Status::malformed_data;
    const TileWindow win = window_of(roi, m.tile);
    const bool mapped = source.mappings != nullptr || source.mapping_count != 0;
    if (mapped && (source.mappings == nullptr || source.mapping_count != win.count())) return Status::invalid_argument;
    if (!mapped && m.file_bytes > source.bytes.size) return Status::buffer_too_small;
    const auto mapping_fits = [&](const TileMapping &map, std::uint64_t index, std::uint64_t length) {
        const std::uint64_t size = source.bytes.size;
        return map.tile_index == index && map.input_offset % 4 == 0 && map.input_offset <= size &&
               length <= size - map.input_offset;
    };
    std::vector<Span> staged;
    const TileMapping *next_mapping = source.mappings;
    for (std::uint64_t row = win.top; row <= win.bottom; ++row) {
        for (std::uint64_t col = win.left; col <= win.right; ++col) {
            const std::uint64_t index = row * across + col;
            if (const Status st = check_tile(m, index); st != Status::success) return st;
            const TileMetadata &t = m.tiles[static_cast<std::size_t>(index)];
            std::uint64_t offset = t.offset;
            if (mapped) {
                const TileMapping &map = *next_mapping++;
                if (!mapping_fits(map, index, t.length)) return Status::invalid_argument;
                offset = map.input_offset;
                staged.push_back({static_cast<std::uintptr_t>(offset), static_cast<std::uintptr_t>(offset + t.length)});
            }
            tsk_decode_task task{};
            task.payload = static_cast<const std::uint8_t *>(source.bytes.data) + offset;
            task.column = static_cast<std::uint32_t>(col);
            task.row = static_cast<std::uint32_t>(row);
            task.image = image;
            task.rows = TSK_NO_ROWS;
            task.words = static_cast<std::uint32_t>(t.length / 4);
            task.shape = TSK_TASK_SHAPE(static_cast<std::uint32_t>(t.mode), t.height, t.width);
            if (t.mode == TileMode::predict_bitpack) {
                task.rows = static_cast<std::uint32_t>(d.rows.size());
                for (std::uint64_t r = 0; r < t.height; ++r) {
                    const SegmentMetadata &g = m.segments[static_cast<std::size_t>(t.first_segment + r)];
                    // Word offset of the row descriptor within the tile payload.
                    d.rows.push_back(static_cast<std::uint16_t>((g.payload_offset - format::segment_header_bytes) / 4));
                }
            }
            d.tasks.push_back(task);
        }
    }
    // Staged payloads are all written by the caller's copies: none may overlap another.
    Footprint copies;
    if (!copies.admit(std::move(staged), {})) return Status::invalid_argument;
    d.tiles_per_image.push_back(win.count());
    return Status::success;
}

} // namespace

// One decode item (reference plan.cpp:351-465).
Status draft_decode_item(const BatchItem &item, Operation operation, bool require_one_type, DecodeDraft &d) {
    const CompressedInput &source = item.compressed_input;
    const bool trusted = source.encoder_result != nullptr;
    if (trusted ? source.metadata != nullptr || source.mappings != nullptr || source.mapping_count != 0
                : source.metadata == nullptr)
        return Status::invalid_argument;
    const ImageSpec image = trusted ? source.trusted_image : source.metadata->image;
    const TileShape shape = trusted ? source.trusted_tile : source.metadata->tile;
    std::uint64_t bound = 0;
    if (const Status st = compress_bound(image, shape, bound); st != Status::success) return st;
    if (d.first) {
        d.type = image.type;
        d.shape = shape;
        d.first = false;
    }
    if (require_one_type && image.type != d.type) return Status::invalid_argument;
    d.mixed_shapes = d.mixed_shapes || shape.width != d.shape.width || shape.height != d.shape.height;

    const ImageView &output = item.output;
    if (const Status st = host::check_view(host::as_const(output)); st != Status::success) return st;
    Roi roi{0, 0, image.width, image.height};
    if (operation != Operation::decompress) {
        roi = item.roi;
        if (roi.width == 0 || roi.height == 0 || std::uint64_t{roi.x} + roi.width > image.width ||
            std::uint64_t{roi.y} + roi.height > image.height)
            return Status::invalid_roi;
    }
    const bool calibrated = operation == Operation::decode_calibrate;
    if (output.image.type != (calibrated ? PixelType::float32 : image.type) || output.image.width != roi.width ||
        output.image.height != roi.height)
        return Status::invalid_argument;
    if (calibrated)
        if (const Status st = check_calibration(item.calibration, output.image, d.readable, d.pointers);
            st != Status::success)
            return st;
    if (const Status st = byte_range(source.bytes.data, source.bytes.size); st != Status::success) return st;
    if (source.bytes.data == nullptr || reinterpret_cast<std::uintptr_t>(source.bytes.data) % 4 != 0)
        return Status::invalid_argument;

    const auto index = static_cast<std::uint32_t>(d.images.size());
    if (trusted) {
        // Modes and offsets are known only on the device, after the encoder ran.
        if (reinterpret_cast<std::uintptr_t>(source.encoder_result) % alignof(DeviceResult) != 0)
            return Status::invalid_argument;
        if (source.bytes.size < bound) return Status::buffer_too_small;
        std::uint64_t n = 0;
        for (std::uint64_t row = roi.y / shape.height; row <= (std::uint64_t{roi.y} + roi.height - 1) / shape.height; ++row)
            for (std::uint64_t col = roi.x / shape.width; col <= (std::uint64_t{roi.x} + roi.width - 1) / shape.width;
                 ++col, ++n) {
                tsk_decode_task task{};
                task.payload = nullptr; // resolved on the device from the index
                task.column = static_cast<std::uint32_t>(col);
                task.row = static_cast<std::uint32_t>(row);
                task.image = index;
                task.rows = TSK_NO_ROWS;
                const std::uint64_t tw = std::min<std::uint64_t>(shape.width, image.width - col * shape.width);
                const std::uint64_t th = std::min<std::uint64_t>(shape.height, image.height - row * shape.height);
                task.shape = TSK_TASK_SHAPE(0u, th, tw);
                d.tasks.push_back(task);
            }
        d.tiles_per_image.push_back(n);
        d.readable.push_back(Span::of(source.encoder_result, sizeof(DeviceResult)));
        d.pointers.push_back(source.encoder_result);
    } else {
        if (const Status st = stage_tiles(source, roi, index, d); st != Status::success) return st;
    }

    tsk_decode_image rec{};
    rec.src = static_cast<const std::uint8_t *>(source.bytes.data);
    rec.out = static_cast<std::uint8_t *>(output.data);
    rec.producer = source.encoder_result;
    rec.src_size = source.bytes.size;
    rec.out_pitch = output.pitch_bytes;
    rec.width = image.width;
    rec.height = image.height;
    rec.roi_x = roi.x;
    rec.roi_y = roi.y;
    rec.roi_w = roi.width;
    rec.roi_h = roi.height;
    rec.tile_columns = static_cast<std::uint32_t>(columns(image, shape));
    rec.tile_w = shape.width;
    rec.tile_h = shape.height;
    rec.type = static_cast<std::uint8_t>(image.type);
    rec.calibrated = calibrated ? 1 : 0;
    rec.offset = 0.0F;
    rec.gain = 1.0F;
    if (calibrated) {
        const Parameter &o = item.calibration.offset, &g = item.calibration.gain;
        rec.offset = o.kind == ParameterKind::scalar ? o.scalar : 0.0F;
        rec.gain = g.kind == ParameterKind::scalar ? g.scalar : 1.0F;
        if (o.kind == ParameterKind::map) {
            rec.offset_map = static_cast<const std::uint8_t *>(o.map.data);
            rec.offset_pitch = o.map.pitch_bytes;
        }
        if (g.kind == ParameterKind::map) {
            rec.gain_map = static_cast<const std::uint8_t *>(g.map.data);
            rec.gain_pitch = g.map.pitch_bytes;
        }
    }
    d.images.push_back(rec);
    d.readable.push_back(Span::of(source.bytes.data, source.bytes.size));
    d.writable.push_back(Span::of(output.data, output.capacity_bytes));
    d.pointers.push_back(source.bytes.data);
    d.pointers.push_back(output.data);
    return Status::success;
}

bool plan_blocks(const DecodeDraft &d, std::vector<tsk_block> &blocks) {
    std::uint64_t task = 0;
    for (std::size_t i = 0; i < d.images.size(); ++i) {
        const tsk_decode_image &img = d.images[i];
        const std::uint64_t count = d.tiles_per_image[i];
        if (count == 0) continue;
        // The region's tile grid, row-major in the draft.
        const std::uint32_t left = img.roi_x / img.tile_w, top = img.roi_y / img.tile_h;
        const std::uint32_t right = static_cast<std::uint32_t>((std::uint64_t{img.roi_x} + img.roi_w - 1) / img.tile_w);
        const std::uint32_t cols = right - left + 1;
        const std::uint64_t rows = count / cols;
        const std::uint32_t out_sample = img.calibrated ? 4u : (img.type == TSK_UINT16 ? 2u : 4u);
        const std::uint32_t kw = std::max(1u, 1024u / (std::uint32_t{img.tile_w} * out_sample));
        const std::uint32_t kh = std::max(1u, 64u / std::uint32_t{img.tile_h});
        if (kw * kh > 8) return false; // the allowed geometries give at most 8 tiles
        for (std::uint64_t br = 0; br < rows; br += kh) {
            for (std::uint32_t bc = 0; bc < cols; bc += kw) {
                tsk_block b{};
                b.image = static_cast<std::uint32_t>(i);
                b.across = static_cast<std::uint16_t>(std::min(kw, cols - bc));
                b.down = static_cast<std::uint16_t>(std::min<std::uint64_t>(kh, rows - br));
                b.col0 = left + bc;
                b.row0 = static_cast<std::uint32_t>(top + br);
                b.count = std::uint32_t{b.across} * b.down;
                std::uint32_t staged = 0;
                bool full = b.across == kw && b.down == kh;
                for (std::uint32_t kr = 0; kr < b.down; ++kr) {
                    for (std::uint32_t kc = 0; kc < b.across; ++kc) {
                        const tsk_decode_task &t = d.tasks[task + (br + kr) * cols + bc + kc];
                        if (t.payload == nullptr) return false;
                        const auto p = reinterpret_cast<std::uintptr_t>(t.payload);
                        tsk_block_tile &bt = b.tiles[kr * b.across + kc];
                        bt.payload = t.payload;
                        bt.words = t.words;
                        bt.shape = t.shape;
                        bt.rows = t.rows;
                        bt.stage = staged;
                        staged += static_cast<std::uint32_t>(((p + std::uint64_t{t.words} * 4 + 15) >> 4) - (p >> 4));
                        full = full && t.shape == TSK_TASK_SHAPE(2u, img.tile_h, img.tile_w) && t.rows != TSK_NO_ROWS;
                    }
                }
                if (staged > TSK_BLOCK_STAGE_USED) return false;
                b.blocks = staged;
                b.flags = full ? TSK_BLOCK_FULL : 0u;
                b.image_copy = img;
                blocks.push_back(b);
            }
        }
        task += count;
    }
    return true;
}

namespace {

struct Draft {
    PixelType type = PixelType::uint16;
    TileShape shape;
    DecodeDraft decode;
    std::vector<tsk_encode_image> encode;
    std::vector<tsk_calibrate_image> calibrate;
    std::vector<std::uint64_t> first_rows;
    std::uint64_t total_rows = 0;
    std::vector<std::uint64_t> tiles; // tiles per image, for windows
    std::vector<Span> writable;
    std::vector<Span> readable;
    std::vector<const void *> pointers;
};

Status draft_compress(const BatchRequest &request, Draft &d) {
    d.type = request.items[0].input.image.type;
    d.shape = request.tile;
    for (std::uint64_t i = 0; i < request.count; ++i) {
        const BatchItem &item = request.items[i];
        if (const Status st = host::check_view(item.input); st != Status::success) return st;
        if (item.input.image.type != d.type) return Status::invalid_argument;
        std::uint64_t bound = 0;
        if (const Status st = compress_bound(item.input.image, request.tile, bound); st != Status::success) return st;
        const MutableBytes &out = item.compressed_output;
        if (const Status st = byte_range(out.data, out.size); st != Status::success) return st;
        if (out.data == nullptr || reinterpret_cast<std::uintptr_t>(out.data) % 4 != 0) return Status::invalid_argument;
        if (out.size < bound) return Status::buffer_too_small;
        tsk_encode_image rec{};
        rec.in = static_cast<const std::uint8_t *>(item.input.data);
        rec.out = static_cast<std::uint8_t *>(out.data);
        rec.in_pitch = item.input.pitch_bytes;
        rec.tile_columns = static_cast<std::uint32_t>(columns(item.input.image, request.tile));
        rec.tile_count = columns(item.input.image, request.tile) * tile_rows(item.input.image, request.tile);
        rec.width = item.input.image.width;
        rec.height = item.input.image.height;
        rec.tile_w = request.tile.width;
        rec.tile_h = request.tile.height;
        rec.type = static_cast<std::uint32_t>(item.input.image.type);
        d.encode.push_back(rec);
        d.tiles.push_back(rec.tile_count);
        d.readable.push_back(Span::of(item.input.data, item.input.capacity_bytes));
        d.writable.push_back(Span::of(out.data, out.size));
        d.pointers.push_back(item.input.data);
        d.pointers.push_back(out.data);
    }
    return Status::success;
}

Status draft_calibrate(const BatchRequest &request, Draft &d) {
    d.type = request.items[0].input.image.type;
    for (std::uint64_t i = 0; i < request.count; ++i) {
        const BatchItem &item = request.items[i];
        Status st = host::check_view(item.input);
        if (st == Status::success) st = host::check_view(host::as_const(item.output));
        if (st != Status::success) return st;
        if (item.input.image.type != d.type || item.output.image.type != PixelType::float32 ||
            item.output.image.width != item.input.image.width || item.output.image.height != item.input.image.height)
            return Status::invalid_argument;
        if ((st = check_calibration(item.calibration, item.output.image, d.readable, d.pointers)) != Status::success)
            return st;
        tsk_calibrate_image rec{};
        rec.in = static_cast<const std::uint8_t *>(item.input.data);
        rec.out = static_cast<std::uint8_t *>(item.output.data);
        rec.in_pitch = item.input.pitch_bytes;
        rec.out_pitch = item.output.pitch_bytes;
        rec.width = item.input.image.width;
        rec.height = item.input.image.height;
        const Parameter &o = item.calibration.offset, &g = item.calibration.gain;
        rec.offset = o.kind == ParameterKind::scalar ? o.scalar : 0.0F;
        rec.gain = g.kind == ParameterKind::scalar ? g.scalar : 1.0F;
        if (o.kind == ParameterKind::map) {
            rec.offset_map = static_cast<const std::uint8_t *>(o.map.data);
            rec.offset_pitch = o.map.pitch_bytes;
        }
        if (g.kind == ParameterKind::map) {
            rec.gain_map = static_cast<const std::uint8_t *>(g.map.data);
            rec.gain_pitch = g.map.pitch_bytes;
        }
        d.calibrate.push_back(rec);
        d.first_rows.push_back(d.total_rows);
        d.total_rows += item.input.image.height;
        d.readable.push_back(Span::of(item.input.data, item.input.capacity_bytes));
        d.writable.push_back(Span::of(item.output.data, item.output.capacity_bytes));
        d.pointers.push_back(item.input.data);
        d.pointers.push_back(item.output.data);
    }
    return Status::success;
}

Status draft_decode(const BatchRequest &request, Draft &d) {
    for (std::uint64_t i = 0; i < request.count; ++i)
        if (const Status st = draft_decode_item(request.items[i], request.operation, true, d.decode); st != Status::success)
            return st;
    d.type = d.decode.type;
    d.shape = d.decode.mixed_shapes ? TileShape{64, 64} : d.decode.shape;
    d.tiles = d.decode.tiles_per_image;
    auto move_into = [](auto &from, auto &to) { to.insert(to.end(), from.begin(), from.end()); };
    move_into(d.decode.writable, d.writable);
    move_into(d.decode.readable, d.readable);
    move_into(d.decode.pointers, d.pointers);
    return Status::success;
}

template <typename T> bool upload(void *storage, std::uint64_t at, const std::vector<T> &v, hipStream_t stream) {
    return v.empty() || hipMemcpyAsync(static_cast<std::uint8_t *>(storage) + at, v.data(), v.size() * sizeof(T),
                                       hipMemcpyHostToDevice, stream) == hipSuccess;
}

Status build(const BatchRequest &request, PlanState &plan, hipStream_t stream) {
    plan.operation = request.operation;
    plan.count = request.count;
    if (request.count == 0) return Status::success;

    Draft d;
    Status st = request.operation == Operation::compress    ? draft_compress(request, d)
                : request.operation == Operation::calibrate ? draft_calibrate(request, d)
                                                            : draft_decode(request, d);
    if (st != Status::success) return st;
    plan.type = d.type;
    if (!plan.footprint.admit(std::move(d.writable), std::move(d.readable))) return Status::invalid_argument;

    int pageable = 0;
    if (hipGetDevice(&plan.device) != hipSuccess ||
        hipDeviceGetAttribute(&pageable, hipDeviceAttributePageableMemoryAccess, plan.device) != hipSuccess)
        return hip_failure();
    plan.pageable = pageable != 0;
    if ((st = stream_device(stream, plan.device)) != Status::success) return st;
    for (const void *p : d.pointers)
        if ((st = accessible(p, plan.device, plan.pageable)) != Status::success) return st;

    WindowSplit windows = split_into_windows(d.tiles, window_tiles);
    // One window over the whole batch when its look-back buffers (2 x 8 bytes
    // per tile) fit the residual region the spec reserves (16 KiB per window tile).
    std::uint64_t total_tiles = 0;
    for (std::uint64_t n : d.tiles) total_tiles += n;
    const std::size_t whole_at = windows.entries.size();
    const bool whole = request.operation == Operation::compress && windows.windows.size() > 1 &&
                       total_tiles <= std::uint64_t{INT32_MAX} &&
                       2 * 8 * total_tiles <= std::uint64_t{window_tiles} * 4096 * 4;
    if (whole) {
        std::uint64_t first = 0;
        for (std::size_t image = 0; image < d.tiles.size(); ++image) {
            windows.entries.push_back({static_cast<std::uint32_t>(image), static_cast<std::uint32_t>(first),
                                       static_cast<std::uint32_t>(d.tiles[image]), 2u, 0u});
            first += d.tiles[image];
        }
    }
    // Tile-to-entry tables: one per window, then the whole-batch window's.
    std::vector<std::uint32_t> blocks;
    std::vector<std::size_t> blocks_first;
    for (const WindowSpan &w : windows.windows) {
        blocks_first.push_back(blocks.size());
        append_block_owners(windows.entries.data() + w.first_entry, w.entry_count, w.tile_count, blocks);
    }
    const std::size_t whole_blocks_at = blocks.size();
    if (whole)
        append_block_owners(windows.entries.data() + whole_at, d.tiles.size(), static_cast<std::uint32_t>(total_tiles),
                            blocks);
    plan.spec = {0, 0, d.shape, 0};
    for (const WindowSpan &w : windows.windows) {
        plan.spec.image_count = std::max<std::uint64_t>(plan.spec.image_count, w.entry_count);
        plan.spec.tile_count = std::max<std::uint64_t>(plan.spec.tile_count, w.tile_count);
    }
    if ((st = host::workspace_layout(plan.spec, plan.layout)) != Status::success) return st;

    std::uint64_t total = 0;
    const auto place = [&total](std::uint64_t count, std::uint64_t size, std::uint64_t &at) {
        std::uint64_t bytes = 0;
        return host::round_up(total, 16, at) && host::mul(count, size, bytes) && host::add(at, bytes, total);
    };
    std::uint64_t dec_images_at = 0, tasks_at = 0, rows_at = 0, enc_images_at = 0, entries_at = 0, cal_at = 0, first_at = 0,
                  blocks_at = 0;
    if (!place(d.decode.images.size(), sizeof(tsk_decode_image), dec_images_at) ||
        !place(d.decode.tasks.size(), sizeof(tsk_decode_task), tasks_at) ||
        !place(d.decode.rows.size(), sizeof(std::uint16_t), rows_at) ||
        !place(d.encode.size(), sizeof(tsk_encode_image), enc_images_at) ||
        !place(windows.entries.size(), sizeof(tsk_window_entry), entries_at) ||
        !place(blocks.size(), sizeof(std::uint32_t), blocks_at) ||
        !place(d.calibrate.size(), sizeof(tsk_calibrate_image), cal_at) ||
        !place(d.first_rows.size(), sizeof(std::uint64_t), first_at))
        return Status::size_overflow;
    if (hipMallocAsync(&plan.storage, static_cast<std::size_t>(std::max<std::uint64_t>(total, 16)), stream) != hipSuccess) {
        plan.storage = nullptr;
        return hip_failure();
    }
    auto *base = static_cast<std::uint8_t *>(plan.storage);
    plan.decode_images = reinterpret_cast<const tsk_decode_image *>(base + dec_images_at);
    plan.tasks = reinterpret_cast<const tsk_decode_task *>(base + tasks_at);
    plan.rows = reinterpret_cast<const std::uint16_t *>(base + rows_at);
    plan.task_count = d.decode.tasks.size();
    // Large plans launch 1024 workgroups that stride over the tiles with the
    // next payload in flight while one decodes; one workgroup per tile exposes
    // each workgroup's load latency there.  Measured on single uint16 images
    // with 64x64 tiles: fused calibration from 1536 tiles with 4 columns per
    // lane (4096 tiles: 47 vs 60 us), plain decodes from 2048 tiles with 8
    // (2048 / 4096 / 8192 tiles: 13 / 20 / 35 vs 15 / 21 / 47 us).  Smaller
    // plans keep one workgroup per tile.
    if (request.operation == Operation::decode_calibrate) {
        if (plan.task_count >= 1536) plan.shape = {4, 256, 1024, 3};
    } else if (plan.task_count >= 2048) {
        plan.shape = {8, 256, 1024, 3};
    }
    plan.encode_images = reinterpret_cast<const tsk_encode_image *>(base + enc_images_at);
    plan.calibrate_images = reinterpret_cast<const tsk_calibrate_image *>(base + cal_at);
    plan.first_rows = reinterpret_cast<const std::uint64_t *>(base + first_at);
    plan.total_rows = d.total_rows;
    for (std::size_t k = 0; k < windows.windows.size(); ++k) {
        const WindowSpan &w = windows.windows[k];
        EncodeWindow placed;
        placed.entries = reinterpret_cast<const tsk_window_entry *>(base + entries_at) + w.first_entry;
        placed.blocks = reinterpret_cast<const std::uint32_t *>(base + blocks_at) + blocks_first[k];
        placed.entry_count = w.entry_count;
        placed.tile_count = w.tile_count;
        plan.windows.push_back(placed);
    }
    if (whole) {
        plan.whole.entries = reinterpret_cast<const tsk_window_entry *>(base + entries_at) + whole_at;
        plan.whole.blocks = reinterpret_cast<const std::uint32_t *>(base + blocks_at) + whole_blocks_at;
        plan.whole.entry_count = static_cast<std::uint32_t>(d.tiles.size());
        plan.whole.tile_count = static_cast<std::uint32_t>(total_tiles);
    }
    const std::uint32_t op = request.operation == Operation::compress    ? 0u
                             : request.operation == Operation::calibrate ? 2u
                             : request.operation == Operation::decode_calibrate ? 3u
                                                                                 : 1u;
    if (tsk_preload(op) != hipSuccess) return hip_failure();
    if (!upload(plan.storage, dec_images_at, d.decode.images, stream) ||
        !upload(plan.storage, tasks_at, d.decode.tasks, stream) || !upload(plan.storage, rows_at, d.decode.rows, stream) ||
        !upload(plan.storage, enc_images_at, d.encode, stream) ||
        !upload(plan.storage, entries_at, windows.entries, stream) ||
        !upload(plan.storage, blocks_at, blocks, stream) ||
        !upload(plan.storage, cal_at, d.calibrate, stream) || !upload(plan.storage, first_at, d.first_rows, stream) ||
        hipStreamSynchronize(stream) != hipSuccess)
        return hip_failure();
    plan.host_images = std::move(d.decode.images);
    plan.host_tasks = std::move(d.decode.tasks);
    plan.host_rows = std::move(d.decode.rows);
    return Status::success;
}

} // namespace

namespace {

// Evaluates the checks in order and returns the first failure.
template <typename... Checks> Status first_failure(Checks &&...checks) {
    Status st = Status::success;
    static_cast<void>(((st = checks(), st == Status::success) && ...));
    return st;
}

} // namespace

Status check_call(const CallNeeds &needs, MutableBytes workspace, ResultBuffer results, hipStream_t stream) noexcept {
    std::uint64_t result_bytes = 0;
    const Span scratch = Span::of(workspace.data, workspace.size);
    return first_failure(
        // Both byte ranges are representable, and non-null when non-empty.
        [&] { return host::mul(results.count, sizeof(DeviceResult), result_bytes) ? Status::success : Status::size_overflow; },
        [&] { return byte_range(workspace.data, workspace.size); },
        [&] { return byte_range(results.data, result_bytes); },
        // Room for the plan's workspace and one record per image.
        [&] {
            const bool room = workspace.size >= needs.workspace && results.count >= needs.results;
            return room ? Status::success : Status::buffer_too_small;
        },
        // Workspace on a 256-byte boundary, records naturally aligned.
        [&] {
            const bool aligned = reinterpret_cast<std::uintptr_t>(workspace.data) % workspace_alignment == 0 &&
                                 reinterpret_cast<std::uintptr_t>(results.data) % alignof(DeviceResult) == 0;
            return aligned ? Status::success : Status::invalid_argument;
        },
        // Both are written: they overlap no buffer of the batch and not each other.
        [&] {
            const Span records = Span::of(results.data, result_bytes);
            const bool apart = needs.footprint->avoids(scratch) && needs.footprint->avoids(records) &&
                               !host::overlaps(workspace.data, workspace.size, results.data, result_bytes);
            return apart ? Status::success : Status::invalid_argument;
        },
        [&] { return stream_device(stream, needs.device); },
        // An empty workspace is never touched, so its pointer need not be device memory.
        [&] { return workspace.size == 0 ? Status::success : accessible(workspace.data, needs.device, needs.pageable); },
        [&] { return accessible(results.data, needs.device, needs.pageable); },
        // Launch failures surface only as the runtime's last error: one left by
        // the caller's earlier work would be reported as this call's, so it is
        // left for the caller and nothing is enqueued.
        [] { return hipPeekAtLastError() == hipSuccess ? Status::success : Status::hip_error; });
}

// Host checks of every execution call, before anything is enqueued
// (reference interface plan.cpp:627-674, docs/api.md "Status and completion").
Status check_execution(const PlanState *plan, Operation operation, MutableBytes workspace, ResultBuffer results,
                       hipStream_t stream) noexcept {
    if (plan == nullptr || plan->operation != operation) return Status::invalid_argument;
    if (plan->count == 0) return Status::success; // an empty batch touches no buffer
    const CallNeeds needs{&plan->footprint, plan->device, plan->pageable, plan->count, plan->layout.total};
    return check_call(needs, workspace, results, stream);
}

Status run_decode(const PlanState &plan, DeviceResult *results, hipStream_t stream, const DecodeShape &shape) noexcept {
    const auto count = static_cast<std::uint32_t>(plan.count);
    if (tsk_results_begin(results, count, nullptr, 0, stream) != hipSuccess) return hip_failure();
    tsk_decode_args a{};
    a.images = plan.decode_images;
    a.tasks = plan.tasks;
    a.rows = plan.rows;
    a.results = results;
    a.task_count = plan.task_count;
    a.row_count = plan.host_rows.size();
    a.image_count = count;
    const bool calibrated = plan.operation == Operation::decode_calibrate;
    a.kind = plan.type == PixelType::uint16 ? (calibrated ? TSK_KIND_U16_CAL : TSK_KIND_U16)
                                            : (calibrated ? TSK_KIND_F32_CAL : TSK_KIND_F32);
    a.group = shape.group;
    a.block = shape.block;
    a.depth = shape.depth;
    const std::uint64_t grid = shape.grid != 0 ? shape.grid : std::min<std::uint64_t>(plan.task_count, 1u << 20);
    a.grid = static_cast<std::uint32_t>(std::max<std::uint64_t>(grid, 1));
    if (tsk_decode(&a, stream) != hipSuccess) return hip_failure();
    if (tsk_results_complete(plan.decode_images, count, results, stream) != hipSuccess) return hip_failure();
    return Status::success;
}

} // namespace detail

BatchPlan::BatchPlan(std::unique_ptr<Impl> impl) noexcept : impl_(std::move(impl)) {}
BatchPlan::~BatchPlan() = default;
BatchPlan::BatchPlan(BatchPlan &&) noexcept = default;
BatchPlan &BatchPlan::operator=(BatchPlan &&) noexcept = default;

WorkspaceSpec BatchPlan::workspace_spec() const noexcept { return impl_ != nullptr ? impl_->spec : WorkspaceSpec{}; }

Status prepare_batch(const BatchRequest &request, std::unique_ptr<BatchPlan> &plan, hipStream_t stream) noexcept {
    if (request.count > UINT32_MAX) return Status::resource_limit;
    if (request.count != 0 && request.items == nullptr) return Status::invalid_argument;
    if (request.operation > Operation::decode_calibrate) return Status::invalid_argument;
    try {
        auto impl = std::make_unique<BatchPlan::Impl>();
        if (const Status st = detail::build(request, *impl, stream); st != Status::success) return st;
        plan.reset(new BatchPlan(std::move(impl)));
        return Status::success;
    } catch (const std::bad_alloc &) {
        return Status::resource_limit;
    } catch (const std::length_error &) {
        return Status::resource_limit;
    }
}

namespace {

Status decode_entry(const BatchPlan &plan, Operation op, MutableBytes workspace, ResultBuffer results,
                    hipStream_t stream) noexcept {
    const detail::PlanState *state = plan.impl();
    const Status st = detail::check_execution(state, op, workspace, results, stream);
    if (st != Status::success || state->count == 0) return st;
    return detail::run_decode(*state, results.data, stream, state->shape);
}

} // namespace

Status decompress_async(const BatchPlan &plan, MutableBytes workspace, ResultBuffer results, hipStream_t stream) noexcept {
    return decode_entry(plan, Operation::decompress, workspace, results, stream);
}

Status decode_roi_async(const BatchPlan &plan, MutableBytes workspace, ResultBuffer results, hipStream_t stream) noexcept {
    return decode_entry(plan, Operation::decode_roi, workspace, results, stream);
}

Status decode_calibrate_async(const BatchPlan &plan, MutableBytes workspace, ResultBuffer results,
                              hipStream_t stream) noexcept {
    return decode_entry(plan, Operation::decode_calibrate, workspace, results, stream);
}

namespace {

// Enqueues the encoder over a plan's tiles: one launch over the whole batch
// when `whole`, else one per window.  `state` holds the two look-back buffers
// (2 x capacity words) and `carry` the two carried cursors.
Status enqueue_encode(const detail::PlanState &plan, std::uint64_t *state, std::uint64_t *carry, bool whole,
                      DeviceResult *results, hipStream_t stream) noexcept {
    tsk_encode_args a{};
    a.run = detail::encode_run();
    a.images = plan.encode_images;
    a.results = results;
    a.state = state;
    a.carry = carry;
    a.type = static_cast<std::uint32_t>(plan.type);
    if (whole) {
        a.capacity = plan.whole.tile_count;
        a.entries = plan.whole.entries;
        a.blocks = plan.whole.blocks;
        a.entry_count = plan.whole.entry_count;
        a.tile_count = plan.whole.tile_count;
        a.parity = 0;
        if (tsk_encode_reset(&a, stream) != hipSuccess || tsk_encode_window(&a, stream) != hipSuccess)
            return detail::hip_failure();
        return Status::success;
    }
    a.capacity = static_cast<std::uint32_t>(plan.spec.tile_count);
    if (tsk_encode_reset(&a, stream) != hipSuccess) return detail::hip_failure();
    for (std::size_t k = 0; k < plan.windows.size(); ++k) {
        const detail::EncodeWindow &w = plan.windows[k];
        a.entries = w.entries;
        a.blocks = w.blocks;
        a.entry_count = w.entry_count;
        a.tile_count = w.tile_count;
        a.parity = static_cast<std::uint32_t>(k & 1u);
        if (tsk_encode_window(&a, stream) != hipSuccess) return detail::hip_failure();
    }
    return Status::success;
}

} // namespace

Status compress_async(const BatchPlan &plan, MutableBytes workspace, ResultBuffer results, hipStream_t stream) noexcept {
    const detail::PlanState *state = plan.impl();
    const Status st = detail::check_execution(state, Operation::compress, workspace, results, stream);
    if (st != Status::success || state->count == 0) return st;
    auto *ws = static_cast<std::uint8_t *>(workspace.data);
    // Look-back buffers in the residual region (16 KiB per window tile), the
    // carried cursor in the image region.
    return enqueue_encode(*state, reinterpret_cast<std::uint64_t *>(ws + state->layout.residuals),
                          reinterpret_cast<std::uint64_t *>(ws + state->layout.images), state->whole.tile_count != 0,
                          results.data, stream);
}

Status compress_async_static(const BatchPlan &plan, ResultBuffer results, hipStream_t stream) noexcept {
    const detail::PlanState *state = plan.impl();
    // The checks of compress_async, with no workspace.
    if (state == nullptr || state->operation != Operation::compress) return Status::invalid_argument;
    if (state->count == 0) return Status::success;
    const detail::CallNeeds needs{&state->footprint, state->device, state->pageable, state->count, 0};
    if (const Status st = detail::check_call(needs, MutableBytes{nullptr, 0}, results, stream); st != Status::success)
        return st;

    void *scratch = nullptr;
    std::uint64_t scratch_bytes = 0;
    if (tsk_scratch_region(&scratch, &scratch_bytes) != hipSuccess) return detail::hip_failure();
    // Look-back buffers first (2 x 8 bytes per tile of the largest launch),
    // then the two carried cursors.  The whole batch runs as one launch when
    // its buffers fit; otherwise by windows of at most 4096 tiles.
    const auto need = [](std::uint64_t capacity) { return 2 * 8 * capacity + 2 * 8; };
    const bool whole = state->whole.tile_count != 0 && need(state->whole.tile_count) <= scratch_bytes;
    const std::uint64_t capacity = whole ? state->whole.tile_count : state->spec.tile_count;
    if (need(capacity) > scratch_bytes) return Status::resource_limit;
    auto *words = static_cast<std::uint64_t *>(scratch);
    return enqueue_encode(*state, words, words + 2 * capacity, whole, results.data, stream);
}

Status calibrate_async(const BatchPlan &plan, MutableBytes workspace, ResultBuffer results, hipStream_t stream) noexcept {
    const detail::PlanState *state = plan.impl();
    const Status st = detail::check_execution(state, Operation::calibrate, workspace, results, stream);
    if (st != Status::success || state->count == 0) return st;
    tsk_calibrate_args a{};
    a.images = state->calibrate_images;
    a.first_rows = state->first_rows;
    a.results = results.data;
    a.total_rows = state->total_rows;
    a.image_count = static_cast<std::uint32_t>(state->count);
    a.type = static_cast<std::uint32_t>(state->type);
    return detail::hip_status(tsk_calibrate(&a, stream));
}

// DeviceBuffer: stream-ordered allocations (reference src/cuda/buffer.cpp:8-61).
DeviceBuffer::~DeviceBuffer() {
    if (data_ != nullptr) static_cast<void>(detail::hip_status(hipFreeAsync(data_, hipStreamPerThread)));
}

DeviceBuffer::DeviceBuffer(DeviceBuffer &&other) noexcept
    : data_(std::exchange(other.data_, nullptr)), size_(std::exchange(other.size_, 0)) {}

DeviceBuffer &DeviceBuffer::operator=(DeviceBuffer &&other) noexcept {
    if (this != &other) {
        DeviceBuffer old(std::move(*this));
        data_ = std::exchange(other.data_, nullptr);
        size_ = std::exchange(other.size_, 0);
    }
    return *this;
}

Status DeviceBuffer::allocate(std::uint64_t bytes, hipStream_t stream, DeviceBuffer &buffer) noexcept {
    if (bytes > max_bytes) return Status::size_overflow;
    DeviceBuffer fresh;
    if (bytes != 0) {
        if (hipMallocAsync(&fresh.data_, static_cast<std::size_t>(bytes), stream) != hipSuccess) {
            fresh.data_ = nullptr;
            return detail::hip_failure();
        }
        fresh.size_ = bytes;
        if (reinterpret_cast<std::uintptr_t>(fresh.data_) % workspace_alignment != 0) {
            static_cast<void>(fresh.release(stream));
            return Status::hip_error;
        }
    }
    const Status st = buffer.release(stream);
    buffer = std::move(fresh);
    return st;
}

Status DeviceBuffer::release(hipStream_t stream) noexcept {
    if (data_ == nullptr) return Status::success;
    const hipError_t e = hipFreeAsync(data_, stream);
    data_ = nullptr;
    size_ = 0;
    return detail::hip_status(e);
}

Status synchronize(hipStream_t stream) noexcept { return detail::hip_status(hipStreamSynchronize(stream)); }

} // namespace tilestream::hip
