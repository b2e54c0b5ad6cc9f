// Validating host parser of .tsc 1.0 archives (reference src/cpu/parser.cpp:27-242;
// docs/format.md; status mapping of docs/api.md "Host metadata versus device
// metadata").  Checks fire in the reference's order so that every malformed
// input maps to the same Status.
#include "checked.h"

#include <tilestream/api.h>
#include <tilestream/format.h>

#include <algorithm>
#include <new>
#include <stdexcept>
#include <system_error>
#include <thread>
#include <utility>
#include <vector>

namespace tilestream {
namespace {

namespace hdr = format::header;
namespace ent = format::entry;

struct Geometry {
    ImageSpec image;
    TileShape tile;
    std::uint64_t across = 0;       // tile columns
    std::uint64_t count = 0;        // tiles
    std::uint64_t payload_at = 0;   // first payload byte
};

Status read_fixed_header(const std::uint8_t *d, std::uint64_t size, Geometry &g) noexcept {
    using host::load_le;
    if (size < format::header_bytes) return Status::malformed_data;
    if (!std::equal(format::magic.begin(), format::magic.end(), d)) return Status::malformed_data;
    if (load_le<std::uint16_t>(d + hdr::version_major) != format::major_version ||
        load_le<std::uint16_t>(d + hdr::version_minor) != format::minor_version || d[hdr::flags] != 0)
        return Status::unsupported_format;
    if (load_le<std::uint16_t>(d + hdr::size) != format::header_bytes ||
        load_le<std::uint16_t>(d + hdr::entry_size) != format::index_entry_bytes ||
        !host::zeros(d + hdr::reserved, 2))
        return Status::malformed_data;
    const std::uint8_t type = d[hdr::pixel_type];
    if (type != static_cast<std::uint8_t>(PixelType::uint16) &&
        type != static_cast<std::uint8_t>(PixelType::float32))
        return Status::unsupported_type;
    g.image = {static_cast<PixelType>(type), load_le<std::uint32_t>(d + hdr::width),
               load_le<std::uint32_t>(d + hdr::height)};
    g.tile = {load_le<std::uint16_t>(d + hdr::tile_width), load_le<std::uint16_t>(d + hdr::tile_height)};
    if (g.image.width == 0 || g.image.height == 0) return Status::malformed_data;
    if (!host::allowed_tile(g.tile)) return Status::unsupported_format;
    std::uint64_t bound = 0;
    if (const Status st = compress_bound(g.image, g.tile, bound); st != Status::success) return st;
    g.across = host::ceil_div(g.image.width, g.tile.width);
    std::uint64_t index_bytes = 0;
    if (!host::mul(g.across, host::ceil_div(g.image.height, g.tile.height), g.count) ||
        !host::mul(g.count, format::index_entry_bytes, index_bytes) ||
        !host::add(format::header_bytes, index_bytes, g.payload_at))
        return Status::size_overflow;
    if (load_le<std::uint64_t>(d + hdr::tile_count) != g.count ||
        load_le<std::uint64_t>(d + hdr::index_offset) != format::header_bytes ||
        load_le<std::uint64_t>(d + hdr::payload_offset) != g.payload_at ||
        load_le<std::uint64_t>(d + hdr::file_size) != size || size > bound || g.payload_at > size)
        return Status::malformed_data;
    return Status::success;
}

// The index alone: every entry's fields, payloads laid end to end in tile
// order up to the end of the file, and the first row slot of each predictive
// tile (a predictive tile has one row per tile row).  Every failure is
// malformed_data, as in one sequential walk, so the passes may be split.
Status walk_index(const std::uint8_t *d, std::uint64_t size, const Geometry &g, std::vector<std::uint64_t> &first,
                  std::uint64_t &rows) noexcept {
    using host::load_le;
    const std::uint64_t s = g.image.type == PixelType::uint16 ? 2 : 4;
    std::uint64_t next = g.payload_at;
    rows = 0;
    for (std::uint64_t i = 0; i < g.count; ++i) {
        const std::uint8_t *e = d + format::header_bytes + i * format::index_entry_bytes;
        const std::uint64_t x = i % g.across * g.tile.width, y = i / g.across * g.tile.height;
        const std::uint64_t w = std::min<std::uint64_t>(g.tile.width, g.image.width - x);
        const std::uint64_t h = std::min<std::uint64_t>(g.tile.height, g.image.height - y);
        const std::uint64_t at = load_le<std::uint64_t>(e + ent::offset);
        const std::uint64_t len = load_le<std::uint64_t>(e + ent::length);
        const std::uint64_t raw4 = (w * h * s + 3) / 4 * 4;
        const std::uint8_t mode = e[ent::mode];
        if (mode > static_cast<std::uint8_t>(TileMode::predict_bitpack) || !host::zeros(e + ent::reserved, 3) ||
            !host::zeros(e + ent::reserved_tail, 8) || load_le<std::uint16_t>(e + ent::width) != w ||
            load_le<std::uint16_t>(e + ent::height) != h || at != next || at > size ||
            len > size - at || len == 0 || len % 4 != 0 || len > raw4)
            return Status::malformed_data;
        first[static_cast<std::size_t>(i)] = rows;
        if (mode == static_cast<std::uint8_t>(TileMode::predict_bitpack)) rows += h;
        next += len;
    }
    return next == size ? Status::success : Status::malformed_data;
}

// Payloads of tiles [begin, end), whose index entries walk_index accepted.
// With `out` null it only validates; otherwise it fills the preallocated tables.
Status walk_payloads(const std::uint8_t *d, const Geometry &g, const std::uint64_t *first, std::uint64_t begin,
                     std::uint64_t end, ImageMetadata *out) noexcept {
    using host::load_le;
    const std::uint64_t s = g.image.type == PixelType::uint16 ? 2 : 4;
    const std::uint32_t widest = g.image.type == PixelType::uint16 ? 16 : 32;
    for (std::uint64_t i = begin; i < end; ++i) {
        const std::uint8_t *e = d + format::header_bytes + i * format::index_entry_bytes;
        const std::uint64_t x = i % g.across * g.tile.width, y = i / g.across * g.tile.height;
        const std::uint64_t w = std::min<std::uint64_t>(g.tile.width, g.image.width - x);
        const std::uint64_t h = std::min<std::uint64_t>(g.tile.height, g.image.height - y);
        const std::uint64_t at = load_le<std::uint64_t>(e + ent::offset);
        const std::uint64_t len = load_le<std::uint64_t>(e + ent::length);
        const std::uint64_t raw = w * h * s, raw4 = (raw + 3) / 4 * 4;
        const std::uint8_t mode = e[ent::mode];
        const std::uint8_t *p = d + at;
        std::uint64_t row = 0;
        switch (static_cast<TileMode>(mode)) {
        case TileMode::raw:
            if (len != raw4 || !host::zeros(p + raw, len - raw)) return Status::malformed_data;
            break;
        case TileMode::constant:
            if (len != 4 || (s == 2 && !host::zeros(p + 2, 2))) return Status::malformed_data;
            break;
        case TileMode::predict_bitpack: {
            row = first[static_cast<std::size_t>(i)];
            std::uint64_t pos = 0;
            for (std::uint64_t r = 0; r < h; ++r) {
                if (len - pos < format::segment_header_bytes) return Status::malformed_data;
                const std::uint32_t seed = load_le<std::uint32_t>(p + pos + format::segment::seed);
                const std::uint32_t b = p[pos + format::segment::bit_width];
                if ((s == 2 && seed > 0xffffu) || b > widest || (w == 1 && b != 0) ||
                    !host::zeros(p + pos + format::segment::reserved, 3))
                    return Status::malformed_data;
                pos += format::segment_header_bytes;
                const std::uint64_t bits = (w - 1) * b, packed = (bits + 31) / 32 * 4;
                if (packed > len - pos) return Status::malformed_data;
                // High bits of the last word beyond the final residual must be zero.
                if (bits % 32 != 0 && (load_le<std::uint32_t>(p + pos + packed - 4) >> (bits % 32)) != 0)
                    return Status::malformed_data;
                if (out != nullptr) out->segments[static_cast<std::size_t>(row + r)] = {seed, b, pos, packed};
                pos += packed;
            }
            if (pos != len) return Status::malformed_data;
            break;
        }
        }
        if (out != nullptr)
            out->tiles[static_cast<std::size_t>(i)] = {static_cast<TileMode>(mode), static_cast<std::uint16_t>(w),
                                                       static_cast<std::uint16_t>(h), at, len, row};
    }
    return Status::success;
}

// walk_payloads over every tile; archives of a few MiB and more are split
// into contiguous tile ranges over host threads (the ranges touch disjoint
// payloads and table slots).
Status walk_all(const std::uint8_t *d, std::uint64_t size, const Geometry &g, const std::uint64_t *first,
                ImageMetadata *out) noexcept {
    constexpr std::uint64_t kBytesPerThread = 4ull << 20;
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const std::uint64_t threads =
        std::min<std::uint64_t>({std::uint64_t{hw}, 16, size / kBytesPerThread + 1, (g.count + 63) / 64});
    if (threads <= 1) return walk_payloads(d, g, first, 0, g.count, out);
    const std::uint64_t per = (g.count + threads - 1) / threads;
    const auto range = [&](std::uint64_t k) {
        return walk_payloads(d, g, first, std::min(g.count, k * per), std::min(g.count, (k + 1) * per), out);
    };
    std::vector<Status> status;
    std::vector<std::thread> pool;
    try {
        // Sized up front: starting a worker never reallocates what the
        // running ones write to.
        status.assign(static_cast<std::size_t>(threads), Status::success);
        pool.reserve(static_cast<std::size_t>(threads - 1));
    } catch (const std::bad_alloc &) {
        return Status::resource_limit;
    } catch (const std::length_error &) {
        return Status::resource_limit;
    }
    try {
        for (std::uint64_t k = 1; k < threads; ++k)
            pool.emplace_back([&status, &range, k] { status[static_cast<std::size_t>(k)] = range(k); });
    } catch (const std::system_error &) {
        // Fewer threads than asked for: the ranges left without one run below.
    }
    const std::uint64_t started = pool.size();
    status[0] = range(0);
    for (std::uint64_t k = started + 1; k < threads; ++k) status[static_cast<std::size_t>(k)] = range(k);
    for (std::thread &t : pool) t.join();
    for (const Status st : status)
        if (st != Status::success) return st;
    return Status::success;
}

Status native_bytes(std::uint64_t tiles, std::uint64_t segments, std::uint64_t &bytes) noexcept {
    std::uint64_t a = 0, b = 0;
    if (!host::mul(tiles, sizeof(TileMetadata), a) || !host::mul(segments, sizeof(SegmentMetadata), b) ||
        !host::add(sizeof(ImageMetadata), a, bytes) || !host::add(bytes, b, bytes))
        return Status::size_overflow;
    return Status::success;
}

} // namespace

Status parse_metadata(ConstBytes source, const ParseLimits &limits, ImageMetadata &metadata) noexcept {
    if (const Status st = host::byte_range(source.data, source.size); st != Status::success) return st;
    if (source.size > limits.file_bytes) return Status::resource_limit;
    const auto *d = static_cast<const std::uint8_t *>(source.data);
    Geometry g;
    if (const Status st = read_fixed_header(d, source.size, g); st != Status::success) return st;
    if (g.count > limits.tile_count) return Status::resource_limit;
    std::uint64_t need = 0;
    if (const Status st = native_bytes(g.count, 0, need); st != Status::success) return st;
    if (need > limits.metadata_bytes) return Status::resource_limit;
    try {
        // The index is at most 1/32 of the file, so its row table is small.
        std::vector<std::uint64_t> first(static_cast<std::size_t>(g.count));
        std::uint64_t rows = 0;
        if (const Status st = walk_index(d, source.size, g, first, rows); st != Status::success) return st;
        if (const Status st = walk_all(d, source.size, g, first.data(), nullptr); st != Status::success) return st;
        if (const Status st = native_bytes(g.count, rows, need); st != Status::success) return st;
        if (need > limits.metadata_bytes) return Status::resource_limit;
        ImageMetadata parsed;
        parsed.image = g.image;
        parsed.tile = g.tile;
        parsed.file_bytes = source.size;
        parsed.tiles.resize(static_cast<std::size_t>(g.count));
        parsed.segments.resize(static_cast<std::size_t>(rows));
        if (const Status st = walk_all(d, source.size, g, first.data(), &parsed); st != Status::success) return st;
        metadata = std::move(parsed);
        return Status::success;
    } catch (const std::bad_alloc &) {
        return Status::resource_limit;
    } catch (const std::length_error &) {
        return Status::resource_limit;
    }
}

} // namespace tilestream
