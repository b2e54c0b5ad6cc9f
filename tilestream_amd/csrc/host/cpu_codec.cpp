// Scalar host codec of the reference API (reference include/tilestream/cpu.h:9-12,
// src/cpu/encoder.cpp:144-197, src/cpu/decoder.cpp:86-110).  Argument checks and
// their statuses follow the reference; the bit packing uses a 64-bit reservoir.
#include "checked.h"

#include <tilestream/cpu.h>
#include <tilestream/format.h>

#include <algorithm>
#include <array>
#include <memory>
#include <vector>

namespace tilestream::cpu {
namespace {

struct TileWork {
    std::array<std::uint32_t, 4096> px{};
    std::array<std::uint32_t, 4096> res{};
    std::array<std::uint8_t, 64> bits{};
};

inline std::uint32_t sample_at(const std::uint8_t *base, std::uint64_t pitch, std::uint64_t x,
                               std::uint64_t y, std::uint64_t s) {
    const std::uint8_t *p = base + y * pitch + x * s;
    return s == 2 ? host::load_le<std::uint16_t>(p) : host::load_le<std::uint32_t>(p);
}

// Zigzag of the signed 16-bit difference (docs/format.md "uint16 transform").
inline std::uint32_t zigzag(std::uint32_t cur, std::uint32_t prev) {
    const std::uint32_t d = (cur - prev) & 0xffffu;
    return d < 32768 ? 2 * d : 2 * (65536 - d) - 1;
}

// Chooses the canonical mode and writes the payload; returns its length.
std::uint64_t encode_tile(const ConstImageView &in, std::uint64_t x0, std::uint64_t y0, std::uint32_t w,
                          std::uint32_t h, TileWork &t, std::uint8_t *out, TileMode &mode) {
    const auto *base = static_cast<const std::uint8_t *>(in.data);
    const std::uint64_t s = in.image.type == PixelType::uint16 ? 2 : 4;
    const bool u16 = s == 2;
    bool same = true;
    std::uint64_t predictive = 0;
    for (std::uint32_t r = 0; r < h; ++r) {
        std::uint32_t any = 0;
        for (std::uint32_t c = 0; c < w; ++c) {
            const std::uint32_t v = sample_at(base, in.pitch_bytes, x0 + c, y0 + r, s);
            t.px[r * w + c] = v;
            same = same && v == t.px[0];
            if (c != 0) {
                const std::uint32_t prev = t.px[r * w + c - 1];
                t.res[r * w + c] = u16 ? zigzag(v, prev) : (v ^ prev);
                any |= t.res[r * w + c];
            }
        }
        std::uint8_t b = 0;
        while (any != 0) {
            ++b;
            any >>= 1;
        }
        t.bits[r] = b;
        predictive += format::segment_header_bytes + (std::uint64_t{w - 1u} * b + 31) / 32 * 4;
    }
    // Candidate costs differ only by their payload; ties keep RAW, then CONSTANT.
    std::uint64_t length = (std::uint64_t{w} * h * s + 3) / 4 * 4;
    mode = TileMode::raw;
    if (same && length > 4) {
        mode = TileMode::constant;
        length = 4;
    }
    if (predictive < length) {
        mode = TileMode::predict_bitpack;
        length = predictive;
    }
    std::fill(out, out + length, std::uint8_t{0});
    if (mode == TileMode::raw) {
        for (std::uint32_t i = 0; i < w * h; ++i) {
            if (u16)
                host::store_le<std::uint16_t>(out + 2 * i, static_cast<std::uint16_t>(t.px[i]));
            else
                host::store_le<std::uint32_t>(out + 4 * i, t.px[i]);
        }
    } else if (mode == TileMode::constant) {
        host::store_le<std::uint32_t>(out, t.px[0]);
    } else {
        std::uint8_t *q = out;
        for (std::uint32_t r = 0; r < h; ++r) {
            const std::uint32_t b = t.bits[r];
            host::store_le<std::uint32_t>(q, t.px[r * w]);
            q[format::segment::bit_width] = static_cast<std::uint8_t>(b);
            q += format::segment_header_bytes;
            std::uint64_t pending = 0;
            std::uint32_t filled = 0;
            for (std::uint32_t c = 1; c < w && b != 0; ++c) {
                pending |= std::uint64_t{t.res[r * w + c]} << filled;
                filled += b;
                if (filled >= 32) {
                    host::store_le<std::uint32_t>(q, static_cast<std::uint32_t>(pending));
                    q += 4;
                    pending >>= 32;
                    filled -= 32;
                }
            }
            if (filled != 0) {
                host::store_le<std::uint32_t>(q, static_cast<std::uint32_t>(pending));
                q += 4;
            }
        }
    }
    return length;
}

void write_header(std::uint8_t *d, const ImageSpec &image, TileShape tile, std::uint64_t count,
                  std::uint64_t size) {
    namespace h = format::header;
    std::fill(d, d + format::header_bytes, std::uint8_t{0});
    std::copy(format::magic.begin(), format::magic.end(), d);
    host::store_le<std::uint16_t>(d + h::version_major, format::major_version);
    host::store_le<std::uint16_t>(d + h::version_minor, format::minor_version);
    host::store_le<std::uint16_t>(d + h::size, format::header_bytes);
    d[h::pixel_type] = static_cast<std::uint8_t>(image.type);
    host::store_le<std::uint32_t>(d + h::width, image.width);
    host::store_le<std::uint32_t>(d + h::height, image.height);
    host::store_le<std::uint16_t>(d + h::tile_width, tile.width);
    host::store_le<std::uint16_t>(d + h::tile_height, tile.height);
    host::store_le<std::uint16_t>(d + h::entry_size, format::index_entry_bytes);
    host::store_le<std::uint64_t>(d + h::tile_count, count);
    host::store_le<std::uint64_t>(d + h::index_offset, format::header_bytes);
    host::store_le<std::uint64_t>(d + h::payload_offset, format::header_bytes + count * format::index_entry_bytes);
    host::store_le<std::uint64_t>(d + h::file_size, size);
}

std::uint32_t residual_at(const std::uint8_t *words, std::uint64_t j, std::uint32_t b) {
    if (b == 0) return 0;
    const std::uint64_t pos = j * b;
    std::uint64_t window = host::load_le<std::uint32_t>(words + pos / 32 * 4);
    if (pos % 32 + b > 32) window |= std::uint64_t{host::load_le<std::uint32_t>(words + pos / 32 * 4 + 4)} << 32;
    window >>= pos % 32;
    return b == 32 ? static_cast<std::uint32_t>(window)
                   : static_cast<std::uint32_t>(window & ((std::uint64_t{1} << b) - 1));
}

} // namespace

Status compress(ConstImageView input, TileShape tile, MutableBytes output,
                std::uint64_t &bytes_written) noexcept {
    if (const Status st = host::check_view(input); st != Status::success) return st;
    std::uint64_t bound = 0;
    if (const Status st = compress_bound(input.image, tile, bound); st != Status::success) return st;
    if (const Status st = host::byte_range(output.data, output.size); st != Status::success) return st;
    if (output.size < bound) return Status::buffer_too_small;
    if (host::overlaps(input.data, input.capacity_bytes, output.data, output.size) ||
        host::overlaps(&bytes_written, sizeof bytes_written, output.data, output.size) ||
        host::overlaps(&bytes_written, sizeof bytes_written, input.data, input.capacity_bytes))
        return Status::invalid_argument;
    const std::uint64_t across = host::ceil_div(input.image.width, tile.width);
    const std::uint64_t count = across * host::ceil_div(input.image.height, tile.height);
    auto *d = static_cast<std::uint8_t *>(output.data);
    std::uint64_t cursor = format::header_bytes + count * format::index_entry_bytes;
    auto work = std::make_unique<TileWork>();
    for (std::uint64_t i = 0; i < count; ++i) {
        const std::uint64_t x = i % across * tile.width, y = i / across * tile.height;
        const auto w = static_cast<std::uint32_t>(std::min<std::uint64_t>(tile.width, input.image.width - x));
        const auto h = static_cast<std::uint32_t>(std::min<std::uint64_t>(tile.height, input.image.height - y));
        TileMode mode = TileMode::raw;
        const std::uint64_t len = encode_tile(input, x, y, w, h, *work, d + cursor, mode);
        std::uint8_t *e = d + format::header_bytes + i * format::index_entry_bytes;
        std::fill(e, e + format::index_entry_bytes, std::uint8_t{0});
        e[format::entry::mode] = static_cast<std::uint8_t>(mode);
        host::store_le<std::uint16_t>(e + format::entry::width, static_cast<std::uint16_t>(w));
        host::store_le<std::uint16_t>(e + format::entry::height, static_cast<std::uint16_t>(h));
        host::store_le<std::uint64_t>(e + format::entry::offset, cursor);
        host::store_le<std::uint64_t>(e + format::entry::length, len);
        cursor += len;
    }
    write_header(d, input.image, tile, count, cursor);
    bytes_written = cursor;
    return Status::success;
}

Status decompress(ConstBytes input, ImageView output, const ParseLimits &limits) noexcept {
    if (const Status st = host::check_view(host::as_const(output)); st != Status::success) return st;
    if (const Status st = host::byte_range(input.data, input.size); st != Status::success) return st;
    if (host::overlaps(input.data, input.size, output.data, output.capacity_bytes))
        return Status::invalid_argument;
    ImageMetadata m;
    if (const Status st = parse_metadata(input, limits, m); st != Status::success) return st;
    if (output.image.width != m.image.width || output.image.height != m.image.height ||
        output.image.type != m.image.type)
        return Status::invalid_argument;
    const auto *src = static_cast<const std::uint8_t *>(input.data);
    auto *dst = static_cast<std::uint8_t *>(output.data);
    const bool u16 = m.image.type == PixelType::uint16;
    const std::uint64_t s = u16 ? 2 : 4;
    const std::uint64_t across = host::ceil_div(m.image.width, m.tile.width);
    for (std::size_t i = 0; i < m.tiles.size(); ++i) {
        const TileMetadata &t = m.tiles[i];
        const std::uint64_t x = i % across * m.tile.width, y = i / across * m.tile.height;
        const std::uint8_t *p = src + t.offset;
        for (std::uint32_t r = 0; r < t.height; ++r) {
            std::uint8_t *line = dst + (y + r) * output.pitch_bytes + x * s;
            std::uint32_t v = 0;
            for (std::uint32_t c = 0; c < t.width; ++c) {
                if (t.mode == TileMode::raw) {
                    const std::uint8_t *q = p + (std::uint64_t{r} * t.width + c) * s;
                    v = u16 ? host::load_le<std::uint16_t>(q) : host::load_le<std::uint32_t>(q);
                } else if (t.mode == TileMode::constant) {
                    v = u16 ? host::load_le<std::uint16_t>(p) : host::load_le<std::uint32_t>(p);
                } else {
                    const SegmentMetadata &g = m.segments[t.first_segment + r];
                    if (c == 0) {
                        v = g.seed;
                    } else {
                        const std::uint32_t z = residual_at(p + g.payload_offset, c - 1, g.bit_width);
                        // Inverse zigzag added modulo 2^16, or XOR of FP32 bit words.
                        v = u16 ? ((v + ((z & 1u) ? 0xffffu - (z >> 1) : (z >> 1))) & 0xffffu) : (v ^ z);
                    }
                }
                if (u16)
                    host::store_le<std::uint16_t>(line + 2 * c, static_cast<std::uint16_t>(v));
                else
                    host::store_le<std::uint32_t>(line + 4 * c, v);
            }
        }
    }
    return Status::success;
}

} // namespace tilestream::cpu
