// TIFF reader and writer over zlib for the file tier.  The reference adapts
// libtiff (src/io/tiff.cpp:59-345); the accepted subset and the statuses
// follow it: one directory, one sample per pixel, 16-bit unsigned or 32-bit
// IEEE samples, min-is-black, top-left, chunky; tiles or strips; no
// compression or Deflate (codes 8 and 32946), with the horizontal predictor.
#include "checked.h"

#include <tilestream/tiff.h>

#include <zlib.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <functional>
#include <new>
#include <stdexcept>
#include <system_error>
#include <thread>
#include <vector>

namespace tilestream::io {
namespace {

enum : std::uint16_t {
    kWidth = 256, kLength = 257, kBits = 258, kCompression = 259, kPhotometric = 262, kStripOffsets = 273,
    kOrientation = 274, kSamples = 277, kRowsPerStrip = 278, kStripBytes = 279, kPlanar = 284, kPredictor = 317,
    kTileWidth = 322, kTileLength = 323, kTileOffsets = 324, kTileBytes = 325, kSampleFormat = 339,
};
enum : std::uint16_t { kShort = 3, kLong = 4, kLong8 = 16 };

// Thrown while parsing: a structure that cannot be read (io_error) or a
// valid TIFF outside the subset (unsupported_format).
struct Malformed {};
struct Unsupported {};

std::uint64_t type_size(std::uint16_t type) noexcept {
    switch (type) {
    case 3: case 8: return 2;
    case 4: case 9: case 11: case 13: return 4;
    case 5: case 10: case 12: case 16: case 17: case 18: return 8;
    default: return 1;
    }
}

struct Entry {
    std::uint16_t tag = 0, type = 0;
    std::uint64_t count = 0;
    std::uint64_t at = 0; // file offset of the values
};

class TiffFile {
  public:
    explicit TiffFile(std::vector<std::uint8_t> bytes) : b_(std::move(bytes)) {
        if (b_.size() < 8) throw Malformed{};
        if (b_[0] == 'I' && b_[1] == 'I') big_endian_ = false;
        else if (b_[0] == 'M' && b_[1] == 'M') big_endian_ = true;
        else throw Malformed{};
        const std::uint64_t magic = get(2, 2);
        std::uint64_t ifd = 0;
        if (magic == 42) {
            ifd = get(4, 4);
        } else if (magic == 43) {
            if (get(4, 2) != 8 || get(6, 2) != 0) throw Malformed{};
            bigtiff_ = true;
            ifd = get(8, 8);
        } else {
            throw Malformed{};
        }
        const std::uint64_t count = get(ifd, bigtiff_ ? 8 : 2);
        const std::uint64_t first = ifd + (bigtiff_ ? 8 : 2), size = bigtiff_ ? 20 : 12, inline_bytes = bigtiff_ ? 8 : 4;
        if (count > (b_.size() - std::min<std::uint64_t>(first, b_.size())) / size) throw Malformed{};
        for (std::uint64_t k = 0; k < count; ++k) {
            const std::uint64_t e = first + k * size;
            Entry x;
            x.tag = static_cast<std::uint16_t>(get(e, 2));
            x.type = static_cast<std::uint16_t>(get(e + 2, 2));
            x.count = get(e + 4, bigtiff_ ? 8 : 4);
            const std::uint64_t value = e + (bigtiff_ ? 12 : 8);
            std::uint64_t bytes = 0;
            if (!host::mul(x.count, type_size(x.type), bytes)) throw Malformed{};
            x.at = bytes <= inline_bytes ? value : get(value, inline_bytes);
            if (x.at > b_.size() || bytes > b_.size() - x.at) throw Malformed{};
            entries_.push_back(x);
        }
        // One image directory only (reference tiff.cpp:69).
        if (get(first + count * size, bigtiff_ ? 8 : 4) != 0) throw Unsupported{};
    }

    // Little- or big-endian unsigned value of n bytes at `at`.
    std::uint64_t get(std::uint64_t at, unsigned n) const {
        if (at > b_.size() || n > b_.size() - at) throw Malformed{};
        std::uint64_t v = 0;
        for (unsigned k = 0; k < n; ++k) {
            const std::uint64_t byte = b_[static_cast<std::size_t>(at + (big_endian_ ? k : n - 1 - k))];
            v = (v << 8) | byte;
        }
        return v;
    }

    const Entry *find(std::uint16_t tag) const noexcept {
        for (const Entry &e : entries_)
            if (e.tag == tag) return &e;
        return nullptr;
    }

    std::vector<std::uint64_t> values(const Entry &e) const {
        if (e.type != kShort && e.type != kLong && e.type != kLong8) throw Unsupported{};
        const unsigned n = static_cast<unsigned>(type_size(e.type));
        std::vector<std::uint64_t> v(static_cast<std::size_t>(e.count));
        for (std::uint64_t k = 0; k < e.count; ++k) v[static_cast<std::size_t>(k)] = get(e.at + k * n, n);
        return v;
    }

    // A one-value field, or `fallback` when absent (libtiff's defaulted reads).
    std::uint64_t scalar(std::uint16_t tag, std::uint64_t fallback, bool required = false) const {
        const Entry *e = find(tag);
        if (e == nullptr) {
            if (required) throw Unsupported{};
            return fallback;
        }
        const std::vector<std::uint64_t> v = values(*e);
        if (v.size() != 1) throw Unsupported{};
        return v[0];
    }

    bool big_endian() const noexcept { return big_endian_; }
    const std::uint8_t *data() const noexcept { return b_.data(); }
    std::uint64_t size() const noexcept { return b_.size(); }

  private:
    std::vector<std::uint8_t> b_;
    std::vector<Entry> entries_;
    bool big_endian_ = false, bigtiff_ = false;
};

bool read_file(const char *path, std::vector<std::uint8_t> &out) {
    std::FILE *f = std::fopen(path, "rb");
    if (f == nullptr) return false;
    bool ok = std::fseek(f, 0, SEEK_END) == 0;
    const long size = ok ? std::ftell(f) : -1;
    ok = ok && size >= 0 && static_cast<std::uint64_t>(size) <= max_bytes && std::fseek(f, 0, SEEK_SET) == 0;
    if (ok) {
        out.resize(static_cast<std::size_t>(size));
        ok = size == 0 || std::fread(out.data(), 1, out.size(), f) == out.size();
    }
    std::fclose(f);
    return ok;
}

// One tile or strip decoded into `out` (expected bytes), in host byte order
// with the predictor undone over rows of `row_samples` samples.
void decode_chunk(const TiffFile &t, std::uint64_t offset, std::uint64_t count, bool deflate, bool predictor,
                  unsigned sample, std::uint64_t row_samples, std::vector<std::uint8_t> &out) {
    const std::uint64_t expected = out.size();
    if (offset > t.size() || count > t.size() - offset) throw Malformed{};
    const std::uint8_t *src = t.data() + offset;
    if (!deflate) {
        if (count < expected) throw Malformed{};
        std::memcpy(out.data(), src, static_cast<std::size_t>(expected));
    } else {
        z_stream z{};
        if (inflateInit(&z) != Z_OK) throw std::bad_alloc{};
        std::uint64_t in_left = count, out_done = 0;
        int rc = Z_OK;
        while (rc == Z_OK && out_done < expected) {
            const auto in_now = static_cast<uInt>(std::min<std::uint64_t>(in_left, 1u << 30));
            const auto out_now = static_cast<uInt>(std::min<std::uint64_t>(expected - out_done, 1u << 30));
            z.next_in = const_cast<Bytef *>(src + (count - in_left));
            z.avail_in = in_now;
            z.next_out = out.data() + out_done;
            z.avail_out = out_now;
            rc = inflate(&z, Z_NO_FLUSH);
            in_left -= in_now - z.avail_in;
            out_done += out_now - z.avail_out;
            if (rc == Z_BUF_ERROR && in_left == 0) break;
            if (rc == Z_BUF_ERROR) rc = Z_OK;
        }
        inflateEnd(&z);
        if (out_done != expected || (rc != Z_OK && rc != Z_STREAM_END)) throw Malformed{};
    }
    if (t.big_endian()) {
        for (std::uint64_t k = 0; k + sample <= expected; k += sample)
            std::reverse(out.data() + k, out.data() + k + sample);
    }
    if (predictor) {
        const std::uint64_t rows = expected / (row_samples * sample);
        for (std::uint64_t r = 0; r < rows; ++r) {
            std::uint8_t *row = out.data() + r * row_samples * sample;
            if (sample == 2) {
                std::uint16_t prev = 0, v = 0;
                for (std::uint64_t x = 0; x < row_samples; ++x) {
                    std::memcpy(&v, row + 2 * x, 2);
                    v = static_cast<std::uint16_t>(v + (x == 0 ? 0 : prev));
                    std::memcpy(row + 2 * x, &v, 2);
                    prev = v;
                }
            } else {
                std::uint32_t prev = 0, v = 0;
                for (std::uint64_t x = 0; x < row_samples; ++x) {
                    std::memcpy(&v, row + 4 * x, 4);
                    v = v + (x == 0 ? 0u : prev);
                    std::memcpy(row + 4 * x, &v, 4);
                    prev = v;
                }
            }
        }
    }
}

// Runs body(begin, end) over contiguous ranges of [0, n) on up to 16 host
// threads (one per 4 MiB of file: tiles and strips are independent); returns
// the status of a failing range (every chunk failure is io_error).
template <typename Body> Status parallel_ranges(std::uint64_t n, std::uint64_t file_bytes, const Body &body) {
    auto run = [&body](std::uint64_t begin, std::uint64_t end, Status &status) {
        try {
            body(begin, end);
        } catch (const Malformed &) {
            status = Status::io_error;
        } catch (const std::bad_alloc &) {
            status = Status::resource_limit;
        } catch (const std::length_error &) {
            status = Status::resource_limit;
        }
    };
    const std::uint64_t hw = std::max(1u, std::thread::hardware_concurrency());
    const std::uint64_t threads = std::max<std::uint64_t>(1, std::min<std::uint64_t>({hw, 16, file_bytes / (4u << 20) + 1, n}));
    std::vector<Status> status(static_cast<std::size_t>(threads), Status::success);
    if (threads == 1) {
        run(0, n, status[0]);
    } else {
        std::vector<std::thread> pool;
        try {
            for (std::uint64_t j = 0; j < threads; ++j)
                pool.emplace_back(run, n * j / threads, n * (j + 1) / threads, std::ref(status[j]));
        } catch (const std::system_error &) {
            for (std::thread &worker : pool) worker.join();
            return Status::resource_limit;
        }
        for (std::thread &worker : pool) worker.join();
    }
    for (Status st : status)
        if (st != Status::success) return st;
    return Status::success;
}

Status decode(const TiffFile &t, const TiffReadLimits &limits, TiffImage &img) {
    const std::uint64_t width = t.scalar(kWidth, 0, true), height = t.scalar(kLength, 0, true);
    const std::uint64_t samples = t.scalar(kSamples, 1), bits = t.scalar(kBits, 1), format = t.scalar(kSampleFormat, 1);
    const std::uint64_t planar = t.scalar(kPlanar, 1), orientation = t.scalar(kOrientation, 1);
    const std::uint64_t photometric = t.scalar(kPhotometric, 0, true), compression = t.scalar(kCompression, 1);
    const std::uint64_t predictor = t.scalar(kPredictor, 1);
    if (samples != 1 || planar != 1 || orientation != 1 || photometric != 1) return Status::unsupported_format;
    PixelType type;
    if (bits == 16 && format == 1) type = PixelType::uint16;
    else if (bits == 32 && format == 3) type = PixelType::float32;
    else return Status::unsupported_format;
    const bool deflate = compression == 8 || compression == 32946;
    if (!deflate && compression != 1) return Status::unsupported_format;
    if (predictor != 1 && predictor != 2) return Status::unsupported_format;
    const bool horizontal = deflate && predictor == 2;
    // Image size checks (reference tiff.cpp:42-57).
    if (width == 0 || height == 0 || width > UINT32_MAX || height > UINT32_MAX) return Status::unsupported_format;
    if (width > limits.max_width || height > limits.max_height) return Status::resource_limit;
    const unsigned sample = type == PixelType::uint16 ? 2 : 4;
    std::uint64_t row_bytes = 0, total = 0;
    if (!host::mul(width, sample, row_bytes) || !host::mul(row_bytes, height, total)) return Status::size_overflow;
    if (total > limits.max_decoded_bytes) return Status::resource_limit;

    TiffImage out;
    out.image = {type, static_cast<std::uint32_t>(width), static_cast<std::uint32_t>(height)};
    out.pitch_bytes = row_bytes;
    if (t.find(kTileWidth) != nullptr) {
        const std::uint64_t tw = t.scalar(kTileWidth, 0), th = t.scalar(kTileLength, 0, true);
        if (tw == 0 || th == 0) return Status::unsupported_format;
        std::uint64_t tile_bytes = 0;
        if (!host::mul(tw, th, tile_bytes) || !host::mul(tile_bytes, sample, tile_bytes)) return Status::size_overflow;
        const std::uint64_t across = host::ceil_div(width, tw), down = host::ceil_div(height, th);
        std::uint64_t tiles = 0;
        if (!host::mul(across, down, tiles) || tiles > UINT32_MAX) return Status::size_overflow;
        const Entry *oe = t.find(kTileOffsets), *ce = t.find(kTileBytes);
        if (oe == nullptr || ce == nullptr) return Status::unsupported_format;
        const std::vector<std::uint64_t> offsets = t.values(*oe), counts = t.values(*ce);
        if (offsets.size() != tiles || counts.size() != tiles) return Status::unsupported_format;
        out.pixels.resize(static_cast<std::size_t>(total));
        const Status st = parallel_ranges(tiles, t.size(), [&](std::uint64_t begin, std::uint64_t end) {
            std::vector<std::uint8_t> chunk(static_cast<std::size_t>(tile_bytes));
            for (std::uint64_t k = begin; k < end; ++k) {
                decode_chunk(t, offsets[k], counts[k], deflate, horizontal, sample, tw, chunk);
                const std::uint64_t x = (k % across) * tw, y = (k / across) * th;
                const std::uint64_t cw = std::min(tw, width - x), ch = std::min(th, height - y);
                for (std::uint64_t r = 0; r < ch; ++r)
                    std::memcpy(out.pixels.data() + (y + r) * row_bytes + x * sample,
                                chunk.data() + r * tw * sample, static_cast<std::size_t>(cw * sample));
            }
        });
        if (st != Status::success) return st;
    } else {
        const std::uint64_t rps = std::min<std::uint64_t>(t.scalar(kRowsPerStrip, UINT32_MAX), height);
        if (rps == 0) return Status::unsupported_format;
        const std::uint64_t strips = host::ceil_div(height, rps);
        const Entry *oe = t.find(kStripOffsets), *ce = t.find(kStripBytes);
        if (oe == nullptr || ce == nullptr) return Status::unsupported_format;
        const std::vector<std::uint64_t> offsets = t.values(*oe), counts = t.values(*ce);
        if (offsets.size() != strips || counts.size() != strips) return Status::unsupported_format;
        out.pixels.resize(static_cast<std::size_t>(total));
        const Status st = parallel_ranges(strips, t.size(), [&](std::uint64_t begin, std::uint64_t end) {
            std::vector<std::uint8_t> chunk;
            for (std::uint64_t k = begin; k < end; ++k) {
                const std::uint64_t rows = std::min(rps, height - k * rps);
                chunk.resize(static_cast<std::size_t>(rows * row_bytes));
                decode_chunk(t, offsets[k], counts[k], deflate, horizontal, sample, width, chunk);
                std::memcpy(out.pixels.data() + k * rps * row_bytes, chunk.data(), chunk.size());
            }
        });
        if (st != Status::success) return st;
    }
    img = std::move(out);
    return Status::success;
}

// The writer: little-endian, tiles, the directory after the data.
class Writer {
  public:
    Writer(std::FILE *f, bool big) : f_(f), big_(big) {}

    bool put(const void *p, std::size_t n) {
        if (n != 0 && std::fwrite(p, 1, n, f_) != n) return false;
        pos_ += n;
        return true;
    }
    bool put_le(std::uint64_t v, unsigned n) {
        std::uint8_t b[8];
        for (unsigned k = 0; k < n; ++k) b[k] = static_cast<std::uint8_t>(v >> (8 * k));
        return put(b, n);
    }
    std::uint64_t pos() const noexcept { return pos_; }
    bool big() const noexcept { return big_; }

  private:
    std::FILE *f_;
    bool big_;
    std::uint64_t pos_ = 0;
};

struct OutEntry {
    std::uint16_t tag, type;
    std::vector<std::uint64_t> values;
};

Status write_file(std::FILE *f, ConstImageView image, const TiffWriteOptions &o, std::uint64_t tile_bytes,
                  std::uint64_t across, std::uint64_t down) {
    Writer w(f, o.bigtiff);
    const unsigned sample = image.image.type == PixelType::uint16 ? 2 : 4;
    const bool deflate = o.compression == TiffCompression::deflate;
    // Header; the directory offset is patched at the end.
    if (!w.put(o.bigtiff ? "II\x2b\x00\x08\x00\x00\x00" : "II\x2a\x00", o.bigtiff ? 8 : 4) ||
        !w.put_le(0, o.bigtiff ? 8 : 4))
        return Status::io_error;
    std::vector<std::uint64_t> offsets, counts;
    std::vector<std::uint8_t> tile(static_cast<std::size_t>(tile_bytes)), packed;
    const auto *src = static_cast<const std::uint8_t *>(image.data);
    for (std::uint64_t ty = 0; ty < down; ++ty)
        for (std::uint64_t tx = 0; tx < across; ++tx) {
            std::fill(tile.begin(), tile.end(), std::uint8_t{0});
            const std::uint64_t x = tx * o.tile_width, y = ty * o.tile_height;
            const std::uint64_t cw = std::min<std::uint64_t>(o.tile_width, image.image.width - x);
            const std::uint64_t ch = std::min<std::uint64_t>(o.tile_height, image.image.height - y);
            for (std::uint64_t r = 0; r < ch; ++r)
                std::memcpy(tile.data() + r * o.tile_width * sample, src + (y + r) * image.pitch_bytes + x * sample,
                            static_cast<std::size_t>(cw * sample));
            const std::uint8_t *data = tile.data();
            std::uint64_t n = tile_bytes;
            if (deflate) {
                if (o.horizontal_predictor) {
                    for (std::uint64_t r = 0; r < o.tile_height; ++r) {
                        std::uint8_t *row = tile.data() + r * o.tile_width * 2;
                        for (std::uint64_t c = o.tile_width - 1; c > 0; --c) {
                            std::uint16_t a = 0, b = 0;
                            std::memcpy(&a, row + 2 * c, 2);
                            std::memcpy(&b, row + 2 * (c - 1), 2);
                            a = static_cast<std::uint16_t>(a - b);
                            std::memcpy(row + 2 * c, &a, 2);
                        }
                    }
                }
                uLongf out_len = compressBound(static_cast<uLong>(tile_bytes));
                packed.resize(out_len);
                if (compress2(packed.data(), &out_len, tile.data(), static_cast<uLong>(tile_bytes),
                              static_cast<int>(o.deflate_level)) != Z_OK)
                    return Status::io_error;
                data = packed.data();
                n = out_len;
            }
            offsets.push_back(w.pos());
            counts.push_back(n);
            if (!w.put(data, static_cast<std::size_t>(n))) return Status::io_error;
        }
    if (w.pos() % 2 != 0 && !w.put("", 1)) return Status::io_error; // word-aligned directory
    const std::uint16_t offset_type = o.bigtiff ? kLong8 : kLong;
    if (!o.bigtiff && w.pos() > UINT32_MAX) return Status::io_error; // classic TIFF offsets are 32-bit
    std::vector<OutEntry> entries = {
        {kWidth, kLong, {image.image.width}},
        {kLength, kLong, {image.image.height}},
        {kBits, kShort, {sample * 8u}},
        {kCompression, kShort, {deflate ? 8u : 1u}},
        {kPhotometric, kShort, {1}},
        {kOrientation, kShort, {1}},
        {kSamples, kShort, {1}},
        {kPlanar, kShort, {1}},
    };
    if (deflate) entries.push_back({kPredictor, kShort, {o.horizontal_predictor ? 2u : 1u}});
    entries.push_back({kTileWidth, kLong, {o.tile_width}});
    entries.push_back({kTileLength, kLong, {o.tile_height}});
    entries.push_back({kTileOffsets, offset_type, offsets});
    entries.push_back({kTileBytes, offset_type, counts});
    entries.push_back({kSampleFormat, kShort, {image.image.type == PixelType::uint16 ? 1u : 3u}});
    const std::uint64_t inline_bytes = o.bigtiff ? 8 : 4, entry_bytes = o.bigtiff ? 20 : 12;
    const std::uint64_t ifd = w.pos();
    // Arrays that do not fit an entry follow the directory.
    std::uint64_t extra = ifd + (o.bigtiff ? 8 : 2) + entries.size() * entry_bytes + (o.bigtiff ? 8 : 4);
    std::vector<std::uint64_t> at(entries.size(), 0);
    for (std::size_t k = 0; k < entries.size(); ++k) {
        const std::uint64_t bytes = entries[k].values.size() * type_size(entries[k].type);
        if (bytes > inline_bytes) {
            at[k] = extra;
            extra += bytes;
        }
    }
    if (!o.bigtiff && extra > UINT32_MAX) return Status::io_error;
    if (!w.put_le(entries.size(), o.bigtiff ? 8 : 2)) return Status::io_error;
    for (std::size_t k = 0; k < entries.size(); ++k) {
        const OutEntry &e = entries[k];
        const unsigned n = static_cast<unsigned>(type_size(e.type));
        if (!w.put_le(e.tag, 2) || !w.put_le(e.type, 2) || !w.put_le(e.values.size(), o.bigtiff ? 8 : 4))
            return Status::io_error;
        if (at[k] != 0) {
            if (!w.put_le(at[k], static_cast<unsigned>(inline_bytes))) return Status::io_error;
        } else {
            std::uint64_t used = 0;
            for (std::uint64_t v : e.values) {
                if (!w.put_le(v, n)) return Status::io_error;
                used += n;
            }
            if (!w.put_le(0, static_cast<unsigned>(inline_bytes - used))) return Status::io_error;
        }
    }
    if (!w.put_le(0, o.bigtiff ? 8 : 4)) return Status::io_error; // no next directory
    for (std::size_t k = 0; k < entries.size(); ++k) {
        if (at[k] == 0) continue;
        const unsigned n = static_cast<unsigned>(type_size(entries[k].type));
        for (std::uint64_t v : entries[k].values)
            if (!w.put_le(v, n)) return Status::io_error;
    }
    // Patch the directory offset into the header.
    std::uint8_t b[8];
    for (unsigned k = 0; k < 8; ++k) b[k] = static_cast<std::uint8_t>(ifd >> (8 * k));
    if (std::fseek(f, o.bigtiff ? 8 : 4, SEEK_SET) != 0 || std::fwrite(b, 1, o.bigtiff ? 8 : 4, f) != (o.bigtiff ? 8u : 4u))
        return Status::io_error;
    return Status::success;
}

} // namespace

Status read_tiff(const char *path, TiffImage &image, const TiffReadLimits &limits) noexcept {
    if (path == nullptr || limits.max_width == 0 || limits.max_height == 0 || limits.max_decoded_bytes > max_bytes)
        return Status::invalid_argument;
    try {
        std::vector<std::uint8_t> bytes;
        if (!read_file(path, bytes)) return Status::io_error;
        const TiffFile t(std::move(bytes));
        return decode(t, limits, image);
    } catch (const Malformed &) {
        return Status::io_error;
    } catch (const Unsupported &) {
        return Status::unsupported_format;
    } catch (const std::bad_alloc &) {
        return Status::resource_limit;
    } catch (const std::length_error &) {
        return Status::resource_limit;
    }
}

Status write_tiff(const char *path, ConstImageView image, const TiffWriteOptions &options) noexcept {
    if (path == nullptr) return Status::invalid_argument;
    // Option rules of reference tiff.cpp:199-223.
    if (const Status st = host::check_view(image); st != Status::success) return st;
    if (options.compression != TiffCompression::none && options.compression != TiffCompression::deflate)
        return Status::invalid_argument;
    if (options.tile_width == 0 || options.tile_height == 0 || options.tile_width % 16 != 0 ||
        options.tile_height % 16 != 0)
        return Status::invalid_argument;
    if (options.horizontal_predictor &&
        (image.image.type != PixelType::uint16 || options.compression != TiffCompression::deflate))
        return Status::invalid_argument;
    if (options.compression == TiffCompression::deflate && (options.deflate_level < 1 || options.deflate_level > 9))
        return Status::invalid_argument;
    const std::uint64_t sample = image.image.type == PixelType::uint16 ? 2 : 4;
    std::uint64_t tile_bytes = 0, tiles = 0;
    if (!host::mul(options.tile_width, options.tile_height, tile_bytes) || !host::mul(tile_bytes, sample, tile_bytes) ||
        tile_bytes > UINT32_MAX)
        return Status::size_overflow;
    const std::uint64_t across = host::ceil_div(image.image.width, options.tile_width);
    const std::uint64_t down = host::ceil_div(image.image.height, options.tile_height);
    if (!host::mul(across, down, tiles) || tiles > UINT32_MAX) return Status::size_overflow;
    std::FILE *f = std::fopen(path, "wb");
    if (f == nullptr) return Status::io_error;
    Status st = Status::io_error;
    try {
        st = write_file(f, image, options, tile_bytes, across, down);
    } catch (const std::bad_alloc &) {
        st = Status::resource_limit;
    } catch (const std::length_error &) {
        st = Status::resource_limit;
    }
    if (std::fclose(f) != 0 && st == Status::success) st = Status::io_error;
    if (st != Status::success) std::remove(path);
    return st;
}

} // namespace tilestream::io
