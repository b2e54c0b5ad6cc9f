// TIFF import and export for the file tier (reference include/tilestream/tiff.h:1-42,
// same names, options and statuses).  The reference adapts libtiff; this is
// a self-contained reader and writer over zlib for the subset the codec
// needs: one image directory, one uint16 or float32 sample per pixel
// (min-is-black, top-left, chunky), tiled or stripped, uncompressed or
// Deflate with an optional horizontal predictor; classic TIFF or BigTIFF in
// either byte order on input, little-endian on output.
#pragma once

#include <tilestream/sizing.h>
#include <tilestream/types.h>

#include <cstdint>
#include <vector>

namespace tilestream::io {

enum class TiffCompression : std::uint32_t { none, deflate };

struct TiffReadLimits {
    std::uint32_t max_width = UINT32_MAX;
    std::uint32_t max_height = UINT32_MAX;
    std::uint64_t max_decoded_bytes = max_bytes;
};

struct TiffImage {
    ImageSpec image;
    std::uint64_t pitch_bytes = 0; // width x sample bytes: rows are packed
    std::vector<std::uint8_t> pixels;
};

struct TiffWriteOptions {
    bool bigtiff = false;
    TiffCompression compression = TiffCompression::none;
    bool horizontal_predictor = false; // uint16 with Deflate only
    std::uint32_t tile_width = 128;    // multiples of 16
    std::uint32_t tile_height = 32;
    std::uint32_t deflate_level = 6;   // 1..9
};

// Statuses: invalid_argument (null path, zero limits), io_error (cannot
// open, not a TIFF, truncated data, a tile or strip that decodes short),
// unsupported_format (anything outside the subset above), size_overflow,
// resource_limit (beyond the limits or out of memory).  `image` is left
// unchanged on failure.
[[nodiscard]] Status read_tiff(const char *path, TiffImage &image, const TiffReadLimits &limits = {}) noexcept;
// Writes tiles of tile_width x tile_height, zero-padded at the image edges.
// A failed write removes the file.
[[nodiscard]] Status write_tiff(const char *path, ConstImageView image, const TiffWriteOptions &options = {}) noexcept;

} // namespace tilestream::io
