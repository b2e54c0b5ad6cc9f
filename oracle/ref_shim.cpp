// C-ABI shim over the reference CPU codec, compiled together with the
// reference's own sources (never copied into this repository) by
// oracle/Makefile into oracle/_ref/libtsref.so.  TEST INFRASTRUCTURE ONLY:
// used by tests/ to pin the oracle, and by bench.py as the "reference" CPU
// baseline.  Wraps reference include/tilestream/{sizing,api,cpu}.h.
#include <tilestream/api.h>
#include <tilestream/cpu.h>
#include <tilestream/sizing.h>

#include <cstdint>

using namespace tilestream;

extern "C" {

int ref_compress_bound(std::uint32_t type, std::uint32_t width, std::uint32_t height,
                       std::uint32_t tile_width, std::uint32_t tile_height, std::uint64_t *bytes) {
    return static_cast<int>(compress_bound({static_cast<PixelType>(type), width, height},
                                           {static_cast<std::uint16_t>(tile_width),
                                            static_cast<std::uint16_t>(tile_height)},
                                           *bytes));
}

int ref_image_span(std::uint32_t type, std::uint32_t width, std::uint32_t height,
                   std::uint64_t pitch, std::uint64_t *bytes) {
    return static_cast<int>(
        image_span({static_cast<PixelType>(type), width, height}, pitch, *bytes));
}

int ref_workspace_size(std::uint64_t images, std::uint64_t tiles, std::uint32_t tile_width,
                       std::uint32_t tile_height, std::uint64_t temp, std::uint64_t *bytes) {
    WorkspaceSpec spec{images, tiles,
                       {static_cast<std::uint16_t>(tile_width),
                        static_cast<std::uint16_t>(tile_height)},
                       temp};
    return static_cast<int>(workspace_size(spec, *bytes));
}

// Returns the parser status; on success fills tile and segment counts.
int ref_parse(const std::uint8_t *data, std::uint64_t size, std::uint64_t limit_file,
              std::uint64_t limit_tiles, std::uint64_t limit_meta, std::uint64_t *tiles,
              std::uint64_t *segments) {
    ImageMetadata metadata;
    const auto status =
        parse_metadata({data, size}, {limit_file, limit_tiles, limit_meta}, metadata);
    if (status == Status::success) {
        *tiles = metadata.tiles.size();
        *segments = metadata.segments.size();
    }
    return static_cast<int>(status);
}

int ref_compress(std::uint32_t type, std::uint32_t width, std::uint32_t height,
                 std::uint64_t pitch, const void *pixels, std::uint64_t capacity_in,
                 std::uint32_t tile_width, std::uint32_t tile_height, std::uint8_t *out,
                 std::uint64_t capacity, std::uint64_t *written) {
    ConstImageView view{pixels, {static_cast<PixelType>(type), width, height}, pitch, capacity_in};
    return static_cast<int>(cpu::compress(view,
                                          {static_cast<std::uint16_t>(tile_width),
                                           static_cast<std::uint16_t>(tile_height)},
                                          {out, capacity}, *written));
}

int ref_decompress(const std::uint8_t *data, std::uint64_t size, std::uint32_t type,
                   std::uint32_t width, std::uint32_t height, void *out, std::uint64_t pitch,
                   std::uint64_t capacity) {
    ImageView view{out, {static_cast<PixelType>(type), width, height}, pitch, capacity};
    return static_cast<int>(cpu::decompress({data, size}, view));
}

} // extern "C"
