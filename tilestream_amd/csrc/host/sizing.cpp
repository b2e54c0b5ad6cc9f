// Checked size queries (reference src/sizing.cpp:9-109; docs/format.md
// "Checked output bound"; docs/api.md "Workspace derivation").
#include "checked.h"

#include <tilestream/format.h>

namespace tilestream {

Status image_span(const ImageSpec &image, std::uint64_t pitch_bytes, std::uint64_t &bytes) noexcept {
    std::uint64_t s = 0;
    if (const Status st = host::sample_size(image, s); st != Status::success) return st;
    std::uint64_t row = 0;
    if (!host::mul(image.width, s, row) || pitch_bytes > max_bytes) return Status::size_overflow;
    if (pitch_bytes < row || pitch_bytes % s != 0) return Status::invalid_argument;
    std::uint64_t span = 0;
    if (!host::mul(pitch_bytes, image.height - 1, span) || !host::add(span, row, span))
        return Status::size_overflow;
    bytes = span;
    return Status::success;
}

// B = 64 + 32 N + W H s + pad, pad = 2 only for uint16 with odd W and odd H.
Status compress_bound(const ImageSpec &image, TileShape tile, std::uint64_t &bytes) noexcept {
    std::uint64_t s = 0;
    if (const Status st = host::sample_size(image, s); st != Status::success) return st;
    if (!host::allowed_tile(tile)) return Status::invalid_argument;
    std::uint64_t n = 0, index = 0, raw = 0, b = 0;
    const std::uint64_t pad = (s == 2 && (image.width & 1) && (image.height & 1)) ? 2 : 0;
    if (!host::mul(host::ceil_div(image.width, tile.width), host::ceil_div(image.height, tile.height), n) ||
        !host::mul(n, format::index_entry_bytes, index) || !host::mul(image.width, image.height, raw) ||
        !host::mul(raw, s, raw) || !host::add(format::header_bytes, index, b) ||
        !host::add(b, raw, b) || !host::add(b, pad, b))
        return Status::size_overflow;
    bytes = b;
    return Status::success;
}

Status workspace_size(const WorkspaceSpec &spec, std::uint64_t &bytes) noexcept {
    host::WorkspaceLayout layout;
    const Status st = host::workspace_layout(spec, layout);
    if (st == Status::success) bytes = layout.total;
    return st;
}

namespace host {

Status workspace_layout(const WorkspaceSpec &spec, WorkspaceLayout &layout) noexcept {
    if (spec.image_count == 0) {
        if (spec.tile_count != 0 || spec.scan_temp_bytes != 0) return Status::invalid_argument;
        layout = {};
        return Status::success;
    }
    if (!allowed_tile(spec.tile) || spec.tile_count < spec.image_count) return Status::invalid_argument;
    if (spec.image_count > UINT32_MAX || spec.tile_count > max_window_tiles) return Status::resource_limit;
    WorkspaceLayout out;
    std::uint64_t cursor = 0;
    // Each region starts on a 256-byte boundary and owns its own padding.
    const auto region = [&cursor](std::uint64_t count, std::uint64_t size, std::uint64_t &at) {
        std::uint64_t bytes = 0;
        at = cursor;
        return mul(count, size, bytes) && round_up(bytes, workspace_alignment, bytes) &&
               add(cursor, bytes, cursor);
    };
    if (!region(spec.tile_count, 4096 * 4, out.residuals) ||
        !region(spec.tile_count, std::uint64_t{spec.tile.height} * format::segment_header_bytes, out.rows) ||
        !region(spec.tile_count, 32, out.tiles) || !region(spec.tile_count, 8, out.lengths) ||
        !region(spec.tile_count + 1, 8, out.offsets) || !region(spec.image_count, 16, out.images) ||
        !region(spec.scan_temp_bytes, 1, out.temp))
        return Status::size_overflow;
    out.total = cursor;
    layout = out;
    return Status::success;
}

} // namespace host
} // namespace tilestream
