// Checked 64-bit size arithmetic, view rules and little-endian byte access
// shared by the host side of the library.
#pragma once

#include <tilestream/sizing.h>

#include <cstdint>
#include <cstring>
#include <limits>

namespace tilestream::host {

// a + b, a * b and rounding up, each refused if the result or an operand
// exceeds L = max_bytes (reference src/checked_size.h:9-32).
[[nodiscard]] inline bool add(std::uint64_t a, std::uint64_t b, std::uint64_t &out) noexcept {
    if (a > max_bytes || b > max_bytes - a) return false;
    out = a + b;
    return true;
}
[[nodiscard]] inline bool mul(std::uint64_t a, std::uint64_t b, std::uint64_t &out) noexcept {
    if (a > max_bytes || b > max_bytes || (b != 0 && a > max_bytes / b)) return false;
    out = a * b;
    return true;
}
[[nodiscard]] inline bool round_up(std::uint64_t v, std::uint64_t a, std::uint64_t &out) noexcept {
    if (a == 0 || a > max_bytes || v > max_bytes) return false;
    const std::uint64_t r = v % a;
    return add(v, r == 0 ? 0 : a - r, out);
}

inline constexpr bool allowed_tile(TileShape t) noexcept {
    return (t.width == 128 && t.height == 32) || (t.width == 64 && t.height == 64) ||
           (t.width == 256 && t.height == 16);
}

inline constexpr std::uint64_t ceil_div(std::uint64_t n, std::uint64_t d) noexcept {
    return n / d + (n % d != 0);
}

// Bytes per sample; unsupported type reported before zero dimensions.
inline Status sample_size(const ImageSpec &image, std::uint64_t &bytes) noexcept {
    if (image.type != PixelType::uint16 && image.type != PixelType::float32)
        return Status::unsupported_type;
    if (image.width == 0 || image.height == 0) return Status::invalid_argument;
    bytes = image.type == PixelType::uint16 ? 2 : 4;
    return Status::success;
}

// A caller byte range [data, data + size) (reference src/cpu/views.h:10-23).
inline Status byte_range(const void *data, std::uint64_t size) noexcept {
    if (size > max_bytes) return Status::size_overflow;
    if (data == nullptr && size != 0) return Status::invalid_argument;
    if (size > std::numeric_limits<std::uintptr_t>::max() - reinterpret_cast<std::uintptr_t>(data))
        return Status::size_overflow;
    return Status::success;
}

inline bool overlaps(const void *a, std::uint64_t an, const void *b, std::uint64_t bn) noexcept {
    const auto x = reinterpret_cast<std::uintptr_t>(a);
    const auto y = reinterpret_cast<std::uintptr_t>(b);
    return an != 0 && bn != 0 && x < y + bn && y < x + an;
}

// Pixel view rule of docs/api.md "Images and views" (reference src/cpu/views.h:32-47).
inline Status check_view(const ConstImageView &view) noexcept {
    std::uint64_t span = 0;
    Status status = image_span(view.image, view.pitch_bytes, span);
    if (status != Status::success) return status;
    status = byte_range(view.data, view.capacity_bytes);
    if (status != Status::success) return status;
    const std::uintptr_t align = view.image.type == PixelType::uint16 ? 2 : 4;
    if (view.data == nullptr || reinterpret_cast<std::uintptr_t>(view.data) % align != 0)
        return Status::invalid_argument;
    return span <= view.capacity_bytes ? Status::success : Status::buffer_too_small;
}

inline ConstImageView as_const(const ImageView &v) noexcept {
    return {v.data, v.image, v.pitch_bytes, v.capacity_bytes};
}

template <typename T> inline T load_le(const std::uint8_t *p) noexcept {
    T v = 0;
    for (unsigned i = 0; i < sizeof(T); ++i) v |= static_cast<T>(T{p[i]} << (8 * i));
    return v;
}

template <typename T> inline void store_le(std::uint8_t *p, T v) noexcept {
    for (unsigned i = 0; i < sizeof(T); ++i) p[i] = static_cast<std::uint8_t>(v >> (8 * i));
}

inline bool zeros(const std::uint8_t *p, std::uint64_t n) noexcept {
    for (std::uint64_t i = 0; i < n; ++i)
        if (p[i] != 0) return false;
    return true;
}

// Byte offsets of the workspace regions (reference docs/api.md "Workspace derivation").
struct WorkspaceLayout {
    std::uint64_t residuals = 0, rows = 0, tiles = 0, lengths = 0, offsets = 0, images = 0,
                  temp = 0, total = 0;
};

[[nodiscard]] Status workspace_layout(const WorkspaceSpec &spec, WorkspaceLayout &layout) noexcept;

} // namespace tilestream::host
