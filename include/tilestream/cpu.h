// Scalar host codec (reference include/tilestream/cpu.h:7-14).  Buffers are
// caller-owned host memory and must not overlap; failures leave outputs unchanged.
#pragma once

#include <tilestream/api.h>

namespace tilestream::cpu {

[[nodiscard]] Status compress(ConstImageView input, TileShape tile, MutableBytes output,
                              std::uint64_t &bytes_written) noexcept;
[[nodiscard]] Status decompress(ConstBytes input, ImageView output,
                                const ParseLimits &limits = {}) noexcept;

} // namespace tilestream::cpu
