// Parse hook for oracle/ref_fuzz_shim.cpp (compiled without the rename): keeps
// the bytes of the most recent parse the fuzz harness makes, then calls the
// reference parser.  TEST INFRASTRUCTURE ONLY.
#include <tilestream/api.h>

#include <cstdint>
#include <vector>

namespace tilestream {
Status ref_capturing_parse(ConstBytes source, const ParseLimits &limits,
                           ImageMetadata &metadata) noexcept;
}

namespace {
std::vector<std::uint8_t> last;
}

tilestream::Status tilestream::ref_capturing_parse(ConstBytes source, const ParseLimits &limits,
                                                   ImageMetadata &metadata) noexcept {
    try {
        const auto *p = static_cast<const std::uint8_t *>(source.data);
        last.assign(p, p + (p ? source.size : 0));
    } catch (...) {
        last.clear();
    }
    return parse_metadata(source, limits, metadata);
}

extern "C" const std::uint8_t *ref_fuzz_last_parsed(std::uint64_t *size) {
    *size = last.size();
    return last.data();
}
