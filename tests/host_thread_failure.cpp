// Parsing with host thread creation failing after `n` workers started (an
// executable-level pthread_create that fails with EAGAIN): the parse must
// still return the same metadata as with every thread available, for n = 0,
// 1, 2, ...  Built and run by test_host_sanitizers.py (without a sanitizer,
// which would intercept pthread_create itself).  TEST INFRASTRUCTURE ONLY.
#include <tilestream/api.h>
#include <tilestream/cpu.h>
#include <tilestream/sizing.h>

#include <dlfcn.h>
#include <pthread.h>

#include <atomic>
#include <cerrno>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

static std::atomic<int> allowed{1 << 30};

extern "C" int pthread_create(pthread_t *thread, const pthread_attr_t *attr, void *(*start)(void *), void *arg) {
    using Create = int (*)(pthread_t *, const pthread_attr_t *, void *(*)(void *), void *);
    static const Create real = reinterpret_cast<Create>(dlsym(RTLD_NEXT, "pthread_create"));
    if (allowed.fetch_sub(1) <= 0) return EAGAIN;
    return real(thread, attr, start, arg);
}

using namespace tilestream;

static bool same(const ImageMetadata &a, const ImageMetadata &b) {
    if (a.tiles.size() != b.tiles.size() || a.segments.size() != b.segments.size()) return false;
    return std::memcmp(a.tiles.data(), b.tiles.data(), a.tiles.size() * sizeof(TileMetadata)) == 0 &&
           std::memcmp(a.segments.data(), b.segments.data(), a.segments.size() * sizeof(SegmentMetadata)) == 0;
}

int main() {
    // A 4096 x 4096 uint16 image with noisy rows: an archive of about 20 MiB,
    // so the payload walk is split over several threads (one per 4 MiB).
    const std::uint32_t w = 4096, h = 4096;
    std::vector<std::uint16_t> px(std::size_t{w} * h);
    std::mt19937 rng(3);
    for (std::size_t i = 0; i < px.size(); ++i) px[i] = static_cast<std::uint16_t>((i % w) * 7 + (rng() & 1023));
    const ImageSpec spec{PixelType::uint16, w, h};
    std::uint64_t bound = 0, written = 0;
    if (compress_bound(spec, TileShape{64, 64}, bound) != Status::success) return 1;
    std::vector<std::uint8_t> file(bound);
    if (cpu::compress(ConstImageView{px.data(), spec, w * 2ull, px.size() * 2}, TileShape{64, 64},
                      MutableBytes{file.data(), file.size()}, written) != Status::success)
        return 2;
    file.resize(written);
    ImageMetadata want;
    if (parse_metadata(ConstBytes{file.data(), file.size()}, ParseLimits{}, want) != Status::success) return 3;
    int bad = 0;
    for (int n = 0; n <= 4; ++n) {
        allowed = n;
        ImageMetadata got;
        const Status st = parse_metadata(ConstBytes{file.data(), file.size()}, ParseLimits{}, got);
        allowed = 1 << 30;
        if (st != Status::success || !same(got, want)) {
            std::printf("threads failing after %d: status %u\n", n, static_cast<unsigned>(st));
            ++bad;
        }
    }
    std::printf("archive %zu bytes, %zu tiles\nrc %d\n", file.size(), want.tiles.size(), bad);
    return bad == 0 ? 0 : 4;
}
