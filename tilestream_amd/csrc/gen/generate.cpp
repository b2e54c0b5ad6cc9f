// Seeded synthetic inputs for the TileStream workloads (host C++, multithreaded).
//
// Every random number is a pure function of (seed, stream, counter), so the
// output does not depend on the thread count or on the order rows are filled.
//
//  * detector frame (uint16): background 1000 + 20 Gaussian blobs, amplitude
//    U[500, 20000], sigma U[3, 40] px, centres U(image); a blob contributes
//    only within 6 sigma of its centre in x and in y.  Pixel =
//    min(Poisson(lambda), 65535), Poisson drawn with Hoermann's transformed
//    rejection (PTRS, 1993) for lambda >= 10 and Knuth's product method below.
//  * field (float32 bit words): sum over octaves k = 0..5 of
//    (1000 / 2^k) sin(2 pi f_k (x cos t_k + y sin t_k) + p_k), f_k = 2^k / 512
//    cycles per pixel, random direction t_k and phase p_k, plus N(0, 0.5)
//    noise (Box-Muller); then 0.1% of words are replaced by one of five
//    special classes with equal probability: NaN with a random payload (random
//    sign, quiet or signalling), +inf, -inf, -0.0, a random subnormal.
//  * dark map: uint16 uniform integers in [90, 110].
//  * gain map: float32 N(1, 0.02).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {

constexpr double kPi = 3.14159265358979323846;

inline std::uint64_t mix(std::uint64_t z) {
    z += 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

struct Stream {
    std::uint64_t key;
    Stream(std::uint64_t seed, std::uint64_t stream) : key(mix(seed ^ mix(stream + 0x51ed27))) {}
    std::uint64_t bits(std::uint64_t counter) const {
        return mix(key + counter * 0xd2b74407b1ce6e93ULL);
    }
    // Uniform in [0, 1) with 53 random bits.
    double uniform(std::uint64_t counter) const {
        return static_cast<double>(bits(counter) >> 11) * 0x1.0p-53;
    }
};

template <typename Fn> void parallel_rows(std::uint32_t height, int threads, Fn fn) {
    if (threads <= 1 || height < 64) {
        fn(0u, height);
        return;
    }
    const auto n = static_cast<std::uint32_t>(std::min<int>(threads, 64));
    std::vector<std::thread> pool;
    const std::uint32_t chunk = (height + n - 1) / n;
    for (std::uint32_t t = 0; t < n; ++t) {
        const std::uint32_t begin = t * chunk;
        const std::uint32_t end = std::min(height, begin + chunk);
        if (begin >= end) break;
        pool.emplace_back([=] { fn(begin, end); });
    }
    for (auto &th : pool) th.join();
}

// Draws per pixel come from counters pixel * 64 + j.
struct PixelDraws {
    const Stream &s;
    std::uint64_t base;
    std::uint32_t used = 0;
    double next() { return s.uniform(base + (used++ & 63)); }
};

// ln(k!) through the reentrant lgamma (no write to the global signgam).
double log_factorial(double k) {
    int sign = 0;
    return ::lgamma_r(k + 1.0, &sign);
}

// Poisson variate.  PTRS: W. Hoermann, "The transformed rejection method for
// generating Poisson random variables", Insurance: Mathematics and Economics
// 12 (1993) 39-45.  Knuth's multiplication method for small means.
std::uint32_t poisson(double lambda, PixelDraws &d) {
    if (lambda < 10.0) {
        const double limit = std::exp(-lambda);
        double product = d.next();
        std::uint32_t k = 0;
        while (product > limit && k < 1000) {
            product *= d.next();
            ++k;
        }
        return k;
    }
    const double slam = std::sqrt(lambda);
    const double loglam = std::log(lambda);
    const double b = 0.931 + 2.53 * slam;
    const double a = -0.059 + 0.02483 * b;
    const double invalpha = 1.1239 + 1.1328 / (b - 3.4);
    const double vr = 0.9277 - 3.6224 / (b - 2.0);
    for (int attempt = 0; attempt < 30; ++attempt) {
        const double u = d.next() - 0.5;
        const double v = d.next();
        const double us = 0.5 - std::fabs(u);
        const double k = std::floor((2.0 * a / us + b) * u + lambda + 0.43);
        if (us >= 0.07 && v <= vr) return static_cast<std::uint32_t>(std::max(k, 0.0));
        if (k < 0.0 || (us < 0.013 && v > us)) continue;
        if (std::log(v) + std::log(invalpha) - std::log(a / (us * us) + b) <=
            -lambda + k * loglam - log_factorial(k))
            return static_cast<std::uint32_t>(k);
    }
    return static_cast<std::uint32_t>(std::llround(lambda));
}

struct Blob {
    double amplitude, sigma, cx, cy;
};

} // namespace

extern "C" {

void tsg_detector_u16(std::uint64_t seed, std::uint32_t width, std::uint32_t height,
                      std::uint16_t *out, std::uint64_t pitch_elems, int threads) {
    const Stream blobs(seed, 1), noise(seed, 2);
    std::vector<Blob> list(20);
    for (std::uint64_t k = 0; k < list.size(); ++k) {
        list[k] = {500.0 + 19500.0 * blobs.uniform(4 * k), 3.0 + 37.0 * blobs.uniform(4 * k + 1),
                   width * blobs.uniform(4 * k + 2), height * blobs.uniform(4 * k + 3)};
    }
    parallel_rows(height, threads, [&](std::uint32_t y0, std::uint32_t y1) {
        std::vector<double> lambda(width);
        for (std::uint32_t y = y0; y < y1; ++y) {
            std::fill(lambda.begin(), lambda.end(), 1000.0);
            for (const auto &blob : list) {
                const double reach = 6.0 * blob.sigma;
                const double dy = y - blob.cy;
                if (std::fabs(dy) > reach) continue;
                const double inv = 1.0 / (2.0 * blob.sigma * blob.sigma);
                const double ey = blob.amplitude * std::exp(-dy * dy * inv);
                const auto xa = static_cast<std::int64_t>(std::ceil(blob.cx - reach));
                const auto xb = static_cast<std::int64_t>(std::floor(blob.cx + reach));
                for (std::int64_t x = std::max<std::int64_t>(xa, 0);
                     x <= std::min<std::int64_t>(xb, width - 1); ++x) {
                    const double dx = x - blob.cx;
                    lambda[x] += ey * std::exp(-dx * dx * inv);
                }
            }
            std::uint16_t *row = out + y * pitch_elems;
            for (std::uint32_t x = 0; x < width; ++x) {
                PixelDraws d{noise, (std::uint64_t{y} * width + x) * 64};
                row[x] = static_cast<std::uint16_t>(std::min<std::uint32_t>(poisson(lambda[x], d), 65535));
            }
        }
    });
}

void tsg_field_f32(std::uint64_t seed, std::uint32_t width, std::uint32_t height,
                   std::uint32_t *out, std::uint64_t pitch_elems, int threads) {
    const Stream waves(seed, 3), noise(seed, 4), special(seed, 5);
    double amp[6], fx[6], fy[6], phase[6];
    for (int k = 0; k < 6; ++k) {
        const double theta = 2.0 * kPi * waves.uniform(2 * k);
        const double f = std::ldexp(1.0, k) / 512.0;
        amp[k] = 1000.0 / std::ldexp(1.0, k);
        fx[k] = 2.0 * kPi * f * std::cos(theta);
        fy[k] = 2.0 * kPi * f * std::sin(theta);
        phase[k] = 2.0 * kPi * waves.uniform(2 * k + 1);
    }
    parallel_rows(height, threads, [&](std::uint32_t y0, std::uint32_t y1) {
        for (std::uint32_t y = y0; y < y1; ++y) {
            std::uint32_t *row = out + y * pitch_elems;
            for (std::uint32_t x = 0; x < width; ++x) {
                const std::uint64_t i = std::uint64_t{y} * width + x;
                double v = 0.0;
                for (int k = 0; k < 6; ++k) v += amp[k] * std::sin(fx[k] * x + fy[k] * y + phase[k]);
                const double u1 = 1.0 - noise.uniform(2 * i), u2 = noise.uniform(2 * i + 1);
                v += 0.5 * std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kPi * u2);
                float f = static_cast<float>(v);
                std::uint32_t bits;
                std::memcpy(&bits, &f, 4);
                if (special.uniform(3 * i) < 0.001) {
                    const std::uint64_t r = special.bits(3 * i + 2);
                    const std::uint32_t sign = static_cast<std::uint32_t>(r >> 63) << 31;
                    switch (static_cast<int>(5.0 * special.uniform(3 * i + 1))) {
                    case 0: { // NaN: quiet or signalling, payload never all zero
                        const std::uint32_t quiet = static_cast<std::uint32_t>(r >> 62) & 1u;
                        const std::uint32_t payload =
                            1u + static_cast<std::uint32_t>(r % ((1u << 22) - 1));
                        bits = sign | 0x7f800000u | (quiet << 22) | payload;
                        break;
                    }
                    case 1: bits = 0x7f800000u; break;
                    case 2: bits = 0xff800000u; break;
                    case 3: bits = 0x80000000u; break;
                    default: bits = sign | (1u + static_cast<std::uint32_t>(r % ((1u << 23) - 1))); break;
                    }
                }
                row[x] = bits;
            }
        }
    });
}

void tsg_dark_u16(std::uint64_t seed, std::uint32_t width, std::uint32_t height,
                  std::uint16_t *out, int threads) {
    const Stream s(seed, 6);
    parallel_rows(height, threads, [&](std::uint32_t y0, std::uint32_t y1) {
        for (std::uint32_t y = y0; y < y1; ++y)
            for (std::uint32_t x = 0; x < width; ++x)
                out[std::uint64_t{y} * width + x] = static_cast<std::uint16_t>(
                    90 + s.bits(std::uint64_t{y} * width + x) % 21);
    });
}

void tsg_gain_f32(std::uint64_t seed, std::uint32_t width, std::uint32_t height, float *out,
                  int threads) {
    const Stream s(seed, 7);
    parallel_rows(height, threads, [&](std::uint32_t y0, std::uint32_t y1) {
        for (std::uint32_t y = y0; y < y1; ++y)
            for (std::uint32_t x = 0; x < width; ++x) {
                const std::uint64_t i = std::uint64_t{y} * width + x;
                const double u1 = 1.0 - s.uniform(2 * i), u2 = s.uniform(2 * i + 1);
                out[i] = static_cast<float>(
                    1.0 + 0.02 * std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kPi * u2));
            }
    });
}

// Uniform integers / reals for workload layouts (image sizes, ROI rectangles).
std::uint64_t tsg_bits(std::uint64_t seed, std::uint64_t stream, std::uint64_t counter) {
    return Stream(seed, stream).bits(counter);
}

} // extern "C"
