// Host code under ThreadSanitizer / AddressSanitizer + UBSan (built and run by
// test_host_sanitizers.py with g++): the CPU encoder, the parser (several MiB
// per archive, so its validation runs on several threads), the CPU decoder,
// and parses of archives with one flipped bit in the index or the payloads,
// which must fail cleanly or succeed; TIFF export and import of the pixels.
#include <tilestream/api.h>
#include <tilestream/cpu.h>
#include <tilestream/sizing.h>
#include <tilestream/tiff.h>

#include <cstdio>
#include <cstdlib>
#include <string>
#include <cstring>
#include <random>
#include <vector>

using namespace tilestream;

static int run(PixelType type, std::uint32_t w, std::uint32_t h, TileShape tile, unsigned seed) {
    const std::uint32_t s = type == PixelType::uint16 ? 2 : 4;
    std::vector<std::uint8_t> px(std::size_t{w} * h * s);
    std::mt19937 rng(seed);
    // Smooth rows with small noise: predictive tiles with narrow widths.
    for (std::size_t i = 0; i < px.size(); ++i)
        px[i] = static_cast<std::uint8_t>(i % s == 0 ? ((i / s) % w) / 16 + (rng() & 3) : (i % s == 1 ? 3 : 0));
    ImageSpec spec{type, w, h};
    std::uint64_t bound = 0;
    if (compress_bound(spec, tile, bound) != Status::success) return 1;
    std::vector<std::uint8_t> file(bound);
    std::uint64_t written = 0;
    ConstImageView in{px.data(), spec, std::uint64_t{w} * s, px.size()};
    if (cpu::compress(in, tile, MutableBytes{file.data(), file.size()}, written) != Status::success) return 2;
    file.resize(written);
    ImageMetadata meta;
    if (parse_metadata(ConstBytes{file.data(), file.size()}, ParseLimits{}, meta) != Status::success) return 3;
    std::vector<std::uint8_t> out(px.size());
    ImageView ov{out.data(), spec, std::uint64_t{w} * s, out.size()};
    if (cpu::decompress(ConstBytes{file.data(), file.size()}, ov) != Status::success) return 4;
    if (out != px) return 5;
    // One flipped bit, alternately in the index and in the payloads.
    int failures = 0;
    for (int k = 0; k < 64; ++k) {
        std::vector<std::uint8_t> bad = file;
        const std::size_t index_end = 64 + 32 * meta.tiles.size();
        const std::size_t at = k % 2 == 0 ? 64 + std::size_t{rng()} % (index_end - 64)
                                          : index_end + std::size_t{rng()} % (bad.size() - index_end);
        bad[at] ^= static_cast<std::uint8_t>(1u << (k % 8));
        ImageMetadata m2;
        if (parse_metadata(ConstBytes{bad.data(), bad.size()}, ParseLimits{}, m2) != Status::success) ++failures;
    }
    // TIFF export and import of the decoded pixels (Deflate, predictor for uint16),
    // then a copy with one flipped bit must fail cleanly or succeed.
    const char *dir = std::getenv("TMPDIR");
    const std::string path = std::string(dir != nullptr ? dir : "/tmp") + "/host_sanitize_" + std::to_string(seed) + ".tif";
    io::TiffWriteOptions options;
    options.compression = io::TiffCompression::deflate;
    options.horizontal_predictor = type == PixelType::uint16;
    if (io::write_tiff(path.c_str(), ConstImageView{out.data(), spec, std::uint64_t{w} * s, out.size()}, options) !=
        Status::success)
        return 6;
    io::TiffImage tiff;
    if (io::read_tiff(path.c_str(), tiff) != Status::success || tiff.pixels != px) return 7;
    std::FILE *f = std::fopen(path.c_str(), "r+b");
    if (f == nullptr) return 8;
    std::fseek(f, 0, SEEK_END);
    const long tiff_bytes = std::ftell(f);
    std::fseek(f, tiff_bytes / 2, SEEK_SET);
    const int byte = std::fgetc(f);
    std::fseek(f, tiff_bytes / 2, SEEK_SET);
    std::fputc(byte ^ 0x10, f);
    std::fclose(f);
    io::TiffImage broken;
    const Status tiff_status = io::read_tiff(path.c_str(), broken);
    std::remove(path.c_str());
    std::printf("%s %ux%u tile %ux%u: %llu bytes, %d of 64 corruptions rejected; corrupted TIFF status %u\n",
                type == PixelType::uint16 ? "u16" : "f32", w, h, tile.width, tile.height,
                static_cast<unsigned long long>(written), failures, static_cast<unsigned>(tiff_status));
    return 0;
}

int main() {
    int rc = 0;
    rc |= run(PixelType::uint16, 4096, 2048, TileShape{64, 64}, 1);
    rc |= run(PixelType::float32, 2048, 2048, TileShape{128, 32}, 2);
    rc |= run(PixelType::uint16, 3001, 1777, TileShape{256, 16}, 3);
    std::printf("rc %d\n", rc);
    return rc;
}
