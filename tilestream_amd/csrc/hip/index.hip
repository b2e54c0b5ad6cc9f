// Resident tile table of a .tsc archive built on the device from the archive's
// own header and index, for archives already in device memory (for example a
// GPU encoder's output): no copy to the host and no host parse.
//
// The checks are the host parser's (reference src/cpu/parser.cpp, restated from
// docs/format.md "Limits and validation order"), in the same order for the
// header, so a rejection carries the parser's status; every later failure is
// malformed_data there too.  The header is checked by one thread; then one
// thread per tile checks its index entry and payload and writes its table
// entry, as the trusted decode path checks the entries and row descriptors of
// encoder output (reference src/cuda/decoder.cu:86-110,191-220), but for every
// tile of the archive and with the full payload rules:
//   index    mode, reserved bytes, the tile's geometry, payload within the
//            file, word multiple, at most the raw size, and payloads laid end
//            to end: tile 0 starts after the index, tile i where tile i - 1
//            ends, the last one ends the file (each tile reads its
//            predecessor's entry instead of a serial walk);
//   payload  RAW: exactly the raw size, padding zero; CONSTANT: one word,
//            uint16 high half zero; PREDICT_BITPACK: one descriptor per tile
//            row (seed range, width at most the sample bits, zero width for a
//            one-column tile, reserved bytes zero), packed words inside the
//            payload with the bits past the last residual zero, rows ending
//            exactly at the payload's end.
#include <hip/hip_runtime.h>

#include <tilestream_kernels.h>

#include <cstdint>

namespace {

template <typename T> using gptr = T __attribute__((address_space(1))) *;

enum : std::uint32_t {
    kUnsupportedType = 3,
    kUnsupportedFormat = 4,
    kMalformed = 5,
    kBufferTooSmall = 7,
};

struct ResultBits {
    std::uint32_t status;
    std::uint32_t reserved;
    std::uint64_t bytes;
};

// 32-bit word at a 4-byte aligned offset of the file.
__device__ inline std::uint32_t word(const std::uint8_t *src, std::uint64_t at) {
    return *reinterpret_cast<gptr<const std::uint32_t>>(reinterpret_cast<std::uintptr_t>(src + at));
}

__device__ inline std::uint64_t dword(const std::uint8_t *src, std::uint64_t at) {
    return word(src, at) | (std::uint64_t{word(src, at + 4)} << 32);
}

__device__ inline bool allowed_tile(std::uint32_t w, std::uint32_t h) {
    return (w == 128 && h == 32) || (w == 64 && h == 64) || (w == 256 && h == 16);
}

// The header (one thread).  Writes the record: the header's status (0 when
// it is valid and describes the expected image), no tile failures yet, and
// the file size.  A valid header bounds the tile count by the file size
// (the index lies inside the file), so the caller sizes the tables after it.
__global__ void index_header_kernel(tsk_index_args a) {
    tsk_index_record rec{0u, 0u, 0u};
    std::uint64_t size = a.capacity;
    if (a.producer != nullptr) {
        const ResultBits produced = *static_cast<const ResultBits *>(a.producer);
        size = produced.bytes;
        if (produced.status != 0) rec.header_status = produced.status;
        else if (size > a.capacity) rec.header_status = kBufferTooSmall;
    }
    const std::uint8_t *s = a.src;
    const auto check = [&]() -> std::uint32_t {
        if (size < 64) return kMalformed;
        if (word(s, 0) != 0x00435354u || word(s, 4) != 0x0a1a0a0du) return kMalformed;
        const std::uint32_t versions = word(s, 8), w12 = word(s, 12);
        if (versions != 1u || (w12 >> 24) != 0u) return kUnsupportedFormat; // major 1, minor 0, flags 0
        const std::uint32_t w28 = word(s, 28);
        if ((w12 & 0xffffu) != 64u || (w28 & 0xffffu) != 32u || (w28 >> 16) != 0u) return kMalformed;
        const std::uint32_t type = (w12 >> 16) & 0xffu;
        if (type != TSK_UINT16 && type != TSK_FLOAT32) return kUnsupportedType;
        const std::uint32_t width = word(s, 16), height = word(s, 20), tiles = word(s, 24);
        const std::uint32_t tw = tiles & 0xffffu, th = tiles >> 16;
        if (width == 0 || height == 0) return kMalformed;
        if (!allowed_tile(tw, th)) return kUnsupportedFormat;
        // A valid header of another image than the caller described.
        if (type != a.type || width != a.width || height != a.height || tw != a.tile_w || th != a.tile_h)
            return TSK_INDEX_OTHER_IMAGE;
        const std::uint64_t payload_at = 64u + 32u * a.tile_count;
        if (dword(s, 32) != a.tile_count || dword(s, 40) != 64u || dword(s, 48) != payload_at ||
            dword(s, 56) != size || size > a.bound || payload_at > size)
            return kMalformed;
        return 0u;
    };
    if (rec.header_status == 0) rec.header_status = check();
    rec.file_bytes = rec.header_status == 0 ? size : 0u;
    *a.record = rec;
}

// Checks tile i's entry and payload; fills its table entry and row offsets.
__device__ inline bool index_tile(const tsk_index_args &a, std::uint64_t i, std::uint64_t size) {
    const std::uint8_t *s = a.src;
    const std::uint32_t across = a.tile_columns;
    const std::uint32_t column = static_cast<std::uint32_t>(i % across), row = static_cast<std::uint32_t>(i / across);
    const std::uint32_t x = column * a.tile_w, y = row * a.tile_h;
    const std::uint32_t w = a.width - x < a.tile_w ? a.width - x : a.tile_w;
    const std::uint32_t h = a.height - y < a.tile_h ? a.height - y : a.tile_h;
    const bool u16 = a.type == TSK_UINT16;
    const std::uint64_t raw = std::uint64_t{w} * h * (u16 ? 2u : 4u), raw4 = (raw + 3u) / 4u * 4u;

    const std::uint64_t e = 64u + 32u * i;
    const std::uint32_t w0 = word(s, e), dims = word(s, e + 4);
    const std::uint64_t at = dword(s, e + 8), len = dword(s, e + 16), tail = dword(s, e + 24);
    const std::uint32_t mode = w0 & 0xffu;
    std::uint64_t expected = 64u + 32u * a.tile_count;
    if (i > 0) {
        const std::uint64_t prev_at = dword(s, e - 24), prev_len = dword(s, e - 16);
        if (prev_len > ~std::uint64_t{0} - prev_at) return false;
        expected = prev_at + prev_len;
    }
    if (mode > 2u || (w0 >> 8) != 0u || tail != 0u || dims != (w | (h << 16)) || at != expected || at % 4u != 0u ||
        at > size || len > size - at || len == 0u || len % 4u != 0u || len > raw4)
        return false;
    if (i + 1u == a.tile_count && at + len != size) return false;

    std::uint32_t rows = TSK_NO_ROWS;
    if (mode == 0u) {
        if (len != raw4) return false;
        if (raw4 != raw && (word(s, at + raw4 - 4u) >> 16) != 0u) return false; // uint16 padding half
    } else if (mode == 1u) {
        if (len != 4u || (u16 && (word(s, at) >> 16) != 0u)) return false;
    } else {
        const std::uint32_t widest = u16 ? 16u : 32u;
        std::uint64_t pos = 0;
        gptr<std::uint16_t> offsets =
            reinterpret_cast<gptr<std::uint16_t>>(reinterpret_cast<std::uintptr_t>(a.rows)) + i * a.tile_h;
        for (std::uint32_t r = 0; r < h; ++r) {
            if (len - pos < 8u) return false;
            const std::uint32_t seed = word(s, at + pos), wb = word(s, at + pos + 4u);
            const std::uint32_t b = wb & 0xffu;
            if ((u16 && seed > 0xffffu) || b > widest || (w == 1u && b != 0u) || (wb >> 8) != 0u) return false;
            offsets[r] = static_cast<std::uint16_t>(pos / 4u);
            pos += 8u;
            const std::uint64_t bits = std::uint64_t{w - 1u} * b, packed = (bits + 31u) / 32u * 4u;
            if (packed > len - pos) return false;
            if (bits % 32u != 0u && (word(s, at + pos + packed - 4u) >> (bits % 32u)) != 0u) return false;
            pos += packed;
        }
        if (pos != len) return false;
        rows = a.rows_base + static_cast<std::uint32_t>(i) * a.tile_h;
    }
    tsk_decode_task t;
    t.payload = s + at;
    t.column = column;
    t.row = row;
    t.image = 0;
    t.rows = rows;
    t.words = static_cast<std::uint32_t>(len / 4u);
    t.shape = TSK_TASK_SHAPE(mode, h, w);
    a.tasks[i] = t;
    return true;
}

__global__ void __launch_bounds__(256) index_tiles_kernel(tsk_index_args a) {
    const tsk_index_record rec = *a.record;
    if (rec.header_status != 0u) return;
    std::uint32_t failures = 0;
    const std::uint64_t stride = std::uint64_t{gridDim.x} * blockDim.x;
    for (std::uint64_t i = std::uint64_t{blockIdx.x} * blockDim.x + threadIdx.x; i < a.tile_count; i += stride)
        if (!index_tile(a, i, rec.file_bytes)) ++failures;
    if (failures != 0u) atomicAdd(&a.record->tile_failures, failures);
}

} // namespace

int tsk_index_header(const tsk_index_args *a, void *stream) {
    hipLaunchKernelGGL(index_header_kernel, dim3(1), dim3(1), 0, static_cast<hipStream_t>(stream), *a);
    return hipGetLastError();
}

int tsk_index_tiles(const tsk_index_args *a, void *stream) {
    const hipStream_t s = static_cast<hipStream_t>(stream);
    if (a->tile_count == 0) return hipSuccess;
    const std::uint64_t groups = (a->tile_count + 255u) / 256u;
    hipLaunchKernelGGL(index_tiles_kernel, dim3(static_cast<std::uint32_t>(groups < 65536u ? groups : 65536u)),
                       dim3(256), 0, s, *a);
    return hipGetLastError();
}
