// Tile decoder for gfx950 (CDNA4, wave64): kernel templates.
//
// Behaviour follows the reference decoder (reference src/cuda/decoder.cu:176-305):
// every tile intersecting the region is reconstructed from column 0, only
// region pixels are written, and the trusted path validates the device-resident
// index and row descriptors first.  The execution scheme is this library's:
//
//  * A workgroup decodes one tile at a time and loops over tasks, with a static
//    stride or by drawing chunks of tasks from a device work counter (the next
//    chunk is requested while the current one is decoded), so one launch drains
//    a whole mixed batch.
//  * The payload is staged into LDS with 16-byte loads from the 16-byte block
//    that contains its first word (such a block never crosses a page).  With
//    depth 2 the next tile's payload is loaded into registers while the current
//    tile decodes, and committed to LDS after it.
//  * A row of tile width W is split into segments of S = min(64, W / G) lanes;
//    each lane decodes G consecutive columns sequentially in registers (one
//    funnel shift per residual), the S lane totals are scanned with DPP
//    row_shr 1/2/4/8 and row_bcast 15/31, and rows wider than 64 G columns
//    carry the running value across passes.
//  * uint16 reconstruction is a 32-bit add of inverse-zigzag terms truncated to
//    16 bits at the store (exact modulo 2^16); float32 is a running XOR of bit
//    words.  Groups fully inside the region are written with 8- and 16-byte
//    stores at any word boundary; 16-bit outputs at odd halfword boundaries and
//    groups on the region's edge store sample by sample.
#pragma once

#include <hip/hip_runtime.h>

#include <tilestream_kernels.h>

#include <cstdint>

namespace tsk {

// Largest payload in words: a tile's payload is at most its raw size (the
// parser and the device-side index checks reject longer ones), 4096 words for
// a 4096-sample float32 tile and 2048 for a uint16 one.  With up to 3 words of
// 16-byte misalignment ahead of it, at most 1025 (float32) or 513 (uint16)
// 16-byte blocks are staged.  The slack covers the farthest read past a row
// end: a lane of the last segment reads up to (G S - 1) b / 32 + 1 <= 256
// words beyond the row's first packed word.  Kernels of the uint16 kinds use
// the smaller stage, so twice as many of their workgroups fit in a CU's LDS.
constexpr std::uint32_t stage_blocks_for(std::uint32_t payload_words) { return (payload_words + 3 + 3) / 4; }
constexpr std::uint32_t kPayloadWords = 4096;
template <int KIND> constexpr std::uint32_t stage_blocks() {
    return stage_blocks_for(KIND == TSK_KIND_U16 || KIND == TSK_KIND_U16_CAL ? kPayloadWords / 2 : kPayloadWords);
}
template <int KIND> constexpr std::uint32_t stage_words() { return stage_blocks<KIND>() * 4 + 280; }

// The descriptor tables never change while a decode runs.  Reading them through
// the constant address space lets the compiler use scalar loads (the scalar
// cache) instead of vector loads whose waits would also wait for the payload
// prefetch in flight.
// Global-memory pointers.  Pointers read from descriptor tables are generic to
// the compiler, and generic (FLAT) accesses also count against the LDS wait
// counter, so every LDS wait would drain the payload prefetch and the output
// stores.  Casting them to the global address space keeps the two apart.
template <typename T> using gptr = T __attribute__((address_space(1))) *;

// Native vector types: unlike HIP's uint4/float4 classes they load and store
// through address-space-qualified pointers.
typedef std::uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef std::uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
// Multi-word global accesses need only word alignment on CDNA: these types let
// output runs at any 4-byte boundary use 8- and 16-byte stores.
typedef std::uint32_t u32x4w __attribute__((ext_vector_type(4), aligned(4)));
typedef std::uint32_t u32x2w __attribute__((ext_vector_type(2), aligned(4)));

// LDS words addressed by a computed byte address.
using lds_word = const __attribute__((address_space(3))) std::uint32_t;

__device__ inline std::uint32_t lds_address(const std::uint32_t *p) {
    return static_cast<std::uint32_t>(reinterpret_cast<std::uintptr_t>((lds_word *)p));
}

__device__ inline lds_word *lds_at(std::uint32_t byte_address) {
    return reinterpret_cast<lds_word *>(static_cast<std::uintptr_t>(byte_address));
}

template <typename T> __device__ inline gptr<T> global(T *p) {
    return reinterpret_cast<gptr<T>>(reinterpret_cast<std::uintptr_t>(p));
}

template <typename T> __device__ inline T load_constant(const T *table, std::uint64_t index) {
    static_assert(sizeof(T) % 4 == 0, "descriptor records are whole words");
    using Word = const __attribute__((address_space(4))) std::uint32_t;
    Word *src = reinterpret_cast<Word *>(reinterpret_cast<std::uintptr_t>(table + index));
    std::uint32_t words[sizeof(T) / 4];
#pragma unroll
    for (unsigned k = 0; k < sizeof(T) / 4; ++k) words[k] = src[k];
    T out;
    __builtin_memcpy(&out, words, sizeof(T));
    return out;
}

struct ResultBits {
    std::uint32_t status;
    std::uint32_t reserved;
    std::uint64_t bytes;
};

constexpr std::uint32_t kMalformed = 5;

__device__ inline void fail(void *results, std::uint32_t image, std::uint32_t status) {
    atomicCAS(&(static_cast<ResultBits *>(results) + image)->status, 0u, status);
}

template <bool U16> struct Ring {
    // uint16: integers modulo 2^32 (truncated to 16 bits at the end); float32:
    // the XOR group of bit words.
    static __device__ inline std::uint32_t add(std::uint32_t a, std::uint32_t b) { return U16 ? a + b : a ^ b; }
    static __device__ inline std::uint32_t sub(std::uint32_t a, std::uint32_t b) { return U16 ? a - b : a ^ b; }
    // Inverse zigzag as an unsigned delta: even z -> z / 2, odd z -> ~(z / 2),
    // which is 65535 - z / 2 modulo 2^16.
    static __device__ inline std::uint32_t term(std::uint32_t z) { return U16 ? (z >> 1) ^ (0u - (z & 1u)) : z; }
};

template <bool U16, std::uint32_t Ctrl, std::uint32_t RowMask>
__device__ inline std::uint32_t dpp_combine(std::uint32_t v) {
    // Lanes whose source is outside the DPP row or masked out read 0, the identity.
    const std::uint32_t t = static_cast<std::uint32_t>(
        __builtin_amdgcn_update_dpp(0, static_cast<int>(v), Ctrl, RowMask, 0xf, true));
    return Ring<U16>::add(v, t);
}

template <bool U16, std::uint32_t Ctrl>
__device__ inline std::uint32_t dpp_combine_if(std::uint32_t v, bool take) {
    const std::uint32_t t = static_cast<std::uint32_t>(
        __builtin_amdgcn_update_dpp(0, static_cast<int>(v), Ctrl, 0xf, 0xf, true));
    return take ? Ring<U16>::add(v, t) : v;
}

// Inclusive scan of v over segments of S consecutive lanes (S a power of two,
// 4 <= S <= 64, uniform).  p is the lane's position in its segment.
template <bool U16> __device__ inline std::uint32_t segment_scan(std::uint32_t v, std::uint32_t p, std::uint32_t S) {
    if (S >= 16) {
        v = dpp_combine<U16, 0x111, 0xf>(v); // row_shr:1
        v = dpp_combine<U16, 0x112, 0xf>(v); // row_shr:2
        v = dpp_combine<U16, 0x114, 0xf>(v); // row_shr:4
        v = dpp_combine<U16, 0x118, 0xf>(v); // row_shr:8
        if (S >= 32) v = dpp_combine<U16, 0x142, 0xa>(v); // row_bcast:15 into rows 1 and 3
        if (S == 64) v = dpp_combine<U16, 0x143, 0xc>(v); // row_bcast:31 into rows 2 and 3
    } else {
        // Segments narrower than a DPP row: keep lanes from pulling across.
        v = dpp_combine_if<U16, 0x111>(v, p >= 1);
        v = dpp_combine_if<U16, 0x112>(v, p >= 2);
        if (S == 8) v = dpp_combine_if<U16, 0x114>(v, p >= 4);
    }
    return v;
}

template <int G> struct Group {
    std::uint32_t v[G];
};

struct Region {
    gptr<std::uint8_t> line; // output row of this tile row
    std::int64_t ex_tile; // output column of tile column 0 (may be negative)
    std::uint64_t ey;     // output row index
    std::uint32_t left, right;
};

__device__ inline float sample_float(std::uint32_t bits, bool u16) {
    return u16 ? static_cast<float>(bits & 0xffffu) : __uint_as_float(bits);
}

// Stores G consecutive 32-bit words at a 4-byte aligned address.  Outputs are
// written once and never read back by the decoder.  When the tile's output
// rows cover whole 128-byte lines, no other tile writes those lines, and a
// group of at most 16 bytes is stored non-temporally (consecutive lanes then
// cover whole lines in one instruction).  Otherwise lines at the tile's edges
// are shared with a neighbouring tile: ordinary stores let the L2 merge both
// halves into one write-back instead of two partial ones.
template <int G, bool NT> __device__ inline void store_words(gptr<std::uint32_t> dst, const std::uint32_t (&w)[G]) {
    if constexpr (G > 4 || !NT) {
        if constexpr (G >= 4) {
#pragma unroll
            for (int k = 0; k + 4 <= G; k += 4)
                *reinterpret_cast<gptr<u32x4w>>(dst + k) = u32x4w{w[k], w[k + 1], w[k + 2], w[k + 3]};
        } else if constexpr (G == 2) {
            *reinterpret_cast<gptr<u32x2w>>(dst) = u32x2w{w[0], w[1]};
        } else {
            dst[0] = w[0];
        }
    } else if constexpr (G == 4) {
        __builtin_nontemporal_store(u32x4w{w[0], w[1], w[2], w[3]}, reinterpret_cast<gptr<u32x4w>>(dst));
    } else if constexpr (G == 2) {
        __builtin_nontemporal_store(u32x2w{w[0], w[1]}, reinterpret_cast<gptr<u32x2w>>(dst));
    } else {
        __builtin_nontemporal_store(w[0], dst);
    }
}

// Loads G consecutive FP32 map values at a 4-byte aligned address.
template <int G> __device__ inline void load_floats(gptr<const float> src, float (&v)[G]) {
    if constexpr (G >= 4) {
#pragma unroll
        for (int k = 0; k + 4 <= G; k += 4) {
            const u32x4w q = *reinterpret_cast<gptr<const u32x4w>>(src + k);
            v[k] = __uint_as_float(q.x);
            v[k + 1] = __uint_as_float(q.y);
            v[k + 2] = __uint_as_float(q.z);
            v[k + 3] = __uint_as_float(q.w);
        }
    } else {
#pragma unroll
        for (int k = 0; k < G; ++k) v[k] = src[k];
    }
}

// Whether an image's calibration reads per-pixel maps (else both operands
// are scalars).
__device__ inline bool has_maps(const tsk_decode_image &img) {
    return img.offset_map != nullptr || img.gain_map != nullptr;
}

// Stores G samples (or their calibrated values) for tile columns col0..col0+G-1.
// Groups wholly inside the region store whole runs; groups on its left or
// right edge store only the region's columns.  MAPS: the image may have
// per-pixel calibration maps (tested per group); without, the operands are
// the scalars.  Kernels choose MAPS per image outside their store loops, so
// that the stores of an image without maps issue no loads: a wait for a load
// also waits for every store issued before it.
template <int G, bool U16, bool CAL, bool NT, bool MAPS>
__device__ inline void emit(const tsk_decode_image &img, const Region &r, std::uint32_t col0, const Group<G> &g) {
    const bool inside = col0 >= r.left && col0 + G <= r.right;
    const std::int64_t ex0 = r.ex_tile + col0;
    if constexpr (CAL && !MAPS) {
        gptr<float> dst = reinterpret_cast<gptr<float>>(r.line) + ex0;
        if (inside) {
            std::uint32_t out[G];
#pragma unroll
            for (int k = 0; k < G; ++k)
                // Two roundings to nearest even, never contracted into an FMA.
                out[k] = __float_as_uint(__fmul_rn(__fsub_rn(sample_float(g.v[k], U16), img.offset), img.gain));
            store_words<G, NT>(reinterpret_cast<gptr<std::uint32_t>>(dst), out);
        } else {
#pragma unroll
            for (int k = 0; k < G; ++k) {
                const std::uint32_t c = col0 + k;
                if (c >= r.left && c < r.right) dst[k] = __fmul_rn(__fsub_rn(sample_float(g.v[k], U16), img.offset), img.gain);
            }
        }
    } else if constexpr (CAL) {
        gptr<float> dst = reinterpret_cast<gptr<float>>(r.line) + ex0;
        const bool omap = img.offset_map != nullptr, gmap = img.gain_map != nullptr;
        const gptr<const float> orow =
            omap ? reinterpret_cast<gptr<const float>>(global(img.offset_map) + r.ey * img.offset_pitch) + ex0 : nullptr;
        const gptr<const float> grow =
            gmap ? reinterpret_cast<gptr<const float>>(global(img.gain_map) + r.ey * img.gain_pitch) + ex0 : nullptr;
        if (inside) {
            float o[G], gn[G];
            if (omap) {
                load_floats<G>(orow, o);
            } else {
#pragma unroll
                for (int k = 0; k < G; ++k) o[k] = img.offset;
            }
            if (gmap) {
                load_floats<G>(grow, gn);
            } else {
#pragma unroll
                for (int k = 0; k < G; ++k) gn[k] = img.gain;
            }
            std::uint32_t out[G];
#pragma unroll
            for (int k = 0; k < G; ++k)
                out[k] = __float_as_uint(__fmul_rn(__fsub_rn(sample_float(g.v[k], U16), o[k]), gn[k]));
            store_words<G, NT>(reinterpret_cast<gptr<std::uint32_t>>(dst), out);
        } else {
#pragma unroll
            for (int k = 0; k < G; ++k) {
                const std::uint32_t c = col0 + k;
                if (c >= r.left && c < r.right) {
                    const float o = omap ? orow[k] : img.offset;
                    const float gn = gmap ? grow[k] : img.gain;
                    dst[k] = __fmul_rn(__fsub_rn(sample_float(g.v[k], U16), o), gn);
                }
            }
        }
    } else if constexpr (U16) {
        gptr<std::uint16_t> dst = reinterpret_cast<gptr<std::uint16_t>>(r.line) + ex0;
        if (inside) {
            if (G >= 2 && (reinterpret_cast<std::uintptr_t>(dst) & 3u) == 0) {
                std::uint32_t w[G / 2 > 0 ? G / 2 : 1];
#pragma unroll
                for (int i = 0; i < G / 2; ++i) w[i] = __builtin_amdgcn_perm(g.v[2 * i + 1], g.v[2 * i], 0x05040100u);
                if constexpr (G >= 2) store_words<G / 2, NT>(reinterpret_cast<gptr<std::uint32_t>>(dst), w);
            } else {
#pragma unroll
                for (int k = 0; k < G; ++k) dst[k] = static_cast<std::uint16_t>(g.v[k]);
            }
        } else {
#pragma unroll
            for (int k = 0; k < G; ++k)
                if (col0 + k >= r.left && col0 + k < r.right) dst[k] = static_cast<std::uint16_t>(g.v[k]);
        }
    } else {
        gptr<std::uint32_t> dst = reinterpret_cast<gptr<std::uint32_t>>(r.line) + ex0;
        if (inside) {
            store_words<G, NT>(dst, g.v);
        } else {
#pragma unroll
            for (int k = 0; k < G; ++k)
                if (col0 + k >= r.left && col0 + k < r.right) dst[k] = g.v[k];
        }
    }
}

struct TileView {
    std::uint64_t tx, ty;
    std::uint32_t w, h;
    std::uint32_t top, bottom, left, right;
    std::uint32_t mode;
    std::uint32_t base; // LDS word of payload word 0
};

// Decodes the staged tile and writes its region pixels.
// Per-row constants of a predictive row.
struct RowBits {
    std::uint32_t seed;
    std::uint32_t b;
    std::uint32_t mask;      // low b bits
    std::uint32_t high, low; // uint16 extract widths: b - 1 and 1 (both 0 for b = 0)
    std::uint32_t words_bit; // LDS bit address of the row's first packed word
};

__device__ inline RowBits row_bits(const std::uint32_t *payload, std::uint32_t start) {
    RowBits r;
    r.seed = payload[start];
    r.b = payload[start + 1];
    r.mask = r.b >= 32 ? 0xffffffffu : ((1u << r.b) - 1u);
    // uint16 residuals are split by two bit-field extracts: bits 1..b-1
    // (z >> 1) and bit 0 sign-extended (-(z & 1)); both widths are 0 for
    // b = 0, which extracts 0.
    r.high = r.b > 0 ? r.b - 1 : 0u;
    r.low = r.b > 0 ? 1u : 0u;
    // A multiple of 32, so the low 5 bits of a bit address are the in-word offset.
    r.words_bit = lds_address(payload + start + 2) << 3;
    return r;
}

// Reconstructs columns col0..col0+G-1 of a predictive row: the lane's G
// residuals are summed sequentially, the lane totals are scanned over the
// segment of S lanes, and `carry` is the value before column 0 of this pass.
// Returns the segment's inclusive total in `incl`.
template <int G, bool U16>
__device__ inline void predict_group(const RowBits &rb, std::uint32_t col0, std::uint32_t p, std::uint32_t S,
                                     std::uint32_t carry, Group<G> &v, std::uint32_t &incl) {
    // Residual j of this lane is residual col0 + j of the row: it moves from
    // column col0 + j to col0 + j + 1.  Bit addresses are below 2^21, so 24-bit
    // multiplies (full rate) suffice.
    const std::uint32_t bit0 = __umul24(col0, rb.b) + rb.words_bit;
    std::uint32_t acc = 0;
    Group<G> pre;
    if constexpr (U16 && G % 2 == 0) {
        // Two uint16 residuals take 2b <= 32 bits, which always lie in the
        // 64-bit window of the word holding the first one and the next: one
        // LDS read and one funnel shift serve the pair, the second residual
        // starting at bit b of it.
        std::uint32_t bit = bit0;
        const std::uint32_t step = rb.b * 2u;
#pragma unroll
        for (int j = 0; j < G; j += 2) {
            lds_word *w = lds_at((bit >> 3) & ~3u);
            const std::uint32_t x = __builtin_amdgcn_alignbit(w[1], w[0], bit);
            const std::uint32_t half0 = __builtin_amdgcn_ubfe(x, 1u, rb.high);
            const std::uint32_t sign0 = static_cast<std::uint32_t>(__builtin_amdgcn_sbfe(static_cast<int>(x), 0u, rb.low));
            acc += half0 ^ sign0;
            pre.v[j] = acc;
            const std::uint32_t half1 = __builtin_amdgcn_ubfe(x, rb.b + 1u, rb.high);
            const std::uint32_t sign1 =
                static_cast<std::uint32_t>(__builtin_amdgcn_sbfe(static_cast<int>(x), rb.b, rb.low));
            acc += half1 ^ sign1;
            pre.v[j + 1] = acc;
            bit += step;
        }
    } else {
#pragma unroll
        for (int j = 0; j < G; ++j) {
            const std::uint32_t bit = __umul24(static_cast<std::uint32_t>(j), rb.b) + bit0;
            lds_word *w = lds_at((bit >> 3) & ~3u);
            const std::uint32_t x = __builtin_amdgcn_alignbit(w[1], w[0], bit);
            if constexpr (U16) {
                const std::uint32_t half = __builtin_amdgcn_ubfe(x, 1u, rb.high);
                const std::uint32_t sign =
                    static_cast<std::uint32_t>(__builtin_amdgcn_sbfe(static_cast<int>(x), 0u, rb.low));
                acc += half ^ sign;
            } else {
                acc ^= x & rb.mask;
            }
            pre.v[j] = acc;
        }
    }
    incl = segment_scan<U16>(acc, p, S);
    const std::uint32_t first = Ring<U16>::add(carry, Ring<U16>::sub(incl, acc));
    v.v[0] = first;
#pragma unroll
    for (int k = 1; k < G; ++k) v.v[k] = Ring<U16>::add(first, pre.v[k - 1]);
}

// Columns col0..col0+G-1 of a RAW (mode 0) or CONSTANT (mode 1) tile row.
template <int G, bool U16>
__device__ inline void stored_group(const TileView &t, const std::uint32_t *payload, std::uint32_t row,
                                    std::uint32_t col0, Group<G> &v) {
    if (t.mode == 0) {
        const std::uint32_t idx0 = __umul24(row, t.w) + col0;
#pragma unroll
        for (int k = 0; k < G; ++k) {
            const std::uint32_t idx = idx0 + k;
            v.v[k] = U16 ? (payload[idx >> 1] >> ((idx & 1u) * 16)) : payload[idx];
        }
    } else {
#pragma unroll
        for (int k = 0; k < G; ++k) v.v[k] = payload[0];
    }
}

__device__ inline Region region_of(const tsk_decode_image &img, const TileView &t, std::uint32_t row) {
    Region r;
    r.ey = t.ty + row - img.roi_y;
    r.line = global(img.out) + r.ey * img.out_pitch;
    r.ex_tile = static_cast<std::int64_t>(t.tx) - static_cast<std::int64_t>(img.roi_x);
    r.left = t.left;
    r.right = t.right;
    return r;
}

// Whether every output row of the tile starts on a 128-byte line and spans
// whole lines, so that no other tile writes into them.
__device__ inline bool whole_lines(const tsk_decode_image &img, const TileView &t) {
    const std::uint32_t s = img.calibrated || img.type != TSK_UINT16 ? 4u : 2u;
    const std::int64_t ex = static_cast<std::int64_t>(t.tx) - static_cast<std::int64_t>(img.roi_x);
    const std::uintptr_t first = reinterpret_cast<std::uintptr_t>(img.out) + static_cast<std::uintptr_t>(ex) * s;
    return t.left == 0 && t.right == img.tile_w && img.out_pitch % 128 == 0 && first % 128 == 0 &&
           (img.tile_w * s) % 128 == 0;
}

// Decodes the staged tile and writes its region pixels as it goes.
template <int G, bool U16, bool CAL, bool MAPS, int BLOCK>
__device__ void decode_tile(const tsk_decode_image &img, const TileView &t, const std::uint32_t *stage,
                            const std::uint16_t *row_start) {
    const std::uint32_t S = (img.tile_w / G) < 64u ? img.tile_w / G : 64u;
    const std::uint32_t passes = img.tile_w / (S * G);
    const std::uint32_t p = threadIdx.x % S;
    const std::uint32_t rows_per_step = BLOCK / S;
    const std::uint32_t *payload = stage + t.base;

    for (std::uint32_t row = t.top + threadIdx.x / S; row < t.bottom; row += rows_per_step) {
        const Region r = region_of(img, t, row);
        if (t.mode == 2) {
            const RowBits rb = row_bits(payload, row_start[row]);
            std::uint32_t carry = rb.seed;
            for (std::uint32_t pass = 0; pass < passes; ++pass) {
                const std::uint32_t col0 = (pass * S + p) * G;
                Group<G> v;
                std::uint32_t incl;
                predict_group<G, U16>(rb, col0, p, S, carry, v, incl);
                if (passes > 1) // S = 64: lane 63 holds the row total so far
                    carry = Ring<U16>::add(carry, static_cast<std::uint32_t>(
                                                      __builtin_amdgcn_readlane(static_cast<int>(incl), 63)));
                emit<G, U16, CAL, false, MAPS>(img, r, col0, v);
            }
        } else {
            for (std::uint32_t pass = 0; pass < passes; ++pass) {
                const std::uint32_t col0 = (pass * S + p) * G;
                Group<G> v;
                stored_group<G, U16>(t, payload, row, col0, v);
                emit<G, U16, CAL, false, MAPS>(img, r, col0, v);
            }
        }
    }
}

// Groups per thread of a 4096-pixel tile when G >= 4 (one pass per row); the
// narrower groups store as they decode and hold nothing.
template <int G, int BLOCK> constexpr int held_groups() {
    return G >= 4 ? (4096 + BLOCK * G - 1) / (BLOCK * G) : 1;
}

// A thread's decoded groups of one tile, kept in registers until they are
// stored (row >= the tile's bottom: no group).
template <int G, int K> struct Held {
    Group<G> v[K];
    std::uint32_t row[K];
};

// Decodes the staged tile into registers (G >= 4, so S <= 64 and one pass per
// row); held_store writes them afterwards.
template <int G, bool U16, int BLOCK, int K>
__device__ inline void decode_tile_held(const tsk_decode_image &img, const TileView &t, const std::uint32_t *stage,
                                        const std::uint16_t *row_start, Held<G, K> &h) {
    static_assert(G >= 4, "one pass per row");
    const std::uint32_t S = img.tile_w / G;
    const std::uint32_t p = threadIdx.x % S;
    const std::uint32_t rows_per_step = BLOCK / S;
    const std::uint32_t col0 = p * G;
    const std::uint32_t *payload = stage + t.base;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const std::uint32_t row = t.top + threadIdx.x / S + k * rows_per_step;
        h.row[k] = row;
        if (row < t.bottom) {
            if (t.mode == 2) {
                const RowBits rb = row_bits(payload, row_start[row]);
                std::uint32_t incl;
                predict_group<G, U16>(rb, col0, p, S, rb.seed, h.v[k], incl);
            } else {
                stored_group<G, U16>(t, payload, row, col0, h.v[k]);
            }
        }
    }
}

template <int G, bool U16, bool CAL, int BLOCK, int K, bool NT, bool MAPS>
__device__ inline void held_store_as(const tsk_decode_image &img, const TileView &t, const Held<G, K> &h) {
    const std::uint32_t col0 = (threadIdx.x % (img.tile_w / G)) * G;
#pragma unroll
    for (int k = 0; k < K; ++k)
        if (h.row[k] < t.bottom) emit<G, U16, CAL, NT, MAPS>(img, region_of(img, t, h.row[k]), col0, h.v[k]);
}

// Tiles whose output rows span whole lines store them non-temporally (where a
// group is one store instruction).  SPLIT: calibrated images with maps and
// without take separate paths, so that the stores of an image without maps
// issue no loads (a wait for a load also waits for every store issued before
// it).  The mixed kind keeps one path, which tests for maps per group: a
// second copy of its calibrated paths costs its plain paths registers.
template <int G, bool U16, bool CAL, int BLOCK, int K, bool SPLIT>
__device__ inline void held_store(const tsk_decode_image &img, const TileView &t, const Held<G, K> &h) {
    const bool nt = whole_lines(img, t);
    if constexpr (CAL && SPLIT) {
        if (has_maps(img)) {
            if (nt) held_store_as<G, U16, CAL, BLOCK, K, true, true>(img, t, h);
            else held_store_as<G, U16, CAL, BLOCK, K, false, true>(img, t, h);
            return;
        }
        if (nt) held_store_as<G, U16, CAL, BLOCK, K, true, false>(img, t, h);
        else held_store_as<G, U16, CAL, BLOCK, K, false, false>(img, t, h);
    } else {
        if (nt) held_store_as<G, U16, CAL, BLOCK, K, true, CAL>(img, t, h);
        else held_store_as<G, U16, CAL, BLOCK, K, false, CAL>(img, t, h);
    }
}

// A tile whose payload is being fetched.
struct Pending {
    std::uint32_t image;
    std::uint32_t column, row;
    std::uint32_t rows;
    std::uint32_t mode, h, w;
    std::uint32_t words;
    std::uint32_t state; // 0 decode, 1 producer failed (skip), 2 malformed
    std::uint32_t mis;   // words ahead of the payload in its first 16-byte block
    std::uint32_t nblocks;
    gptr<const u32x4> blocks;
};

// Request mode: the request holding task i (one table load) and the tile of
// its region.
__device__ inline tsk_decode_task request_task(const tsk_decode_args &a, std::uint64_t i, std::uint32_t &request) {
    const std::uint32_t r = load_constant(a.request_index, i);
    const tsk_request q = load_constant(a.requests, r);
    const std::uint32_t local = static_cast<std::uint32_t>(i - q.first_task);
    const std::uint32_t dr = local / q.cols;
    const std::uint32_t col = q.col0 + (local - dr * q.cols);
    const std::uint32_t row = q.row0 + dr;
    tsk_decode_task t = load_constant(a.tasks, std::uint64_t{q.tile_base} + std::uint64_t{row} * q.tile_columns + col);
    t.column = col;
    t.row = row;
    request = r;
    if (a.staged != nullptr) {
        // A tile of an archive in host memory, copied into device staging.
        const std::uint32_t block = load_constant(a.staged, i);
        if (block != TSK_NOT_STAGED)
            t.payload = a.staging + std::uint64_t{block} * 16 + (reinterpret_cast<std::uintptr_t>(t.payload) & 15u);
    }
    return t;
}

// Reads a task descriptor and, on the trusted path, the producer's result and
// the device-resident index entry (reference decoder.cu:191-220).
__device__ inline Pending describe_task(const tsk_decode_args &a, const tsk_decode_task &task, std::uint32_t image) {
    Pending p;
    p.image = image;
    p.column = task.column;
    p.row = task.row;
    p.rows = task.rows;
    p.mode = task.shape & 0xffu;
    p.h = (task.shape >> 8) & 0xffu;
    p.w = task.shape >> 16;
    p.words = task.words;
    p.state = 0;
    const std::uint8_t *addr = task.payload;
    if (addr == nullptr) {
        const tsk_decode_image img = load_constant(a.images, task.image);
        const gptr<const ResultBits> prod = global(static_cast<const ResultBits *>(img.producer));
        if (prod->status != 0) {
            p.state = 1;
        } else {
            const std::uint64_t index = std::uint64_t{task.row} * img.tile_columns + task.column;
            const gptr<const std::uint32_t> s32 =
                global(reinterpret_cast<const std::uint32_t *>(img.src + 64 + index * 32));
            const std::uint32_t mode = s32[0] & 0xffu;
            const std::uint32_t dims = s32[1];
            const std::uint64_t off = s32[2] | (std::uint64_t{s32[3]} << 32);
            const std::uint64_t len = s32[4] | (std::uint64_t{s32[5]} << 32);
            const std::uint64_t end = prod->bytes < img.src_size ? prod->bytes : img.src_size;
            const std::uint64_t raw4 = (std::uint64_t{p.w} * p.h * (img.type == TSK_UINT16 ? 2 : 4) + 3) / 4 * 4;
            if (mode > 2 || (dims & 0xffffu) != p.w || (dims >> 16) != p.h || off % 4 != 0 || len % 4 != 0 ||
                len == 0 || len > raw4 || off > end || len > end - off) {
                p.state = 2;
            } else {
                p.mode = mode;
                p.words = static_cast<std::uint32_t>(len / 4);
                addr = img.src + off;
            }
        }
    }
    if (p.state != 0) {
        p.words = 0;
        addr = reinterpret_cast<const std::uint8_t *>(a.tasks); // never dereferenced: no blocks
    }
    const std::uintptr_t u = reinterpret_cast<std::uintptr_t>(addr);
    p.mis = static_cast<std::uint32_t>((u >> 2) & 3u);
    p.blocks = reinterpret_cast<gptr<const u32x4>>(u - 4u * p.mis);
    p.nblocks = p.state == 0 ? (p.words + p.mis + 3) >> 2 : 0;
    return p;
}

__device__ inline Pending describe(const tsk_decode_args &a, std::uint64_t i) {
    if (a.request_count != 0) {
        std::uint32_t request = 0;
        const tsk_decode_task task = request_task(a, i, request);
        return describe_task(a, task, request);
    }
    const tsk_decode_task task = load_constant(a.tasks, i);
    return describe_task(a, task, task.image);
}

// describe() for a kernel of kind KIND: a payload longer than the kind's
// stage (which the checks above already exclude) is malformed, never staged.
template <int KIND> __device__ inline Pending describe_for(const tsk_decode_args &a, std::uint64_t i) {
    Pending p = describe(a, i);
    if (p.nblocks > stage_blocks<KIND>()) {
        p.state = 2;
        p.nblocks = 0;
    }
    return p;
}

// Work assignment.  Without a work counter: a static stride over the tasks.
// With one: the first seven eighths of the tasks by static stride (every
// workgroup busy on neighbouring tasks, so payload reads and output writes
// stay local, with no atomics), the rest in chunks drawn from the counter
// (the next chunk requested one chunk ahead) to even out the finish.
struct Cursor {
    std::uint64_t next, end;
    std::uint64_t ahead; // the chunk requested one chunk early (thread 0)
    bool dynamic;
    std::uint32_t left;  // tasks left in the current static run
};

__device__ inline std::uint32_t run_length(const tsk_decode_args &a) { return a.run > 1 ? a.run : 1u; }

__device__ inline std::uint64_t static_end(const tsk_decode_args &a) { return a.task_count - a.task_count / 8; }

// The stride position of this workgroup.  Workgroups are dispatched to the
// eight XCDs (each with its own L2) round robin, so workgroup b runs on XCD
// b % 8.  Giving each XCD a contiguous run of grid / 8 neighbouring tasks per
// stride step keeps neighbouring tiles, whose output rows share cache lines
// at tile edges, in one L2, which then writes those lines back whole.
__device__ inline std::uint64_t stride_slot() {
    const std::uint32_t b = blockIdx.x, g = gridDim.x;
    return g % 8 == 0 ? std::uint64_t{b % 8} * (g / 8) + b / 8 : b;
}

// The cursor of a workgroup: slot s takes tasks [s run, s run + run), then
// strides on by grid x run.
__device__ inline Cursor start_cursor(const tsk_decode_args &a) {
    return Cursor{stride_slot() * run_length(a), 0, 0, false, run_length(a)};
}

// Thread 0 requests the first dynamic chunk at kernel start.
__device__ inline void queue_start(const tsk_decode_args &a, Cursor &c) {
    if (threadIdx.x == 0) c.ahead = atomicAdd(reinterpret_cast<unsigned long long *>(a.queue), a.chunk);
}

template <int BLOCK>
__device__ inline std::uint64_t draw(const tsk_decode_args &a, Cursor &c, std::uint64_t *shared_next) {
    if (a.queue == nullptr || !c.dynamic) {
        if (a.queue == nullptr || c.next < static_end(a)) {
            const std::uint64_t i = c.next;
            if (--c.left == 0) {
                c.left = run_length(a);
                c.next += std::uint64_t{gridDim.x - 1} * c.left + 1;
            } else {
                ++c.next;
            }
            return i;
        }
        c.dynamic = true;
        c.next = c.end = 0;
    }
    if (c.next >= c.end) {
        // Publish the chunk requested earlier and request the one after it.
        __syncthreads();
        if (threadIdx.x == 0) {
            shared_next[0] = c.ahead;
            c.ahead = atomicAdd(reinterpret_cast<unsigned long long *>(a.queue), a.chunk);
        }
        __syncthreads();
        // The chunk start is the same for every lane: keep it (and every
        // descriptor read through it) in scalar registers.
        const std::uint64_t start = shared_next[0] + static_end(a);
        c.next = (static_cast<std::uint64_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(start >> 32))) << 32) |
                 static_cast<std::uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(start)));
        c.end = c.next + a.chunk;
    }
    return c.next++;
}

// Validates the row descriptors of a trusted (encoder-produced) predictive
// tile staged in LDS and fills row_start: a serial walk in thread 0
// (reference decoder.cu:86-110).  Returns whether the tile is well formed.
__device__ inline bool walk_rows(const Pending &c, bool u16, const std::uint32_t *stage, std::uint16_t *row_start,
                                 std::uint32_t *shared_flag) {
    if (threadIdx.x == 0) {
        const std::uint32_t widest = u16 ? 16u : 32u;
        std::uint32_t cur = 0;
        bool ok = true;
        for (std::uint32_t r = 0; r < c.h && ok; ++r) {
            if (c.words - cur < 2) {
                ok = false;
                break;
            }
            const std::uint32_t seed = stage[c.mis + cur];
            const std::uint32_t bits = stage[c.mis + cur + 1];
            const std::uint32_t packed = ((c.w - 1) * bits + 31) / 32;
            if (bits > widest || (u16 && seed > 65535u) || c.words - cur - 2 < packed) {
                ok = false;
                break;
            }
            row_start[r] = static_cast<std::uint16_t>(cur);
            cur += 2 + packed;
        }
        *shared_flag = ok && cur == c.words;
    }
    __syncthreads();
    return *shared_flag != 0;
}

__device__ inline TileView tile_view(const Pending &c, const tsk_decode_image &img) {
    TileView t;
    t.tx = std::uint64_t{c.column} * img.tile_w;
    t.ty = std::uint64_t{c.row} * img.tile_h;
    t.w = c.w;
    t.h = c.h;
    t.mode = c.mode;
    t.base = c.mis;
    const std::uint64_t rx = img.roi_x, ry = img.roi_y;
    t.top = ry > t.ty ? static_cast<std::uint32_t>(ry - t.ty) : 0u;
    t.bottom = ry + img.roi_h < t.ty + t.h ? static_cast<std::uint32_t>(ry + img.roi_h - t.ty) : t.h;
    t.left = rx > t.tx ? static_cast<std::uint32_t>(rx - t.tx) : 0u;
    t.right = rx + img.roi_w < t.tx + t.w ? static_cast<std::uint32_t>(rx + img.roi_w - t.tx) : t.w;
    return t;
}

// Makes the prefetched payload registers and the look-ahead queue draw
// available: the wait lands here, before this tile's output stores are
// issued, so it never waits for them (vector memory counters retire in order).
template <int KMAX>
__device__ inline void await_prefetch(const u32x4 (&reg)[KMAX], std::uint32_t row_reg, std::uint64_t ahead) {
#pragma unroll
    for (int k = 0; k < KMAX; ++k) asm volatile("" ::"v"(reg[k]) : "memory");
    asm volatile("" ::"v"(row_reg), "v"(static_cast<std::uint32_t>(ahead)), "v"(static_cast<std::uint32_t>(ahead >> 32))
                 : "memory");
}

// Loads a tile's staged blocks and row offsets into registers.  A slot is
// loaded when any lane of the workgroup needs it (a uniform branch); within
// it, lanes past the tile's last block load that block again, so no load is
// masked per lane and the register array stays in place.
template <int BLOCK, int KMAX>
__device__ inline void prefetch(const tsk_decode_args &a, const Pending &p, u32x4 (&reg)[KMAX], std::uint32_t &row_reg) {
    const gptr<const u32x4> base = p.blocks;
    const std::uint32_t last = p.nblocks != 0 ? p.nblocks - 1 : 0u;
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
        const std::uint32_t b = threadIdx.x + k * BLOCK;
        if (k * BLOCK < p.nblocks) reg[k] = __builtin_nontemporal_load(base + (b < last ? b : last));
    }
    // The row offsets load is unconditional too (from the task table when the
    // tile has none), so every prefetch issues the same number of loads.
    const bool rows = p.mode == 2 && p.rows != TSK_NO_ROWS && p.rows < a.row_count;
    const std::uint64_t r = rows ? p.rows + (threadIdx.x < p.h ? threadIdx.x : 0u) : 0u;
    const std::uint64_t r_clamped = r < a.row_count ? r : p.rows;
    const gptr<const std::uint16_t> rp =
        rows ? global(a.rows) + r_clamped
             : reinterpret_cast<gptr<const std::uint16_t>>(reinterpret_cast<std::uintptr_t>(a.tasks));
    const std::uint32_t v = rp[0];
    row_reg = rows ? v : 0u;
}

// Decodes the staged tile with the kernel's kind and stores it as it goes.
template <int G, int KIND, int BLOCK>
__device__ inline void decode_tile_as(const tsk_decode_image &img, const TileView &t, const std::uint32_t *stage,
                                      const std::uint16_t *row_start) {
    const bool maps = has_maps(img);
    if constexpr (KIND == TSK_KIND_U16) {
        decode_tile<G, true, false, false, BLOCK>(img, t, stage, row_start);
    } else if constexpr (KIND == TSK_KIND_F32) {
        decode_tile<G, false, false, false, BLOCK>(img, t, stage, row_start);
    } else if constexpr (KIND == TSK_KIND_U16_CAL) {
        if (maps) decode_tile<G, true, true, true, BLOCK>(img, t, stage, row_start);
        else decode_tile<G, true, true, false, BLOCK>(img, t, stage, row_start);
    } else if constexpr (KIND == TSK_KIND_F32_CAL) {
        if (maps) decode_tile<G, false, true, true, BLOCK>(img, t, stage, row_start);
        else decode_tile<G, false, true, false, BLOCK>(img, t, stage, row_start);
    } else if (img.type == TSK_UINT16) {
        // The mixed kind: one calibrated path that tests for maps per group.
        if (img.calibrated) decode_tile<G, true, true, true, BLOCK>(img, t, stage, row_start);
        else decode_tile<G, true, false, false, BLOCK>(img, t, stage, row_start);
    } else {
        if (img.calibrated) decode_tile<G, false, true, true, BLOCK>(img, t, stage, row_start);
        else decode_tile<G, false, false, false, BLOCK>(img, t, stage, row_start);
    }
}

// Depth 1: the payload of a tile is loaded after the previous one is decoded.
template <int G, int KIND, int BLOCK>
__device__ inline __attribute__((always_inline)) void decode_body_staged(const tsk_decode_args &a) {
    __shared__ __attribute__((aligned(16))) std::uint32_t stage[stage_words<KIND>()];
    __shared__ std::uint16_t row_start[64];
    __shared__ std::uint32_t shared_flag;
    __shared__ std::uint64_t shared_next[1];

    const std::uint32_t tid = threadIdx.x;
    Cursor cursor = start_cursor(a);
    if (a.queue != nullptr) queue_start(a, cursor);
    for (;;) {
        const std::uint64_t i = draw<BLOCK>(a, cursor, shared_next);
        if (i >= a.task_count) return;
        const Pending c = describe_for<KIND>(a, i);
        __syncthreads(); // the previous tile is done with the LDS
        for (std::uint32_t k = tid; k < c.nblocks; k += BLOCK) *reinterpret_cast<u32x4 *>(stage + 4 * k) = c.blocks[k];
        if (c.mode == 2 && c.rows != TSK_NO_ROWS && tid < c.h && c.rows + tid < a.row_count)
            row_start[tid] = (global(a.rows) + c.rows)[tid];
        __syncthreads();
        if (c.state == 2) {
            if (tid == 0) fail(a.results, c.image, kMalformed);
            continue;
        }
        if (c.state != 0) continue;
        // The descriptor is copied by value: output stores cannot alias it.
        const tsk_decode_image img = load_constant(a.images, c.image);
        const bool u16 = img.type == TSK_UINT16;
        if (c.mode == 2 && c.rows == TSK_NO_ROWS && !walk_rows(c, u16, stage, row_start, &shared_flag)) {
            if (tid == 0) fail(a.results, c.image, kMalformed);
            continue;
        }
        decode_tile_as<G, KIND, BLOCK>(img, tile_view(c, img), stage, row_start);
    }
}

// Decodes staged tile c into registers; false when nothing is to be stored.
template <int G, int KIND, int BLOCK, int K>
__device__ inline __attribute__((always_inline)) bool decode_phase(const tsk_decode_args &a, const Pending &c,
                                                                   const std::uint32_t *stage, std::uint16_t *row_start,
                                                                   std::uint32_t *shared_flag, tsk_decode_image &img,
                                                                   TileView &t, Held<G, K> &h) {
    if (c.state == 2) {
        if (threadIdx.x == 0) fail(a.results, c.image, kMalformed);
        return false;
    }
    if (c.state != 0) return false;
    // The descriptor is copied by value: output stores cannot alias it.
    img = load_constant(a.images, c.image);
    const bool u16 = img.type == TSK_UINT16;
    if (c.mode == 2 && c.rows == TSK_NO_ROWS && !walk_rows(c, u16, stage, row_start, shared_flag)) {
        if (threadIdx.x == 0) fail(a.results, c.image, kMalformed);
        return false;
    }
    t = tile_view(c, img);
    if constexpr (G < 4) {
        // A row may take several passes: decoded and stored as it goes.
        decode_tile_as<G, KIND, BLOCK>(img, t, stage, row_start);
        return false;
    } else if constexpr (KIND == TSK_KIND_U16 || KIND == TSK_KIND_U16_CAL) {
        decode_tile_held<G, true, BLOCK, K>(img, t, stage, row_start, h);
    } else if constexpr (KIND == TSK_KIND_F32 || KIND == TSK_KIND_F32_CAL) {
        decode_tile_held<G, false, BLOCK, K>(img, t, stage, row_start, h);
    } else if (u16) {
        decode_tile_held<G, true, BLOCK, K>(img, t, stage, row_start, h);
    } else {
        decode_tile_held<G, false, BLOCK, K>(img, t, stage, row_start, h);
    }
    return true;
}

template <int G, int KIND, int BLOCK, int K>
__device__ inline __attribute__((always_inline)) void store_phase(const tsk_decode_image &img, const TileView &t,
                                                                  const Held<G, K> &h) {
    if constexpr (KIND == TSK_KIND_U16) {
        held_store<G, true, false, BLOCK, K, true>(img, t, h);
    } else if constexpr (KIND == TSK_KIND_F32) {
        held_store<G, false, false, BLOCK, K, true>(img, t, h);
    } else if constexpr (KIND == TSK_KIND_U16_CAL) {
        held_store<G, true, true, BLOCK, K, true>(img, t, h);
    } else if constexpr (KIND == TSK_KIND_F32_CAL) {
        held_store<G, false, true, BLOCK, K, true>(img, t, h);
    } else if (img.type == TSK_UINT16) {
        if (img.calibrated) held_store<G, true, true, BLOCK, K, false>(img, t, h);
        else held_store<G, true, false, BLOCK, K, false>(img, t, h);
    } else {
        if (img.calibrated) held_store<G, false, true, BLOCK, K, false>(img, t, h);
        else held_store<G, false, false, BLOCK, K, false>(img, t, h);
    }
}

template <int BLOCK, int KMAX>
__device__ inline void commit(const tsk_decode_args &a, const Pending &q, const u32x4 (&reg)[KMAX], std::uint32_t row_reg,
                              std::uint32_t *stage, std::uint16_t *row_start) {
    const std::uint32_t tid = threadIdx.x;
#pragma unroll
    for (int k = 0; k < KMAX; ++k)
        if (tid + k * BLOCK < q.nblocks) *reinterpret_cast<u32x4 *>(stage + 4 * (tid + k * BLOCK)) = reg[k];
    if (q.mode == 2 && q.rows != TSK_NO_ROWS && tid < q.h && q.rows + tid < a.row_count)
        row_start[tid] = static_cast<std::uint16_t>(row_reg);
}

// Depth 2 and 3: two LDS stage buffers and DEPTH - 1 register sets.  Tile n
// is decoded (into registers) from one buffer while the payloads of the next
// DEPTH - 1 tiles are in flight into registers.  Then the oldest set is
// awaited (the wait covers only loads issued before it and the previous
// tile's stores), committed to the other buffer, tile n's groups are stored,
// the set is reloaded with the tile DEPTH - 1 ahead, and one barrier ends the
// round.
template <int G, int KIND, int BLOCK, int DEPTH>
struct Pipeline {
    static constexpr int KMAX = static_cast<int>((stage_blocks<KIND>() + BLOCK - 1) / BLOCK);
    static constexpr int K = held_groups<G, BLOCK>();
    static constexpr int SETS = DEPTH - 1;

    const tsk_decode_args &a;
    std::uint32_t (*stage)[stage_words<KIND>()];
    std::uint16_t (*row_start)[64];
    std::uint32_t *shared_flag;
    std::uint64_t *shared_next;
    Cursor cursor;
    u32x4 reg[SETS][KMAX];
    std::uint32_t row_reg[SETS];
    Pending c;         // staged in stage[buf]
    Pending p[SETS];   // p[0]: next tile, in flight in set `next`; p[1]: the one after
    bool more[SETS];
    std::uint32_t buf;

    __device__ inline __attribute__((always_inline)) void load(int set, std::uint32_t slot) {
        const std::uint64_t n = draw<BLOCK>(a, cursor, shared_next);
        more[slot] = n < a.task_count;
        if (more[slot]) p[slot] = describe_for<KIND>(a, n);
        else p[slot].nblocks = 0, p[slot].mode = 0;
        prefetch<BLOCK>(a, p[slot], reg[set], row_reg[set]);
    }

    // One round with the next tile's payload in register set S.
    template <int S> __device__ inline __attribute__((always_inline)) bool round() {
        tsk_decode_image img{};
        TileView t{};
        Held<G, K> h;
        const bool store = decode_phase<G, KIND, BLOCK, K>(a, c, stage[buf], row_start[buf], shared_flag, img, t, h);
        const bool had_next = more[0];
        if (had_next) {
            await_prefetch(reg[S], row_reg[S], cursor.ahead);
            commit<BLOCK>(a, p[0], reg[S], row_reg[S], stage[buf ^ 1u], row_start[buf ^ 1u]);
        }
        if constexpr (G >= 4) {
            if (store) store_phase<G, KIND, BLOCK, K>(img, t, h);
        }
        if (!had_next) return false;
        c = p[0];
        if constexpr (SETS == 2) {
            p[0] = p[1];
            more[0] = more[1];
            load(S, 1);
        } else {
            load(S, 0);
        }
        __syncthreads(); // the committed tile is visible; the decoded buffer is free
        buf ^= 1u;
        return true;
    }

    __device__ inline __attribute__((always_inline)) void run() {
        if (a.queue != nullptr) queue_start(a, cursor);
        const std::uint64_t i = draw<BLOCK>(a, cursor, shared_next);
        if (i >= a.task_count) return;
        c = describe_for<KIND>(a, i);
        prefetch<BLOCK>(a, c, reg[0], row_reg[0]);
        await_prefetch(reg[0], row_reg[0], cursor.ahead);
        commit<BLOCK>(a, c, reg[0], row_reg[0], stage[0], row_start[0]);
        buf = 0;
        if constexpr (SETS == 2) {
            load(1, 0);
            load(0, 1);
            __syncthreads();
            for (;;) {
                if (!round<1>()) break;
                if (!round<0>()) break;
            }
        } else {
            load(0, 0);
            __syncthreads();
            while (round<0>()) {
            }
        }
    }
};

template <int G, int KIND, int BLOCK, int DEPTH>
__device__ inline __attribute__((always_inline)) void decode_body_prefetch(const tsk_decode_args &a) {
    __shared__ __attribute__((aligned(16))) std::uint32_t stage[2][stage_words<KIND>()];
    __shared__ std::uint16_t row_start[2][64];
    __shared__ std::uint32_t shared_flag;
    __shared__ std::uint64_t shared_next[1];
    Pipeline<G, KIND, BLOCK, DEPTH> pipe{a, stage, row_start, &shared_flag, shared_next, start_cursor(a)};
    pipe.run();
}

template <int G, int KIND, int BLOCK, int DEPTH>
__device__ inline __attribute__((always_inline)) void decode_body(const tsk_decode_args &a) {
    if constexpr (DEPTH > 1) decode_body_prefetch<G, KIND, BLOCK, DEPTH>(a);
    else decode_body_staged<G, KIND, BLOCK>(a);
}

template <int G, int KIND, int BLOCK, int DEPTH>
__global__ void __launch_bounds__(BLOCK) decode_kernel(tsk_decode_args a) {
    decode_body<G, KIND, BLOCK, DEPTH>(a);
}

using DecodeFn = void (*)(tsk_decode_args);

// Every (group, block, depth) combination is compiled: groups 1, 2, 4, 8, 16
// lanes per row, 64 to 512 threads, depth 1 to 3 (each register set of depth
// 2 and 3 holds ceil(1025 / block) 16-byte blocks per thread).
constexpr bool compiled(std::uint32_t group, std::uint32_t block, std::uint32_t depth) {
    const bool g_ok = group == 1 || group == 2 || group == 4 || group == 8 || group == 16;
    const bool b_ok = block == 64 || block == 128 || block == 256 || block == 512;
    return g_ok && b_ok && depth >= 1 && depth <= 3;
}

template <int KIND, int G, int BLOCK> DecodeFn pick_depth(std::uint32_t depth) {
    switch (depth) {
    case 1: return decode_kernel<G, KIND, BLOCK, 1>;
    case 2: return decode_kernel<G, KIND, BLOCK, 2>;
    case 3: return decode_kernel<G, KIND, BLOCK, 3>;
    default: return nullptr;
    }
}

template <int KIND, int G> DecodeFn pick_block(std::uint32_t block, std::uint32_t depth) {
    switch (block) {
    case 64: return pick_depth<KIND, G, 64>(depth);
    case 128: return pick_depth<KIND, G, 128>(depth);
    case 256: return pick_depth<KIND, G, 256>(depth);
    case 512: return pick_depth<KIND, G, 512>(depth);
    default: return nullptr;
    }
}

} // namespace tsk

// One translation unit per kind and lane-group width instantiates its kernels
// and exports their lookup by workgroup size and depth: decode_<kind>.hip,
// compiled once for each width with -DTSK_DECODE_GROUP=1, 2, 4, 8 and 16, so
// the builds of the 75 kernels of a kind run side by side.
#define TSK_TABLE_NAME2(KIND_NAME, G) tsk_decode_table_##KIND_NAME##_g##G
#define TSK_TABLE_NAME(KIND_NAME, G) TSK_TABLE_NAME2(KIND_NAME, G)
#define TSK_DEFINE_DECODE_TABLE(KIND_NAME, KIND)                                                             \
    extern "C" tsk::DecodeFn TSK_TABLE_NAME(KIND_NAME, TSK_DECODE_GROUP)(std::uint32_t block, std::uint32_t depth) { \
        return tsk::pick_block<KIND, TSK_DECODE_GROUP>(block, depth);                                         \
    }
