// Canonical GPU encoder for gfx950: the same bytes as the CPU encoder, in one
// pass over the input.
//
// One launch per window of tiles (reference src/cuda/encoder.cu:290-327 runs
// analyze, a device scan, pack and finish per window; here they are one
// kernel).  Workgroup b encodes window tile b.  Each XCD dispatches its
// workgroups in increasing order, so the smallest unfinished tile is always
// running and waits only on finished tiles.  A workgroup:
//   stage    the tile's samples into LDS (16-byte loads when rows are aligned);
//   analyze  residuals, per-row minimal widths, the three candidate costs and
//            the mode with tie order RAW < CONSTANT < PREDICT_BITPACK
//            (reference src/cuda/encoder.cu:62-138, docs/format.md
//            "Canonical encoder"), with lane-group reductions per row;
//   place    publishes its payload length, then finds its payload offset by a
//            decoupled look-back over the tiles before it, restarting at
//            the first tile of its image (a segmented scan; reference
//            src/cuda/encoder.cu:284-288 uses a device-wide scan instead);
//   pack     builds the payload in LDS while the offset is still unknown
//            (each lane ORs its residuals, held in registers since the
//            analysis, into the row bit streams), then writes the index entry
//            and the payload (reference src/cuda/encoder.cu:175-238); the
//            last tile of an image writes
//            the 64-byte header and the result record (reference
//            src/cuda/encoder.cu:240-280).
// The look-back waits are bounded: a workgroup that waits too long marks its
// published prefix as failed, the mark travels with every prefix and carried
// cursor built on it, and the image's result reports hip_error instead of a
// hang.  Two look-back state buffers alternate by window parity: a window's
// workgroups clear the other buffer for the next window, so no workgroup
// needs a device-wide fence (which writes back the L2 of its XCD).
#include <hip/hip_runtime.h>

#include <tilestream_kernels.h>

#include <atomic>
#include <cstdint>
#include <mutex>
#include <type_traits>

namespace {

// Per pixel type: threads per workgroup (one tile each) and the LDS sample.
// uint16 tiles keep 16-bit samples and use 128 threads, so twice as many
// tiles fit a CU; each lane holds 4096 / threads columns of one tile row.
template <bool U16> struct Cfg {
    static constexpr std::uint32_t threads = U16 ? 128u : 256u;
    // Each lane owns `columns` consecutive columns of one tile row.
    static constexpr std::uint32_t columns = 4096u / threads;
    // Waves per SIMD the registers are budgeted for: 14 uint16 or 8 float32 tiles per CU.
    static constexpr std::uint32_t waves = U16 ? 7u : 8u;
    using sample = std::conditional_t<U16, std::uint16_t, std::uint32_t>;
    // In LDS a lane's columns are a segment of columns + pad samples and a row
    // is `lanes` segments: every lane of a wave then reads a distinct bank at
    // each step (pad 1 for 32-bit, 2 for 16-bit samples).
    static constexpr std::uint32_t segment = columns + (U16 ? 2u : 1u);
    static constexpr std::uint32_t words = threads * segment * sizeof(sample) / 4u;
};

// LDS place of sample (r, c): row stride = lanes x segment samples.
template <bool U16> __device__ inline std::uint32_t place(std::uint32_t r, std::uint32_t c, std::uint32_t row) {
    return r * row + (c / Cfg<U16>::columns) * Cfg<U16>::segment + c % Cfg<U16>::columns;
}
constexpr std::uint64_t kFlagAggregate = 1ull << 62;    // value: this tile's length
constexpr std::uint64_t kFlagPrefix = 2ull << 62;       // value: lengths from the image's first tile here through this one
constexpr std::uint64_t kFlagMask = 3ull << 62;
constexpr std::uint64_t kFailed = 1ull << 61;           // a wait before this value timed out
constexpr std::uint64_t kValueMask = kFailed - 1;
constexpr std::uint32_t kSpinLimit = 1u << 20;
constexpr std::uint32_t kHipError = 9;

// Global memory through address-space-1 pointers: generic (FLAT) accesses
// count in the LDS wait counter too, so every LDS wait would drain them.
template <typename T> using gptr = T __attribute__((address_space(1))) *;
typedef std::uint32_t u32x4 __attribute__((ext_vector_type(4)));
// Multi-word global stores need only word alignment on CDNA.
typedef std::uint32_t u32x4w __attribute__((ext_vector_type(4), aligned(4)));
template <typename T> __device__ inline gptr<T> global(T *p) { return (gptr<T>)p; }

struct DeviceResultBits {
    std::uint32_t status;
    std::uint32_t reserved;
    std::uint64_t bytes;
};

struct Geometry {
    std::uint64_t x, y;
    std::uint32_t w, h;
};

__device__ inline Geometry geometry(const tsk_encode_image &img, std::uint64_t tile) {
    Geometry g;
    std::uint64_t row;
    if (tile >> 32 == 0) {
        row = static_cast<std::uint32_t>(tile) / img.tile_columns; // a 32-bit division where it suffices
    } else {
        row = tile / img.tile_columns;
    }
    g.x = (tile - row * img.tile_columns) * img.tile_w;
    g.y = row * img.tile_h;
    g.w = static_cast<std::uint32_t>(img.width - g.x < img.tile_w ? img.width - g.x : img.tile_w);
    g.h = static_cast<std::uint32_t>(img.height - g.y < img.tile_h ? img.height - g.y : img.tile_h);
    return g;
}

// Window entries are ordered by their first tile (scalar loads: every lane
// of the workgroup asks for the same entries).
__device__ inline std::uint32_t owner(const tsk_window_entry *e, std::uint32_t n, std::uint32_t tile) {
    std::uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const std::uint32_t mid = lo + (hi - lo) / 2;
        if (e[mid].first <= tile) lo = mid;
        else hi = mid;
    }
    return lo;
}

__device__ inline std::uint64_t payload_start(const tsk_encode_image &img) { return 64 + 32 * img.tile_count; }

template <bool U16> __device__ inline std::uint32_t residual(std::uint32_t cur, std::uint32_t prev) {
    if constexpr (U16) {
        const std::uint32_t d = (cur - prev) & 0xffffu;
        return d < 32768u ? 2u * d : 2u * (65536u - d) - 1u;
    } else {
        return cur ^ prev;
    }
}

typedef std::uint16_t u16x2 __attribute__((ext_vector_type(2)));
typedef std::int16_t i16x2 __attribute__((ext_vector_type(2)));

// residual<true> of both 16-bit halves at once (packed 16-bit integer ops).
__device__ inline std::uint32_t residual_pair(std::uint32_t cur, std::uint32_t prev) {
    const u16x2 d = __builtin_bit_cast(u16x2, cur) - __builtin_bit_cast(u16x2, prev);
    const u16x2 sign = __builtin_bit_cast(u16x2, __builtin_bit_cast(i16x2, d) >> static_cast<std::int16_t>(15));
    return __builtin_bit_cast(std::uint32_t, static_cast<u16x2>((d << static_cast<std::uint16_t>(1)) ^ sign));
}

// ORs `bits` into an LSB-first bit stream at bit `pos`: two LDS ORs, the
// second of zero when the bits end inside the first word.
__device__ inline void or_bits(std::uint32_t *stream, std::uint32_t pos, std::uint32_t bits) {
    const std::uint64_t v = std::uint64_t{bits} << (pos & 31u);
    std::uint32_t *w = stream + (pos >> 5);
    atomicOr(w, static_cast<std::uint32_t>(v));
    atomicOr(w + 1, static_cast<std::uint32_t>(v >> 32));
}

// floor(p / d) as a multiply-high by ceil(2^32 / d): exact while p * d < 2^32
// (p below 2^17 and d at most 256 here).
__device__ inline std::uint32_t magic_of(std::uint32_t d) { return static_cast<std::uint32_t>(0xffffffffu / d) + 1u; }
__device__ inline std::uint32_t div_magic(std::uint32_t p, std::uint32_t magic, std::uint32_t d) {
    return d == 1u ? p : __umulhi(p, magic);
}

__device__ inline std::uint64_t load_state(const std::uint64_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void store_state(std::uint64_t *p, std::uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline std::uint64_t wave_sum(std::uint64_t v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// Copies the tile into LDS at place(r, c), one 16- or 32-bit word per
// sample.  Every load of a thread is issued before the first one is waited
// for (a tile is at most 4 chunks of 16 bytes, or 16 samples, per thread).
template <bool U16>
__device__ inline void stage_tile(const tsk_encode_image &img, const Geometry &g, std::uint32_t row,
                                  typename Cfg<U16>::sample *px) {
    constexpr std::uint32_t kThreads = Cfg<U16>::threads;
    constexpr std::uint32_t s = U16 ? 2u : 4u;
    const std::uint8_t *origin = img.in + g.y * img.in_pitch + g.x * s;
    const bool vector = ((reinterpret_cast<std::uintptr_t>(img.in) | img.in_pitch) & 15u) == 0;
    if (vector) {
        // Rows start 16-byte aligned (tile widths are multiples of 64 samples).
        constexpr std::uint32_t per = 16u / s;
        constexpr std::uint32_t most = 4096u / per / kThreads; // chunks per thread
        const std::uint32_t across = (g.w + per - 1u) / per;
        const std::uint32_t chunks = across * g.h;
        const std::uint32_t magic = magic_of(across);
        u32x4 v[most];
#pragma unroll
        for (std::uint32_t q = 0; q < most; ++q) {
            const std::uint32_t k = threadIdx.x + q * kThreads;
            const std::uint32_t r = div_magic(k, magic, across), c = k - r * across;
            // Samples are read once: non-temporal loads.
            if (k < chunks && (c + 1u) * per <= g.w)
                v[q] = __builtin_nontemporal_load(reinterpret_cast<gptr<const u32x4>>(
                    global(origin + std::uint64_t{r} * img.in_pitch + c * 16u)));
        }
#pragma unroll
        for (std::uint32_t q = 0; q < most; ++q) {
            const std::uint32_t k = threadIdx.x + q * kThreads;
            if (k >= chunks) continue;
            const std::uint32_t r = div_magic(k, magic, across), c = k - r * across;
            // A chunk never crosses a segment (segments hold a multiple of `per` columns).
            typename Cfg<U16>::sample *dst = px + place<U16>(r, c * per, row);
            if ((c + 1u) * per <= g.w) {
                // 32-bit writes: 8 uint16 samples start at a 4-byte boundary (even places).
                std::uint32_t *d32 = reinterpret_cast<std::uint32_t *>(dst);
                d32[0] = v[q].x;
                d32[1] = v[q].y;
                d32[2] = v[q].z;
                d32[3] = v[q].w;
            } else {
                // The image's right edge: only the samples inside it.
                const std::uint8_t *p = origin + std::uint64_t{r} * img.in_pitch + c * 16u;
                for (std::uint32_t e = 0; c * per + e < g.w; ++e)
                    dst[e] = U16 ? reinterpret_cast<gptr<const std::uint16_t>>(global(p))[e]
                                 : reinterpret_cast<gptr<const std::uint32_t>>(global(p))[e];
            }
        }
    } else {
        // Rounds of 16 samples per thread, all loads of a round issued first.
        constexpr std::uint32_t round = 16;
        const std::uint32_t n = g.w * g.h;
        const std::uint32_t magic = magic_of(g.w);
        for (std::uint32_t k0 = 0; k0 < 4096u; k0 += round * kThreads) {
            std::uint32_t v[round];
#pragma unroll
            for (std::uint32_t q = 0; q < round; ++q) {
                const std::uint32_t k = k0 + threadIdx.x + q * kThreads;
                const std::uint32_t r = div_magic(k, magic, g.w), c = k - r * g.w;
                const std::uint8_t *p = origin + std::uint64_t{r} * img.in_pitch;
                if (k < n)
                    v[q] = U16 ? reinterpret_cast<gptr<const std::uint16_t>>(global(p))[c]
                               : reinterpret_cast<gptr<const std::uint32_t>>(global(p))[c];
            }
#pragma unroll
            for (std::uint32_t q = 0; q < round; ++q) {
                const std::uint32_t k = k0 + threadIdx.x + q * kThreads;
                const std::uint32_t r = div_magic(k, magic, g.w), c = k - r * g.w;
                if (k < n) px[place<U16>(r, c, row)] = static_cast<typename Cfg<U16>::sample>(v[q]);
            }
        }
    }
}

// Window tile of workgroup b.  Workgroups are dispatched to the eight XCDs
// round robin (b on XCD b % 8), and each XCD dispatches its own in increasing
// order.  In every full group of 8 x run workgroups, XCD x takes the run of
// tiles [x run, (x + 1) run) of the group, in increasing order; the tail
// beyond the last full group is taken in order.  The rows of neighbouring
// tiles share cache lines, which then stay in one L2.  Look-back safety: the
// smallest unfinished tile T follows, in its XCD's order, only finished tiles
// (earlier groups, and the earlier tiles of its own run), so its XCD has room
// for it; workgroups that wait (tiles above T) and come before T in the
// dispatch order belong to T's group and number fewer than run per XCD, so
// with run at most 32 (one workgroup per CU of an XCD) they cannot fill an XCD
// and stall the dispatch of T.
__device__ inline std::uint32_t window_tile(std::uint32_t b, std::uint32_t n, std::uint32_t run) {
    const std::uint32_t group = 8u * run;
    if (run <= 1u || b >= n / group * group) return b;
    const std::uint32_t base = b / group * group, r = b - base;
    return base + (r % 8u) * run + r / 8u;
}

template <bool U16>
__global__ void __launch_bounds__(Cfg<U16>::threads) __attribute__((amdgpu_waves_per_eu(Cfg<U16>::waves))) encode_kernel(tsk_encode_args a) {
    constexpr std::uint32_t kThreads = Cfg<U16>::threads;
    constexpr std::uint32_t kColumns = Cfg<U16>::columns;
    using Sample = typename Cfg<U16>::sample;
    // The tile's samples; after the analysis of a predictive tile, its payload words.
    __shared__ __attribute__((aligned(16))) std::uint32_t lds[Cfg<U16>::words];
    Sample *px = reinterpret_cast<Sample *>(lds);
    __shared__ std::uint32_t row_word[65];
    __shared__ std::uint32_t row_bits[64];
    __shared__ std::uint32_t wave_cost[kThreads / 64];
    __shared__ std::uint64_t shared_offset;

    const std::uint32_t lane = threadIdx.x % 64u;
    const std::uint32_t t = window_tile(blockIdx.x, a.tile_count, a.run);
    std::uint64_t *state = a.state + std::uint64_t{a.parity} * a.capacity;
    std::uint32_t o;
    if (a.blocks != nullptr) {
        // The block's entry, then forward over entries starting later in the block.
        o = a.blocks[t >> 6];
        while (o + 1u < a.entry_count && a.entries[o + 1u].first <= t) ++o;
    } else {
        o = owner(a.entries, a.entry_count, t);
    }
    const tsk_window_entry e = a.entries[o];
    const tsk_encode_image &img = a.images[e.image];
    const std::uint64_t index = e.start + (t - e.first);
    const Geometry g = geometry(img, index);
    // Lanes per tile row: threads / tile_h; lane sl of row r owns columns
    // [sl x kColumns, (sl + 1) x kColumns).
    const std::uint32_t lanes = kThreads / img.tile_h;
    const std::uint32_t row = lanes * Cfg<U16>::segment;
    const std::uint32_t r = threadIdx.x / lanes, sl = threadIdx.x % lanes;
    const std::uint32_t c0 = sl * kColumns;

    stage_tile<U16>(img, g, row, px);
    __syncthreads();

    // ---- analyze -----------------------------------------------------------------
    const std::uint32_t first = px[0];
    bool same = true;
    std::uint32_t any = 0, seed = 0;
    // Residuals stay in registers until the packing; uint16 residuals (at
    // most 16 bits) two to a register.  The sample before a lane's first
    // column is the previous lane's last; column 0 is taken against itself
    // (residual 0, never packed: the row's seed stands for it).
    constexpr std::uint32_t kPer = U16 ? 2u : 1u;
    std::uint32_t z[kColumns / kPer];
#pragma unroll
    for (std::uint32_t i = 0; i < kColumns / kPer; ++i) z[i] = 0;
    const bool full = r < g.h && c0 + kColumns <= g.w;
    if (r < g.h && c0 < g.w) {
        const Sample *lane_px = px + place<U16>(r, c0, row);
        const std::uint32_t before = lane_px[c0 != 0 ? -static_cast<int>(Cfg<U16>::segment - kColumns + 1u) : 0];
        seed = lane_px[0];
        if (full) {
            // Every column valid: no per-column conditions.
            std::uint32_t diff = 0;
            if constexpr (U16) {
                const std::uint32_t *lane_w = reinterpret_cast<const std::uint32_t *>(lane_px); // even place
                const std::uint32_t first2 = first * 0x10001u;
                std::uint32_t prev_w = before << 16;
#pragma unroll
                for (std::uint32_t j = 0; j < kColumns / 2u; ++j) {
                    const std::uint32_t w = lane_w[j];
                    // Previous samples of the two halves: prev_w's high half and w's low half.
                    z[j] = residual_pair(w, __builtin_amdgcn_alignbit(w, prev_w, 16));
                    any |= z[j];
                    diff |= w ^ first2;
                    prev_w = w;
                }
                any = (any | (any >> 16)) & 0xffffu;
            } else {
                std::uint32_t prev = before;
#pragma unroll
                for (std::uint32_t i = 0; i < kColumns; ++i) {
                    const std::uint32_t v = lane_px[i];
                    z[i] = v ^ prev;
                    any |= z[i];
                    diff |= v ^ first;
                    prev = v;
                }
            }
            same = diff == 0;
        } else {
            std::uint32_t prev = before;
#pragma unroll
            for (std::uint32_t i = 0; i < kColumns; ++i) {
                if (c0 + i < g.w) {
                    const std::uint32_t v = lane_px[i];
                    same = same && v == first;
                    const std::uint32_t zi = residual<U16>(v, prev);
                    z[i / kPer] |= zi << (U16 ? 16u * (i % 2u) : 0u);
                    any |= zi;
                    prev = v;
                }
            }
        }
    }
    for (std::uint32_t off = lanes / 2u; off > 0; off >>= 1) any |= __shfl_xor(any, static_cast<int>(off));
    std::uint32_t cost = 0;
    if (sl == 0 && r < g.h) {
        const std::uint32_t b = any == 0 ? 0u : 32u - static_cast<std::uint32_t>(__builtin_clz(any));
        row_bits[r] = b;
        cost = 8u + (((g.w - 1u) * b + 31u) / 32u) * 4u;
    }
    for (int off = 32; off > 0; off >>= 1) cost += __shfl_xor(cost, off);
    if (lane == 0) wave_cost[threadIdx.x / 64u] = cost;
    const bool uniform = __syncthreads_and(same) != 0;
    std::uint32_t total = 0;
#pragma unroll
    for (std::uint32_t w = 0; w < kThreads / 64u; ++w) total += wave_cost[w];
    const std::uint32_t raw4 = (g.w * g.h * (U16 ? 2u : 4u) + 3u) / 4u * 4u;
    std::uint32_t mode = 0, length = raw4;
    if (uniform && length > 4u) {
        mode = 1;
        length = 4;
    }
    if (total < length) {
        mode = 2;
        length = total;
    }

    // ---- place: publish the length, then look back for the offset ----------------
    const bool segment_start = t == e.first;
    // A continued image starts at the cursor the previous window left.
    const std::uint64_t carried = (e.flags & 1u) ? a.carry[a.parity] : payload_start(img);
    if (threadIdx.x == 0)
        store_state(&state[t], segment_start ? (kFlagPrefix | (carried & kFailed) | length) : (kFlagAggregate | length));

    if (mode == 2) {
        // The payload is built in LDS before the offset is known.  Words of each
        // row: seed, width, packed residuals; row_word is their exclusive prefix.
        if (threadIdx.x < 64) {
            const std::uint32_t words = lane < g.h ? 2u + ((g.w - 1u) * row_bits[lane] + 31u) / 32u : 0u;
            std::uint32_t incl = words;
            for (std::uint32_t off = 1; off < 64; off <<= 1) {
                const std::uint32_t other = __shfl_up(incl, off);
                if (lane >= off) incl += other;
            }
            if (lane == 0) row_word[0] = 0;
            row_word[lane + 1] = incl;
        }
        // The samples were last read before the barrier above: their area
        // becomes the payload, zeroed first.
        for (std::uint32_t k = threadIdx.x * 4u; k < length / 4u; k += kThreads * 4u)
            *reinterpret_cast<u32x4 *>(lds + k) = u32x4{0u, 0u, 0u, 0u};
        __syncthreads();
        if (r < g.h) {
            // Each lane ORs its residuals (still in registers) into the row's
            // LSB-first bit stream, residual of column c at bit (c - 1) x b.
            // Every write to the zeroed payload is an OR, so lanes sharing a
            // word (or an OR of zero past a row's end) need no ordering.
            const std::uint32_t b = row_bits[r], base = row_word[r] + 2u;
            if (sl == 0) {
                atomicOr(&lds[base - 2u], seed);
                atomicOr(&lds[base - 1u], b);
            }
            std::uint32_t *stream = lds + base;
            if (b != 0 && full) {
                if constexpr (U16) {
                    // Residual pairs (2 b <= 32 bits) of columns c0 + 2j, c0 + 2j + 1.
#pragma unroll
                    for (std::uint32_t j = 0; j < kColumns / 2u; ++j) {
                        const std::uint32_t hi = z[j] >> 16;
                        std::uint32_t bits = (z[j] & 0xffffu) | (hi << b);
                        std::uint32_t pos = (c0 + 2u * j - 1u) * b;
                        if (j == 0 && c0 == 0) {
                            bits = hi; // column 0 carries no residual
                            pos = 0;
                        }
                        or_bits(stream, pos, bits);
                    }
                } else {
#pragma unroll
                    for (std::uint32_t i = 0; i < kColumns; ++i)
                        if (i != 0 || c0 != 0) or_bits(stream, (c0 + i - 1u) * b, z[i]);
                }
            } else if (b != 0 && c0 < g.w) {
#pragma unroll
                for (std::uint32_t i = 0; i < kColumns; ++i) {
                    const std::uint32_t c = c0 + i;
                    if (c != 0 && c < g.w)
                        or_bits(stream, (c - 1u) * b, U16 ? (z[i / kPer] >> (16u * (i % 2u))) & 0xffffu : z[i]);
                }
            }
        }
    }
    if (threadIdx.x < 64) {
        std::uint64_t exclusive = 0;
        bool failed = (carried & kFailed) != 0;
        if (!segment_start) {
            std::int64_t top = static_cast<std::int64_t>(t) - 1;
            std::uint32_t spins = 0;
            for (;;) {
                // One 128-byte line of look-back words per round trip (a poll's
                // latency grows with the lines it touches): lane k reads tile
                // top - k, down to the first word of top's line.
                const std::uint32_t width =
                    static_cast<std::uint32_t>((reinterpret_cast<std::uintptr_t>(&state[top]) >> 3) & 15u) + 1u;
                const bool active = lane < width;
                const std::int64_t j = top - static_cast<std::int64_t>(lane);
                const std::uint64_t s =
                    active && j >= static_cast<std::int64_t>(e.first) ? load_state(&state[j]) : 0ull;
                const std::uint64_t flag = s & kFlagMask;
                const std::uint64_t prefix_lanes = __ballot(active && flag == kFlagPrefix);
                const std::uint64_t empty_lanes = __ballot(active && flag == 0ull);
                // The nearest published prefix, and whether every tile before it has published.
                const std::uint32_t stop =
                    prefix_lanes != 0 ? static_cast<std::uint32_t>(__builtin_ctzll(prefix_lanes)) : width;
                const std::uint64_t needed = stop == width ? ((1ull << width) - 1ull) : ((2ull << stop) - 1ull);
                if ((empty_lanes & needed) == 0) {
                    const bool mine = active && lane <= stop;
                    exclusive += wave_sum(mine ? (s & kValueMask) : 0ull);
                    failed = failed || __ballot(mine && (s & kFailed) != 0) != 0;
                    if (stop != width) break;
                    top -= width;
                    continue;
                }
                if (++spins > kSpinLimit) {
                    failed = true;
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
        }
        if (lane == 0) {
            const std::uint64_t mark = failed ? kFailed : 0ull;
            if (!segment_start) store_state(&state[t], kFlagPrefix | mark | ((exclusive + length) & kValueMask));
            shared_offset = mark | ((carried & kValueMask) + exclusive);
        }
    }
    __syncthreads();
    const bool failed = (shared_offset & kFailed) != 0;
    const std::uint64_t offset = shared_offset & kValueMask;

    // ---- pack ----------------------------------------------------------------------
    if (threadIdx.x < 8) {
        // 32-byte index entry: mode, reserved, width, height, offset, length, reserved.
        std::uint32_t v = 0;
        switch (threadIdx.x) {
        case 0: v = mode; break;
        case 1: v = g.w | (g.h << 16); break;
        case 2: v = static_cast<std::uint32_t>(offset); break;
        case 3: v = static_cast<std::uint32_t>(offset >> 32); break;
        case 4: v = length; break;
        default: v = 0; break;
        }
        reinterpret_cast<gptr<std::uint32_t>>(global(img.out + 64 + index * 32))[threadIdx.x] = v;
    }
    gptr<std::uint32_t> dst = reinterpret_cast<gptr<std::uint32_t>>(global(img.out + offset));
    const std::uint32_t words = length / 4u;
    if (mode == 2) {
        // 16-byte stores at the payload's word alignment, then the last words.
        const std::uint32_t quads = words / 4u;
        for (std::uint32_t k = threadIdx.x; k < quads; k += kThreads)
            *reinterpret_cast<gptr<u32x4w>>(dst + 4u * k) = *reinterpret_cast<const u32x4 *>(lds + 4u * k);
        if (threadIdx.x < words - 4u * quads) dst[4u * quads + threadIdx.x] = lds[4u * quads + threadIdx.x];
    } else if (mode == 0) {
        const std::uint32_t magic = magic_of(g.w);
        const std::uint32_t n = g.w * g.h;
        for (std::uint32_t k = threadIdx.x; k < words; k += kThreads) {
            std::uint32_t v;
            if constexpr (U16) {
                const std::uint32_t p0 = 2u * k, r0 = div_magic(p0, magic, g.w);
                v = px[place<U16>(r0, p0 - r0 * g.w, row)];
                if (p0 + 1u < n) {
                    const std::uint32_t r1 = div_magic(p0 + 1u, magic, g.w);
                    v |= std::uint32_t{px[place<U16>(r1, p0 + 1u - r1 * g.w, row)]} << 16;
                }
            } else {
                const std::uint32_t r0 = div_magic(k, magic, g.w);
                v = px[place<U16>(r0, k - r0 * g.w, row)];
            }
            dst[k] = v;
        }
    } else if (threadIdx.x == 0) {
        dst[0] = first;
    }

    // ---- the image's end: header and result, or the cursor for the next window ----
    if (threadIdx.x == 0 && t == e.first + e.count - 1u) {
        const std::uint64_t end = offset + length;
        if ((e.flags & 2u) == 0) {
            a.carry[a.parity ^ 1u] = end | (failed ? kFailed : 0ull);
        } else if (failed) {
            static_cast<DeviceResultBits *>(a.results)[e.image] = {kHipError, 0u, 0u};
        } else {
            // 64-byte file header (docs/format.md "File header").
            const std::uint64_t count = img.tile_count;
            const std::uint64_t payload = payload_start(img);
            std::uint32_t h[16];
            h[0] = 0x00435354u; // 'T' 'S' 'C' 0
            h[1] = 0x0a1a0a0du; // 13 10 26 10
            h[2] = 1u;          // major 1, minor 0
            h[3] = 64u | (img.type << 16); // header size, pixel type, flags 0
            h[4] = img.width;
            h[5] = img.height;
            h[6] = img.tile_w | (std::uint32_t{img.tile_h} << 16);
            h[7] = 32u; // index entry size, reserved 0
            h[8] = static_cast<std::uint32_t>(count);
            h[9] = static_cast<std::uint32_t>(count >> 32);
            h[10] = 64u;
            h[11] = 0u;
            h[12] = static_cast<std::uint32_t>(payload);
            h[13] = static_cast<std::uint32_t>(payload >> 32);
            h[14] = static_cast<std::uint32_t>(end);
            h[15] = static_cast<std::uint32_t>(end >> 32);
            gptr<std::uint32_t> out = reinterpret_cast<gptr<std::uint32_t>>(global(img.out));
            for (int i = 0; i < 16; ++i) out[i] = h[i];
            static_cast<DeviceResultBits *>(a.results)[e.image] = {0u, 0u, end};
        }
    }

    {
        // Clear the other buffer for the next window (the previous one is done
        // with it), last: a store in flight would hold up this tile's first loads.
        std::uint64_t *other = a.state + std::uint64_t{a.parity ^ 1u} * a.capacity;
        for (std::uint32_t k = t + threadIdx.x * a.tile_count; k < a.capacity; k += kThreads * a.tile_count) other[k] = 0;
    }
}

// Device memory of the allocation-free encode entry point: both look-back
// buffers and the two carried cursors, held by every GPU the library runs on
// until the process ends.  Host code never passes this name to a kernel (on
// the host it is only a stand-in for the symbol); it uses the per-device
// address returned by tsk_scratch_region.
constexpr std::uint64_t kScratchBytes = std::uint64_t{TSK_SCRATCH_BYTES};
static __device__ __align__(256) std::uint8_t scratch[kScratchBytes];

// Device address of `scratch` on each device, looked up once per device with
// that device current.
constexpr int kMaxDevices = 64;
std::atomic<void *> scratch_address[kMaxDevices];
std::mutex scratch_lookup;

__global__ void clear_kernel(std::uint64_t *state, std::uint64_t words) {
    const std::uint64_t stride = std::uint64_t{gridDim.x} * blockDim.x;
    for (std::uint64_t k = std::uint64_t{blockIdx.x} * blockDim.x + threadIdx.x; k < words; k += stride) state[k] = 0;
}

} // namespace

extern "C" {

int tsk_encode_reset(const tsk_encode_args *a, void *stream) {
    const std::uint64_t words = 2ull * a->capacity;
    const std::uint32_t blocks = static_cast<std::uint32_t>(words / 256u + 1u < 1024u ? words / 256u + 1u : 1024u);
    hipLaunchKernelGGL(clear_kernel, dim3(blocks), dim3(256), 0, static_cast<hipStream_t>(stream), a->state, words);
    return hipGetLastError();
}

int tsk_encode_window(const tsk_encode_args *a, void *stream) {
    if (a->tile_count == 0) return hipSuccess;
    if (a->type == TSK_UINT16)
        hipLaunchKernelGGL(encode_kernel<true>, dim3(a->tile_count), dim3(Cfg<true>::threads), 0,
                           static_cast<hipStream_t>(stream), *a);
    else
        hipLaunchKernelGGL(encode_kernel<false>, dim3(a->tile_count), dim3(Cfg<false>::threads), 0,
                           static_cast<hipStream_t>(stream), *a);
    return hipGetLastError();
}

int tsk_scratch_region(void **address, std::uint64_t *bytes) {
    int device = 0;
    const hipError_t e = hipGetDevice(&device);
    if (e != hipSuccess) return e;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    void *p = scratch_address[device].load(std::memory_order_acquire);
    if (p == nullptr) {
        const std::lock_guard<std::mutex> hold(scratch_lookup);
        p = scratch_address[device].load(std::memory_order_relaxed);
        if (p == nullptr) {
            const hipError_t found = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (found != hipSuccess || p == nullptr) return found != hipSuccess ? found : hipErrorInvalidSymbol;
            scratch_address[device].store(p, std::memory_order_release);
        }
    }
    *address = p;
    *bytes = kScratchBytes;
    return hipSuccess;
}

int tsk_preload_encode(void) {
    hipFuncAttributes attr;
    const void *fns[] = {reinterpret_cast<const void *>(encode_kernel<true>), reinterpret_cast<const void *>(encode_kernel<false>),
                         reinterpret_cast<const void *>(clear_kernel)};
    for (const void *f : fns) {
        const hipError_t e = hipFuncGetAttributes(&attr, f);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

} // extern "C"
