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
//    words.  Groups fully inside the region are written with 16-byte stores
//    when aligned.
#pragma once

#include <hip/hip_runtime.h>

#include <tilestream_kernels.h>

#include <cstdint>

namespace tsk {

// Largest payload (64 x 64 float32 RAW) in words; with up to 3 words of 16-byte
// misalignment ahead of it, at most 1025 16-byte blocks are staged.  The slack
// covers the farthest read past a row end: a lane of the last segment reads up
// to (G S - 1) b / 32 + 1 <= 256 words beyond the row's first packed word.
constexpr std::uint32_t kPayloadWords = 4096;
constexpr std::uint32_t kMaxBlocks = (kPayloadWords + 3 + 3) / 4;
constexpr std::uint32_t kStageWords = kMaxBlocks * 4 + 280;

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

// Stores G samples (or their calibrated values) for tile columns col0..col0+G-1.
template <int G, bool U16, bool CAL>
__device__ inline void emit(const tsk_decode_image &img, const Region &r, std::uint32_t col0, const Group<G> &g) {
    const bool inside = col0 >= r.left && col0 + G <= r.right;
    const std::int64_t ex0 = r.ex_tile + col0;
    if constexpr (CAL) {
        float out[G];
        const gptr<const std::uint8_t> orow = img.offset_map ? global(img.offset_map) + r.ey * img.offset_pitch : nullptr;
        const gptr<const std::uint8_t> grow = img.gain_map ? global(img.gain_map) + r.ey * img.gain_pitch : nullptr;
#pragma unroll
        for (int k = 0; k < G; ++k) {
            const std::uint32_t c = col0 + k;
            float o = img.offset, gn = img.gain;
            if (c >= r.left && c < r.right) {
                if (orow) o = reinterpret_cast<gptr<const float>>(orow)[ex0 + k];
                if (grow) gn = reinterpret_cast<gptr<const float>>(grow)[ex0 + k];
            }
            // Two roundings to nearest even, never contracted into an FMA.
            out[k] = __fmul_rn(__fsub_rn(sample_float(g.v[k], U16), o), gn);
        }
        gptr<float> dst = reinterpret_cast<gptr<float>>(r.line) + ex0;
        if (inside && G >= 4 && (reinterpret_cast<std::uintptr_t>(dst) & 15u) == 0) {
#pragma unroll
            for (int k = 0; k + 4 <= G; k += 4)
                *reinterpret_cast<gptr<f32x4>>(dst + k) = f32x4{out[k], out[k + 1], out[k + 2], out[k + 3]};
        } else {
#pragma unroll
            for (int k = 0; k < G; ++k)
                if (col0 + k >= r.left && col0 + k < r.right) dst[k] = out[k];
        }
    } else if constexpr (U16) {
        gptr<std::uint16_t> dst = reinterpret_cast<gptr<std::uint16_t>>(r.line) + ex0;
        constexpr int kBytes = 2 * G;
        if (G >= 2 && inside && (reinterpret_cast<std::uintptr_t>(dst) & ((kBytes >= 16 ? 16 : kBytes) - 1)) == 0) {
            std::uint32_t w[G / 2 > 0 ? G / 2 : 1];
#pragma unroll
            for (int i = 0; i < G / 2; ++i) w[i] = (g.v[2 * i] & 0xffffu) | (g.v[2 * i + 1] << 16);
            if constexpr (G >= 8) {
#pragma unroll
                for (int i = 0; i + 4 <= G / 2; i += 4)
                    *reinterpret_cast<gptr<u32x4>>(dst + 2 * i) = u32x4{w[i], w[i + 1], w[i + 2], w[i + 3]};
            } else if constexpr (G == 4) {
                *reinterpret_cast<gptr<u32x2>>(dst) = u32x2{w[0], w[1]};
            } else if constexpr (G == 2) {
                *reinterpret_cast<gptr<std::uint32_t>>(dst) = w[0];
            }
        } else {
#pragma unroll
            for (int k = 0; k < G; ++k)
                if (col0 + k >= r.left && col0 + k < r.right) dst[k] = static_cast<std::uint16_t>(g.v[k]);
        }
    } else {
        gptr<std::uint32_t> dst = reinterpret_cast<gptr<std::uint32_t>>(r.line) + ex0;
        constexpr int kBytes = 4 * G;
        if (inside && (reinterpret_cast<std::uintptr_t>(dst) & ((kBytes >= 16 ? 16 : kBytes) - 1)) == 0) {
            if constexpr (G >= 4) {
#pragma unroll
                for (int k = 0; k + 4 <= G; k += 4)
                    *reinterpret_cast<gptr<u32x4>>(dst + k) = u32x4{g.v[k], g.v[k + 1], g.v[k + 2], g.v[k + 3]};
            } else if constexpr (G == 2) {
                *reinterpret_cast<gptr<u32x2>>(dst) = u32x2{g.v[0], g.v[1]};
            } else {
                dst[0] = g.v[0];
            }
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
template <int G, bool U16, bool CAL, int BLOCK>
__device__ void decode_tile(const tsk_decode_image &img, const TileView &t, const std::uint32_t *stage,
                            const std::uint16_t *row_start) {
    const std::uint32_t S = (img.tile_w / G) < 64u ? img.tile_w / G : 64u;
    const std::uint32_t passes = img.tile_w / (S * G);
    const std::uint32_t p = threadIdx.x % S;
    const std::uint32_t rows_per_step = BLOCK / S;
    const std::uint32_t *payload = stage + t.base;

    for (std::uint32_t row = t.top + threadIdx.x / S; row < t.bottom; row += rows_per_step) {
        Region r;
        r.ey = t.ty + row - img.roi_y;
        r.line = global(img.out) + r.ey * img.out_pitch;
        r.ex_tile = static_cast<std::int64_t>(t.tx) - static_cast<std::int64_t>(img.roi_x);
        r.left = t.left;
        r.right = t.right;
        if (t.mode == 2) {
            const std::uint32_t start = row_start[row];
            const std::uint32_t seed = payload[start];
            const std::uint32_t b = payload[start + 1];
            const std::uint32_t mask = b >= 32 ? 0xffffffffu : ((1u << b) - 1u);
            const std::uint32_t *words = payload + start + 2;
            std::uint32_t carry = seed;
            for (std::uint32_t pass = 0; pass < passes; ++pass) {
                const std::uint32_t col0 = (pass * S + p) * G;
                // Residual j of this lane is residual col0 + j of the row: it moves
                // from column col0 + j to col0 + j + 1.  Bit offsets are below
                // 2^13, so 24-bit multiplies (full rate) suffice.
                const std::uint32_t bit0 = __umul24(col0, b);
                std::uint32_t acc = 0;
                Group<G> pre;
#pragma unroll
                for (int j = 0; j < G; ++j) {
                    const std::uint32_t bit = __umul24(static_cast<std::uint32_t>(j), b) + bit0;
                    const std::uint32_t w = bit >> 5;
                    const std::uint32_t z = __builtin_amdgcn_alignbit(words[w + 1], words[w], bit) & mask;
                    acc = Ring<U16>::add(acc, Ring<U16>::term(z));
                    pre.v[j] = acc;
                }
                const std::uint32_t incl = segment_scan<U16>(acc, p, S);
                const std::uint32_t first = Ring<U16>::add(carry, Ring<U16>::sub(incl, acc));
                if (passes > 1)
                    carry = Ring<U16>::add(carry, static_cast<std::uint32_t>(
                                                      __builtin_amdgcn_readlane(static_cast<int>(incl), 63)));
                Group<G> v;
                v.v[0] = first;
#pragma unroll
                for (int k = 1; k < G; ++k) v.v[k] = Ring<U16>::add(first, pre.v[k - 1]);
                emit<G, U16, CAL>(img, r, col0, v);
            }
        } else {
            for (std::uint32_t pass = 0; pass < passes; ++pass) {
                const std::uint32_t col0 = (pass * S + p) * G;
                Group<G> v;
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
                emit<G, U16, CAL>(img, r, col0, v);
            }
        }
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

// Request mode: the request holding task i, found by a binary search between
// the sampled index entries around i (at most TSK_REQUEST_STRIDE requests, all
// loads scalar), and the tile of its region.
__device__ inline tsk_decode_task request_task(const tsk_decode_args &a, std::uint64_t i, std::uint32_t &request) {
    const std::uint64_t slot = i / TSK_REQUEST_STRIDE;
    std::uint32_t r = load_constant(a.request_index, slot);
    std::uint32_t hi = (slot + 1) * TSK_REQUEST_STRIDE < a.task_count ? load_constant(a.request_index, slot + 1)
                                                                      : a.request_count - 1;
    // Invariant: requests[r].first_task <= i < requests[hi + 1].first_task.
    while (r < hi) {
        const std::uint32_t mid = r + (hi - r + 1) / 2;
        if (load_constant(a.requests, mid).first_task <= i)
            r = mid;
        else
            hi = mid - 1;
    }
    const tsk_request q = load_constant(a.requests, r);
    const std::uint32_t local = static_cast<std::uint32_t>(i - q.first_task);
    const std::uint32_t dr = local / q.cols;
    const std::uint32_t col = q.col0 + (local - dr * q.cols);
    const std::uint32_t row = q.row0 + dr;
    tsk_decode_task t = load_constant(a.tasks, std::uint64_t{q.tile_base} + std::uint64_t{row} * q.tile_columns + col);
    t.column = col;
    t.row = row;
    request = r;
    return t;
}

// Reads a task descriptor and, on the trusted path, the producer's result and
// the device-resident index entry (reference decoder.cu:191-220).
__device__ inline Pending describe(const tsk_decode_args &a, std::uint64_t i) {
    std::uint32_t request = 0;
    const tsk_decode_task task = a.request_count != 0 ? request_task(a, i, request) : load_constant(a.tasks, i);
    Pending p;
    p.image = a.request_count != 0 ? request : task.image;
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

// Work assignment: a static stride, or chunks drawn from a device counter with
// the next chunk requested one chunk ahead.
struct Cursor {
    std::uint64_t next, end, ahead;
};

template <int BLOCK>
__device__ inline std::uint64_t draw(const tsk_decode_args &a, Cursor &c, std::uint64_t *shared_next) {
    if (a.queue == nullptr) {
        const std::uint64_t i = c.next;
        c.next += gridDim.x;
        return i;
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
        const std::uint64_t start = shared_next[0];
        c.next = (static_cast<std::uint64_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(start >> 32))) << 32) |
                 static_cast<std::uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(start)));
        c.end = c.next + a.chunk;
    }
    return c.next++;
}

template <int G, int KIND, int BLOCK, int DEPTH>
__global__ void __launch_bounds__(BLOCK) decode_kernel(tsk_decode_args a) {
    __shared__ __attribute__((aligned(16))) std::uint32_t stage[kStageWords];
    __shared__ std::uint16_t row_start[64];
    __shared__ std::uint32_t shared_flag;
    __shared__ std::uint64_t shared_next[1];
    constexpr int KMAX = DEPTH > 1 ? static_cast<int>((kMaxBlocks + BLOCK - 1) / BLOCK) : 1;

    const std::uint32_t tid = threadIdx.x;
    Cursor cursor{blockIdx.x, 0, 0};
    if (a.queue != nullptr) {
        if (tid == 0) cursor.ahead = atomicAdd(reinterpret_cast<unsigned long long *>(a.queue), a.chunk);
        cursor.next = cursor.end = 0;
    }
    std::uint64_t i = draw<BLOCK>(a, cursor, shared_next);
    if (i >= a.task_count) return;

    Pending p = describe(a, i);
    u32x4 reg[KMAX];
    std::uint32_t row_reg = 0;
    if constexpr (DEPTH > 1) {
#pragma unroll
        for (int k = 0; k < KMAX; ++k)
            if (tid + k * BLOCK < p.nblocks) reg[k] = p.blocks[tid + k * BLOCK];
        if (p.mode == 2 && p.rows != TSK_NO_ROWS && tid < p.h && p.rows + tid < a.row_count)
            row_reg = (global(a.rows) + p.rows)[tid];
    }

    for (;;) {
        const Pending c = p;
        __syncthreads(); // the previous tile is done with the LDS
        if constexpr (DEPTH > 1) {
#pragma unroll
            for (int k = 0; k < KMAX; ++k)
                if (tid + k * BLOCK < c.nblocks) *reinterpret_cast<u32x4 *>(stage + 4 * (tid + k * BLOCK)) = reg[k];
            if (c.mode == 2 && c.rows != TSK_NO_ROWS && tid < c.h && c.rows + tid < a.row_count)
                row_start[tid] = static_cast<std::uint16_t>(row_reg);
        } else {
            for (std::uint32_t k = tid; k < c.nblocks; k += BLOCK)
                *reinterpret_cast<u32x4 *>(stage + 4 * k) = c.blocks[k];
            if (c.mode == 2 && c.rows != TSK_NO_ROWS && tid < c.h && c.rows + tid < a.row_count)
                row_start[tid] = (global(a.rows) + c.rows)[tid];
        }
        __syncthreads();

        // Start fetching the next tile; its loads are in flight while this one decodes.
        const std::uint64_t n = draw<BLOCK>(a, cursor, shared_next);
        const bool more = n < a.task_count;
        if (more) {
            p = describe(a, n);
            if constexpr (DEPTH > 1) {
#pragma unroll
                for (int k = 0; k < KMAX; ++k)
                    if (tid + k * BLOCK < p.nblocks) reg[k] = p.blocks[tid + k * BLOCK];
                if (p.mode == 2 && p.rows != TSK_NO_ROWS && tid < p.h && p.rows + tid < a.row_count)
                    row_reg = (global(a.rows) + p.rows)[tid];
            }
        }

        if (c.state == 2) {
            if (tid == 0) fail(a.results, c.image, kMalformed);
        } else if (c.state == 0) {
            // The descriptor is copied by value: output stores cannot alias it.
            const tsk_decode_image img = load_constant(a.images, c.image);
            const bool u16 = img.type == TSK_UINT16;
            bool good = true;
            if (c.mode == 2 && c.rows == TSK_NO_ROWS) {
                // Trusted encoder output: serial walk of the row descriptors in LDS
                // (reference decoder.cu:86-110).
                if (tid == 0) {
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
                    shared_flag = ok && cur == c.words;
                }
                __syncthreads();
                good = shared_flag != 0;
                if (!good && tid == 0) fail(a.results, c.image, kMalformed);
            }
            if (good) {
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
                if constexpr (KIND == TSK_KIND_U16) {
                    decode_tile<G, true, false, BLOCK>(img, t, stage, row_start);
                } else if constexpr (KIND == TSK_KIND_F32) {
                    decode_tile<G, false, false, BLOCK>(img, t, stage, row_start);
                } else if constexpr (KIND == TSK_KIND_U16_CAL) {
                    decode_tile<G, true, true, BLOCK>(img, t, stage, row_start);
                } else if constexpr (KIND == TSK_KIND_F32_CAL) {
                    decode_tile<G, false, true, BLOCK>(img, t, stage, row_start);
                } else {
                    if (u16) {
                        if (img.calibrated) decode_tile<G, true, true, BLOCK>(img, t, stage, row_start);
                        else decode_tile<G, true, false, BLOCK>(img, t, stage, row_start);
                    } else {
                        if (img.calibrated) decode_tile<G, false, true, BLOCK>(img, t, stage, row_start);
                        else decode_tile<G, false, false, BLOCK>(img, t, stage, row_start);
                    }
                }
            }
        }
        if (!more) break;
    }
}

using DecodeFn = void (*)(tsk_decode_args);

// The (group, block, depth) combinations that are compiled: every group with
// 256 threads and depth 1; groups 4, 8, 16 with every block size and depth,
// depth 2 only for 256 and 512 threads (its register prefetch holds
// ceil(1025 / block) 16-byte blocks per thread).
constexpr bool compiled(std::uint32_t group, std::uint32_t block, std::uint32_t depth) {
    const bool g_ok = group == 1 || group == 2 || group == 4 || group == 8 || group == 16;
    const bool b_ok = block == 64 || block == 128 || block == 256 || block == 512;
    if (!g_ok || !b_ok || depth < 1 || depth > 2) return false;
    if (group <= 2) return block == 256 && depth == 1;
    return depth == 1 || block >= 256;
}

template <int KIND, int G> DecodeFn pick_block(std::uint32_t block, std::uint32_t depth) {
    if constexpr (G <= 2) {
        return block == 256 && depth == 1 ? decode_kernel<G, KIND, 256, 1> : nullptr;
    } else {
        if (depth == 1) {
            switch (block) {
            case 64: return decode_kernel<G, KIND, 64, 1>;
            case 128: return decode_kernel<G, KIND, 128, 1>;
            case 256: return decode_kernel<G, KIND, 256, 1>;
            case 512: return decode_kernel<G, KIND, 512, 1>;
            default: return nullptr;
            }
        }
        if (depth == 2) {
            switch (block) {
            case 256: return decode_kernel<G, KIND, 256, 2>;
            case 512: return decode_kernel<G, KIND, 512, 2>;
            default: return nullptr;
            }
        }
        return nullptr;
    }
}

template <int KIND> DecodeFn pick(std::uint32_t group, std::uint32_t block, std::uint32_t depth) {
    switch (group) {
    case 1: return pick_block<KIND, 1>(block, depth);
    case 2: return pick_block<KIND, 2>(block, depth);
    case 4: return pick_block<KIND, 4>(block, depth);
    case 8: return pick_block<KIND, 8>(block, depth);
    case 16: return pick_block<KIND, 16>(block, depth);
    default: return nullptr;
    }
}

} // namespace tsk

// One translation unit per kind instantiates its kernels and exports this lookup.
#define TSK_DEFINE_DECODE_TABLE(NAME, KIND)                                                          \
    extern "C" tsk::DecodeFn NAME(std::uint32_t group, std::uint32_t block, std::uint32_t depth) { \
        return tsk::pick<KIND>(group, block, depth);                                               \
    }
