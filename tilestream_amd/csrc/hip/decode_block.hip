// Block decoder for gfx950: the batched decode of planned archives with a
// write pattern the HBM serves at full rate.
//
// Output written one tile at a time costs HBM efficiency: a streaming kernel
// that has each workgroup move one 64x64 uint16 tile (8 KiB of output as 64
// rows of 128 bytes) per step reaches about 4.6 TB/s of reads plus writes on
// this hardware; the same kernel moving eight neighbouring tiles per step
// reaches 5.3-6.0 TB/s, eight tiles far apart only 4.4-4.6.  A block is
// therefore up to eight neighbouring tiles of one image, `across` x `down`,
// whose output is at most 64 rows of at most 1 KiB (uint16 output: 512
// columns, 32-bit output: 256):
//
//  * a workgroup stages the payloads of all the block's tiles in LDS with
//    16-byte loads straight into LDS (global_load_lds_dwordx4, no registers
//    held), the row offsets of its predictive rows beside them;
//  * every wave decodes whole block rows: lane l owns G = 16 / (output sample
//    bytes) consecutive columns of one tile row (the residual extraction,
//    segment scan and calibration of the tile decoder, decode_core.h), so
//    each store instruction writes one contiguous 1 KiB run of an output row;
//  * two workgroups share a CU (the stage is 67 KiB): while one waits for its
//    block's payloads the other decodes; the descriptor of a workgroup's next
//    block (its tile records and image record inline, 416 bytes) is loaded
//    into LDS while the current block decodes.
//
// Blocks are drawn like the tile decoder's tasks: seven eighths by static
// stride with each XCD on a contiguous run, the rest in chunks from a device
// counter.  Only host-planned tiles are decoded here (parsed archives, staged
// payloads); encoder output decoded without host metadata uses the tile
// decoder, which validates the device-resident index.
#include "decode_core.h"

namespace tsk {
namespace {

static_assert(sizeof(tsk_block) == 416, "a block record is 26 16-byte blocks");
static_assert(sizeof(tsk_decode_image) % 4 == 0, "image records are whole words");
static_assert(sizeof(tsk_block_tile) == 32, "tile records are 32 bytes");

constexpr std::uint32_t kStageBlocks = TSK_BLOCK_STAGE;
constexpr std::uint32_t kMaxTiles = 8;
constexpr std::uint32_t kRowSlots = 64;     // row offsets per tile in LDS
constexpr std::uint32_t kRecordBlocks = 26; // 16-byte blocks of a block record
constexpr std::uint32_t kRecordSlot = 64;   // 16-byte blocks reserved per record (one wave's load)

using lds_u32x4 = __attribute__((address_space(3))) u32x4;
using lds_void = __attribute__((address_space(3))) void;
using global_void = const __attribute__((address_space(1))) void;

// One wave loads 16 bytes per active lane from `src` (per lane) into
// consecutive 16-byte LDS blocks starting at `dst` (the same for every lane):
// lane l writes block l.
__device__ inline void load_to_lds(const void *src, u32x4 *dst) {
    __builtin_amdgcn_global_load_lds((global_void *)reinterpret_cast<std::uintptr_t>(src), (lds_void *)(lds_u32x4 *)dst,
                                     16, 0, 0);
}

__device__ inline std::uint32_t uniform(std::uint32_t v) {
    return static_cast<std::uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(v)));
}

// LDS of a workgroup.
struct Shared {
    u32x4 stage[kStageBlocks];
    u32x4 record[2][kRecordSlot]; // block records, alternating by round
    std::uint16_t rows[kMaxTiles * kRowSlots];
    std::uint64_t next[1];
};

// Requests block i's record into LDS slot `slot` (lanes below 26 of wave 0).
__device__ inline void request_record(const tsk_decode_args &a, std::uint64_t i, Shared &sh, std::uint32_t slot) {
    if (threadIdx.x < kRecordBlocks)
        load_to_lds(reinterpret_cast<const u32x4 *>(a.blocks + i) + threadIdx.x, &sh.record[slot][0]);
}

// The uniform fields of a staged block record.
struct BlockJob {
    std::uint32_t image;
    std::uint32_t across, down;
    std::uint32_t col0, row0;
    std::uint32_t count;
    std::uint32_t blocks;
    std::uint32_t flags;
};

// The staged copy of the image record, in scalar registers.
__device__ inline tsk_decode_image read_image(const tsk_block &b) {
    constexpr unsigned kWords = sizeof(tsk_decode_image) / 4;
    const std::uint32_t *src = reinterpret_cast<const std::uint32_t *>(&b.image_copy);
    std::uint32_t words[kWords];
#pragma unroll
    for (unsigned k = 0; k < kWords; ++k) words[k] = uniform(src[k]);
    tsk_decode_image img;
    __builtin_memcpy(&img, words, sizeof(img));
    return img;
}

__device__ inline BlockJob read_job(const tsk_block &b) {
    BlockJob j;
    j.image = uniform(b.image);
    j.across = uniform(b.across);
    j.down = uniform(b.down);
    j.col0 = uniform(b.col0);
    j.row0 = uniform(b.row0);
    const std::uint32_t count = uniform(b.count), blocks = uniform(b.blocks);
    j.count = count < kMaxTiles ? count : kMaxTiles;
    j.blocks = blocks < TSK_BLOCK_STAGE_USED ? blocks : TSK_BLOCK_STAGE_USED;
    j.flags = uniform(b.flags);
    return j;
}

// Issues the loads of the staged block's payloads into the stage (slot s of
// wave w covers stage blocks [(s * waves + w) * 64, + 64); a wave's slot is
// loaded when its first block is staged, and lanes past the last block
// reload it, so at most TSK_BLOCK_STAGE_USED rounded up to 64 blocks are
// written) and returns this thread's row offset entry, loaded into a
// register: entry e = 64 t + r is row r of tile t.
template <int BLOCK>
__device__ inline std::uint32_t request_payloads(const tsk_decode_args &a, const tsk_block &b, const BlockJob &j,
                                                 Shared &sh, bool &row_live) {
    constexpr std::uint32_t WAVES = BLOCK / 64;
    constexpr std::uint32_t SLOTS = (kStageBlocks + BLOCK - 1) / BLOCK;
    const std::uint32_t wave = threadIdx.x / 64, lane = threadIdx.x & 63u;
    const std::uint32_t last = j.blocks != 0 ? j.blocks - 1 : 0u;
    // The tile staged at stage block k: the last one whose stage starts at or
    // before k.
    std::uint32_t stage_at[kMaxTiles];
    std::uintptr_t origin[kMaxTiles];
#pragma unroll
    for (std::uint32_t t = 0; t < kMaxTiles; ++t) {
        const std::uintptr_t p = reinterpret_cast<std::uintptr_t>(b.tiles[t].payload);
        const std::uint32_t at = uniform(b.tiles[t].stage);
        stage_at[t] = t < j.count ? at : 0xffffffffu;
        origin[t] = (p & ~std::uintptr_t{15}) - std::uintptr_t{at} * 16u;
    }
#pragma unroll
    for (std::uint32_t s = 0; s < SLOTS; ++s) {
        const std::uint32_t first = (s * WAVES + wave) * 64u;
        if (first >= j.blocks) break;
        const std::uint32_t k = first + lane;
        const std::uint32_t blk = k < last ? k : last;
        std::uintptr_t src = origin[0] + std::uintptr_t{blk} * 16u;
#pragma unroll
        for (std::uint32_t t = 1; t < kMaxTiles; ++t)
            if (blk >= stage_at[t]) src = origin[t] + std::uintptr_t{blk} * 16u;
        load_to_lds(reinterpret_cast<const void *>(src), &sh.stage[first]);
    }
    // Row offsets of predictive rows (512 entries, one per thread).
    static_assert(BLOCK >= static_cast<int>(kMaxTiles * kRowSlots), "one row entry per thread");
    std::uint32_t out = 0;
    row_live = false;
    const std::uint32_t e = threadIdx.x;
    if (e < kMaxTiles * kRowSlots) {
        const std::uint32_t t = e / kRowSlots, r = e % kRowSlots;
        const tsk_block_tile &bt = b.tiles[t];
        const std::uint32_t mode = bt.shape & 0xffu, h = (bt.shape >> 8) & 0xffu;
        row_live = t < j.count && mode == 2 && bt.rows != TSK_NO_ROWS && r < h && std::uint64_t{bt.rows} + r < a.row_count;
        if (row_live) out = global(a.rows)[bt.rows + r];
    }
    return out;
}

// Whether every output row of the block starts on a 128-byte line and spans
// whole lines, with no region edge inside it: then no other block writes
// into its lines and the stores are non-temporal.
__device__ inline bool block_whole_lines(const tsk_decode_image &img, const BlockJob &j) {
    const std::uint32_t s = img.calibrated || img.type != TSK_UINT16 ? 4u : 2u;
    const std::uint64_t x0 = std::uint64_t{j.col0} * img.tile_w;
    const std::uint64_t x1 = x0 + std::uint64_t{j.across} * img.tile_w;
    const bool inside = x0 >= img.roi_x && x1 <= std::uint64_t{img.roi_x} + img.roi_w;
    const std::uintptr_t first = reinterpret_cast<std::uintptr_t>(img.out) + (x0 - img.roi_x) * s;
    return inside && img.out_pitch % 128 == 0 && first % 128 == 0 && ((x1 - x0) * s) % 128 == 0;
}

__device__ inline std::uint32_t log2_of(std::uint32_t power_of_two) {
    return 31u - static_cast<std::uint32_t>(__builtin_clz(power_of_two));
}

__device__ inline std::uint32_t payload_base(const tsk_block_tile &bt) {
    return bt.stage * 4u + static_cast<std::uint32_t>((reinterpret_cast<std::uintptr_t>(bt.payload) >> 2) & 3u);
}

// Decodes and stores every region row of the staged block (any tile modes,
// partial tiles, region edges).
template <bool U16, bool CAL, int BLOCK, bool NT, bool MAPS>
__device__ inline void block_rows(const tsk_decode_image &img, const BlockJob &j, const tsk_block &b,
                                  const std::uint32_t *stage, const std::uint16_t *rows) {
    constexpr int G = (U16 && !CAL) ? 8 : 4;
    const std::uint32_t W = img.tile_w, H = img.tile_h;
    const std::uint32_t lg_s = log2_of(W / G), lg_h = log2_of(H); // lanes per tile row: 8 .. 64
    const std::uint32_t S = 1u << lg_s;
    const std::uint32_t lane = threadIdx.x & 63u;
    const std::uint32_t kc = lane >> lg_s, p = lane & (S - 1u);
    const std::uint32_t col0 = p * G;
    const bool column_live = kc < j.across;
    const std::uint64_t tx = std::uint64_t{j.col0 + (column_live ? kc : 0u)} * W;
    // The region's columns within this lane's tile.
    const std::uint64_t rx = img.roi_x, rx1 = rx + img.roi_w;
    const std::uint32_t left = rx > tx ? static_cast<std::uint32_t>(rx - tx) : 0u;
    const std::uint32_t right = rx1 < tx + W ? static_cast<std::uint32_t>(rx1 - tx) : W;
    const std::uint64_t ry0 = img.roi_y, ry1 = ry0 + img.roi_h;
    const std::uint32_t total = j.down << lg_h;
    for (std::uint32_t y = threadIdx.x / 64u; y < total; y += BLOCK / 64) {
        const std::uint32_t kr = y >> lg_h, ry = y & (H - 1u);
        const std::uint64_t ty = std::uint64_t{j.row0 + kr} * H + ry; // archive row
        const std::uint32_t t = kr * j.across + (column_live ? kc : 0u);
        const tsk_block_tile &bt = b.tiles[t];
        const std::uint32_t shape = bt.shape;
        const std::uint32_t mode = shape & 0xffu, h = (shape >> 8) & 0xffu, w = shape >> 16;
        if (!column_live || ry >= h || ty < ry0 || ty >= ry1 || left >= right) continue;
        Region r;
        r.ey = ty - ry0;
        r.line = global(img.out) + r.ey * img.out_pitch;
        r.ex_tile = static_cast<std::int64_t>(tx) - static_cast<std::int64_t>(rx);
        r.left = left;
        r.right = right < w ? right : w;
        const std::uint32_t *payload = stage + payload_base(bt);
        Group<G> v;
        if (mode == 2) {
            const RowBits rb = row_bits(payload, rows[t * kRowSlots + ry]);
            std::uint32_t incl;
            predict_group<G, U16>(rb, col0, p, S, rb.seed, v, incl);
        } else if (mode == 0) {
            const std::uint32_t idx0 = __umul24(ry, w) + col0;
#pragma unroll
            for (int k = 0; k < G; ++k) {
                const std::uint32_t idx = idx0 + k;
                v.v[k] = U16 ? (payload[idx >> 1] >> ((idx & 1u) * 16)) : payload[idx];
            }
        } else {
#pragma unroll
            for (int k = 0; k < G; ++k) v.v[k] = payload[0];
        }
        if (col0 < r.right && col0 + G > r.left) emit<G, U16, CAL, NT, MAPS>(img, r, col0, v);
    }
}

// Stores a group of G samples (or calibrated values) at a 16-byte aligned
// output address: one 16-byte store.  MAPS: the image has calibration maps.
template <int G, bool U16, bool CAL, bool NT, bool MAPS>
__device__ inline void store_group(const tsk_decode_image &img, gptr<std::uint8_t> line, std::uint64_t ex0,
                                   std::uint64_t ey, const Group<G> &g) {
    if constexpr (CAL) {
        float o[G], gn[G];
        if (MAPS && img.offset_map != nullptr) {
            load_floats<G>(reinterpret_cast<gptr<const float>>(global(img.offset_map) + ey * img.offset_pitch) + ex0, o);
        } else {
#pragma unroll
            for (int k = 0; k < G; ++k) o[k] = img.offset;
        }
        if (MAPS && img.gain_map != nullptr) {
            load_floats<G>(reinterpret_cast<gptr<const float>>(global(img.gain_map) + ey * img.gain_pitch) + ex0, gn);
        } else {
#pragma unroll
            for (int k = 0; k < G; ++k) gn[k] = img.gain;
        }
        std::uint32_t out[G];
#pragma unroll
        for (int k = 0; k < G; ++k)
            // Two roundings to nearest even, never contracted into an FMA.
            out[k] = __float_as_uint(__fmul_rn(__fsub_rn(sample_float(g.v[k], U16), o[k]), gn[k]));
        store_words<G, NT>(reinterpret_cast<gptr<std::uint32_t>>(line) + ex0, out);
    } else if constexpr (U16) {
        std::uint32_t w[G / 2];
#pragma unroll
        for (int i = 0; i < G / 2; ++i) w[i] = __builtin_amdgcn_perm(g.v[2 * i + 1], g.v[2 * i], 0x05040100u);
        store_words<G / 2, NT>(reinterpret_cast<gptr<std::uint32_t>>(reinterpret_cast<gptr<std::uint16_t>>(line) + ex0), w);
    } else {
        store_words<G, NT>(reinterpret_cast<gptr<std::uint32_t>>(line) + ex0, g.v);
    }
}

// A full block (TSK_BLOCK_FULL) inside the region whose output rows start
// on words: every lane decodes and stores one group in every row it
// covers, with no mode, edge or region tests.  S (lanes per tile row) is a
// constant, so the segment scan has no branches and the unrolled rows are
// straight-line code whose LDS reads the compiler can overlap.
template <bool U16, bool CAL, int BLOCK, bool NT, bool MAPS, std::uint32_t S>
__device__ inline void block_rows_full(const tsk_decode_image &img, const BlockJob &j, const tsk_block &b,
                                       const std::uint32_t *stage, const std::uint16_t *rows) {
    constexpr int G = (U16 && !CAL) ? 8 : 4;
    constexpr std::uint32_t WAVES = BLOCK / 64;
    constexpr int PER_WAVE = static_cast<int>(64 / WAVES);
    constexpr std::uint32_t W = S * G;
    const std::uint32_t H = img.tile_h;
    const std::uint32_t lg_h = log2_of(H);
    const std::uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x / 64u;
    const std::uint32_t kc = lane / S, p = lane % S;
    const std::uint32_t col0 = p * G;
    const std::uint64_t ex0 = std::uint64_t{j.col0 + kc} * W - img.roi_x + col0;
    const std::uint64_t ey0 = std::uint64_t{j.row0} * H - img.roi_y;
    // The output row of this wave, advanced by `waves` rows per step.
    gptr<std::uint8_t> line = global(img.out) + (ey0 + wave) * img.out_pitch;
    const std::uint64_t line_step = std::uint64_t{WAVES} * img.out_pitch;
#pragma unroll
    for (int i = 0; i < PER_WAVE; ++i) {
        const std::uint32_t y = wave + static_cast<std::uint32_t>(i) * WAVES;
        const std::uint32_t kr = y >> lg_h, ry = y & (H - 1u);
        const std::uint32_t t = kr * j.across + kc;
        const tsk_block_tile &bt = b.tiles[t];
        const RowBits rb = row_bits(stage + payload_base(bt), rows[t * kRowSlots + ry]);
        Group<G> v;
        std::uint32_t incl;
        predict_group<G, U16>(rb, col0, p, S, rb.seed, v, incl);
        store_group<G, U16, CAL, NT, MAPS>(img, line, ex0, ey0 + y, v);
        line += line_step;
    }
}

// Whether the block takes the full-block path: all its tiles predictive and
// full size, 64 rows, inside the region both ways, output rows on words (a
// group's store is one multi-word access at word alignment, so a region at
// any column keeps its inner blocks on this path).
__device__ inline bool block_is_full(const tsk_decode_image &img, const BlockJob &j) {
    if ((j.flags & TSK_BLOCK_FULL) == 0) return false;
    const std::uint32_t s = img.calibrated || img.type != TSK_UINT16 ? 4u : 2u;
    const std::uint64_t x0 = std::uint64_t{j.col0} * img.tile_w, y0 = std::uint64_t{j.row0} * img.tile_h;
    const std::uint64_t x1 = x0 + std::uint64_t{j.across} * img.tile_w, y1 = y0 + std::uint64_t{j.down} * img.tile_h;
    if (x0 < img.roi_x || y0 < img.roi_y || x1 > std::uint64_t{img.roi_x} + img.roi_w ||
        y1 > std::uint64_t{img.roi_y} + img.roi_h || y1 - y0 != 64)
        return false;
    const std::uintptr_t first = reinterpret_cast<std::uintptr_t>(img.out) + (x0 - img.roi_x) * s;
    return first % 4 == 0 && img.out_pitch % 4 == 0;
}

template <bool U16, bool CAL, int BLOCK, bool MAPS>
__device__ inline __attribute__((always_inline)) void block_rows_as(const tsk_decode_image &img, const BlockJob &j, const tsk_block &b,
                                     const std::uint32_t *stage, const std::uint16_t *rows) {
    constexpr std::uint32_t G = (U16 && !CAL) ? 8 : 4;
    const bool nt = block_whole_lines(img, j);
    if (block_is_full(img, j)) {
        switch (img.tile_w / G) {
        case 8:
            if constexpr (G == 8) {
                if (nt) block_rows_full<U16, CAL, BLOCK, true, MAPS, 8>(img, j, b, stage, rows);
                else block_rows_full<U16, CAL, BLOCK, false, MAPS, 8>(img, j, b, stage, rows);
            }
            return;
        case 16:
            if (nt) block_rows_full<U16, CAL, BLOCK, true, MAPS, 16>(img, j, b, stage, rows);
            else block_rows_full<U16, CAL, BLOCK, false, MAPS, 16>(img, j, b, stage, rows);
            return;
        case 32:
            if (nt) block_rows_full<U16, CAL, BLOCK, true, MAPS, 32>(img, j, b, stage, rows);
            else block_rows_full<U16, CAL, BLOCK, false, MAPS, 32>(img, j, b, stage, rows);
            return;
        default:
            if constexpr (G == 4) {
                if (nt) block_rows_full<U16, CAL, BLOCK, true, MAPS, 64>(img, j, b, stage, rows);
                else block_rows_full<U16, CAL, BLOCK, false, MAPS, 64>(img, j, b, stage, rows);
            }
            return;
        }
    } else if (nt) {
        block_rows<U16, CAL, BLOCK, true, MAPS>(img, j, b, stage, rows);
    } else {
        block_rows<U16, CAL, BLOCK, false, MAPS>(img, j, b, stage, rows);
    }
}

// SPLIT: calibrated images with maps and without take separate paths (chosen
// per block), so that the store loops of an image without maps issue no
// loads (a wait for a load also waits for every store issued before it).
// The mixed kind keeps one path, which tests for maps per group: its kernel
// holds every pixel type and calibration, and a second copy of the
// calibrated paths would push it past its register budget.
template <bool U16, bool CAL, int BLOCK, bool SPLIT>
__device__ inline __attribute__((always_inline)) void block_rows_any(const tsk_decode_image &img, const BlockJob &j,
                                                                     const tsk_block &b, const std::uint32_t *stage,
                                                                     const std::uint16_t *rows) {
    if constexpr (CAL && SPLIT) {
        if (has_maps(img)) block_rows_as<U16, CAL, BLOCK, true>(img, j, b, stage, rows);
        else block_rows_as<U16, CAL, BLOCK, false>(img, j, b, stage, rows);
    } else {
        block_rows_as<U16, CAL, BLOCK, CAL>(img, j, b, stage, rows);
    }
}

template <int KIND, int BLOCK>
__device__ inline __attribute__((always_inline)) void block_decode(const tsk_decode_image &img, const BlockJob &j,
                                                                   const tsk_block &b, const std::uint32_t *stage,
                                                                   const std::uint16_t *rows) {
    if constexpr (KIND == TSK_KIND_U16) {
        block_rows_any<true, false, BLOCK, true>(img, j, b, stage, rows);
    } else if constexpr (KIND == TSK_KIND_F32) {
        block_rows_any<false, false, BLOCK, true>(img, j, b, stage, rows);
    } else if constexpr (KIND == TSK_KIND_U16_CAL) {
        block_rows_any<true, true, BLOCK, true>(img, j, b, stage, rows);
    } else if constexpr (KIND == TSK_KIND_F32_CAL) {
        block_rows_any<false, true, BLOCK, true>(img, j, b, stage, rows);
    } else if (img.type == TSK_UINT16) {
        if (img.calibrated) block_rows_any<true, true, BLOCK, false>(img, j, b, stage, rows);
        else block_rows_any<true, false, BLOCK, false>(img, j, b, stage, rows);
    } else {
        if (img.calibrated) block_rows_any<false, true, BLOCK, false>(img, j, b, stage, rows);
        else block_rows_any<false, false, BLOCK, false>(img, j, b, stage, rows);
    }
}

// Two workgroups per CU (the LDS holds two stages): waves per SIMD are
// 2 x BLOCK / 256, which at 512 threads leaves 128 registers per lane.
template <int KIND, int BLOCK>
__global__ void __launch_bounds__(BLOCK, 2 * BLOCK / 256) block_decode_kernel(tsk_decode_args a) {
    __shared__ __attribute__((aligned(16))) Shared sh;
    const std::uint32_t *words = reinterpret_cast<const std::uint32_t *>(sh.stage);

    Cursor cursor = start_cursor(a);
    if (a.queue != nullptr) queue_start(a, cursor);
    std::uint64_t i = draw<BLOCK>(a, cursor, sh.next);
    if (i >= a.task_count) return;
    std::uint32_t slot = 0;
    request_record(a, i, sh, slot);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    for (;;) {
        const tsk_block &b = *reinterpret_cast<const tsk_block *>(&sh.record[slot][0]);
        const BlockJob j = read_job(b);
        bool row_live = false;
        const std::uint32_t row = request_payloads<BLOCK>(a, b, j, sh, row_live);
        // The next block's record arrives while this block decodes.
        const std::uint64_t n = draw<BLOCK>(a, cursor, sh.next);
        if (n < a.task_count) request_record(a, n, sh, slot ^ 1u);
        // The image record is copied by value: output stores cannot alias it.
        const tsk_decode_image img = read_image(b);
        __builtin_amdgcn_s_waitcnt(0); // this wave's payload, row and record loads
        if (row_live) sh.rows[threadIdx.x] = static_cast<std::uint16_t>(row);
        __syncthreads();
        block_decode<KIND, BLOCK>(img, j, b, words, sh.rows);
        if (n >= a.task_count) return;
        __syncthreads(); // every wave is done with the stage and this record
        i = n;
        slot ^= 1u;
    }
}

// Block records of a batch of region requests, one wave per block (four per
// workgroup): the wave finds its request and its place in the request's
// block grid, lane k reads tile k's entry of the resident tile table, the
// stage offsets are the lanes' running sum of staged 16-byte blocks, and the
// 104 words of the record are stored by consecutive lanes.  The same record
// as the host's block planning (plan_blocks) makes for a planned batch.
static_assert(offsetof(tsk_block, tiles) == 32 && offsetof(tsk_block, image_copy) == 288 &&
                  sizeof(tsk_decode_image) == 120 && sizeof(tsk_block_tile) == 32,
              "record words: 8 header, 8 per tile, 30 image, 2 pad");

__global__ void __launch_bounds__(256) build_blocks_kernel(tsk_block_build_args a) {
    const std::uint32_t lane = threadIdx.x & 63u;
    const std::uint64_t b = std::uint64_t{blockIdx.x} * 4u + threadIdx.x / 64u;
    if (b >= a.block_count) return; // uniform per wave
    const std::uint32_t r = a.block_index[b];
    const tsk_request q = a.requests[r];
    const std::uint32_t qa = q.across, qd = q.down;
    const std::uint32_t local = static_cast<std::uint32_t>(b - q.first_block);
    const std::uint32_t nbx = (q.cols + qa - 1) / qa;
    const std::uint32_t by = local / nbx, bx = local - by * nbx;
    const std::uint32_t across = min(qa, q.cols - bx * qa), down = min(qd, q.rows - by * qd);
    const std::uint32_t col0 = q.col0 + bx * qa, row0 = q.row0 + by * qd;
    const std::uint32_t count = across * down;
    const tsk_decode_image *img = a.images + r;
    const std::uint32_t tile_w = img->tile_w, tile_h = img->tile_h;

    // Tile `lane` of the block (lanes below count): its record words.
    std::uint32_t f[6] = {0, 0, 0, 0, 0, 0}; // payload low, high, words, shape, rows, stage
    std::uint32_t staged = 0;
    bool full = true;
    if (lane < count) {
        const std::uint32_t kr = lane / across, kc = lane - kr * across;
        const tsk_decode_task t =
            a.tiles[std::uint64_t{q.tile_base} + std::uint64_t{row0 + kr} * q.tile_columns + col0 + kc];
        const std::uintptr_t p = reinterpret_cast<std::uintptr_t>(t.payload);
        staged = static_cast<std::uint32_t>(((p + std::uint64_t{t.words} * 4u + 15u) >> 4) - (p >> 4));
        full = t.shape == TSK_TASK_SHAPE(2u, tile_h, tile_w) && t.rows != TSK_NO_ROWS;
        f[0] = static_cast<std::uint32_t>(p);
        f[1] = static_cast<std::uint32_t>(p >> 32);
        f[2] = t.words;
        f[3] = t.shape;
        f[4] = t.rows;
    }
    std::uint32_t stage = 0, total = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const std::uint32_t v = static_cast<std::uint32_t>(__shfl(static_cast<int>(staged), k));
        if (static_cast<std::uint32_t>(k) < lane) stage += v;
        total += v;
    }
    f[5] = lane < count ? stage : 0u;
    const bool all_full = __ballot(lane < count && !full) == 0;
    const std::uint32_t flags = across == qa && down == qd && all_full ? TSK_BLOCK_FULL : 0u;

    auto *out = reinterpret_cast<std::uint32_t *>(a.blocks + b);
    const auto *image_words = reinterpret_cast<const std::uint32_t *>(img);
#pragma unroll
    for (std::uint32_t pass = 0; pass < 2; ++pass) {
        const std::uint32_t w = lane + 64u * pass;
        // Tile words: field (w - 8) % 8 of tile (w - 8) / 8, from that tile's lane.
        const std::uint32_t src = w >= 8u ? ((w - 8u) / 8u) & 7u : 0u, field = (w - 8u) & 7u;
        std::uint32_t tile_word = 0;
#pragma unroll
        for (std::uint32_t k = 0; k < 6; ++k) {
            const std::uint32_t v = static_cast<std::uint32_t>(__shfl(static_cast<int>(f[k]), static_cast<int>(src)));
            if (field == k) tile_word = v;
        }
        std::uint32_t v = 0;
        if (w < 8u) {
            v = w == 0 ? r : w == 1 ? (across | (down << 16)) : w == 2 ? col0 : w == 3 ? row0 : w == 4 ? count
                : w == 5 ? total : w == 6 ? flags : 0u;
        } else if (w < 72u) {
            v = tile_word;
        } else if (w < 102u) {
            v = image_words[w - 72u];
        }
        if (w < sizeof(tsk_block) / 4) out[w] = v;
    }
}

template <int BLOCK> DecodeFn block_function(std::uint32_t kind) {
    switch (kind) {
    case TSK_KIND_U16: return block_decode_kernel<TSK_KIND_U16, BLOCK>;
    case TSK_KIND_F32: return block_decode_kernel<TSK_KIND_F32, BLOCK>;
    case TSK_KIND_U16_CAL: return block_decode_kernel<TSK_KIND_U16_CAL, BLOCK>;
    case TSK_KIND_F32_CAL: return block_decode_kernel<TSK_KIND_F32_CAL, BLOCK>;
    case TSK_KIND_MIXED: return block_decode_kernel<TSK_KIND_MIXED, BLOCK>;
    default: return nullptr;
    }
}

} // namespace
} // namespace tsk

extern "C" {

int tsk_build_blocks(const tsk_block_build_args *args, void *stream) {
    if (args->block_count == 0) return hipSuccess;
    const std::uint64_t grid = (args->block_count + 3) / 4;
    if (grid > 0x7fffffffu) return hipErrorInvalidValue;
    hipLaunchKernelGGL(tsk::build_blocks_kernel, dim3(static_cast<std::uint32_t>(grid)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), *args);
    return hipGetLastError();
}

// The block decoder of a kind for 512 threads per workgroup (one row offset
// entry per thread).
tsk::DecodeFn tsk_block_table(std::uint32_t kind, std::uint32_t block) {
    if (block == 512) return tsk::block_function<512>(kind);
    return nullptr;
}

} // extern "C"
