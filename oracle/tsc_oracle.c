/*
 * TileStream .tsc v1.0 CPU oracle -- TEST INFRASTRUCTURE ONLY (see tsc_oracle.h).
 *
 * Every function restates the normative text of reference docs/format.md and
 * follows the reference CPU codec's observable behaviour (status values and the
 * order in which checks fire), cited per function.  It is deliberately simple:
 * a 64-bit reservoir bit packer/reader, scalar loops, no SIMD.
 *
 * Pinned by: the reference golden files (tests/golden hex files, hand-calculated),
 * the closed forms of the output bound and workspace size, brute force over all
 * uint16 values and all widths, decode(encode(x)) == x, and byte/status equality
 * with the compiled reference CPU codec in oracle/_ref on seeded images and on
 * every truncation / field mutation of small archives (tests/test_oracle.py).
 */
#include "tsc_oracle.h"

#include <stdlib.h>
#include <string.h>

/* L = min(INT64_MAX, SIZE_MAX) -- reference include/tilestream/sizing.h:11-14. */
#define ORC_L ((uint64_t)INT64_MAX)

/* Native sizes the reference counts against ParseLimits::metadata_bytes:
 * sizeof(ImageMetadata), sizeof(TileMetadata), sizeof(SegmentMetadata) on LP64
 * (reference include/tilestream/api.h:11-33, src/cpu/parser.cpp:173-184). */
#define ORC_IMAGE_METADATA_BYTES 72u
#define ORC_TILE_METADATA_BYTES 32u
#define ORC_SEGMENT_METADATA_BYTES 24u

static const uint8_t orc_magic[8] = {0x54, 0x53, 0x43, 0x00, 0x0d, 0x0a, 0x1a, 0x0a};

/* ---- checked arithmetic (reference src/checked_size.h:9-32) ---- */
static int ck_add(uint64_t a, uint64_t b, uint64_t *r) {
    if (a > ORC_L || b > ORC_L - a) return 0;
    *r = a + b;
    return 1;
}
static int ck_mul(uint64_t a, uint64_t b, uint64_t *r) {
    if (a > ORC_L || b > ORC_L) return 0;
    if (b != 0 && a > ORC_L / b) return 0;
    *r = a * b;
    return 1;
}
static int ck_align(uint64_t v, uint64_t a, uint64_t *r) {
    uint64_t rem;
    if (a == 0 || a > ORC_L || v > ORC_L) return 0;
    rem = v % a;
    return ck_add(v, rem == 0 ? 0 : a - rem, r);
}

static int geometry_ok(uint32_t tw, uint32_t th) {
    return (tw == 128 && th == 32) || (tw == 64 && th == 64) || (tw == 256 && th == 16);
}

static uint64_t div_up(uint64_t n, uint64_t d) { return n / d + (n % d != 0); }

static uint32_t get16(const uint8_t *p) { return (uint32_t)p[0] | ((uint32_t)p[1] << 8); }
static uint32_t get32(const uint8_t *p) {
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) |
           ((uint32_t)p[3] << 24);
}
static uint64_t get64(const uint8_t *p) {
    return (uint64_t)get32(p) | ((uint64_t)get32(p + 4) << 32);
}
static void put16(uint8_t *p, uint32_t v) {
    p[0] = (uint8_t)v;
    p[1] = (uint8_t)(v >> 8);
}
static void put32(uint8_t *p, uint32_t v) {
    put16(p, v & 0xffffu);
    put16(p + 2, v >> 16);
}
static void put64(uint8_t *p, uint64_t v) {
    put32(p, (uint32_t)v);
    put32(p + 4, (uint32_t)(v >> 32));
}
static int all_zero(const uint8_t *p, uint64_t n) {
    uint64_t i;
    for (i = 0; i < n; ++i)
        if (p[i]) return 0;
    return 1;
}

/* Sample size, type checked before dimensions (reference src/checked_size.h:51-60). */
static int sample_bytes(uint32_t type, uint32_t w, uint32_t h, uint64_t *s) {
    if (type != ORC_UINT16 && type != ORC_FLOAT32) return ORC_UNSUPPORTED_TYPE;
    if (w == 0 || h == 0) return ORC_INVALID_ARGUMENT;
    *s = type == ORC_UINT16 ? 2 : 4;
    return ORC_SUCCESS;
}

/* B = 64 + 32 N + W H s + pad  (docs/format.md "Checked output bound";
 * reference src/sizing.cpp:32-59). */
int orc_compress_bound(uint32_t type, uint32_t width, uint32_t height, uint32_t tile_width,
                       uint32_t tile_height, uint64_t *bytes) {
    uint64_t s, n, index, raw, b;
    int st = sample_bytes(type, width, height, &s);
    if (st) return st;
    if (!geometry_ok(tile_width, tile_height)) return ORC_INVALID_ARGUMENT;
    if (!ck_mul(div_up(width, tile_width), div_up(height, tile_height), &n) ||
        !ck_mul(n, 32, &index) || !ck_mul(width, height, &raw) || !ck_mul(raw, s, &raw) ||
        !ck_add(64, index, &b) || !ck_add(b, raw, &b) ||
        !ck_add(b, (s == 2 && (width & 1) && (height & 1)) ? 2 : 0, &b))
        return ORC_SIZE_OVERFLOW;
    *bytes = b;
    return ORC_SUCCESS;
}

/* span = pitch (h - 1) + w s  (reference docs/api.md "Images and views";
 * src/sizing.cpp:9-30). */
int orc_image_span(uint32_t type, uint32_t width, uint32_t height, uint64_t pitch,
                   uint64_t *bytes) {
    uint64_t s, row, span;
    int st = sample_bytes(type, width, height, &s);
    if (st) return st;
    if (!ck_mul(width, s, &row) || pitch > ORC_L) return ORC_SIZE_OVERFLOW;
    if (pitch < row || pitch % s != 0) return ORC_INVALID_ARGUMENT;
    if (!ck_mul(pitch, (uint64_t)height - 1, &span) || !ck_add(span, row, &span))
        return ORC_SIZE_OVERFLOW;
    *bytes = span;
    return ORC_SUCCESS;
}

/* Seven regions, each rounded to 256 bytes (reference docs/api.md "Workspace
 * derivation"; src/sizing.cpp:72-109). */
int orc_workspace_size(uint64_t image_count, uint64_t tile_count, uint32_t tile_width,
                       uint32_t tile_height, uint64_t scan_temp_bytes, uint64_t *bytes) {
    uint64_t counts[7], sizes[7], total = 0;
    int i;
    if (image_count == 0) {
        if (tile_count != 0 || scan_temp_bytes != 0) return ORC_INVALID_ARGUMENT;
        *bytes = 0;
        return ORC_SUCCESS;
    }
    if (!geometry_ok(tile_width, tile_height) || tile_count < image_count)
        return ORC_INVALID_ARGUMENT;
    if (image_count > UINT32_MAX || tile_count > (uint64_t)INT32_MAX) return ORC_RESOURCE_LIMIT;
    counts[0] = tile_count;     sizes[0] = 4096 * 4;
    counts[1] = tile_count;     sizes[1] = (uint64_t)tile_height * 8;
    counts[2] = tile_count;     sizes[2] = 32;
    counts[3] = tile_count;     sizes[3] = 8;
    counts[4] = tile_count + 1; sizes[4] = 8;
    counts[5] = image_count;    sizes[5] = 16;
    counts[6] = scan_temp_bytes; sizes[6] = 1;
    for (i = 0; i < 7; ++i) {
        uint64_t region;
        if (!ck_mul(counts[i], sizes[i], &region) || !ck_align(region, 256, &region) ||
            !ck_add(total, region, &total))
            return ORC_SIZE_OVERFLOW;
    }
    *bytes = total;
    return ORC_SUCCESS;
}

/* ------------------------------------------------------------------------- */
/* Parser                                                                    */
/* ------------------------------------------------------------------------- */

typedef struct {
    uint32_t type, width, height, tw, th;
    uint64_t columns, count, payload, size;
} layout_t;

/* Fixed header checks in the reference order (reference src/cpu/parser.cpp:27-84;
 * docs/format.md "File header"). */
static int parse_header(const uint8_t *d, uint64_t size, layout_t *L) {
    uint64_t bound, rows, index_bytes;
    int st;
    if (size < 64) return ORC_MALFORMED_DATA;
    if (memcmp(d, orc_magic, 8) != 0) return ORC_MALFORMED_DATA;
    if (get16(d + 8) != 1 || get16(d + 10) != 0 || d[15] != 0) return ORC_UNSUPPORTED_FORMAT;
    if (get16(d + 12) != 64 || get16(d + 28) != 32 || d[30] != 0 || d[31] != 0)
        return ORC_MALFORMED_DATA;
    if (d[14] != ORC_UINT16 && d[14] != ORC_FLOAT32) return ORC_UNSUPPORTED_TYPE;
    L->type = d[14];
    L->width = get32(d + 16);
    L->height = get32(d + 20);
    L->tw = get16(d + 24);
    L->th = get16(d + 26);
    L->size = size;
    if (L->width == 0 || L->height == 0) return ORC_MALFORMED_DATA;
    if (!geometry_ok(L->tw, L->th)) return ORC_UNSUPPORTED_FORMAT;
    st = orc_compress_bound(L->type, L->width, L->height, L->tw, L->th, &bound);
    if (st) return st;
    L->columns = div_up(L->width, L->tw);
    rows = div_up(L->height, L->th);
    if (!ck_mul(L->columns, rows, &L->count) || !ck_mul(L->count, 32, &index_bytes) ||
        !ck_add(64, index_bytes, &L->payload))
        return ORC_SIZE_OVERFLOW;
    if (get64(d + 32) != L->count || get64(d + 40) != 64 || get64(d + 48) != L->payload ||
        get64(d + 56) != size || size > bound || L->payload > size)
        return ORC_MALFORMED_DATA;
    return ORC_SUCCESS;
}

/* Index and payload walk (reference src/cpu/parser.cpp:86-171; docs/format.md
 * "Tile index", "Payload modes"). */
static int parse_tiles(const uint8_t *d, const layout_t *L, uint64_t *segments, orc_tile *tiles,
                       orc_segment *segs) {
    const uint64_t s = L->type == ORC_UINT16 ? 2 : 4;
    const uint32_t max_bits = L->type == ORC_UINT16 ? 16 : 32;
    uint64_t expect = L->payload, nseg = 0, i;
    for (i = 0; i < L->count; ++i) {
        const uint8_t *e = d + 64 + 32 * i;
        const uint64_t x = (i % L->columns) * L->tw, y = (i / L->columns) * L->th;
        const uint64_t w = (L->width - x) < L->tw ? (L->width - x) : L->tw;
        const uint64_t h = (L->height - y) < L->th ? (L->height - y) : L->th;
        const uint64_t off = get64(e + 8), len = get64(e + 16);
        const uint64_t raw = w * h * s, raw4 = (raw + 3) / 4 * 4;
        const uint32_t mode = e[0];
        const uint8_t *p;
        uint64_t first = 0;
        if (mode > ORC_PREDICT || !all_zero(e + 1, 3) || !all_zero(e + 24, 8) ||
            get16(e + 4) != w || get16(e + 6) != h || off != expect || off > L->size ||
            len > L->size - off || len == 0 || len % 4 != 0 || len > raw4)
            return ORC_MALFORMED_DATA;
        p = d + off;
        if (mode == ORC_RAW) {
            if (len != raw4 || !all_zero(p + raw, len - raw)) return ORC_MALFORMED_DATA;
        } else if (mode == ORC_CONSTANT) {
            if (len != 4 || (L->type == ORC_UINT16 && (p[2] | p[3]) != 0))
                return ORC_MALFORMED_DATA;
        } else {
            uint64_t cur = 0, r;
            first = nseg;
            for (r = 0; r < h; ++r) {
                uint32_t seed, bits;
                uint64_t nbits, words;
                if (len - cur < 8) return ORC_MALFORMED_DATA;
                seed = get32(p + cur);
                bits = p[cur + 4];
                if ((L->type == ORC_UINT16 && seed > 0xffffu) || bits > max_bits ||
                    (w == 1 && bits != 0) || !all_zero(p + cur + 5, 3))
                    return ORC_MALFORMED_DATA;
                cur += 8;
                nbits = (w - 1) * bits;
                words = (nbits + 31) / 32;
                if (words * 4 > len - cur) return ORC_MALFORMED_DATA;
                /* Unused high bits of the final word must be zero. */
                if (nbits % 32 != 0 && (get32(p + cur + words * 4 - 4) >> (nbits % 32)) != 0)
                    return ORC_MALFORMED_DATA;
                if (segs) {
                    segs[nseg].seed = seed;
                    segs[nseg].bit_width = bits;
                    segs[nseg].payload_offset = cur;
                    segs[nseg].payload_bytes = words * 4;
                }
                ++nseg;
                cur += words * 4;
            }
            if (cur != len) return ORC_MALFORMED_DATA;
        }
        if (tiles) {
            tiles[i].mode = mode;
            tiles[i].width = (uint16_t)w;
            tiles[i].height = (uint16_t)h;
            tiles[i].offset = off;
            tiles[i].length = len;
            tiles[i].first_segment = first;
        }
        expect += len;
    }
    *segments = nseg;
    return expect == L->size ? ORC_SUCCESS : ORC_MALFORMED_DATA;
}

static int meta_bytes(uint64_t tiles, uint64_t segs, uint64_t *bytes) {
    uint64_t a, b;
    if (!ck_mul(tiles, ORC_TILE_METADATA_BYTES, &a) ||
        !ck_mul(segs, ORC_SEGMENT_METADATA_BYTES, &b) ||
        !ck_add(ORC_IMAGE_METADATA_BYTES, a, bytes) || !ck_add(*bytes, b, bytes))
        return ORC_SIZE_OVERFLOW;
    return ORC_SUCCESS;
}

/* Byte-range rule for caller spans (reference src/cpu/views.h:10-23). */
static int span_ok(const void *p, uint64_t n) {
    if (n > ORC_L) return ORC_SIZE_OVERFLOW;
    if (p == NULL && n != 0) return ORC_INVALID_ARGUMENT;
    if (n > UINTPTR_MAX - (uintptr_t)p) return ORC_SIZE_OVERFLOW;
    return ORC_SUCCESS;
}

/* parse_metadata: budgets before allocation, two walks (reference
 * src/cpu/parser.cpp:187-242; docs/api.md "Host metadata versus device metadata"). */
int orc_parse(const uint8_t *file, uint64_t size, uint64_t limit_file_bytes,
              uint64_t limit_tile_count, uint64_t limit_metadata_bytes, orc_info *info,
              orc_tile *tiles, orc_segment *segments) {
    layout_t L;
    uint64_t need, nseg = 0;
    int st = span_ok(file, size);
    if (st) return st;
    if (size > limit_file_bytes) return ORC_RESOURCE_LIMIT;
    memset(&L, 0, sizeof L);
    st = parse_header(file, size, &L);
    if (st) return st;
    if (L.count > limit_tile_count) return ORC_RESOURCE_LIMIT;
    st = meta_bytes(L.count, 0, &need);
    if (st) return st;
    if (need > limit_metadata_bytes) return ORC_RESOURCE_LIMIT;
    st = parse_tiles(file, &L, &nseg, NULL, NULL);
    if (st) return st;
    st = meta_bytes(L.count, nseg, &need);
    if (st) return st;
    if (need > limit_metadata_bytes) return ORC_RESOURCE_LIMIT;
    if (tiles) {
        st = parse_tiles(file, &L, &nseg, tiles, segments);
        if (st) return st;
    }
    if (info) {
        info->type = L.type;
        info->width = L.width;
        info->height = L.height;
        info->tile_width = (uint16_t)L.tw;
        info->tile_height = (uint16_t)L.th;
        info->file_bytes = size;
        info->tile_count = L.count;
        info->segment_count = nseg;
    }
    return ORC_SUCCESS;
}

/* ------------------------------------------------------------------------- */
/* Encoder                                                                   */
/* ------------------------------------------------------------------------- */

/* Zigzag of the signed 16-bit delta (docs/format.md "uint16 transform"). */
static uint32_t zigzag16(uint32_t cur, uint32_t prev) {
    int32_t d = (int32_t)((cur - prev) & 0xffffu);
    if (d >= 32768) d -= 65536;
    return d >= 0 ? (uint32_t)(2 * d) : (uint32_t)(-2 * d - 1);
}

static uint32_t bit_length(uint32_t v) {
    uint32_t n = 0;
    while (v) {
        ++n;
        v >>= 1;
    }
    return n;
}

static uint32_t load_sample(const uint8_t *base, uint64_t pitch, uint64_t x, uint64_t y,
                            uint64_t s) {
    const uint8_t *p = base + y * pitch + x * s;
    return s == 2 ? get16(p) : get32(p);
}

/* Canonical tile encoding: per-row minimal widths, complete candidate costs and
 * the tie order RAW < CONSTANT < PREDICT_BITPACK (docs/format.md "Canonical
 * encoder and decoder acceptance"; reference src/cpu/encoder.cpp:39-118). */
static uint64_t encode_tile(const uint8_t *px, uint64_t pitch, uint64_t s, uint32_t type,
                            uint64_t x0, uint64_t y0, uint32_t w, uint32_t h, uint8_t *out,
                            uint32_t *mode_out) {
    uint32_t v[4096], res[4096], bits[64];
    uint32_t r, c, mode;
    int constant = 1;
    uint64_t pred = 0, raw4 = ((uint64_t)w * h * s + 3) / 4 * 4, len;
    for (r = 0; r < h; ++r) {
        uint32_t acc = 0;
        for (c = 0; c < w; ++c) {
            uint32_t cur = load_sample(px, pitch, x0 + c, y0 + r, s);
            v[r * w + c] = cur;
            if (cur != v[0]) constant = 0;
            if (c > 0) {
                uint32_t prev = v[r * w + c - 1];
                res[r * w + c] = type == ORC_UINT16 ? zigzag16(cur, prev) : (cur ^ prev);
                acc |= res[r * w + c];
            }
        }
        bits[r] = bit_length(acc);
        pred += 8 + 4 * (((uint64_t)(w - 1) * bits[r] + 31) / 32);
    }
    mode = ORC_RAW;
    len = raw4;
    if (constant && len > 4) {
        mode = ORC_CONSTANT;
        len = 4;
    }
    if (pred < len) {
        mode = ORC_PREDICT;
        len = pred;
    }
    memset(out, 0, len);
    if (mode == ORC_RAW) {
        for (c = 0; c < w * h; ++c) {
            if (s == 2)
                put16(out + 2 * (uint64_t)c, v[c]);
            else
                put32(out + 4 * (uint64_t)c, v[c]);
        }
    } else if (mode == ORC_CONSTANT) {
        put32(out, v[0]);
    } else {
        uint8_t *q = out;
        for (r = 0; r < h; ++r) {
            /* 64-bit reservoir: residuals enter at the top of the pending bits,
             * full 32-bit words leave from the bottom (LSB-first packing). */
            uint64_t acc = 0;
            uint32_t pending = 0, b = bits[r];
            put32(q, v[r * w]);
            q[4] = (uint8_t)b;
            q += 8;
            for (c = 1; c < w && b > 0; ++c) {
                acc |= (uint64_t)res[r * w + c] << pending;
                pending += b;
                if (pending >= 32) {
                    put32(q, (uint32_t)acc);
                    q += 4;
                    acc >>= 32;
                    pending -= 32;
                }
            }
            if (pending > 0) {
                put32(q, (uint32_t)acc);
                q += 4;
            }
        }
    }
    *mode_out = mode;
    return len;
}

/* Whole-image encode: header, 32-byte index entries, consecutive payloads
 * (docs/format.md "File header", "Tile index"; reference src/cpu/encoder.cpp:120-197). */
int orc_compress(uint32_t type, uint32_t width, uint32_t height, uint64_t pitch,
                 const void *pixels, uint32_t tile_width, uint32_t tile_height, uint8_t *out,
                 uint64_t capacity, uint64_t *written) {
    uint64_t bound, s, cols, rows, count, cursor, i, span;
    int st = orc_image_span(type, width, height, pitch, &span);
    if (st) return st;
    st = orc_compress_bound(type, width, height, tile_width, tile_height, &bound);
    if (st) return st;
    if (capacity < bound) return ORC_BUFFER_TOO_SMALL;
    s = type == ORC_UINT16 ? 2 : 4;
    cols = div_up(width, tile_width);
    rows = div_up(height, tile_height);
    count = cols * rows;
    cursor = 64 + 32 * count;
    for (i = 0; i < count; ++i) {
        const uint64_t x = (i % cols) * tile_width, y = (i / cols) * tile_height;
        const uint32_t w = (uint32_t)((width - x) < tile_width ? (width - x) : tile_width);
        const uint32_t h = (uint32_t)((height - y) < tile_height ? (height - y) : tile_height);
        uint8_t *e = out + 64 + 32 * i;
        uint32_t mode;
        uint64_t len = encode_tile((const uint8_t *)pixels, pitch, s, type, x, y, w, h,
                                   out + cursor, &mode);
        memset(e, 0, 32);
        e[0] = (uint8_t)mode;
        put16(e + 4, w);
        put16(e + 6, h);
        put64(e + 8, cursor);
        put64(e + 16, len);
        cursor += len;
    }
    memset(out, 0, 64);
    memcpy(out, orc_magic, 8);
    put16(out + 8, 1);
    put16(out + 10, 0);
    put16(out + 12, 64);
    out[14] = (uint8_t)type;
    put32(out + 16, width);
    put32(out + 20, height);
    put16(out + 24, tile_width);
    put16(out + 26, tile_height);
    put16(out + 28, 32);
    put64(out + 32, count);
    put64(out + 40, 64);
    put64(out + 48, 64 + 32 * count);
    put64(out + 56, cursor);
    *written = cursor;
    return ORC_SUCCESS;
}

/* ------------------------------------------------------------------------- */
/* Decoder                                                                   */
/* ------------------------------------------------------------------------- */

/* Residual j of a row: bits [j b, j b + b) of the LSB-first word stream; a second
 * word is read only when the symbol crosses into it (docs/format.md
 * "PREDICT_BITPACK"; reference src/cpu/decoder.cpp:11-27). */
static uint32_t read_residual(const uint8_t *words, uint64_t j, uint32_t b) {
    uint64_t pos, window;
    uint32_t sh;
    if (b == 0) return 0;
    pos = j * b;
    sh = (uint32_t)(pos % 32);
    window = get32(words + pos / 32 * 4);
    if (sh + b > 32) window |= (uint64_t)get32(words + pos / 32 * 4 + 4) << 32;
    return (uint32_t)((window >> sh) & ((b == 32) ? 0xffffffffull : ((1ull << b) - 1)));
}

/* Reconstruct tile rows [r0, r1) into `emit`: a callback-free design that writes
 * full tile rows into a 4096-word scratch, row-major w wide. */
static void decode_tile(const uint8_t *file, const orc_tile *t, const orc_segment *segs,
                        uint32_t type, uint32_t *tile_px) {
    const uint8_t *p = file + t->offset;
    const uint32_t w = t->width, h = t->height;
    uint32_t r, c;
    for (r = 0; r < h; ++r) {
        uint32_t *row = tile_px + r * w;
        if (t->mode == ORC_RAW) {
            for (c = 0; c < w; ++c)
                row[c] = type == ORC_UINT16 ? get16(p + 2 * ((uint64_t)r * w + c))
                                            : get32(p + 4 * ((uint64_t)r * w + c));
        } else if (t->mode == ORC_CONSTANT) {
            uint32_t k = type == ORC_UINT16 ? get16(p) : get32(p);
            for (c = 0; c < w; ++c) row[c] = k;
        } else {
            const orc_segment *g = &segs[t->first_segment + r];
            uint32_t prev = g->seed;
            row[0] = prev;
            for (c = 1; c < w; ++c) {
                uint32_t z = read_residual(p + g->payload_offset, c - 1, g->bit_width);
                if (type == ORC_UINT16) {
                    /* inverse zigzag: even z -> +z/2, odd z -> -(z/2) - 1, added mod 2^16 */
                    int32_t delta = (z & 1u) ? -(int32_t)(z >> 1) - 1 : (int32_t)(z >> 1);
                    prev = (uint32_t)((int32_t)prev + delta) & 0xffffu;
                } else {
                    prev ^= z;
                }
                row[c] = prev;
            }
        }
    }
}

typedef struct {
    orc_info info;
    orc_tile *tiles;
    orc_segment *segs;
} parsed_t;

static int parse_all(const uint8_t *file, uint64_t size, parsed_t *P) {
    int st = orc_parse(file, size, ORC_L, ORC_L, ORC_L, &P->info, NULL, NULL);
    if (st) return st;
    P->tiles = (orc_tile *)malloc(sizeof(orc_tile) * (P->info.tile_count ? P->info.tile_count : 1));
    P->segs = (orc_segment *)malloc(sizeof(orc_segment) *
                                    (P->info.segment_count ? P->info.segment_count : 1));
    if (!P->tiles || !P->segs) {
        free(P->tiles);
        free(P->segs);
        return ORC_RESOURCE_LIMIT;
    }
    return orc_parse(file, size, ORC_L, ORC_L, ORC_L, &P->info, P->tiles, P->segs);
}

typedef void (*sink_fn)(void *ctx, uint64_t ox, uint64_t oy, uint32_t value);

/* Decode every tile intersecting [x, x+w) x [y, y+h) and hand each in-region
 * pixel to `sink` with region-local coordinates (ROI = crop of the full decode,
 * reference docs/api.md "decode_roi_async"). */
static int decode_region(const uint8_t *file, uint64_t size, uint32_t x, uint32_t y, uint32_t w,
                         uint32_t h, int whole, sink_fn sink, void *ctx) {
    parsed_t P;
    uint32_t *tile_px;
    uint64_t cols, i;
    int st = parse_all(file, size, &P);
    if (st) return st;
    if (whole) {
        x = 0;
        y = 0;
        w = P.info.width;
        h = P.info.height;
    } else if (w == 0 || h == 0 || (uint64_t)x + w > P.info.width ||
               (uint64_t)y + h > P.info.height) {
        free(P.tiles);
        free(P.segs);
        return ORC_INVALID_ROI;
    }
    tile_px = (uint32_t *)malloc(4096 * sizeof(uint32_t));
    cols = div_up(P.info.width, P.info.tile_width);
    for (i = 0; i < P.info.tile_count; ++i) {
        const orc_tile *t = &P.tiles[i];
        const uint64_t tx = (i % cols) * P.info.tile_width, ty = (i / cols) * P.info.tile_height;
        uint64_t r, c;
        if (tx >= (uint64_t)x + w || tx + t->width <= x || ty >= (uint64_t)y + h ||
            ty + t->height <= y)
            continue;
        decode_tile(file, t, P.segs, P.info.type, tile_px);
        for (r = 0; r < t->height; ++r) {
            const uint64_t gy = ty + r;
            if (gy < y || gy >= (uint64_t)y + h) continue;
            for (c = 0; c < t->width; ++c) {
                const uint64_t gx = tx + c;
                if (gx < x || gx >= (uint64_t)x + w) continue;
                sink(ctx, gx - x, gy - y, tile_px[r * t->width + c]);
            }
        }
    }
    free(tile_px);
    free(P.tiles);
    free(P.segs);
    return ORC_SUCCESS;
}

typedef struct {
    uint8_t *out;
    uint64_t pitch, s;
} raw_sink_t;

static void raw_sink(void *ctx, uint64_t ox, uint64_t oy, uint32_t v) {
    raw_sink_t *k = (raw_sink_t *)ctx;
    uint8_t *p = k->out + oy * k->pitch + ox * k->s;
    if (k->s == 2)
        put16(p, v);
    else
        put32(p, v);
}

int orc_decompress(const uint8_t *file, uint64_t size, void *out, uint64_t pitch) {
    raw_sink_t k;
    orc_info info;
    int st = orc_parse(file, size, ORC_L, ORC_L, ORC_L, &info, NULL, NULL);
    if (st) return st;
    k.out = (uint8_t *)out;
    k.pitch = pitch;
    k.s = info.type == ORC_UINT16 ? 2 : 4;
    return decode_region(file, size, 0, 0, 0, 0, 1, raw_sink, &k);
}

int orc_decode_roi(const uint8_t *file, uint64_t size, uint32_t x, uint32_t y, uint32_t w,
                   uint32_t h, void *out, uint64_t pitch) {
    raw_sink_t k;
    orc_info info;
    int st = orc_parse(file, size, ORC_L, ORC_L, ORC_L, &info, NULL, NULL);
    if (st) return st;
    k.out = (uint8_t *)out;
    k.pitch = pitch;
    k.s = info.type == ORC_UINT16 ? 2 : 4;
    return decode_region(file, size, x, y, w, h, 0, raw_sink, &k);
}

/* ------------------------------------------------------------------------- */
/* Calibration                                                               */
/* ------------------------------------------------------------------------- */

/* Sample value: uint16 converts exactly, FP32 keeps its stored bits
 * (reference docs/numerical-analysis.md "Implemented semantics"). */
static float sample_float(uint32_t type, uint32_t bits) {
    float f;
    if (type == ORC_UINT16) return (float)bits;
    memcpy(&f, &bits, 4);
    return f;
}

/* r = RN(RN(x - o) * g): two separately rounded binary32 operations.  This file
 * is compiled with -ffp-contract=off so no FMA can merge them. */
static float calibrate_one(float x, float o, float g) {
    volatile float d = x - o;
    return d * g;
}

static float param(float scalar, const float *map, uint64_t pitch, uint64_t x, uint64_t y) {
    if (!map) return scalar;
    return *(const float *)((const uint8_t *)map + y * pitch + x * 4);
}

void orc_calibrate(uint32_t type, uint32_t width, uint32_t height, const void *in,
                   uint64_t in_pitch, float offset, const float *offset_map,
                   uint64_t offset_pitch, float gain, const float *gain_map, uint64_t gain_pitch,
                   float *out, uint64_t out_pitch) {
    const uint64_t s = type == ORC_UINT16 ? 2 : 4;
    uint64_t x, y;
    for (y = 0; y < height; ++y) {
        for (x = 0; x < width; ++x) {
            const uint8_t *p = (const uint8_t *)in + y * in_pitch + x * s;
            const uint32_t bits = s == 2 ? get16(p) : get32(p);
            const float r = calibrate_one(sample_float(type, bits),
                                          param(offset, offset_map, offset_pitch, x, y),
                                          param(gain, gain_map, gain_pitch, x, y));
            *(float *)((uint8_t *)out + y * out_pitch + x * 4) = r;
        }
    }
}

typedef struct {
    uint32_t type;
    float offset, gain;
    const float *omap, *gmap;
    uint64_t opitch, gpitch;
    float *out;
    uint64_t pitch;
} cal_sink_t;

static void cal_sink(void *ctx, uint64_t ox, uint64_t oy, uint32_t v) {
    cal_sink_t *k = (cal_sink_t *)ctx;
    *(float *)((uint8_t *)k->out + oy * k->pitch + ox * 4) =
        calibrate_one(sample_float(k->type, v), param(k->offset, k->omap, k->opitch, ox, oy),
                      param(k->gain, k->gmap, k->gpitch, ox, oy));
}

int orc_decode_roi_calibrate(const uint8_t *file, uint64_t size, uint32_t x, uint32_t y,
                             uint32_t w, uint32_t h, float offset, const float *offset_map,
                             uint64_t offset_pitch, float gain, const float *gain_map,
                             uint64_t gain_pitch, float *out, uint64_t out_pitch) {
    cal_sink_t k;
    orc_info info;
    int st = orc_parse(file, size, ORC_L, ORC_L, ORC_L, &info, NULL, NULL);
    if (st) return st;
    k.type = info.type;
    k.offset = offset;
    k.gain = gain;
    k.omap = offset_map;
    k.gmap = gain_map;
    k.opitch = offset_pitch;
    k.gpitch = gain_pitch;
    k.out = out;
    k.pitch = out_pitch;
    return decode_region(file, size, x, y, w, h, 0, cal_sink, &k);
}
