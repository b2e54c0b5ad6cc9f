/*
 * TileStream .tsc v1.0 CPU oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference CPU codec, written from the normative
 * format description (reference docs/format.md) and checked against the
 * reference's golden files and, where it builds, against the compiled
 * reference CPU codec (oracle/_ref).  It is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may
 * load it.  The product library never links, loads or calls it.
 *
 * Status values are the reference's `tilestream::Status` numbers
 * (reference include/tilestream/types.h:7-19).
 */
#ifndef TSC_ORACLE_H
#define TSC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    ORC_SUCCESS = 0,
    ORC_INVALID_ARGUMENT = 1,
    ORC_INVALID_ROI = 2,
    ORC_UNSUPPORTED_TYPE = 3,
    ORC_UNSUPPORTED_FORMAT = 4,
    ORC_MALFORMED_DATA = 5,
    ORC_SIZE_OVERFLOW = 6,
    ORC_BUFFER_TOO_SMALL = 7,
    ORC_RESOURCE_LIMIT = 8
};

enum { ORC_UINT16 = 1, ORC_FLOAT32 = 2 };
enum { ORC_RAW = 0, ORC_CONSTANT = 1, ORC_PREDICT = 2 };

typedef struct {
    uint32_t type, width, height;
    uint16_t tile_width, tile_height;
    uint64_t file_bytes;
    uint64_t tile_count;
    uint64_t segment_count;
} orc_info;

typedef struct {
    uint32_t mode;
    uint16_t width, height;
    uint64_t offset;        /* absolute file offset of the payload */
    uint64_t length;        /* padded payload bytes */
    uint64_t first_segment; /* index into the segment table, 0 unless predictive */
} orc_tile;

typedef struct {
    uint32_t seed;
    uint32_t bit_width;
    uint64_t payload_offset; /* packed residual words, relative to the tile payload */
    uint64_t payload_bytes;
} orc_segment;

int orc_compress_bound(uint32_t type, uint32_t width, uint32_t height, uint32_t tile_width,
                       uint32_t tile_height, uint64_t *bytes);
int orc_image_span(uint32_t type, uint32_t width, uint32_t height, uint64_t pitch,
                   uint64_t *bytes);
int orc_workspace_size(uint64_t image_count, uint64_t tile_count, uint32_t tile_width,
                       uint32_t tile_height, uint64_t scan_temp_bytes, uint64_t *bytes);

/* Validating parser.  With tiles == NULL only `info` is filled (counts).  With
 * tables of info->tile_count / info->segment_count entries it fills them. */
int orc_parse(const uint8_t *file, uint64_t size, uint64_t limit_file_bytes,
              uint64_t limit_tile_count, uint64_t limit_metadata_bytes, orc_info *info,
              orc_tile *tiles, orc_segment *segments);

/* Canonical encoder.  `pixels` has `pitch` bytes between rows. */
int orc_compress(uint32_t type, uint32_t width, uint32_t height, uint64_t pitch,
                 const void *pixels, uint32_t tile_width, uint32_t tile_height, uint8_t *out,
                 uint64_t capacity, uint64_t *written);

/* Whole-image decode into a pitched output of the archive's size and type. */
int orc_decompress(const uint8_t *file, uint64_t size, void *out, uint64_t pitch);

/* Region decode: every intersecting tile is reconstructed and cropped. */
int orc_decode_roi(const uint8_t *file, uint64_t size, uint32_t x, uint32_t y, uint32_t w,
                   uint32_t h, void *out, uint64_t pitch);

/* Host binary32 calibration r = RN(RN(x - offset) * gain), compiled with
 * -ffp-contract=off.  Samples: uint16 (type 1) or float32 bit words (type 2).
 * offset/gain maps may be NULL (scalar used), else pitched FP32 maps. */
void orc_calibrate(uint32_t type, uint32_t width, uint32_t height, const void *in,
                   uint64_t in_pitch, float offset, const float *offset_map,
                   uint64_t offset_pitch, float gain, const float *gain_map, uint64_t gain_pitch,
                   float *out, uint64_t out_pitch);

/* Decode + calibrate a region with scalar parameters, the CPU baseline of the
 * fused path. */
int orc_decode_roi_calibrate(const uint8_t *file, uint64_t size, uint32_t x, uint32_t y,
                             uint32_t w, uint32_t h, float offset, const float *offset_map,
                             uint64_t offset_pitch, float gain, const float *gain_map,
                             uint64_t gain_pitch, float *out, uint64_t out_pitch);

#ifdef __cplusplus
}
#endif

#endif
