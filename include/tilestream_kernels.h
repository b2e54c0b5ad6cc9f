/*
 * TileStream kernel layer: the C-ABI between the host planner (C++) and the
 * hand-written HIP kernels for gfx950.  Plain POD descriptor tables, device
 * pointers, sizes and a hipStream_t passed as `void *`; every launcher returns
 * a hipError_t value (0 = hipSuccess) and never allocates, synchronizes or
 * loads code.
 *
 * These launchers replace the reference's kernel launches:
 *   tsk_results_begin / tsk_results_complete  <- reference src/cuda/decoder.cu:307-326
 *   tsk_decode                                 <- reference src/cuda/decoder.cu:176-305, 328-339
 *   tsk_encode_window                          <- reference src/cuda/encoder.cu:62-138 (analyze),
 *                                                 284-288 (scan), 175-238 (pack), 240-280 (finish):
 *                                                 one single-pass kernel per window
 *   tsk_encode_reset                           <- (no counterpart: clears the window state)
 *   tsk_calibrate / tsk_calibrate_report       <- reference src/cuda/calibrate.cu:13-62
 *   tsk_preload                                <- reference src/cuda/decoder.cu:341-349,
 *                                                 encoder.cu:329-349, calibrate.cu:64-71
 */
#ifndef TILESTREAM_KERNELS_H
#define TILESTREAM_KERNELS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TSK_NO_ROWS 0xffffffffu
#define TSK_NOT_STAGED 0xffffffffu

/* Pixel kinds of a decode image / task. */
enum { TSK_UINT16 = 1, TSK_FLOAT32 = 2 };

/* Decode kernel variants: fixed pixel type and calibration, or mixed. */
enum {
    TSK_KIND_U16 = 0,
    TSK_KIND_F32 = 1,
    TSK_KIND_U16_CAL = 2,
    TSK_KIND_F32_CAL = 3,
    TSK_KIND_MIXED = 4
};

/* Per-output descriptor of a decode (or fused decode + calibrate) job. */
typedef struct {
    const uint8_t *src;          /* compressed bytes (full file, or staged payloads) */
    uint8_t *out;                /* output pixel at the region's top-left corner */
    const void *producer;        /* trusted path: the encoder's DeviceResult, else NULL */
    const uint8_t *offset_map;   /* FP32 maps in region coordinates, NULL = scalar */
    const uint8_t *gain_map;
    uint64_t src_size;           /* readable bytes at src */
    uint64_t out_pitch;
    uint64_t offset_pitch;
    uint64_t gain_pitch;
    uint32_t width, height;      /* archive image */
    uint32_t roi_x, roi_y, roi_w, roi_h;
    uint32_t tile_columns;       /* ceil(width / tile_w) */
    uint16_t tile_w, tile_h;
    uint8_t type;                /* TSK_UINT16 / TSK_FLOAT32 */
    uint8_t calibrated;          /* output FP32 (x - offset) * gain */
    uint16_t reserved;
    float offset, gain;          /* scalar parameters */
} tsk_decode_image;

/* One tile to decode.  On the trusted path the payload address, mode and
 * length are read from the device-resident index instead (payload = 0). */
typedef struct {
    const uint8_t *payload;      /* absolute device address of the tile payload */
    uint32_t column, row;        /* tile position in the archive's tile grid */
    uint32_t image;              /* index into the image table */
    uint32_t rows;               /* first entry of the row table, TSK_NO_ROWS if none */
    uint32_t words;              /* payload length / 4 */
    uint32_t shape;              /* mode | actual height << 8 | actual width << 16 */
} tsk_decode_task;

#define TSK_TASK_SHAPE(mode, height, width) ((uint32_t)(mode) | ((uint32_t)(height) << 8) | ((uint32_t)(width) << 16))

/* Request mode: a batch of region requests over resident archives.  Request r
 * covers tasks [first_task, next request's first_task): the tiles of columns
 * col0.. col0+cols-1 and rows row0.., row-major, of the archive whose tiles
 * start at tile_base in the resident tile table (tsk_decode_args.tasks).
 * tsk_decode_args.request_index holds the request of every task.  For the
 * block decoder (tsk_build_blocks) request r also covers blocks
 * [first_block, next request's first_block): its tile grid cut into blocks
 * of across x down tiles, row-major. */
typedef struct {
    uint64_t first_task;
    uint32_t tile_base;          /* index of the archive's tile 0 in the tile table */
    uint32_t tile_columns;       /* tile columns of the archive */
    uint32_t col0, row0;         /* first tile of the region */
    uint32_t cols;               /* tiles per region row */
    uint32_t rows;               /* tile rows of the region */
    uint32_t first_block;        /* block decoding: the request's first block */
    uint16_t across, down;       /* block decoding: tiles of a full block */
} tsk_request;

/* Block decode: a block is up to 8 tiles of one image, `across` x `down`
 * neighbouring tiles of its region whose output is at most 64 rows of at
 * most 1 KiB: a workgroup stages all their payloads in LDS at once and every
 * wave writes whole 1 KiB runs of one output row. */
typedef struct {
    const uint8_t *payload;      /* absolute device address of the tile payload */
    uint32_t words;              /* payload length / 4 */
    uint32_t shape;              /* mode | actual height << 8 | actual width << 16 */
    uint32_t rows;               /* first entry of the row table, TSK_NO_ROWS if none */
    uint32_t stage;              /* first 16-byte LDS block of the payload's 16-byte blocks */
    uint32_t reserved[2];
} tsk_block_tile;

/* 416 bytes (26 16-byte blocks, read by one wave straight into LDS). */
typedef struct {
    uint32_t image;
    uint16_t across, down;       /* tiles of this block */
    uint32_t col0, row0;         /* archive tile of its top-left tile */
    uint32_t count;              /* across x down */
    uint32_t blocks;             /* 16-byte LDS blocks staged for all its tiles */
    uint32_t flags;              /* TSK_BLOCK_FULL */
    uint32_t reserved;
    tsk_block_tile tiles[8];     /* row-major; entries past count are zero */
    tsk_decode_image image_copy; /* the image record, so that a block needs one read */
    uint64_t pad;
} tsk_block;

/* 16-byte LDS blocks of a block's stage: eight RAW 64x64 uint16 tiles, each
 * 16-byte misaligned, fill 4104; planning rejects more than 4135, so the
 * farthest read past a row's end (258 words) and the rounding of the last
 * wave's loads to 64 blocks stay inside the stage. */
#define TSK_BLOCK_STAGE 4200u
#define TSK_BLOCK_STAGE_USED 4135u

/* Every tile of the block is predictive and full size, and the block has
 * its full number of tiles both ways. */
#define TSK_BLOCK_FULL 1u

/* Block records of a batch of region requests, built on the device from the
 * requests and the resident tile table (one record per block, in request
 * order): the block decoder's input for request batches. */
typedef struct {
    const tsk_request *requests;
    const uint32_t *block_index;     /* block -> request */
    const tsk_decode_image *images;  /* one per request */
    const tsk_decode_task *tiles;    /* the resident tile table */
    tsk_block *blocks;               /* out: block_count records */
    uint64_t block_count;
} tsk_block_build_args;

typedef struct {
    const tsk_decode_image *images;
    const tsk_decode_task *tasks;
    const uint16_t *rows;        /* word offset of each predictive row in its payload */
    void *results;               /* DeviceResult[image_count] */
    uint64_t *queue;             /* work counter for the dynamic queue, NULL = static stride */
    uint64_t task_count;
    uint64_t row_count;          /* entries in the row table */
    uint32_t image_count;
    uint32_t kind;               /* TSK_KIND_* */
    uint32_t group;              /* columns per lane: 1, 2, 4, 8 or 16; 0: block decode */
    uint32_t block;              /* threads per workgroup: 64, 128, 256 or 512 */
    uint32_t grid;               /* workgroups; tiles are strided or drawn from the queue */
    uint32_t chunk;              /* tiles per queue draw (the last eighth of the tasks; the rest by stride) */
    uint32_t depth;              /* tiles a workgroup holds at once: 1 staged; 2, 3: 1 or 2 more in flight */
    uint32_t request_count;      /* request mode when non-zero (requests, request_index set) */
    const tsk_request *requests;
    const uint32_t *request_index; /* task -> request */
    /* Request mode, archives in host memory: per task, the tile's place in
     * `staging` in 16-byte blocks (TSK_NOT_STAGED: read the payload where it
     * is); NULL when no task is staged.  A staged payload keeps its offset
     * within its 16-byte block. */
    const uint32_t *staged;
    const uint8_t *staging;
    /* Static stride in runs: each workgroup takes `run` consecutive tasks
     * (neighbouring tiles, whose outputs form one compact block) before it
     * strides on; 0 or 1 = single tasks. */
    uint32_t run;
    uint32_t reserved;
    /* Block decode (group 0): task_count counts blocks. */
    const tsk_block *blocks;
} tsk_decode_args;

/* Per-image descriptor of an encode job. */
typedef struct {
    const uint8_t *in;
    uint8_t *out;                /* .tsc file, capacity >= compress_bound */
    uint64_t in_pitch;
    uint64_t tile_count;
    uint32_t width, height;
    uint32_t tile_columns;
    uint16_t tile_w, tile_h;
    uint32_t type;
} tsk_encode_image;

/* The part of one image that falls into one encode window. */
typedef struct {
    uint32_t image;
    uint32_t first;              /* first window tile of this image */
    uint32_t count;
    uint32_t flags;              /* 1: continues an earlier window, 2: finishes the image */
    uint64_t start;              /* image tile number of the first window tile */
} tsk_window_entry;

typedef struct {
    const tsk_encode_image *images;
    const tsk_window_entry *entries;
    /* Per block of 64 window tiles: the entry holding the block's first tile
     * (count ceil(tile_count / 64)); NULL: search the entries. */
    const uint32_t *blocks;
    void *results;               /* DeviceResult per image */
    uint64_t *state;             /* workspace: 2 x capacity look-back words, one buffer per window parity */
    uint64_t *carry;             /* workspace: 2 slots, output cursor of a continued image */
    uint32_t capacity;           /* tiles per look-back buffer (the largest window) */
    uint32_t entry_count;
    uint32_t tile_count;         /* tiles in this window */
    uint32_t parity;             /* window number & 1 */
    uint32_t type;               /* TSK_UINT16 / TSK_FLOAT32 */
    /* Consecutive window tiles per XCD: workgroups b with one b % 8 run on one
     * XCD, so runs of neighbouring tiles share that XCD's L2 (rows of
     * neighbouring tiles share cache lines).  0 or 1: workgroup b, tile b. */
    uint32_t run;
} tsk_encode_args;

/* Per-image descriptor of standalone calibration. */
typedef struct {
    const uint8_t *in;
    uint8_t *out;
    const uint8_t *offset_map;
    const uint8_t *gain_map;
    uint64_t in_pitch, out_pitch, offset_pitch, gain_pitch;
    uint32_t width, height;
    float offset, gain;
} tsk_calibrate_image;

typedef struct {
    const tsk_calibrate_image *images;
    const uint64_t *first_rows;  /* first row of each image in the concatenated rows */
    void *results;
    uint64_t total_rows;
    uint32_t image_count;
    uint32_t type;
} tsk_calibrate_args;

/* Device-built resident tile table of one archive in device memory. */
typedef struct {
    uint32_t header_status;      /* 0, a parser status, or TSK_INDEX_OTHER_IMAGE */
    uint32_t tile_failures;      /* tiles whose entry or payload failed a check */
    uint64_t file_bytes;         /* the file's size when the header is valid */
} tsk_index_record;

/* The header is valid but describes another image than the caller's. */
#define TSK_INDEX_OTHER_IMAGE 1u

typedef struct {
    const uint8_t *src;          /* the .tsc file, 4-byte aligned */
    const void *producer;        /* the encoder's DeviceResult (status, file size), or NULL */
    tsk_decode_task *tasks;      /* out: tile_count request-mode task entries (image 0) */
    uint16_t *rows;              /* out: tile_h row-descriptor word offsets per tile, at rows + i x tile_h */
    tsk_index_record *record;    /* out */
    uint64_t capacity;           /* readable bytes at src; the file size when producer is NULL */
    uint64_t bound;              /* compress bound of the expected image and tile shape */
    uint64_t tile_count;         /* tiles of the expected image */
    uint32_t width, height;      /* expected image */
    uint32_t tile_columns;
    uint32_t rows_base;          /* row-table index of rows[0] in the resident table */
    uint16_t tile_w, tile_h;
    uint32_t type;               /* TSK_UINT16 / TSK_FLOAT32 */
} tsk_index_args;

/* Validates a .tsc file in device memory as the host parser does and writes
 * its request-mode tile table.  tsk_index_header: one thread checks the
 * header and writes the record (tile_failures zeroed; tasks and rows unused).
 * tsk_index_tiles: one thread per tile checks its entry and payload and
 * writes its table entry, when the record's header status is 0. */
int tsk_index_header(const tsk_index_args *args, void *stream);
int tsk_index_tiles(const tsk_index_args *args, void *stream);

/* Zeroes `count` result records and, when `counter` is not NULL, `words`
 * (at most 256) work-counter words, in one kernel. */
int tsk_results_begin(void *results, uint32_t count, uint32_t *counter, uint32_t words, void *stream);
int tsk_results_complete(const tsk_decode_image *images, uint32_t count, void *results,
                         void *stream);
/* Request mode: copies the 16-byte blocks holding each staged task's payload
 * from where the resident tile table points (mapped host memory) into
 * `staging` (args->staged, args->staging set). */
int tsk_stage_payloads(const tsk_decode_args *args, void *stream);
/* tsk_results_complete that writes the final records to `out` (for example
 * mapped host memory) and leaves `results` as the decode left them. */
int tsk_results_publish(const tsk_decode_image *images, uint32_t count, const void *results, void *out,
                        void *stream);
int tsk_decode(const tsk_decode_args *args, void *stream);
/* Builds the block records of a batch of region requests (one wave per
 * block).  Every request's archive is device-resident, and a block of its
 * geometry stages at most TSK_BLOCK_STAGE_USED 16-byte blocks even with every
 * tile RAW (the caller checks it). */
int tsk_build_blocks(const tsk_block_build_args *args, void *stream);
/* Resident workgroups per CU of a decode kernel variant (for persistent grids). */
int tsk_decode_occupancy(uint32_t kind, uint32_t group, uint32_t block, uint32_t depth, int *blocks);
/* Whether a (group, block, depth) combination has a compiled kernel. */
int tsk_decode_supported(uint32_t group, uint32_t block, uint32_t depth);
/* Clears both look-back buffers; once before the first window of an execution
 * (each window clears the buffer the next one uses). */
int tsk_encode_reset(const tsk_encode_args *args, void *stream);
/* Encodes one window: index entries, payloads, and for each image that ends
 * in it the header and result.  Windows of one execution run in order. */
int tsk_encode_window(const tsk_encode_args *args, void *stream);
int tsk_calibrate(const tsk_calibrate_args *args, void *stream);
/* Size of the encoder's static device scratch (look-back buffers and carried
 * cursors of the allocation-free encode entry point): 1 MiB per GPU. */
#define TSK_SCRATCH_BYTES (1u << 20)
/* Device address and size of that scratch on the current device, looked up
 * on the first call for each device and kept until the process ends. */
int tsk_scratch_region(void **address, uint64_t *bytes);
/* Loads every kernel of an operation so that no execution loads code.
 * op: 0 compress, 1 decode, 2 calibrate, 3 fused decode. */
int tsk_preload(uint32_t op);
/* Copies `blocks` 16-byte blocks with a plain streaming kernel (the known-size
 * reference for scaling hardware byte counters). */
int tsk_copy_blocks(void *dst, const void *src, uint64_t blocks, void *stream);
/* Copies `blocks` 16-byte blocks in the fastest plain streaming shape measured
 * on this hardware: the practical ceiling for kernels that read and write. */
int tsk_stream_copy(void *dst, const void *src, uint64_t blocks, void *stream);
/* Reads `blocks` 16-byte blocks (sink: one word, written only if the XOR of
 * all data is all-ones) and writes `blocks` 16-byte blocks of `value`, each in
 * its fastest streaming shape measured here: the one-sided HBM ceilings. */
int tsk_stream_read(const void *src, uint64_t blocks, void *sink, void *stream);
int tsk_stream_write(void *dst, uint64_t blocks, uint32_t value, void *stream);
/* Zero-fills `words` 32-bit words (a kernel, safe under graph capture). */
int tsk_fill_u32(uint32_t *data, uint64_t words, uint32_t value, void *stream);

#ifdef __cplusplus
}
#endif

#endif
