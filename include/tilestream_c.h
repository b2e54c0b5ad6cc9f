/*
 * TileStream C-ABI: the drop-in boundary of this library.
 *
 * Flat `extern "C"` entry points over the C++ API (tilestream/hip.h and
 * friends) with plain pointers and sizes; any FFI (ctypes, cgo, JNI, N-API) can
 * bind them.  Every function returns a tilestream Status value (0 = success,
 * numbering of reference include/tilestream/types.h:7-19, 9 = HIP runtime
 * failure) unless stated otherwise.  Record types have exactly the layout of
 * their C++ counterparts.  Streams are `void *` (hipStream_t).
 *
 * Reference interface each entry replaces:
 *   ts_image_span / ts_compress_bound / ts_workspace_size
 *                          <- reference include/tilestream/sizing.h:26-30
 *   ts_parse_metadata      <- reference include/tilestream/api.h:42-43
 *   ts_cpu_compress / ts_cpu_decompress <- reference include/tilestream/cpu.h:9-12
 *   ts_read_tiff / ts_write_tiff <- reference include/tilestream/tiff.h:37-40
 *   ts_prepare_batch       <- reference include/tilestream/cuda.h:176-177
 *   ts_plan_workspace_spec <- reference include/tilestream/cuda.h:156
 *   ts_compress_async, ts_decompress_async, ts_decode_roi_async,
 *   ts_calibrate_async, ts_decode_calibrate_async
 *                          <- reference include/tilestream/cuda.h:178-187
 *   ts_buffer_allocate / ts_buffer_release <- reference include/tilestream/cuda.h:113-140
 *   ts_synchronize         <- reference include/tilestream/cuda.h:145
 * Extensions without a reference counterpart (batched persistent decode,
 * autotuning, serving) are declared in the second half.
 */
#ifndef TILESTREAM_C_H
#define TILESTREAM_C_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { uint32_t type, width, height; } ts_image_spec;          /* PixelType 1/2 */
typedef struct { uint16_t width, height; } ts_tile_shape;
typedef struct { uint32_t x, y, width, height; } ts_roi;
typedef struct { const void *data; uint64_t size; } ts_const_bytes;
typedef struct { void *data; uint64_t size; } ts_mutable_bytes;
typedef struct { const void *data; ts_image_spec image; uint64_t pitch_bytes, capacity_bytes; } ts_const_image_view;
typedef struct { void *data; ts_image_spec image; uint64_t pitch_bytes, capacity_bytes; } ts_image_view;
typedef struct { uint32_t status, reserved; uint64_t bytes_written; } ts_device_result;
typedef struct { uint64_t image_count, tile_count; ts_tile_shape tile; uint64_t scan_temp_bytes; } ts_workspace_spec;
typedef struct { uint64_t file_bytes, tile_count, metadata_bytes; } ts_parse_limits;

/* Host metadata of a parsed archive (tilestream::ImageMetadata), opaque. */
typedef struct ts_metadata ts_metadata;
typedef struct {
    ts_image_spec image;
    ts_tile_shape tile;
    uint64_t file_bytes, tile_count, segment_count;
} ts_metadata_summary;
typedef struct { uint32_t mode; uint16_t width, height; uint64_t offset, length, first_segment; } ts_tile_metadata;
typedef struct { uint32_t seed, bit_width; uint64_t payload_offset, payload_bytes; } ts_segment_metadata;

typedef struct { uint32_t kind; float scalar; ts_const_image_view map; } ts_parameter; /* kind 0 scalar, 1 map */
typedef struct { ts_parameter offset, gain; } ts_calibration;
typedef struct { uint64_t tile_index, input_offset; } ts_tile_mapping;
typedef struct {
    ts_const_bytes bytes;
    const ts_metadata *metadata;
    const ts_tile_mapping *mappings;
    uint64_t mapping_count;
    const ts_device_result *encoder_result;
    ts_image_spec trusted_image;
    ts_tile_shape trusted_tile;
} ts_compressed_input;
typedef struct {
    ts_const_image_view input;
    ts_image_view output;
    ts_mutable_bytes compressed_output;
    ts_compressed_input compressed_input;
    ts_roi roi;
    ts_calibration calibration;
} ts_batch_item;
/* operation: 0 compress, 1 decompress, 2 decode_roi, 3 calibrate, 4 decode_calibrate */
typedef struct { uint32_t operation; const ts_batch_item *items; uint64_t count; ts_tile_shape tile; } ts_batch_request;

typedef struct ts_plan ts_plan;

/* ---- sizing, parser, host codec ---------------------------------------- */
uint32_t ts_image_span(const ts_image_spec *image, uint64_t pitch_bytes, uint64_t *bytes);
uint32_t ts_compress_bound(const ts_image_spec *image, ts_tile_shape tile, uint64_t *bytes);
uint32_t ts_workspace_size(const ts_workspace_spec *spec, uint64_t *bytes);
/* On success *out owns a new metadata object (free with ts_metadata_free). */
uint32_t ts_parse_metadata(ts_const_bytes source, const ts_parse_limits *limits, ts_metadata **out);
void ts_metadata_free(ts_metadata *metadata);
void ts_metadata_get_summary(const ts_metadata *metadata, ts_metadata_summary *out);
/* Copies / overwrites tile and segment records (tests fabricate metadata). */
uint32_t ts_metadata_get_tile(const ts_metadata *metadata, uint64_t index, ts_tile_metadata *out);
uint32_t ts_metadata_set_tile(ts_metadata *metadata, uint64_t index, const ts_tile_metadata *in);
uint32_t ts_metadata_get_segment(const ts_metadata *metadata, uint64_t index, ts_segment_metadata *out);
uint32_t ts_metadata_set_segment(ts_metadata *metadata, uint64_t index, const ts_segment_metadata *in);
/* Keeps the first `tiles` tile and `segments` segment records (no more than
 * there are), as resizing the C++ ImageMetadata vectors would. */
uint32_t ts_metadata_truncate(ts_metadata *metadata, uint64_t tiles, uint64_t segments);
void ts_metadata_set_file_bytes(ts_metadata *metadata, uint64_t file_bytes);
uint32_t ts_cpu_compress(ts_const_image_view input, ts_tile_shape tile, ts_mutable_bytes output,
                         uint64_t *bytes_written);
uint32_t ts_cpu_decompress(ts_const_bytes input, ts_image_view output, const ts_parse_limits *limits);

/* ---- TIFF files (one uint16 or float32 sample per pixel) ----------------- */
typedef struct ts_tiff_image ts_tiff_image;
typedef struct {
    uint32_t max_width, max_height;
    uint64_t max_decoded_bytes;
} ts_tiff_read_limits;
typedef struct {
    uint32_t bigtiff;              /* 0: classic TIFF, 1: BigTIFF */
    uint32_t compression;          /* 0: none, 1: Deflate */
    uint32_t horizontal_predictor; /* uint16 with Deflate only */
    uint32_t tile_width, tile_height;
    uint32_t deflate_level;
} ts_tiff_write_options;
void ts_tiff_default_options(ts_tiff_read_limits *limits, ts_tiff_write_options *options);
/* On success *out owns the decoded image (free with ts_tiff_image_free); limits may be NULL. */
uint32_t ts_read_tiff(const char *path, const ts_tiff_read_limits *limits, ts_tiff_image **out);
/* The image, its row pitch (packed rows) and its pixels (host memory owned by `image`). */
void ts_tiff_image_get(const ts_tiff_image *image, ts_image_spec *spec, uint64_t *pitch_bytes,
                       const void **pixels);
void ts_tiff_image_free(ts_tiff_image *image);
uint32_t ts_write_tiff(const char *path, ts_const_image_view image, const ts_tiff_write_options *options);

/* ---- GPU plans and execution ------------------------------------------- */
/* On success *plan is replaced (the previous plan, if any, is destroyed). */
uint32_t ts_prepare_batch(const ts_batch_request *request, ts_plan **plan, void *stream);
void ts_plan_destroy(ts_plan *plan);
void ts_plan_workspace_spec(const ts_plan *plan, ts_workspace_spec *out);
uint32_t ts_compress_async(const ts_plan *plan, ts_mutable_bytes workspace, ts_device_result *results,
                           uint64_t result_count, void *stream);
/* ts_compress_async without a workspace.  No counterpart in the reference:
 * for callers that cannot allocate device memory (graph capture, serving
 * loops).  The encoder state lives in a static 1 MiB device array per GPU
 * (held until the process ends); a plan whose state does not fit returns
 * resource_limit (8).  Calls that overlap on one GPU share that array and
 * race.  Allocates and frees nothing. */
uint32_t ts_compress_async_static(const ts_plan *plan, ts_device_result *results, uint64_t result_count,
                                  void *stream);
uint32_t ts_decompress_async(const ts_plan *plan, ts_mutable_bytes workspace, ts_device_result *results,
                             uint64_t result_count, void *stream);
uint32_t ts_decode_roi_async(const ts_plan *plan, ts_mutable_bytes workspace, ts_device_result *results,
                             uint64_t result_count, void *stream);
uint32_t ts_calibrate_async(const ts_plan *plan, ts_mutable_bytes workspace, ts_device_result *results,
                            uint64_t result_count, void *stream);
uint32_t ts_decode_calibrate_async(const ts_plan *plan, ts_mutable_bytes workspace, ts_device_result *results,
                                   uint64_t result_count, void *stream);
/* DeviceBuffer::allocate / release: stream-ordered, 256-byte aligned. */
uint32_t ts_buffer_allocate(uint64_t bytes, void *stream, void **data);
uint32_t ts_buffer_release(void *data, void *stream);
uint32_t ts_synchronize(void *stream);

/* ---- extension: launch shape of the decode kernel ---------------------- */
/* group: columns per lane (1,2,4,8,16); block: 64..512; grid: 0 = one workgroup per tile;
 * depth: 1 staged; 2 or 3: one or two more tiles in flight.  invalid_argument for a combination
 * without a kernel.  A new plan's shape: group 4, block 256, grid 1024, depth 3 for fused
 * calibration plans of 1536 tiles or more; group 8, block 256, grid 1024, depth 3 for other
 * decode plans of 2048 or more; group 8, block 256, grid 0, depth 1 otherwise. */
uint32_t ts_plan_set_decode_shape(ts_plan *plan, uint32_t group, uint32_t block, uint32_t grid, uint32_t depth);

/* ---- extension: batched persistent decode (tilestream/batch.h) ---------
 * No counterpart in the reference.  Items may mix pixel types, tile shapes,
 * regions and fused calibration; operations[i] is 1 decompress, 2 decode_roi
 * or 4 decode_calibrate. */
typedef struct ts_queue ts_queue;
typedef struct { uint32_t group, block, grid, chunk, depth, run; } ts_launch_shape;
typedef struct {
    uint64_t images, tasks, decoded_bytes, output_bytes, payload_bytes, map_bytes;
} ts_queue_stats;
uint32_t ts_queue_prepare(const ts_batch_item *items, const uint32_t *operations, uint64_t count,
                          ts_queue **queue, void *stream);
void ts_queue_destroy(ts_queue *queue);
uint64_t ts_queue_workspace_bytes(const ts_queue *queue);
void ts_queue_get_stats(const ts_queue *queue, ts_queue_stats *out);
void ts_queue_get_shape(const ts_queue *queue, ts_launch_shape *out);
uint32_t ts_queue_set_shape(ts_queue *queue, const ts_launch_shape *shape);
uint32_t ts_queue_run(ts_queue *queue, ts_mutable_bytes workspace, ts_device_result *results,
                      uint64_t result_count, void *stream);
uint32_t ts_queue_capture(ts_queue *queue, ts_mutable_bytes workspace, ts_device_result *results,
                          uint64_t result_count, void *stream);
uint32_t ts_queue_launch(ts_queue *queue, void *stream);
/* Times each candidate (median of `repeats`), keeps the fastest, writes times in ms. */
uint32_t ts_queue_autotune(ts_queue *queue, ts_mutable_bytes workspace, ts_device_result *results,
                           uint64_t result_count, void *stream, const ts_launch_shape *candidates,
                           uint64_t candidate_count, uint32_t repeats, float *milliseconds);

/* ---- extension: serving resident archives (tilestream/serve.h) ---------
 * No counterpart in the reference.  Archives stay in device memory with tile
 * tables planned once; region decodes and encodes are batched continuously
 * over several streams (hip::ServeRequest).  Archives may also live in mapped
 * pinned host memory: their tiles are staged into device memory per batch. */
typedef struct ts_archive_set ts_archive_set;
typedef struct ts_server ts_server;
typedef struct {
    ts_image_spec image;
    ts_tile_shape tile;
    uint64_t tile_count, file_bytes;
    uint32_t host, reserved;       /* host: 1 when the archive is in pinned host memory (tiles staged per batch) */
} ts_archive_info;
typedef struct {
    uint32_t kind;                 /* 0 region decode, 1 encode */
    uint32_t archive, calibrated;  /* region; calibrated: FP32 (x - offset) * gain output */
    ts_tile_shape tile;            /* encode */
    ts_roi roi;                    /* region */
    ts_image_view output;          /* region: device pixels */
    ts_const_image_view input;     /* encode: device pixels */
    ts_mutable_bytes compressed;   /* encode: device, at least compress_bound bytes */
    float offset, gain;
} ts_serve_request;
typedef struct {
    uint32_t streams, max_batch, max_tasks, encode_window;
    uint64_t admission_bytes;
    uint64_t staging_bytes;        /* per stream, device staging for tiles of host archives */
    ts_launch_shape shape;
    uint64_t large_bytes;          /* requests with this many output bytes batch apart, on streams - 1 at most */
    uint64_t block_tasks;          /* region batches of this many tiles (device archives) decode by blocks; 0: never */
} ts_server_options;
typedef struct {
    uint64_t requests, batches, decoded_bytes, failed;
    double seconds;
    uint64_t admission_holds; /* batches closed early because admission_bytes was reached */
    uint64_t peak_in_flight;  /* most output bytes admitted and not yet finished */
} ts_server_stats;
uint32_t ts_archive_set_create(ts_archive_set **set);
void ts_archive_set_destroy(ts_archive_set *set);
/* `metadata` is the parse of the same bytes; the bytes stay resident and unchanged. */
uint32_t ts_archive_set_add(ts_archive_set *set, ts_const_bytes device_file, const ts_metadata *metadata,
                            uint32_t *id);
/* An archive in device memory without host metadata: validated and tabled on
 * the device (tilestream::hip::ArchiveSet::add_device; no reference
 * counterpart).  `producer` (may be NULL) is the encode's device result. */
uint32_t ts_archive_set_add_device(ts_archive_set *set, ts_const_bytes device_file, ts_image_spec image,
                                   ts_tile_shape tile, const ts_device_result *producer, void *stream, uint32_t *id);
uint32_t ts_archive_set_commit(ts_archive_set *set, void *stream);
uint32_t ts_archive_set_size(const ts_archive_set *set);
uint32_t ts_archive_set_info(const ts_archive_set *set, uint32_t id, ts_archive_info *out);
void ts_server_default_options(ts_server_options *out);
/* The archive set must outlive the server. */
uint32_t ts_server_create(const ts_archive_set *set, const ts_server_options *options, ts_server **server);
void ts_server_destroy(ts_server *server);
uint32_t ts_server_check(const ts_server *server, const ts_serve_request *request);
/* Open-loop replay; arrival_ns non-decreasing; latency_ns and results (host
 * memory: status, and bytes_written, the file size of an encode) may be NULL. */
uint32_t ts_server_replay(ts_server *server, const ts_serve_request *requests, const uint64_t *arrival_ns,
                          uint64_t count, uint64_t *latency_ns, ts_device_result *results, ts_server_stats *stats);
uint32_t ts_server_run_batch(ts_server *server, const ts_serve_request *requests, uint64_t count,
                             ts_device_result *results);

/* ---- HIP runtime helpers for hosts without their own HIP binding --------
 * These return hipError_t values (0 = hipSuccess). */
int ts_hip_malloc(uint64_t bytes, void **data);
int ts_hip_free(void *data);
int ts_hip_host_malloc(uint64_t bytes, void **data); /* pinned, mapped */
int ts_hip_host_free(void *data);
int ts_hip_memcpy(void *dst, const void *src, uint64_t bytes);  /* any direction; complete on return */
int ts_hip_memcpy_async(void *dst, const void *src, uint64_t bytes, void *stream);
int ts_hip_memset(void *dst, int value, uint64_t bytes);          /* complete on return */
int ts_hip_stream_create(void **stream);
int ts_hip_stream_destroy(void *stream);
int ts_hip_stream_synchronize(void *stream);
int ts_hip_device_synchronize(void);
int ts_hip_get_last_error(void);
int ts_hip_peek_last_error(void);
int ts_hip_device_count(int *count);
int ts_hip_event_create(void **event);
int ts_hip_event_destroy(void *event);
int ts_hip_event_record(void *event, void *stream);
int ts_hip_event_synchronize(void *event);
int ts_hip_stream_wait_event(void *stream, void *event);
int ts_hip_event_elapsed_ms(void *start, void *stop, float *ms);
/* Sizes of the record types, for binding checks: name -> sizeof, 0 if unknown. */
uint64_t ts_abi_sizeof(const char *name);

#ifdef __cplusplus
}
#endif

#endif
