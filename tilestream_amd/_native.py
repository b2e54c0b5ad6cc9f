"""ctypes binding of the TileStream C-ABI (include/tilestream_c.h).

Loads tilestream_amd/_native/libtilestream_amd.so built in-tree by
`python -m tilestream_amd.build`.  There is no fallback: if the library is
missing, every call raises.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_native", "libtilestream_amd.so")
GEN_PATH = os.path.join(HERE, "_native", "libtsgen.so")


class NativeLibraryMissing(RuntimeError):
    pass


# ---- record types (layout-identical to tilestream_c.h) ------------------------

class ImageSpec(C.Structure):
    _fields_ = [("type", C.c_uint32), ("width", C.c_uint32), ("height", C.c_uint32)]


class TileShape(C.Structure):
    _fields_ = [("width", C.c_uint16), ("height", C.c_uint16)]

    def __init__(self, width=128, height=32):
        super().__init__(width, height)


class Roi(C.Structure):
    _fields_ = [("x", C.c_uint32), ("y", C.c_uint32), ("width", C.c_uint32), ("height", C.c_uint32)]


class ConstBytes(C.Structure):
    _fields_ = [("data", C.c_void_p), ("size", C.c_uint64)]


class MutableBytes(C.Structure):
    _fields_ = [("data", C.c_void_p), ("size", C.c_uint64)]


class ConstImageView(C.Structure):
    _fields_ = [("data", C.c_void_p), ("image", ImageSpec), ("pitch_bytes", C.c_uint64),
                ("capacity_bytes", C.c_uint64)]


class ImageView(C.Structure):
    _fields_ = [("data", C.c_void_p), ("image", ImageSpec), ("pitch_bytes", C.c_uint64),
                ("capacity_bytes", C.c_uint64)]


class DeviceResult(C.Structure):
    _fields_ = [("status", C.c_uint32), ("reserved", C.c_uint32), ("bytes_written", C.c_uint64)]


class WorkspaceSpec(C.Structure):
    _fields_ = [("image_count", C.c_uint64), ("tile_count", C.c_uint64), ("tile", TileShape),
                ("scan_temp_bytes", C.c_uint64)]


class ParseLimits(C.Structure):
    _fields_ = [("file_bytes", C.c_uint64), ("tile_count", C.c_uint64), ("metadata_bytes", C.c_uint64)]


class MetadataSummary(C.Structure):
    _fields_ = [("image", ImageSpec), ("tile", TileShape), ("file_bytes", C.c_uint64),
                ("tile_count", C.c_uint64), ("segment_count", C.c_uint64)]


class TileMetadata(C.Structure):
    _fields_ = [("mode", C.c_uint32), ("width", C.c_uint16), ("height", C.c_uint16),
                ("offset", C.c_uint64), ("length", C.c_uint64), ("first_segment", C.c_uint64)]


class SegmentMetadata(C.Structure):
    _fields_ = [("seed", C.c_uint32), ("bit_width", C.c_uint32), ("payload_offset", C.c_uint64),
                ("payload_bytes", C.c_uint64)]


class Parameter(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("scalar", C.c_float), ("map", ConstImageView)]


class Calibration(C.Structure):
    _fields_ = [("offset", Parameter), ("gain", Parameter)]

    def __init__(self, *args, **kw):
        super().__init__(*args, **kw)
        if not args and "gain" not in kw:
            self.gain.scalar = 1.0


class TileMapping(C.Structure):
    _fields_ = [("tile_index", C.c_uint64), ("input_offset", C.c_uint64)]


class CompressedInput(C.Structure):
    _fields_ = [("bytes", ConstBytes), ("metadata", C.c_void_p), ("mappings", C.POINTER(TileMapping)),
                ("mapping_count", C.c_uint64), ("encoder_result", C.c_void_p),
                ("trusted_image", ImageSpec), ("trusted_tile", TileShape)]


class BatchItem(C.Structure):
    _fields_ = [("input", ConstImageView), ("output", ImageView), ("compressed_output", MutableBytes),
                ("compressed_input", CompressedInput), ("roi", Roi), ("calibration", Calibration)]

    def __init__(self, *args, **kw):
        super().__init__(*args, **kw)
        if "calibration" not in kw:
            self.calibration.gain.scalar = 1.0


class BatchRequest(C.Structure):
    _fields_ = [("operation", C.c_uint32), ("items", C.POINTER(BatchItem)), ("count", C.c_uint64),
                ("tile", TileShape)]


class LaunchShape(C.Structure):
    _fields_ = [("group", C.c_uint32), ("block", C.c_uint32), ("grid", C.c_uint32), ("chunk", C.c_uint32),
                ("depth", C.c_uint32), ("run", C.c_uint32)]


class QueueStats(C.Structure):
    _fields_ = [("images", C.c_uint64), ("tasks", C.c_uint64), ("decoded_bytes", C.c_uint64),
                ("output_bytes", C.c_uint64), ("payload_bytes", C.c_uint64), ("map_bytes", C.c_uint64)]


class ArchiveInfo(C.Structure):
    _fields_ = [("image", ImageSpec), ("tile", TileShape), ("tile_count", C.c_uint64), ("file_bytes", C.c_uint64),
                ("host", C.c_uint32), ("reserved", C.c_uint32)]


class ServeRequest(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("archive", C.c_uint32), ("calibrated", C.c_uint32), ("tile", TileShape),
                ("roi", Roi), ("output", ImageView), ("input", ConstImageView), ("compressed", MutableBytes),
                ("offset", C.c_float), ("gain", C.c_float)]


class ServerOptions(C.Structure):
    _fields_ = [("streams", C.c_uint32), ("max_batch", C.c_uint32), ("max_tasks", C.c_uint32),
                ("encode_window", C.c_uint32), ("admission_bytes", C.c_uint64), ("staging_bytes", C.c_uint64),
                ("shape", LaunchShape), ("large_bytes", C.c_uint64), ("block_tasks", C.c_uint64)]


class TiffReadLimits(C.Structure):
    _fields_ = [("max_width", C.c_uint32), ("max_height", C.c_uint32), ("max_decoded_bytes", C.c_uint64)]


class TiffWriteOptions(C.Structure):
    _fields_ = [("bigtiff", C.c_uint32), ("compression", C.c_uint32), ("horizontal_predictor", C.c_uint32),
                ("tile_width", C.c_uint32), ("tile_height", C.c_uint32), ("deflate_level", C.c_uint32)]


class ServerStats(C.Structure):
    _fields_ = [("requests", C.c_uint64), ("batches", C.c_uint64), ("decoded_bytes", C.c_uint64),
                ("failed", C.c_uint64), ("seconds", C.c_double), ("admission_holds", C.c_uint64),
                ("peak_in_flight", C.c_uint64)]


RECORDS = {
    "ts_image_spec": ImageSpec, "ts_tile_shape": TileShape, "ts_roi": Roi,
    "ts_const_bytes": ConstBytes, "ts_mutable_bytes": MutableBytes,
    "ts_const_image_view": ConstImageView, "ts_image_view": ImageView,
    "ts_device_result": DeviceResult, "ts_workspace_spec": WorkspaceSpec,
    "ts_parse_limits": ParseLimits, "ts_metadata_summary": MetadataSummary,
    "ts_tile_metadata": TileMetadata, "ts_segment_metadata": SegmentMetadata,
    "ts_parameter": Parameter, "ts_calibration": Calibration, "ts_tile_mapping": TileMapping,
    "ts_compressed_input": CompressedInput, "ts_batch_item": BatchItem,
    "ts_batch_request": BatchRequest, "ts_launch_shape": LaunchShape, "ts_queue_stats": QueueStats,
    "ts_archive_info": ArchiveInfo, "ts_serve_request": ServeRequest, "ts_server_options": ServerOptions,
    "ts_server_stats": ServerStats, "ts_tiff_read_limits": TiffReadLimits,
    "ts_tiff_write_options": TiffWriteOptions,
}

_u32, _u64, _vp, _int = C.c_uint32, C.c_uint64, C.c_void_p, C.c_int
_P = C.POINTER

# name -> (restype, argtypes)
PROTOTYPES = {
    "ts_image_span": (_u32, [_P(ImageSpec), _u64, _P(_u64)]),
    "ts_compress_bound": (_u32, [_P(ImageSpec), TileShape, _P(_u64)]),
    "ts_workspace_size": (_u32, [_P(WorkspaceSpec), _P(_u64)]),
    "ts_parse_metadata": (_u32, [ConstBytes, _P(ParseLimits), _P(_vp)]),
    "ts_metadata_free": (None, [_vp]),
    "ts_metadata_get_summary": (None, [_vp, _P(MetadataSummary)]),
    "ts_metadata_get_tile": (_u32, [_vp, _u64, _P(TileMetadata)]),
    "ts_metadata_set_tile": (_u32, [_vp, _u64, _P(TileMetadata)]),
    "ts_metadata_get_segment": (_u32, [_vp, _u64, _P(SegmentMetadata)]),
    "ts_metadata_set_segment": (_u32, [_vp, _u64, _P(SegmentMetadata)]),
    "ts_metadata_truncate": (_u32, [_vp, _u64, _u64]),
    "ts_metadata_set_file_bytes": (None, [_vp, _u64]),
    "ts_cpu_compress": (_u32, [ConstImageView, TileShape, MutableBytes, _P(_u64)]),
    "ts_cpu_decompress": (_u32, [ConstBytes, ImageView, _P(ParseLimits)]),
    "ts_tiff_default_options": (None, [_P(TiffReadLimits), _P(TiffWriteOptions)]),
    "ts_read_tiff": (_u32, [C.c_char_p, _P(TiffReadLimits), _P(_vp)]),
    "ts_tiff_image_get": (None, [_vp, _P(ImageSpec), _P(_u64), _P(_vp)]),
    "ts_tiff_image_free": (None, [_vp]),
    "ts_write_tiff": (_u32, [C.c_char_p, ConstImageView, _P(TiffWriteOptions)]),
    "ts_prepare_batch": (_u32, [_P(BatchRequest), _P(_vp), _vp]),
    "ts_plan_destroy": (None, [_vp]),
    "ts_plan_workspace_spec": (None, [_vp, _P(WorkspaceSpec)]),
    "ts_compress_async": (_u32, [_vp, MutableBytes, _vp, _u64, _vp]),
    "ts_compress_async_static": (_u32, [_vp, _vp, _u64, _vp]),
    "ts_decompress_async": (_u32, [_vp, MutableBytes, _vp, _u64, _vp]),
    "ts_decode_roi_async": (_u32, [_vp, MutableBytes, _vp, _u64, _vp]),
    "ts_calibrate_async": (_u32, [_vp, MutableBytes, _vp, _u64, _vp]),
    "ts_decode_calibrate_async": (_u32, [_vp, MutableBytes, _vp, _u64, _vp]),
    "ts_buffer_allocate": (_u32, [_u64, _vp, _P(_vp)]),
    "ts_buffer_release": (_u32, [_vp, _vp]),
    "ts_synchronize": (_u32, [_vp]),
    "ts_plan_set_decode_shape": (_u32, [_vp, _u32, _u32, _u32, _u32]),
    "ts_queue_prepare": (_u32, [_P(BatchItem), _P(_u32), _u64, _P(_vp), _vp]),
    "ts_queue_destroy": (None, [_vp]),
    "ts_queue_workspace_bytes": (_u64, [_vp]),
    "ts_queue_get_stats": (None, [_vp, _P(QueueStats)]),
    "ts_queue_get_shape": (None, [_vp, _P(LaunchShape)]),
    "ts_queue_set_shape": (_u32, [_vp, _P(LaunchShape)]),
    "ts_queue_run": (_u32, [_vp, MutableBytes, _vp, _u64, _vp]),
    "ts_queue_capture": (_u32, [_vp, MutableBytes, _vp, _u64, _vp]),
    "ts_queue_launch": (_u32, [_vp, _vp]),
    "ts_queue_autotune": (_u32, [_vp, MutableBytes, _vp, _u64, _vp, _P(LaunchShape), _u64, _u32,
                                 _P(C.c_float)]),
    "ts_archive_set_create": (_u32, [_P(_vp)]),
    "ts_archive_set_destroy": (None, [_vp]),
    "ts_archive_set_add": (_u32, [_vp, ConstBytes, _vp, _P(_u32)]),
    "ts_archive_set_add_device": (_u32, [_vp, ConstBytes, ImageSpec, TileShape, _vp, _vp, _P(_u32)]),
    "ts_archive_set_commit": (_u32, [_vp, _vp]),
    "ts_archive_set_size": (_u32, [_vp]),
    "ts_archive_set_info": (_u32, [_vp, _u32, _P(ArchiveInfo)]),
    "ts_server_default_options": (None, [_P(ServerOptions)]),
    "ts_server_create": (_u32, [_vp, _P(ServerOptions), _P(_vp)]),
    "ts_server_destroy": (None, [_vp]),
    "ts_server_check": (_u32, [_vp, _P(ServeRequest)]),
    "ts_server_replay": (_u32, [_vp, _P(ServeRequest), _P(_u64), _u64, _P(_u64), _P(DeviceResult),
                                _P(ServerStats)]),
    "ts_server_run_batch": (_u32, [_vp, _P(ServeRequest), _u64, _P(DeviceResult)]),
    "ts_hip_malloc": (_int, [_u64, _P(_vp)]),
    "ts_hip_free": (_int, [_vp]),
    "ts_hip_host_malloc": (_int, [_u64, _P(_vp)]),
    "ts_hip_host_free": (_int, [_vp]),
    "ts_hip_memcpy": (_int, [_vp, _vp, _u64]),
    "ts_hip_memcpy_async": (_int, [_vp, _vp, _u64, _vp]),
    "ts_hip_memset": (_int, [_vp, _int, _u64]),
    "ts_hip_stream_create": (_int, [_P(_vp)]),
    "ts_hip_stream_destroy": (_int, [_vp]),
    "ts_hip_stream_synchronize": (_int, [_vp]),
    "ts_hip_device_synchronize": (_int, []),
    "ts_hip_get_last_error": (_int, []),
    "ts_hip_peek_last_error": (_int, []),
    "ts_hip_device_count": (_int, [_P(_int)]),
    "ts_hip_event_create": (_int, [_P(_vp)]),
    "ts_hip_event_destroy": (_int, [_vp]),
    "ts_hip_event_record": (_int, [_vp, _vp]),
    "ts_hip_event_synchronize": (_int, [_vp]),
    "ts_hip_stream_wait_event": (_int, [_vp, _vp]),
    "ts_hip_event_elapsed_ms": (_int, [_vp, _vp, _P(C.c_float)]),
    "ts_abi_sizeof": (_u64, [C.c_char_p]),
    "tsk_copy_blocks": (_int, [_vp, _vp, _u64, _vp]),
    "tsk_stream_copy": (_int, [_vp, _vp, _u64, _vp]),
    "tsk_stream_read": (_int, [_vp, _u64, _vp, _vp]),
    "tsk_stream_write": (_int, [_vp, _u64, _u32, _vp]),
    "tsk_fill_u32": (_int, [_vp, _u64, _u32, _vp]),
    "tsk_decode_supported": (_int, [_u32, _u32, _u32]),
    "tsk_decode_occupancy": (_int, [_u32, _u32, _u32, _u32, _P(_int)]),
    "tsk_preload": (_int, [_u32]),
}

_lib = None
_gen = None


def lib():
    """The loaded product library; raises if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NativeLibraryMissing(
                f"{LIB_PATH} is missing: build it with `python -m tilestream_amd.build`")
        handle = C.CDLL(LIB_PATH)
        for name, (res, args) in PROTOTYPES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def gen():
    global _gen
    if _gen is None:
        if not os.path.exists(GEN_PATH):
            raise NativeLibraryMissing(f"{GEN_PATH} is missing: run `python -m tilestream_amd.build`")
        g = C.CDLL(GEN_PATH)
        g.tsg_detector_u16.argtypes = [_u64, _u32, _u32, _vp, _u64, _int]
        g.tsg_field_f32.argtypes = [_u64, _u32, _u32, _vp, _u64, _int]
        g.tsg_dark_u16.argtypes = [_u64, _u32, _u32, _vp, _int]
        g.tsg_gain_f32.argtypes = [_u64, _u32, _u32, _vp, _int]
        g.tsg_bits.restype = _u64
        g.tsg_bits.argtypes = [_u64, _u64, _u64]
        for f in (g.tsg_detector_u16, g.tsg_field_f32, g.tsg_dark_u16, g.tsg_gain_f32):
            f.restype = None
        _gen = g
    return _gen
