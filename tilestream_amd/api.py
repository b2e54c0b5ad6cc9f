"""Python mirror of the TileStream API over the C-ABI.

Names, argument meaning and Status results follow the reference C++ API
(reference include/tilestream/{types,sizing,api,cpu,cuda}.h); calls return the
Status (and a value where the C++ API fills an out-argument).  Device memory is
managed by this module through the library's HIP helpers; numpy arrays are the
host side.
"""
from __future__ import annotations

import ctypes as C
import os
from enum import IntEnum
from typing import Optional, Sequence

import numpy as np

from . import _native as N
from ._native import (ArchiveInfo, BatchItem, BatchRequest, Calibration, CompressedInput, ConstBytes,
                      ConstImageView, DeviceResult, ImageSpec, ImageView, LaunchShape, MutableBytes,
                      Parameter, ParseLimits, QueueStats, Roi, SegmentMetadata, ServeRequest,
                      ServerOptions, ServerStats, TileMapping, TileMetadata, TileShape, WorkspaceSpec)

MAX_BYTES = (1 << 63) - 1


class Status(IntEnum):
    success = 0
    invalid_argument = 1
    invalid_roi = 2
    unsupported_type = 3
    unsupported_format = 4
    malformed_data = 5
    size_overflow = 6
    buffer_too_small = 7
    resource_limit = 8
    hip_error = 9
    io_error = 10


class PixelType(IntEnum):
    uint16 = 1
    float32 = 2


class TileMode(IntEnum):
    raw = 0
    constant = 1
    predict_bitpack = 2


class Operation(IntEnum):
    compress = 0
    decompress = 1
    decode_roi = 2
    calibrate = 3
    decode_calibrate = 4


class ParameterKind(IntEnum):
    scalar = 0
    map = 1


class HipError(RuntimeError):
    pass


def _check_hip(err: int, what: str) -> None:
    if err != 0:
        raise HipError(f"{what} failed with hipError {err}")


def pixel_type_of(dtype) -> PixelType:
    dt = np.dtype(dtype)
    if dt == np.uint16:
        return PixelType.uint16
    if dt in (np.dtype(np.uint32), np.dtype(np.float32)):
        return PixelType.float32
    raise TypeError(f"no pixel type for {dt}")


def sample_bytes(ptype) -> int:
    return 2 if int(ptype) == PixelType.uint16 else 4


# ---- sizing -------------------------------------------------------------------

def image_span(image: ImageSpec, pitch_bytes: int):
    out = C.c_uint64(0)
    st = N.lib().ts_image_span(C.byref(image), pitch_bytes, C.byref(out))
    return Status(st), out.value


def compress_bound(image: ImageSpec, tile: TileShape = None):
    out = C.c_uint64(0)
    st = N.lib().ts_compress_bound(C.byref(image), tile or TileShape(), C.byref(out))
    return Status(st), out.value


def workspace_size(spec: WorkspaceSpec):
    out = C.c_uint64(0)
    st = N.lib().ts_workspace_size(C.byref(spec), C.byref(out))
    return Status(st), out.value


# ---- host parser and codec ----------------------------------------------------

class Metadata:
    """Owned tilestream::ImageMetadata (from parse_metadata)."""

    def __init__(self, handle: int, keepalive=None):
        self.handle = handle
        self._keep = keepalive

    def __del__(self):
        if getattr(self, "handle", None):
            N.lib().ts_metadata_free(self.handle)
            self.handle = None

    @property
    def summary(self) -> N.MetadataSummary:
        s = N.MetadataSummary()
        N.lib().ts_metadata_get_summary(self.handle, C.byref(s))
        return s

    def tile(self, i: int) -> TileMetadata:
        t = TileMetadata()
        if N.lib().ts_metadata_get_tile(self.handle, i, C.byref(t)):
            raise IndexError(i)
        return t

    def set_tile(self, i: int, t: TileMetadata) -> None:
        if N.lib().ts_metadata_set_tile(self.handle, i, C.byref(t)):
            raise IndexError(i)

    def segment(self, i: int) -> SegmentMetadata:
        s = SegmentMetadata()
        if N.lib().ts_metadata_get_segment(self.handle, i, C.byref(s)):
            raise IndexError(i)
        return s

    def set_segment(self, i: int, s: SegmentMetadata) -> None:
        if N.lib().ts_metadata_set_segment(self.handle, i, C.byref(s)):
            raise IndexError(i)

    def truncate(self, tiles: int, segments: int) -> None:
        if N.lib().ts_metadata_truncate(self.handle, tiles, segments):
            raise IndexError((tiles, segments))

    def set_file_bytes(self, file_bytes: int) -> None:
        N.lib().ts_metadata_set_file_bytes(self.handle, file_bytes)


def _host_bytes(data) -> tuple:
    if isinstance(data, (bytes, bytearray)):
        arr = np.frombuffer(bytes(data), np.uint8)
    else:
        arr = np.ascontiguousarray(data).view(np.uint8).reshape(-1)
    ptr = arr.ctypes.data if arr.size else None
    return arr, ConstBytes(ptr, arr.size)


def parse_metadata(data, limits: Optional[ParseLimits] = None):
    arr, cb = _host_bytes(data)
    handle = C.c_void_p(0)
    lim = limits or ParseLimits(MAX_BYTES, MAX_BYTES, MAX_BYTES)
    st = Status(N.lib().ts_parse_metadata(cb, C.byref(lim), C.byref(handle)))
    return st, (Metadata(handle.value) if st == Status.success else None)


def host_view(arr: np.ndarray, const=True):
    """Image view of a 2-D host array (uint16, or float32/uint32 bit words)."""
    h, w = arr.shape
    spec = ImageSpec(int(pixel_type_of(arr.dtype)), w, h)
    cls = ConstImageView if const else ImageView
    pitch = arr.strides[0] if h > 1 else w * arr.itemsize
    return cls(arr.ctypes.data, spec, pitch, pitch * (h - 1) + w * arr.itemsize)


def cpu_compress(pixels: np.ndarray, tile: TileShape = None):
    px = np.ascontiguousarray(pixels)
    st, bound = compress_bound(host_view(px).image, tile)
    if st != Status.success:
        return st, b""
    out = np.zeros(bound, np.uint8)
    written = C.c_uint64(0)
    st = Status(N.lib().ts_cpu_compress(host_view(px), tile or TileShape(),
                                        MutableBytes(out.ctypes.data, bound), C.byref(written)))
    return st, out[: written.value].tobytes() if st == Status.success else b""


def cpu_decompress(data, out: np.ndarray, limits: Optional[ParseLimits] = None) -> Status:
    arr, cb = _host_bytes(data)
    lim = limits or ParseLimits(MAX_BYTES, MAX_BYTES, MAX_BYTES)
    return Status(N.lib().ts_cpu_decompress(cb, host_view(out, const=False), C.byref(lim)))


# ---- TIFF files ------------------------------------------------------------------

def tiff_write_options(**fields) -> "N.TiffWriteOptions":
    """The library's default TIFF write options with the given fields replaced
    (bigtiff, compression 0 none / 1 Deflate, horizontal_predictor, tile_width,
    tile_height, deflate_level)."""
    o = N.TiffWriteOptions()
    N.lib().ts_tiff_default_options(None, C.byref(o))
    for k, v in fields.items():
        setattr(o, k, int(v))
    return o


def write_tiff(path: str, pixels: np.ndarray, options: "Optional[N.TiffWriteOptions]" = None) -> Status:
    """Writes a uint16 or float32 (stored as its bit pattern) 2-D image (reference io::write_tiff)."""
    opts = options if options is not None else tiff_write_options()
    return Status(N.lib().ts_write_tiff(os.fsencode(path), host_view(pixels), C.byref(opts)))


def read_tiff(path: str, limits: "Optional[N.TiffReadLimits]" = None):
    """(status, pixels): the decoded image as uint16, or float32 samples as uint32
    bit patterns (reference io::read_tiff)."""
    handle = C.c_void_p()
    st = Status(N.lib().ts_read_tiff(os.fsencode(path), C.byref(limits) if limits is not None else None,
                                     C.byref(handle)))
    if st != Status.success:
        return st, None
    try:
        spec, pitch, data = ImageSpec(), C.c_uint64(), C.c_void_p()
        N.lib().ts_tiff_image_get(handle, C.byref(spec), C.byref(pitch), C.byref(data))
        dtype = np.uint16 if spec.type == int(PixelType.uint16) else np.uint32
        out = np.empty((spec.height, spec.width), dtype)
        if out.nbytes:
            C.memmove(out.ctypes.data, data.value, out.nbytes)
        return st, out
    finally:
        N.lib().ts_tiff_image_free(handle)


# ---- device memory, streams, events --------------------------------------------

class DeviceMemory:
    """A hipMalloc allocation, freed on close()/garbage collection."""

    def __init__(self, nbytes: int):
        self.size = int(nbytes)
        ptr = C.c_void_p(0)
        if self.size:
            _check_hip(N.lib().ts_hip_malloc(self.size, C.byref(ptr)), "hipMalloc")
        self.ptr = ptr.value or 0

    @classmethod
    def from_host(cls, arr) -> "DeviceMemory":
        a = np.ascontiguousarray(arr) if not isinstance(arr, (bytes, bytearray)) else np.frombuffer(arr, np.uint8)
        m = cls(a.nbytes)
        if a.nbytes:
            _check_hip(N.lib().ts_hip_memcpy(m.ptr, a.ctypes.data, a.nbytes), "hipMemcpy H2D")
        return m

    def upload(self, arr, offset: int = 0) -> None:
        a = np.ascontiguousarray(arr) if not isinstance(arr, (bytes, bytearray)) else np.frombuffer(arr, np.uint8)
        assert offset + a.nbytes <= self.size
        if a.nbytes:
            _check_hip(N.lib().ts_hip_memcpy(self.ptr + offset, a.ctypes.data, a.nbytes), "hipMemcpy H2D")

    def download(self, nbytes: Optional[int] = None, offset: int = 0) -> np.ndarray:
        n = self.size - offset if nbytes is None else nbytes
        out = np.empty(n, np.uint8)
        if n:
            _check_hip(N.lib().ts_hip_memcpy(out.ctypes.data, self.ptr + offset, n), "hipMemcpy D2H")
        return out

    def fill(self, byte: int) -> None:
        if self.size:
            _check_hip(N.lib().ts_hip_memset(self.ptr, byte, self.size), "hipMemset")

    def close(self) -> None:
        if self.ptr:
            N.lib().ts_hip_free(self.ptr)
            self.ptr = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HostMemory:
    """Pinned, mapped host memory (hipHostMalloc): device kernels read it over
    the host link at the same address; freed on close()/garbage collection."""

    def __init__(self, nbytes: int):
        self.size = int(nbytes)
        ptr = C.c_void_p(0)
        if self.size:
            _check_hip(N.lib().ts_hip_host_malloc(self.size, C.byref(ptr)), "hipHostMalloc")
        self.ptr = ptr.value or 0

    def array(self) -> np.ndarray:
        """A uint8 NumPy view of the memory (valid while this object lives)."""
        if not self.size:
            return np.zeros(0, np.uint8)
        return np.ctypeslib.as_array((C.c_uint8 * self.size).from_address(self.ptr))

    def upload(self, arr, offset: int = 0) -> None:
        a = np.ascontiguousarray(arr) if not isinstance(arr, (bytes, bytearray)) else np.frombuffer(arr, np.uint8)
        assert offset + a.nbytes <= self.size
        C.memmove(self.ptr + offset, a.ctypes.data, a.nbytes)

    def download(self, nbytes: Optional[int] = None, offset: int = 0) -> np.ndarray:
        n = self.size - offset if nbytes is None else nbytes
        return self.array()[offset:offset + n].copy()

    def close(self) -> None:
        if self.ptr:
            N.lib().ts_hip_host_free(self.ptr)
            self.ptr = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Stream:
    def __init__(self):
        s = C.c_void_p(0)
        _check_hip(N.lib().ts_hip_stream_create(C.byref(s)), "hipStreamCreate")
        self.handle = s.value

    def synchronize(self) -> None:
        _check_hip(N.lib().ts_hip_stream_synchronize(self.handle), "hipStreamSynchronize")

    def wait(self, event: "Event") -> None:
        """Later work on this stream waits for `event`."""
        _check_hip(N.lib().ts_hip_stream_wait_event(self.handle, event.handle), "hipStreamWaitEvent")

    def close(self) -> None:
        if self.handle:
            N.lib().ts_hip_stream_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Event:
    def __init__(self):
        e = C.c_void_p(0)
        _check_hip(N.lib().ts_hip_event_create(C.byref(e)), "hipEventCreate")
        self.handle = e.value

    def record(self, stream) -> None:
        _check_hip(N.lib().ts_hip_event_record(self.handle, _stream(stream)), "hipEventRecord")

    def synchronize(self) -> None:
        _check_hip(N.lib().ts_hip_event_synchronize(self.handle), "hipEventSynchronize")

    def elapsed_ms(self, stop: "Event") -> float:
        ms = C.c_float(0)
        _check_hip(N.lib().ts_hip_event_elapsed_ms(self.handle, stop.handle, C.byref(ms)), "hipEventElapsedTime")
        return ms.value

    def __del__(self):
        try:
            if self.handle:
                N.lib().ts_hip_event_destroy(self.handle)
        except Exception:
            pass


def memcpy_async(dst: int, src: int, nbytes: int, stream=None) -> None:
    """hipMemcpyAsync between any two addresses (pinned host memory makes H2D asynchronous)."""
    _check_hip(N.lib().ts_hip_memcpy_async(dst, src, nbytes, _stream(stream)), "hipMemcpyAsync")


def _stream(stream):
    if stream is None:
        return None
    return stream.handle if isinstance(stream, Stream) else stream


def decode_shape_supported(group: int, block: int, depth: int) -> bool:
    """Whether the library has a decode kernel for this (group, block, depth)."""
    return bool(N.lib().tsk_decode_supported(group, block, depth))


def device_count() -> int:
    n = C.c_int(0)
    err = N.lib().ts_hip_device_count(C.byref(n))
    return n.value if err == 0 else 0


def device_synchronize() -> None:
    _check_hip(N.lib().ts_hip_device_synchronize(), "hipDeviceSynchronize")


def device_view(mem: DeviceMemory, ptype, width: int, height: int, pitch: Optional[int] = None,
                offset: int = 0, capacity: Optional[int] = None, const=False):
    s = sample_bytes(ptype)
    pitch = pitch if pitch is not None else width * s
    cap = capacity if capacity is not None else mem.size - offset
    cls = ConstImageView if const else ImageView
    return cls(mem.ptr + offset, ImageSpec(int(ptype), width, height), pitch, cap)


# ---- GPU plans ------------------------------------------------------------------

class Plan:
    """Owning handle of a prepared BatchPlan."""

    def __init__(self, handle: int, keepalive):
        self.handle = handle
        self._keep = keepalive  # host arrays the request points to stay alive with the plan

    def workspace_spec(self) -> WorkspaceSpec:
        spec = WorkspaceSpec()
        N.lib().ts_plan_workspace_spec(self.handle, C.byref(spec))
        return spec

    def workspace_bytes(self) -> int:
        st, n = workspace_size(self.workspace_spec())
        assert st == Status.success
        return n

    def set_decode_shape(self, group=8, block=256, grid=0, depth=1) -> Status:
        return Status(N.lib().ts_plan_set_decode_shape(self.handle, group, block, grid, depth))

    def close(self) -> None:
        if self.handle:
            N.lib().ts_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def prepare_batch(operation: Operation, items: Sequence[BatchItem], tile: TileShape = None,
                  stream=None, previous: Optional[Plan] = None):
    """Returns (Status, Plan).  On failure `previous` is returned unchanged."""
    arr = (BatchItem * max(1, len(items)))(*items)
    req = BatchRequest(int(operation), C.cast(arr, C.POINTER(BatchItem)) if items else None,
                       len(items), tile or TileShape())
    handle = C.c_void_p(0)
    st = Status(N.lib().ts_prepare_batch(C.byref(req), C.byref(handle), _stream(stream)))
    if st != Status.success:
        return st, previous
    return st, Plan(handle.value, (arr, items))


def _run(fn, plan: Plan, workspace, results, count, stream) -> Status:
    ws = MutableBytes(workspace.ptr, workspace.size) if isinstance(workspace, DeviceMemory) else (
        workspace if workspace is not None else MutableBytes(None, 0))
    rptr = results.ptr if isinstance(results, DeviceMemory) else results
    return Status(fn(plan.handle if plan else None, ws, rptr, count, _stream(stream)))


def compress_async(plan, workspace, results, count, stream=None) -> Status:
    return _run(N.lib().ts_compress_async, plan, workspace, results, count, stream)


def decompress_async(plan, workspace, results, count, stream=None) -> Status:
    return _run(N.lib().ts_decompress_async, plan, workspace, results, count, stream)


def decode_roi_async(plan, workspace, results, count, stream=None) -> Status:
    return _run(N.lib().ts_decode_roi_async, plan, workspace, results, count, stream)


def calibrate_async(plan, workspace, results, count, stream=None) -> Status:
    return _run(N.lib().ts_calibrate_async, plan, workspace, results, count, stream)


def decode_calibrate_async(plan, workspace, results, count, stream=None) -> Status:
    return _run(N.lib().ts_decode_calibrate_async, plan, workspace, results, count, stream)


def synchronize(stream=None) -> Status:
    return Status(N.lib().ts_synchronize(_stream(stream)))


def read_results(mem: DeviceMemory, count: int):
    raw = mem.download(16 * count)
    rec = raw.view(np.uint32).reshape(count, 4)
    status = rec[:, 0].copy()
    written = raw.view(np.uint64).reshape(count, 2)[:, 1].copy()
    return status, written


def scalar(value: float) -> Parameter:
    return Parameter(int(ParameterKind.scalar), float(value), ConstImageView())


def map_parameter(view: ConstImageView) -> Parameter:
    return Parameter(int(ParameterKind.map), 0.0, view)


# ---- batched persistent decode (extension) ------------------------------------------

class DecodeQueue:
    """A prepared mixed batch drained by one persistent decode launch."""

    def __init__(self, handle: int, keepalive):
        self.handle = handle
        self._keep = keepalive
        self.count = len(keepalive[1])

    @classmethod
    def prepare(cls, items: Sequence[BatchItem], operations: Sequence[int], stream=None):
        arr = (BatchItem * max(1, len(items)))(*items)
        ops = (C.c_uint32 * max(1, len(items)))(*[int(o) for o in operations])
        handle = C.c_void_p(0)
        st = Status(N.lib().ts_queue_prepare(arr, ops, len(items), C.byref(handle), _stream(stream)))
        if st != Status.success:
            return st, None
        return st, cls(handle.value, (arr, items, ops))

    def workspace_bytes(self) -> int:
        return int(N.lib().ts_queue_workspace_bytes(self.handle))

    def stats(self) -> QueueStats:
        s = QueueStats()
        N.lib().ts_queue_get_stats(self.handle, C.byref(s))
        return s

    @property
    def shape(self) -> LaunchShape:
        s = LaunchShape()
        N.lib().ts_queue_get_shape(self.handle, C.byref(s))
        return s

    def set_shape(self, group=8, block=256, grid=0, chunk=8, depth=2, run=1) -> Status:
        return Status(N.lib().ts_queue_set_shape(self.handle,
                                                 C.byref(LaunchShape(group, block, grid, chunk, depth, run))))

    def _ws(self, workspace):
        return MutableBytes(workspace.ptr, workspace.size) if workspace is not None else MutableBytes(None, 0)

    def run(self, workspace, results, stream=None) -> Status:
        return Status(N.lib().ts_queue_run(self.handle, self._ws(workspace), results.ptr, self.count,
                                           _stream(stream)))

    def capture(self, workspace, results, stream) -> Status:
        return Status(N.lib().ts_queue_capture(self.handle, self._ws(workspace), results.ptr, self.count,
                                               _stream(stream)))

    def launch(self, stream) -> Status:
        return Status(N.lib().ts_queue_launch(self.handle, _stream(stream)))

    def autotune(self, workspace, results, stream, candidates, repeats=5):
        arr = (LaunchShape * len(candidates))(*[LaunchShape(*c) for c in candidates])
        ms = (C.c_float * len(candidates))()
        st = Status(N.lib().ts_queue_autotune(self.handle, self._ws(workspace), results.ptr, self.count,
                                              _stream(stream), arr, len(candidates), repeats, ms))
        return st, list(ms)

    def close(self) -> None:
        if self.handle:
            N.lib().ts_queue_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- serving resident archives ----------------------------------------------------


class ArchiveSet:
    """Archives resident in device or pinned host memory with their tile tables (tilestream/serve.h)."""

    def __init__(self):
        handle = C.c_void_p(0)
        st = Status(N.lib().ts_archive_set_create(C.byref(handle)))
        if st != Status.success:
            raise RuntimeError(f"ts_archive_set_create: {st.name}")
        self.handle = handle.value
        self._keep = []

    def add(self, device_file, metadata: Metadata, size: Optional[int] = None, offset: int = 0):
        """Adds an archive already in device memory (DeviceMemory) or in pinned
        host memory (HostMemory; its tiles are staged per batch), at `offset`
        in the allocation; returns (status, id)."""
        ident = C.c_uint32(0)
        n = device_file.size - offset if size is None else size
        st = Status(N.lib().ts_archive_set_add(self.handle, ConstBytes(device_file.ptr + offset, n),
                                               metadata.handle, C.byref(ident)))
        if st == Status.success:
            self._keep.append((device_file, metadata))
        return st, ident.value

    def add_device(self, device_file, image: ImageSpec, tile, producer=None, size: Optional[int] = None,
                   offset: int = 0, stream=None):
        """Adds an archive in device memory (DeviceMemory, at `offset`) without
        host metadata: validated and tabled on the device.  `producer`: the
        encode's device result (a DeviceMemory or an address), or None when the
        file is exactly `size` bytes.  Returns (status, id)."""
        ident = C.c_uint32(0)
        n = device_file.size - offset if size is None else size
        prod = None if producer is None else (producer if isinstance(producer, int) else producer.ptr)
        t = tile if isinstance(tile, TileShape) else TileShape(*tile)
        st = Status(N.lib().ts_archive_set_add_device(self.handle, ConstBytes(device_file.ptr + offset, n), image, t,
                                                      prod, _stream(stream), C.byref(ident)))
        if st == Status.success:
            self._keep.append((device_file, producer))
        return st, ident.value

    def commit(self, stream=None) -> Status:
        return Status(N.lib().ts_archive_set_commit(self.handle, _stream(stream)))

    def __len__(self) -> int:
        return int(N.lib().ts_archive_set_size(self.handle))

    def info(self, ident: int):
        out = ArchiveInfo()
        st = Status(N.lib().ts_archive_set_info(self.handle, ident, C.byref(out)))
        return st, out

    def close(self) -> None:
        if self.handle:
            N.lib().ts_archive_set_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def server_options(streams=None, max_batch=None, max_tasks=None, admission_bytes=None, shape=None,
                   encode_window=None, staging_bytes=None, large_bytes=None, block_tasks=None) -> ServerOptions:
    """The library's default server options with the given fields replaced."""
    o = ServerOptions()
    N.lib().ts_server_default_options(C.byref(o))
    for name, value in (("streams", streams), ("max_batch", max_batch), ("max_tasks", max_tasks),
                        ("admission_bytes", admission_bytes), ("encode_window", encode_window),
                        ("staging_bytes", staging_bytes), ("large_bytes", large_bytes),
                        ("block_tasks", block_tasks)):
        if value is not None:
            setattr(o, name, value)
    if shape is not None:
        o.shape = LaunchShape(*shape)
    return o


def region_request(archive: int, roi, output: ImageView, calibrated=False, offset=0.0, gain=1.0) -> ServeRequest:
    """Decode of `roi` of a resident archive into `output` (FP32 (x - offset) * gain when calibrated)."""
    return ServeRequest(kind=0, archive=archive, calibrated=1 if calibrated else 0, roi=Roi(*roi), output=output,
                        offset=offset, gain=gain)


def encode_request(input: ConstImageView, tile, compressed: MutableBytes) -> ServeRequest:
    """Canonical encode of the device image `input` with tile shape `tile` into `compressed`."""
    return ServeRequest(kind=1, tile=TileShape(*tile), input=input, compressed=compressed, gain=1.0)


class Server:
    """Continuous batching of region decodes and encodes over resident archives."""

    def __init__(self, archives: ArchiveSet, options: Optional[ServerOptions] = None):
        handle = C.c_void_p(0)
        opts = options or server_options()
        st = Status(N.lib().ts_server_create(archives.handle, C.byref(opts), C.byref(handle)))
        if st != Status.success:
            raise RuntimeError(f"ts_server_create: {st.name}")
        self.handle = handle.value
        self._archives = archives
        self.options = opts

    def check(self, request: ServeRequest) -> Status:
        return Status(N.lib().ts_server_check(self.handle, C.byref(request)))

    @staticmethod
    def _results(n: int):
        res = np.zeros(max(1, n), dtype=[("status", "<u4"), ("reserved", "<u4"), ("bytes_written", "<u8")])
        return res, res.ctypes.data_as(C.POINTER(DeviceResult))

    def replay(self, requests: Sequence[ServeRequest], arrival_ns):
        """Open-loop replay; returns (status, latency_ns, statuses, bytes_written, stats)."""
        n = len(requests)
        arr = (ServeRequest * max(1, n))(*requests)
        arrivals = np.ascontiguousarray(arrival_ns, dtype=np.uint64)
        assert arrivals.size == n
        latency = np.zeros(max(1, n), np.uint64)
        res, res_ptr = self._results(n)
        stats = ServerStats()
        st = Status(N.lib().ts_server_replay(self.handle, arr, arrivals.ctypes.data_as(C.POINTER(C.c_uint64)), n,
                                             latency.ctypes.data_as(C.POINTER(C.c_uint64)), res_ptr,
                                             C.byref(stats)))
        return st, latency[:n], res["status"][:n].copy(), res["bytes_written"][:n].copy(), stats

    def run_batch(self, requests: Sequence[ServeRequest]):
        """All requests at once; returns (status, statuses, bytes_written)."""
        n = len(requests)
        arr = (ServeRequest * max(1, n))(*requests)
        res, res_ptr = self._results(n)
        st = Status(N.lib().ts_server_run_batch(self.handle, arr, n, res_ptr))
        return st, res["status"][:n].copy(), res["bytes_written"][:n].copy()

    def close(self) -> None:
        if self.handle:
            N.lib().ts_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
