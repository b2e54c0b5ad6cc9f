"""ctypes front end of the C oracle (tsc_oracle.c) -- TEST INFRASTRUCTURE ONLY.

Numpy in, numpy/bytes out.  Status values are the reference's
`tilestream::Status` numbers (reference include/tilestream/types.h:7-19).
`reference()` returns the compiled reference CPU codec (oracle/_ref) when it
was built, else None.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "_build", "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libtsref.so")

UINT16, FLOAT32 = 1, 2
RAW, CONSTANT, PREDICT = 0, 1, 2
L = (1 << 63) - 1


class Info(C.Structure):
    _fields_ = [("type", C.c_uint32), ("width", C.c_uint32), ("height", C.c_uint32),
                ("tile_width", C.c_uint16), ("tile_height", C.c_uint16),
                ("file_bytes", C.c_uint64), ("tile_count", C.c_uint64),
                ("segment_count", C.c_uint64)]


class Tile(C.Structure):
    _fields_ = [("mode", C.c_uint32), ("width", C.c_uint16), ("height", C.c_uint16),
                ("offset", C.c_uint64), ("length", C.c_uint64), ("first_segment", C.c_uint64)]


class Segment(C.Structure):
    _fields_ = [("seed", C.c_uint32), ("bit_width", C.c_uint32),
                ("payload_offset", C.c_uint64), ("payload_bytes", C.c_uint64)]


def build() -> None:
    """Compile the oracle (and the reference CPU codec when its sources exist)."""
    targets = ["oracle"]
    if os.path.isdir("/root/reference/src/cpu"):
        targets.append("ref")
    subprocess.run(["make", "-s", "-C", HERE] + targets, check=True)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build()
        _lib = C.CDLL(ORACLE_SO)
        u64p = C.POINTER(C.c_uint64)
        _lib.orc_compress_bound.argtypes = [C.c_uint32] * 5 + [u64p]
        _lib.orc_image_span.argtypes = [C.c_uint32] * 3 + [C.c_uint64, u64p]
        _lib.orc_workspace_size.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32,
                                            C.c_uint64, u64p]
        _lib.orc_parse.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64,
                                   C.POINTER(Info), C.c_void_p, C.c_void_p]
        _lib.orc_compress.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64, C.c_void_p,
                                      C.c_uint32, C.c_uint32, C.c_void_p, C.c_uint64, u64p]
        _lib.orc_decompress.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64]
        _lib.orc_decode_roi.argtypes = [C.c_void_p, C.c_uint64] + [C.c_uint32] * 4 + [
            C.c_void_p, C.c_uint64]
        _lib.orc_calibrate.restype = None
        _lib.orc_calibrate.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p,
                                       C.c_uint64, C.c_float, C.c_void_p, C.c_uint64,
                                       C.c_float, C.c_void_p, C.c_uint64, C.c_void_p,
                                       C.c_uint64]
        _lib.orc_decode_roi_calibrate.argtypes = [C.c_void_p, C.c_uint64] + [C.c_uint32] * 4 + [
            C.c_float, C.c_void_p, C.c_uint64, C.c_float, C.c_void_p, C.c_uint64, C.c_void_p,
            C.c_uint64]
    return _lib


def reference():
    """The compiled reference CPU codec (oracle/_ref/libtsref.so) or None."""
    global _ref
    if _ref is None and os.path.exists(REF_SO):
        _ref = C.CDLL(REF_SO)
        u64p = C.POINTER(C.c_uint64)
        _ref.ref_compress_bound.argtypes = [C.c_uint32] * 5 + [u64p]
        _ref.ref_image_span.argtypes = [C.c_uint32] * 3 + [C.c_uint64, u64p]
        _ref.ref_workspace_size.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32,
                                            C.c_uint64, u64p]
        _ref.ref_parse.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64,
                                   u64p, u64p]
        _ref.ref_compress.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64,
                                      C.c_void_p, C.c_uint64, C.c_uint32, C.c_uint32,
                                      C.c_void_p, C.c_uint64, u64p]
        _ref.ref_decompress.argtypes = [C.c_void_p, C.c_uint64, C.c_uint32, C.c_uint32,
                                        C.c_uint32, C.c_void_p, C.c_uint64, C.c_uint64]
    return _ref


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def type_of(pixels: np.ndarray) -> int:
    if pixels.dtype == np.uint16:
        return UINT16
    if pixels.dtype in (np.uint32, np.float32):
        return FLOAT32
    raise TypeError(pixels.dtype)


def compress_bound(ptype, width, height, tile=(128, 32)):
    out = C.c_uint64(0)
    st = lib().orc_compress_bound(ptype, width, height, tile[0], tile[1], C.byref(out))
    return st, out.value


def image_span(ptype, width, height, pitch):
    out = C.c_uint64(0)
    st = lib().orc_image_span(ptype, width, height, pitch, C.byref(out))
    return st, out.value


def workspace_size(images, tiles, tile, temp):
    out = C.c_uint64(0)
    st = lib().orc_workspace_size(images, tiles, tile[0], tile[1], temp, C.byref(out))
    return st, out.value


def compress(pixels: np.ndarray, tile=(128, 32)) -> bytes:
    """Canonical encode of a 2-D uint16 or float32/uint32-bit-word array."""
    ptype = type_of(pixels)
    px = np.ascontiguousarray(pixels)
    h, w = px.shape
    st, bound = compress_bound(ptype, w, h, tile)
    if st:
        raise ValueError(f"compress_bound status {st}")
    out = np.empty(bound, np.uint8)
    written = C.c_uint64(0)
    pitch = px.strides[0] if h > 1 else w * px.itemsize
    st = lib().orc_compress(ptype, w, h, pitch, _ptr(px), tile[0], tile[1], _ptr(out),
                            bound, C.byref(written))
    if st:
        raise ValueError(f"compress status {st}")
    return out[: written.value].tobytes()


def parse(file: bytes, limits=(L, L, L)):
    """(status, info, tiles, segments) of the validating parser."""
    buf = np.frombuffer(file, np.uint8) if len(file) else np.zeros(1, np.uint8)
    info = Info()
    st = lib().orc_parse(_ptr(buf) if len(file) else None, len(file), *limits, C.byref(info),
                         None, None)
    if st:
        return st, None, None, None
    tiles = (Tile * max(1, info.tile_count))()
    segs = (Segment * max(1, info.segment_count))()
    st = lib().orc_parse(_ptr(buf), len(file), *limits, C.byref(info), tiles, segs)
    return st, info, list(tiles)[: info.tile_count], list(segs)[: info.segment_count]


def parse_status(file: bytes, limits=(L, L, L)) -> int:
    buf = np.frombuffer(file, np.uint8) if len(file) else None
    info = Info()
    return lib().orc_parse(_ptr(buf), len(file), *limits, C.byref(info), None, None)


def decompress(file: bytes) -> np.ndarray:
    st, info, _, _ = parse(file)
    if st:
        raise ValueError(f"parse status {st}")
    dt = np.uint16 if info.type == UINT16 else np.uint32
    out = np.empty((info.height, info.width), dt)
    buf = np.frombuffer(file, np.uint8)
    st = lib().orc_decompress(_ptr(buf), len(file), _ptr(out), out.strides[0])
    if st:
        raise ValueError(f"decompress status {st}")
    return out


def decode_roi(file: bytes, roi) -> np.ndarray:
    x, y, w, h = roi
    st, info, _, _ = parse(file)
    if st:
        raise ValueError(f"parse status {st}")
    dt = np.uint16 if info.type == UINT16 else np.uint32
    out = np.empty((h, w), dt)
    buf = np.frombuffer(file, np.uint8)
    st = lib().orc_decode_roi(_ptr(buf), len(file), x, y, w, h, _ptr(out), out.strides[0])
    if st:
        raise ValueError(f"decode_roi status {st}")
    return out


def calibrate(samples: np.ndarray, offset, gain) -> np.ndarray:
    """Host binary32 (x - offset) * gain; offset/gain are floats or FP32 maps."""
    ptype = type_of(samples)
    s = np.ascontiguousarray(samples)
    h, w = s.shape
    out = np.empty((h, w), np.float32)
    om = np.ascontiguousarray(offset, np.float32) if isinstance(offset, np.ndarray) else None
    gm = np.ascontiguousarray(gain, np.float32) if isinstance(gain, np.ndarray) else None
    lib().orc_calibrate(ptype, w, h, _ptr(s), s.strides[0],
                        0.0 if om is not None else float(offset), _ptr(om),
                        om.strides[0] if om is not None else 0,
                        1.0 if gm is not None else float(gain), _ptr(gm),
                        gm.strides[0] if gm is not None else 0, _ptr(out), out.strides[0])
    return out


def decode_roi_calibrate(file: bytes, roi, offset, gain) -> np.ndarray:
    x, y, w, h = roi
    out = np.empty((h, w), np.float32)
    om = np.ascontiguousarray(offset, np.float32) if isinstance(offset, np.ndarray) else None
    gm = np.ascontiguousarray(gain, np.float32) if isinstance(gain, np.ndarray) else None
    buf = np.frombuffer(file, np.uint8)
    st = lib().orc_decode_roi_calibrate(
        _ptr(buf), len(file), x, y, w, h, 0.0 if om is not None else float(offset), _ptr(om),
        om.strides[0] if om is not None else 0, 1.0 if gm is not None else float(gain),
        _ptr(gm), gm.strides[0] if gm is not None else 0, _ptr(out), out.strides[0])
    if st:
        raise ValueError(f"decode_roi_calibrate status {st}")
    return out


# ---- compiled reference CPU codec (oracle/_ref) -------------------------------

def ref_compress(pixels: np.ndarray, tile=(128, 32)) -> bytes:
    r = reference()
    ptype = type_of(pixels)
    px = np.ascontiguousarray(pixels)
    h, w = px.shape
    st, bound = compress_bound(ptype, w, h, tile)
    out = np.empty(bound, np.uint8)
    written = C.c_uint64(0)
    pitch = px.strides[0] if h > 1 else w * px.itemsize
    st = r.ref_compress(ptype, w, h, pitch, _ptr(px), px.nbytes, tile[0], tile[1],
                        _ptr(out), bound, C.byref(written))
    if st:
        raise ValueError(f"reference compress status {st}")
    return out[: written.value].tobytes()


def ref_decompress(file: bytes, ptype, width, height) -> np.ndarray:
    r = reference()
    dt = np.uint16 if ptype == UINT16 else np.uint32
    out = np.empty((height, width), dt)
    buf = np.frombuffer(file, np.uint8)
    st = r.ref_decompress(_ptr(buf), len(file), ptype, width, height, _ptr(out), out.strides[0],
                          out.nbytes)
    if st:
        raise ValueError(f"reference decompress status {st}")
    return out


def ref_parse_status(file: bytes, limits=(L, L, L)) -> int:
    r = reference()
    buf = np.frombuffer(file, np.uint8) if len(file) else None
    t, s = C.c_uint64(0), C.c_uint64(0)
    return r.ref_parse(_ptr(buf), len(file), *limits, C.byref(t), C.byref(s))
