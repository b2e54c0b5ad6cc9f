"""Host side of the library without a GPU: the C-ABI exports, record layouts,
sizing, the product parser and CPU codec against the oracle, and the request
validation that returns before any device call."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import tilestream_amd as ts
from tilestream_amd import _native as N
from tilestream_amd import api, generators as g
from tilestream_amd._native import BatchItem, ConstBytes, ImageSpec, Roi, TileShape, WorkspaceSpec

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b((?:ts|tsk)_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.parametrize("header", ["tilestream_c.h", "tilestream_kernels.h"])
def test_library_exports_every_declared_symbol(header):
    lib = C.CDLL(N.LIB_PATH)
    names = declared(header)
    assert len(names) >= 10
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_record_layouts_match():
    for name, cls in N.RECORDS.items():
        assert N.lib().ts_abi_sizeof(name.encode()) == C.sizeof(cls), name
    assert C.sizeof(N.DeviceResult) == 16
    assert C.sizeof(N.TileMetadata) == 32 and C.sizeof(N.SegmentMetadata) == 24


def test_sizing_matches_oracle(orc):
    rng = np.random.default_rng(12)
    for _ in range(3000):
        t = int(rng.integers(0, 4))
        w = int(rng.choice([0, 1, 3, 129, 4096, 2**31, 2**32 - 1]))
        h = int(rng.choice([0, 1, 33, 4096, 2**32 - 1]))
        tile = [(128, 32), (64, 64), (256, 16), (32, 128)][int(rng.integers(0, 4))]
        st, b = api.compress_bound(ImageSpec(t, w, h), TileShape(*tile))
        ost, ob = orc.compress_bound(t, w, h, tile)
        assert st == ost and (st or b == ob)
        pitch = int(rng.choice([0, 2, 4, 6, 1000, 2**40, 2**63]))
        st, b = api.image_span(ImageSpec(t, w, h), pitch)
        ost, ob = orc.image_span(t, w, h, pitch)
        assert st == ost and (st or b == ob)
        imgs, tl = int(rng.choice([0, 1, 2, 2**32])), int(rng.choice([0, 1, 2, 4096, 2**31]))
        temp = int(rng.choice([0, 1, 257, 2**62]))
        st, b = api.workspace_size(WorkspaceSpec(imgs, tl, TileShape(*tile), temp))
        ost, ob = orc.workspace_size(imgs, tl, tile, temp)
        assert st == ost and (st or b == ob)


def _archives(orc):
    rng = np.random.default_rng(5)
    yield orc.compress(rng.integers(0, 40, (33, 129)).astype(np.uint16))
    yield orc.compress(rng.integers(0, 2**32, (5, 7), dtype=np.uint64).astype(np.uint32))
    yield orc.compress(np.full((9, 3), 77, np.uint16))
    yield orc.compress(np.cumsum(rng.integers(0, 3, (17, 70)), axis=1).astype(np.uint32), (64, 64))
    yield orc.compress(rng.integers(0, 4, (3, 257)).astype(np.uint16), (256, 16))


def test_parser_statuses_match_oracle(orc):
    L = (1 << 63) - 1
    for data in _archives(orc):
        for n in range(len(data)):
            assert ts.parse_metadata(data[:n])[0] == orc.parse_status(data[:n])
        b = bytearray(data)
        for pos in range(len(b)):
            for v in {0, 1, 0xFF, b[pos] ^ 0x40}:
                if v == data[pos]:
                    continue
                b[pos] = v
                assert ts.parse_metadata(bytes(b))[0] == orc.parse_status(bytes(b)), (pos, v)
                b[pos] = data[pos]
        for lim in [(len(data) - 1, L, L), (L, 0, L), (L, L, 72), (L, L, 150)]:
            assert ts.parse_metadata(data, N.ParseLimits(*lim))[0] == orc.parse_status(data, lim)


def test_parser_tables_match_oracle(orc):
    for data in _archives(orc):
        st, meta = ts.parse_metadata(data)
        ost, info, tiles, segs = orc.parse(data)
        assert st == ost == 0
        s = meta.summary
        assert (s.tile_count, s.segment_count, s.file_bytes) == (info.tile_count, info.segment_count, len(data))
        for i, t in enumerate(tiles):
            m = meta.tile(i)
            assert (m.mode, m.width, m.height, m.offset, m.length, m.first_segment) == (
                t.mode, t.width, t.height, t.offset, t.length, t.first_segment)
        for i, sg in enumerate(segs):
            m = meta.segment(i)
            assert (m.seed, m.bit_width, m.payload_offset, m.payload_bytes) == (
                sg.seed, sg.bit_width, sg.payload_offset, sg.payload_bytes)


def test_threaded_parser_matches_oracle(orc):
    """Archives of several MiB are validated and tabulated by several host
    threads over contiguous tile ranges: same tables and statuses as the
    oracle's sequential walk, for mutations in every range and the index."""
    rng = np.random.default_rng(71)
    archives = [orc.compress(g.field_f32(2048, 1536, 72), (64, 64)),
                orc.compress(g.detector_u16(4096, 2048, 73), (256, 16))]
    for data in archives:
        assert len(data) > (8 << 20)
        st, meta = ts.parse_metadata(data)
        ost, info, tiles, segs = orc.parse(data)
        assert st == ost == 0
        assert (meta.summary.tile_count, meta.summary.segment_count) == (info.tile_count, info.segment_count)
        for i in list(range(0, len(tiles), 7)) + [len(tiles) - 1]:
            m, t = meta.tile(i), tiles[i]
            assert (m.mode, m.width, m.height, m.offset, m.length, m.first_segment) == (
                t.mode, t.width, t.height, t.offset, t.length, t.first_segment), i
        for i in list(range(0, len(segs), 31)) + [len(segs) - 1]:
            m, sg = meta.segment(i), segs[i]
            assert (m.seed, m.bit_width, m.payload_offset, m.payload_bytes) == (
                sg.seed, sg.bit_width, sg.payload_offset, sg.payload_bytes), i
        b = bytearray(data)
        index_end = 64 + 32 * info.tile_count
        positions = list(rng.integers(64, index_end, 8)) + list(rng.integers(index_end, len(data), 40))
        positions += [len(data) - 1, len(data) - 5, index_end, index_end + 4]
        for pos in positions:
            pos = int(pos)
            for v in {0, 0xFF, b[pos] ^ 0x40}:
                if v == data[pos]:
                    continue
                b[pos] = v
                assert ts.parse_metadata(bytes(b))[0] == orc.parse_status(bytes(b)), (pos, v)
                b[pos] = data[pos]
        for n in (len(data) - 1, len(data) // 2, index_end + 1):
            assert ts.parse_metadata(data[:n])[0] == orc.parse_status(data[:n])


@pytest.mark.parametrize("tile", [(128, 32), (64, 64), (256, 16)])
def test_product_cpu_codec_matches_oracle(orc, tile):
    for w, h, seed in [(1024, 512, 1), (333, 517, 2), (1, 1, 3), (129, 33, 4)]:
        for px in (g.detector_u16(w, h, seed), g.field_f32(w, h, seed)):
            st, data = ts.cpu_compress(px, TileShape(*tile))
            assert st == ts.Status.success
            assert data == orc.compress(px, tile)
            out = np.zeros_like(px)
            assert ts.cpu_decompress(data, out) == ts.Status.success
            assert np.array_equal(out, px)


def test_cpu_codec_argument_checks():
    px = np.zeros((4, 4), np.uint16)
    st, bound = api.compress_bound(ImageSpec(1, 4, 4))
    out = np.zeros(bound - 1, np.uint8)
    w = C.c_uint64(0)
    st = N.lib().ts_cpu_compress(api.host_view(px), TileShape(), N.MutableBytes(out.ctypes.data, out.size),
                                 C.byref(w))
    assert st == ts.Status.buffer_too_small
    bad = api.host_view(px)
    bad.pitch_bytes = 3
    assert N.lib().ts_cpu_compress(bad, TileShape(), N.MutableBytes(out.ctypes.data, out.size),
                                   C.byref(w)) == ts.Status.invalid_argument


# ---- request validation (host checks run before any device call) -----------------

def _decode_item(data, out, roi=(0, 0, 0, 0)):
    st, meta = ts.parse_metadata(data)
    assert st == ts.Status.success
    buf = np.frombuffer(data, np.uint8).copy()
    item = BatchItem()
    item.output = api.host_view(out, const=False)
    item.compressed_input.bytes = ConstBytes(buf.ctypes.data, buf.size)
    item.compressed_input.metadata = meta.handle
    item.roi = Roi(*roi)
    return item, (meta, buf, out)


def _prepare(op, items):
    return api.prepare_batch(op, items)[0]


def test_validation_matrix_before_device(orc):
    """Each malformed request returns its Status before touching the device; a
    valid request here (no GPU) reaches the runtime and reports hip_error."""
    px = g.detector_u16(200, 100, 3)
    data = orc.compress(px)
    out = np.zeros_like(px)
    item, keep = _decode_item(data, out)
    assert _prepare(ts.Operation.decompress, [item]) == ts.Status.hip_error
    # region checks
    for roi, expect in [((0, 0, 0, 5), ts.Status.invalid_roi), ((0, 0, 201, 1), ts.Status.invalid_roi),
                        ((2**32 - 1, 0, 2, 1), ts.Status.invalid_roi), ((0, 99, 1, 2), ts.Status.invalid_roi)]:
        it, k = _decode_item(data, np.zeros((max(roi[3], 1), max(roi[2], 1)), np.uint16), roi)
        assert _prepare(ts.Operation.decode_roi, [it]) == expect
    # output shape / type
    it, k = _decode_item(data, np.zeros((100, 199), np.uint16))
    assert _prepare(ts.Operation.decompress, [it]) == ts.Status.invalid_argument
    it, k = _decode_item(data, np.zeros((100, 200), np.uint32))
    assert _prepare(ts.Operation.decompress, [it]) == ts.Status.invalid_argument
    # fused decode needs FP32 output
    it, k = _decode_item(data, np.zeros((100, 200), np.uint16), (0, 0, 200, 100))
    assert _prepare(ts.Operation.decode_calibrate, [it]) == ts.Status.invalid_argument
    # a map of the wrong size
    it, k = _decode_item(data, np.zeros((10, 20), np.float32), (5, 5, 20, 10))
    m = np.zeros((10, 19), np.float32)
    it.calibration.offset = api.map_parameter(api.host_view(m))
    assert _prepare(ts.Operation.decode_calibrate, [it]) == ts.Status.invalid_argument
    it.calibration.offset = api.scalar(1.0)
    it.calibration.gain.kind = 7
    assert _prepare(ts.Operation.decode_calibrate, [it]) == ts.Status.invalid_argument
    # missing metadata, unaligned source, too-small source
    it, k = _decode_item(data, np.zeros_like(px))
    it.compressed_input.metadata = None
    assert _prepare(ts.Operation.decompress, [it]) == ts.Status.invalid_argument
    it, k = _decode_item(data, np.zeros_like(px))
    it.compressed_input.bytes = ConstBytes(it.compressed_input.bytes.data + 2, len(data) - 2)
    assert _prepare(ts.Operation.decompress, [it]) == ts.Status.invalid_argument
    it, k = _decode_item(data, np.zeros_like(px))
    it.compressed_input.bytes = ConstBytes(it.compressed_input.bytes.data, len(data) - 4)
    assert _prepare(ts.Operation.decompress, [it]) == ts.Status.buffer_too_small
    # aliasing: the output overlaps the compressed input
    it, k = _decode_item(data, np.zeros_like(px))
    it.output.data = it.compressed_input.bytes.data
    it.output.capacity_bytes = px.nbytes
    assert _prepare(ts.Operation.decompress, [it]) in (ts.Status.invalid_argument,)
    # two outputs overlap
    out2 = np.zeros_like(px)
    a, k1 = _decode_item(data, out2)
    b, k2 = _decode_item(data, out2)
    assert _prepare(ts.Operation.decompress, [a, b]) == ts.Status.invalid_argument
    # one pixel type per plan
    f = g.field_f32(30, 30, 1)
    c, k3 = _decode_item(orc.compress(f), np.zeros_like(f))
    assert _prepare(ts.Operation.decompress, [a, c]) == ts.Status.invalid_argument
    # fabricated metadata is rejected host-side
    it, k = _decode_item(data, np.zeros_like(px))
    meta = k[0]
    seg = meta.segment(0)
    seg.bit_width = 17  # wider than uint16 residuals allow
    meta.set_segment(0, seg)
    assert _prepare(ts.Operation.decompress, [it]) == ts.Status.malformed_data
    it, k = _decode_item(data, np.zeros_like(px))
    meta = k[0]
    t = meta.tile(1)
    t.mode = 3
    meta.set_tile(1, t)
    assert _prepare(ts.Operation.decompress, [it]) == ts.Status.malformed_data
    it, k = _decode_item(data, np.zeros_like(px))
    meta = k[0]
    t = meta.tile(0)
    t.width -= 1
    meta.set_tile(0, t)
    assert _prepare(ts.Operation.decompress, [it]) == ts.Status.malformed_data


def test_request_level_checks():
    req = N.BatchRequest(1, None, 1, TileShape())
    h = C.c_void_p(0)
    assert N.lib().ts_prepare_batch(C.byref(req), C.byref(h), None) == ts.Status.invalid_argument
    req = N.BatchRequest(9, None, 0, TileShape())
    assert N.lib().ts_prepare_batch(C.byref(req), C.byref(h), None) == ts.Status.invalid_argument
    req = N.BatchRequest(1, None, 2**32, TileShape())
    assert N.lib().ts_prepare_batch(C.byref(req), C.byref(h), None) == ts.Status.resource_limit
    # an empty batch needs no device and no workspace
    req = N.BatchRequest(1, None, 0, TileShape())
    assert N.lib().ts_prepare_batch(C.byref(req), C.byref(h), None) == ts.Status.success
    plan = api.Plan(h.value, None)
    assert plan.workspace_bytes() == 0
    assert api.decompress_async(plan, None, None, 0) == ts.Status.success
    assert api.decode_roi_async(plan, None, None, 0) == ts.Status.invalid_argument


def test_compress_validation(orc):
    px = g.detector_u16(64, 64, 1)
    item = BatchItem()
    item.input = api.host_view(px)
    st, bound = api.compress_bound(item.input.image)
    out = np.zeros(bound, np.uint8)
    item.compressed_output = N.MutableBytes(out.ctypes.data, bound - 1)
    assert _prepare(ts.Operation.compress, [item]) == ts.Status.buffer_too_small
    item.compressed_output = N.MutableBytes(out.ctypes.data + 2, bound)
    assert _prepare(ts.Operation.compress, [item]) == ts.Status.invalid_argument
    item.compressed_output = N.MutableBytes(out.ctypes.data, bound)
    assert _prepare(ts.Operation.compress, [item]) == ts.Status.hip_error
    assert api.prepare_batch(ts.Operation.compress, [item], TileShape(32, 32))[0] == ts.Status.invalid_argument
