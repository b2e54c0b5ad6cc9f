"""GPU decode parity with the CPU oracle: whole image, regions, every width,
special words, partial tiles, lane-group widths, staged payloads."""
import numpy as np
import pytest

import tilestream_amd as ts
from tilestream_amd import generators as g

from helpers import Job, decode_on_gpu, upload_archive

pytestmark = pytest.mark.gpu

SHAPES = [(128, 32), (64, 64), (256, 16)]


def test_detector_frame_1024_bit_exact(orc):
    """uint16 1024x1024 detector frame, 128x32 tiles: CPU encode, one GPU decode."""
    px = g.detector_u16(1024, 1024, 2026517)
    data = orc.compress(px, (128, 32))
    out, written = decode_on_gpu(data)
    assert written == px.nbytes
    assert np.array_equal(out, px)


@pytest.mark.parametrize("tile", SHAPES)
@pytest.mark.parametrize("kind", ["u16", "f32"])
def test_whole_image_shapes(orc, tile, kind):
    w, h = 517, 301  # partial tiles on both edges
    px = g.detector_u16(w, h, 5) if kind == "u16" else g.field_f32(w, h, 5)
    data = orc.compress(px, tile)
    out, _ = decode_on_gpu(data)
    assert np.array_equal(out, px)


SHAPE_GRID = [(g_, b_, d_) for g_ in (1, 2, 4, 8, 16) for b_ in (64, 128, 256, 512) for d_ in (1, 2, 3)]


@pytest.mark.parametrize("group,block,depth", SHAPE_GRID)
def test_lane_group_widths(orc, group, block, depth):
    assert ts.api.decode_shape_supported(group, block, depth)
    for kind, tile in [("u16", (128, 32)), ("f32", (64, 64)), ("u16", (256, 16))]:
        px = g.detector_u16(300, 200, 9) if kind == "u16" else g.field_f32(300, 200, 9)
        data = orc.compress(px, tile)
        for grid in (0, 3):  # one workgroup per tile, and 3 workgroups looping over all tiles
            out, _ = decode_on_gpu(data, shape=(group, block, grid, depth))
            assert np.array_equal(out, px), (kind, tile, group, block, depth, grid)


def test_every_width_and_partial_rows(orc):
    from test_oracle import _row_with_width
    rng = np.random.default_rng(4)
    for ptype, top in [(1, 16), (2, 32)]:
        for w in (1, 2, 31, 32, 33, 63, 64, 65, 127, 128, 129, 255, 256):
            rows = []
            for bits in range(top + 1):
                if w == 1 and bits:
                    continue
                rows.append(_row_with_width(rng, w, bits, ptype))
            px = np.stack(rows)
            for tile in SHAPES:
                data = orc.compress(px, tile)
                out, _ = decode_on_gpu(data, shape=(8, 256, 0, 1))
                assert np.array_equal(out, px), (ptype, w, tile)


def test_all_uint16_values_and_special_words(orc):
    px = np.arange(65536, dtype=np.uint32).astype(np.uint16).reshape(256, 256)
    for tile in SHAPES:
        out, _ = decode_on_gpu(orc.compress(px, tile))
        assert np.array_equal(out, px)
    specials = np.array([0, 0x80000000, 0x7F800000, 0xFF800000, 0x7FC00000, 0xFFC00001, 0x7F800001,
                         0xFF812345, 0x00000001, 0x807FFFFF, 0x00400000, 0x3F800000, 0xBF800000,
                         0x7F7FFFFF, 0xFF7FFFFF, 0x00800000], np.uint32)
    rng = np.random.default_rng(8)
    f = specials[rng.integers(0, 16, (70, 130))]
    for tile in SHAPES:
        out, _ = decode_on_gpu(orc.compress(f, tile))
        assert np.array_equal(out, f)


@pytest.mark.parametrize("kind", ["u16", "f32"])
def test_random_rois_match_crop(orc, kind):
    w, h = 700, 389
    px = g.detector_u16(w, h, 21) if kind == "u16" else g.field_f32(w, h, 21)
    rng = np.random.default_rng(22)
    for tile in SHAPES:
        data = orc.compress(px, tile)
        for _ in range(6):
            rw, rh = int(rng.integers(1, w + 1)), int(rng.integers(1, h + 1))
            x, y = int(rng.integers(0, w - rw + 1)), int(rng.integers(0, h - rh + 1))
            out, written = decode_on_gpu(data, roi=(x, y, rw, rh))
            assert np.array_equal(out, px[y:y + rh, x:x + rw]), (tile, x, y, rw, rh)
            assert written == rw * rh * px.itemsize
            assert np.array_equal(out, orc.decode_roi(data, (x, y, rw, rh)))


def test_mixed_geometries_in_one_batch(orc):
    job = Job()
    expected = []
    for i, tile in enumerate(SHAPES * 2):
        px = g.detector_u16(100 + 37 * i, 60 + 29 * i, 30 + i)
        job.add_decode(orc.compress(px, tile))
        expected.append(px)
    status, written, outs = job.run(ts.Operation.decompress)
    assert (status == 0).all()
    for o, e in zip(outs, expected):
        assert np.array_equal(o, e)


def test_staged_payloads_in_reverse_order(orc):
    """Only the tiles of the region are staged, in reverse order, into a fresh buffer."""
    px = g.field_f32(600, 300, 40)
    data = orc.compress(px, (128, 32))
    st, meta = ts.parse_metadata(data)
    roi = (100, 50, 300, 120)
    across = (600 + 127) // 128
    sel = [r * across + c for r in range(50 // 32, (50 + 120 - 1) // 32 + 1)
           for c in range(100 // 128, (100 + 300 - 1) // 128 + 1)]
    tiles = [meta.tile(i) for i in sel]
    total = sum(t.length for t in tiles)
    staged = bytearray(total + 12)
    pos = 12
    mappings = []
    for i, t in reversed(list(zip(sel, tiles))):
        staged[pos:pos + t.length] = data[t.offset:t.offset + t.length]
        mappings.append((i, pos))
        pos += t.length
    mappings.sort()
    mem = upload_archive(bytes(staged))
    job = Job()
    job.add_decode(data, roi=roi, mem=mem, metadata=meta, mappings=mappings)
    status, written, outs = job.run(ts.Operation.decode_roi)
    assert status[0] == 0
    x, y, w, h = roi
    assert np.array_equal(outs[0], px[y:y + h, x:x + w])


def test_unselected_payloads_poisoned(orc):
    """An ROI decode reads only the tiles it needs: poison every other payload."""
    px = g.detector_u16(640, 256, 41)
    data = bytearray(orc.compress(px, (64, 64)))
    st, meta = ts.parse_metadata(bytes(data))
    roi = (70, 70, 100, 50)
    keep = set()
    for r in range(70 // 64, (70 + 50 - 1) // 64 + 1):
        for c in range(70 // 64, (70 + 100 - 1) // 64 + 1):
            keep.add(r * 10 + c)
    for i in range(meta.summary.tile_count):
        if i not in keep:
            t = meta.tile(i)
            data[t.offset:t.offset + t.length] = b"\xff" * t.length
    job = Job()
    job.add_decode(bytes(data), roi=roi, metadata=meta)
    status, _, outs = job.run(ts.Operation.decode_roi)
    assert status[0] == 0
    assert np.array_equal(outs[0], px[70:120, 70:170])


@pytest.mark.parametrize("kind", ["u16", "f32"])
def test_large_plan_striding_launch(orc, kind):
    """Decode plans of 2048 tiles or more launch 1024 workgroups that stride over
    the tiles; the image equals the one-workgroup-per-tile launch's and the
    source pixels."""
    w, h = 4100, 2090  # 65 x 33 = 2145 tiles of 64x64, ragged last column and row
    px = g.detector_u16(w, h, 71) if kind == "u16" else g.field_f32(w, h, 71)
    data = orc.compress(px, (64, 64))
    striding, _ = decode_on_gpu(data)
    per_tile, n = decode_on_gpu(data, shape=(8, 256, 0, 1))
    assert n == px.nbytes
    word = np.uint16 if kind == "u16" else np.uint32
    assert np.array_equal(np.ascontiguousarray(striding).view(word), np.ascontiguousarray(per_tile).view(word))
    assert np.array_equal(np.ascontiguousarray(striding).view(word), np.ascontiguousarray(px).view(word))
