"""GPU canonical encoder: archive bytes identical to the CPU oracle; the trusted
encode -> decode chain without host metadata; device-side corruption detection."""
import json
import os

import numpy as np
import pytest

import tilestream_amd as ts
from tilestream_amd import api, generators as g
from tilestream_amd._native import BatchItem, ConstBytes, Roi, TileShape

from helpers import DeviceOutput

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
SHAPES = [(128, 32), (64, 64), (256, 16)]


class Encoded:
    def __init__(self, images, tile=(128, 32), guard=64):
        self.images = images
        self.keep = []
        items, self.outs, self.bounds = [], [], []
        for px in images:
            h, w = px.shape
            pitch = (w + 5) * px.itemsize
            host = np.zeros((h, w + 5), px.dtype)
            host[:, :w] = px
            src = api.DeviceMemory.from_host(host)
            st, bound = api.compress_bound(api.host_view(px).image, TileShape(*tile))
            out = api.DeviceMemory(bound + guard)
            out.fill(0xA5)
            it = BatchItem()
            it.input = api.device_view(src, api.pixel_type_of(px.dtype), w, h, pitch, const=True)
            it.compressed_output = api.MutableBytes(out.ptr, bound)
            items.append(it)
            self.outs.append(out)
            self.bounds.append(bound)
            self.keep.append(src)
        st, self.plan = api.prepare_batch(ts.Operation.compress, items, TileShape(*tile))
        assert st == ts.Status.success, st
        self.items = items
        self.ws = api.DeviceMemory(self.plan.workspace_bytes())
        self.results = api.DeviceMemory(16 * len(images))

    def run(self, stream=None):
        assert api.compress_async(self.plan, self.ws, self.results, len(self.images), stream) == ts.Status.success
        assert api.synchronize(stream) == ts.Status.success
        status, written = api.read_results(self.results, len(self.images))
        assert (status == 0).all(), status
        files = []
        for out, bound, n in zip(self.outs, self.bounds, written):
            raw = out.download()
            assert (raw[bound:] == 0xA5).all(), "guard bytes after the bound were written"
            files.append(raw[:n].tobytes())
        return files


def test_golden_files_on_gpu():
    kat = json.load(open(os.path.join(HERE, "golden", "kat.json")))["cases"]
    for c in kat:
        dt = np.uint16 if c["type"] == 1 else np.uint32
        px = np.array(c["pixels"], dt).reshape(c["height"], c["width"])
        assert Encoded([px]).run()[0].hex() == c["file"], c["name"]


@pytest.mark.parametrize("tile", SHAPES)
def test_random_batches_match_cpu(orc, tile):
    rng = np.random.default_rng(60)
    imgs = []
    for i in range(12):
        w, h = int(rng.integers(1, 400)), int(rng.integers(1, 300))
        imgs.append(g.detector_u16(w, h, 60 + i))
    for px, data in zip(imgs, Encoded(imgs, tile).run()):
        assert data == orc.compress(px, tile)
    imgs = [g.field_f32(int(rng.integers(1, 400)), int(rng.integers(1, 300)), 70 + i) for i in range(12)]
    for px, data in zip(imgs, Encoded(imgs, tile).run()):
        assert data == orc.compress(px, tile)


def test_every_width(orc):
    from test_oracle import _row_with_width
    rng = np.random.default_rng(61)
    for ptype, top in [(1, 16), (2, 32)]:
        for w in (1, 2, 31, 32, 33, 127, 128, 129, 255, 256):
            rows = [_row_with_width(rng, w, b, ptype) for b in range(top + 1) if w > 1 or b == 0]
            px = np.stack(rows)
            for tile in SHAPES:
                assert Encoded([px], tile).run()[0] == orc.compress(px, tile), (ptype, w, tile)


def test_image_split_across_windows(orc):
    """14,005 tiles of one image in one plan (the plan's whole-batch window; windows that carry
    the output cursor are covered by the serving tests with 64- and 200-tile windows)."""
    px = g.detector_u16(128 * 5, 32 * 2801, 62)
    data = Encoded([px], (128, 32)).run()[0]
    assert data == orc.compress(px, (128, 32))


def test_many_tiny_images(orc):
    rng = np.random.default_rng(63)
    imgs = [rng.integers(0, 1 << int(rng.integers(0, 17)), (int(rng.integers(1, 40)), int(rng.integers(1, 40))))
            .astype(np.uint16) for _ in range(300)]
    for px, data in zip(imgs, Encoded(imgs, (64, 64)).run()):
        assert data == orc.compress(px, (64, 64))


def _trusted_decode(enc, index, roi=None):
    px = enc.images[index]
    h, w = px.shape
    x, y, rw, rh = roi if roi is not None else (0, 0, w, h)
    out = DeviceOutput(api.pixel_type_of(px.dtype), rw, rh)
    it = BatchItem()
    it.output = out.view()
    it.compressed_input.bytes = ConstBytes(enc.outs[index].ptr, enc.bounds[index])
    it.compressed_input.encoder_result = enc.results.ptr + 16 * index
    it.compressed_input.trusted_image = api.host_view(px).image
    it.compressed_input.trusted_tile = TileShape(*enc.tile)
    it.roi = Roi(x, y, rw, rh)
    op = ts.Operation.decode_roi if roi is not None else ts.Operation.decompress
    st, plan = api.prepare_batch(op, [it])
    assert st == ts.Status.success, st
    ws = api.DeviceMemory(plan.workspace_bytes())
    res = api.DeviceMemory(16)
    fn = api.decode_roi_async if roi is not None else api.decompress_async
    assert fn(plan, ws, res, 1) == ts.Status.success
    assert api.synchronize() == ts.Status.success
    status, written = api.read_results(res, 1)
    return int(status[0]), int(written[0]), out.read()


@pytest.mark.parametrize("tile", SHAPES)
def test_trusted_encoder_output_decodes(tile):
    for imgs in ([g.detector_u16(517, 301, 64), g.detector_u16(64, 70, 67)],
                 [g.field_f32(300, 250, 65), g.field_f32(129, 33, 68)]):
        enc = Encoded(imgs, tile)
        enc.tile = tile
        enc.run()
        for i, px in enumerate(imgs):
            st, n, out = _trusted_decode(enc, i)
            assert st == 0 and n == px.nbytes and np.array_equal(out, px)
            rw, rh = min(60, px.shape[1] - 3), min(25, px.shape[0] - 2)
            st, n, out = _trusted_decode(enc, i, (3, 2, rw, rh))
            assert st == 0 and np.array_equal(out, px[2:2 + rh, 3:3 + rw])


def test_trusted_path_detects_corruption_and_producer_failure():
    px = g.detector_u16(300, 200, 66)
    enc = Encoded([px], (64, 64))
    enc.tile = (64, 64)
    enc.run()
    # A wrong mode byte in the device index is caught on the device.
    api._check_hip(api.N.lib().ts_hip_memset(enc.outs[0].ptr + 64, 7, 1), "memset")
    st, n, _ = _trusted_decode(enc, 0)
    assert st == ts.Status.malformed_data and n == 0
    # A failed producer record is propagated with zero bytes written.
    enc.run()
    bad = np.array([ts.Status.size_overflow, 0, 0, 0], np.uint32)
    enc.results.upload(bad)
    st, n, _ = _trusted_decode(enc, 0)
    assert st == ts.Status.size_overflow and n == 0
