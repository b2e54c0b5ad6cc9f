"""The allocation-free encode entry point (ts_compress_async_static): the same
archive bytes and results as ts_compress_async with a workspace, on calls in a
row whose plans shrink (the static device scratch keeps the previous call's
contents and must be cleared by each call)."""
import ctypes as C

import numpy as np
import pytest

import tilestream_amd as ts
from tilestream_amd import _native as N
from tilestream_amd import api, generators as g
from tilestream_amd._native import BatchItem, TileShape

pytestmark = pytest.mark.gpu


class Batch:
    def __init__(self, images, tile):
        self.images, self.keep, items, self.outs, self.bounds = images, [], [], [], []
        for px in images:
            h, w = px.shape
            src = api.DeviceMemory.from_host(np.ascontiguousarray(px))
            st, bound = api.compress_bound(api.host_view(px).image, TileShape(*tile))
            assert st == ts.Status.success, st
            out = api.DeviceMemory(bound)
            it = BatchItem()
            it.input = api.device_view(src, api.pixel_type_of(px.dtype), w, h, w * px.itemsize, const=True)
            it.compressed_output = api.MutableBytes(out.ptr, bound)
            items.append(it)
            self.outs.append(out)
            self.bounds.append(bound)
            self.keep.append(src)
        st, self.plan = api.prepare_batch(ts.Operation.compress, items, TileShape(*tile))
        assert st == ts.Status.success, st
        self.items = items
        self.results = api.DeviceMemory(16 * len(images))

    def collect(self):
        assert api.synchronize(None) == ts.Status.success
        status, written = api.read_results(self.results, len(self.images))
        assert (status == 0).all(), status
        return [out.download(int(n)).tobytes() for out, n in zip(self.outs, written)], list(written)

    def with_workspace(self):
        for out in self.outs:
            out.fill(0x11)
        self.results.fill(0x5A)
        ws = api.DeviceMemory(self.plan.workspace_bytes())
        assert api.compress_async(self.plan, ws, self.results, len(self.images)) == ts.Status.success
        got = self.collect()
        del ws
        return got

    def static(self):
        for out in self.outs:
            out.fill(0x22)
        self.results.fill(0x5A)
        st = N.lib().ts_compress_async_static(self.plan.handle, self.results.ptr, len(self.images), None)
        assert st == ts.Status.success, ts.Status(st)
        return self.collect()


def tiles(batch, tile):
    return sum(-(-px.shape[1] // tile[0]) * -(-px.shape[0] // tile[1]) for px in batch)


def test_static_entry_matches_workspace_entry_on_shrinking_calls(orc):
    rng = np.random.default_rng(7)
    # 1: 67,584 tiles, more than one launch's state fits the scratch: runs by windows.
    big = [rng.integers(0, 1 << 12, (4096, 4096 + 64 * (i % 2)), dtype=np.uint16) for i in range(16)]
    big.append(g.detector_u16(2048, 512, 3))
    assert tiles(big, (64, 64)) > 65535
    # 2: about 5,000 tiles over several images: one launch over the whole batch.
    mid = [g.field_f32(int(rng.integers(300, 1500)), int(rng.integers(300, 1500)), 10 + i) for i in range(12)]
    # 3: one small image, fewer tiles than the calls before it left in the scratch.
    small = [g.detector_u16(300, 200, 5)]
    for images, tile in ((big, (64, 64)), (mid, (128, 32)), (small, (256, 16))):
        b = Batch(images, tile)
        ref_files, ref_written = b.with_workspace()
        files, written = b.static()
        assert written == ref_written
        assert files == ref_files, f"{len(images)} images, {tiles(images, tile)} tiles"
        del b
    # The small archive is also the oracle's, byte for byte.
    assert files[0] == orc.compress(small[0], (256, 16))


def test_static_entry_checks():
    b = Batch([g.detector_u16(200, 100, 1)], (128, 32))
    lib = N.lib()
    # Results misaligned, too few results, a plan of another operation.
    assert lib.ts_compress_async_static(b.plan.handle, b.results.ptr + 4, 1, None) == ts.Status.invalid_argument
    assert lib.ts_compress_async_static(b.plan.handle, b.results.ptr, 0, None) == ts.Status.buffer_too_small
    assert lib.ts_compress_async_static(None, b.results.ptr, 1, None) == ts.Status.invalid_argument
    # Results inside the plan's own output range overlap it.
    assert lib.ts_compress_async_static(b.plan.handle, b.outs[0].ptr, 1, None) == ts.Status.invalid_argument
    assert b.static()[1] == b.with_workspace()[1]
