"""Execution API behaviour on a GPU (reference tests/cuda_api_test.cpp:209-246,
365-378): independent plans on two streams and on two host threads give the
same results as alone, and a pending HIP error from the caller's earlier work
blocks enqueue (hip_error) and is left for the caller."""
import threading

import numpy as np
import pytest

import tilestream_amd as ts
from tilestream_amd import _native as N
from tilestream_amd import api, generators as g

from helpers import DeviceOutput

pytestmark = pytest.mark.gpu


class Decode:
    """One decompress plan with its buffers, for an archive of `px`."""

    def __init__(self, orc, px, tile, stream=None):
        self.px = px
        data = orc.compress(px, tile)
        st, self.meta = api.parse_metadata(data)
        assert st == ts.Status.success
        self.src = api.DeviceMemory.from_host(np.frombuffer(data, np.uint8))
        h, w = px.shape
        self.out = DeviceOutput(api.pixel_type_of(px.dtype), w, h)
        item = api.BatchItem()
        item.output = self.out.view()
        item.compressed_input.bytes = api.ConstBytes(self.src.ptr, self.src.size)
        item.compressed_input.metadata = self.meta.handle
        st, self.plan = api.prepare_batch(ts.Operation.decompress, [item], stream=stream)
        assert st == ts.Status.success, st
        self.item = item
        self.ws = api.DeviceMemory(max(256, self.plan.workspace_bytes()))
        self.res = api.DeviceMemory(16)

    def enqueue(self, stream):
        return api.decompress_async(self.plan, self.ws, self.res, 1, stream)

    def check(self):
        status, written = api.read_results(self.res, 1)
        assert status[0] == 0 and written[0] == self.px.nbytes
        assert np.array_equal(self.out.read(), self.px.view(np.uint16 if self.px.dtype == np.uint16 else np.uint32))


def test_two_streams(orc):
    s1, s2 = api.Stream(), api.Stream()
    a = Decode(orc, g.detector_u16(1500, 900, 90), (128, 32), s1)
    b = Decode(orc, g.field_f32(1100, 700, 91), (64, 64), s2)
    for _ in range(3):
        assert a.enqueue(s1) == ts.Status.success
        assert b.enqueue(s2) == ts.Status.success
    s1.synchronize()
    s2.synchronize()
    a.check()
    b.check()


def test_two_host_threads(orc):
    errors = []

    def work(seed):
        try:
            s = api.Stream()
            d = Decode(orc, g.detector_u16(900 + seed, 700, seed), (256, 16), s)
            for _ in range(5):
                assert d.enqueue(s) == ts.Status.success
            s.synchronize()
            d.check()
        except Exception as exc:  # noqa: BLE001 - reported below
            errors.append(repr(exc))

    threads = [threading.Thread(target=work, args=(92 + k,)) for k in range(2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors


def test_pending_caller_error_blocks_enqueue(orc):
    d = Decode(orc, g.detector_u16(300, 200, 95), (64, 64))
    assert N.lib().ts_hip_get_last_error() == 0
    ptr = api.C.c_void_p(0)
    assert N.lib().ts_hip_malloc(1 << 60, api.C.byref(ptr)) != 0  # the caller's failed allocation
    assert N.lib().ts_hip_peek_last_error() != 0
    assert d.enqueue(None) == ts.Status.hip_error
    assert N.lib().ts_hip_peek_last_error() != 0  # left for the caller
    assert N.lib().ts_hip_get_last_error() != 0   # the caller clears it
    assert d.enqueue(None) == ts.Status.success
    assert api.synchronize(None) == ts.Status.success
    d.check()


def test_enqueue_returns_before_the_work_completes(orc):
    """Execution never waits for the device (reference tests/cuda_api_test.cpp:263-333):
    enqueueing milliseconds of decode work returns long before it completes."""
    import time
    s = api.Stream()
    d = Decode(orc, g.field_f32(8192, 4096, 96), (64, 64), s)  # 128 MiB decoded per call
    assert d.enqueue(s) == ts.Status.success  # first run outside the timing
    s.synchronize()
    t0 = time.perf_counter()
    for _ in range(100):
        assert d.enqueue(s) == ts.Status.success
    t1 = time.perf_counter()
    s.synchronize()
    t2 = time.perf_counter()
    assert t2 - t0 > 0.002, "too little work to tell enqueue from completion"
    assert t1 - t0 < 0.5 * (t2 - t0), (t1 - t0, t2 - t0)
    d.check()
