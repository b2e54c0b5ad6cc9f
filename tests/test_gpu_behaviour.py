"""API behaviour on the GPU that the reference tests check around execution
(reference tests/cuda_decoder_test.cpp:466-656, tests/cuda_api_test.cpp:248-363,
tests/cuda_loading_test.cpp:48-77), and parity at full size:

* every rejection of an execution call (workspace too small or misaligned,
  results misaligned or too few, either overlapping a buffer of the batch or
  the other, pageable host memory where the device cannot reach it, the wrong
  operation, an unrepresentable result count, the trusted producer's record
  as results) returns its status before anything is enqueued, and leaves the
  results and the output untouched;
* ts_buffer_allocate / ts_buffer_release ownership rules, and 50 prepare /
  allocate / execute / release / destroy cycles return every byte to the
  stream-ordered pool;
* preparing, executing and releasing on one stream never waits for another,
  busy stream, and the first execution of a plan loads nothing;
* config 2 at full size (4096 x 4096 uint16 frame and float32 field, 64 x 64
  tiles, fused dark and gain maps) and regions of one 4096 x 4096 uint16
  archive fused with (x - 100) * 1.05, equal to the oracle word for word.

A stream of a second device cannot be made on a one-GPU machine, so the
foreign-stream rejection is covered by the CPU harness's stream check only."""
import ctypes as C
import time

import numpy as np
import pytest

import tilestream_amd as ts
from tilestream_amd import _native as N
from tilestream_amd import api, generators as g
from tilestream_amd._native import BatchItem, CompressedInput, ConstBytes, MutableBytes, Roi, TileShape

from helpers import DeviceOutput

pytestmark = pytest.mark.gpu

HIP = C.CDLL("libamdhip64.so")
HIP.hipStreamQuery.argtypes = [C.c_void_p]
HIP.hipDeviceGetDefaultMemPool.argtypes = [C.POINTER(C.c_void_p), C.c_int]
HIP.hipMemPoolGetAttribute.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
HIP.hipDeviceGetAttribute.argtypes = [C.POINTER(C.c_int), C.c_int, C.c_int]
HIP_ERROR_NOT_READY = 600
POOL_USED_MEM_CURRENT = 7


def _pool_in_use() -> int:
    pool = C.c_void_p(0)
    assert HIP.hipDeviceGetDefaultMemPool(C.byref(pool), 0) == 0
    used = C.c_uint64(0)
    assert HIP.hipMemPoolGetAttribute(pool, POOL_USED_MEM_CURRENT, C.byref(used)) == 0
    return used.value


class RoiJob:
    """A decode_roi plan over one 200 x 70 uint16 archive with its buffers."""

    def __init__(self, orc):
        self.px = g.detector_u16(200, 70, 23)
        self.data = orc.compress(self.px)
        st, self.meta = api.parse_metadata(self.data)
        self.src = api.DeviceMemory.from_host(np.frombuffer(self.data, np.uint8))
        self.out = DeviceOutput(ts.PixelType.uint16, 120, 50)
        it = BatchItem()
        it.output = self.out.view()
        it.roi = Roi(30, 10, 120, 50)
        it.compressed_input.bytes = ConstBytes(self.src.ptr, self.src.size)
        it.compressed_input.metadata = self.meta.handle
        st, self.plan = api.prepare_batch(ts.Operation.decode_roi, [it])
        assert st == ts.Status.success, st
        self.ws_bytes = self.plan.workspace_bytes()
        self.ws = api.DeviceMemory(self.ws_bytes + 512)
        self.res = api.DeviceMemory(64)
        self.res.fill(0x5A)

    def call(self, ws_ptr, ws_size, res_ptr, count, fn=None):
        fn = fn or N.lib().ts_decode_roi_async
        return ts.Status(fn(self.plan.handle, MutableBytes(ws_ptr, ws_size), res_ptr, count, None))


def test_execution_rejections_leave_buffers_untouched(orc):
    j = RoiJob(orc)
    host_ws = np.zeros(j.ws_bytes + 512, np.uint8)
    host_res = np.zeros(64, np.uint8)
    ws, res = j.ws.ptr, j.res.ptr
    cases = [
        ("workspace one byte short", (ws, j.ws_bytes - 1, res, 1), ts.Status.buffer_too_small),
        ("workspace off its 256-byte boundary", (ws + 16, j.ws_bytes, res, 1), ts.Status.invalid_argument),
        ("results misaligned", (ws, j.ws_bytes, res + 4, 1), ts.Status.invalid_argument),
        ("no result records", (ws, j.ws_bytes, res, 0), ts.Status.buffer_too_small),
        ("results overlap the compressed input", (ws, j.ws_bytes, j.src.ptr, 1), ts.Status.invalid_argument),
        ("results overlap the output", (ws, j.ws_bytes, j.out.mem.ptr, 1), ts.Status.invalid_argument),
        ("results overlap the workspace", (ws, j.ws_bytes, ws + 256, 1), ts.Status.invalid_argument),
        ("workspace overlaps the compressed input", (j.src.ptr, j.ws_bytes, res, 1), ts.Status.invalid_argument),
        ("null workspace of nonzero size", (None, j.ws_bytes, res, 1), ts.Status.invalid_argument),
        ("result count too large to size", (ws, j.ws_bytes, res, 1 << 62), ts.Status.size_overflow),
        ("results in plain host memory", (ws, j.ws_bytes, host_res.ctypes.data, 1), None),
        ("workspace in plain host memory", (host_ws.ctypes.data, j.ws_bytes, res, 1), None),
    ]
    out_before = j.out.mem.download().copy()
    for name, args, want in cases:
        st = j.call(*args)
        if want is None:
            # Pageable host memory is accepted only where the device reports
            # coherent pageable memory access.
            assert st in (ts.Status.invalid_argument, ts.Status.success), (name, st)
            if st == ts.Status.success:
                assert api.synchronize(None) == ts.Status.success
                j.res.fill(0x5A)
                j.out.mem.fill(0xA5)
                out_before = j.out.mem.download().copy()
                continue
        else:
            assert st == want, (name, st)
        assert (j.res.download() == 0x5A).all(), name
        assert np.array_equal(j.out.mem.download(), out_before), name
        assert N.lib().ts_hip_peek_last_error() == 0, name
    # The wrong operation for the plan.
    assert j.call(ws, j.ws_bytes, res, 1, N.lib().ts_decompress_async) == ts.Status.invalid_argument
    assert j.call(ws, j.ws_bytes, res, 1, N.lib().ts_decode_calibrate_async) == ts.Status.invalid_argument
    assert (j.res.download() == 0x5A).all()
    # And the valid call decodes the region.
    assert j.call(ws, j.ws_bytes, res, 1) == ts.Status.success
    assert api.synchronize(None) == ts.Status.success
    status, written = api.read_results(j.res, 1)
    assert status[0] == 0 and written[0] == 120 * 50 * 2
    assert np.array_equal(j.out.read(), j.px[10:60, 30:150])


def test_trusted_results_must_not_alias_the_producer():
    px = g.detector_u16(300, 100, 5)
    src = api.DeviceMemory.from_host(px)
    st, bound = api.compress_bound(api.host_view(px).image, TileShape(128, 32))
    archive = api.DeviceMemory(bound)
    enc = BatchItem()
    enc.input = api.device_view(src, ts.PixelType.uint16, 300, 100, const=True)
    enc.compressed_output = MutableBytes(archive.ptr, bound)
    st, eplan = api.prepare_batch(ts.Operation.compress, [enc], TileShape(128, 32))
    assert st == ts.Status.success
    produced = api.DeviceMemory(16)
    ews = api.DeviceMemory(eplan.workspace_bytes())
    assert api.compress_async(eplan, ews, produced, 1) == ts.Status.success
    out = DeviceOutput(ts.PixelType.uint16, 300, 100)
    it = BatchItem()
    it.output = out.view()
    it.roi = Roi(0, 0, 300, 100)
    ci = CompressedInput()
    ci.bytes = ConstBytes(archive.ptr, bound)
    ci.encoder_result = produced.ptr
    ci.trusted_image = api.host_view(px).image
    ci.trusted_tile = TileShape(128, 32)
    it.compressed_input = ci
    st, plan = api.prepare_batch(ts.Operation.decode_roi, [it])
    assert st == ts.Status.success, st
    ws = api.DeviceMemory(plan.workspace_bytes())
    # The producer's record is an input of the plan: it cannot take results.
    assert api.decode_roi_async(plan, ws, produced, 1) == ts.Status.invalid_argument
    res = api.DeviceMemory(16)
    assert api.decode_roi_async(plan, ws, res, 1) == ts.Status.success
    assert api.synchronize(None) == ts.Status.success
    assert np.array_equal(out.read(), px)


def test_device_buffer_ownership(orc):
    lib = N.lib()
    s = api.Stream()
    p = C.c_void_p(1)
    assert lib.ts_buffer_allocate(0, s.handle, C.byref(p)) == ts.Status.success and p.value is None
    assert lib.ts_buffer_allocate(1000, s.handle, C.byref(p)) == ts.Status.success
    first = p.value
    assert first is not None and first % 256 == 0
    q = C.c_void_p(0)
    assert lib.ts_buffer_allocate(2**64 - 1, s.handle, C.byref(q)) == ts.Status.size_overflow
    assert lib.ts_buffer_allocate(1 << 52, s.handle, C.byref(q)) == ts.Status.hip_error
    # A failed allocation is not reported again by the next execution.
    assert lib.ts_hip_peek_last_error() == 0
    assert lib.ts_buffer_release(first, s.handle) == ts.Status.success
    assert lib.ts_buffer_release(None, s.handle) == ts.Status.success
    s.synchronize()
    j = RoiJob(orc)
    assert j.call(j.ws.ptr, j.ws_bytes, j.res.ptr, 1) == ts.Status.success
    assert api.synchronize(None) == ts.Status.success


class Cycle:
    """Archives, outputs and a pinned result area set up beforehand, so that
    one cycle makes library calls only: prepare on the stream, allocate the
    workspace and results with the library's allocator, execute, copy the
    results back on the stream, wait for the stream, release, destroy."""

    def __init__(self, orc, images):
        self.items, self.keep = [], []
        for px in images:
            data = orc.compress(px, (64, 64))
            st, meta = api.parse_metadata(data)
            src = api.DeviceMemory.from_host(np.frombuffer(data, np.uint8))
            out = api.DeviceMemory(px.nbytes)
            it = BatchItem()
            it.output = api.device_view(out, api.pixel_type_of(px.dtype), px.shape[1], px.shape[0])
            it.compressed_input.bytes = ConstBytes(src.ptr, src.size)
            it.compressed_input.metadata = meta.handle
            self.items.append(it)
            self.keep += [meta, src, out]
        self.pinned = C.c_void_p(0)
        assert N.lib().ts_hip_host_malloc(16 * len(self.items), C.byref(self.pinned)) == 0

    def run(self, stream):
        lib, n = N.lib(), len(self.items)
        st, plan = api.prepare_batch(ts.Operation.decompress, self.items, stream=stream)
        assert st == ts.Status.success
        ws, res = C.c_void_p(0), C.c_void_p(0)
        assert lib.ts_buffer_allocate(plan.workspace_bytes(), stream.handle, C.byref(ws)) == 0
        assert lib.ts_buffer_allocate(16 * n, stream.handle, C.byref(res)) == 0
        st = ts.Status(lib.ts_decompress_async(plan.handle, MutableBytes(ws.value, plan.workspace_bytes()), res.value,
                                               n, stream.handle))
        assert st == ts.Status.success
        assert lib.ts_hip_memcpy_async(self.pinned, res.value, 16 * n, stream.handle) == 0
        stream.synchronize()
        assert lib.ts_buffer_release(ws.value, stream.handle) == 0
        assert lib.ts_buffer_release(res.value, stream.handle) == 0
        plan.close()

    def statuses(self):
        raw = (C.c_uint32 * (4 * len(self.items))).from_address(self.pinned.value)
        return [raw[4 * i] for i in range(len(self.items))]

    def close(self):
        N.lib().ts_hip_host_free(self.pinned)


def test_repeated_cycles_return_all_memory(orc):
    images = [g.detector_u16(300 + 37 * k, 200 + 11 * k, 600 + k) for k in range(5)]
    cycle = Cycle(orc, images)
    s = api.Stream()
    cycle.run(s)
    s.synchronize()
    api.device_synchronize()
    before = _pool_in_use()
    for _ in range(50):
        cycle.run(s)
        assert cycle.statuses() == [0] * len(images)
    s.synchronize()
    api.device_synchronize()
    assert _pool_in_use() == before
    cycle.close()


def _busy(stream, seconds_of_work=1.0):
    """Enqueues about `seconds_of_work` of streaming copies on `stream`."""
    nbytes = 1 << 30
    a, b = api.DeviceMemory(nbytes), api.DeviceMemory(nbytes)
    reps = max(1, int(seconds_of_work / (2 * nbytes / 5.0e12)))
    for _ in range(reps):
        assert N.lib().tsk_stream_copy(b.ptr, a.ptr, nbytes // 16, stream.handle) == 0
    return a, b


def test_no_hidden_device_synchronization(orc):
    images = [g.field_f32(400 + 50 * k, 300, 700 + k) for k in range(6)]
    cycle = Cycle(orc, images)
    worker, blocked = api.Stream(), api.Stream()
    cycle.run(worker)  # the first preparation in a process loads kernels
    keep = _busy(blocked)
    t0 = time.perf_counter()
    cycle.run(worker)
    elapsed = time.perf_counter() - t0
    assert HIP.hipStreamQuery(blocked.handle) == HIP_ERROR_NOT_READY, "the busy stream finished first"
    blocked.synchronize()
    assert elapsed < 5.0, elapsed
    assert cycle.statuses() == [0] * len(images)
    cycle.close()
    del keep


@pytest.mark.parametrize("op", ["decompress", "compress", "calibrate"])
def test_first_execution_does_not_wait_for_another_stream(orc, op):
    work, blocked = api.Stream(), api.Stream()
    px = g.detector_u16(700, 500, 31)
    keep = []
    if op == "decompress":
        data = orc.compress(px)
        st, meta = api.parse_metadata(data)
        src = api.DeviceMemory.from_host(np.frombuffer(data, np.uint8))
        out = api.DeviceMemory(px.nbytes)
        it = BatchItem()
        it.output = api.device_view(out, ts.PixelType.uint16, 700, 500)
        it.compressed_input.bytes = ConstBytes(src.ptr, src.size)
        it.compressed_input.metadata = meta.handle
        keep += [meta, src, out]
        st, plan = api.prepare_batch(ts.Operation.decompress, [it], stream=work)
        fn = api.decompress_async
    elif op == "compress":
        src = api.DeviceMemory.from_host(px)
        st, bound = api.compress_bound(api.host_view(px).image, TileShape(128, 32))
        out = api.DeviceMemory(bound)
        it = BatchItem()
        it.input = api.device_view(src, ts.PixelType.uint16, 700, 500, const=True)
        it.compressed_output = MutableBytes(out.ptr, bound)
        keep += [src, out]
        st, plan = api.prepare_batch(ts.Operation.compress, [it], TileShape(128, 32), stream=work)
        fn = api.compress_async
    else:
        src = api.DeviceMemory.from_host(px)
        out = api.DeviceMemory(px.size * 4)
        it = BatchItem()
        it.input = api.device_view(src, ts.PixelType.uint16, 700, 500, const=True)
        it.output = api.device_view(out, ts.PixelType.float32, 700, 500)
        it.calibration.offset = api.scalar(100.0)
        it.calibration.gain = api.scalar(1.05)
        keep += [src, out]
        st, plan = api.prepare_batch(ts.Operation.calibrate, [it], stream=work)
        fn = api.calibrate_async
    assert st == ts.Status.success, st
    ws = api.DeviceMemory(max(256, plan.workspace_bytes()))
    res = api.DeviceMemory(16)
    busy = _busy(blocked)
    t0 = time.perf_counter()
    assert fn(plan, ws if plan.workspace_bytes() else None, res, 1, work) == ts.Status.success
    work.synchronize()
    elapsed = time.perf_counter() - t0
    assert HIP.hipStreamQuery(blocked.handle) == HIP_ERROR_NOT_READY, "the busy stream finished first"
    blocked.synchronize()
    status, _ = api.read_results(res, 1)
    assert status[0] == 0 and elapsed < 5.0
    del busy


def _same_words(got_bits, want_bits):
    gf, wf = got_bits.view(np.float32), want_bits.view(np.float32)
    return bool(np.all((got_bits == want_bits) | (np.isnan(gf) & np.isnan(wf))))


@pytest.mark.parametrize("dtype", [np.uint16, np.uint32])
def test_config2_full_size_fused_maps(orc, dtype):
    side = 4096
    px = g.image_for(side, side, dtype, 2026517 + 11)
    data = orc.compress(px, (64, 64))
    dark = g.dark_u16(side, side, 2026517 + 12).astype(np.float32)
    gain = g.gain_f32(side, side, 2026517 + 13)
    st, meta = api.parse_metadata(data)
    src = api.DeviceMemory.from_host(np.frombuffer(data, np.uint8))
    om, gm = api.DeviceMemory.from_host(dark), api.DeviceMemory.from_host(gain)
    out = DeviceOutput(ts.PixelType.float32, side, side)
    it = BatchItem()
    it.output = out.view()
    it.roi = Roi(0, 0, side, side)
    it.compressed_input.bytes = ConstBytes(src.ptr, src.size)
    it.compressed_input.metadata = meta.handle
    it.calibration.offset = api.map_parameter(api.device_view(om, ts.PixelType.float32, side, side, const=True))
    it.calibration.gain = api.map_parameter(api.device_view(gm, ts.PixelType.float32, side, side, const=True))
    st, plan = api.prepare_batch(ts.Operation.decode_calibrate, [it])
    assert st == ts.Status.success
    ws = api.DeviceMemory(plan.workspace_bytes())
    res = api.DeviceMemory(16)
    assert api.decode_calibrate_async(plan, ws, res, 1) == ts.Status.success
    assert api.synchronize(None) == ts.Status.success
    status, written = api.read_results(res, 1)
    assert status[0] == 0 and written[0] == side * side * 4
    want = orc.decode_roi_calibrate(data, (0, 0, side, side), dark, gain).view(np.uint32)
    assert _same_words(out.read(), want)


def test_regions_of_a_full_size_archive(orc):
    side = 4096
    px = g.detector_u16(side, side, 2026517 + 40)
    data = orc.compress(px, (64, 64))
    st, meta = api.parse_metadata(data)
    src = api.DeviceMemory.from_host(np.frombuffer(data, np.uint8))
    rois = [r[1:] for r in g.roi_requests(60, 2026517 + 41, side, side)]
    items, outs = [], []
    for (x, y, w, h) in rois:
        o = DeviceOutput(ts.PixelType.float32, w, h)
        it = BatchItem()
        it.output = o.view()
        it.roi = Roi(x, y, w, h)
        it.compressed_input.bytes = ConstBytes(src.ptr, src.size)
        it.compressed_input.metadata = meta.handle
        it.calibration.offset = api.scalar(100.0)
        it.calibration.gain = api.scalar(1.05)
        items.append(it)
        outs.append(o)
    st, plan = api.prepare_batch(ts.Operation.decode_calibrate, items)
    assert st == ts.Status.success
    ws = api.DeviceMemory(plan.workspace_bytes())
    res = api.DeviceMemory(16 * len(items))
    assert api.decode_calibrate_async(plan, ws, res, len(items)) == ts.Status.success
    assert api.synchronize(None) == ts.Status.success
    status, _ = api.read_results(res, len(items))
    assert (status == 0).all()
    for roi, o in zip(rois, outs):
        want = orc.decode_roi_calibrate(data, roi, 100.0, 1.05).view(np.uint32)
        assert _same_words(o.read(), want), roi
