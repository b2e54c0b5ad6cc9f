"""Seeded robustness sweep on the GPU (reference tests/cuda_robustness_test.cpp:89-306).

* Every case of seeds 13,000,000 ... 13,009,999 (tests/robustness_cases.py), in
  batches of 240 images of one pixel type and all three tile shapes: the GPU
  encoder's archives equal the oracle's byte for byte; the GPU decode of the
  oracle's archives equals the pixels; the region decode of the GPU archives
  on the trusted path (no host metadata) equals the oracle's crop; the fused
  region decode with calibration equals the oracle's host binary32 result
  word for word (NaNs by class).
* Fabricated metadata (one field of a parse changed, 3,000 trials over both
  pixel types and all tile shapes): preparation rejects it with a format
  status, or the decode it plans runs to completion."""
import numpy as np
import pytest

import tilestream_amd as ts
from tilestream_amd import api
from tilestream_amd._native import BatchItem, CompressedInput, ConstBytes, Roi, TileShape

from helpers import DeviceOutput
import robustness_cases as R

pytestmark = pytest.mark.gpu

BATCH = 240


def _device_input(px, pad):
    h, w = px.shape
    host = np.zeros((h, w + pad), px.dtype)
    host[:, :w] = px
    mem = api.DeviceMemory.from_host(host)
    view = api.device_view(mem, api.pixel_type_of(px.dtype), w, h, (w + pad) * px.itemsize, const=True)
    return mem, view


def _same_words(got_bits, want_bits):
    gf, wf = got_bits.view(np.float32), want_bits.view(np.float32)
    return bool(np.all((got_bits == want_bits) | (np.isnan(gf) & np.isnan(wf))))


def _run(op, items, fn):
    st, plan = api.prepare_batch(op, items)
    assert st == ts.Status.success, st
    ws = api.DeviceMemory(plan.workspace_bytes()) if plan.workspace_bytes() else None
    res = api.DeviceMemory(16 * len(items))
    res.fill(0x5A)
    assert fn(plan, ws, res, len(items), None) == ts.Status.success
    assert api.synchronize(None) == ts.Status.success
    return api.read_results(res, len(items))


def _check_batch(orc, cases, failures):
    files = [orc.compress(c.pixels, c.tile) for c in cases]
    keep = []
    # GPU encode, one plan per tile shape (inputs with padded rows).
    gpu = {}
    for shape in R.SHAPES:
        idx = [i for i, c in enumerate(cases) if c.tile == shape]
        if not idx:
            continue
        items, outs = [], []
        for i in idx:
            mem, view = _device_input(cases[i].pixels, cases[i].pad)
            st, bound = api.compress_bound(view.image, TileShape(*shape))
            out = api.DeviceMemory(bound)
            it = BatchItem()
            it.input = view
            it.compressed_output = api.MutableBytes(out.ptr, bound)
            items.append(it)
            outs.append((out, bound))
            keep.append(mem)
        st, plan = api.prepare_batch(ts.Operation.compress, items, TileShape(*shape))
        assert st == ts.Status.success, st
        ws = api.DeviceMemory(plan.workspace_bytes())
        res = api.DeviceMemory(16 * len(items))
        assert api.compress_async(plan, ws, res, len(items)) == ts.Status.success
        assert api.synchronize(None) == ts.Status.success
        status, written = api.read_results(res, len(items))
        for k, i in enumerate(idx):
            gpu[i] = (outs[k][0], outs[k][1], res, k, int(status[k]), int(written[k]))
        keep += [plan, ws, res]
    for i, c in enumerate(cases):
        out, bound, res, k, st, n = gpu[i]
        if st != 0 or out.download(n).tobytes() != files[i]:
            failures.append((c.seed, "GPU archive differs from the oracle's"))

    # Whole decode of the oracle's archives.
    items, outs = [], []
    for c, f in zip(cases, files):
        st, meta = api.parse_metadata(f)
        assert st == ts.Status.success
        src = api.DeviceMemory.from_host(np.frombuffer(f, np.uint8))
        h, w = c.pixels.shape
        o = DeviceOutput(api.pixel_type_of(c.dtype), w, h)
        it = BatchItem()
        it.output = o.view()
        it.compressed_input.bytes = ConstBytes(src.ptr, src.size)
        it.compressed_input.metadata = meta.handle
        items.append(it)
        outs.append(o)
        keep += [meta, src]
    status, _ = _run(ts.Operation.decompress, items, api.decompress_async)
    for c, o, st in zip(cases, outs, status):
        if st != 0 or not np.array_equal(o.read(), c.pixels):
            failures.append((c.seed, "GPU decode of the oracle's archive differs"))

    # Region decode of the GPU archives on the trusted path.
    items, outs = [], []
    for i, c in enumerate(cases):
        out, bound, res, k, _, _ = gpu[i]
        x, y, w, h = c.roi
        o = DeviceOutput(api.pixel_type_of(c.dtype), w, h)
        it = BatchItem()
        it.output = o.view()
        it.roi = Roi(x, y, w, h)
        ci = CompressedInput()
        ci.bytes = ConstBytes(out.ptr, bound)
        ci.encoder_result = res.ptr + 16 * k
        ci.trusted_image = api.host_view(c.pixels).image
        ci.trusted_tile = TileShape(*c.tile)
        it.compressed_input = ci
        items.append(it)
        outs.append(o)
    status, _ = _run(ts.Operation.decode_roi, items, api.decode_roi_async)
    for c, o, st in zip(cases, outs, status):
        x, y, w, h = c.roi
        if st != 0 or not np.array_equal(o.read(), c.pixels[y:y + h, x:x + w]):
            failures.append((c.seed, "trusted region decode of the GPU archive differs"))

    # Fused region decode with calibration (scalars; maps for every third case).
    items, outs, params = [], [], []
    for i, (c, f) in enumerate(zip(cases, files)):
        x, y, w, h = c.roi
        st, meta = api.parse_metadata(f)
        src = api.DeviceMemory.from_host(np.frombuffer(f, np.uint8))
        o = DeviceOutput(ts.PixelType.float32, w, h)
        it = BatchItem()
        it.output = o.view()
        it.roi = Roi(x, y, w, h)
        it.compressed_input.bytes = ConstBytes(src.ptr, src.size)
        it.compressed_input.metadata = meta.handle
        rng = np.random.default_rng(c.seed)
        if i % 3 == 0:
            off = rng.uniform(90, 110, (h, w)).astype(np.float32)
            gain = rng.normal(1, 0.02, (h, w)).astype(np.float32)
            om, gm = api.DeviceMemory.from_host(off), api.DeviceMemory.from_host(gain)
            it.calibration.offset = api.map_parameter(api.device_view(om, ts.PixelType.float32, w, h, const=True))
            it.calibration.gain = api.map_parameter(api.device_view(gm, ts.PixelType.float32, w, h, const=True))
            keep += [om, gm]
        else:
            off, gain = float(rng.uniform(-50, 150)), float(rng.uniform(0.5, 2.0))
            it.calibration.offset = api.scalar(off)
            it.calibration.gain = api.scalar(gain)
        items.append(it)
        outs.append(o)
        params.append((off, gain))
        keep += [meta, src]
    status, _ = _run(ts.Operation.decode_calibrate, items, api.decode_calibrate_async)
    for c, f, o, st, (off, gain) in zip(cases, files, outs, status, params):
        want = orc.decode_roi_calibrate(f, c.roi, off, gain).view(np.uint32)
        if st != 0 or not _same_words(o.read(), want):
            failures.append((c.seed, "fused region decode differs from host binary32"))


def test_random_round_trips_on_every_path(orc):
    seeds = R.selected_seeds()
    failures = []
    for dtype in (np.uint16, np.uint32):
        batch = []
        for seed in seeds:
            c = R.make_case(seed)
            if c.dtype == dtype:
                batch.append(c)
            if len(batch) == BATCH or (seed == seeds[-1] and batch):
                _check_batch(orc, batch, failures)
                batch = []
    assert not failures, failures[:20]


def _mixed_image(dtype, rng):
    """300 x 192 with a constant, a smooth and a noisy band of 64 rows, so
    every tile shape gives all three tile modes."""
    img = np.empty((192, 300), dtype)
    img[0:64] = 7
    y, x = np.mgrid[0:64, 0:300]
    smooth = (1000 + 3 * x + 5 * y).astype(np.float64)
    img[64:128] = smooth.astype(dtype) if dtype == np.uint16 else smooth.astype(np.float32).view(np.uint32)
    mask = 0xFFFF if dtype == np.uint16 else 0xFFFFFFFF
    img[128:192] = rng.integers(0, mask + 1, (64, 300), dtype=np.uint64).astype(dtype)
    return img


def _nudge(value, rng, step=1):
    if rng.integers(0, 3) == 0:
        return int(rng.integers(0, 1 << 63)) if rng.integers(0, 3) == 0 else int(rng.integers(0, 65536))
    return max(0, int(value) + int(rng.integers(-8, 9)) * step)


def _fabricate(meta, rng):
    """Changes one field of parsed metadata (the mutation classes of reference
    tests/cuda_robustness_test.cpp:196-240)."""
    summ = meta.summary
    ti = int(rng.integers(0, summ.tile_count))
    si = int(rng.integers(0, summ.segment_count)) if summ.segment_count else None
    t = meta.tile(ti)
    kind = int(rng.integers(0, 12))
    u16 = 0xFFFF
    if kind == 0:
        t.mode = int(rng.integers(0, 4))
    elif kind == 1:
        t.width = _nudge(t.width, rng) & u16
    elif kind == 2:
        t.height = _nudge(t.height, rng) & u16
    elif kind == 3:
        t.offset = _nudge(t.offset, rng, 4)
    elif kind == 4:
        t.length = _nudge(t.length, rng, 4)
    elif kind == 5:
        t.first_segment = _nudge(t.first_segment, rng)
    elif kind in (6, 7, 8, 9) and si is not None:
        s = meta.segment(si)
        if kind == 6:
            s.seed = _nudge(s.seed, rng) & 0xFFFFFFFF
        elif kind == 7:
            s.bit_width = int(rng.integers(0, 40))
        elif kind == 8:
            s.payload_offset = _nudge(s.payload_offset, rng, 4)
        else:
            s.payload_bytes = _nudge(s.payload_bytes, rng, 4)
        meta.set_segment(si, s)
        return
    elif kind == 10:
        meta.set_file_bytes(_nudge(summ.file_bytes, rng, 4))
        return
    elif kind == 11:
        if rng.integers(0, 2) == 0:
            meta.truncate(summ.tile_count - 1, summ.segment_count)
        elif summ.segment_count:
            meta.truncate(summ.tile_count, summ.segment_count - 1)
        return
    meta.set_tile(ti, t)


def test_fabricated_metadata_is_rejected_or_stays_in_bounds(orc):
    rng = np.random.default_rng(4242)
    statuses = {}
    accepted = 0
    for dtype in (np.uint16, np.uint32):
        image = _mixed_image(dtype, rng)
        for shape in R.SHAPES:
            data = orc.compress(image, shape)
            st, meta = api.parse_metadata(data)
            modes = {meta.tile(i).mode for i in range(meta.summary.tile_count)}
            assert modes == {0, 1, 2}, modes
            src = api.DeviceMemory.from_host(np.frombuffer(data, np.uint8))
            out = api.DeviceMemory(300 * 192 * 4)
            res = api.DeviceMemory(16)
            for _ in range(500):
                st, meta = api.parse_metadata(data)
                _fabricate(meta, rng)
                x, y = int(rng.integers(0, 300)), int(rng.integers(0, 192))
                w, h = int(rng.integers(1, 301 - x)), int(rng.integers(1, 193 - y))
                it = BatchItem()
                it.output = api.device_view(out, api.pixel_type_of(dtype), w, h, w * image.itemsize)
                it.roi = Roi(x, y, w, h)
                it.compressed_input.bytes = ConstBytes(src.ptr, src.size)
                it.compressed_input.metadata = meta.handle
                st, plan = api.prepare_batch(ts.Operation.decode_roi, [it])
                statuses[int(st)] = statuses.get(int(st), 0) + 1
                if st != ts.Status.success:
                    assert st in (ts.Status.malformed_data, ts.Status.unsupported_format,
                                  ts.Status.buffer_too_small), st
                    continue
                accepted += 1
                ws = api.DeviceMemory(plan.workspace_bytes()) if plan.workspace_bytes() else None
                assert api.decode_roi_async(plan, ws, res, 1) == ts.Status.success
                assert api.synchronize(None) == ts.Status.success
    assert accepted > 0 and statuses.get(int(ts.Status.malformed_data), 0) > 0, statuses
