"""Calibration r = RN(RN(x - offset) * gain): standalone and fused with decode.

Tolerance: FP32 results equal the host binary32 computation bit for bit
(compiled with -ffp-contract=off), NaN compared by class; against the FP64
reference the error stays below 1.5 ulp for normal results and below 2^-149
absolute for subnormal ones (reference docs/numerical-analysis.md "Error bounds").
"""
import json
import os

import numpy as np
import pytest

import tilestream_amd as ts
from tilestream_amd import api, generators as g
from tilestream_amd._native import BatchItem

from helpers import DeviceOutput, Job

pytestmark = pytest.mark.gpu

SPECIAL = np.array([0x00000000, 0x80000000, 0x7F800000, 0xFF800000, 0x7FC00000, 0xFFC00001, 0x7F800001,
                    0x00000001, 0x807FFFFF, 0x00800000, 0x3F800000, 0xBF800000, 0x7F7FFFFF, 0xFF7FFFFF,
                    0x3F800001, 0x4B800000, 0x33800000, 0x00400000, 0x80000001, 0x42C80000, 0xC2C80000],
                   np.uint32)


def same_or_both_nan(a, b):
    fa, fb = a.view(np.float32), b.view(np.float32)
    both_nan = np.isnan(fa) & np.isnan(fb)
    return bool(np.all((a == b) | both_nan)), int(np.sum(~((a == b) | both_nan)))


HERE = os.path.dirname(os.path.abspath(__file__))


def contract_rows():
    """The reference's 42 calibration contract rows (tests/golden/calibration_contract.json):
    float32 sample words, offsets, gains and the exact expected words, as a 1 x 42 image."""
    rows = json.load(open(os.path.join(HERE, "golden", "calibration_contract.json")))["rows"]
    col = lambda k: np.array([r[k] for r in rows], np.uint32).reshape(1, len(rows))  # noqa: E731
    return col("x"), col("offset").view(np.float32), col("gain").view(np.float32), col("expected")


def calibrate_on_gpu(samples, offset, gain):
    h, w = samples.shape
    ptype = api.pixel_type_of(samples.dtype)
    src = api.DeviceMemory.from_host(samples)
    out = DeviceOutput(ts.PixelType.float32, w, h)
    job = Job()
    item = BatchItem()
    item.input = api.device_view(src, ptype, w, h, samples.strides[0], const=True)
    item.output = out.view()
    item.calibration.offset = job._param(offset, w, h, 0.0)
    item.calibration.gain = job._param(gain, w, h, 1.0)
    st, plan = api.prepare_batch(ts.Operation.calibrate, [item])
    assert st == ts.Status.success
    assert plan.workspace_bytes() == 0
    results = api.DeviceMemory(16)
    assert api.calibrate_async(plan, None, results, 1) == ts.Status.success
    assert api.synchronize() == ts.Status.success
    status, written = api.read_results(results, 1)
    assert status[0] == 0 and written[0] == w * h * 4
    return out.read()


def test_standalone_matches_host_binary32(orc):
    px = g.detector_u16(333, 97, 3)
    dark = g.dark_u16(333, 97, 4).astype(np.float32)
    gain = g.gain_f32(333, 97, 5)
    got = calibrate_on_gpu(px, dark, gain)
    want = orc.calibrate(px, dark, gain).view(np.uint32)
    assert np.array_equal(got, want)
    f = g.field_f32(333, 97, 6)
    got = calibrate_on_gpu(f, 100.0, 1.05)
    ok, bad = same_or_both_nan(got, orc.calibrate(f, 100.0, 1.05).view(np.uint32))
    assert ok, bad


@pytest.mark.parametrize("w", [1031, 4099, 9001])
def test_standalone_wide_rows(orc, w):
    """Rows of more than 1024 pixels: a lane loads up to four groups of a row
    (1024 pixels apart) before it stores them, and rows wider than 4096
    pixels take more than one such pass, then the last w % 4 pixels one per
    lane.  Odd widths put every other uint16 row off a 4-byte boundary."""
    h = 3
    px = g.detector_u16(w, h, 21)
    dark = g.dark_u16(w, h, 22).astype(np.float32)
    gain = g.gain_f32(w, h, 23)
    assert np.array_equal(calibrate_on_gpu(px, dark, gain), orc.calibrate(px, dark, gain).view(np.uint32))
    assert np.array_equal(calibrate_on_gpu(px, dark, 1.25), orc.calibrate(px, dark, 1.25).view(np.uint32))
    f = g.field_f32(w, h, 24)
    ok, bad = same_or_both_nan(calibrate_on_gpu(f, 100.0, gain), orc.calibrate(f, 100.0, gain).view(np.uint32))
    assert ok, bad
    ok, bad = same_or_both_nan(calibrate_on_gpu(f, 100.0, 1.05), orc.calibrate(f, 100.0, 1.05).view(np.uint32))
    assert ok, bad


def test_every_uint16_value():
    px = np.arange(65536, dtype=np.uint32).astype(np.uint16).reshape(256, 256)
    got = calibrate_on_gpu(px, 100.25, 1.5).view(np.float32)
    want = ((px.astype(np.float32) - np.float32(100.25)).astype(np.float32) * np.float32(1.5)).astype(np.float32)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_special_word_grid(orc):
    """21 x 21 x 21 grid of special sample, offset and gain words."""
    n = len(SPECIAL)
    x = np.repeat(SPECIAL, n * n).reshape(n, n * n)
    o = np.tile(np.repeat(SPECIAL, n), n).reshape(n, n * n).view(np.float32)
    gn = np.tile(SPECIAL, n * n).reshape(n, n * n).view(np.float32)
    got = calibrate_on_gpu(x, o, gn)
    ok, bad = same_or_both_nan(got, orc.calibrate(x, o, gn).view(np.uint32))
    assert ok, bad


def _fp64_errors(samples, offset, gain, result, u16):
    x = samples.astype(np.float64) if u16 else samples.view(np.float32).astype(np.float64)
    ref = (x - np.broadcast_to(np.float64(offset), x.shape)) * np.broadcast_to(np.float64(gain), x.shape)
    r = result.view(np.float32).astype(np.float64)
    finite = np.isfinite(ref) & np.isfinite(r) & (np.abs(ref) < 2.0 ** 128 - 2.0 ** 103)
    e = np.frexp(np.where(ref == 0, 1.0, np.abs(ref)))[1] - 1
    ulp = np.where(np.abs(ref) >= 2.0 ** -126, np.ldexp(1.0, e - 23), 2.0 ** -149)
    err = np.abs(r - ref)[finite] / ulp[finite]
    normal = (np.abs(ref) >= 2.0 ** -126)[finite]
    return err[normal], np.abs(r - ref)[finite][~normal]


@pytest.mark.parametrize("kind", ["u16", "f32"])
def test_fused_equals_separate_and_bound(orc, kind):
    w, h = 700, 389
    px = g.detector_u16(w, h, 31) if kind == "u16" else g.field_f32(w, h, 31)
    data = orc.compress(px, (64, 64))
    roi = (37, 21, 500, 300)
    x0, y0, rw, rh = roi
    dark = g.dark_u16(rw, rh, 32).astype(np.float32)
    gain = g.gain_f32(rw, rh, 33)
    job = Job()
    job.add_decode(data, roi=roi, calibrated=True, offset=dark, gain=gain)
    status, written, outs = job.run(ts.Operation.decode_calibrate)
    assert status[0] == 0 and written[0] == rw * rh * 4
    fused = outs[0]
    crop = np.ascontiguousarray(px[y0:y0 + rh, x0:x0 + rw])
    separate = calibrate_on_gpu(crop, dark, gain)
    assert np.array_equal(fused, separate)  # same device operations, identical words, NaNs included
    ok, bad = same_or_both_nan(fused, orc.decode_roi_calibrate(data, roi, dark, gain).view(np.uint32))
    assert ok, bad
    normal, sub = _fp64_errors(crop, dark, gain, fused, kind == "u16")
    assert normal.max() < 1.5 + 2.0 ** -20
    assert sub.size == 0 or sub.max() < 2.0 ** -149 * (1 + 2.0 ** -20)


@pytest.mark.parametrize("tile", [(128, 32), (64, 64), (256, 16)])
def test_fused_scalar_region(orc, tile):
    px = g.detector_u16(1000, 600, 51)
    data = orc.compress(px, tile)
    rng = np.random.default_rng(52)
    for _ in range(4):
        rw, rh = int(rng.integers(1, 1000)), int(rng.integers(1, 600))
        x, y = int(rng.integers(0, 1000 - rw + 1)), int(rng.integers(0, 600 - rh + 1))
        job = Job()
        job.add_decode(data, roi=(x, y, rw, rh), calibrated=True, offset=100.0, gain=1.05)
        status, _, outs = job.run(ts.Operation.decode_calibrate)
        assert status[0] == 0
        want = orc.decode_roi_calibrate(data, (x, y, rw, rh), 100.0, 1.05).view(np.uint32)
        assert np.array_equal(outs[0], want)


def test_contract_rows_standalone_with_maps():
    x, o, gn, want = contract_rows()
    ok, bad = same_or_both_nan(calibrate_on_gpu(x, o, gn), want)
    assert ok, bad


def test_contract_rows_standalone_with_scalars():
    x, o, gn, want = contract_rows()
    for i in range(x.shape[1]):
        got = calibrate_on_gpu(np.ascontiguousarray(x[:, i:i + 1]), float(o[0, i]), float(gn[0, i]))
        ok, _ = same_or_both_nan(got, want[:, i:i + 1])
        assert ok, (i, hex(int(got[0, 0])), hex(int(want[0, i])))


def test_contract_rows_fused_with_decode(orc):
    """The same rows as float32 samples of an archive, decoded fused with maps."""
    x, o, gn, want = contract_rows()
    data = orc.compress(x, (64, 64))
    job = Job()
    job.add_decode(data, calibrated=True, offset=o, gain=gn)
    status, written, outs = job.run(ts.Operation.decode_calibrate)
    assert status[0] == 0 and written[0] == x.size * 4
    ok, bad = same_or_both_nan(outs[0], want)
    assert ok, bad


@pytest.mark.parametrize("kind", ["u16", "f32"])
def test_tiff_file_to_calibrated_output(orc, tmp_path, kind):
    """The file tier's TIFF path: pixels exported as a Deflate TIFF, imported
    on the host and calibrated on the GPU equal the host binary32 calibration
    of the original pixels word for word (NaN by class)."""
    px = g.detector_u16(300, 211, 41) if kind == "u16" else g.field_f32(300, 211, 41)
    path = str(tmp_path / "f.tif")
    opts = api.tiff_write_options(compression=1, horizontal_predictor=1 if kind == "u16" else 0)
    assert api.write_tiff(path, px, opts) == ts.Status.success
    st, back = api.read_tiff(path)
    assert st == ts.Status.success and np.array_equal(back, px)
    got = calibrate_on_gpu(back, 100.0, 1.25)
    want = orc.calibrate(px, 100.0, 1.25)
    ok, bad = same_or_both_nan(got.reshape(-1), np.asarray(want).view(np.uint32).reshape(-1))
    assert ok, bad


@pytest.mark.parametrize("kind", ["u16", "f32"])
def test_large_fused_plan_striding_launch(orc, kind):
    """Fused plans of 1536 tiles or more launch 1024 workgroups that stride over
    the tiles; their words equal the one-workgroup-per-tile launch's and the
    oracle's (NaNs by class)."""
    w, h = 2600, 2400  # 41 x 38 = 1558 tiles of 64x64, ragged last column and row
    px = g.detector_u16(w, h, 34) if kind == "u16" else g.field_f32(w, h, 34)
    data = orc.compress(px, (64, 64))
    dark = g.dark_u16(w, h, 35).astype(np.float32)
    gain = g.gain_f32(w, h, 36)
    outs = []
    for shape in (None, (8, 256, 0, 1)):
        job = Job()
        job.add_decode(data, calibrated=True, offset=dark, gain=gain)
        status, written, o = job.run(ts.Operation.decode_calibrate, shape)
        assert status[0] == 0 and written[0] == w * h * 4
        outs.append(np.ascontiguousarray(o[0]).view(np.uint32))
    assert np.array_equal(outs[0], outs[1])
    ok, bad = same_or_both_nan(outs[0], orc.decode_roi_calibrate(data, (0, 0, w, h), dark, gain).view(np.uint32))
    assert ok, bad
