"""Serving resident archives: region requests decoded in request mode (tiles
derived on the device from each request's region) and encode requests in the
same batches, continuous batching over several streams, each result identical
alone and in the mix, request checks."""
import numpy as np
import pytest

import tilestream_amd as ts
from tilestream_amd import api, generators as g

pytestmark = pytest.mark.gpu

ARCHIVES = [("u16", 700, 389, (64, 64)), ("f32", 517, 301, (128, 32)), ("u16", 1000, 600, (256, 16)),
            ("f32", 300, 200, (64, 64)), ("u16", 129, 33, (128, 32))]
OFFSET, GAIN = 100.0, 1.05


class Resident:
    def __init__(self, orc):
        self.pixels, self.files, self.mems = [], [], []
        self.set = api.ArchiveSet()
        for i, (kind, w, h, tile) in enumerate(ARCHIVES):
            px = g.detector_u16(w, h, 300 + i) if kind == "u16" else g.field_f32(w, h, 300 + i)
            data = orc.compress(px, tile)
            st, meta = api.parse_metadata(data)
            assert st == ts.Status.success
            mem = api.DeviceMemory.from_host(np.frombuffer(data, np.uint8))
            st, ident = self.set.add(mem, meta)
            assert st == ts.Status.success and ident == i
            self.pixels.append(px)
            self.files.append(data)
            self.mems.append(mem)
        assert self.set.commit() == ts.Status.success


@pytest.fixture(scope="module")
def resident(orc):
    return Resident(orc)


def _requests(resident, rng, count, big_every=7):
    """Random regions: mostly small (one or a few tiles), some large."""
    specs = []
    for k in range(count):
        a = int(rng.integers(0, len(ARCHIVES)))
        h, w = resident.pixels[a].shape
        if k % big_every == 0:
            rw, rh = int(rng.integers(w // 2, w + 1)), int(rng.integers(h // 2, h + 1))
        else:
            rw, rh = int(rng.integers(1, min(w, 90) + 1)), int(rng.integers(1, min(h, 70) + 1))
        x, y = int(rng.integers(0, w - rw + 1)), int(rng.integers(0, h - rh + 1))
        specs.append((a, (x, y, rw, rh), bool(rng.integers(0, 2))))
    return specs


class Outputs:
    """Disjoint device outputs for a list of request specs, in one allocation."""

    def __init__(self, resident, specs):
        self.offsets, total = [], 0
        for a, (x, y, w, h), cal in specs:
            s = 4 if cal or resident.pixels[a].dtype != np.uint16 else 2
            self.offsets.append((total, s))
            total += (w * h * s + 255) // 256 * 256
        self.mem = api.DeviceMemory(max(256, total))
        self.mem.fill(0xA5)
        self.requests = []
        for (a, roi, cal), (off, s) in zip(specs, self.offsets):
            ptype = ts.PixelType.float32 if s == 4 else ts.PixelType.uint16
            view = api.device_view(self.mem, ptype, roi[2], roi[3], offset=off, capacity=roi[2] * roi[3] * s)
            self.requests.append(api.region_request(a, roi, view, cal, OFFSET, GAIN))

    def read(self, specs, i):
        off, s = self.offsets[i]
        w, h = specs[i][1][2], specs[i][1][3]
        raw = self.mem.download(w * h * s, off)
        return raw.view(np.uint16 if s == 2 else np.uint32).reshape(h, w)


def _expected(orc, resident, spec):
    a, roi, cal = spec
    if cal:
        return orc.decode_roi_calibrate(resident.files[a], roi, OFFSET, GAIN).view(np.uint32)
    x, y, w, h = roi
    px = resident.pixels[a][y:y + h, x:x + w]
    return px if px.dtype == np.uint16 else px.view(np.uint32)


def _same(got, want):
    if got.dtype == np.uint32:
        fg, fw = got.view(np.float32), want.view(np.float32)
        return bool(np.all((got == want) | (np.isnan(fg) & np.isnan(fw))))
    return bool(np.array_equal(got, want))


@pytest.mark.parametrize("shape", [(8, 256, 0, 4, 2), (4, 128, 0, 0, 1), (16, 512, 0, 2, 2), (8, 256, 3, 1, 1),
                                   (4, 256, 0, 4, 3), (8, 512, 0, 1, 3)])
def test_batch_matches_oracle(orc, resident, shape):
    rng = np.random.default_rng(500 + shape[0] + shape[1])
    specs = _requests(resident, rng, 90)
    out = Outputs(resident, specs)
    server = api.Server(resident.set, api.server_options(streams=3, max_batch=16, shape=shape))
    st, statuses, _ = server.run_batch(out.requests)
    assert st == ts.Status.success
    assert (statuses == 0).all()
    for i, spec in enumerate(specs):
        assert _same(out.read(specs, i), _expected(orc, resident, spec)), (shape, i, spec)


@pytest.mark.parametrize("max_batch", [16, 1024])
def test_block_batches_match_oracle(orc, resident, max_batch):
    """Every region batch decoded by blocks (block records built on the device
    from the requests): both pixel types, all three tile shapes, plain and
    calibrated, regions cutting blocks on every side."""
    rng = np.random.default_rng(610 + max_batch)
    specs = _requests(resident, rng, 90)
    out = Outputs(resident, specs)
    server = api.Server(resident.set, api.server_options(streams=3, max_batch=max_batch, block_tasks=1))
    st, statuses, _ = server.run_batch(out.requests)
    assert st == ts.Status.success
    assert (statuses == 0).all()
    for i, spec in enumerate(specs):
        assert _same(out.read(specs, i), _expected(orc, resident, spec)), (max_batch, i, spec)


def test_blocks_and_tiles_give_the_same_words(resident):
    """The same batch decoded by blocks and tile by tile, replayed with Poisson
    arrivals: identical outputs word for word."""
    rng = np.random.default_rng(613)
    specs = _requests(resident, rng, 120, big_every=3)
    by_tiles, by_blocks = Outputs(resident, specs), Outputs(resident, specs)
    arrivals = np.cumsum(rng.exponential(2e4, len(specs))).astype(np.uint64)
    for out, block_tasks in ((by_tiles, 0), (by_blocks, 1)):
        server = api.Server(resident.set, api.server_options(streams=3, max_batch=32, block_tasks=block_tasks))
        st, _, statuses, _, _ = server.replay(out.requests, arrivals)
        assert st == ts.Status.success and (statuses == 0).all()
    for i in range(len(specs)):
        assert np.array_equal(by_tiles.read(specs, i), by_blocks.read(specs, i)), (i, specs[i])


def test_many_single_tile_requests_in_one_batch(orc, resident):
    """Hundreds of one-tile requests share each sampled request-index stride."""
    rng = np.random.default_rng(77)
    specs = []
    for k in range(700):
        a = k % len(ARCHIVES)
        h, w = resident.pixels[a].shape
        tw, th = ARCHIVES[a][3]
        cx, cy = int(rng.integers(0, (w + tw - 1) // tw)), int(rng.integers(0, (h + th - 1) // th))
        x0, y0 = cx * tw, cy * th
        rw, rh = min(tw, w - x0), min(th, h - y0)
        specs.append((a, (x0, y0, rw, rh), k % 3 == 0))
    specs.append((2, (0, 0, 1000, 600), False))  # 4 x 38 = 152 tiles after them
    out = Outputs(resident, specs)
    server = api.Server(resident.set, api.server_options(streams=2, max_batch=1024))
    st, statuses, _ = server.run_batch(out.requests)
    assert st == ts.Status.success and (statuses == 0).all()
    for i, spec in enumerate(specs):
        assert _same(out.read(specs, i), _expected(orc, resident, spec)), (i, spec)


def test_alone_and_in_the_mix_are_identical(resident):
    rng = np.random.default_rng(91)
    specs = _requests(resident, rng, 60)
    mix = Outputs(resident, specs)
    server = api.Server(resident.set, api.server_options(streams=3, max_batch=8))
    st, statuses, _ = server.run_batch(mix.requests)
    assert st == ts.Status.success and (statuses == 0).all()
    for i in range(0, len(specs), 4):
        alone = Outputs(resident, [specs[i]])
        st, statuses, _ = server.run_batch(alone.requests)
        assert st == ts.Status.success and statuses[0] == 0
        assert np.array_equal(alone.read([specs[i]], 0), mix.read(specs, i)), (i, specs[i])


@pytest.mark.parametrize("large", [None, 64 << 10])
def test_replay_with_poisson_arrivals(orc, resident, large):
    """At the default size-class bound every region is small; at 64 KiB the
    larger ones batch apart on at most two of the three streams."""
    rng = np.random.default_rng(12)
    specs = _requests(resident, rng, 200)
    out = Outputs(resident, specs)
    arrivals = np.cumsum(rng.exponential(20_000.0, len(specs))).astype(np.uint64)  # 50k requests/s
    server = api.Server(resident.set, api.server_options(streams=3, max_batch=32,
                                                         admission_bytes=4 << 20, large_bytes=large))
    st, latency, statuses, _, stats = server.replay(out.requests, arrivals)
    assert st == ts.Status.success
    assert (statuses == 0).all() and stats.failed == 0
    assert stats.requests == len(specs) and 1 <= stats.batches <= len(specs)
    assert (latency > 0).all() and stats.seconds > 0
    want_bytes = sum(r[1][2] * r[1][3] * resident.pixels[r[0]].itemsize for r in specs)
    assert stats.decoded_bytes == want_bytes
    for i, spec in enumerate(specs):
        assert _same(out.read(specs, i), _expected(orc, resident, spec)), (i, spec)


def test_admission_smaller_than_one_request_still_progresses(orc, resident):
    rng = np.random.default_rng(13)
    specs = _requests(resident, rng, 20, big_every=2)
    out = Outputs(resident, specs)
    server = api.Server(resident.set, api.server_options(streams=2, admission_bytes=1024))
    st, latency, statuses, _, stats = server.replay(out.requests, np.zeros(len(specs), np.uint64))
    assert st == ts.Status.success and (statuses == 0).all()
    assert stats.batches >= len(specs) // 2  # the large regions go one at a time
    assert stats.admission_holds > 0
    # Requests above the cap run alone: the peak is one request's output.
    assert stats.peak_in_flight <= max(_output_bytes(resident, specs))
    for i, spec in enumerate(specs):
        assert _same(out.read(specs, i), _expected(orc, resident, spec))


def _output_bytes(resident, specs):
    return [w * h * (4 if cal or resident.pixels[a].dtype != np.uint16 else 2) for a, (x, y, w, h), cal in specs]


@pytest.mark.parametrize("streams", [1, 3])
def test_admission_cap_binds_and_bounds_bytes_in_flight(orc, resident, streams):
    """A cap well below the outputs of the whole replay but above any single
    request: batches close early at the cap, the output bytes in flight never
    exceed it, and every result still equals the oracle."""
    rng = np.random.default_rng(29 + streams)
    specs = _requests(resident, rng, 80, big_every=3)
    sizes = _output_bytes(resident, specs)
    cap = max(max(sizes), sum(sizes) // 6)
    out = Outputs(resident, specs)
    server = api.Server(resident.set, api.server_options(streams=streams, admission_bytes=cap))
    st, latency, statuses, _, stats = server.replay(out.requests, np.zeros(len(specs), np.uint64))
    assert st == ts.Status.success and (statuses == 0).all()
    assert stats.admission_holds > 0
    assert 0 < stats.peak_in_flight <= cap
    for i, spec in enumerate(specs):
        assert _same(out.read(specs, i), _expected(orc, resident, spec)), (streams, i, spec)
    # Without a binding cap the same replay never holds.
    free = api.Server(resident.set, api.server_options(streams=streams, admission_bytes=sum(sizes) + 1))
    st, _, statuses, _, stats = free.replay(out.requests, np.zeros(len(specs), np.uint64))
    assert st == ts.Status.success and (statuses == 0).all() and stats.admission_holds == 0


def test_request_checks(resident):
    server = api.Server(resident.set, api.server_options(max_tasks=64))
    mem = api.DeviceMemory(2 << 20)
    ok = api.region_request(0, (0, 0, 64, 64), api.device_view(mem, ts.PixelType.uint16, 64, 64))
    assert server.check(ok) == ts.Status.success
    bad_id = api.region_request(len(ARCHIVES), (0, 0, 64, 64), api.device_view(mem, ts.PixelType.uint16, 64, 64))
    assert server.check(bad_id) == ts.Status.invalid_argument
    out_of_image = api.region_request(0, (650, 0, 64, 64), api.device_view(mem, ts.PixelType.uint16, 64, 64))
    assert server.check(out_of_image) == ts.Status.invalid_roi
    empty = api.region_request(0, (0, 0, 0, 64), api.device_view(mem, ts.PixelType.uint16, 1, 64))
    assert server.check(empty) == ts.Status.invalid_roi
    wrong_type = api.region_request(0, (0, 0, 64, 64), api.device_view(mem, ts.PixelType.float32, 64, 64))
    assert server.check(wrong_type) == ts.Status.invalid_argument
    cal_u16 = api.region_request(0, (0, 0, 64, 64), api.device_view(mem, ts.PixelType.uint16, 64, 64), True)
    assert server.check(cal_u16) == ts.Status.invalid_argument
    wrong_size = api.region_request(0, (0, 0, 64, 64), api.device_view(mem, ts.PixelType.uint16, 63, 64))
    assert server.check(wrong_size) == ts.Status.invalid_argument
    too_many_tiles = api.region_request(2, (0, 0, 1000, 600), api.device_view(mem, ts.PixelType.uint16, 1000, 600))
    assert server.check(too_many_tiles) == ts.Status.resource_limit
    # outputs of one replay must not overlap
    a = api.region_request(0, (0, 0, 64, 64), api.device_view(mem, ts.PixelType.uint16, 64, 64))
    b = api.region_request(0, (64, 0, 64, 64), api.device_view(mem, ts.PixelType.uint16, 64, 64, offset=4096))
    st, *_ = server.run_batch([a, b])
    assert st == ts.Status.invalid_argument
    # nor write into an archive
    arch = resident.mems[0]
    c = api.region_request(0, (0, 0, 8, 8), api.device_view(arch, ts.PixelType.uint16, 8, 8, offset=128))
    st, *_ = server.run_batch([c])
    assert st == ts.Status.invalid_argument
    # arrivals must not decrease
    d = api.region_request(0, (64, 0, 64, 64), api.device_view(mem, ts.PixelType.uint16, 64, 64, offset=16384))
    st, *_ = server.replay([ok, d], np.array([5, 1], np.uint64))
    assert st == ts.Status.invalid_argument


def test_archive_set_rejects_bad_archives(orc):
    px = g.detector_u16(200, 100, 3)
    data = bytearray(orc.compress(px, (64, 64)))
    st, meta = api.parse_metadata(bytes(data))
    assert st == ts.Status.success
    s = api.ArchiveSet()
    mem = api.DeviceMemory.from_host(np.frombuffer(bytes(data), np.uint8))
    st, _ = s.add(mem, meta, size=len(data) - 4)
    assert st == ts.Status.buffer_too_small
    t = meta.tile(1)
    t.width -= 1  # no longer the tile's geometry
    meta.set_tile(1, t)
    st, _ = s.add(mem, meta)
    assert st == ts.Status.malformed_data
    assert len(s) == 0
    with pytest.raises(RuntimeError):
        api.Server(s)  # nothing committed


ENCODES = [("u16", 300, 200, (128, 32)), ("f32", 129, 77, (64, 64)), ("u16", 1, 1, (256, 16)),
           ("f32", 700, 90, (256, 16)), ("u16", 640, 32 * 900, (128, 32)), ("u16", 64, 64, (64, 64))]


class EncodeJobs:
    """Device inputs (rows padded past the width) and bound-sized outputs with guard bytes."""

    def __init__(self, specs, seed, guard=64):
        self.pixels, self.inputs, self.outs, self.bounds, self.requests = [], [], [], [], []
        for i, (kind, w, h, tile) in enumerate(specs):
            px = g.detector_u16(w, h, seed + i) if kind == "u16" else g.field_f32(w, h, seed + i)
            host = np.zeros((h, w + 3), px.dtype)
            host[:, :w] = px
            src = api.DeviceMemory.from_host(host)
            view = api.device_view(src, api.pixel_type_of(px.dtype), w, h, (w + 3) * px.itemsize, const=True)
            st, bound = api.compress_bound(view.image, api.TileShape(*tile))
            assert st == ts.Status.success
            out = api.DeviceMemory(bound + guard)
            out.fill(0xA5)
            self.pixels.append(px)
            self.inputs.append(src)
            self.outs.append(out)
            self.bounds.append(bound)
            self.requests.append(api.encode_request(view, tile, api.MutableBytes(out.ptr, bound)))

    def file(self, i, written):
        raw = self.outs[i].download()
        assert (raw[self.bounds[i]:] == 0xA5).all(), "guard bytes after the bound were written"
        return raw[:written].tobytes()


def test_encode_requests_match_oracle_in_a_mixed_batch(orc, resident):
    """Encodes (both pixel types, one image over 4096 tiles) interleaved with regions."""
    rng = np.random.default_rng(404)
    specs = _requests(resident, rng, 30)
    out = Outputs(resident, specs)
    jobs = EncodeJobs(ENCODES, 410)
    requests = list(out.requests)
    where = []
    for k, r in enumerate(jobs.requests):
        at = 1 + 5 * k
        requests.insert(at, r)
        where.append(at)
    server = api.Server(resident.set, api.server_options(streams=2, max_batch=64))
    st, statuses, written = server.run_batch(requests)
    assert st == ts.Status.success and (statuses == 0).all()
    for k, at in enumerate(where):
        want = orc.compress(jobs.pixels[k], ENCODES[k][3])
        assert written[at] == len(want), (k, ENCODES[k])
        assert jobs.file(k, int(written[at])) == want, (k, ENCODES[k])
    regions = [i for i in range(len(requests)) if i not in where]
    for i, spec in enumerate(specs):
        assert _same(out.read(specs, i), _expected(orc, resident, spec)), (i, spec)
        assert written[regions[i]] == spec[1][2] * spec[1][3] * (4 if spec[2] else resident.pixels[spec[0]].itemsize)


@pytest.mark.parametrize("large", [None, 256 << 10])
def test_encode_alone_and_in_the_mix_are_identical(resident, large):
    rng = np.random.default_rng(405)
    specs = _requests(resident, rng, 24)
    out = Outputs(resident, specs)
    jobs = EncodeJobs(ENCODES, 420)
    requests = jobs.requests + list(out.requests)
    order = np.random.default_rng(406).permutation(len(requests))
    mixed = [requests[i] for i in order]
    arrivals = np.cumsum(np.random.default_rng(407).exponential(30_000.0, len(mixed))).astype(np.uint64)
    server = api.Server(resident.set, api.server_options(streams=3, max_batch=6, admission_bytes=8 << 20,
                                                         large_bytes=large))
    st, latency, statuses, written, stats = server.replay(mixed, arrivals)
    assert st == ts.Status.success and (statuses == 0).all() and stats.failed == 0
    pos = {int(i): k for k, i in enumerate(order)}
    mix_files = [jobs.file(k, int(written[pos[k]])) for k in range(len(ENCODES))]
    want_bytes = sum(p.nbytes for p in jobs.pixels) + sum(
        r[1][2] * r[1][3] * resident.pixels[r[0]].itemsize for r in specs)
    assert stats.decoded_bytes == want_bytes
    for k in range(len(ENCODES)):
        jobs.outs[k].fill(0xA5)
        st, statuses, written = server.run_batch([jobs.requests[k]])
        assert st == ts.Status.success and statuses[0] == 0
        assert jobs.file(k, int(written[0])) == mix_files[k], (k, ENCODES[k])


def test_encode_request_checks(resident):
    server = api.Server(resident.set, api.server_options(max_tasks=64))
    src = api.DeviceMemory(1 << 20)
    dst = api.DeviceMemory(1 << 20)
    view = api.device_view(src, ts.PixelType.uint16, 100, 100, const=True)
    st, bound = api.compress_bound(view.image, api.TileShape(64, 64))
    ok = api.encode_request(view, (64, 64), api.MutableBytes(dst.ptr, bound))
    assert server.check(ok) == ts.Status.success
    small = api.encode_request(view, (64, 64), api.MutableBytes(dst.ptr, bound - 4))
    assert server.check(small) == ts.Status.buffer_too_small
    misaligned = api.encode_request(view, (64, 64), api.MutableBytes(dst.ptr + 2, bound))
    assert server.check(misaligned) == ts.Status.invalid_argument
    bad_tile = api.encode_request(view, (0, 64), api.MutableBytes(dst.ptr, bound))
    assert server.check(bad_tile) == ts.Status.invalid_argument
    big = api.device_view(src, ts.PixelType.uint16, 700, 700, const=True)
    st, big_bound = api.compress_bound(big.image, api.TileShape(64, 64))
    too_many = api.encode_request(big, (64, 64), api.MutableBytes(dst.ptr, big_bound))
    assert server.check(too_many) == ts.Status.resource_limit  # 121 tiles > 64
    # an encode output must not overlap its own input or another output
    into_input = api.encode_request(view, (64, 64), api.MutableBytes(src.ptr + 4096, bound))
    st, *_ = server.run_batch([into_input])
    assert st == ts.Status.invalid_argument
    st, *_ = server.run_batch([ok, ok])
    assert st == ts.Status.invalid_argument


@pytest.mark.parametrize("window", [64, 200])
def test_encode_windows_carry_cursors(orc, resident, window):
    """Small encoder windows: images split across windows, middle windows that
    both continue and are continued, parity alternating across pixel types."""
    jobs = EncodeJobs(ENCODES, 430)
    rng = np.random.default_rng(431)
    specs = _requests(resident, rng, 10)
    out = Outputs(resident, specs)
    requests = jobs.requests[:3] + list(out.requests) + jobs.requests[3:]
    server = api.Server(resident.set, api.server_options(streams=2, max_batch=64, encode_window=window))
    st, statuses, written = server.run_batch(requests)
    assert st == ts.Status.success and (statuses == 0).all()
    where = [0, 1, 2] + [3 + len(specs) + k for k in range(len(ENCODES) - 3)]
    for k, at in enumerate(where):
        want = orc.compress(jobs.pixels[k], ENCODES[k][3])
        assert jobs.file(k, int(written[at])) == want, (window, k, ENCODES[k])
    for i, spec in enumerate(specs):
        assert _same(out.read(specs, i), _expected(orc, resident, spec)), (i, spec)


def test_server_rejects_bad_encode_window(resident):
    for window in (0, 63, (1 << 20) + 1):
        with pytest.raises(RuntimeError):
            api.Server(resident.set, api.server_options(encode_window=window))


class Mixed:
    """The archives once in device memory and once in pinned host memory (at
    offsets that are 4 mod 16, so staged payloads keep an odd block phase)."""

    def __init__(self, orc, resident):
        self.pixels = resident.pixels + resident.pixels
        self.files = resident.files + resident.files
        offsets, total = [], 0
        for data in resident.files:
            offsets.append(total + 4)
            total += (len(data) + 4 + 255) // 256 * 256
        self.host = api.HostMemory(total)
        for data, off in zip(resident.files, offsets):
            self.host.upload(np.frombuffer(data, np.uint8), off)
        self.set = api.ArchiveSet()
        for i, (data, mem) in enumerate(zip(resident.files, resident.mems)):
            st, meta = api.parse_metadata(data)
            st, ident = self.set.add(mem, meta)
            assert st == ts.Status.success and ident == i
        for i, (data, off) in enumerate(zip(resident.files, offsets)):
            st, meta = api.parse_metadata(data)
            st, ident = self.set.add(self.host, meta, len(data), off)
            assert st == ts.Status.success and ident == len(ARCHIVES) + i, st
        assert self.set.commit() == ts.Status.success


@pytest.fixture(scope="module")
def mixed(orc, resident):
    return Mixed(orc, resident)


def _mixed_requests(mixed, rng, count, big=True):
    specs = []
    for k in range(count):
        a = int(rng.integers(0, 2 * len(ARCHIVES)))
        h, w = mixed.pixels[a].shape
        if big and k % 5 == 0:
            rw, rh = int(rng.integers(w // 2, w + 1)), int(rng.integers(h // 2, h + 1))
        else:
            rw, rh = int(rng.integers(1, min(w, 150) + 1)), int(rng.integers(1, min(h, 120) + 1))
        x, y = int(rng.integers(0, w - rw + 1)), int(rng.integers(0, h - rh + 1))
        specs.append((a, (x, y, rw, rh), bool(rng.integers(0, 2))))
    return specs


def test_host_archives_are_staged(mixed):
    for i in range(2 * len(ARCHIVES)):
        st, info = mixed.set.info(i)
        assert st == ts.Status.success and info.host == (1 if i >= len(ARCHIVES) else 0)


@pytest.mark.parametrize("staging,big", [(256 << 20, True), (512 << 10, False)])
def test_host_and_device_archives_in_one_batch(orc, mixed, staging, big):
    """Tiles of host archives staged per batch next to tiles read in place;
    a small staging area splits the batches."""
    rng = np.random.default_rng(600 + staging % 97)
    specs = _mixed_requests(mixed, rng, 80, big)
    out = Outputs(mixed, specs)
    # With every request arriving at once, one batch takes them all unless the
    # staging area (small requests only) splits it.
    server = api.Server(mixed.set, api.server_options(streams=3, max_batch=32 if big else 1024,
                                                      staging_bytes=staging))
    st, latency, statuses, _, stats = server.replay(out.requests, np.zeros(len(specs), np.uint64))
    assert st == ts.Status.success and (statuses == 0).all()
    if not big:
        assert stats.batches >= 2
    for i, spec in enumerate(specs):
        assert _same(out.read(specs, i), _expected(orc, mixed, spec)), (staging, i, spec)


def test_staged_alone_and_in_the_mix_are_identical(mixed):
    rng = np.random.default_rng(610)
    specs = _mixed_requests(mixed, rng, 40)
    mix = Outputs(mixed, specs)
    server = api.Server(mixed.set, api.server_options(streams=2, max_batch=16))
    arrivals = np.cumsum(rng.exponential(20_000.0, len(specs))).astype(np.uint64)
    st, _, statuses, _, _ = server.replay(mix.requests, arrivals)
    assert st == ts.Status.success and (statuses == 0).all()
    for i in range(0, len(specs), 3):
        alone = Outputs(mixed, [specs[i]])
        st, statuses, _ = server.run_batch(alone.requests)
        assert st == ts.Status.success and statuses[0] == 0
        assert np.array_equal(alone.read([specs[i]], 0), mix.read(specs, i)), (i, specs[i])


def test_staging_checks(mixed):
    server = api.Server(mixed.set, api.server_options(staging_bytes=64 << 10))
    mem = api.DeviceMemory(4 << 20)
    h, w = mixed.pixels[len(ARCHIVES) + 2].shape
    whole = api.region_request(len(ARCHIVES) + 2, (0, 0, w, h), api.device_view(mem, ts.PixelType.uint16, w, h))
    assert server.check(whole) == ts.Status.resource_limit  # its payload exceeds the staging area
    on_device = api.region_request(2, (0, 0, w, h), api.device_view(mem, ts.PixelType.uint16, w, h))
    assert server.check(on_device) == ts.Status.success
    small = api.region_request(len(ARCHIVES) + 2, (0, 0, 64, 16), api.device_view(mem, ts.PixelType.uint16, 64, 16))
    assert server.check(small) == ts.Status.success
    for bad in (0, 8):
        with pytest.raises(RuntimeError):
            api.Server(mixed.set, api.server_options(staging_bytes=bad))
