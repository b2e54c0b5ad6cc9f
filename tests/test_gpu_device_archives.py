"""Archives added to a serving set straight from device memory (ArchiveSet
add_device): the GPU encoder's output is validated and tabled on the device
and served without a host parse.  Results equal the oracle and the host-tabled
path; the device validation gives the host parser's statuses on the reference
fuzz corpus and on planted corruptions."""
import os

import numpy as np
import pytest

import tilestream_amd as ts
from tilestream_amd import api, generators as g
from tilestream_amd._native import BatchItem, ImageSpec, TileShape

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
OFFSET, GAIN = 100.0, 1.05
IMAGES = [("u16", 700, 389, (64, 64)), ("f32", 517, 301, (128, 32)), ("u16", 1000, 600, (256, 16)),
          ("f32", 300, 200, (64, 64)), ("u16", 129, 33, (128, 32)), ("f32", 1, 1, (256, 16)),
          ("u16", 255, 1, (64, 64)), ("u16", 64, 64, (64, 64))]


def _pixels(kind, w, h, seed):
    return g.detector_u16(w, h, seed) if kind == "u16" else g.field_f32(w, h, seed)


class DeviceEncoded:
    """Images encoded on the GPU, one plan per (tile shape, pixel type); the
    archives stay in device memory next to each encode's device result."""

    def __init__(self, specs, seed=900):
        self.pixels, self.outs, self.results, self.index, self.keep = [], [], [], [], []
        by_tile = {}
        for i, (kind, w, h, tile) in enumerate(specs):
            px = _pixels(kind, w, h, seed + i)
            self.pixels.append(px)
            by_tile.setdefault((tile, kind), []).append(i)
        self.outs = [None] * len(specs)
        self.index = [None] * len(specs)
        stream = api.Stream()
        for (tile, _), members in by_tile.items():
            items = []
            for i in members:
                px = self.pixels[i]
                h, w = px.shape
                src = api.DeviceMemory.from_host(np.ascontiguousarray(px))
                st, bound = api.compress_bound(api.host_view(px).image, TileShape(*tile))
                assert st == ts.Status.success
                out = api.DeviceMemory(bound)
                it = BatchItem()
                it.input = api.device_view(src, api.pixel_type_of(px.dtype), w, h, const=True)
                it.compressed_output = api.MutableBytes(out.ptr, bound)
                items.append(it)
                self.outs[i] = out
                self.keep.append(src)
            st, plan = api.prepare_batch(ts.Operation.compress, items, TileShape(*tile), stream=stream)
            assert st == ts.Status.success, st
            ws = api.DeviceMemory(max(256, plan.workspace_bytes()))
            res = api.DeviceMemory(16 * len(members))
            assert api.compress_async(plan, ws, res, len(members), stream) == ts.Status.success
            for k, i in enumerate(members):
                self.index[i] = (res, 16 * k)
            self.keep += [plan, ws, items]
            self.results.append(res)
        stream.synchronize()
        self.specs = specs

    def producer(self, i):
        res, off = self.index[i]
        return res.ptr + off

    def spec(self, i):
        px = self.pixels[i]
        return ImageSpec(int(api.pixel_type_of(px.dtype)), px.shape[1], px.shape[0])

    def file(self, i):
        res, off = self.index[i]
        status, written = api.read_results(res, len(res.download()) // 16)
        n = int(written[off // 16])
        return self.outs[i].download(n).tobytes()


@pytest.fixture(scope="module")
def encoded():
    return DeviceEncoded(IMAGES)


def _requests(pixels, rng, count):
    specs = []
    for k in range(count):
        a = int(rng.integers(0, len(pixels)))
        h, w = pixels[a].shape
        rw, rh = int(rng.integers(1, w + 1)), int(rng.integers(1, h + 1))
        if k % 5 == 0:
            rw, rh = w, h
        x, y = int(rng.integers(0, w - rw + 1)), int(rng.integers(0, h - rh + 1))
        specs.append((a, (x, y, rw, rh), bool(rng.integers(0, 2))))
    return specs


class Outputs:
    def __init__(self, pixels, specs):
        self.offsets, total = [], 0
        for a, (x, y, w, h), cal in specs:
            s = 4 if cal or pixels[a].dtype != np.uint16 else 2
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
        return self.mem.download(w * h * s, off).view(np.uint16 if s == 2 else np.uint32).reshape(h, w)


def _expected(pixels, spec):
    a, (x, y, w, h), cal = spec
    px = pixels[a][y:y + h, x:x + w]
    if cal:
        with np.errstate(all="ignore"):
            f = px.astype(np.float32) if px.dtype == np.uint16 else px.view(np.float32)
            return ((f - np.float32(OFFSET)) * np.float32(GAIN)).view(np.uint32)
    return px


def _same(got, want):
    if got.dtype == np.uint32:
        fg, fw = got.view(np.float32), want.view(np.float32)
        return bool(np.all((got == want) | (np.isnan(fg) & np.isnan(fw))))
    return bool(np.array_equal(got, want))


def test_gpu_encoded_archives_are_canonical(orc, encoded):
    for i, (kind, w, h, tile) in enumerate(IMAGES):
        assert encoded.file(i) == orc.compress(encoded.pixels[i], tile), i


@pytest.mark.parametrize("with_producer", [True, False])
def test_serve_device_archives_without_host_parse(encoded, with_producer):
    aset = api.ArchiveSet()
    for i in range(len(IMAGES)):
        if with_producer:
            st, ident = aset.add_device(encoded.outs[i], encoded.spec(i), IMAGES[i][3], encoded.producer(i))
        else:
            st, ident = aset.add_device(encoded.outs[i], encoded.spec(i), IMAGES[i][3], size=len(encoded.file(i)))
        assert st == ts.Status.success and ident == i, (i, st)
        st, info = aset.info(ident)
        assert info.file_bytes == len(encoded.file(i)) and info.tile_count > 0 and info.host == 0
    assert aset.commit() == ts.Status.success
    rng = np.random.default_rng(91 + with_producer)
    specs = _requests(encoded.pixels, rng, 120)
    out = Outputs(encoded.pixels, specs)
    server = api.Server(aset, api.server_options(streams=2, max_batch=32))
    st, statuses, _ = server.run_batch(out.requests)
    assert st == ts.Status.success and (statuses == 0).all()
    for i, spec in enumerate(specs):
        assert _same(out.read(specs, i), _expected(encoded.pixels, spec)), (i, spec)


def test_device_and_host_tables_serve_identically(encoded):
    """The same archives tabled on the device and on the host (parse of the
    downloaded bytes), in one set: every request gives the same words."""
    aset = api.ArchiveSet()
    n = len(IMAGES)
    for i in range(n):
        st, _ = aset.add_device(encoded.outs[i], encoded.spec(i), IMAGES[i][3], encoded.producer(i))
        assert st == ts.Status.success
    metas = []
    for i in range(n):
        data = encoded.file(i)
        st, meta = api.parse_metadata(data)
        assert st == ts.Status.success
        metas.append(meta)
        st, ident = aset.add(encoded.outs[i], meta, len(data))
        assert st == ts.Status.success and ident == n + i
    assert aset.commit() == ts.Status.success
    rng = np.random.default_rng(17)
    specs = _requests(encoded.pixels, rng, 80)
    twin = [(a + n, roi, cal) for a, roi, cal in specs]
    pixels = encoded.pixels + encoded.pixels
    out, out2 = Outputs(pixels, specs), Outputs(pixels, twin)
    server = api.Server(aset, api.server_options(streams=1, max_batch=16))
    for o in (out, out2):
        st, statuses, _ = server.run_batch(o.requests)
        assert st == ts.Status.success and (statuses == 0).all()
    for i in range(len(specs)):
        assert np.array_equal(out.read(specs, i), out2.read(twin, i)), i


def test_add_after_commit_keeps_device_tables(encoded):
    """A later commit re-uploads the host tables and copies the device-built
    ones again: archives added before and after both serve correctly."""
    aset = api.ArchiveSet()
    assert aset.add_device(encoded.outs[0], encoded.spec(0), IMAGES[0][3], encoded.producer(0))[0] == ts.Status.success
    assert aset.commit() == ts.Status.success
    assert aset.add_device(encoded.outs[1], encoded.spec(1), IMAGES[1][3], encoded.producer(1))[0] == ts.Status.success
    assert aset.commit() == ts.Status.success
    pixels = encoded.pixels[:2]
    specs = [(0, (0, 0, pixels[0].shape[1], pixels[0].shape[0]), False),
             (1, (0, 0, pixels[1].shape[1], pixels[1].shape[0]), False), (0, (5, 7, 100, 50), True)]
    out = Outputs(pixels, specs)
    st, statuses, _ = api.Server(aset).run_batch(out.requests)
    assert st == ts.Status.success and (statuses == 0).all()
    for i, spec in enumerate(specs):
        assert _same(out.read(specs, i), _expected(pixels, spec)), i


def _device_copy(data: bytes, extra=0):
    mem = api.DeviceMemory(max(256, len(data) + extra))
    if data:
        mem.upload(np.frombuffer(data, np.uint8))
    return mem


def test_device_validation_rejections(orc):
    px = g.detector_u16(300, 200, 5)
    tile = (64, 64)
    data = orc.compress(px, tile)
    spec = ImageSpec(1, 300, 200)
    st, meta = api.parse_metadata(data)
    first = meta.tile(0)
    aset = api.ArchiveSet()
    ok = _device_copy(data)
    assert aset.add_device(ok, spec, tile, size=len(data))[0] == ts.Status.success
    # Another image than the caller expects.
    assert aset.add_device(ok, ImageSpec(1, 301, 200), tile, size=len(data))[0] == ts.Status.invalid_argument
    assert aset.add_device(ok, ImageSpec(2, 300, 200), tile, size=len(data))[0] == ts.Status.invalid_argument
    assert aset.add_device(ok, spec, (128, 32), size=len(data))[0] == ts.Status.invalid_argument
    # Not the whole file, or more than the file.
    assert aset.add_device(ok, spec, tile, size=len(data) - 4)[0] == ts.Status.malformed_data
    big = _device_copy(data, 16)
    assert aset.add_device(big, spec, tile, size=len(data) + 16)[0] == ts.Status.malformed_data
    # Header fields: magic, version, pixel type.
    for at, value, want in ((0, 0x55, ts.Status.malformed_data), (8, 2, ts.Status.unsupported_format),
                            (14, 3, ts.Status.unsupported_type), (15, 1, ts.Status.unsupported_format)):
        bad = bytearray(data)
        bad[at] = value
        assert api.parse_metadata(bytes(bad))[0] == want
        assert aset.add_device(_device_copy(bytes(bad)), spec, tile, size=len(data))[0] == want, at
    # Payload corruptions the host parser rejects: a row descriptor's reserved
    # byte, a packed tail bit, an index entry's reserved byte.
    mutations = []
    for k in range(meta.summary.tile_count):
        t = meta.tile(k)
        if t.mode == 2:
            seg = meta.segment(t.first_segment)
            mutations.append(t.offset + seg.payload_offset - 3)      # descriptor reserved byte
            if seg.payload_bytes and ((t.width - 1) * seg.bit_width) % 32:
                mutations.append(t.offset + seg.payload_offset + seg.payload_bytes - 1)  # bits past the last residual
            break
    mutations.append(64 + 1)  # index entry 0 reserved byte
    for at in mutations:
        bad = bytearray(data)
        bad[at] ^= 0x80
        assert api.parse_metadata(bytes(bad))[0] == ts.Status.malformed_data, at
        assert aset.add_device(_device_copy(bytes(bad)), spec, tile, size=len(data))[0] == ts.Status.malformed_data, at
    # A producer that failed: its status is the answer.
    failed = api.DeviceMemory(16)
    failed.upload(np.array([7, 0, 0, 0], np.uint32))
    assert aset.add_device(ok, spec, tile, failed.ptr)[0] == ts.Status.buffer_too_small
    # A producer that claims more bytes than the buffer holds.
    claim = api.DeviceMemory(16)
    claim.upload(np.array([0, 0, len(data) + 64, 0], np.uint32))
    assert aset.add_device(ok, spec, tile, claim.ptr, size=len(data))[0] == ts.Status.buffer_too_small
    # Host memory is not accepted, nor an unaligned start.
    host = api.HostMemory(len(data) + 256) if hasattr(api, "HostMemory") else None
    if host is not None:
        assert aset.add_device(host, spec, tile, size=len(data))[0] == ts.Status.invalid_argument
    assert aset.add_device(ok, spec, tile, size=len(data) - 2, offset=2)[0] == ts.Status.invalid_argument
    # Only the valid archive was added.
    assert len(aset) == 1 and first.length > 0


def test_fuzz_corpus_statuses_match_host_parser():
    """Every input of the reference fuzz corpus (tests/golden/fuzz_replay.npz)
    whose header names a valid image: the device validation's status equals
    the host parser's (no limits); the others are rejected too."""
    z = np.load(os.path.join(HERE, "golden", "fuzz_replay.npz"))
    offsets, blob = z["offsets"], z["data"]
    aset = api.ArchiveSet()
    compared = rejected = accepted = 0
    for k in range(len(offsets) - 1):
        data = blob[offsets[k]:offsets[k + 1]].tobytes()
        want, _ = api.parse_metadata(data)
        accepted += want == ts.Status.success
        spec, tile = ImageSpec(1, 1, 1), (64, 64)
        if len(data) >= 28:
            ptype = data[14]
            width, height = int.from_bytes(data[16:20], "little"), int.from_bytes(data[20:24], "little")
            tw, th = int.from_bytes(data[24:26], "little"), int.from_bytes(data[26:28], "little")
            spec, tile = ImageSpec(ptype, width, height), (tw, th)
        mem = _device_copy(data)
        st_bound, _ = api.compress_bound(spec, TileShape(*tile))
        st, _ = aset.add_device(mem, spec, tile, size=len(data))
        if st_bound == ts.Status.success and len(data) >= 28:
            assert st == want, (str(z["name"][k]), st, want)
            compared += 1
        else:
            assert st != ts.Status.success, str(z["name"][k])
            rejected += 1
    assert compared >= 500 and accepted >= 300
    assert len(aset) == accepted
