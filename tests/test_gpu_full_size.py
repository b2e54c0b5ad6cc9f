"""Parity at the measured workloads' full sizes (the 4096 x 4096 fused frames
are covered in test_gpu_behaviour.py):

* the batched persistent decode of the 256-image uneven batch (sides 256 to
  4096 in steps of 16, half uint16 detector frames and half float32 fields,
  tiles 128 x 32, 64 x 64 or 256 x 16, output rows on 256-byte boundaries,
  3.65 GB decoded), through both decoders and a captured graph: every image
  equal to the pixels it was encoded from, every result record success with
  the image's byte count;
* the GPU encoder over the same images (rows on 256 bytes), every archive
  byte-identical to the host codec's;
* region serving over 64 resident 4096 x 4096 uint16 archives (64 x 64
  tiles): 10,000 region requests with log-uniform sides in [32, 2048], each
  fused with (x - 100) * 1.05 to FP32, served by the continuous-batching
  server with 3 streams, every output equal word for word to binary32
  (x - 100) * 1.05 of the generated pixels (numpy rounds each float32
  operation once, as the kernels do).

The pixels come from the seeded generators the archives were encoded from
(by the host codec, which equals the oracle on the golden vectors), so the
comparison needs no oracle pass over gigabytes."""
import numpy as np
import pytest

from tilestream_amd import api, generators as g
from tilestream_amd.workloads import ResidentBatch

pytestmark = pytest.mark.gpu

SEED = 2026517


@pytest.fixture(scope="module")
def full_batch():
    return ResidentBatch.build(g.batch_layout(256, SEED), SEED, pitch_align=256)


@pytest.mark.parametrize("shape", [(0, 512, 0, 1, 2), (4, 256, 0, 4, 3)])
def test_full_batch_decodes_every_image(full_batch, shape):
    batch = full_batch
    n = len(batch.entries)
    assert n == 256 and batch.decoded_bytes > 3_500_000_000
    items = [batch.item(i) for i in range(n)]
    st, queue = api.DecodeQueue.prepare(items, [api.Operation.decompress] * n)
    assert st == api.Status.success, st
    assert queue.set_shape(*shape) == api.Status.success
    ws = api.DeviceMemory(max(256, queue.workspace_bytes()))
    results = api.DeviceMemory(16 * n)
    stream = api.Stream()
    assert queue.capture(ws, results, stream) == api.Status.success
    for replay in range(2):
        batch.clear_outputs()
        assert queue.launch(stream) == api.Status.success
        stream.synchronize()
        status, written = api.read_results(results, n)
        assert (status == 0).all(), (shape, replay)
        for i, e in enumerate(batch.entries):
            assert int(written[i]) == e.raw_bytes
            assert np.array_equal(batch.read_output(i), e.pixels), (shape, replay, i, e.width, e.height, e.tile)


def test_full_batch_encodes_every_image(full_batch):
    """The GPU encoder over the same 256 images, rows on 256 bytes (as the
    decode above writes them), one compress plan per pixel type and tile
    shape: every archive byte-identical to the host codec's, which equals the
    oracle on the golden vectors."""
    batch = full_batch
    offsets, pitches, total = [], [], 0
    for e in batch.entries:
        pitch = (e.width * e.sample + 255) // 256 * 256
        offsets.append(total)
        pitches.append(pitch)
        total += (pitch * (e.height - 1) + e.width * e.sample + 255) // 256 * 256
    raw = api.DeviceMemory(total)
    for e, off, pitch in zip(batch.entries, offsets, pitches):
        rows = np.zeros((e.height, pitch), np.uint8)
        rows[:, :e.width * e.sample] = e.pixels.view(np.uint8).reshape(e.height, -1)
        raw.upload(rows.reshape(-1)[:pitch * (e.height - 1) + e.width * e.sample], off)
    groups = {}
    for i, e in enumerate(batch.entries):
        groups.setdefault((int(e.ptype), tuple(e.tile)), []).append(i)
    stream = api.Stream()
    for (ptype, tile), members in sorted(groups.items()):
        items, outs = [], []
        for i in members:
            e = batch.entries[i]
            st, bound = api.compress_bound(api.ImageSpec(ptype, e.width, e.height), api.TileShape(*tile))
            assert st == api.Status.success
            out = api.DeviceMemory(bound)
            it = api.BatchItem()
            it.input = api.device_view(raw, e.ptype, e.width, e.height, pitches[i], offset=offsets[i],
                                       capacity=pitches[i] * (e.height - 1) + e.width * e.sample, const=True)
            it.compressed_output = api.MutableBytes(out.ptr, bound)
            items.append(it)
            outs.append(out)
        st, plan = api.prepare_batch(api.Operation.compress, items, api.TileShape(*tile), stream=stream)
        assert st == api.Status.success, st
        ws = api.DeviceMemory(max(256, plan.workspace_bytes()))
        res = api.DeviceMemory(16 * len(members))
        assert api.compress_async(plan, ws, res, len(members), stream) == api.Status.success
        stream.synchronize()
        status, written = api.read_results(res, len(members))
        assert (status == 0).all(), (ptype, tile)
        for i, out, n in zip(members, outs, written):
            archive = batch.entries[i].archive
            assert int(n) == len(archive), (i, ptype, tile)
            assert out.download(len(archive)).tobytes() == archive, (i, ptype, tile)


def test_full_region_serving_matches_pixels():
    side, tile, archives, count = 4096, (64, 64), 64, 10_000
    offset, gain = 100.0, 1.05
    rb = ResidentBatch.build([(side, side, np.uint16, tile)] * archives, SEED + 4)
    stream = api.Stream()
    aset = api.ArchiveSet()
    for e in rb.entries:
        st, _ = aset.add(rb.src, e.meta, len(e.archive), e.src_offset)
        assert st == api.Status.success, st
    assert aset.commit(stream) == api.Status.success
    rois = g.roi_requests(count, SEED + 5, side, side, archives=archives)
    offsets, total = [], 0
    for _, _, _, w, h in rois:
        offsets.append(total)
        total += (w * h * 4 + 255) // 256 * 256
    out = api.DeviceMemory(total)
    out.fill(0)
    requests = [api.region_request(a, (x, y, w, h),
                                   api.device_view(out, api.PixelType.float32, w, h, w * 4, offset=off,
                                                   capacity=w * h * 4), True,
                                   offset, gain)
                for (a, x, y, w, h), off in zip(rois, offsets)]
    server = api.Server(aset, api.server_options(streams=3, max_batch=1024, shape=(4, 256, 0, 4, 3)))
    st, _, statuses, written, stats = server.replay(requests, np.zeros(count, np.uint64))
    assert st == api.Status.success and (statuses == 0).all(), st
    assert stats.batches > 1
    o32, g32 = np.float32(offset), np.float32(gain)
    for i, ((a, x, y, w, h), off) in enumerate(zip(rois, offsets)):
        assert int(written[i]) == w * h * 4
        want = ((rb.entries[a].pixels[y:y + h, x:x + w].astype(np.float32) - o32) * g32).view(np.uint32)
        got = out.download(w * h * 4, off).view(np.uint32).reshape(h, w)
        assert np.array_equal(got, want), (i, a, x, y, w, h)
