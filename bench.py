"""Benchmark: decoded GB/s of batched persistent decode vs the plain per-image path.

Workload (one "step"): decode a batch of 256 seeded images whose width and
height are uniform in {256, 272, ..., 4096} (not tile-aligned), half uint16
detector frames (background + Gaussian blobs + Poisson noise), half float32
octave fields with 0.1% special words, tile shape drawn from {128x32, 64x64,
256x16}.  Archives and outputs are resident in HBM before timing starts.

  batched   one DecodeQueue launch (persistent kernel draining a device work
            queue) captured in a HIP graph; launch shape from the autotuner
  per-image 256 decompress_async calls, one plan per image, same stream

Further legs in the same JSON line (each can be skipped or run alone):
  roi        10,000 region requests over 64 resident 4096x4096 archives, fused
             calibration, through the serving loop: saturation, p50/p99 latency
  fused      4096x4096 whole decode fused with dark and gain maps, FP64 errors
  encode     the batch's pixels encoded on the GPU (six compress plans per step)
  serve_mix  Poisson mix of region decodes, whole decodes and GPU encodes over
             images with sides from {512..8192}: throughput and p50/p99 per kind

Prints one JSON line.  Outputs are verified bit-exact after timing.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "decoded GB/s (uncompressed) batched-persistent vs per-image; p99 ROI latency"
HBM_PEAK_GBPS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def rank_stats(per_call) -> dict:
    """Nearest-rank p05/p50/p95 (ms) of per-call times."""
    v = sorted(per_call)
    rank = lambda p: v[max(0, int(np.ceil(p * len(v))) - 1)]  # noqa: E731
    return {"p05_ms": round(rank(0.05), 4), "p50_ms": round(rank(0.50), 4), "p95_ms": round(rank(0.95), 4)}


def sampled(stream, fn, samples: int = 30, min_ms: float = 2.0):
    """The reference's sample rule (benchmarks/runner.h:197-246) in one process:
    each sample repeats fn until it lasts at least min_ms (at most 1000 calls),
    timed with events on the work's stream; nearest-rank p05/p50/p95 of the
    per-call times in ms, and the per-call time of every sample."""
    one = timed(stream, fn, 1)
    reps = int(min(1000, max(1, np.ceil(min_ms / max(one, 1e-6)))))
    per_call = [timed(stream, fn, reps) for _ in range(samples)]
    return {"samples": samples, "calls_per_sample": reps, **rank_stats(per_call),
            "per_call_ms": [round(t, 5) for t in per_call]}


def per_image_plans(batch, items, stream):
    """One decompress plan per image (the reference's per-image method), the
    workspace they share and one result record each."""
    from tilestream_amd import api
    plans, ws_bytes = [], 0
    for it in items:
        st, plan = api.prepare_batch(api.Operation.decompress, [it], stream=stream)
        assert st == api.Status.success, st
        plans.append(plan)
        ws_bytes = max(ws_bytes, plan.workspace_bytes())
    return plans, api.DeviceMemory(ws_bytes), [api.DeviceMemory(16) for _ in plans]


def protocol_round(options: dict, shape: tuple, batched_first: bool) -> dict:
    """One fresh-process round of the sample rule: build the seeded batch,
    prepare the queue at the given launch shape and the per-image plans, warm
    up, then sample both methods (in the given order).  Runs in a process
    forked from a server that never touched the GPU."""
    from tilestream_amd import api, generators
    from tilestream_amd.workloads import ResidentBatch
    layout = generators.batch_layout(options["images"], options["seed"])
    batch = ResidentBatch.build(layout, options["seed"], options["threads"], keep_pixels=False,
                                pitch_align=options["pitch_align"])
    stream = api.Stream()
    n = len(batch.entries)
    items = [batch.item(i) for i in range(n)]
    st, queue = api.DecodeQueue.prepare(items, [api.Operation.decompress] * n, stream)
    assert st == api.Status.success, st
    ws = api.DeviceMemory(max(queue.workspace_bytes(), 256))
    results = api.DeviceMemory(16 * n)
    assert queue.set_shape(*shape) == api.Status.success
    assert queue.capture(ws, results, stream) == api.Status.success
    plans, pws, rec = per_image_plans(batch, items, stream)
    for plan in plans:
        assert plan.set_decode_shape(8, 256, 0, 1) == api.Status.success

    def per_image():
        for plan, r in zip(plans, rec):
            api.decompress_async(plan, pws, r, 1, stream)

    def batched():
        queue.launch(stream)

    for _ in range(options["warmup"]):
        batched()
        per_image()
    order = [("batched", batched), ("per_image", per_image)]
    if not batched_first:
        order.reverse()
    out = {name: sampled(stream, fn) for name, fn in order}
    stream.synchronize()
    return out


def _protocol_child(options, shape, batched_first, conn):
    try:
        conn.send(("ok", protocol_round(options, shape, batched_first)))
    except BaseException as exc:  # reported by the parent
        conn.send(("error", repr(exc)))
    conn.close()


def protocol_rounds(ctx, rounds: int, options: dict, shape: tuple, timeout_s: float = 240.0) -> dict:
    """The reference's timing protocol (docs/benchmarking.md "Rounds"): the
    sample rule in `rounds` separate processes, the order of the two methods
    alternating between rounds, all samples merged (nearest-rank statistics
    over every sample) with each round's median kept."""
    merged = {"batched": [], "per_image": []}
    medians = {"batched": [], "per_image": []}
    reps = {"batched": [], "per_image": []}
    for r in range(rounds):
        parent, child = ctx.Pipe(duplex=False)
        proc = ctx.Process(target=_protocol_child, args=(options, shape, r % 2 == 0, child))
        proc.start()
        child.close()
        if not parent.poll(timeout_s):
            proc.kill()
            proc.join()
            raise RuntimeError(f"protocol round {r + 1} sent nothing within {timeout_s:.0f} s")
        try:
            kind, payload = parent.recv()
        except EOFError:
            proc.join(60)
            raise RuntimeError(f"protocol round {r + 1} ended with exit code {proc.exitcode} and no result")
        proc.join(60)
        if kind != "ok":
            raise RuntimeError(f"protocol round {r + 1} failed: {payload}")
        for name in merged:
            merged[name] += payload[name]["per_call_ms"]
            medians[name].append(payload[name]["p50_ms"])
            reps[name].append(payload[name]["calls_per_sample"])
        log(f"protocol round {r + 1}/{rounds}: batched p50 {payload['batched']['p50_ms']} ms, "
            f"per-image p50 {payload['per_image']['p50_ms']} ms")
    out = {name: {"samples": len(merged[name]), "calls_per_sample": min(reps[name]), **rank_stats(merged[name]),
                  "round_p50_ms": medians[name]} for name in merged}
    out["rounds"] = rounds
    return out


def timed(stream, fn, steps: int):
    from tilestream_amd import api
    start, stop = api.Event(), api.Event()
    stream.synchronize()
    start.record(stream)
    for _ in range(steps):
        fn()
    stop.record(stream)
    stop.synchronize()
    return start.elapsed_ms(stop) / steps


def host_cores() -> dict:
    """The host's logical CPU count and the CPUs this process may run on (the
    baseline's thread count is min(16, the latter))."""
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count() or 1
    return {"host_logical_cpus": os.cpu_count(), "host_cpus_usable": usable}


def cpu_baseline(batch, threads: int, single_budget_bytes: int):
    """The reference CPU codec (oracle/_ref, compiled from the reference sources)
    or, where it is absent, the C oracle: whole-image decode of the same archives."""
    from oracle import oracle
    ref = oracle.reference() is not None
    kind = "reference" if ref else "port"

    def decode(e):
        if ref:
            oracle.ref_decompress(e.archive, int(e.ptype), e.width, e.height)
        else:
            oracle.decompress(e.archive)
        return e.raw_bytes

    entries = batch.entries
    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(threads) as pool:
        total = sum(pool.map(decode, entries))
    multi = time.perf_counter() - t0
    sample, acc = [], 0
    for e in entries:
        if acc >= single_budget_bytes:
            break
        sample.append(e)
        acc += e.raw_bytes
    t0 = time.perf_counter()
    single_bytes = sum(decode(e) for e in sample)
    single = time.perf_counter() - t0
    return {
        "value": round(total / multi / 1e9, 3), "unit": "GB/s", "cores": threads, "kind": kind, **host_cores(),
        "sample": f"whole-image decode of all {len(entries)} archives of the batch "
                  f"({total / 1e9:.2f} GB decoded) on {threads} threads, one image per thread",
        "single_thread": {"value": round(single_bytes / single / 1e9, 3), "unit": "GB/s", "cores": 1,
                          "sample": f"first {len(sample)} images ({single_bytes / 1e9:.2f} GB decoded)"},
    }


def profiled_traffic(name="decode_batch256_kernel_trace.md", key="traffic_bytes_per_launch"):
    """HBM bytes per launch (or step) of a default workload, measured with
    rocprofv3 FETCH_SIZE and WRITE_SIZE counters (scaled by a copy kernel of
    known size) and recorded in the committed profile summary; None when the
    summary is absent."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        with open(path) as f:
            for line in f:
                if line.startswith(key + ":"):
                    return int(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def copy_ceiling(stream, total_bytes: int, steps: int = 5):
    """HBM rate of a plain streaming copy moving `total_bytes` (half read, half
    written) in the fastest copy shape measured here (tsk_stream_copy): the
    practical ceiling for a kernel that reads and writes about as much."""
    from tilestream_amd import api
    from tilestream_amd import _native as N
    half = (total_bytes // 2) // 4096 * 4096
    a, b = api.DeviceMemory(half), api.DeviceMemory(half)
    fn = lambda: N.lib().tsk_stream_copy(b.ptr, a.ptr, half // 16, stream.handle)  # noqa: E731
    fn()
    ms = timed(stream, fn, steps)
    return 2 * half / (ms * 1e-3) / 1e9


def stream_ceilings(stream, read_bytes: int, write_bytes: int, steps: int = 5) -> dict:
    """The one-sided HBM ceilings beside the copy: a read-only and a write-only
    stream (tsk_stream_read / tsk_stream_write, each in its fastest shape
    measured here) over 4 GiB, and the rate a kernel with this mix of reads
    and writes would reach if its reads ran at the read-only rate and its
    writes at the write-only rate one after the other (`mix_bound`)."""
    from tilestream_amd import api
    from tilestream_amd import _native as N
    size = 4 << 30
    a, sink = api.DeviceMemory(size), api.DeviceMemory(256)
    a.fill(1)
    rd = lambda: N.lib().tsk_stream_read(a.ptr, size // 16, sink.ptr, stream.handle)  # noqa: E731
    wr = lambda: N.lib().tsk_stream_write(a.ptr, size // 16, 7, stream.handle)  # noqa: E731
    rd()
    wr()
    read_gbps = size / (timed(stream, rd, steps) * 1e-3) / 1e9
    write_gbps = size / (timed(stream, wr, steps) * 1e-3) / 1e9
    seconds = read_bytes / (read_gbps * 1e9) + write_bytes / (write_gbps * 1e9)
    return {"read_only": round(read_gbps, 1), "write_only": round(write_gbps, 1),
            "mix_bound": round((read_bytes + write_bytes) / seconds / 1e9, 1)}


WARM_MS = 0.0  # --warm-ms: settle() before each measured section


def settle() -> None:
    if WARM_MS > 0:
        gpu_warm(WARM_MS)


def gpu_warm(ms: float) -> None:
    """Runs a plain streaming copy (1 GiB each way) for `ms` of wall time so
    that clocks and power settle before anything is measured or profiled: on a
    cold GPU the first launches of a kernel run up to 12% slower for about 10
    launches (clock and power transient)."""
    from tilestream_amd import api
    from tilestream_amd import _native as N
    size = 1 << 30
    a, b, s = api.DeviceMemory(size), api.DeviceMemory(size), api.Stream()
    a.fill(0)
    end = time.perf_counter() + ms * 1e-3
    while time.perf_counter() < end:
        for _ in range(8):
            assert N.lib().tsk_stream_copy(b.ptr, a.ptr, size // 16, s.handle) == 0
        s.synchronize()


def percentile(values, q: float) -> float:
    """Nearest-rank percentile (the reference harness's convention)."""
    v = np.sort(np.asarray(values))
    k = max(1, int(np.ceil(q / 100.0 * len(v))))
    return float(v[k - 1])


def roi_leg(args, threads: int, stream):
    """ROI batch over resident archives: 4096x4096 uint16 detector archives with
    64x64 tiles, region requests with log-uniform sides in [32, 2048] at uniform
    origins, each fused with scalar calibration (x - 100) * 1.05 to FP32.
    Served by the continuous-batching server (request-mode decode)."""
    from tilestream_amd import api, generators
    from tilestream_amd.workloads import ResidentBatch

    side, tile = 4096, (64, 64)
    offset, gain = 100.0, 1.05
    t0 = time.perf_counter()
    layout = [(side, side, np.uint16, tile)] * args.roi_archives
    rb = ResidentBatch.build(layout, args.seed + 4, threads)
    aset = api.ArchiveSet()
    for e in rb.entries:
        st, _ = aset.add(rb.src, e.meta, len(e.archive), e.src_offset)
        assert st == api.Status.success, st
    assert aset.commit(stream) == api.Status.success
    rois = generators.roi_requests(args.roi_requests, args.seed + 5, side, side, archives=args.roi_archives)
    align = args.pitch_align
    offsets, pitches, total = [], [], 0
    for a, x, y, w, h in rois:
        pitch = (w * 4 + align - 1) // align * align
        offsets.append(total)
        pitches.append(pitch)
        total += (pitch * (h - 1) + w * 4 + 255) // 256 * 256
    out = api.DeviceMemory(total)
    requests = []
    for (a, x, y, w, h), off, pitch in zip(rois, offsets, pitches):
        view = api.device_view(out, api.PixelType.float32, w, h, pitch, offset=off,
                               capacity=pitch * (h - 1) + w * 4)
        requests.append(api.region_request(a, (x, y, w, h), view, True, offset, gain))
    # Algorithmic bytes: the payload of every tile a region touches (whole tiles
    # are decoded) plus the FP32 region written.
    lengths = []
    for e in rb.entries:
        meta = e.meta
        cols, rows_ = (side + tile[0] - 1) // tile[0], (side + tile[1] - 1) // tile[1]
        lengths.append(np.array([meta.tile(k).length for k in range(cols * rows_)], np.int64).reshape(rows_, cols))
    touched = 0
    for a, x, y, w, h in rois:
        touched += int(lengths[a][y // tile[1]:(y + h - 1) // tile[1] + 1, x // tile[0]:(x + w - 1) // tile[0] + 1].sum())
    decoded = sum(w * h * 2 for _, _, _, w, h in rois)
    log(f"roi: {len(rb.entries)} archives ({rb.archive_bytes / 1e9:.2f} GB), {len(rois)} requests, "
        f"{decoded / 1e9:.2f} GB decoded, {total / 1e9:.2f} GB FP32 out, set up in {time.perf_counter() - t0:.1f} s")

    shape = tuple(int(v) for v in args.roi_shape.split(","))
    # One server per configured stream count.  The first count is the headline
    # configuration; it is fixed by the command line, never chosen from the
    # results.
    stream_counts = [int(v) for v in str(args.roi_streams).split(",")]
    servers = {k: api.Server(aset, api.server_options(streams=k, max_batch=args.roi_max_batch, shape=shape))
               for k in stream_counts}
    server = servers[stream_counts[0]]
    n = len(requests)
    # Saturation: every request arrives at once; continuous batching drains them.
    sat = None
    settle()
    for _ in range(2):
        st, lat, statuses, _, stats = server.replay(requests, np.zeros(n, np.uint64))
        assert st == api.Status.success and (statuses == 0).all(), st
        sat = stats
    sat_qps = n / sat.seconds
    log(f"roi saturation: {sat.seconds * 1e3:.2f} ms, {sat_qps:.0f} requests/s, "
        f"{sat.decoded_bytes / sat.seconds / 1e9:.1f} GB/s decoded, {sat.batches} batches")
    rng = np.random.default_rng(args.seed + 6)
    latency = []
    for q in args.qps:
        arrivals = np.cumsum(rng.exponential(1e9 / q, n)).astype(np.uint64)
        by_streams = []
        for k in stream_counts:
            st, lat, statuses, _, stats = servers[k].replay(requests, arrivals)
            assert st == api.Status.success and (statuses == 0).all(), st
            rec = {"qps": q, "streams": k, "achieved_qps": round(n / (arrivals[-1] * 1e-9), 1),
                   "p50_us": round(percentile(lat, 50) / 1e3, 1), "p99_us": round(percentile(lat, 99) / 1e3, 1),
                   "max_us": round(float(lat.max()) / 1e3, 1), "batches": int(stats.batches),
                   "mean_batch": round(n / stats.batches, 1)}
            log(f"roi latency at {q} req/s, {k} streams: p50 {rec['p50_us']} us, p99 {rec['p99_us']} us, "
                f"{stats.batches} batches")
            by_streams.append(rec)
        head = dict(by_streams[0])
        if len(by_streams) > 1:
            head["other_stream_counts"] = [{kk: r[kk] for kk in ("streams", "p50_us", "p99_us", "max_us", "batches")}
                                           for r in by_streams[1:]]
        latency.append(head)

    # The same requests into outputs whose column 0 sits at the 128-byte phase
    # of its archive column in a full-frame FP32 buffer, so that tile
    # boundaries fall on cache-line boundaries (no line written by two tiles).
    phased_offsets, phased_total = [], 0
    for (a, x, y, w, h), pitch in zip(rois, pitches):
        phase = (x * 4) % 128
        base = (phased_total + 255) // 256 * 256 + phase
        phased_offsets.append(base)
        phased_total = base + pitch * (h - 1) + w * 4
    out2 = api.DeviceMemory(phased_total + 256)
    phased = []
    for (a, x, y, w, h), off, pitch in zip(rois, phased_offsets, pitches):
        view = api.device_view(out2, api.PixelType.float32, w, h, pitch, offset=off, capacity=pitch * (h - 1) + w * 4)
        phased.append(api.region_request(a, (x, y, w, h), view, True, offset, gain))
    for _ in range(2):
        st, _, statuses, _, psat = server.replay(phased, np.zeros(n, np.uint64))
        assert st == api.Status.success and (statuses == 0).all(), st
    arrivals = np.cumsum(np.random.default_rng(args.seed + 7).exponential(1e9 / args.qps[0], n)).astype(np.uint64)
    st, plat, statuses, _, _ = server.replay(phased, arrivals)
    assert st == api.Status.success and (statuses == 0).all(), st
    phased_rec = {"saturation_ms": round(psat.seconds * 1e3, 3),
                  "requests_per_s": round(n / psat.seconds, 1),
                  "algorithmic_GBps": round((touched + 4 * sum(w * h for *_, w, h in rois)) / psat.seconds / 1e9, 1),
                  "qps": args.qps[0], "p50_us": round(percentile(plat, 50) / 1e3, 1),
                  "p99_us": round(percentile(plat, 99) / 1e3, 1),
                  "note": "outputs placed so that column 0 has its archive column's 128-byte phase"}
    log(f"roi with phase-aligned outputs: {phased_rec['saturation_ms']} ms saturated, p50 {phased_rec['p50_us']} us, "
        f"p99 {phased_rec['p99_us']} us at {args.qps[0]:.0f} req/s")
    del out2, phased

    # Reference-style path on a bounded sample: one prepare_batch + one
    # decode_calibrate_async per request, synchronized (closed loop).
    plain_lat = []
    ws = api.DeviceMemory(1 << 20)
    res = api.DeviceMemory(16)
    sample = min(args.roi_plain_sample, n)
    for i in range(sample):
        a, x, y, w, h = rois[i]
        e = rb.entries[a]
        it = rb.item(a, (x, y, w, h))
        it.output = requests[i].output
        it.calibration.offset = api.scalar(offset)
        it.calibration.gain = api.scalar(gain)
        t1 = time.perf_counter()
        st, plan = api.prepare_batch(api.Operation.decode_calibrate, [it], stream=stream)
        assert st == api.Status.success, st
        if plan.workspace_bytes() > ws.size:
            ws = api.DeviceMemory(plan.workspace_bytes())
        assert api.decode_calibrate_async(plan, ws, res, 1, stream) == api.Status.success
        stream.synchronize()
        plain_lat.append(time.perf_counter() - t1)
    plain = None
    if sample > 0:
        plain = {"requests": sample, "p50_us": round(percentile(plain_lat, 50) * 1e6, 1),
                 "p99_us": round(percentile(plain_lat, 99) * 1e6, 1),
                 "requests_per_s": round(sample / sum(plain_lat), 1),
                 "note": "one prepare_batch + decode_calibrate_async + stream sync per request, host planning included"}
        log(f"roi per-request path: p50 {plain['p50_us']} us, p99 {plain['p99_us']} us, "
            f"{plain['requests_per_s']} requests/s")

    verified = None
    if not args.no_verify:
        # Re-run the saturation replay into cleared outputs, then check a sample
        # against host binary32 (x - 100) * 1.05 of the generated pixels.
        out.fill(0)
        st, _, statuses, _, _ = server.replay(requests, np.zeros(n, np.uint64))
        ok = st == api.Status.success and bool((statuses == 0).all())
        big = max(range(n), key=lambda i: rois[i][3] * rois[i][4])
        o32, g32 = np.float32(offset), np.float32(gain)
        for i in sorted(set(range(0, n, max(1, n // 200))) | {big, n - 1}):
            a, x, y, w, h = rois[i]
            want = ((rb.entries[a].pixels[y:y + h, x:x + w].astype(np.float32) - o32) * g32).view(np.uint32)
            pitch = pitches[i]
            raw = np.zeros(pitch * h, np.uint8)
            raw[:pitch * (h - 1) + w * 4] = out.download(pitch * (h - 1) + w * 4, offsets[i])
            got = np.ascontiguousarray(raw.reshape(h, pitch)[:, :w * 4]).view(np.uint32)
            ok = ok and np.array_equal(got, want)
        verified = ok
        log(f"roi verification bit-exact: {ok}")
    return {
        "workload": f"{len(rb.entries)} resident {side}x{side} uint16 archives ({tile[0]}x{tile[1]} tiles), "
                    f"{n} region requests, sides log-uniform in [32, 2048], uniform origins, fused "
                    f"(x - {offset:g}) * {gain:g} to FP32; Poisson arrivals",
        "requests": n, "decoded_bytes": decoded, "output_bytes": sum(w * h * 4 for *_, w, h in rois),
        "saturation": {"ms": round(sat.seconds * 1e3, 3), "requests_per_s": round(sat_qps, 1),
                       "decoded_GBps": round(sat.decoded_bytes / sat.seconds / 1e9, 2),
                       "batches": int(sat.batches),
                       "algorithmic_GBps": round((touched + 4 * sum(w * h for *_, w, h in rois)) / sat.seconds / 1e9, 1),
                       "touched_payload_bytes": touched},
        "output_row_alignment_bytes": align,
        "latency": latency,
        "per_request_path": plain,
        "phase_aligned_outputs": phased_rec,
        "server": {"streams": stream_counts, "max_batch": args.roi_max_batch, "shape": list(shape),
                   "note": f"every headline figure uses {stream_counts[0]} streams, fixed before the run; "
                           "other counts (if any) are listed under other_stream_counts, same arrivals"},
        "verified_bit_exact": verified,
    }


def encode_cpu_baseline(batch, threads: int, sample_bytes: int):
    """The reference CPU encoder (oracle/_ref; the C oracle where it is absent)
    on the first images of the batch up to about sample_bytes of pixels: one
    thread, then `threads` threads with one image each; the archives must equal
    the batch's."""
    from oracle import oracle
    ref = oracle.reference() is not None
    enc = oracle.ref_compress if ref else oracle.compress
    sample, acc = [], 0
    for e in batch.entries:
        if acc >= sample_bytes:
            break
        sample.append(e)
        acc += e.raw_bytes

    def one(e):
        tile = (e.tile.width, e.tile.height) if hasattr(e.tile, "width") else tuple(e.tile)
        data = enc(e.pixels, tile)
        return bytes(data) == bytes(e.archive)

    t0 = time.perf_counter()
    same = all(one(e) for e in sample)
    single = time.perf_counter() - t0
    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(threads) as pool:
        same = all(pool.map(one, sample)) and same
    multi = time.perf_counter() - t0
    return {"value": round(acc / multi / 1e9, 3), "unit": "GB/s of pixels", "cores": threads, **host_cores(),
            "kind": "reference" if ref else "port", "single_thread_GBps": round(acc / single / 1e9, 3),
            "identical_archives": same,
            "sample": f"first {len(sample)} images of the batch ({acc / 1e9:.2f} GB of pixels), one image per thread"}


def encode_leg(args, batch, stream, align=None, secondary=False):
    """GPU canonical encode of the decode batch's images (pixels resident in
    HBM, rows starting every `align` bytes: by default on 256 bytes, as the
    decode leg writes them): one compress plan per (pixel type, tile shape),
    all six run back to back per step.  Archive bytes are checked against the
    host codec's.  `secondary`: a compact record without the CPU baseline."""
    from tilestream_amd import api
    from tilestream_amd.workloads import _align
    groups = {}
    for i, e in enumerate(batch.entries):
        groups.setdefault((int(e.ptype), tuple(e.tile)), []).append(i)
    # Every input row starts on `align` bytes (1: rows packed, pitch = width x
    # sample).
    align = max(1, args.encode_pitch_align if align is None else align)
    raw_off, pitches, total = {}, {}, 0
    for e_i, e in enumerate(batch.entries):
        raw_off[e_i] = total
        pitches[e_i] = (e.width * e.sample + align - 1) // align * align
        total += _align(pitches[e_i] * (e.height - 1) + e.width * e.sample)
    raw = api.DeviceMemory(total)
    for e_i, e in enumerate(batch.entries):
        if pitches[e_i] == e.width * e.sample:
            raw.upload(e.pixels, raw_off[e_i])
        else:
            rows = np.zeros((e.height, pitches[e_i]), np.uint8)
            rows[:, :e.width * e.sample] = e.pixels.view(np.uint8).reshape(e.height, -1)
            raw.upload(rows.reshape(-1)[:pitches[e_i] * (e.height - 1) + e.width * e.sample], raw_off[e_i])
    plans, keep, out_total = [], [], 0
    bounds = {}
    for (ptype, tile), members in sorted(groups.items()):
        for i in members:
            e = batch.entries[i]
            st, bound = api.compress_bound(api.ImageSpec(ptype, e.width, e.height), api.TileShape(*tile))
            assert st == api.Status.success
            bounds[i] = (out_total, bound)
            out_total += _align(bound)
    out = api.DeviceMemory(out_total)
    ws_bytes = 0
    for (ptype, tile), members in sorted(groups.items()):
        items = []
        for i in members:
            e = batch.entries[i]
            it = api.BatchItem()
            it.input = api.device_view(raw, e.ptype, e.width, e.height, pitches[i], offset=raw_off[i],
                                       capacity=pitches[i] * (e.height - 1) + e.width * e.sample, const=True)
            off, bound = bounds[i]
            it.compressed_output = api.MutableBytes(out.ptr + off, bound)
            items.append(it)
        st, plan = api.prepare_batch(api.Operation.compress, items, api.TileShape(*tile), stream=stream)
        assert st == api.Status.success, st
        ws_bytes = max(ws_bytes, plan.workspace_bytes())
        res = api.DeviceMemory(16 * len(members))
        plans.append((plan, res, members))
        keep.append(items)
    ws = api.DeviceMemory(ws_bytes)

    def step():
        for plan, res, members in plans:
            api.compress_async(plan, ws, res, len(members), stream)

    settle()
    for _ in range(args.warmup):
        step()
    ms = timed(stream, step, args.steps)
    stream.synchronize()
    ok, written_total = True, 0
    for plan, res, members in plans:
        status, written = api.read_results(res, len(members))
        ok = ok and bool((status == 0).all())
        for i, n in zip(members, written):
            written_total += int(n)
            ok = ok and int(n) == len(batch.entries[i].archive)
    if not args.no_verify:
        for i in range(0, len(batch.entries), 16):
            off, _ = bounds[i]
            ok = ok and out.download(len(batch.entries[i].archive), off).tobytes() == batch.entries[i].archive
    raw_bytes = batch.decoded_bytes
    alg = raw_bytes + written_total
    log(f"encode (input rows on {align} bytes): {ms:.4f} ms per step, {raw_bytes / (ms * 1e-3) / 1e9:.1f} GB/s "
        f"of pixels, {alg / (ms * 1e-3) / 1e9:.1f} GB/s algorithmic, verified {ok}")
    achieved = alg / (ms * 1e-3) / 1e9
    if secondary:
        return {"input_row_alignment_bytes": align, "ms_per_step": round(ms, 4),
                "pixels_GBps": round(raw_bytes / (ms * 1e-3) / 1e9, 1), "algorithmic_GBps": round(achieved, 1),
                "frac": round(achieved / HBM_PEAK_GBPS, 4), "verified_bit_exact": ok}
    cpu = None if args.no_cpu else encode_cpu_baseline(batch, 16, 500_000_000)
    default_workload = args.images == 256 and args.seed == 2026517 and align == 256
    return {"workload": f"encode of the {len(batch.entries)} batch images, one plan per (pixel type, tile shape), "
                        f"{len(plans)} plans per step", "ms_per_step": round(ms, 4),
            "input_row_alignment_bytes": align,
            "pixels_GBps": round(raw_bytes / (ms * 1e-3) / 1e9, 1),
            "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                         "frac": round(achieved / HBM_PEAK_GBPS, 4),
                         "traffic": profiled_traffic("encode_batch256_kernel_trace.md", "traffic_bytes_per_step")
                         if default_workload else None,
                         "algorithmic_bytes_per_step": alg},
            "archive_bytes": written_total,
            "cpu_baseline": cpu,
            "note": "algorithmic bytes = pixels read once + archive bytes written; traffic = HBM bytes per step "
                    "from rocprofv3 FETCH_SIZE/WRITE_SIZE (profiles/encode_batch256_kernel_trace.md)",
            "verified_bit_exact": ok}


def serve_mix_leg(args, threads: int, stream):
    """Serving mix: Poisson arrivals of region decodes (70%: sides log-uniform in
    [32, 2048] clipped to the image, uniform origins), whole-image decodes (20%)
    and GPU encodes (10%) over images whose width and height are drawn from
    {512, 1024, 2048, 4096, 8192}, 50/50 uint16 / float32, tiles from {128x32,
    64x64, 256x16}; archives and encode inputs resident in HBM; 200 GB
    admission cap on the output bytes in flight."""
    from tilestream_amd import api
    from tilestream_amd.workloads import ResidentBatch, _align

    rng = np.random.default_rng(args.seed + 20)
    sides = (512, 1024, 2048, 4096, 8192)
    tiles = ((128, 32), (64, 64), (256, 16))
    t0 = time.perf_counter()
    layout = [(int(rng.choice(sides)), int(rng.choice(sides)), np.uint16 if k % 2 == 0 else np.uint32,
               tiles[int(rng.integers(0, len(tiles)))]) for k in range(args.mix_images)]
    rb = ResidentBatch.build(layout, args.seed + 21, threads)
    aset = api.ArchiveSet()
    for e in rb.entries:
        st, _ = aset.add(rb.src, e.meta, len(e.archive), e.src_offset)
        assert st == api.Status.success, st
    assert aset.commit(stream) == api.Status.success
    raw_offsets, total = [], 0
    for e in rb.entries:
        raw_offsets.append(total)
        total += _align(e.raw_bytes)
    raw = api.DeviceMemory(total)
    for e, off in zip(rb.entries, raw_offsets):
        raw.upload(e.pixels, off)

    n = args.mix_requests
    kinds = rng.choice(3, size=n, p=[0.7, 0.2, 0.1])  # 0 region, 1 whole, 2 encode
    images = rng.integers(0, len(rb.entries), n)
    lo, hi = np.log(32.0), np.log(2048.0)
    specs, out_total, enc_total = [], 0, 0
    for k, a in zip(kinds, images):
        e = rb.entries[int(a)]
        if k == 2:
            st, bound = api.compress_bound(api.ImageSpec(int(e.ptype), e.width, e.height), api.TileShape(*e.tile))
            assert st == api.Status.success
            specs.append((2, int(a), None, enc_total, bound, 0))
            enc_total += _align(bound)
            continue
        if k == 1:
            x, y, w, h = 0, 0, e.width, e.height
        else:
            w = min(e.width, int(np.exp(rng.uniform(lo, hi))))
            h = min(e.height, int(np.exp(rng.uniform(lo, hi))))
            x, y = int(rng.integers(0, e.width - w + 1)), int(rng.integers(0, e.height - h + 1))
        pitch = (w * e.sample + args.pitch_align - 1) // args.pitch_align * args.pitch_align
        span = pitch * (h - 1) + w * e.sample
        specs.append((int(k), int(a), (x, y, w, h), out_total, span, pitch))
        out_total += _align(span)
    out = api.DeviceMemory(max(256, out_total))
    enc = api.DeviceMemory(max(256, enc_total))
    requests = []
    for k, a, roi, off, size, pitch in specs:
        e = rb.entries[a]
        if k == 2:
            view = api.device_view(raw, e.ptype, e.width, e.height, offset=raw_offsets[a], capacity=e.raw_bytes,
                                   const=True)
            requests.append(api.encode_request(view, e.tile, api.MutableBytes(enc.ptr + off, size)))
        else:
            view = api.device_view(out, e.ptype, roi[2], roi[3], pitch, offset=off, capacity=size)
            requests.append(api.region_request(a, roi, view))
    counts = [int(np.sum(kinds == k)) for k in range(3)]
    log(f"serve mix: {len(rb.entries)} images ({rb.decoded_bytes / 1e9:.2f} GB raw, {rb.archive_bytes / 1e9:.2f} GB "
        f"archives), {n} requests ({counts[0]} region, {counts[1]} whole, {counts[2]} encode), "
        f"{out_total / 1e9:.1f} GB decode outputs, {enc_total / 1e9:.1f} GB encode outputs, "
        f"set up in {time.perf_counter() - t0:.1f} s")

    shape = tuple(int(v) for v in args.roi_shape.split(","))
    opts = api.server_options(streams=args.mix_streams, max_batch=args.roi_max_batch,
                              admission_bytes=200 * 10**9, shape=shape)
    server = api.Server(aset, opts)
    sat = None
    for _ in range(2):
        st, lat, statuses, written, stats = server.replay(requests, np.zeros(n, np.uint64))
        assert st == api.Status.success and (statuses == 0).all(), st
        sat = stats
    log(f"serve mix saturation: {sat.seconds * 1e3:.2f} ms, {n / sat.seconds:.0f} requests/s, "
        f"{sat.decoded_bytes / sat.seconds / 1e9:.1f} GB/s, {sat.batches} batches")
    arrivals_rng = np.random.default_rng(args.seed + 22)
    latency = []
    for q in args.mix_qps:
        arrivals = np.cumsum(arrivals_rng.exponential(1e9 / q, n)).astype(np.uint64)
        st, lat, statuses, _, stats = server.replay(requests, arrivals)
        assert st == api.Status.success and (statuses == 0).all(), st
        rec = {"qps": q, "achieved_qps": round(n / (arrivals[-1] * 1e-9), 1),
               "p50_us": round(percentile(lat, 50) / 1e3, 1), "p99_us": round(percentile(lat, 99) / 1e3, 1),
               "batches": int(stats.batches), "GBps": round(stats.decoded_bytes / stats.seconds / 1e9, 1)}
        for k, name in enumerate(("region", "whole", "encode")):
            sel = lat[kinds == k]
            if sel.size:
                rec[name] = {"p50_us": round(percentile(sel, 50) / 1e3, 1),
                             "p99_us": round(percentile(sel, 99) / 1e3, 1)}
        log(f"serve mix at {q} req/s: p50 {rec['p50_us']} us, p99 {rec['p99_us']} us; region p99 "
            f"{rec['region']['p99_us']} us, whole p99 {rec['whole']['p99_us']} us, "
            f"encode p99 {rec['encode']['p99_us']} us")
        latency.append(rec)

    # The same requests under an admission cap that binds: an eighth of the
    # replay's output bytes (never below the largest single request).
    sizes = [size for *_, size, _ in specs]
    cap = max(max(sizes), (out_total + enc_total) // 8)
    bounded = api.Server(aset, api.server_options(streams=args.mix_streams, max_batch=args.roi_max_batch,
                                                  admission_bytes=cap, shape=shape))
    for _ in range(2):
        st, _, statuses, _, bsat = bounded.replay(requests, np.zeros(n, np.uint64))
        assert st == api.Status.success and (statuses == 0).all(), st
    arrivals = np.cumsum(np.random.default_rng(args.seed + 23).exponential(1e9 / args.mix_qps[0], n)).astype(np.uint64)
    st, blat, statuses, _, bstats = bounded.replay(requests, arrivals)
    assert st == api.Status.success and (statuses == 0).all(), st
    admission = {"admission_bytes": int(cap), "saturation_ms": round(bsat.seconds * 1e3, 3),
                 "requests_per_s": round(n / bsat.seconds, 1), "GBps": round(bsat.decoded_bytes / bsat.seconds / 1e9, 2),
                 "batches": int(bsat.batches), "admission_holds": int(bsat.admission_holds),
                 "peak_in_flight_bytes": int(bsat.peak_in_flight),
                 "qps": args.mix_qps[0], "p50_us": round(percentile(blat, 50) / 1e3, 1),
                 "p99_us": round(percentile(blat, 99) / 1e3, 1), "holds_at_qps": int(bstats.admission_holds),
                 "note": "admission cap = max(largest request, 1/8 of the replay's output bytes); saturation "
                         "and Poisson replays of the same requests"}
    log(f"serve mix with a {cap / 1e9:.2f} GB admission cap: {admission['saturation_ms']} ms saturated, "
        f"{bsat.admission_holds} holds, peak {bsat.peak_in_flight / 1e9:.2f} GB in flight; p99 "
        f"{admission['p99_us']} us at {args.mix_qps[0]:.0f} req/s")

    verified = None
    if not args.no_verify:
        # Replay into cleared outputs (once without and once with the binding
        # cap); check every encode's size, a sample of encodes byte for byte
        # against the host codec's archive, and a sample of decodes against
        # the generated pixels.
        ok = True
        for srv in (server, bounded):
            ok = ok and verify_mix(srv, requests, specs, rb, out, enc)
        verified = bool(ok)
        log(f"serve mix verification bit-exact: {verified}")
    return {
        "workload": f"{n} requests (70% region decode, sides log-uniform in [32, 2048]; 20% whole decode; "
                    f"10% GPU encode) over {len(rb.entries)} images with sides from {{512..8192}}, 50/50 "
                    f"uint16/float32, tiles from {{128x32, 64x64, 256x16}}; Poisson arrivals; admission cap "
                    f"200 GB of output bytes in flight (never binding); admission_bound: the same with a cap "
                    f"that binds",
        "requests": {"region": counts[0], "whole": counts[1], "encode": counts[2]},
        "images_raw_bytes": rb.decoded_bytes,
        "saturation": {"ms": round(sat.seconds * 1e3, 3), "requests_per_s": round(n / sat.seconds, 1),
                       "GBps": round(sat.decoded_bytes / sat.seconds / 1e9, 2), "batches": int(sat.batches),
                       "bytes": int(sat.decoded_bytes), "admission_holds": int(sat.admission_holds),
                       "note": "bytes = pixels decoded (stored sample size) + pixels encoded"},
        "latency": latency,
        "admission_bound": admission,
        "server": {"streams": args.mix_streams, "max_batch": args.roi_max_batch, "shape": list(shape),
                   "large_bytes": int(opts.large_bytes),
                   "note": "requests with at least large_bytes of output batch apart, on streams - 1 at most, "
                           "with (streams - 1) / streams of the decode grid"},
        "verified_bit_exact": verified,
    }


def verify_mix(server, requests, specs, rb, out, enc) -> bool:
    """One saturation replay of the mix into cleared outputs, checked against
    the generated pixels and the host codec's archives (a sample of each kind)."""
    from tilestream_amd import api
    n = len(requests)
    out.fill(0)
    enc.fill(0)
    st, _, statuses, written, _ = server.replay(requests, np.zeros(n, np.uint64))
    ok = st == api.Status.success and bool((statuses == 0).all())
    checked = {0: 0, 1: 0, 2: 0}
    limit = {0: 200, 1: 12, 2: 12}
    for i, (k, a, roi, off, size, pitch) in enumerate(specs):
        e = rb.entries[a]
        if k == 2:
            ok = ok and int(written[i]) == len(e.archive)
            if checked[2] < limit[2]:
                ok = ok and enc.download(len(e.archive), off).tobytes() == e.archive
                checked[2] += 1
            continue
        if checked[k] >= limit[k]:
            continue
        x, y, w, h = roi
        buf = np.zeros(pitch * h, np.uint8)
        buf[:size] = out.download(size, off)
        got = np.ascontiguousarray(buf.reshape(h, pitch)[:, :w * e.sample]).view(e.pixels.dtype)
        ok = ok and np.array_equal(got, e.pixels[y:y + h, x:x + w])
        checked[k] += 1
    log(f"serve mix check (admission {server.options.admission_bytes / 1e9:.2f} GB): {ok} (checked {checked})")
    return bool(ok)


def staged_leg(args, threads: int, stream):
    """Region requests over archives in pinned host memory: 4096x4096 uint16
    archives (64x64 tiles), sides log-uniform in [32, 2048], fused (x - 100) *
    1.05 to FP32.  The server stages the tiles each batch touches into device
    memory with a kernel reading over the host link while other batches
    decode.  Beside it, the reference's protocol on the same requests (host
    gather of the touched tiles into pinned staging, one H2D copy per request
    on a copy stream, prepare_batch with tile mappings, decode on a compute
    stream that waits for the copy) and a plain 1 GiB H2D copy."""
    from tilestream_amd import api, generators
    from tilestream_amd.workloads import ResidentBatch, _align

    side, tile = 4096, (64, 64)
    offset, gain = 100.0, 1.05
    t0 = time.perf_counter()
    rb = ResidentBatch.build([(side, side, np.uint16, tile)] * args.staged_archives, args.seed + 30, threads)
    host_off, total = [], 0
    for e in rb.entries:
        host_off.append(total)
        total += _align(len(e.archive))
    host = api.HostMemory(total)
    for e, off in zip(rb.entries, host_off):
        host.upload(np.frombuffer(e.archive, np.uint8), off)
    aset = api.ArchiveSet()
    for e, off in zip(rb.entries, host_off):
        st, _ = aset.add(host, e.meta, len(e.archive), off)
        assert st == api.Status.success, st
    assert aset.commit(stream) == api.Status.success
    n = args.staged_requests
    rois = generators.roi_requests(n, args.seed + 31, side, side, archives=args.staged_archives)
    align = args.pitch_align
    offsets, pitches, out_total = [], [], 0
    for a, x, y, w, h in rois:
        pitch = (w * 4 + align - 1) // align * align
        offsets.append(out_total)
        pitches.append(pitch)
        out_total += (pitch * (h - 1) + w * 4 + 255) // 256 * 256
    out = api.DeviceMemory(out_total)
    requests = []
    for (a, x, y, w, h), off, pitch in zip(rois, offsets, pitches):
        view = api.device_view(out, api.PixelType.float32, w, h, pitch, offset=off, capacity=pitch * (h - 1) + w * 4)
        requests.append(api.region_request(a, (x, y, w, h), view, True, offset, gain))
    # Tiles each request touches, their payload bytes, and the 16-byte blocks staged.
    cols = rows_ = side // tile[0]
    tile_off = [np.array([e.meta.tile(k).offset for k in range(cols * rows_)], np.int64).reshape(rows_, cols)
                for e in rb.entries]
    tile_len = [np.array([e.meta.tile(k).length for k in range(cols * rows_)], np.int64).reshape(rows_, cols)
                for e in rb.entries]
    touched, blocks = [], 0
    for a, x, y, w, h in rois:
        r0, r1, c0, c1 = y // tile[1], (y + h - 1) // tile[1] + 1, x // tile[0], (x + w - 1) // tile[0] + 1
        touched.append((r0, r1, c0, c1))
        o = tile_off[a][r0:r1, c0:c1] + host_off[a]
        ln = tile_len[a][r0:r1, c0:c1]
        blocks += int(((o + ln + 15) // 16 - o // 16).sum())
    payload = sum(int(tile_len[a][r0:r1, c0:c1].sum()) for (a, *_), (r0, r1, c0, c1) in zip(rois, touched))
    decoded = sum(w * h * 2 for *_, w, h in rois)
    log(f"staged: {len(rb.entries)} archives in pinned host memory ({rb.archive_bytes / 1e9:.2f} GB), {n} requests, "
        f"{decoded / 1e9:.2f} GB decoded, {payload / 1e9:.3f} GB of touched payload, set up in "
        f"{time.perf_counter() - t0:.1f} s")
    del rb.src  # the device copy of the archives is not used here

    shape = tuple(int(v) for v in args.roi_shape.split(","))
    server = api.Server(aset, api.server_options(streams=args.mix_streams, max_batch=args.roi_max_batch,
                                                 staging_bytes=args.staging_mib << 20, shape=shape))
    sat = None
    settle()
    for _ in range(2):
        st, lat, statuses, _, stats = server.replay(requests, np.zeros(n, np.uint64))
        assert st == api.Status.success and (statuses == 0).all(), st
        sat = stats
    log(f"staged saturation: {sat.seconds * 1e3:.2f} ms, {n / sat.seconds:.0f} requests/s, "
        f"{sat.decoded_bytes / sat.seconds / 1e9:.1f} GB/s decoded, {blocks * 16 / sat.seconds / 1e9:.1f} GB/s "
        f"over the host link, {sat.batches} batches")
    rng = np.random.default_rng(args.seed + 32)
    latency = []
    for q in args.staged_qps:
        arrivals = np.cumsum(rng.exponential(1e9 / q, n)).astype(np.uint64)
        st, lat, statuses, _, stats = server.replay(requests, arrivals)
        assert st == api.Status.success and (statuses == 0).all(), st
        rec = {"qps": q, "achieved_qps": round(n / (arrivals[-1] * 1e-9), 1),
               "p50_us": round(percentile(lat, 50) / 1e3, 1), "p99_us": round(percentile(lat, 99) / 1e3, 1),
               "batches": int(stats.batches)}
        log(f"staged latency at {q} req/s: p50 {rec['p50_us']} us, p99 {rec['p99_us']} us")
        latency.append(rec)

    # Plain H2D copy of 1 GiB of pinned memory: the host link's ceiling here.
    big = api.HostMemory(1 << 30)
    dev = api.DeviceMemory(1 << 30)
    copy_ms = timed(stream, lambda: api.memcpy_async(dev.ptr, big.ptr, 1 << 30, stream), 3)
    del big, dev
    link = (1 << 30) / (copy_ms * 1e-3) / 1e9

    # The reference's protocol on a sample (benchmarks/codec_benchmark.cpp archive_roi).
    sample = min(args.staged_plain_sample, n)
    sizes = [int(tile_len[a][r0:r1, c0:c1].sum()) for (a, *_), (r0, r1, c0, c1) in zip(rois[:sample], touched[:sample])]
    st_off = np.concatenate([[0], np.cumsum([_align(z) for z in sizes])]).astype(np.int64)
    staging_host = api.HostMemory(int(st_off[-1]))
    staging_dev = api.DeviceMemory(int(st_off[-1]))
    hview, sview = host.array(), staging_host.array()
    copy_stream = api.Stream()
    res = api.DeviceMemory(16 * sample)
    ws = None
    t_start = time.perf_counter()
    items, keep = [], []
    for i in range(sample):
        a, x, y, w, h = rois[i]
        r0, r1, c0, c1 = touched[i]
        at, maps = int(st_off[i]), []
        for r in range(r0, r1):
            for c in range(c0, c1):
                o, ln = int(tile_off[a][r, c]) + host_off[a], int(tile_len[a][r, c])
                sview[at:at + ln] = hview[o:o + ln]
                maps.append(api.TileMapping(r * cols + c, at - int(st_off[i])))
                at += ln
        it = api.BatchItem()
        it.roi = api.Roi(x, y, w, h)
        it.output = requests[i].output
        it.calibration.offset = api.scalar(offset)
        it.calibration.gain = api.scalar(gain)
        arr = (api.TileMapping * len(maps))(*maps)
        it.compressed_input.bytes = api.ConstBytes(staging_dev.ptr + int(st_off[i]), sizes[i])
        it.compressed_input.metadata = rb.entries[a].meta.handle
        it.compressed_input.mappings = C_cast_mappings(arr)
        it.compressed_input.mapping_count = len(maps)
        items.append(it)
        keep.append(arr)
    t_gather = time.perf_counter()
    plans = []
    for it in items:
        st, plan = api.prepare_batch(api.Operation.decode_calibrate, [it], stream=stream)
        assert st == api.Status.success, st
        plans.append(plan)
        if ws is None or plan.workspace_bytes() > ws.size:
            ws = api.DeviceMemory(max(256, plan.workspace_bytes()))
    t_prepare = time.perf_counter()
    copied = [api.Event() for _ in range(sample)]
    first, last = api.Event(), api.Event()
    first.record(copy_stream)
    for i in range(sample):
        api.memcpy_async(staging_dev.ptr + int(st_off[i]), staging_host.ptr + int(st_off[i]), sizes[i], copy_stream)
        copied[i].record(copy_stream)
        stream.wait(copied[i])
        assert api.decode_calibrate_async(plans[i], ws, res.ptr + 16 * i, 1, stream) == api.Status.success
    last.record(stream)
    last.synchronize()
    t_end = time.perf_counter()
    plain_decoded = sum(w * h * 2 for *_, w, h in rois[:sample])
    plain = {"requests": sample, "host_end_to_end_ms": round((t_end - t_start) * 1e3, 2),
             "gather_ms": round((t_gather - t_start) * 1e3, 2), "prepare_ms": round((t_prepare - t_gather) * 1e3, 2),
             "device_ms": round(first.elapsed_ms(last), 3),
             "requests_per_s": round(sample / (t_end - t_start), 1),
             "decoded_GBps_end_to_end": round(plain_decoded / (t_end - t_start) / 1e9, 2),
             "decoded_GBps_device": round(plain_decoded / (first.elapsed_ms(last) * 1e-3) / 1e9, 2),
             "note": "host gather (NumPy slices) + prepare_batch with tile mappings + per-request H2D on a copy "
                     "stream overlapped with decode on the compute stream"}
    status, _ = api.read_results(res, sample)
    log(f"staged reference protocol ({sample} requests): {plain['host_end_to_end_ms']} ms end to end "
        f"({plain['requests_per_s']} req/s), device span {plain['device_ms']} ms; host link {link:.1f} GB/s")

    verified = None
    if not args.no_verify:
        out.fill(0)
        st, _, statuses, _, _ = server.replay(requests, np.zeros(n, np.uint64))
        ok = st == api.Status.success and bool((statuses == 0).all()) and bool((status == 0).all())
        o32, g32 = np.float32(offset), np.float32(gain)
        for i in sorted(set(range(0, n, max(1, n // 100))) | {n - 1}):
            a, x, y, w, h = rois[i]
            want = ((rb.entries[a].pixels[y:y + h, x:x + w].astype(np.float32) - o32) * g32).view(np.uint32)
            pitch = pitches[i]
            raw = np.zeros(pitch * h, np.uint8)
            raw[:pitch * (h - 1) + w * 4] = out.download(pitch * (h - 1) + w * 4, offsets[i])
            got = np.ascontiguousarray(raw.reshape(h, pitch)[:, :w * 4]).view(np.uint32)
            ok = ok and np.array_equal(got, want)
        verified = bool(ok)
        log(f"staged verification bit-exact: {verified}")
    return {
        "workload": f"{len(rb.entries)} 4096x4096 uint16 archives (64x64 tiles) in pinned host memory, {n} region "
                    f"requests, sides log-uniform in [32, 2048], fused (x - {offset:g}) * {gain:g} to FP32",
        "requests": n, "decoded_bytes": decoded, "touched_payload_bytes": payload, "staged_bytes": blocks * 16,
        "saturation": {"ms": round(sat.seconds * 1e3, 3), "requests_per_s": round(n / sat.seconds, 1),
                       "decoded_GBps": round(sat.decoded_bytes / sat.seconds / 1e9, 2),
                       "host_link_GBps": round(blocks * 16 / sat.seconds / 1e9, 2), "batches": int(sat.batches)},
        "host_link_copy_GBps": round(link, 2),
        "latency": latency,
        "reference_protocol": plain,
        "server": {"streams": args.mix_streams, "max_batch": args.roi_max_batch, "staging_mib": args.staging_mib},
        "verified_bit_exact": verified,
    }


def file_leg(args, stream):
    """File in the page cache to calibrated FP32 on the device, per the
    reference's file tier (raw file: read, H2D, calibrate; .tsc file: read,
    parse and plan, H2D, decode fused with (x - 100) * 1.25; TIFF file: read
    and decode on one host thread, H2D from pageable memory, calibrate): a
    4096x4096 uint16 detector frame and a 4096x4096 float32 field, 128x32
    tiles (TIFF: Deflate level 6, the horizontal predictor for uint16 only).
    Each stage is timed on the host clock (device stages end with a stream
    synchronize), median of the repetitions; the results are compared word
    for word."""
    import tempfile
    from tilestream_amd import api, generators
    side, tile = 4096, (128, 32)
    offset, gain = 100.0, 1.25
    reps = max(3, args.file_reps)
    out = {}
    with tempfile.TemporaryDirectory(dir="/tmp") as d:
        for kind, dtype in (("u16", np.uint16), ("f32", np.uint32)):
            px = generators.image_for(side, side, dtype, args.seed + 40)
            st, archive = api.cpu_compress(px, api.TileShape(*tile))
            assert st == api.Status.success, st
            raw_path, tsc_path = os.path.join(d, f"{kind}.raw"), os.path.join(d, f"{kind}.tsc")
            px.tofile(raw_path)
            with open(tsc_path, "wb") as f:
                f.write(archive)
            tiff_path = os.path.join(d, f"{kind}.tif")
            assert api.write_tiff(tiff_path, px, api.tiff_write_options(
                compression=1, horizontal_predictor=1 if kind == "u16" else 0, tile_width=tile[0],
                tile_height=tile[1], deflate_level=6)) == api.Status.success
            tiff_bytes = os.path.getsize(tiff_path)
            ptype = api.pixel_type_of(px.dtype)
            host = api.HostMemory(max(px.nbytes, len(archive)))
            hv = host.array()
            pixels = api.DeviceMemory(px.nbytes)
            src = api.DeviceMemory(len(archive))
            out_raw, out_tsc = api.DeviceMemory(side * side * 4), api.DeviceMemory(side * side * 4)
            res = api.DeviceMemory(16)

            def read(path, n):
                with open(path, "rb", buffering=0) as f:
                    got = f.readinto(memoryview(hv)[:n])
                assert got == n

            read(raw_path, px.nbytes)  # warm page cache
            read(tsc_path, len(archive))
            # Raw path: the calibration plan is made once, outside the timed stages.
            it = api.BatchItem()
            it.input = api.device_view(pixels, ptype, side, side, const=True)
            it.output = api.device_view(out_raw, api.PixelType.float32, side, side)
            it.calibration.offset = api.scalar(offset)
            it.calibration.gain = api.scalar(gain)
            st, cal_plan = api.prepare_batch(api.Operation.calibrate, [it], stream=stream)
            assert st == api.Status.success, st
            ws = api.DeviceMemory(max(256, cal_plan.workspace_bytes()))
            raw_t = {"read": [], "h2d": [], "calibrate": []}
            for _ in range(reps):
                t0 = time.perf_counter()
                read(raw_path, px.nbytes)
                t1 = time.perf_counter()
                api.memcpy_async(pixels.ptr, host.ptr, px.nbytes, stream)
                stream.synchronize()
                t2 = time.perf_counter()
                assert api.calibrate_async(cal_plan, ws, res, 1, stream) == api.Status.success
                stream.synchronize()
                t3 = time.perf_counter()
                raw_t["read"].append(t1 - t0)
                raw_t["h2d"].append(t2 - t1)
                raw_t["calibrate"].append(t3 - t2)
            tsc_t = {"read": [], "parse_prepare": [], "h2d": [], "decode_calibrate": []}
            ws2 = None
            for _ in range(reps):
                t0 = time.perf_counter()
                read(tsc_path, len(archive))
                t1 = time.perf_counter()
                st, meta = api.parse_metadata(hv[:len(archive)])
                assert st == api.Status.success, st
                it2 = api.BatchItem()
                it2.compressed_input.bytes = api.ConstBytes(src.ptr, len(archive))
                it2.compressed_input.metadata = meta.handle
                it2.roi = api.Roi(0, 0, side, side)
                it2.output = api.device_view(out_tsc, api.PixelType.float32, side, side)
                it2.calibration.offset = api.scalar(offset)
                it2.calibration.gain = api.scalar(gain)
                st, plan = api.prepare_batch(api.Operation.decode_calibrate, [it2], stream=stream)
                assert st == api.Status.success, st
                if ws2 is None:
                    ws2 = api.DeviceMemory(max(256, plan.workspace_bytes()))
                t2 = time.perf_counter()
                api.memcpy_async(src.ptr, host.ptr, len(archive), stream)
                stream.synchronize()
                t3 = time.perf_counter()
                assert api.decode_calibrate_async(plan, ws2, res, 1, stream) == api.Status.success
                stream.synchronize()
                t4 = time.perf_counter()
                tsc_t["read"].append(t1 - t0)
                tsc_t["parse_prepare"].append(t2 - t1)
                tsc_t["h2d"].append(t3 - t2)
                tsc_t["decode_calibrate"].append(t4 - t3)
            tiff_t = {"read_decode": [], "h2d": [], "calibrate": []}
            out_tiff = api.DeviceMemory(side * side * 4)
            it3 = api.BatchItem()
            it3.input = api.device_view(pixels, ptype, side, side, const=True)
            it3.output = api.device_view(out_tiff, api.PixelType.float32, side, side)
            it3.calibration.offset = api.scalar(offset)
            it3.calibration.gain = api.scalar(gain)
            st, tiff_plan = api.prepare_batch(api.Operation.calibrate, [it3], stream=stream)
            assert st == api.Status.success, st
            for _ in range(reps):
                t0 = time.perf_counter()
                st, decoded = api.read_tiff(tiff_path)
                assert st == api.Status.success, st
                t1 = time.perf_counter()
                pixels.upload(decoded)  # from pageable memory, as the reference's TiffImage
                t2 = time.perf_counter()
                assert api.calibrate_async(tiff_plan, ws, res, 1, stream) == api.Status.success
                stream.synchronize()
                t3 = time.perf_counter()
                tiff_t["read_decode"].append(t1 - t0)
                tiff_t["h2d"].append(t2 - t1)
                tiff_t["calibrate"].append(t3 - t2)
            same = bool(np.array_equal(out_raw.download(), out_tsc.download())) and bool(
                np.array_equal(out_raw.download(), out_tiff.download()))
            med = lambda v: round(float(np.median(v)) * 1e3, 3)  # noqa: E731
            rec = {"ratio": round(px.nbytes / len(archive), 3),
                   "tiff_ratio": round(px.nbytes / tiff_bytes, 3),
                   "raw_file": {k: med(v) for k, v in raw_t.items()},
                   "tsc_file": {k: med(v) for k, v in tsc_t.items()},
                   "tiff_file": {k: med(v) for k, v in tiff_t.items()},
                   "identical_outputs": same}
            for tier in ("raw_file", "tsc_file", "tiff_file"):
                rec[tier]["host_end_to_end"] = round(sum(rec[tier].values()), 3)
            log(f"file tier {kind}: raw {rec['raw_file']['host_end_to_end']} ms, tsc {rec['tsc_file']['host_end_to_end']} "
                f"ms ({rec['tsc_file']}), tiff {rec['tiff_file']['host_end_to_end']} ms (ratio {rec['tiff_ratio']}), "
                f"identical {same}")
            out[kind] = rec
    return {"workload": "4096x4096 uint16 frame and float32 field, 128x32 tiles, file in the page cache to FP32 "
                        "(x - 100) * 1.25 on the device; raw, .tsc and TIFF (Deflate 6) files; stage medians in ms",
            **out,
            "verified_bit_exact": all(r["identical_outputs"] for r in out.values())}


def C_cast_mappings(arr):
    import ctypes
    from tilestream_amd import api
    return ctypes.cast(arr, ctypes.POINTER(api.TileMapping))


def fp64_errors(got_bits: np.ndarray, x: np.ndarray, offset: np.ndarray, gain: np.ndarray):
    """ULP error of FP32 (x - offset) * gain against the FP64 value, over
    finite normal references: max / p99 / median, and results beyond the
    1.5 ulp bound of two roundings; non-finite results must match by class."""
    xf = x.astype(np.float64) if x.dtype == np.uint16 else x.view(np.float32).astype(np.float64)
    ref = (xf - offset.astype(np.float64)) * gain.astype(np.float64)
    got = got_bits.view(np.float32).astype(np.float64)
    finite = np.isfinite(ref)
    with np.errstate(invalid="ignore"):
        ref32 = ref.astype(np.float32)
    normal = finite & (np.abs(ref32) >= np.finfo(np.float32).tiny) & np.isfinite(ref32)
    ulp = np.spacing(np.abs(ref32[normal])).astype(np.float64)
    err = np.abs(got[normal] - ref[normal]) / ulp
    nan_mismatch = int(np.sum(np.isnan(ref) != np.isnan(got)))
    inf_mismatch = int(np.sum(np.isinf(ref) & (got != ref)))
    return {"max_ulp": round(float(err.max()), 4), "p99_ulp": round(float(np.percentile(err, 99)), 4),
            "median_ulp": round(float(np.median(err)), 4), "violations_over_1_5_ulp": int(np.sum(err > 1.5)),
            "nonfinite_class_mismatches": nan_mismatch + inf_mismatch, "pixels": int(x.size)}


def fused_cpu_baseline(archive, px, dark, gain, gpu_bits):
    """The fused workload on one host thread, as the reference measures its CPU
    path for fused configs: the reference CPU codec decodes the archive
    (oracle/_ref; the C oracle where it is absent), then host binary32
    (x - offset) * gain with the same maps (two roundings, as on the GPU)."""
    from oracle import oracle
    ref = oracle.reference() is not None
    h, w = px.shape
    ptype = 1 if px.dtype == np.uint16 else 2
    reps, best = 3, None
    for _ in range(reps):
        t0 = time.perf_counter()
        dec = oracle.ref_decompress(archive, ptype, w, h) if ref else oracle.decompress(archive)
        x = dec.astype(np.float32) if dec.dtype == np.uint16 else dec.view(np.float32)
        with np.errstate(invalid="ignore", over="ignore"):
            res = (x - dark) * gain
        t = time.perf_counter() - t0
        best = t if best is None else min(best, t)
    host_bits = res.astype(np.float32).view(np.uint32).reshape(-1)
    gpu_bits = np.asarray(gpu_bits).view(np.uint32).reshape(-1)
    nan_host, nan_gpu = np.isnan(host_bits.view(np.float32)), np.isnan(gpu_bits.view(np.float32))
    same = bool(np.array_equal(nan_host, nan_gpu) and np.array_equal(host_bits[~nan_host], gpu_bits[~nan_host]))
    return {"ms": round(best * 1e3, 2), "decoded_GBps": round(px.nbytes / best / 1e9, 3), "cores": 1,
            "kind": "reference" if ref else "port",
            "equals_gpu": same, "equality": "word for word, NaN by class",
            "sample": "whole decode of the 4096x4096 archive + host binary32 calibration with both maps, "
                      "best of 3 on one thread"}


def fused_leg(args, stream):
    """Plain method at real shape: one 4096x4096 uint16 detector frame and one
    4096x4096 float32 field (64x64 tiles), whole decode fused with a per-pixel
    dark map (uint16 U[90,110], exact in FP32) and an FP32 gain map N(1, 0.02).
    One plan per image; four buffer sets rotate so the 256 MB last-level
    cache cannot hold a set between steps."""
    from tilestream_amd import api, generators
    side, tile, sets = 4096, (64, 64), 4
    out = {}
    for kind in ("u16", "f32"):
        px = generators.image_for(side, side, np.uint16 if kind == "u16" else np.uint32, args.seed + 11)
        st, archive = api.cpu_compress(px, api.TileShape(*tile))
        assert st == api.Status.success, st
        st, meta = api.parse_metadata(archive)
        dark = generators.dark_u16(side, side, args.seed + 12).astype(np.float32)
        gain = generators.gain_f32(side, side, args.seed + 13)
        keep, plans = [], []
        ws_bytes = 0
        for k in range(sets):
            src = api.DeviceMemory.from_host(np.frombuffer(archive, np.uint8))
            om, gm = api.DeviceMemory.from_host(dark), api.DeviceMemory.from_host(gain)
            dst = api.DeviceMemory(side * side * 4)
            it = api.BatchItem()
            it.compressed_input.bytes = api.ConstBytes(src.ptr, len(archive))
            it.compressed_input.metadata = meta.handle
            it.roi = api.Roi(0, 0, side, side)
            it.output = api.device_view(dst, api.PixelType.float32, side, side)
            it.calibration.offset = api.map_parameter(api.device_view(om, api.PixelType.float32, side, side, const=True))
            it.calibration.gain = api.map_parameter(api.device_view(gm, api.PixelType.float32, side, side, const=True))
            st, plan = api.prepare_batch(api.Operation.decode_calibrate, [it], stream=stream)
            assert st == api.Status.success, st
            ws_bytes = max(ws_bytes, plan.workspace_bytes())
            plans.append(plan)
            keep.append((src, om, gm, dst, it))
        ws = api.DeviceMemory(max(256, ws_bytes))
        res = api.DeviceMemory(16)
        step = [0]

        def run():
            api.decode_calibrate_async(plans[step[0] % sets], ws, res, 1, stream)
            step[0] += 1

        settle()
        for _ in range(args.warmup):
            run()
        ms = timed(stream, run, args.steps)
        stream.synchronize()
        status, written = api.read_results(res, 1)
        got = keep[(step[0] - 1) % sets][3].download().view(np.uint32).reshape(side, side)
        payload = sum(meta.tile(k).length for k in range(meta.summary.tile_count))
        alg = payload + side * side * (4 + 8)
        out[kind] = {
            "ms": round(ms, 4), "decoded_GBps": round(px.nbytes / (ms * 1e-3) / 1e9, 1),
            "algorithmic_GBps": round(alg / (ms * 1e-3) / 1e9, 1), "roofline_frac": round(alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4),
            "algorithmic_bytes": alg, "status": int(status[0]),
            "fp64_error": fp64_errors(got, px, dark, gain),
        }
        if not args.no_cpu:
            out[kind]["cpu_baseline"] = fused_cpu_baseline(archive, px, dark, gain, got)
        log(f"fused {kind}: {ms:.4f} ms, {out[kind]['algorithmic_GBps']} GB/s algorithmic, "
            f"max {out[kind]['fp64_error']['max_ulp']} ulp")
        # Standalone calibration of the raw image with the same maps (reference
        # calibrate_async): the same rotating sets; its output must equal the
        # fused decode's word for word.
        cal_plans, cal_keep = [], []
        cal_ws = 256
        ptype = api.PixelType.uint16 if kind == "u16" else api.PixelType.float32
        for k in range(sets):
            _, om, gm, _, _ = keep[k]
            raw = api.DeviceMemory.from_host(px)
            dst = api.DeviceMemory(side * side * 4)
            it = api.BatchItem()
            it.input = api.device_view(raw, ptype, side, side, const=True)
            it.output = api.device_view(dst, api.PixelType.float32, side, side)
            it.calibration.offset = api.map_parameter(api.device_view(om, api.PixelType.float32, side, side, const=True))
            it.calibration.gain = api.map_parameter(api.device_view(gm, api.PixelType.float32, side, side, const=True))
            st, plan = api.prepare_batch(api.Operation.calibrate, [it], stream=stream)
            assert st == api.Status.success, st
            cal_ws = max(cal_ws, plan.workspace_bytes())
            cal_plans.append(plan)
            cal_keep.append((raw, dst, it))
        ws2 = api.DeviceMemory(cal_ws)
        step2 = [0]

        def run_cal():
            api.calibrate_async(cal_plans[step2[0] % sets], ws2, res, 1, stream)
            step2[0] += 1

        settle()
        for _ in range(args.warmup):
            run_cal()
        cal_ms = timed(stream, run_cal, args.steps)
        stream.synchronize()
        cal_status, _ = api.read_results(res, 1)
        cal_got = cal_keep[(step2[0] - 1) % sets][1].download().view(np.uint32).reshape(side, side)
        cal_alg = side * side * (px.itemsize + 4 + 8)
        out[kind]["standalone_calibrate"] = {
            "ms": round(cal_ms, 4), "algorithmic_GBps": round(cal_alg / (cal_ms * 1e-3) / 1e9, 1),
            "roofline_frac": round(cal_alg / (cal_ms * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4), "algorithmic_bytes": cal_alg,
            "status": int(cal_status[0]), "equals_fused_bitwise": bool(np.array_equal(cal_got, got)),
            "note": "calibrate_async of the raw image with the same maps: pixels + FP32 output + two FP32 maps"}
        log(f"calibrate {kind}: {cal_ms:.4f} ms, {out[kind]['standalone_calibrate']['algorithmic_GBps']} GB/s "
            f"algorithmic, equals fused {out[kind]['standalone_calibrate']['equals_fused_bitwise']}")
        del keep, plans, cal_keep, cal_plans
    return {"workload": "4096x4096 uint16 detector frame and 4096x4096 float32 field, 64x64 tiles, whole "
                        "decode fused with dark map U[90,110] and gain map N(1,0.02), one plan per image, "
                        "4 rotating buffer sets; beside it, standalone calibration of the raw images", **out}


def main() -> int:
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--protocol-rounds", type=int, default=5,
                   help="fresh processes for the batched vs per-image sample rule (0: sample in this process)")
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--images", type=int, default=256)
    p.add_argument("--seed", type=int, default=2026517)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-verify", action="store_true")
    p.add_argument("--tune-repeats", type=int, default=3)
    p.add_argument("--profile", action="store_true",
                   help="only the batched launches (for rocprofv3): no autotune, per-image path, "
                        "CPU baseline or verification")
    p.add_argument("--shape", default="8,256,8,2", help="group,block,chunk,depth used with --profile")
    p.add_argument("--warm-ms", type=float, default=0.0,
                   help="run a streaming copy for this long before anything else (settles clocks for profiles)")
    p.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                   help="only run a plain copy kernel of this many MiB (scales HBM byte counters)")
    p.add_argument("--pitch-align", type=int, default=256,
                   help="output rows start at multiples of this many bytes, as hipMallocPitch "
                        "allocates them (1: packed rows)")
    p.add_argument("--no-roi", action="store_true", help="skip the ROI serving leg")
    p.add_argument("--fused-only", action="store_true", help="only the fused and standalone calibration leg")
    p.add_argument("--no-fused", action="store_true", help="skip the fused-calibration real-shape leg")
    p.add_argument("--roi-only", action="store_true", help="only the ROI serving leg")
    p.add_argument("--roi-archives", type=int, default=64)
    p.add_argument("--roi-requests", type=int, default=10000)
    p.add_argument("--roi-streams", default="3",
                   help="server stream counts for the ROI leg (comma-separated).  The first is fixed before "
                        "the run and gives every headline figure (saturation, p50/p99 at each rate, "
                        "verification); further counts are replayed on the same arrivals for comparison only")
    p.add_argument("--mix-streams", type=int, default=3,
                   help="server streams for the serving-mix and staged legs (3 gave the most throughput)")
    p.add_argument("--roi-max-batch", type=int, default=1024)
    p.add_argument("--roi-plain-sample", type=int, default=300)
    p.add_argument("--roi-shape", default="4,256,0,4,3", help="server launch shape group,block,grid,chunk,depth")
    p.add_argument("--qps", type=lambda v: [float(x) for x in v.split(",")], default=[100000.0, 400000.0],
                   help="Poisson arrival rates (requests/s) for the ROI latency runs")
    p.add_argument("--no-mix", action="store_true", help="skip the serving-mix leg")
    p.add_argument("--no-staged", action="store_true", help="skip the host-archive (staged) leg")
    p.add_argument("--no-file", action="store_true", help="skip the file-to-result leg")
    p.add_argument("--file-only", action="store_true", help="only the file-to-result leg")
    p.add_argument("--file-reps", type=int, default=7)
    p.add_argument("--staged-only", action="store_true", help="only the host-archive (staged) leg")
    p.add_argument("--staged-archives", type=int, default=16)
    p.add_argument("--staged-requests", type=int, default=3000)
    p.add_argument("--staged-plain-sample", type=int, default=300)
    p.add_argument("--staging-mib", type=int, default=256, help="server staging per stream")
    p.add_argument("--staged-qps", type=lambda v: [float(x) for x in v.split(",")], default=[20000.0, 80000.0],
                   help="Poisson arrival rates (requests/s) for the staged latency runs")
    p.add_argument("--no-encode", action="store_true", help="skip the encode leg")
    p.add_argument("--encode-only", action="store_true", help="only the encode leg (of the decode batch's images)")
    p.add_argument("--encode-pitch-align", type=int, default=256,
                   help="encode input rows start on this many bytes (1: packed rows)")
    p.add_argument("--mix-only", action="store_true", help="only the serving-mix leg")
    p.add_argument("--mix-images", type=int, default=24)
    p.add_argument("--mix-requests", type=int, default=3000)
    p.add_argument("--mix-qps", type=lambda v: [float(x) for x in v.split(",")], default=[20000.0, 80000.0],
                   help="Poisson arrival rates (requests/s) for the serving-mix latency runs")
    args = p.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (this codec has no "
              "multi-GPU path: shard images or requests across processes, one per GPU)", file=sys.stderr)
        return 2

    forkserver = None
    only = (args.file_only or args.staged_only or args.mix_only or args.fused_only or args.roi_only
            or args.encode_only or args.profile or args.copy_probe)
    if args.protocol_rounds > 0 and not only:
        # The protocol's fresh processes fork from this server, started before
        # anything here touches the GPU.
        import multiprocessing as mp
        from multiprocessing import forkserver as mp_forkserver
        forkserver = mp.get_context("forkserver")
        mp_forkserver.ensure_running()

    from tilestream_amd import api, generators
    from tilestream_amd.workloads import ResidentBatch

    if api.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        return 1
    global WARM_MS
    WARM_MS = args.warm_ms
    if args.copy_probe:
        from tilestream_amd import _native as N
        nbytes = args.copy_probe << 20
        a, b, s = api.DeviceMemory(nbytes), api.DeviceMemory(nbytes), api.Stream()
        a.fill(1)
        for _ in range(max(1, args.steps)):
            assert N.lib().tsk_copy_blocks(b.ptr, a.ptr, nbytes // 16, s.handle) == 0
        s.synchronize()
        log(f"copy probe: {args.steps} copies of {nbytes} bytes (read {nbytes}, write {nbytes} each)")
        return 0
    threads = max(1, min(16, host_cores()["host_cpus_usable"]))
    if args.file_only:
        fl = file_leg(args, api.Stream())
        print(json.dumps({"metric": "file in page cache to calibrated FP32", "file": fl}), flush=True)
        return 0 if fl["verified_bit_exact"] else 3
    if args.staged_only:
        staged = staged_leg(args, threads, api.Stream())
        print(json.dumps({"metric": "staged region serving from pinned host memory", "staged": staged}), flush=True)
        return 0 if staged["verified_bit_exact"] in (None, True) else 3
    if args.mix_only:
        mix = serve_mix_leg(args, threads, api.Stream())
        print(json.dumps({"metric": "serving mix throughput and latency", "serve_mix": mix}), flush=True)
        return 0 if mix["verified_bit_exact"] in (None, True) else 3
    if args.fused_only:
        fu = fused_leg(args, api.Stream())
        ok = all(v["standalone_calibrate"]["equals_fused_bitwise"] for k, v in fu.items() if k in ("u16", "f32"))
        print(json.dumps({"metric": "fused decode and calibration", "fused_real_shape": fu}), flush=True)
        return 0 if ok else 3
    if args.roi_only:
        roi = roi_leg(args, threads, api.Stream())
        print(json.dumps({"metric": "p99 ROI latency", "roi": roi}), flush=True)
        return 0 if roi["verified_bit_exact"] in (None, True) else 3
    t0 = time.perf_counter()
    layout = generators.batch_layout(args.images, args.seed)
    batch = ResidentBatch.build(layout, args.seed, threads, pitch_align=args.pitch_align)
    log(f"built {len(layout)} images, {batch.decoded_bytes / 1e9:.2f} GB decoded, "
        f"{batch.archive_bytes / 1e9:.2f} GB archives in {time.perf_counter() - t0:.1f} s")

    stream = api.Stream()
    n = len(batch.entries)
    results = api.DeviceMemory(16 * n)
    if args.encode_only:
        enc = encode_leg(args, batch, stream)
        print(json.dumps({"metric": "GPU encode GB/s", "encode": enc}), flush=True)
        return 0 if enc["verified_bit_exact"] else 3

    # ---- batched persistent path -------------------------------------------------
    items = [batch.item(i) for i in range(n)]
    st, queue = api.DecodeQueue.prepare(items, [api.Operation.decompress] * n, stream)
    assert st == api.Status.success, st
    qws = api.DeviceMemory(max(queue.workspace_bytes(), 256))
    # The tile decoder's whole space: lane-group width x workgroup size x depth,
    # each at queue chunks of 4 and 16 tiles, plus two static-stride shapes.
    candidates = [(g, b, 0, c, d) for g in (1, 2, 4, 8, 16) for b in (64, 128, 256, 512) for d in (1, 2, 3)
                  for c in (4, 16)]
    candidates += [(8, 256, 0, 0, 2), (4, 256, 0, 0, 3)]
    # The block decoder (group 0): 512 threads, two workgroups per CU.
    candidates += [(0, 512, 0, c, 2) for c in (1, 2, 4)]
    if args.profile:
        grp, blk, chk, dep = (int(v) for v in args.shape.split(","))
        assert queue.set_shape(grp, blk, 0, chk, dep) == api.Status.success
    else:
        st, times = queue.autotune(qws, results, stream, candidates, args.tune_repeats)
        assert st == api.Status.success, st
        log("autotune " + " ".join(f"{c[0]}/{c[1]}/{c[3]}/{c[4]}:{t:.3f}" for c, t in zip(candidates, times)))
    best = queue.shape
    log(f"shape group={best.group} block={best.block} chunk={best.chunk} depth={best.depth}")
    assert queue.capture(qws, results, stream) == api.Status.success
    launch = lambda: queue.launch(stream)  # noqa: E731
    settle()
    for _ in range(args.warmup):
        launch()
    ms_batched = timed(stream, launch, args.steps)
    stats = queue.stats()
    if args.profile:
        log(f"batched {ms_batched:.4f} ms per launch, "
            f"{batch.decoded_bytes / (ms_batched * 1e-3) / 1e9:.1f} GB/s decoded")
        print(json.dumps({"metric": "batched decode launches for a profiler", "ms_per_launch": round(ms_batched, 4),
                          "steps": args.steps, "warmup": args.warmup, "warm_ms": args.warm_ms,
                          "shape": [best.group, best.block, best.grid, best.chunk, best.depth, best.run],
                          "tiles": int(stats.tasks), "payload_bytes": int(stats.payload_bytes),
                          "output_bytes": int(stats.output_bytes), "map_bytes": int(stats.map_bytes)}), flush=True)
        return 0

    # ---- the same launch into packed output rows (pitch = width x sample) -----------
    packed = None
    if args.pitch_align != 1:
        from tilestream_amd.workloads import _align
        offsets, total = [], 0
        for e in batch.entries:
            offsets.append(total)
            total += _align(e.raw_bytes)
        out2 = api.DeviceMemory(total)
        items2 = []
        for e, off in zip(batch.entries, offsets):
            it = batch.item(batch.entries.index(e))
            it.output = api.device_view(out2, e.ptype, e.width, e.height, e.width * e.sample, off,
                                        capacity=e.raw_bytes)
            items2.append(it)
        st, q2 = api.DecodeQueue.prepare(items2, [api.Operation.decompress] * n, stream)
        assert st == api.Status.success, st
        assert q2.set_shape(best.group, best.block, best.grid, best.chunk, best.depth) == api.Status.success
        assert q2.capture(qws, results, stream) == api.Status.success
        for _ in range(args.warmup):
            q2.launch(stream)
        ms_packed = timed(stream, lambda: q2.launch(stream), args.steps)
        packed = {"value": round(batch.decoded_bytes / (ms_packed * 1e-3) / 1e9, 2), "unit": "GB/s",
                  "ms_per_step": round(ms_packed, 4), "note": "same launch shape, output rows packed"}
        log(f"packed rows: {ms_packed:.4f} ms per launch")
        del q2, out2, items2

    # ---- plain per-image path ------------------------------------------------------
    # The reference's per-image method: one launch per image, one workgroup per
    # tile.  The plans' own default (a persistent stride grid for plain decodes
    # of 2048 tiles or more) is timed beside it.
    plans, pws, rec = per_image_plans(batch, items, stream)
    default_shapes = [-(-e.width // e.tile[0]) * -(-e.height // e.tile[1]) >= 2048 for e in batch.entries]

    def per_image():
        for plan, r in zip(plans, rec):
            api.decompress_async(plan, pws, r, 1, stream)

    for plan, big in zip(plans, default_shapes):
        if big:
            assert plan.set_decode_shape(8, 256, 1024, 3) == api.Status.success
    for _ in range(args.warmup):
        per_image()
    ms_plain_default = timed(stream, per_image, max(3, args.steps // 4))
    for plan in plans:
        assert plan.set_decode_shape(8, 256, 0, 1) == api.Status.success
    for _ in range(args.warmup):
        per_image()
    ms_plain = timed(stream, per_image, max(3, args.steps // 4))
    if args.protocol_rounds > 0:
        opts = {"images": args.images, "seed": args.seed, "threads": threads, "pitch_align": args.pitch_align,
                "warmup": args.warmup}
        shape = (best.group, best.block, best.grid, best.chunk, best.depth, best.run)
        protocol = protocol_rounds(forkserver, args.protocol_rounds, opts, shape)
        protocol["note"] = ("reference timing protocol: the sample rule (30 samples of at least 2 ms, events on "
                            "the stream) in fresh processes, alternating the order of the two methods, samples "
                            "of every round merged, nearest-rank percentiles of the per-launch time")
    else:
        protocol = {"batched": sampled(stream, launch), "per_image": sampled(stream, per_image),
                    "note": "reference sample rule in this process only (30 samples of at least 2 ms)"}
        for name in ("batched", "per_image"):
            protocol[name].pop("per_call_ms")
    protocol["speedup_p50"] = round(protocol["per_image"]["p50_ms"] / protocol["batched"]["p50_ms"], 3)
    log(f"protocol: batched p50 {protocol['batched']['p50_ms']} ms, per-image p50 {protocol['per_image']['p50_ms']} "
        f"ms, speedup {protocol['speedup_p50']}")

    # ---- verification --------------------------------------------------------------
    verified = None
    if not args.no_verify:
        batch.clear_outputs()
        queue.launch(stream)
        stream.synchronize()
        status, written = api.read_results(results, n)
        ok = bool((status == 0).all()) and all(int(w) == e.raw_bytes for w, e in zip(written, batch.entries))
        for i, e in enumerate(batch.entries):
            ok = ok and np.array_equal(batch.read_output(i), e.pixels)
        verified = ok
        log(f"verification bit-exact: {ok}")

    ceiling = copy_ceiling(stream, stats.payload_bytes + stats.output_bytes + stats.map_bytes)
    log(f"copy ceiling for the same bytes: {ceiling:.1f} GB/s")
    one_sided = stream_ceilings(stream, stats.payload_bytes + stats.map_bytes, stats.output_bytes)
    log(f"one-sided ceilings: {one_sided}")
    cpu = None if args.no_cpu else cpu_baseline(batch, threads, 1_000_000_000)
    del plans, pws, rec
    roi = None if args.no_roi else roi_leg(args, threads, stream)
    fused = None if args.no_fused else fused_leg(args, stream)
    enc = None
    if not args.no_encode:
        enc = encode_leg(args, batch, stream)
        # The same images with packed rows, whose tile rows share 128-byte
        # lines with their neighbours.
        enc["packed_rows"] = encode_leg(args, batch, stream, align=1, secondary=True)
        enc["verified_bit_exact"] = enc["verified_bit_exact"] and enc["packed_rows"]["verified_bit_exact"]
    mix = None if args.no_mix else serve_mix_leg(args, threads, stream)
    staged = None if args.no_staged else staged_leg(args, threads, stream)
    fl = None if args.no_file else file_leg(args, stream)
    if cpu:
        log(f"cpu baseline {cpu['value']} GB/s on {cpu['cores']} threads ({cpu['kind']})")

    decoded = batch.decoded_bytes
    gbps = decoded / (ms_batched * 1e-3) / 1e9
    gbps_plain = decoded / (ms_plain * 1e-3) / 1e9
    traffic_alg = stats.payload_bytes + stats.output_bytes + stats.map_bytes
    default_workload = args.images == 256 and args.seed == 2026517 and args.pitch_align == 256
    traffic = profiled_traffic() if default_workload else None
    achieved = traffic_alg / (ms_batched * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": round(gbps, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_batched, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u16+u32 integer decode",
        "data": "synthetic (seeded generators), archives and outputs resident in HBM",
        "config": {
            "workload": f"batch of {n} images, sides U{{256..4096 step 16}}, 50/50 uint16 detector / "
                        f"float32 field, tiles from {{128x32, 64x64, 256x16}}, whole-image decode",
            "images": n, "tiles": int(stats.tasks), "decoded_bytes": decoded,
            "archive_bytes": batch.archive_bytes, "seed": args.seed,
            "output_row_alignment_bytes": args.pitch_align,
        },
        "packed_rows": packed,
        "per_image": {"value": round(gbps_plain, 2), "unit": "GB/s", "ms_per_step": round(ms_plain, 4),
                      "launches_per_step": 3 * n, "launch": "one workgroup per tile (the reference's method)",
                      "plan_default_shape": {"value": round(batch.decoded_bytes / (ms_plain_default * 1e-3) / 1e9, 2),
                                             "unit": "GB/s", "ms_per_step": round(ms_plain_default, 4),
                                             "speedup_vs": round(ms_plain_default / ms_batched, 3),
                                             "note": "1024 striding workgroups for images of 2048 tiles or more"}},
        "speedup_vs_per_image": round(ms_plain / ms_batched, 3),
        "timing_protocol": protocol,
        "roofline": {
            "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": traffic,
            "algorithmic_bytes_per_launch": traffic_alg,
            "copy_ceiling": round(ceiling, 1),
            "frac_of_copy_ceiling": round(achieved / ceiling, 4),
            "stream_ceilings": one_sided,
            "frac_of_mix_bound": round(achieved / one_sided["mix_bound"], 4),
            "note": "payload bytes of every tile read + decoded output bytes written, over the "
                    "graph launch time (results init, counter reset, decode, results completion; the "
                    "decode kernel is 98-99% of it); traffic = HBM bytes per launch from rocprofv3 "
                    "FETCH_SIZE/WRITE_SIZE on this workload (profiles/decode_batch256_kernel_trace.md); "
                    "copy_ceiling = a plain 16-byte streaming copy moving the same number of bytes, "
                    "measured in this run; stream_ceilings = read-only and write-only streams of 4 GiB "
                    "and the rate of this launch's reads and writes at those two rates in sequence",
        },
        "autotune": {"group": best.group, "block": best.block, "chunk": best.chunk, "depth": best.depth,
                     "candidates": len(candidates)},
        "cpu_baseline": cpu,
        "verified_bit_exact": verified,
        "roi": roi,
        "fused_real_shape": fused,
        "encode": enc,
        "serve_mix": mix,
        "staged": staged,
        "file": fl,
    }
    if roi is not None:
        out["p99_roi_latency_us"] = {str(int(r["qps"])): r["p99_us"] for r in roi["latency"]}
        out["p50_roi_latency_us"] = {str(int(r["qps"])): r["p50_us"] for r in roi["latency"]}
        out["roi_streams"] = roi["latency"][0]["streams"] if roi["latency"] else None
    print(json.dumps(out), flush=True)
    ok = verified in (None, True) and all(leg is None or leg["verified_bit_exact"] in (None, True)
                                          for leg in (roi, mix, enc, staged, fl))
    return 0 if ok else 3


if __name__ == "__main__":
    sys.exit(main())
