"""Benchmark: decoded GB/s of batched persistent decode vs the plain per-image path.

Workload (one "step"): decode a batch of 256 seeded images whose width and
height are uniform in {256, 272, ..., 4096} (not tile-aligned), half uint16
detector frames (background + Gaussian blobs + Poisson noise), half float32
octave fields with 0.1% special words, tile shape drawn from {128x32, 64x64,
256x16}.  Archives and outputs are resident in HBM before timing starts.

  batched   one DecodeQueue launch (persistent kernel draining a device work
            queue) captured in a HIP graph; launch shape from the autotuner
  per-image 256 decompress_async calls, one plan per image, same stream

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
        "value": round(total / multi / 1e9, 3), "unit": "GB/s", "cores": threads, "kind": kind,
        "sample": f"whole-image decode of all {len(entries)} archives of the batch "
                  f"({total / 1e9:.2f} GB decoded) on {threads} threads, one image per thread",
        "single_thread": {"value": round(single_bytes / single / 1e9, 3), "unit": "GB/s", "cores": 1,
                          "sample": f"first {len(sample)} images ({single_bytes / 1e9:.2f} GB decoded)"},
    }


def main() -> int:
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
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
    p.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                   help="only run a plain copy kernel of this many MiB (scales HBM byte counters)")
    args = p.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (this codec has no "
              "multi-GPU path: shard images or requests across processes, one per GPU)", file=sys.stderr)
        return 2

    from tilestream_amd import api, generators
    from tilestream_amd.workloads import ResidentBatch

    if api.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        return 1
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
    threads = max(1, min(16, os.cpu_count() or 1))
    t0 = time.perf_counter()
    layout = generators.batch_layout(args.images, args.seed)
    batch = ResidentBatch.build(layout, args.seed, threads)
    log(f"built {len(layout)} images, {batch.decoded_bytes / 1e9:.2f} GB decoded, "
        f"{batch.archive_bytes / 1e9:.2f} GB archives in {time.perf_counter() - t0:.1f} s")

    stream = api.Stream()
    n = len(batch.entries)
    results = api.DeviceMemory(16 * n)

    # ---- batched persistent path -------------------------------------------------
    items = [batch.item(i) for i in range(n)]
    st, queue = api.DecodeQueue.prepare(items, [api.Operation.decompress] * n, stream)
    assert st == api.Status.success, st
    qws = api.DeviceMemory(max(queue.workspace_bytes(), 256))
    candidates = [(g, b, 0, c, d) for g in (4, 8, 16) for (b, d) in ((128, 1), (256, 1), (256, 2), (512, 2))
                  for c in (4, 16)]
    candidates += [(8, 256, 0, 0, 2), (4, 256, 0, 0, 2)]
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
    for _ in range(args.warmup):
        launch()
    ms_batched = timed(stream, launch, args.steps)
    stats = queue.stats()
    if args.profile:
        log(f"batched {ms_batched:.4f} ms per launch, "
            f"{batch.decoded_bytes / (ms_batched * 1e-3) / 1e9:.1f} GB/s decoded")
        return 0

    # ---- plain per-image path ------------------------------------------------------
    plans, ws_bytes = [], 0
    for i in range(n):
        st, plan = api.prepare_batch(api.Operation.decompress, [items[i]], stream=stream)
        assert st == api.Status.success, st
        plans.append(plan)
        ws_bytes = max(ws_bytes, plan.workspace_bytes())
    pws = api.DeviceMemory(ws_bytes)
    rec = [api.DeviceMemory(16) for _ in range(n)]

    def per_image():
        for plan, r in zip(plans, rec):
            api.decompress_async(plan, pws, r, 1, stream)

    for _ in range(args.warmup):
        per_image()
    ms_plain = timed(stream, per_image, max(3, args.steps // 4))

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

    cpu = None if args.no_cpu else cpu_baseline(batch, threads, 1_000_000_000)
    if cpu:
        log(f"cpu baseline {cpu['value']} GB/s on {cpu['cores']} threads ({cpu['kind']})")

    decoded = batch.decoded_bytes
    gbps = decoded / (ms_batched * 1e-3) / 1e9
    gbps_plain = decoded / (ms_plain * 1e-3) / 1e9
    traffic_alg = stats.payload_bytes + stats.output_bytes + stats.map_bytes
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
        },
        "per_image": {"value": round(gbps_plain, 2), "unit": "GB/s", "ms_per_step": round(ms_plain, 4),
                      "launches_per_step": 3 * n},
        "speedup_vs_per_image": round(ms_plain / ms_batched, 3),
        "roofline": {
            "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": None,
            "algorithmic_bytes_per_launch": traffic_alg,
            "note": "payload bytes of every tile read + decoded output bytes written, over the "
                    "graph launch time (results init, counter reset, decode, results completion)",
        },
        "autotune": {"group": best.group, "block": best.block, "chunk": best.chunk, "depth": best.depth,
                     "candidates": len(candidates)},
        "cpu_baseline": cpu,
        "verified_bit_exact": verified,
    }
    print(json.dumps(out), flush=True)
    return 0 if verified in (None, True) else 3


if __name__ == "__main__":
    sys.exit(main())
