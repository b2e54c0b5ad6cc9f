"""Device-resident workloads for benchmarks and tests.

A `ResidentBatch` holds seeded images, their .tsc archives (encoded with the
library's own host codec, many images in parallel threads), one device buffer
with every archive, one device buffer with every output, and the parsed host
metadata.  It builds BatchItems for the plain per-image path and for the
DecodeQueue.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import api, generators
from ._native import BatchItem, ConstBytes, Roi, TileShape

ALIGN = 256


def _align(n: int) -> int:
    return (n + ALIGN - 1) // ALIGN * ALIGN


@dataclass
class Entry:
    width: int
    height: int
    dtype: type
    tile: tuple
    seed: int
    pixels: Optional[np.ndarray] = None
    archive: Optional[bytes] = None
    meta: Optional[api.Metadata] = None
    src_offset: int = 0
    out_offset: int = 0
    out_pitch: int = 0

    @property
    def ptype(self):
        return api.PixelType.uint16 if np.dtype(self.dtype) == np.uint16 else api.PixelType.float32

    @property
    def sample(self) -> int:
        return 2 if np.dtype(self.dtype) == np.uint16 else 4

    @property
    def raw_bytes(self) -> int:
        return self.width * self.height * self.sample


@dataclass
class ResidentBatch:
    entries: List[Entry]
    src: api.DeviceMemory = None
    out: api.DeviceMemory = None
    keep: list = field(default_factory=list)

    @classmethod
    def build(cls, layout, seed: int, threads: Optional[int] = None, keep_pixels: bool = True,
              pitch_align: int = 1):
        """layout: iterable of (width, height, dtype, tile).  Output rows start
        every width * sample bytes rounded up to `pitch_align`."""
        entries = [Entry(w, h, dt, tile, seed + i) for i, (w, h, dt, tile) in enumerate(layout)]
        threads = threads or max(1, min(16, os.cpu_count() or 1))

        def make(e: Entry):
            px = generators.image_for(e.width, e.height, e.dtype, e.seed)
            st, data = api.cpu_compress(px, TileShape(*e.tile))
            if st != api.Status.success:
                raise RuntimeError(f"encode failed: {st}")
            e.archive = data
            e.pixels = px if keep_pixels else None
            return e

        # The generator is itself multithreaded; a few images at a time keep all cores busy.
        with cf.ThreadPoolExecutor(max(1, threads // 4)) as pool:
            list(pool.map(make, entries))
        src_total = out_total = 0
        for e in entries:
            e.src_offset = src_total
            src_total += _align(len(e.archive))
            e.out_offset = out_total
            e.out_pitch = (e.width * e.sample + pitch_align - 1) // pitch_align * pitch_align
            out_total += _align(e.out_pitch * (e.height - 1) + e.width * e.sample)
            st, e.meta = api.parse_metadata(e.archive)
            if st != api.Status.success:
                raise RuntimeError(f"parse failed: {st}")
        b = cls(entries)
        b.src = api.DeviceMemory(src_total)
        b.out = api.DeviceMemory(out_total)
        staging = np.zeros(src_total, np.uint8)
        for e in entries:
            staging[e.src_offset:e.src_offset + len(e.archive)] = np.frombuffer(e.archive, np.uint8)
        b.src.upload(staging)
        return b

    @property
    def decoded_bytes(self) -> int:
        return sum(e.raw_bytes for e in self.entries)

    @property
    def archive_bytes(self) -> int:
        return sum(len(e.archive) for e in self.entries)

    def item(self, i: int, roi=None) -> BatchItem:
        e = self.entries[i]
        it = BatchItem()
        x, y, w, h = roi if roi is not None else (0, 0, e.width, e.height)
        full = roi is None or (w, h) == (e.width, e.height)
        pitch = e.out_pitch if full else w * e.sample
        it.output = api.device_view(self.out, e.ptype, w, h, pitch, e.out_offset,
                                    capacity=pitch * (h - 1) + w * e.sample)
        it.compressed_input.bytes = ConstBytes(self.src.ptr + e.src_offset, len(e.archive))
        it.compressed_input.metadata = e.meta.handle
        it.roi = Roi(x, y, w, h)
        return it

    def read_output(self, i: int) -> np.ndarray:
        e = self.entries[i]
        span = e.out_pitch * (e.height - 1) + e.width * e.sample
        raw = np.zeros(e.out_pitch * e.height, np.uint8)
        raw[:span] = self.out.download(span, e.out_offset)
        rows = raw.reshape(e.height, e.out_pitch)[:, : e.width * e.sample]
        return np.ascontiguousarray(rows).view(np.uint16 if e.sample == 2 else np.uint32)

    def clear_outputs(self) -> None:
        self.out.fill(0)
