"""Seeded synthetic workloads (detector frames, octave fields, calibration maps).

The pixel generators live in libtsgen.so (tilestream_amd/csrc/gen/generate.cpp);
this module shapes them into numpy arrays and lays out the batch and ROI
workloads.  Every value is a pure function of the seed.
"""
from __future__ import annotations

import math
import os

import numpy as np

from . import _native as N

THREADS = max(1, min(16, os.cpu_count() or 1))
SHAPES = [(128, 32), (64, 64), (256, 16)]


def detector_u16(width: int, height: int, seed: int) -> np.ndarray:
    """Background 1000 + 20 Gaussian blobs + Poisson noise, clipped to uint16."""
    out = np.empty((height, width), np.uint16)
    N.gen().tsg_detector_u16(seed, width, height, out.ctypes.data, width, THREADS)
    return out


def field_f32(width: int, height: int, seed: int) -> np.ndarray:
    """Octave sinusoid field + N(0, 0.5) noise with 0.1% special words (uint32 bits)."""
    out = np.empty((height, width), np.uint32)
    N.gen().tsg_field_f32(seed, width, height, out.ctypes.data, width, THREADS)
    return out


def dark_u16(width: int, height: int, seed: int) -> np.ndarray:
    out = np.empty((height, width), np.uint16)
    N.gen().tsg_dark_u16(seed, width, height, out.ctypes.data, THREADS)
    return out


def gain_f32(width: int, height: int, seed: int) -> np.ndarray:
    out = np.empty((height, width), np.float32)
    N.gen().tsg_gain_f32(seed, width, height, out.ctypes.data, THREADS)
    return out


def bits(seed: int, stream: int, counter: int) -> int:
    return int(N.gen().tsg_bits(seed, stream, counter))


def uniform(seed: int, stream: int, counter: int) -> float:
    return (bits(seed, stream, counter) >> 11) * 2.0 ** -53


def batch_layout(count: int, seed: int, lo: int = 256, hi: int = 4096, step: int = 16):
    """Uneven batch: (width, height, dtype, tile) per image, sides uniform on the
    grid lo, lo + step, ..., hi; half uint16, half float32; tile shape uniform."""
    n_sides = (hi - lo) // step + 1
    items = []
    for i in range(count):
        w = lo + step * (bits(seed, 100, 4 * i) % n_sides)
        h = lo + step * (bits(seed, 100, 4 * i + 1) % n_sides)
        dtype = np.uint16 if bits(seed, 100, 4 * i + 2) % 2 == 0 else np.uint32
        tile = SHAPES[bits(seed, 100, 4 * i + 3) % 3]
        items.append((w, h, dtype, tile))
    return items


def image_for(width: int, height: int, dtype, seed: int) -> np.ndarray:
    return detector_u16(width, height, seed) if np.dtype(dtype) == np.uint16 else field_f32(width, height, seed)


def roi_requests(count: int, seed: int, width: int, height: int, lo: int = 32, hi: int = 2048,
                 archives: int = 1):
    """ROI rectangles with log-uniform sides in [lo, hi] at uniform origins."""
    out = []
    for i in range(count):
        w = int(round(math.exp(math.log(lo) + (math.log(hi) - math.log(lo)) * uniform(seed, 200, 5 * i))))
        h = int(round(math.exp(math.log(lo) + (math.log(hi) - math.log(lo)) * uniform(seed, 200, 5 * i + 1))))
        w, h = max(1, min(w, width)), max(1, min(h, height))
        x = bits(seed, 200, 5 * i + 2) % (width - w + 1)
        y = bits(seed, 200, 5 * i + 3) % (height - h + 1)
        a = bits(seed, 200, 5 * i + 4) % archives
        out.append((a, x, y, w, h))
    return out
