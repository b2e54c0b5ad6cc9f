"""Deterministic random round-trip cases for the robustness suites: one case
per seed, seeds from 13,000,000 as in the reference suite
(reference tests/robustness/cases.h:21-22).  The generator is this suite's
own: type and tile shape follow the case number (even cases uint16, odd
float32; shapes in turn), sides are edge-heavy (1, tile edges +-1 or random
up to 300 x 160), a quarter of the images have padded rows, three quarters
decode a random region, and the pixels come from eight patterns: constant,
ramp, a descending ramp that wraps, rows of residuals of a chosen width,
noise, smooth waves, IEEE special words (extremes for uint16), and a
checkerboard with flipped bits.  TEST INFRASTRUCTURE ONLY."""
from __future__ import annotations

import os
from dataclasses import dataclass

import numpy as np

FIRST_SEED = 13_000_000
DEFAULT_COUNT = 10_000
SHAPES = [(128, 32), (64, 64), (256, 16)]
PATTERNS = 8

# One word of every IEEE class: zeros of both signs, subnormals, the smallest
# normal, the largest finite, infinities, quiet and signalling NaNs with
# payloads, and one.
SPECIAL_WORDS = np.array([0x00000000, 0x80000000, 0x00000001, 0x807FFFFF, 0x00800000, 0x7F7FFFFF,
                          0xFF800000, 0x7F800000, 0x7FC00000, 0xFFC12345, 0x7F800001, 0x3F800000], np.uint32)


@dataclass
class Case:
    seed: int
    dtype: type
    tile: tuple
    pixels: np.ndarray      # (h, w)
    pad: int                # extra samples per row of the padded input
    roi: tuple              # (x, y, w, h)
    pattern: int


def _side(rng: np.random.Generator, tile_side: int, limit: int) -> int:
    choice = rng.integers(0, 6)
    if choice == 0:
        return 1
    if choice == 1:
        return max(1, tile_side + int(rng.integers(-1, 2)))
    if choice == 2:
        return max(1, 2 * tile_side + int(rng.integers(-1, 2)))
    return int(rng.integers(1, limit + 1))


def _fill(dtype, w: int, h: int, pattern: int, rng: np.random.Generator) -> np.ndarray:
    u16 = dtype == np.uint16
    mask = 0xFFFF if u16 else 0xFFFFFFFF
    base = int(rng.integers(0, 1 << 32))
    y, x = np.mgrid[0:h, 0:w].astype(np.uint64)
    if pattern == 0:
        v = np.full((h, w), base, np.uint64)
    elif pattern == 1:
        v = base + x * 3 + y * 7
    elif pattern == 2:
        v = (np.uint64(1 << 40) + 65535 - x - y * w).astype(np.uint64)
    elif pattern == 3:
        bits = int(rng.integers(0, 17 if u16 else 33))
        spread = (1 << bits) - 1
        steps = rng.integers(0, spread + 1, size=(h, w), dtype=np.uint64)
        v = base + np.cumsum(steps, axis=1, dtype=np.uint64)
    elif pattern == 4:
        v = rng.integers(0, mask + 1, size=(h, w), dtype=np.uint64)
    elif pattern == 5:
        amp = 2000.0 if u16 else 1000.0
        field = amp * np.sin(x / 17.0 + rng.random()) * np.cos(y / 11.0) + 3 * amp
        if u16:
            v = field.astype(np.uint64)
        else:
            v = field.astype(np.float32).view(np.uint32).astype(np.uint64)
    elif pattern == 6:
        if u16:
            v = rng.choice(np.array([0, 1, 32767, 32768, 65534, 65535], np.uint64), size=(h, w))
        else:
            v = rng.choice(SPECIAL_WORDS.astype(np.uint64), size=(h, w))
    else:
        v = np.where((x + y) % 2 == 0, base, ~np.uint64(base))
        flips = rng.integers(0, 16 if u16 else 32, size=(h, w), dtype=np.uint64)
        hit = rng.random((h, w)) < 0.1
        v = np.where(hit, v ^ (np.uint64(1) << flips), v)
    return (v & np.uint64(mask)).astype(dtype)


def make_case(seed: int) -> Case:
    index = seed - FIRST_SEED
    rng = np.random.default_rng(seed)
    dtype = np.uint16 if index % 2 == 0 else np.uint32
    tile = SHAPES[index // 2 % len(SHAPES)]
    w, h = _side(rng, tile[0], 300), _side(rng, tile[1], 160)
    pad = int(rng.integers(0, 5)) if rng.integers(0, 4) == 0 else 0
    pattern = int(rng.integers(0, PATTERNS))
    pixels = _fill(dtype, w, h, pattern, rng)
    if rng.integers(0, 4) == 0:
        roi = (0, 0, w, h)
    else:
        rx, ry = int(rng.integers(0, w)), int(rng.integers(0, h))
        roi = (rx, ry, int(rng.integers(1, w - rx + 1)), int(rng.integers(1, h - ry + 1)))
    return Case(seed, dtype, tile, pixels, pad, roi, pattern)


def selected_seeds() -> list:
    """TS_ROBUSTNESS_CASE selects one seed; TS_ROBUSTNESS_CASES the number of cases."""
    one = os.environ.get("TS_ROBUSTNESS_CASE")
    if one:
        return [int(one)]
    count = int(os.environ.get("TS_ROBUSTNESS_CASES", DEFAULT_COUNT))
    return list(range(FIRST_SEED, FIRST_SEED + count))
