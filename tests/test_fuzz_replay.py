"""Replay of the reference fuzz corpus through the product parser and CPU codec.

The fixture (tests/golden/fuzz_replay.npz, written by make_fuzz_replay.py)
holds every input of the reference's corpus (fuzz/corpus/{parse,mutation},
replayed there by fuzz/replay.cpp:23-44), as the bytes its harness hands to
the parser, with the reference parser's status under the harness's limits
(fuzz/harness.h:25).  Each input must get the same status from
ts_parse_metadata and from the oracle.  An accepted input must also pass the
harness's own checks (fuzz/harness.h:91-163), run here on the product: the
metadata equals the oracle's field by field and covers the file, the decode
leaves row padding untouched and equals the oracle's, and the canonical
re-encode of the decoded pixels is no larger than the input and decodes to
the same pixels.
"""
import numpy as np
import pytest

from oracle import oracle
from tilestream_amd import api

FIXTURE = "tests/golden/fuzz_replay.npz"
LIMITS = (1 << 20, 1 << 15, 8 << 20)
DECODE_LIMIT = 4 << 20
PAD = 0xA5


@pytest.fixture(scope="module")
def corpus():
    import os
    z = np.load(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), FIXTURE))
    offsets = z["offsets"]
    data = z["data"]
    files = [data[offsets[i]:offsets[i + 1]].tobytes() for i in range(len(offsets) - 1)]
    return {"files": files, "status": z["status"].astype(int), "kind": z["kind"],
            "clean": z["harness_clean"], "name": z["name"]}


def test_fixture_covers_both_corpora(corpus):
    kinds = corpus["kind"]
    assert (kinds == 0).sum() >= 140 and (kinds == 1).sum() >= 420
    assert corpus["clean"].all()
    # Both accepted and rejected inputs occur in both corpora.
    for k in (0, 1):
        st = corpus["status"][kinds == k]
        assert (st == 0).any() and (st != 0).any()


def _product_status(file: bytes):
    st, meta = api.parse_metadata(file, api.ParseLimits(*LIMITS))
    return int(st), meta


def test_statuses_equal_reference_and_oracle(corpus):
    bad = []
    for name, file, want in zip(corpus["name"], corpus["files"], corpus["status"]):
        got, _ = _product_status(file)
        orc = oracle.parse_status(file, LIMITS)
        if got != want or orc != want:
            bad.append((str(name), int(want), got, orc))
    assert not bad, bad[:10]


def _check_structure(meta, file: bytes):
    s = meta.summary
    assert s.file_bytes == len(file)
    st, info, tiles, segs = oracle.parse(file, LIMITS)
    assert st == 0
    assert (s.image.type, s.image.width, s.image.height) == (info.type, info.width, info.height)
    assert (s.tile.width, s.tile.height) == (info.tile_width, info.tile_height)
    assert (s.tile_count, s.segment_count) == (info.tile_count, info.segment_count)
    for i, t in enumerate(tiles):
        p = meta.tile(i)
        assert (p.mode, p.width, p.height, p.offset, p.length, p.first_segment) == \
            (t.mode, t.width, t.height, t.offset, t.length, t.first_segment), i
    for i, g in enumerate(segs):
        p = meta.segment(i)
        assert (p.seed, p.bit_width, p.payload_offset, p.payload_bytes) == \
            (g.seed, g.bit_width, g.payload_offset, g.payload_bytes), i
    # The payloads tile the file after the index: contiguous, word multiples.
    at = tiles[0].offset if tiles else len(file)
    for t in tiles:
        assert t.offset == at and t.length > 0 and t.length % 4 == 0
        at += t.length
    assert at == len(file)
    return s


def test_accepted_inputs_decode_and_reencode(corpus):
    checked = 0
    for name, file, want in zip(corpus["name"], corpus["files"], corpus["status"]):
        if want != 0:
            continue
        got, meta = _product_status(file)
        assert got == 0, name
        s = _check_structure(meta, file)
        w, h = s.image.width, s.image.height
        dt = np.uint16 if s.image.type == 1 else np.uint32
        if w * h * np.dtype(dt).itemsize > DECODE_LIMIT:
            continue
        # One padding sample per row that the decode must leave alone.
        padded = np.full((h, w + 1), PAD * (0x0101 if dt == np.uint16 else 0x01010101), dt)
        view = padded[:, :w]
        assert api.cpu_decompress(file, view, api.ParseLimits(*LIMITS)) == api.Status.success, name
        assert (padded[:, w] == padded.dtype.type(PAD * (0x0101 if dt == np.uint16 else 0x01010101))).all(), name
        pixels = np.ascontiguousarray(view)
        assert np.array_equal(pixels, oracle.decompress(file)), name
        st, again = api.cpu_compress(pixels, api.TileShape(s.tile.width, s.tile.height))
        assert st == api.Status.success, name
        assert len(again) <= len(file), name
        back = np.zeros_like(pixels)
        assert api.cpu_decompress(again, back) == api.Status.success
        assert np.array_equal(back, pixels), name
        checked += 1
    assert checked >= 300
