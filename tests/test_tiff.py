"""TIFF import/export of the file tier (reference tests/tiff_test.cpp,
tests/tiff_interop.py): round trips over both sample types, classic TIFF and
BigTIFF, no compression and Deflate with the horizontal predictor; files of
an independent TIFF implementation (Pillow) read, and ours read by it; the
statuses of unsupported, broken and over-limit files and of invalid writes."""
import os
import struct

import numpy as np
import pytest

from tilestream_amd import api
from tilestream_amd import generators as g
from tilestream_amd.api import Status

Image = pytest.importorskip("PIL.Image")

# Float32 words that must survive bit for bit (reference tests/tiff_interop.py:10-27).
SPECIAL = np.array([0x00000000, 0x80000000, 0x7F800000, 0xFF800000, 0x00000001, 0x80000001, 0x007FFFFF,
                    0x807FFFFF, 0x7F7FFFFF, 0xFF7FFFFF, 0x7FC01234, 0x7FA12345, 0xFFC0ABCD, 0x3F800000,
                    0xBF800000, 0x00800000], np.uint32)


def _u16(w, h, seed=1):
    return g.detector_u16(w, h, seed)


def _f32_words(w, h, seed=2):
    words = g.field_f32(w, h, seed).view(np.uint32).copy()
    words.reshape(-1)[:SPECIAL.size] = SPECIAL
    return words


@pytest.mark.parametrize("kind", ["u16", "f32"])
@pytest.mark.parametrize("bigtiff", [0, 1])
@pytest.mark.parametrize("compression,predictor", [(0, 0), (1, 0), (1, 1)])
@pytest.mark.parametrize("tile", [(128, 32), (64, 64), (16, 48)])
def test_round_trip(tmp_path, kind, bigtiff, compression, predictor, tile):
    if predictor and kind == "f32":
        pytest.skip("the predictor is for uint16 samples")
    px = _u16(333, 217) if kind == "u16" else _f32_words(129, 65)
    path = str(tmp_path / "a.tif")
    opts = api.tiff_write_options(bigtiff=bigtiff, compression=compression, horizontal_predictor=predictor,
                                  tile_width=tile[0], tile_height=tile[1])
    assert api.write_tiff(path, px, opts) == Status.success
    st, back = api.read_tiff(path)
    assert st == Status.success
    assert back.dtype == px.dtype and np.array_equal(back, px)
    with open(path, "rb") as f:
        head = f.read(4)
    assert head == (b"II+\x00" if bigtiff else b"II*\x00")


def test_pitched_source_and_one_pixel(tmp_path):
    wide = _u16(300, 40)
    px = wide[:, 7:7 + 211]  # rows 600 bytes apart
    path = str(tmp_path / "p.tif")
    assert api.write_tiff(path, px, api.tiff_write_options(compression=1)) == Status.success
    st, back = api.read_tiff(path)
    assert st == Status.success and np.array_equal(back, px)
    one = np.array([[0xBEEF]], np.uint16)
    assert api.write_tiff(path, one) == Status.success
    st, back = api.read_tiff(path)
    assert st == Status.success and back.shape == (1, 1) and back[0, 0] == 0xBEEF


@pytest.mark.parametrize("compression", [None, "tiff_deflate", "tiff_adobe_deflate"])
@pytest.mark.parametrize("big_endian", [False, True])
def test_reads_pillow_strips_uint16(tmp_path, compression, big_endian):
    px = _u16(257, 131, 3)
    path = str(tmp_path / "pil.tif")
    im = Image.fromarray(px.astype(">u2") if big_endian else px)
    im.save(path, **({} if compression is None else {"compression": compression}))
    if compression is None:  # Pillow's Deflate path (libtiff) writes little-endian files
        with open(path, "rb") as f:
            assert f.read(2) == (b"MM" if big_endian else b"II")
    st, back = api.read_tiff(path)
    assert st == Status.success and np.array_equal(back, px)


@pytest.mark.parametrize("compression", [None, "tiff_adobe_deflate"])
def test_reads_pillow_float32(tmp_path, compression):
    words = _f32_words(97, 33)
    path = str(tmp_path / "pilf.tif")
    Image.fromarray(words.view(np.float32)).save(path, **({} if compression is None else {"compression": compression}))
    st, back = api.read_tiff(path)
    assert st == Status.success and np.array_equal(back, words)


@pytest.mark.parametrize("kind", ["u16", "f32"])
def test_pillow_reads_ours(tmp_path, kind):
    px = _u16(190, 77, 4) if kind == "u16" else _f32_words(190, 77)
    path = str(tmp_path / "ours.tif")
    opts = api.tiff_write_options(compression=1, horizontal_predictor=1 if kind == "u16" else 0)
    assert api.write_tiff(path, px, opts) == Status.success
    arr = np.array(Image.open(path))
    if kind == "u16":
        assert np.array_equal(arr.astype(np.uint16), px)
    else:
        assert np.array_equal(arr.astype(np.float32).view(np.uint32), px)


def test_tsc_tiff_tsc_preserves_the_archive(tmp_path):
    """Decode an archive, export it as TIFF, import and encode it again: the
    canonical encoder gives the same bytes (reference tiff_test.cpp:265)."""
    px = _u16(301, 177, 9)
    st, archive = api.cpu_compress(px, api.TileShape(64, 64))
    assert st == Status.success
    out = np.zeros_like(px)
    assert api.cpu_decompress(archive, out) == Status.success
    path = str(tmp_path / "r.tif")
    assert api.write_tiff(path, out, api.tiff_write_options(compression=1, horizontal_predictor=1)) == Status.success
    st, back = api.read_tiff(path)
    assert st == Status.success
    st, again = api.cpu_compress(back, api.TileShape(64, 64))
    assert st == Status.success and bytes(again) == bytes(archive)


def test_unsupported_files(tmp_path):
    px = _u16(40, 30)
    path = str(tmp_path / "u.tif")
    Image.fromarray((px >> 8).astype(np.uint8)).save(path)  # 8-bit samples
    assert api.read_tiff(path)[0] == Status.unsupported_format
    Image.fromarray(np.zeros((10, 12, 3), np.uint8)).save(path)  # three samples per pixel
    assert api.read_tiff(path)[0] == Status.unsupported_format
    Image.fromarray(px).save(path, save_all=True, append_images=[Image.fromarray(px)])  # two directories
    assert api.read_tiff(path)[0] == Status.unsupported_format
    Image.fromarray(px.astype(np.int32)).save(path)  # 32-bit signed integers
    assert api.read_tiff(path)[0] == Status.unsupported_format


def _ifd_entries(data):
    """(tag -> (type, count, value field offset)) of a little-endian classic TIFF."""
    assert data[:4] == b"II*\x00"
    ifd = struct.unpack_from("<I", data, 4)[0]
    n = struct.unpack_from("<H", data, ifd)[0]
    out = {}
    for k in range(n):
        tag, typ, count = struct.unpack_from("<HHI", data, ifd + 2 + 12 * k)
        out[tag] = (typ, count, ifd + 2 + 12 * k + 8)
    return out


@pytest.mark.parametrize("compression", [0, 1])
def test_short_tile_is_an_io_error(tmp_path, compression):
    px = _u16(64, 64, 7)
    path = str(tmp_path / "s.tif")
    assert api.write_tiff(path, px, api.tiff_write_options(compression=compression, tile_width=64,
                                                           tile_height=64)) == Status.success
    data = bytearray(open(path, "rb").read())
    typ, count, at = _ifd_entries(bytes(data))[325]  # TileByteCounts, one tile: inline
    assert typ == 4 and count == 1
    length = struct.unpack_from("<I", data, at)[0]
    struct.pack_into("<I", data, at, length - 9)
    open(path, "wb").write(bytes(data))
    assert api.read_tiff(path)[0] == Status.io_error


def test_broken_and_missing_files(tmp_path):
    path = str(tmp_path / "b.tif")
    assert api.read_tiff(str(tmp_path / "missing.tif"))[0] == Status.io_error
    open(path, "wb").write(b"not a tiff file")
    assert api.read_tiff(path)[0] == Status.io_error
    assert api.write_tiff(path, _u16(100, 50), api.tiff_write_options(compression=1)) == Status.success
    data = open(path, "rb").read()
    open(path, "wb").write(data[:len(data) // 2])  # directory lost
    assert api.read_tiff(path)[0] == Status.io_error


def test_limits(tmp_path):
    px = _u16(120, 80)
    path = str(tmp_path / "l.tif")
    assert api.write_tiff(path, px) == Status.success
    from tilestream_amd import _native as N
    full = N.TiffReadLimits(2 ** 32 - 1, 2 ** 32 - 1, 2 ** 48)
    assert api.read_tiff(path, full)[0] == Status.success
    assert api.read_tiff(path, N.TiffReadLimits(119, 2 ** 32 - 1, 2 ** 48))[0] == Status.resource_limit
    assert api.read_tiff(path, N.TiffReadLimits(2 ** 32 - 1, 79, 2 ** 48))[0] == Status.resource_limit
    assert api.read_tiff(path, N.TiffReadLimits(2 ** 32 - 1, 2 ** 32 - 1, 120 * 80 * 2 - 1))[0] == Status.resource_limit
    assert api.read_tiff(path, N.TiffReadLimits(0, 10, 10))[0] == Status.invalid_argument


def test_invalid_writes(tmp_path):
    px, f = _u16(33, 17), _f32_words(33, 17)
    path = str(tmp_path / "w.tif")
    bad = [
        (px, dict(tile_width=24)),                                  # not a multiple of 16
        (px, dict(tile_height=0)),
        (px, dict(horizontal_predictor=1)),                         # predictor without Deflate
        (f, dict(compression=1, horizontal_predictor=1)),           # predictor for float32
        (px, dict(compression=1, deflate_level=0)),
        (px, dict(compression=1, deflate_level=10)),
        (px, dict(compression=2)),
    ]
    for pixels, opts in bad:
        assert api.write_tiff(path, pixels, api.tiff_write_options(**opts)) == Status.invalid_argument, opts
        assert not os.path.exists(path)
    # A path that cannot be created: an I/O error, no file left behind.
    assert api.write_tiff(str(tmp_path / "no-such-dir" / "x.tif"), px) == Status.io_error


def test_large_files_decode_on_several_threads(tmp_path):
    """Files of several MiB split their tiles or strips over host threads; the
    result is the same image as one thread gives."""
    px = _u16(2048, 2304, 11)  # 9 MiB of samples
    strips = str(tmp_path / "strips.tif")
    Image.fromarray(px).save(strips)  # Pillow: uncompressed strips
    st, back = api.read_tiff(strips)
    assert st == Status.success and np.array_equal(back, px)
    tiles = str(tmp_path / "tiles.tif")
    assert api.write_tiff(tiles, px, api.tiff_write_options(tile_width=64, tile_height=64)) == Status.success
    st, back = api.read_tiff(tiles)
    assert st == Status.success and np.array_equal(back, px)
