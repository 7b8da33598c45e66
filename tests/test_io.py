"""GeoTIFF / CSV I/O and the command line's argument handling (CPU only).

The TIFF reader is checked against an independent parse of the Redwood
fixture (an uncompressed float32 tile) and, where the system's libtiff is
present, against libtiff itself in both directions (libtiff reads what the
writer wrote; the reader reads what libtiff wrote with LZW)."""
import ctypes
import ctypes.util
import os
import struct
import subprocess

import numpy as np
import pytest

import mgpukde_tiled_amd as M

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
RED_TIF = os.path.join(GOLD, "redwood_mask.tif")
RED_CSV = os.path.join(GOLD, "pntsRedwood.csv")
CLI = os.path.join(os.path.dirname(HERE), "mgpukde_tiled_amd", "mgpukde_tiled")


def test_redwood_tiff_matches_independent_parse():
    nc, nr, gt, nd, blk = M.tiff_info(RED_TIF)
    assert (nc, nr) == (128, 128) and blk == (128, 128) and nd == -9999.0
    assert np.allclose(gt, [0.0, 1 / 128, 0.0, 1.0, 0.0, -1 / 128], rtol=0, atol=0)
    raw = open(RED_TIF, "rb").read()
    # one uncompressed 128x128 float32 tile at TileOffsets = 254 (tag dump of the fixture)
    ref = np.frombuffer(raw[254:254 + 65536], "<f4").reshape(128, 128)
    got = M.tiff_read(RED_TIF)
    assert np.array_equal(got, ref)
    assert np.array_equal(M.tiff_read(RED_TIF, (5, 7, 30, 20)), ref[7:27, 5:35])


def test_redwood_csv():
    x, y, h, w = M.csv_read_points(RED_CSV)
    lines = open(RED_CSV).read().strip().splitlines()[1:]
    assert len(x) == len(lines) == 62
    ref = np.array([[float(v) for v in ln.split(",")] for ln in lines], np.float64)
    assert np.array_equal(x, ref[:, 0].astype(np.float32)) and np.array_equal(y, ref[:, 1].astype(np.float32))
    assert np.all(w == 1.0)


def test_csv_errors_and_write(tmp_path):
    p = tmp_path / "hdr.csv"
    p.write_text("x,y,w\n")
    with pytest.raises(M.MkdeError):
        M.csv_read_points(str(p))
    p = tmp_path / "noh.csv"
    p.write_text("x,y,w\n1,2,3\n")
    with pytest.raises(M.MkdeError, match="No bandwidth"):
        M.csv_read_points(str(p), with_h=True)
    p = tmp_path / "h.csv"
    p.write_text("x,y,h,w\n1,2,0.5,3\n4,5,0.25,2\n")
    x, y, h, w = M.csv_read_points(str(p), with_h=True, sample_weights=True)
    assert list(h) == [0.5, 0.25] and list(w) == [3.0, 2.0]
    x, y, h, w = M.csv_read_points(str(p), with_h=True, sample_weights=False)
    assert list(w) == [1.0, 1.0]
    out = tmp_path / "o.csv"
    M.csv_write_points(str(out), x, y, h, w)
    assert out.read_text() == "x, y, h, w\n1.000000, 2.000000, 0.500000, 1.000000\n" \
                              "4.000000, 5.000000, 0.250000, 1.000000\n"


@pytest.mark.parametrize("kind", ["random", "runs", "floats", "large"])
def test_lzw_round_trip(kind):
    r = np.random.default_rng(3)
    if kind == "random":
        data = r.integers(0, 256, 10000, dtype=np.uint8).tobytes()
    elif kind == "runs":
        data = bytes(np.repeat(r.integers(0, 4, 3000, dtype=np.uint8), r.integers(1, 50, 3000)))
    elif kind == "floats":
        data = r.normal(size=20000).astype(np.float32).tobytes()
    else:
        data = (np.arange(300000) % 251).astype(np.uint8).tobytes()
    enc = M.lzw_encode(data)
    assert M.lzw_decode(enc, len(data)) == data


def test_writer_round_trip(tmp_path):
    r = np.random.default_rng(9)
    a = r.uniform(0, 5, (300, 517)).astype(np.float32)
    a[10:40, 100:200] = -9999.0
    gt = [1000.0, 2.5, 0.0, 5000.0, 0.0, -2.5]
    p = str(tmp_path / "o.tif")
    M.tiff_write(p, a, gt, -9999.0, georef_from=RED_TIF)
    nc, nr, g2, nd, _ = M.tiff_info(p)
    assert (nc, nr) == (517, 300) and nd == -9999.0 and np.allclose(g2, gt, rtol=0, atol=0)
    assert np.array_equal(M.tiff_read(p), a)


def _libtiff():
    name = ctypes.util.find_library("tiff") or "libtiff.so.5"
    try:
        return ctypes.CDLL(name)
    except OSError:
        return None


def test_libtiff_reads_written_file(tmp_path):
    L = _libtiff()
    if L is None:
        pytest.skip("libtiff not present")
    L.TIFFOpen.restype = ctypes.c_void_p
    L.TIFFOpen.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
    L.TIFFReadEncodedTile.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_ssize_t]
    L.TIFFReadEncodedTile.restype = ctypes.c_ssize_t
    L.TIFFClose.argtypes = [ctypes.c_void_p]
    a = (np.arange(300 * 260, dtype=np.float32) % 97 * 0.25).reshape(300, 260)
    p = str(tmp_path / "w.tif")
    M.tiff_write(p, a, [0, 1, 0, 300, 0, -1], -1.0)
    t = L.TIFFOpen(p.encode(), b"r")
    assert t
    tiles = []
    buf = np.empty(256 * 256, np.float32)
    for k in range(4):  # 2 x 2 tiles of 256
        n = L.TIFFReadEncodedTile(t, k, buf.ctypes.data, buf.nbytes)
        assert n == buf.nbytes
        tiles.append(buf.reshape(256, 256).copy())
    L.TIFFClose(t)
    full = np.vstack([np.hstack(tiles[0:2]), np.hstack(tiles[2:4])])
    assert np.array_equal(full[:300, :260], a)


def test_reader_reads_libtiff_lzw(tmp_path):
    L = _libtiff()
    if L is None:
        pytest.skip("libtiff not present")
    L.TIFFOpen.restype = ctypes.c_void_p
    L.TIFFOpen.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
    L.TIFFWriteEncodedStrip.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_ssize_t]
    L.TIFFWriteEncodedStrip.restype = ctypes.c_ssize_t
    L.TIFFClose.argtypes = [ctypes.c_void_p]
    L.TIFFSetField.restype = ctypes.c_int
    a = (np.random.default_rng(1).integers(0, 1000, (64, 100))).astype(np.uint16)
    p = str(tmp_path / "l.tif")
    t = L.TIFFOpen(p.encode(), b"w")
    for tag, val in ((256, 100), (257, 64), (258, 16), (277, 1), (278, 16), (259, 5), (262, 1), (284, 1),
                     (339, 1), (317, 2)):
        assert L.TIFFSetField(ctypes.c_void_p(t), ctypes.c_uint32(tag), ctypes.c_int(val)) == 1
    for s in range(4):
        blk = np.ascontiguousarray(a[16 * s:16 * (s + 1)])
        assert L.TIFFWriteEncodedStrip(t, s, blk.ctypes.data, blk.nbytes) == blk.nbytes
    L.TIFFClose(t)
    assert np.array_equal(M.tiff_read(p), a.astype(np.float32))


def _run_cli(args, cwd):
    return subprocess.run([CLI] + [str(a) for a in args], cwd=cwd, capture_output=True, text=True, timeout=60)


def test_cli_argument_errors(tmp_path):
    if not os.path.exists(CLI):
        pytest.skip("CLI not built")
    r = _run_cli(["0", "100"], tmp_path)
    assert r.returncode == 1 and "Incorrect arguments" in r.stdout
    r = _run_cli(["0", 100, 1, 0.1, -1, 1, 0, 1, 0, 12, 1, "a.tif", "b.tif", 0], tmp_path)
    assert r.returncode == 1 and "bandwidth option -1 in mode 0" in r.stdout
    r = _run_cli(["5", 100, 1, 0.1, 0, 1, 0, 1, 0, 12, 1, "a.tif", "b.tif", 0], tmp_path)
    assert r.returncode == 1
    r = _run_cli(["1", "missing.csv", RED_TIF, RED_TIF, 0, 1, 0, 1, 0, 12, 1, "a.tif", "b.tif", 0], tmp_path)
    assert r.returncode == 2 and "Error opening file" in r.stdout
