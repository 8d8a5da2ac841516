"""PGM reader / writer (include/iwpp_io.h) and the command-line tool's
argument handling.  CPU only."""
import os
import subprocess

import numpy as np
import pytest

import iwpp

CLI = os.path.join(os.path.dirname(iwpp.LIB_PATH), "iwpp_recon")


@pytest.mark.parametrize("dtype,maxval", [(np.uint8, 255), (np.uint8, 7), (np.uint16, 4095),
                                          (np.uint16, 65535), (np.uint16, 256)])
@pytest.mark.parametrize("shape", [(1, 1), (3, 7), (64, 33)])
def test_round_trip(tmp_path, dtype, maxval, shape):
    rng = np.random.default_rng(shape[0] * 100 + maxval)
    a = rng.integers(0, maxval + 1, shape).astype(dtype)
    p = tmp_path / "a.pgm"
    iwpp.pgm_write(p, a, maxval)
    b, mv = iwpp.pgm_read(p)
    assert mv == maxval and b.dtype == a.dtype and np.array_equal(a, b)


def test_binary_layout(tmp_path):
    # P5, "width height", 16-bit samples big-endian (the Netpbm definition)
    a = np.array([[1, 258, 65535]], dtype=np.uint16)
    p = tmp_path / "b.pgm"
    iwpp.pgm_write(p, a, 65535)
    assert p.read_bytes() == b"P5\n3 1\n65535\n\x00\x01\x01\x02\xff\xff"
    a8 = np.array([[1, 2], [3, 4], [5, 6]], dtype=np.uint8)
    iwpp.pgm_write(p, a8)
    assert p.read_bytes() == b"P5\n2 3\n255\n\x01\x02\x03\x04\x05\x06"


def test_ascii_and_comments(tmp_path):
    p = tmp_path / "c.pgm"
    p.write_bytes(b"P2\n# a comment\n3 2 # trailing\n# more\n  9\n0 1 2\n3 # x\n 4 9\n")
    b, mv = iwpp.pgm_read(p)
    assert mv == 9 and b.tolist() == [[0, 1, 2], [3, 4, 9]]
    q = tmp_path / "d.pgm"
    q.write_bytes(b"P5 #c\n2 1\n255\n\x07\x08")
    b, mv = iwpp.pgm_read(q)
    assert b.tolist() == [[7, 8]]


@pytest.mark.parametrize("content", [b"", b"P6\n1 1\n255\n\x00", b"P5\n2 2\n255\n\x00\x01\x02",
                                     b"P5\n0 2\n255\n", b"P5\n1 1\n70000\n\x00\x00",
                                     b"P2\n2 1\n5\n1 6\n", b"P5\n1 1\n0\n\x00"])
def test_rejects_bad_files(tmp_path, content):
    p = tmp_path / "bad.pgm"
    p.write_bytes(content)
    with pytest.raises(iwpp.PGMError):
        iwpp.pgm_read(p)


def test_missing_file(tmp_path):
    with pytest.raises(iwpp.PGMError):
        iwpp.pgm_read(tmp_path / "nope.pgm")


def test_cli_usage_errors(tmp_path):
    assert os.path.exists(CLI)
    r = subprocess.run([CLI], capture_output=True)
    assert r.returncode == 1 and b"usage" in r.stderr
    a = tmp_path / "a.pgm"
    iwpp.pgm_write(a, np.zeros((4, 4), np.uint8))
    b = tmp_path / "b.pgm"
    iwpp.pgm_write(b, np.zeros((4, 5), np.uint8))
    r = subprocess.run([CLI, str(a), str(b), str(tmp_path / "o.pgm")], capture_output=True)
    assert r.returncode == 1 and b"differ" in r.stderr
    r = subprocess.run([CLI, str(a), str(a), str(tmp_path / "o.pgm"), "--conn", "6"],
                       capture_output=True)
    assert r.returncode == 1
    r = subprocess.run([CLI, str(tmp_path / "x.pgm"), str(a), str(tmp_path / "o.pgm")],
                       capture_output=True)
    assert r.returncode == 1 and b"cannot read" in r.stderr
