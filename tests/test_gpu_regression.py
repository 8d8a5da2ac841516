"""Regression on the reference's own test images: a 2048 x 2048 window of the
50% marker/mask pair (IWPP MP/Imagens_de_Teste/50-percent-*.jpg, decoded with
PIL, stored in tests/golden/pair50.npz with the SHA-256 of the oracle's
reconstruction) through every queue variant, several tile sizes, both
connectivities, the legacy entry point and the command-line tool -- the
counterpart of IWPP MP/test_script.sh, with the hash in place of the
external verifier."""
import hashlib
import os
import subprocess
import sys

import numpy as np
import pytest

import iwpp

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
PAIR = np.load(os.path.join(HERE, "golden", "pair50.npz"))
CLI = os.path.join(os.path.dirname(iwpp.LIB_PATH), "iwpp_recon")


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("conn", [4, 8])
@pytest.mark.parametrize("queue", ["mutex", "dq", "uq"])
def test_pair50_every_variant_and_tile(torch_cuda, conn, queue):
    marker, mask = PAIR["marker"], PAIR["mask"]
    want = str(PAIR[f"out{conn}_sha256"])
    I = iwpp.DeviceImage.from_numpy(mask)
    for tile in (0, 512, 1024, 2048):
        say("pair50", conn, queue, tile)
        J = iwpp.DeviceImage.from_numpy(marker)
        iwpp.reconstruct(J, I, connectivity=conn, tile=tile, queue=queue)
        assert sha(J.numpy()) == want, (conn, queue, tile)


def test_pair50_legacy_entry_point(torch_cuda):
    marker, mask = PAIR["marker"], PAIR["mask"]
    J, I = iwpp.DeviceImage.from_numpy(marker), iwpp.DeviceImage.from_numpy(mask)
    iwpp.MorphologicalReconstruction(J.ptr, I.ptr, J.pitch, I.pitch, marker.shape[0])
    torch_cuda.cuda.synchronize()
    assert sha(J.numpy()) == str(PAIR["out4_sha256"])


@pytest.mark.parametrize("args,conn", [(["--queue", "uq"], 4), (["--conn", "8", "--queue", "mutex",
                                                                 "--tile", "1024"], 8),
                                       (["--legacy"], 4), (["--queue", "dq", "--repeat", "3",
                                                            "--stats"], 4)])
def test_pair50_command_line(torch_cuda, tmp_path, args, conn):
    say("cli", args)
    mk, ms, out = tmp_path / "marker.pgm", tmp_path / "mask.pgm", tmp_path / "out.pgm"
    iwpp.pgm_write(mk, PAIR["marker"])
    iwpp.pgm_write(ms, PAIR["mask"])
    r = subprocess.run([CLI, str(mk), str(ms), str(out)] + args, capture_output=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr.decode()
    times = [l for l in r.stdout.decode().splitlines() if not l.startswith("visits")]
    assert len(times) == (3 if "--repeat" in args else 1) and all(float(t) > 0 for t in times)
    res, mv = iwpp.pgm_read(out)
    assert mv == 255 and sha(res) == str(PAIR[f"out{conn}_sha256"])


def test_command_line_uint16(torch_cuda, tmp_path):
    # 12-bit data through 16-bit PGM files
    import oracle
    rng = np.random.default_rng(5)
    mask = rng.integers(0, 4096, (300, 517)).astype(np.uint16)
    marker = np.maximum(mask.astype(np.int64) - 700, 0).astype(np.uint16)
    mk, ms, out = tmp_path / "m.pgm", tmp_path / "k.pgm", tmp_path / "o.pgm"
    iwpp.pgm_write(mk, marker, 4095)
    iwpp.pgm_write(ms, mask, 4095)
    r = subprocess.run([CLI, str(mk), str(ms), str(out), "--conn", "8"], capture_output=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr.decode()
    res, mv = iwpp.pgm_read(out)
    assert mv == 4095 and np.array_equal(res, oracle.reconstruct(marker, mask, 8))
