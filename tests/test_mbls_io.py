"""The reference's on-disk H-matrix format (savembls / loadmbls,
Include/H.h:47-103; mblock::save / load, Include/mblock.h:737-804).

Pinned by the SHA-256 of the files the compiled reference wrote for two golden
matrices (tests/golden/golden_mbls.npz): the oracle reproduces those matrices
bit for bit, so writing them through ahm_mbls_save must give the reference's
bytes exactly; reading them back must give the same leaves and values, and the
C++ surface (loadmbls -> savembls in tests/cpp/mbls_io) must reproduce the file.
"""
import hashlib
import os
import subprocess

import numpy as np
import pytest

from ahmed_amd import _lib
from conftest import GOLDEN, ROOT, golden_points
from helpers import golden_hmatrix, leaves_of, relerr

CASES = ["ellip_yukawa1800", "pca_yukawa1500"]


def mbls_golden():
    return dict(np.load(os.path.join(GOLDEN, "golden_mbls.npz"), allow_pickle=False))


def written(tmp_path, name):
    g, T, H = golden_hmatrix(name)
    m = mbls_golden()
    leaves, data = leaves_of(H)
    path = tmp_path / f"{name}.mbls"
    _lib.mbls_save(path, float(m[f"{name}_eta"]), int(m[f"{name}_bmin"]), leaves, data)
    return g, m, leaves, data, path


@pytest.mark.parametrize("name", CASES)
def test_writer_gives_the_reference_bytes(tmp_path, name):
    _, m, _, _, path = written(tmp_path, name)
    raw = path.read_bytes()
    assert len(raw) == int(m[f"{name}_bytes"])
    assert hashlib.sha256(raw).hexdigest() == str(m[f"{name}_sha256"])


@pytest.mark.parametrize("name", CASES)
def test_reader_returns_the_leaves(tmp_path, name):
    _, m, leaves, data, path = written(tmp_path, name)
    eta, bmin, rl, rd = _lib.mbls_read(path)
    assert eta == float(m[f"{name}_eta"]) and bmin == int(m[f"{name}_bmin"])
    for f in ("n1", "n2", "rank", "offset"):
        assert np.array_equal(rl[f], leaves[f]), f
    assert np.array_equal(rd, data)
    # FP32 files (mblock<float>): same records with 4-byte values
    p32 = path.with_suffix(".f32")
    _lib.mbls_save(p32, eta, bmin, leaves, data.astype(np.float32))
    _, _, l32, d32 = _lib.mbls_read(p32, np.float32)
    assert np.array_equal(l32["rank"], leaves["rank"]) and np.array_equal(d32, data.astype(np.float32))
    assert p32.stat().st_size == path.stat().st_size - 4 * len(data)


def test_cpp_surface_loadmbls_savembls(tmp_path):
    exe = os.path.join(ROOT, "tests", "cpp", "mbls_io")
    if not os.path.exists(exe):
        pytest.skip("tests/cpp/mbls_io not built (make -C ahmed_amd/csrc)")
    name = "ellip_yukawa1800"        # a bbx tree, as the program builds
    g, m, _, _, path = written(tmp_path, name)
    pts = tmp_path / "pts.bin"
    np.ascontiguousarray(golden_points(g), np.float64).tofile(pts)
    out = tmp_path / "again.mbls"
    p = subprocess.run([exe, str(pts), str(path), str(out)], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    assert f"leaves={len(g['rank'])}" in p.stdout
    assert out.read_bytes() == path.read_bytes()


def test_corrupt_files_are_rejected(tmp_path):
    _, _, leaves, data, path = written(tmp_path, "pca_yukawa1500")
    raw = path.read_bytes()
    cut = tmp_path / "cut.mbls"
    cut.write_bytes(raw[: len(raw) // 2])
    with pytest.raises(_lib.AhmError, match="truncated"):
        _lib.mbls_read(cut)
    # a Hermitian dense record (status 1) of the reference: outside the GeH path
    bad = bytearray(raw)
    status_at = 8 + 4 + 4 + 1 + 8          # header, presence byte, n1, n2
    bad[status_at:status_at + 4] = (1).to_bytes(4, "little")
    hp = tmp_path / "herm.mbls"
    hp.write_bytes(bytes(bad))
    with pytest.raises(_lib.AhmError, match="status 1"):
        _lib.mbls_read(hp)
    absent = bytearray(raw)
    absent[16] = 0                          # first leaf marked absent
    ap = tmp_path / "absent.mbls"
    ap.write_bytes(bytes(absent))
    with pytest.raises(_lib.AhmError, match="absent"):
        _lib.mbls_read(ap)
    with pytest.raises(_lib.AhmError, match="cannot open"):
        _lib.mbls_read(tmp_path / "missing.mbls")


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_gpu_load_matches_reference_product(tmp_path, name):
    g, _, _, _, path = written(tmp_path, name)
    import torch
    ctx = _lib.Context(0)
    H, tree = _lib.HMatrix.load(ctx, path, golden_points(g), int(g["kind"]))
    assert np.array_equal(tree.perm()[0], g["op_perm"])
    x = torch.tensor(g["x"], device="cuda")
    y = torch.zeros(int(g["N"]), dtype=torch.float64, device="cuda")
    H.matvec(x, y)
    err = relerr(y.cpu().numpy(), g["y"])
    print(f"[gpu] {name}: loaded from the reference's format, H x vs reference {err:.2e}")
    assert err <= 1e-12
    H.close()
    tree.close()
