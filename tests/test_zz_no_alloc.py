"""ahm_hmat_matvec_noalloc: the H-matvec on the library's static per-device
workspace, for callers that cannot allocate.  Through the C ABI it must give
the same bits as ahm_hmat_matvec on two calls in a row (the second with a
smaller matrix, so the array still holds the first call's values), for one
and several right-hand sides in FP64 and FP32, and it must refuse a product
whose workspace exceeds the array, leaving Y untouched."""
import numpy as np
import pytest

from helpers import leaves_of, sphere_hmatrix

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def ctx():
    from ahmed_amd import _lib
    assert torch.cuda.is_available()
    return _lib.Context(0)


def upload(ctx, H):
    from ahmed_amd import _lib
    leaves, data = leaves_of(H)
    return _lib.HMatrix.upload(ctx, H.N, H.N, leaves, data)


@pytest.mark.parametrize("p", [1, 8, 32])
def test_noalloc_equals_matvec_two_calls_smaller_second(ctx, p):
    gen = torch.Generator("cuda").manual_seed(p)
    for N in (16384, 6000):
        print(f"[gpu] noalloc p={p} N={N}", flush=True)
        _, _, H = sphere_hmatrix(N, 64 if N > 8000 else 32, 1e-8)
        D = upload(ctx, H)
        for D_ in (D, D.convert()):
            dt = torch.float64 if D_ is D else torch.float32
            shape = (N,) if p == 1 else (p, N)
            X = torch.randn(shape, dtype=torch.float64, device="cuda", generator=gen).to(dt)
            Y0 = torch.randn(shape, dtype=torch.float64, device="cuda", generator=gen).to(dt)
            Ya, Yb = Y0.clone(), Y0.clone()
            D_.matvec(X, Ya, -0.75)
            D_.matvec_noalloc(X, Yb, -0.75)
            torch.cuda.synchronize()
            assert torch.equal(Ya, Yb), f"p={p} N={N} {dt}"


def test_noalloc_refuses_oversized_workspace(ctx):
    from ahmed_amd import _lib
    _, _, H = sphere_hmatrix(6000, 32, 1e-8)
    D = upload(ctx, H)
    p = 4096   # interleaved X alone: 6000 x 4096 x 8 bytes = 197 MB, plus V^T X
    X = torch.zeros((p, 6000), dtype=torch.float64, device="cuda")
    Y = torch.zeros((p, 6000), dtype=torch.float64, device="cuda")
    with pytest.raises(_lib.AhmError, match="workspace"):
        D.matvec_noalloc(X, Y)
    torch.cuda.synchronize()
    assert float(Y.abs().max()) == 0.0
