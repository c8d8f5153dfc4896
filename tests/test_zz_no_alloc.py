"""The one-right-hand-side solve that allocates nothing
(scuffx_lu_solve_noalloc; no counterpart in the reference, meant for
callers that cannot allocate device memory).  On the same factors its
result equals scuffx_lu_solve's (trans 'N', nrhs 1) bit for bit -- both
run the same kernels, only the buffers differ -- on two calls in a row,
the second with a smaller system; n beyond its maximum is an error."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _solve_both(sx, n, seed, prec):
    import torch
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    dtype = "complex128" if prec == 64 else "complex64"
    M = sx.HMatrix.from_numpy(A, dtype=dtype)
    assert M.LUFactorize() == 0
    L = sx.lib()
    sx._sync_stream()
    tdt = torch.complex128 if prec == 64 else torch.complex64
    b_ref = torch.tensor(b, dtype=tdt, device="cuda")
    b_na = b_ref.clone()
    print(f"  noalloc solve n={n} prec={prec}", flush=True)
    assert L.scuffx_lu_solve(sx._ptr(M.t), n, n, sx._ptr(M.ipiv), b"N", 1, sx._ptr(b_ref), n, prec) == 0
    rc = L.scuffx_lu_solve_noalloc(sx._ptr(M.t), n, n, sx._ptr(M.ipiv), sx._ptr(b_na), prec)
    assert rc == 0, L.scuffx_last_error()
    torch.cuda.synchronize()
    x_ref, x_na = b_ref.cpu().numpy(), b_na.cpu().numpy()
    assert np.array_equal(x_ref, x_na)
    x = x_na.astype(np.complex128)
    berr = np.linalg.norm(A @ x - b) / (np.linalg.norm(A) * np.linalg.norm(x))
    assert berr <= (1e-14 if prec == 64 else 1e-6), berr


@pytest.mark.parametrize("prec", [64, 32])
def test_noalloc_solve_matches_lu_solve(sx, prec):
    _solve_both(sx, 3000, 11 + prec, prec)
    _solve_both(sx, 1000, 12 + prec, prec)     # second call, smaller system


def test_noalloc_solve_rejects_large_n(sx):
    import torch
    L = sx.lib()
    sx._sync_stream()
    nmax = 16384
    assert L.scuffx_lu_solve_noalloc_bytes() >= 16 * nmax
    A = torch.zeros(4, dtype=torch.complex128, device="cuda")
    piv = torch.zeros(4, dtype=torch.int32, device="cuda")
    b = torch.zeros(4, dtype=torch.complex128, device="cuda")
    rc = L.scuffx_lu_solve_noalloc(sx._ptr(A), nmax + 1, nmax + 1, sx._ptr(piv), sx._ptr(b), 64)
    assert rc == -1
    assert "16384" in L.scuffx_last_error().decode()
