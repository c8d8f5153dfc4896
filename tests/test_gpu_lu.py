"""GPU parity of the dense LU and triangular solves (HMatrix::LUFactorize
/ LUSolve, libs/libhmat/LBWrappers.cc:151-245) against the oracle's
LAPACK-contract zgetrf/zgetrs, called through the C ABI."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEED = 20261202


def rand(n, seed, m=None):
    rng = np.random.default_rng(seed)
    m = n if m is None else m
    return rng.standard_normal((n, m)) + 1j * rng.standard_normal((n, m))


def gpu_lu(sx, A, dtype="complex128"):
    M = sx.HMatrix.from_numpy(A, dtype=dtype)
    info = M.LUFactorize()
    return M, info


@pytest.mark.parametrize("n", [1, 2, 5, 31, 32, 33, 64, 100, 255, 257, 300, 513, 700])
def test_lu64_matches_oracle(sx, O, n):
    print(f"lu64 n={n}", flush=True)
    A = rand(n, SEED + n)
    M, info = gpu_lu(sx, A)
    LU, ipiv, info_o = O.zgetrf(A)
    assert info == info_o == 0
    piv = M.ipiv.cpu().numpy()
    assert np.array_equal(piv, ipiv), np.nonzero(piv != ipiv)[0][:5]
    F = M.numpy()
    err = np.max(np.abs(F - LU)) / np.max(np.abs(LU))
    assert err <= 1e-12 * max(1.0, n / 100), err


@pytest.mark.parametrize("n", [40, 300, 700])
def test_lu32_backward_error(sx, O, n):
    print(f"lu32 n={n}", flush=True)
    A = rand(n, SEED + 7 * n).astype(np.complex64)
    M, info = gpu_lu(sx, A, dtype="complex64")
    assert info == 0
    F = M.numpy().astype(np.complex128)
    piv = M.ipiv.cpu().numpy()
    L = np.tril(F, -1) + np.eye(n)
    U = np.triu(F)
    PA = A.astype(np.complex128).copy()
    for i in range(n):
        p = piv[i] - 1
        if p != i:
            PA[[i, p], :] = PA[[p, i], :]
    berr = np.linalg.norm(PA - L @ U) / np.linalg.norm(A)
    assert berr <= 2e-6 * np.sqrt(n / 40), berr
    # pivots agree with the oracle on the same (rounded) input in all but
    # near-tie positions
    _, ipiv, _ = O.zgetrf(A.astype(np.complex128))
    assert np.mean(piv == ipiv) > 0.95


@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
def test_lu_large_uses_big_tiles(sx, dtype):
    """n = 2600: the trailing updates run on the large-tile GEMM; factors
    and pivots against LAPACK zgetrf (scipy) for complex128, backward error
    for complex64."""
    import scipy.linalg.lapack as lapack
    n = 2600
    print(f"lu large {dtype}", flush=True)
    A = rand(n, SEED + 11)
    if dtype == "complex64":
        A = A.astype(np.complex64)
    M, info = gpu_lu(sx, A, dtype=dtype)
    assert info == 0
    F = M.numpy().astype(np.complex128)
    piv = M.ipiv.cpu().numpy()
    if dtype == "complex128":
        lu2, piv2, _ = lapack.zgetrf(A)
        assert np.array_equal(piv, piv2 + 1)
        assert np.max(np.abs(F - lu2)) <= 1e-11 * np.max(np.abs(lu2))
    else:
        L = np.tril(F, -1) + np.eye(n)
        U = np.triu(F)
        PA = A.astype(np.complex128)[:, :]
        perm = np.arange(n)
        for i in range(n):
            p = piv[i] - 1
            perm[[i, p]] = perm[[p, i]]
        berr = np.linalg.norm(PA[perm] - L @ U) / np.linalg.norm(A)
        assert berr <= 3e-5, berr


def test_lu_singular_info(sx, O):
    A = rand(70, 5)
    A[:, 40] = 0
    M, info = gpu_lu(sx, A)
    _, _, info_o = O.zgetrf(A)
    assert info == info_o == 41


@pytest.mark.parametrize("trans", ["N", "T", "C"])
@pytest.mark.parametrize("n,nrhs", [(77, 1), (300, 3), (600, 2)])
def test_lu_solve_matches_oracle(sx, O, trans, n, nrhs):
    print(f"solve {trans} n={n} nrhs={nrhs}", flush=True)
    A = rand(n, SEED + 3 * n)
    B = rand(n, SEED + 5 * n, nrhs)
    M, _ = gpu_lu(sx, A)
    X = sx.HMatrix.from_numpy(B)
    M.LUSolve(X, Trans=trans)
    Xg = X.numpy()
    # oracle solve on the GPU's own factors: checks the solve alone
    Xo = O.zgetrs(M.numpy(), M.ipiv.cpu().numpy(), B, trans)
    assert np.linalg.norm(Xg - Xo) <= 1e-11 * np.linalg.norm(Xo)
    op = {"N": A, "T": A.T, "C": A.conj().T}[trans]
    assert np.linalg.norm(op @ Xg - B) <= 1e-10 * np.linalg.norm(B) * n


def test_lu_solve_hvector(sx):
    n = 129
    A = rand(n, 9)
    b = rand(n, 10, 1)[:, 0]
    M, _ = gpu_lu(sx, A)
    import torch
    v = sx.HVector(n, data=torch.from_numpy(b.copy()).cuda())
    M.LUSolve(v)
    assert np.linalg.norm(A @ v.numpy() - b) <= 1e-11 * np.linalg.norm(b)


def test_lu_solve_errors(sx):
    n = 8
    M = sx.HMatrix.from_numpy(rand(n, 1))
    with pytest.raises(sx.SCUFFXError):
        M.LUSolve(sx.HMatrix.from_numpy(rand(n, 2, 1)))     # not factored yet
    M.LUFactorize()
    with pytest.raises(sx.SCUFFXError):
        M.LUSolve(sx.HMatrix.from_numpy(rand(n, 2, 1)), Trans="X")
    with pytest.raises(sx.SCUFFXError):
        M.LUSolve(sx.HMatrix.from_numpy(rand(n, 2, 1), dtype="complex64"))


@pytest.mark.parametrize("n", [200, 1000])
def test_mixed_refinement(sx, n):
    print(f"mixed n={n}", flush=True)
    rng = np.random.default_rng(n)
    A = rand(n, SEED + n) + n * np.eye(n) * 0.2
    x_true = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    b = A @ x_true
    import torch
    M = sx.HMatrix.from_numpy(A)
    bv = sx.HVector(n, data=torch.from_numpy(b.copy()).cuda())
    x, iters, rel = sx.solve_mixed(M, bv, tol=1e-14, maxiter=20)
    xs = x.numpy()
    assert iters >= 1
    assert np.linalg.norm(xs - x_true) <= 1e-12 * np.linalg.norm(x_true)
    # the complex128 matrix is left untouched
    assert np.array_equal(M.numpy(), A)


def tournament_lu(A, w=32, rows=256, fan=8):
    """The complex64 factorization's pivoting, restated in complex128:
    32-column panels; every 256-row block nominates the rows a
    partial-pivoting LU of it pivots on (largest |re|+|im|, first on
    ties); nominee sets are stacked 8 at a time and factored again until
    one stack is left, whose pivot rows, in order, are the panel's.
    Returns ipiv (1-based, LAPACK interchange sequence)."""
    def pivots(S):
        S = S.copy()
        used = np.zeros(S.shape[0], bool)
        out = []
        for j in range(min(S.shape[1], S.shape[0])):
            a = np.where(used, -1.0, np.abs(S[:, j].real) + np.abs(S[:, j].imag))
            p = int(np.argmax(a))
            used[p] = True
            out.append(p)
            r = ~used
            l = S[r, j] / S[p, j] if a[p] != 0 else S[r, j]
            S[np.ix_(r, range(j + 1, S.shape[1]))] -= np.outer(l, S[p, j + 1:])
        return out

    A = A.copy()
    n = A.shape[0]
    ipiv = np.zeros(n, dtype=np.int64)
    for r0 in range(0, n, w):
        ww = min(w, n - r0)
        P = A[r0:, r0:r0 + ww]
        sets = [[r0 + b + p for p in pivots(P[b:b + rows])] for b in range(0, n - r0, rows)]
        while len(sets) > 1:
            sets = [[s[p] for p in pivots(A[s, r0:r0 + ww])]
                    for s in (sum(sets[k:k + fan], []) for k in range(0, len(sets), fan))]
        chosen = [sets[0][p] for p in pivots(A[sets[0], r0:r0 + ww])] if len(sets[0]) > ww else sets[0]
        order = np.arange(n)
        for j, R in enumerate(chosen):
            pos = int(np.where(order == R)[0][0])
            ipiv[r0 + j] = pos + 1
            order[[r0 + j, pos]] = order[[pos, r0 + j]]
            A[[r0 + j, pos]] = A[[pos, r0 + j]]
        # LU of the diagonal block without further interchanges
        B = A[r0:r0 + ww, r0:r0 + ww].copy()
        for j in range(ww):
            B[j + 1:, j] /= B[j, j]
            B[j + 1:, j + 1:] -= np.outer(B[j + 1:, j], B[j, j + 1:])
        L11 = np.tril(B, -1) + np.eye(ww)
        U11 = np.triu(B)
        A[r0:r0 + ww, r0:r0 + ww] = B
        A[r0 + ww:, r0:r0 + ww] = np.linalg.solve(U11.T, A[r0 + ww:, r0:r0 + ww].T).T
        A[r0:r0 + ww, r0 + ww:] = np.linalg.solve(L11, A[r0:r0 + ww, r0 + ww:])
        A[r0 + ww:, r0 + ww:] -= A[r0 + ww:, r0:r0 + ww] @ A[r0:r0 + ww, r0 + ww:]
    return ipiv


@pytest.mark.parametrize("n", [40, 260, 700, 2600])
def test_lu32_mixed_factors(sx, O, n):
    """complex64 factorization of the mixed solver (tournament pivoting,
    block-row solves through the panel blocks' inverses): ipiv is a valid
    interchange sequence, P A = L U to complex64 accuracy, |L| stays small
    (tournament pivoting bounds it less tightly than the sqrt(2) of
    partial pivoting), and the pivots agree with the restatement above on
    the rounded matrix except near ties.  n = 260 leaves a 4-row
    workgroup; n = 2600 takes two merge levels."""
    import ctypes as C
    import torch
    print(f"lu32 mixed n={n}", flush=True)
    A = rand(n, SEED + 13 * n)
    M = sx.HMatrix.from_numpy(A)
    LU32 = torch.zeros((n, n), dtype=torch.complex64, device="cuda")
    ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
    info = sx.lib().scuffx_lu32_from_c128(sx._ptr(M.t), n, n, sx._ptr(LU32), n, sx._ptr(ipiv))
    assert info == 0
    F = LU32.cpu().numpy().T.astype(np.complex128)
    piv = ipiv.cpu().numpy()
    assert np.all((piv >= np.arange(1, n + 1)) & (piv <= n))
    L = np.tril(F, -1) + np.eye(n)
    U = np.triu(F)
    perm = np.arange(n)
    for i in range(n):
        p = piv[i] - 1
        perm[[i, p]] = perm[[p, i]]
    A32 = A.astype(np.complex64).astype(np.complex128)
    berr = np.linalg.norm(A32[perm] - L @ U) / np.linalg.norm(A32)
    print(f"  backward error {berr:.2e}, max |L| {np.abs(L).max():.2f}", flush=True)
    assert berr <= 2e-5 * np.sqrt(n / 40)
    assert np.abs(L).max() <= 4.0
    if n <= 700:
        ipt = tournament_lu(A32)
        agree = np.mean(piv == ipt)
        print(f"  pivots agreeing with the restatement: {agree:.3f}", flush=True)
        assert agree > 0.95


def test_lu32_mixed_partial_pivoting_switch(sx, O, monkeypatch):
    """SCUFFX_MIXED_PIVOTING=partial (read at every factorization): the
    mixed solver's complex64 factorization pivots row by row (k_panel)
    through the same block-row kernels; P A = L U to complex64 accuracy,
    |L| <= sqrt 2, pivots equal the oracle's partial pivoting on the rounded
    matrix except near ties, and a mixed solve refines to 1e-12."""
    import torch
    n = 700
    print("mixed factorization with partial pivoting", flush=True)
    monkeypatch.setenv("SCUFFX_MIXED_PIVOTING", "partial")
    rng = np.random.default_rng(SEED + 41)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    M = sx.HMatrix.from_numpy(A)
    LU32 = torch.zeros((n, n), dtype=torch.complex64, device="cuda")
    ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
    rc = sx.lib().scuffx_lu32_from_c128(sx._ptr(M.t), n, n, sx._ptr(LU32), n, sx._ptr(ipiv))
    assert rc == 0, (rc, sx.lib().scuffx_last_error())
    F = LU32.cpu().numpy().T.astype(np.complex128)
    piv = ipiv.cpu().numpy()
    L = np.tril(F, -1) + np.eye(n)
    U = np.triu(F)
    perm = np.arange(n)
    for i in range(n):
        p = piv[i] - 1
        perm[[i, p]] = perm[[p, i]]
    A32 = A.astype(np.complex64).astype(np.complex128)
    berr = np.linalg.norm(A32[perm] - L @ U) / np.linalg.norm(A32)
    assert berr <= 2e-5 * np.sqrt(n / 40)
    assert np.abs(L).max() <= np.sqrt(2.0) + 1e-5
    _, ipo, _ = O.zgetrf(A32)
    assert np.mean(piv == ipo) > 0.95
    x_true = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    B = A + n * 0.2 * np.eye(n)
    x, _, _ = sx.solve_mixed(sx.HMatrix.from_numpy(B), sx.HVector(n, data=torch.from_numpy(B @ x_true).cuda()),
                             tol=1e-14, maxiter=20)
    assert np.linalg.norm(x.numpy() - x_true) <= 1e-12 * np.linalg.norm(x_true)
