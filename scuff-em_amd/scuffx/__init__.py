"""scuffx -- Python front end of the MI355X BEM system path.

Mirrors the names of the reference's SWIG module for this path
(libs/python/scuff.i:9-53, usage in doc/docs/API/MainFlow.md):

    G  = scuffx.RWGGeometry("Sphere.scuffgeo")        # RWGGeometry.cc:353
    M  = G.AssembleBEMMatrix(Omega)                    # AssembleBEMMatrix.cc:684
    KN = G.AssembleRHSVector(Omega, scuffx.PlaneWave([1,0,0],[0,0,1]))
    M.LUFactorize()                                    # LBWrappers.cc:151
    M.LUSolve(KN)                                      # LBWrappers.cc:175

Matrices and vectors live in GPU memory (torch tensors used as plain
device buffers); all arithmetic runs in the HIP library libscuffx.so
through its C ABI (include/scuffx.h).  There is no CPU fallback: if the
library or a GPU is missing the calls raise.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBPATH = os.path.join(_HERE, "libscuffx.so")
_LIB = None

ZVAC = 376.73031346177
NOT_CONVERGED = 1      # SCUFFX_NOT_CONVERGED


class SCUFFXError(RuntimeError):
    pass


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(_LIBPATH):
            raise ImportError("libscuffx.so not built (run __graft_entry__.build() or make -C scuff-em_amd)")
        # torch bundles its own libamdhip64.so.7 with the same soname as the
        # system one this library links: whichever is loaded first serves
        # the whole process.  torch has to come first so that it runs on
        # the HIP runtime it was built against; the library is ABI
        # compatible with either.
        import torch  # noqa: F401
        L = C.CDLL(_LIBPATH)
        vp, dp, ip, lp = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_long)
        sig = {
            "scuffx_last_error": ([], C.c_char_p),
            "scuffx_version": ([], C.c_int),
            "scuffx_set_stream": ([vp], C.c_int),
            "scuffx_geom_open": ([C.c_char_p, C.POINTER(vp)], C.c_int),
            "scuffx_geom_new": ([C.POINTER(vp)], C.c_int),
            "scuffx_geom_add_mesh": ([vp, vp, C.c_int, vp, C.c_int, C.c_char_p], C.c_int),
            "scuffx_geom_add_msh": ([vp, C.c_char_p, C.c_char_p], C.c_int),
            "scuffx_geom_free": ([vp], C.c_int),
            "scuffx_geom_total_bfs": ([vp], C.c_int),
            "scuffx_geom_num_surfaces": ([vp], C.c_int),
            "scuffx_geom_counts": ([vp, C.c_int, ip, ip, ip], C.c_int),
            "scuffx_geom_bf_range": ([vp, C.c_int, ip, ip], C.c_int),
            "scuffx_geom_edges": ([vp, C.c_int, vp, vp], C.c_int),
            "scuffx_geom_panels": ([vp, C.c_int, vp], C.c_int),
            "scuffx_geom_vertices": ([vp, C.c_int, vp], C.c_int),
            "scuffx_geom_mate": ([vp, C.c_int], C.c_int),
            "scuffx_geom_epsmu": ([vp, C.c_int, C.c_double, C.c_double, vp, vp], C.c_int),
            "scuffx_trans_open": ([C.c_char_p, vp], C.c_int),
            "scuffx_trans_count": ([vp], C.c_int),
            "scuffx_trans_tag": ([vp, C.c_int, C.c_char_p, C.c_int], C.c_int),
            "scuffx_trans_free": ([vp], C.c_int),
            "scuffx_geom_label": ([vp, C.c_int, C.c_char_p, C.c_int], C.c_int),
            "scuffx_geom_set_label": ([vp, C.c_int, C.c_char_p], C.c_int),
            "scuffx_geom_check_trans": ([vp, vp], C.c_int),
            "scuffx_geom_transform": ([vp, vp, C.c_int], C.c_int),
            "scuffx_geom_untransform": ([vp], C.c_int),
            "scuffx_geom_surface_moved": ([vp, C.c_int], C.c_int),
            "scuffx_batch_sweep": ([vp, vp, C.c_int, vp, vp, C.c_int, vp], C.c_int),
            "scuffx_sweep_batched": ([vp, C.c_int, vp, C.c_int, vp, vp, C.c_int, vp, vp, vp, vp], C.c_int),
            "scuffx_adjoint": ([vp, C.c_int, C.c_double, C.c_double, vp, vp, C.c_int, vp, vp], C.c_int),
            "scuffx_geom_transform_surface": ([vp, C.c_int, vp, vp], C.c_int),
            "scuffx_geom_set_vertices": ([vp, C.c_int, vp, C.c_int], C.c_int),
            "scuffx_geom_set_far_precision": ([vp, C.c_int], C.c_int),
            "scuffx_vertex_slots": ([vp], C.c_int),
            "scuffx_vertex_gradient": ([vp, C.c_double, C.c_double, vp, C.c_long, vp, vp, vp, C.c_long, vp],
                                       C.c_int),
            "scuffx_rhs_vertex_gradient": ([vp, C.c_double, C.c_double, vp, vp, vp, vp], C.c_int),
            "scuffx_pabs_vertex_gradient": ([vp, vp, vp, vp], C.c_int),
            "scuffx_hdf5_open": ([C.c_char_p, vp], C.c_int),
            "scuffx_hdf5_close": ([vp], C.c_int),
            "scuffx_hdf5_write_matrix": ([vp, C.c_char_p, vp, C.c_int, C.c_int, C.c_int], C.c_int),
            "scuffx_hdf5_write_vector": ([vp, C.c_char_p, vp, C.c_int, C.c_int], C.c_int),
            "scuffx_hdf5_read_matrix": ([C.c_char_p, C.c_char_p, vp, ip, ip, ip], C.c_int),
            "scuffx_hdf5_read_vector": ([C.c_char_p, C.c_char_p, vp, ip, ip], C.c_int),
            "scuffx_icosphere": ([C.c_int, C.c_double, vp, vp, vp], C.c_int),
            "scuffx_fippi_build": ([vp], C.c_int),
            "scuffx_fippi_count": ([vp], C.c_int),
            "scuffx_fippi_records": ([vp, vp], C.c_int),
            "scuffx_assemble": ([vp, C.c_double, C.c_double, vp, C.c_long], C.c_int),
            "scuffx_assemble_stats": ([vp, vp, vp], C.c_int),
            "scuffx_assemble_block": ([vp, C.c_int, C.c_int, C.c_double, C.c_double, vp, C.c_long, vp, C.c_long,
                                       C.c_long, C.c_int, vp, vp], C.c_int),
            "scuffx_lu_factor": ([vp, C.c_long, C.c_long, vp, C.c_int], C.c_int),
            "scuffx_lu_solve": ([vp, C.c_long, C.c_long, vp, C.c_char, C.c_long, vp, C.c_long, C.c_int], C.c_int),
            "scuffx_lu32_from_c128": ([vp, C.c_long, C.c_long, vp, C.c_long, vp], C.c_int),
            "scuffx_refine": ([vp, C.c_long, C.c_long, vp, C.c_long, vp, vp, vp, C.c_double, C.c_int,
                               ip, dp], C.c_int),
            "scuffx_gmres_ir": ([vp, C.c_long, C.c_long, vp, C.c_long, vp, vp, vp, C.c_double, C.c_int,
                                 ip, dp], C.c_int),
            "scuffx_shutdown": ([], C.c_int),
            "scuffx_assemble_real": ([vp, C.c_double, C.c_double, vp, C.c_long, C.c_int], C.c_int),
            "scuffx_assemble_block_real": ([vp, C.c_int, C.c_int, C.c_double, C.c_double, vp, C.c_long, C.c_int,
                                            vp, C.c_long, C.c_long, C.c_int, vp, vp], C.c_int),
            "scuffx_unpack_symmetric": ([vp, C.c_long, vp, C.c_long, C.c_int], C.c_int),
            "scuffx_lu_invert": ([vp, C.c_long, C.c_long, vp, C.c_int], C.c_int),
            "scuffx_lu_solve_noalloc": ([vp, C.c_long, C.c_long, vp, vp, C.c_int], C.c_int),
            "scuffx_lu_solve_noalloc_bytes": ([], C.c_long),
            "scuffx_rhs_planewave": ([vp, C.c_double, C.c_double, vp, vp, vp], C.c_int),
            "scuffx_power": ([vp, C.c_double, C.c_double, vp, vp, dp, dp], C.c_int),
            "scuffx_power_emt": ([vp, C.c_double, C.c_double, vp, vp, vp, C.c_long, dp, dp], C.c_int),
            "scuffx_assemble_region": ([vp, C.c_double, C.c_double, C.c_int, vp, C.c_long], C.c_int),
            "scuffx_lu_logdet": ([vp, C.c_long, C.c_long, vp, C.c_int, dp, dp], C.c_int),
            "scuffx_trace_solve": ([vp, C.c_long, C.c_long, vp, vp, C.c_long, C.c_long, C.c_int, dp, dp], C.c_int),
            "scuffx_preload_cache": ([C.c_char_p], C.c_long),
            "scuffx_store_cache": ([C.c_char_p], C.c_long),
            "scuffx_cache_size": ([], C.c_long),
            "scuffx_clear_cache": ([], C.c_int),
            "scuffx_profile": ([C.c_int], C.c_int),
            "scuffx_profile_read": ([C.c_char_p, dp, dp, lp], C.c_int),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _LIB = L
        # release device memory while the HIP runtime is still up: at exit,
        # C++ destructors of the library (and __del__ of live objects) run
        # after the runtime -- or a profiler attached to it -- is torn down.
        # torch was imported first, so this handler runs before its own.
        import atexit
        atexit.register(_shutdown)
    return _LIB


def _shutdown():
    # every device call goes through _sync_stream, which initializes torch's
    # HIP context first: without it the library holds no device memory
    import torch
    if _LIB is not None and torch.cuda.is_initialized():
        _LIB.scuffx_shutdown()


def _check(rc):
    if rc < 0:
        raise SCUFFXError(lib().scuffx_last_error().decode() or ("error %d" % rc))
    return rc


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise SCUFFXError("scuffx needs a GPU (torch.cuda.is_available() is False)")
    return torch


def _sync_stream():
    torch = _torch()
    lib().scuffx_set_stream(C.c_void_p(torch.cuda.current_stream().cuda_stream))


def _sync_stream_if_gpu():
    """the caller's stream, where a GPU is present (geometry edits work
    without one)"""
    import torch
    if torch.cuda.is_available():
        _sync_stream()


def _ptr(t):
    return C.c_void_p(t.data_ptr())


def icosphere(level, radius=1.0, center=None):
    """Vertices (nV x 3) and triangles (nT x 3) of a recursively subdivided
    icosahedron projected on a sphere (benchmark meshes)."""
    nV, nT = 10 * 4 ** level + 2, 20 * 4 ** level
    V = np.zeros((nV, 3))
    T = np.zeros((nT, 3), dtype=np.int32)
    c = None if center is None else np.ascontiguousarray(center, dtype=np.float64)
    _check(lib().scuffx_icosphere(level, radius, None if c is None else c.ctypes.data_as(C.c_void_p),
                                  V.ctypes.data_as(C.c_void_p), T.ctypes.data_as(C.c_void_p)))
    return V, T


def CreateGammaMatrix(axis):
    """3x3 generator Gamma of rotations about `axis` (Gamma x = u x x for the
    unit axis u), the matrix of CreateGammaMatrix
    (libs/libscuff/EdgeEdgeInteractions.cc:241-270)."""
    u = np.asarray(axis, dtype=np.float64)
    u = u / np.linalg.norm(u)
    return np.array([[0.0, -u[2], u[1]], [u[2], 0.0, -u[0]], [-u[1], u[0], 0.0]])


_PREC = {"complex128": 64, "complex64": 32, "float64": 8}


def _prec_of(dtype):
    import torch
    if dtype in _PREC:
        return _PREC[dtype]
    return {torch.complex128: 64, torch.complex64: 32, torch.float64: 8}[dtype]


def _torch_dtype(prec):
    import torch
    return {64: torch.complex128, 32: torch.complex64, 8: torch.float64}[prec]


class HVector:
    """Dense vector in GPU memory (libs/libhmat/libhmat.h:71-150):
    complex128 by default, complex64, or float64 (LHM_REAL)."""

    def __init__(self, N, dtype="complex128", data=None):
        torch = _torch()
        self.N = N
        self.prec = _prec_of(data.dtype if data is not None else dtype)
        self.t = torch.zeros(N, dtype=_torch_dtype(self.prec), device="cuda") if data is None else data

    def numpy(self):
        return self.t.cpu().numpy()

    def ExportToHDF5(self, dest, Name):
        """HVector::ExportToHDF5 (HDF5IO.cc:436-490)"""
        ExportToHDF5(dest, Name, self.numpy())

    @classmethod
    def FromHDF5(cls, FileName, Name):
        """HVector(FileName, LHM_HDF5, Name) (HDF5IO.cc:497-540)"""
        torch = _torch()
        v = ImportFromHDF5(FileName, Name, Vector=True)
        return cls(len(v), data=torch.from_numpy(v).cuda())

    def GetEntry(self, n):
        return complex(self.t[n].item())

    def GetEntryD(self, n):
        return complex(self.t[n].item()).real

    def SetEntry(self, n, v):
        self.t[n] = v


class HMatrix:
    """Dense column-major matrix in GPU memory with LU factors and pivots
    (libs/libhmat/libhmat.h:211-394).  dtype "complex128" (LHM_COMPLEX,
    default), "complex64" or "float64" (LHM_REAL).  Storage "normal": a
    torch tensor of shape (NC, NR), t[c, r] is entry (r, c); "symmetric"
    (LHM_SYMMETRIC, square only): packed upper triangle, t[r + c(c+1)/2] for
    r <= c (HMatrix.cc:325-342).  A packed matrix is factored as a full copy
    with partial pivoting (LU), not by dsptrf/zsptrf's Bunch-Kaufman LDL^T;
    LUSolve, LogDet and TraceSolve use that copy."""

    def __init__(self, NR, NC, dtype="complex128", data=None, Storage="normal"):
        torch = _torch()
        self.NR, self.NC = NR, NC
        self.prec = _prec_of(data.dtype if data is not None else dtype)
        self.packed = (Storage == "symmetric")
        if self.packed and NR != NC:
            self.packed = False      # as HMatrix.cc:94-96 does for non-square matrices
        shape = (NR * (NR + 1) // 2,) if self.packed else (NC, NR)
        self.t = torch.zeros(shape, dtype=_torch_dtype(self.prec), device="cuda") if data is None else data
        self.ipiv = None
        self.factored = False
        self._full = None            # full working copy of a packed matrix (its factors)

    @property
    def RealComplex(self):
        return "LHM_REAL" if self.prec == 8 else "LHM_COMPLEX"

    @classmethod
    def from_numpy(cls, A, dtype=None, Storage="normal"):
        torch = _torch()
        A = np.asarray(A)
        if dtype is None:
            dtype = "float64" if not np.iscomplexobj(A) else "complex128"
        M = cls(A.shape[0], A.shape[1], dtype=dtype, Storage=Storage)
        if M.packed:
            iu = np.triu_indices(A.shape[0])
            order = np.lexsort((iu[0], iu[1]))           # column-major upper triangle
            M.t.copy_(torch.from_numpy(np.ascontiguousarray(A[iu[0][order], iu[1][order]])).to(M.t.dtype))
        else:
            M.t.copy_(torch.from_numpy(np.ascontiguousarray(A.T)).to(M.t.dtype))
        return M

    def _index(self, r, c):
        if r > c:
            r, c = c, r
        return r + c * (c + 1) // 2

    def numpy(self):
        if not self.packed:
            return self.t.cpu().numpy().T.copy()
        p = self.t.cpu().numpy()
        n = self.NR
        A = np.zeros((n, n), dtype=p.dtype)
        for c in range(n):
            A[:c + 1, c] = p[c * (c + 1) // 2:c * (c + 1) // 2 + c + 1]
        return np.triu(A) + np.triu(A, 1).T

    def ExportToHDF5(self, dest, Name):
        """HMatrix::ExportToHDF5 (HDF5IO.cc:171-250)"""
        ExportToHDF5(dest, Name, self.numpy())

    @classmethod
    def FromHDF5(cls, FileName, Name):
        """HMatrix(FileName, LHM_HDF5, Name) (HDF5IO.cc:252-418)"""
        return cls.from_numpy(ImportFromHDF5(FileName, Name))

    def GetEntry(self, r, c):
        v = self.t[self._index(r, c)] if self.packed else self.t[c, r]
        return complex(v.item())

    def GetEntryD(self, r, c):
        return self.GetEntry(r, c).real

    def SetEntry(self, r, c, v):
        if self.packed:
            self.t[self._index(r, c)] = v
        else:
            self.t[c, r] = v
        self.factored = False

    def _work(self):
        """the matrix the factorization works on: itself, or the full copy
        of packed storage"""
        if not self.packed:
            return self.t
        torch = _torch()
        if self._full is None:
            self._full = torch.empty((self.NC, self.NR), dtype=self.t.dtype, device="cuda")
        _sync_stream()
        _check(lib().scuffx_unpack_symmetric(_ptr(self.t), self.NR, _ptr(self._full), self.NR, self.prec))
        return self._full

    def LUFactorize(self):
        """In-place LU with partial pivoting; returns LAPACK info."""
        torch = _torch()
        if self.NR != self.NC:
            raise SCUFFXError("LUFactorize needs a square matrix")
        if self.ipiv is None:
            self.ipiv = torch.zeros(self.NR, dtype=torch.int32, device="cuda")
        A = self._work()
        _sync_stream()
        info = _check(lib().scuffx_lu_factor(_ptr(A), self.NR, self.NR, _ptr(self.ipiv), self.prec))
        self.factored = True
        return info

    def _factors(self):
        if not self.factored:
            raise SCUFFXError("LUFactorize() must be called first")
        return self._full if self.packed else self.t

    def LUSolve(self, X, Trans="N", nrhs=None):
        """Overwrites X (HVector or HMatrix) with op(A)^{-1} X."""
        if not self.factored:
            raise SCUFFXError("LUSolve called on an unfactored matrix")
        if isinstance(X, HVector):
            t, n_rhs, ldb = X.t, 1, X.N
        else:
            if X.packed:
                raise SCUFFXError("LUSolve: the right-hand sides must be in normal storage")
            t, n_rhs, ldb = X.t, (X.NC if nrhs is None else nrhs), X.NR
        if t.dtype != self.t.dtype:
            raise SCUFFXError("LUSolve: type mismatch between the factors and the right-hand side")
        if Trans != "N" and self.packed:
            raise SCUFFXError("transposed LU-solves not available for packed matrices")
        _sync_stream()
        _check(lib().scuffx_lu_solve(_ptr(self._factors()), self.NR, self.NR, _ptr(self.ipiv), Trans.encode()[:1],
                                     n_rhs, _ptr(t), ldb, self.prec))
        return 0

    def LUInvert(self):
        """Replaces the matrix with its inverse, after LUFactorize()
        (HMatrix::LUInvert, LBWrappers.cc:251-335); returns LAPACK info."""
        if not self.factored:
            raise SCUFFXError("LUFactorize() must be called before LUInvert()")
        F = self._factors()
        _sync_stream()
        info = _check(lib().scuffx_lu_invert(_ptr(F), self.NR, self.NR, _ptr(self.ipiv), self.prec))
        if info == 0 and self.packed:
            # the inverse of a symmetric matrix is symmetric: back to packed storage
            torch = _torch()
            iu = torch.triu_indices(self.NR, self.NR, device="cuda")
            order = torch.argsort(iu[1] * self.NR + iu[0])
            self.t.copy_(F[iu[1][order], iu[0][order]])
        self.factored = False
        return info

    def LogDet(self):
        """log det of the factored matrix as a complex number
        (log |det| + i arg det, arg modulo 2 pi)"""
        if not self.factored:
            raise SCUFFXError("LogDet needs LUFactorize() first")
        la, ph = C.c_double(), C.c_double()
        _sync_stream()
        _check(lib().scuffx_lu_logdet(_ptr(self._factors()), self.NR, self.NR, _ptr(self.ipiv), self.prec,
                                      C.byref(la), C.byref(ph)))
        return complex(la.value, ph.value)

    def TraceSolve(self, D):
        """sum of the first D.NC diagonal entries of A^{-1} D with the
        factors of A; D (NR x NC with NC <= NR, same type) is overwritten by
        A^{-1} D."""
        if not self.factored:
            raise SCUFFXError("TraceSolve needs LUFactorize() first")
        if D.NR != self.NR or D.NC > self.NR or D.prec != self.prec or D.packed:
            raise SCUFFXError("TraceSolve: D must have NR rows, at most NR columns and the same type")
        tr, ti = C.c_double(), C.c_double()
        _sync_stream()
        _check(lib().scuffx_trace_solve(_ptr(self._factors()), self.NR, self.NR, _ptr(self.ipiv), _ptr(D.t), D.NR,
                                        D.NC, self.prec, C.byref(tr), C.byref(ti)))
        return complex(tr.value, ti.value)


def GetLNDetMInvMInf(M, MInf):
    """energy integrand of the Casimir caller (CasimirIntegrand.cc:51-95,
    method 1): -log|det(M^{-1} M_inf)| / (2 pi).  M is LU-factored; MInf is
    the LU-factored block-diagonal matrix or its log|det| as a number."""
    import math
    la = MInf if isinstance(MInf, (int, float)) else MInf.LogDet().real
    v = la - M.LogDet().real
    return 0.0 if not math.isfinite(v) else -v / (2.0 * math.pi)


def GetTraceMInvdM(M, dM):
    """force integrand of the Casimir caller (CasimirIntegrand.cc:100-146):
    -2 Re Tr(M^{-1} dM) / (2 pi) with M LU-factored; dM is overwritten."""
    import math
    t = 2.0 * M.TraceSolve(dM).real
    return 0.0 if not math.isfinite(t) else -t / (2.0 * math.pi)


def PreloadCache(FileName):
    """Fill the process-wide FIPPI cache from a file in the reference's
    format (PreloadCache, libscuff.h:662; FIPPICache.cc:304-400).  As the
    reference does, an unreadable or malformed file only warns and the
    cache stays as it was.  Returns the number of records read."""
    import warnings
    n = lib().scuffx_preload_cache(str(FileName).encode())
    if n < 0:
        warnings.warn("PreloadCache: " + lib().scuffx_last_error().decode())
        return 0
    return n


def StoreCache(FileName):
    """Write the process-wide FIPPI cache to a file in the reference's
    format (StoreCache, libscuff.h:663; FIPPICache.cc:230-302); returns the
    number of records written."""
    return _check(lib().scuffx_store_cache(str(FileName).encode()))


def CacheSize():
    return lib().scuffx_cache_size()


def ClearCache():
    """Empty the process-wide FIPPI cache (no reference counterpart)."""
    return _check(lib().scuffx_clear_cache())


class PlaneWave:
    """Plane wave E = E0 exp(i k nHat.x) in the exterior medium
    (libs/libIncField/PlaneWave.cc:71-87)."""

    def __init__(self, E0, nHat):
        self.E0 = np.asarray(E0, dtype=np.complex128)
        self.nHat = np.asarray(nHat, dtype=np.float64)


def OpenHDF5Context(FileName):
    """creates (truncates) an HDF5 file for several datasets
    (HMatrix::OpenHDF5Context, libs/libhmat/HDF5IO.cc:120-145)"""
    ctx = C.c_void_p()
    _check(lib().scuffx_hdf5_open(FileName.encode(), C.byref(ctx)))
    return ctx


def CloseHDF5Context(ctx):
    _check(lib().scuffx_hdf5_close(ctx))


def ExportToHDF5(dest, Name, A):
    """writes a host array in the reference's layout: 2-D as an HMatrix
    (column-major, HDF5IO.cc:171-236), 1-D as an HVector (:436-475); dest
    is a context from OpenHDF5Context or a file name (a file of its own)"""
    A = np.asarray(A)
    cplx = np.iscomplexobj(A)
    own = isinstance(dest, str)
    ctx = OpenHDF5Context(dest) if own else dest
    try:
        if A.ndim == 2:
            data = np.asfortranarray(A, dtype=np.complex128 if cplx else np.float64)
            _check(lib().scuffx_hdf5_write_matrix(ctx, Name.encode(), data.ctypes.data_as(C.c_void_p),
                                                  A.shape[0], A.shape[1], int(cplx)))
        elif A.ndim == 1:
            data = np.ascontiguousarray(A, dtype=np.complex128 if cplx else np.float64)
            _check(lib().scuffx_hdf5_write_vector(ctx, Name.encode(), data.ctypes.data_as(C.c_void_p),
                                                  A.shape[0], int(cplx)))
        else:
            raise SCUFFXError("only matrices and vectors are exported")
    finally:
        if own:
            CloseHDF5Context(ctx)


def ImportFromHDF5(FileName, Name, Vector=False):
    """reads a dataset written in the reference's layout: a matrix (2-D
    array) or, with Vector=True, a vector (HDF5IO.cc:252-418, :497-540)"""
    c = C.c_int()
    if Vector:
        n = C.c_int()
        _check(lib().scuffx_hdf5_read_vector(FileName.encode(), Name.encode(), None, C.byref(n), C.byref(c)))
        out = np.zeros(n.value, dtype=np.complex128 if c.value else np.float64)
        _check(lib().scuffx_hdf5_read_vector(FileName.encode(), Name.encode(), out.ctypes.data_as(C.c_void_p),
                                             C.byref(n), C.byref(c)))
        return out
    nr, nc = C.c_int(), C.c_int()
    _check(lib().scuffx_hdf5_read_matrix(FileName.encode(), Name.encode(), None, C.byref(nr), C.byref(nc), C.byref(c)))
    out = np.zeros((nr.value, nc.value), dtype=np.complex128 if c.value else np.float64, order="F")
    _check(lib().scuffx_hdf5_read_matrix(FileName.encode(), Name.encode(), out.ctypes.data_as(C.c_void_p),
                                         C.byref(nr), C.byref(nc), C.byref(c)))
    return out


class GTComplex:
    """one entry of a .trans file: a tag and the transformations of the
    surfaces it names (GTransformation.h); index i into its GTCList"""

    def __init__(self, lst, i):
        self.List, self.Index = lst, i
        buf = C.create_string_buffer(1024)
        _check(lib().scuffx_trans_tag(lst.h, i, buf, len(buf)))
        self.Tag = buf.value.decode()

    def __repr__(self):
        return "GTComplex(%r)" % self.Tag


class GTCList(list):
    """the GTComplex entries of a .trans file, in file order"""

    def __init__(self, h):
        super().__init__()
        self.h = h
        for i in range(_check(lib().scuffx_trans_count(h))):
            self.append(GTComplex(self, i))

    def __del__(self):
        if getattr(self, "h", None) and self.h.value and _LIB is not None:
            _LIB.scuffx_trans_free(self.h)
            self.h = C.c_void_p()


def ReadTransFile(FileName=None):
    """reads a .trans file (GTransformation.cc:531-580): TRANSFORMATION tag
    ... ENDTRANSFORMATION sections and one-line TRANS entries; no file
    name gives a single identity transformation tagged DEFAULT"""
    h = C.c_void_p()
    _check(lib().scuffx_trans_open(FileName.encode() if FileName else None, C.byref(h)))
    return GTCList(h)


class RWGGeometry:
    """Surfaces, RWG edges and regions; BEM assembly on the GPU."""

    def __init__(self, GeoFile=None):
        self.h = C.c_void_p()
        if GeoFile is None:
            _check(lib().scuffx_geom_new(C.byref(self.h)))
        else:
            _check(lib().scuffx_geom_open(GeoFile.encode(), C.byref(self.h)))

    def __del__(self):
        if getattr(self, "h", None) and self.h.value and _LIB is not None:
            _LIB.scuffx_geom_free(self.h)
            self.h = C.c_void_p()

    def AddMesh(self, V, T, Material="PEC", Label=None):
        V = np.ascontiguousarray(V, dtype=np.float64)
        T = np.ascontiguousarray(T, dtype=np.int32)
        ns = _check(lib().scuffx_geom_add_mesh(self.h, V.ctypes.data_as(C.c_void_p), V.size // 3,
                                               T.ctypes.data_as(C.c_void_p), T.size // 3,
                                               Material.encode()))
        if Label is not None:
            self.SetLabel(ns, Label)
        return ns

    def AddMeshFile(self, path, Material="PEC"):
        return _check(lib().scuffx_geom_add_msh(self.h, path.encode(), Material.encode()))

    @property
    def TotalBFs(self):
        return _check(lib().scuffx_geom_total_bfs(self.h))

    @property
    def NumSurfaces(self):
        return _check(lib().scuffx_geom_num_surfaces(self.h))

    def Counts(self, ns=0):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        _check(lib().scuffx_geom_counts(self.h, ns, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def Edges(self, ns=0):
        _, _, ne = self.Counts(ns)
        out = np.zeros((ne, 9), dtype=np.int32)
        lens = np.zeros(ne)
        _check(lib().scuffx_geom_edges(self.h, ns, out.ctypes.data_as(C.c_void_p), lens.ctypes.data_as(C.c_void_p)))
        return out, lens

    def Panels(self, ns=0):
        _, npan, _ = self.Counts(ns)
        out = np.zeros((npan, 8))
        _check(lib().scuffx_geom_panels(self.h, ns, out.ctypes.data_as(C.c_void_p)))
        return out

    @property
    def Mate(self):
        """Mate[ns]: the first earlier surface identical to ns, or -1
        (RWGGeometry.cc:733-770)"""
        return [lib().scuffx_geom_mate(self.h, ns) for ns in range(self.NumSurfaces)]

    def GetEpsMu(self, Omega, Region=0):
        """(eps, mu) of a region at Omega (RegionMPs[r]->GetEpsMu,
        libMatProp.cc:501); region 0 is the exterior"""
        e, m = np.zeros(2), np.zeros(2)
        Omega = complex(Omega)
        _check(lib().scuffx_geom_epsmu(self.h, Region, Omega.real, Omega.imag, e.ctypes.data_as(C.c_void_p),
                                       m.ctypes.data_as(C.c_void_p)))
        return complex(e[0], e[1]), complex(m[0], m[1])

    def Label(self, ns=0):
        """surface label: OBJECT label, mesh file name, or ByHand"""
        buf = C.create_string_buffer(4096)
        _check(lib().scuffx_geom_label(self.h, ns, buf, len(buf)))
        return buf.value.decode()

    def SetLabel(self, ns, Label):
        _check(lib().scuffx_geom_set_label(self.h, ns, Label.encode()))

    def GetSurfaceByLabel(self, Label):
        """index of the surface with this label (case-insensitive), or -1
        (RWGGeometry.cc:868-880)"""
        for ns in range(self.NumSurfaces):
            if self.Label(ns).lower() == Label.lower():
                return ns
        return -1

    def CheckGTCList(self, GTCs):
        """raises if a transformation names no surface or an unknown one
        (RWGGeometry.cc:931-943)"""
        _check(lib().scuffx_geom_check_trans(self.h, GTCs.h))

    def Transform(self, GTC):
        """applies one .trans entry (RWGGeometry.cc:888-912): the named
        surfaces move, SurfaceMoved flags them"""
        _sync_stream_if_gpu()
        _check(lib().scuffx_geom_transform(self.h, GTC.List.h, GTC.Index))

    def UnTransform(self):
        """undoes every transformation since the last UnTransform
        (RWGGeometry.cc:917-921)"""
        _sync_stream_if_gpu()
        _check(lib().scuffx_geom_untransform(self.h))

    def TransformSurface(self, ns, R=None, D=None):
        """moves surface ns by X' = R X + D (R 3x3, default identity; D
        3-vector, default 0) as one Transform entry naming it; UnTransform
        reverts it"""
        Rm = None if R is None else np.ascontiguousarray(R, dtype=np.float64).reshape(9)
        Dv = None if D is None else np.ascontiguousarray(D, dtype=np.float64).reshape(3)
        _sync_stream_if_gpu()
        _check(lib().scuffx_geom_transform_surface(self.h, ns,
                                                   None if Rm is None else Rm.ctypes.data_as(C.c_void_p),
                                                   None if Dv is None else Dv.ctypes.data_as(C.c_void_p)))

    def AdjointContraction(self, ns, Omega, X, Y, TorqueAxes=None):
        """sum_ij conj(X_i) dM_ij/dtheta_m Y_j for a rigid displacement of
        surface ns along x, y, z and rotations about TorqueAxes (through the
        origin): complex array of 3 + len(TorqueAxes) values.  X, Y:
        HVectors of TotalBFs (the GradM / dMdT path, libscuff.h:500-505,
        contracted without storing dM)."""
        nt = 0 if TorqueAxes is None else len(TorqueAxes)
        gam = np.zeros(9 * max(nt, 1))
        for t in range(nt):
            gam[9 * t:9 * t + 9] = CreateGammaMatrix(TorqueAxes[t]).ravel(order="F")
        out = np.zeros(2 * (3 + nt))
        Omega = complex(Omega)
        _sync_stream()
        _check(lib().scuffx_adjoint(self.h, ns, Omega.real, Omega.imag, _ptr(X.t), _ptr(Y.t), nt,
                                    gam.ctypes.data_as(C.c_void_p) if nt else None, out.ctypes.data_as(C.c_void_p)))
        return out[0::2] + 1j * out[1::2]

    @property
    def SurfaceMoved(self):
        return [bool(_check(lib().scuffx_geom_surface_moved(self.h, ns))) for ns in range(self.NumSurfaces)]

    def Vertices(self, ns=0):
        """vertex coordinates of surface ns, shape (NumVertices, 3)"""
        nv, _, _ = self.Counts(ns)
        out = np.zeros((nv, 3))
        _check(lib().scuffx_geom_vertices(self.h, ns, out.ctypes.data_as(C.c_void_p)))
        return out

    def SetFarPrecision(self, bits):
        """64 (default) or 32: precision of the far panel pairs in later
        assemblies (32: the mixed form's single-precision far pairs)"""
        _check(lib().scuffx_geom_set_far_precision(self.h, int(bits)))

    def SetVertices(self, ns, V):
        """new vertex coordinates (NumVertices x 3) for surface ns, same
        topology; the next assembly rebuilds the device geometry"""
        Vc = np.ascontiguousarray(V, dtype=np.float64).reshape(-1, 3)
        _sync_stream_if_gpu()
        _check(lib().scuffx_geom_set_vertices(self.h, ns, Vc.ctypes.data_as(C.c_void_p), Vc.shape[0]))

    def _vertex_split(self, g):
        """(slots,) array -> list of per-surface (NumVertices, 3) arrays"""
        out, o = [], 0
        for ns in range(self.NumSurfaces):
            nv = self.Counts(ns)[0]
            out.append(g[o:o + 3 * nv].reshape(nv, 3))
            o += 3 * nv
        return out

    def VertexGradient(self, Omega, M, u=None, v=None, C=None):
        """sum_ij C_ij dM_ij / d(vertex coordinates), C = u v^T or a dense
        complex128 matrix: a list over surfaces of complex (NumVertices, 3)
        arrays.  M: the matrix assembled at Omega (HMatrix, complex128)."""
        torch = _torch()
        n = _check(lib().scuffx_vertex_slots(self.h))
        g = torch.zeros(n, dtype=torch.complex128, device="cuda")
        Omega = complex(Omega)
        N = self.TotalBFs
        ut = None if u is None else (u.t if isinstance(u, HVector) else u).contiguous()
        vt = None if v is None else (v.t if isinstance(v, HVector) else v).contiguous()
        Ct = None if C is None else (C.t if isinstance(C, HMatrix) else C)
        _sync_stream()
        _check(lib().scuffx_vertex_gradient(self.h, Omega.real, Omega.imag, _ptr(M.t), N,
                                            None if ut is None else _ptr(ut), None if vt is None else _ptr(vt),
                                            None if Ct is None else _ptr(Ct), N, _ptr(g)))
        return self._vertex_split(g.cpu().numpy())

    def RHSVertexGradient(self, Omega, IF, mu):
        """sum_i mu_i d RHS_i / d(vertex coordinates) for the plane wave IF"""
        torch = _torch()
        n = _check(lib().scuffx_vertex_slots(self.h))
        g = torch.zeros(n, dtype=torch.complex128, device="cuda")
        E = np.zeros(6)
        E[0::2] = IF.E0.real
        E[1::2] = IF.E0.imag
        nh = np.ascontiguousarray(IF.nHat)
        Omega = complex(Omega)
        mt = (mu.t if isinstance(mu, HVector) else mu).contiguous()
        _sync_stream()
        _check(lib().scuffx_rhs_vertex_gradient(self.h, Omega.real, Omega.imag, E.ctypes.data_as(C.c_void_p),
                                                nh.ctypes.data_as(C.c_void_p), _ptr(mt), _ptr(g)))
        return self._vertex_split(g.cpu().numpy())

    def PAbsVertexGradient(self, KN):
        """(explicit derivative of GetPower's OPFT PAbs at fixed KN, per
        surface (NumVertices, 3) real arrays; gamma, a complex128 device
        tensor with dPAbs = Re(gamma^T dKN))"""
        torch = _torch()
        n = _check(lib().scuffx_vertex_slots(self.h))
        g = torch.zeros(n, dtype=torch.float64, device="cuda")
        gam = torch.zeros(self.TotalBFs, dtype=torch.complex128, device="cuda")
        _sync_stream()
        _check(lib().scuffx_pabs_vertex_gradient(self.h, _ptr(KN.t), _ptr(g), _ptr(gam)))
        return self._vertex_split(g.cpu().numpy()), gam

    def AllocateBEMMatrix(self, PureImagFreq=False, Packed=False):
        """RWGGeometry::AllocateBEMMatrix (AssembleBEMMatrix.cc:778-783): a
        real matrix (LHM_REAL) for pure imaginary frequencies of a compact
        geometry, else complex; Packed gives symmetric packed storage."""
        N = self.TotalBFs
        return HMatrix(N, N, dtype="float64" if PureImagFreq else "complex128",
                       Storage="symmetric" if Packed else "normal")

    def BuildFIPPITable(self):
        _sync_stream()
        return _check(lib().scuffx_fippi_build(self.h))

    def FIPPIRecords(self):
        n = lib().scuffx_fippi_count(self.h)
        out = np.zeros((n, 36))
        if n:
            _check(lib().scuffx_fippi_records(self.h, out.ctypes.data_as(C.c_void_p)))
        return out

    def AssembleBEMMatrix(self, Omega, M=None):
        """fills M (AssembleBEMMatrix.cc:684-773).  A complex128 matrix gets
        the entries; a real one (AllocateBEMMatrix(PureImagFreq=True)) their
        real parts, as AddEntry on a real HMatrix keeps (HMatrix.cc:419-435);
        packed storage keeps the upper triangle."""
        N = self.TotalBFs
        if M is None:
            M = HMatrix(N, N)
        elif M.NR != N or M.NC != N or M.prec == 32:
            # wrong-size matrix: a new one is returned, as the reference does
            # (AssembleBEMMatrix.cc:700-705)
            M = HMatrix(N, N)
        Omega = complex(Omega)
        _sync_stream()
        if M.prec == 8:
            _check(lib().scuffx_assemble_real(self.h, Omega.real, Omega.imag, _ptr(M.t), N, int(M.packed)))
        elif M.packed:
            # complex symmetric packed storage: assembled full, then packed
            full = HMatrix(N, N)
            _check(lib().scuffx_assemble(self.h, Omega.real, Omega.imag, _ptr(full.t), N))
            torch = _torch()
            iu = torch.triu_indices(N, N, device="cuda")
            order = torch.argsort(iu[1] * N + iu[0])
            M.t.copy_(full.t[iu[1][order], iu[0][order]])
        else:
            _check(lib().scuffx_assemble(self.h, Omega.real, Omega.imag, _ptr(M.t), N))
        M.factored = False
        return M

    def BFOffset(self, ns):
        """(first row of surface ns in the BEM matrix, its number of basis
        functions) -- BFIndexOffset and NumBFs of the reference"""
        off, nbf = C.c_int(), C.c_int()
        _check(lib().scuffx_geom_bf_range(self.h, ns, C.byref(off), C.byref(nbf)))
        return off.value, nbf.value

    def BlockSizes(self, nsa, nsb):
        return self.BFOffset(nsa)[1], self.BFOffset(nsb)[1]

    def AssembleBEMMatrixBlockReal(self, nsa, nsb, Omega, GradM=False, TorqueAxes=None, Packed=False):
        """the real counterpart of AssembleBEMMatrixBlock for pure imaginary
        frequencies (the LHM_REAL T, U and dU blocks of the Casimir caller,
        CreateSC3Data.cc:100-157): real parts of the block and of its
        derivatives; Packed (diagonal blocks only) gives symmetric packed
        storage.  Returns B, or (B, GradB), or (B, GradB, dBdT)."""
        torch = _torch()
        na, nb = self.BlockSizes(nsa, nsb)
        B = HMatrix(na, nb, dtype="float64", Storage="symmetric" if Packed else "normal")
        nt = 0 if TorqueAxes is None else len(TorqueAxes)
        want = GradM or nt > 0
        G3 = torch.zeros((3, nb, na), dtype=torch.float64, device="cuda") if want else None
        T3 = torch.zeros((max(nt, 1), nb, na), dtype=torch.float64, device="cuda") if nt else None
        gam = np.zeros(9 * max(nt, 1))
        for t in range(nt):
            gam[9 * t:9 * t + 9] = CreateGammaMatrix(TorqueAxes[t]).ravel(order="F")
        Omega = complex(Omega)
        _sync_stream()
        _check(lib().scuffx_assemble_block_real(self.h, nsa, nsb, Omega.real, Omega.imag, _ptr(B.t), na,
                                                int(B.packed), _ptr(G3) if want else None, na, na * nb, nt,
                                                gam.ctypes.data_as(C.c_void_p) if nt else None,
                                                _ptr(T3) if nt else None))
        if not want:
            return B
        grads = [HMatrix(na, nb, data=G3[m]) for m in range(3)]
        if nt == 0:
            return B, grads
        return B, grads, [HMatrix(na, nb, data=T3[t]) for t in range(nt)]

    def AssembleBEMMatrixBlock(self, nsa, nsb, Omega, GradM=False, TorqueAxes=None):
        """Block (nsa rows, nsb columns) of the BEM matrix
        (RWGGeometry::AssembleBEMMatrixBlock, libscuff.h:500-505); nsa == nsb
        gives the diagonal block.  For distinct surfaces,
        GradM=True adds its derivatives with respect to a rigid displacement
        of surface nsa along x, y, z; TorqueAxes (up to three 3-vectors) adds
        the derivatives with respect to a rotation of surface nsa about each
        axis through the origin (dMdT).
        Returns B, or (B, [dB/dx, dB/dy, dB/dz]), or (B, GradB, [dB/dtheta...])."""
        torch = _torch()
        na, nb = self.BlockSizes(nsa, nsb)
        B = HMatrix(na, nb)
        nt = 0 if TorqueAxes is None else len(TorqueAxes)
        want = GradM or nt > 0
        G3 = torch.zeros((3, nb, na), dtype=torch.complex128, device="cuda") if want else None
        T3 = torch.zeros((max(nt, 1), nb, na), dtype=torch.complex128, device="cuda") if nt else None
        gam = np.zeros(9 * max(nt, 1))
        for t in range(nt):
            gam[9 * t:9 * t + 9] = CreateGammaMatrix(TorqueAxes[t]).ravel(order="F")
        Omega = complex(Omega)
        _sync_stream()
        _check(lib().scuffx_assemble_block(self.h, nsa, nsb, Omega.real, Omega.imag, _ptr(B.t), na,
                                           _ptr(G3) if want else None, na, na * nb, nt,
                                           gam.ctypes.data_as(C.c_void_p) if nt else None,
                                           _ptr(T3) if nt else None))
        if not want:
            return B
        grads = [HMatrix(na, nb, data=G3[m]) for m in range(3)]
        if nt == 0:
            return B, grads
        return B, grads, [HMatrix(na, nb, data=T3[t]) for t in range(nt)]

    def AssemblyStats(self):
        ms = np.zeros(6)
        cnt = np.zeros(4, dtype=np.int64)
        _check(lib().scuffx_assemble_stats(self.h, ms.ctypes.data_as(C.c_void_p), cnt.ctypes.data_as(C.c_void_p)))
        keys = ["fippi_ms", "self_ms", "far_ms", "near_ms", "stamp_ms", "mirror_ms"]
        d = dict(zip(keys, ms.tolist()))
        d.update(dict(zip(["nfar", "nnear", "nself", "nfippi"], cnt.tolist())))
        return d

    def AssembleRHSVector(self, Omega, IF, RHS=None):
        N = self.TotalBFs
        if RHS is None:
            RHS = HVector(N)
        E = np.zeros(6)
        E[0::2] = IF.E0.real
        E[1::2] = IF.E0.imag
        nh = np.ascontiguousarray(IF.nHat)
        Omega = complex(Omega)
        _sync_stream()
        _check(lib().scuffx_rhs_planewave(self.h, Omega.real, Omega.imag, E.ctypes.data_as(C.c_void_p),
                                          nh.ctypes.data_as(C.c_void_p), _ptr(RHS.t)))
        return RHS

    def GetPower(self, Omega, KN, RHS, Method="OPFT"):
        """{'PAbs', 'PExt', 'PScat'} in the reference's power units.
        Method "OPFT": absorbed power from the crossed overlap integrals
        (OPFT.cc:310-364), PScat = PExt - PAbs.  Method "EMT": scattered
        power by energy/momentum transfer (EMTPFT.cc:468-803, the
        reference's default), PAbs = PExt - PScat; it assembles the
        exterior region's matrix (one more N x N complex128 buffer)."""
        pa, pe = C.c_double(), C.c_double()
        Omega = complex(Omega)
        _sync_stream()
        if Method.upper() == "EMT":
            torch = _torch()
            N = self.TotalBFs
            Mext = torch.empty((N, N), dtype=torch.complex128, device="cuda")
            _check(lib().scuffx_assemble_region(self.h, Omega.real, Omega.imag, 0, _ptr(Mext), N))
            ps = C.c_double()
            _check(lib().scuffx_power_emt(self.h, Omega.real, Omega.imag, _ptr(KN.t), _ptr(RHS.t), _ptr(Mext), N,
                                          C.byref(pa), C.byref(ps)))
            del Mext
            return {"PAbs": pa.value, "PScat": ps.value, "PExt": pa.value + ps.value}
        if Method.upper() != "OPFT":
            raise ValueError("Method must be 'OPFT' or 'EMT'")
        _check(lib().scuffx_power(self.h, Omega.real, Omega.imag, _ptr(KN.t), _ptr(RHS.t), C.byref(pa), C.byref(pe)))
        return {"PAbs": pa.value, "PExt": pe.value, "PScat": pe.value - pa.value}

    def AssembleRegionMatrix(self, Omega, Region=0):
        """the BEM matrix with one region's contributions only (0: the
        exterior), complex128"""
        torch = _torch()
        N = self.TotalBFs
        M = HMatrix(N, N, dtype="complex128")
        Omega = complex(Omega)
        _sync_stream()
        _check(lib().scuffx_assemble_region(self.h, Omega.real, Omega.imag, Region, _ptr(M.t), N))
        return M


def BatchSweep(G, Omegas, IF, prec=64):
    """power at every frequency of Omegas for the plane wave IF (the
    scattering caller's frequency loop, scuff-scatter.cc:373-567):
    array (len(Omegas), 3) of PAbs, PExt, PScat.  prec=32 factorizes in
    complex64 and refines in complex128."""
    om = np.asarray([complex(w) for w in Omegas], dtype=np.complex128)
    E0 = np.ascontiguousarray(np.asarray(IF.E0, dtype=np.complex128)).view(np.float64)
    nh = np.ascontiguousarray(IF.nHat, dtype=np.float64)
    out = np.zeros((len(om), 3))
    _sync_stream()
    _check(lib().scuffx_batch_sweep(G.h, om.view(np.float64).ctypes.data_as(C.c_void_p), len(om),
                                    E0.ctypes.data_as(C.c_void_p), nh.ctypes.data_as(C.c_void_p), prec,
                                    out.ctypes.data_as(C.c_void_p)))
    return out


def SweepBatched(Gs, Omegas, IF, RefineSteps=5, Timing=None):
    """the frequency sweep of several geometries at once: every (geometry,
    frequency) instance assembled, factored in complex64 by one batched LU
    per geometry, solved and refined in complex128, powered -- one captured
    hipGraph (scuffx_sweep_batched).  Returns (P, relres, info): P of shape
    (len(Gs), len(Omegas), 3) = PAbs, PExt, PScat; relres the final
    residuals; info = (instances solved again step by step, LU block width
    chosen for the first geometry)."""
    Gs = list(Gs)
    om = np.asarray([complex(w) for w in Omegas], dtype=np.complex128)
    E0 = np.ascontiguousarray(np.asarray(IF.E0, dtype=np.complex128)).view(np.float64)
    nh = np.ascontiguousarray(IF.nHat, dtype=np.float64)
    hs = (C.c_void_p * len(Gs))(*[G.h.value for G in Gs])
    out = np.zeros((len(Gs), len(om), 3))
    rel = np.zeros((len(Gs), len(om)))
    info = np.zeros(2, dtype=np.int32)
    ms = np.zeros(5)
    _sync_stream()
    _check(lib().scuffx_sweep_batched(hs, len(Gs), om.view(np.float64).ctypes.data_as(C.c_void_p), len(om),
                                      E0.ctypes.data_as(C.c_void_p), nh.ctypes.data_as(C.c_void_p), RefineSteps,
                                      out.ctypes.data_as(C.c_void_p), rel.ctypes.data_as(C.c_void_p),
                                      info.ctypes.data_as(C.c_void_p), ms.ctypes.data_as(C.c_void_p)))
    if Timing is not None:
        Timing.update(zip(("setup_ms", "capture_ms", "instantiate_ms", "run_ms", "results_ms"), ms.tolist()))
    return out, rel, (int(info[0]), int(info[1]))


PROFILED_KERNELS = ("gemm_c128", "gemm_c64", "panel_c128", "panel_c64", "trsm_c128", "trsm_c64",
                    "far", "near", "self", "stamp", "fippi", "resid", "grad", "gemm_f64", "panel_f64")


def profile(on=True):
    """Time every launch of the hot kernels with HIP events on their
    stream (resets the totals)."""
    _check(lib().scuffx_profile(1 if on else 0))


def profile_read():
    """{kernel: (ms, work, launches)} accumulated since profile(True)."""
    out = {}
    for k in PROFILED_KERNELS:
        ms, w, n = C.c_double(), C.c_double(), C.c_long()
        _check(lib().scuffx_profile_read(k.encode(), C.byref(ms), C.byref(w), C.byref(n)))
        out[k] = (ms.value, w.value, n.value)
    return out


def solve_mixed(M, b, tol=1e-13, maxiter=30, method="ir"):
    """Solve M x = b with a complex64 LU of M and complex128 refinement;
    returns (x, iterations, final relative residual |b - M x|_inf / |b|_inf).
    method "ir": classical refinement, continued by GMRES-IR if it stalls
    above tol; "gmres": GMRES-IR from the start (maxiter bounds its inner
    steps).  Raises if the tolerance is not reached."""
    torch = _torch()
    N = M.NR
    LU32 = torch.empty((N, N), dtype=torch.complex64, device="cuda")
    ipiv = torch.zeros(N, dtype=torch.int32, device="cuda")
    _sync_stream()
    info = _check(lib().scuffx_lu32_from_c128(_ptr(M.t), N, N, _ptr(LU32), N, _ptr(ipiv)))
    if info > 0:
        raise SCUFFXError("singular complex64 factor (info=%d)" % info)
    x = HVector(N)
    it, rel = C.c_int(), C.c_double()
    fn = {"ir": lib().scuffx_refine, "gmres": lib().scuffx_gmres_ir}[method]
    rc = _check(fn(_ptr(M.t), N, N, _ptr(LU32), N, _ptr(ipiv), _ptr(b.t), _ptr(x.t),
                   tol, maxiter, C.byref(it), C.byref(rel)))
    if rc == NOT_CONVERGED:
        raise SCUFFXError(lib().scuffx_last_error().decode())
    return x, it.value, rel.value
