"""ctypes front end of the oracle library (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: the CPU restatement of the reference path used
as the checker by tests/, __graft_entry__.smoke() and the cpu_baseline leg
of bench.py.  The product (scuff-em_amd/) never imports this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_zp = np.ctypeslib.ndpointer(dtype=np.complex128, flags="F_CONTIGUOUS")
_zv = np.ctypeslib.ndpointer(dtype=np.complex128, flags="C_CONTIGUOUS")

PPI_ALGS = ["LOCUBATURE", "HOCUBATURE", "TD", "HKTD", "DESING"]
ZVAC = 376.73031346177


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    path = os.path.join(_HERE, "liboracle.so")
    if not os.path.exists(path):
        build()
    L = C.CDLL(path)
    L.o_api_geom_new.restype = C.c_void_p
    L.o_api_geom_free.argtypes = [C.c_void_p]
    L.o_api_add_surface.argtypes = [C.c_void_p, _dp, C.c_int, _ip, C.c_int, C.c_int,
                                    C.c_double, C.c_double, C.c_double, C.c_double]
    L.o_api_add_surface_msh.argtypes = [C.c_void_p, C.c_char_p, C.c_int, C.c_double,
                                        C.c_double, C.c_double, C.c_double]
    for f in ("o_api_total_bfs",):
        getattr(L, f).argtypes = [C.c_void_p]
    for f in ("o_api_num_edges", "o_api_num_panels", "o_api_num_vertices"):
        getattr(L, f).argtypes = [C.c_void_p, C.c_int]
    L.o_api_edges.argtypes = [C.c_void_p, C.c_int, _ip, _dp]
    L.o_api_panels.argtypes = [C.c_void_p, C.c_int, _dp]
    L.o_api_ppi.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                            C.c_double, C.c_int, C.c_int, _dp, _dp]
    L.o_api_assemble.argtypes = [C.c_void_p, C.c_double, C.c_double, _zp, C.c_int]
    L.o_api_assemble_rows.argtypes = [C.c_void_p, C.c_double, C.c_double, _zp, C.c_int,
                                      C.c_int, C.c_int]
    L.o_api_assemble_rows_compact.argtypes = [C.c_void_p, C.c_double, C.c_double, _zp, C.c_int,
                                              C.c_int, C.c_int]
    L.o_api_assemble_region.argtypes = [C.c_void_p, C.c_double, C.c_double, _zp, C.c_int, C.c_int]
    L.o_api_power_emt.argtypes = [C.c_void_p, C.c_double, C.c_double, _zp, _zp, _zp, C.c_int, _dp]
    L.o_api_fippi_count.argtypes = [C.c_void_p]
    L.o_api_fippi_count.restype = C.c_size_t
    L.o_api_qifippi.argtypes = [_dp, _dp, C.c_int, _dp, C.POINTER(C.c_size_t)]
    L.o_api_taylor_duffy.argtypes = [C.c_int, C.c_int, _ip, _ip, _dp, _dp, C.c_void_p, _dp,
                                     C.POINTER(C.c_int), C.c_int]
    L.o_api_zgetrf.argtypes = [C.c_int, _zp, C.c_int, _ip]
    L.o_api_zgetrs.argtypes = [C.c_char, C.c_int, C.c_int, _zp, C.c_int, _ip, _zp, C.c_int]
    L.o_api_rhs.argtypes = [C.c_void_p, C.c_double, C.c_double, _dp, _dp, _zv]
    L.o_api_power.argtypes = [C.c_void_p, C.c_double, C.c_double, _zv, _zv, _dp]
    L.o_api_mie.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int, _dp]
    L.o_api_icosphere.argtypes = [C.c_int, C.c_double, C.c_void_p, _dp, _ip]
    L.o_api_write_gmsh.argtypes = [C.c_char_p, _dp, C.c_int, _ip, C.c_int]
    L.o_api_tcr.argtypes = [C.c_int, C.c_void_p]
    L.o_api_assemble_block_grad.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int, C.c_int,
                                            C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p,
                                            C.c_void_p]
    L.o_api_num_bfs.argtypes = [C.c_void_p, C.c_int]
    L.o_api_set_mate.argtypes = [C.c_void_p, C.c_int, C.c_int]
    _CB = C.CFUNCTYPE(None, C.c_uint, C.POINTER(C.c_double), C.c_void_p, C.c_uint,
                      C.POINTER(C.c_double))
    L._integrand_type = _CB
    for f in ("o_api_hcubature", "o_api_pcubature"):
        getattr(L, f).argtypes = [C.c_uint, _CB, C.c_uint, _dp, _dp, C.c_size_t, C.c_double,
                                  C.c_double, _dp, _dp]
    _LIB = L
    return L


def tcr(order):
    """Triangle cubature rule (order 4: 6 points, order 20: 78 points) as an
    (n, 3) array of (u, v, weight)."""
    n = lib().o_api_tcr(order, None)
    out = np.zeros(3 * n)
    lib().o_api_tcr(order, out.ctypes.data_as(C.c_void_p))
    return out.reshape(n, 3)


def _cubature(which, f, fdim, xmin, xmax, maxeval, abstol, reltol):
    L = lib()
    ncalls = [0]

    def cb(ndim, x, _p, fd, fval):
        xs = np.ctypeslib.as_array(x, shape=(ndim,))
        v = f(xs.copy())
        ncalls[0] += 1
        for i in range(fd):
            fval[i] = v[i]
    cfn = L._integrand_type(cb)
    xmin = np.ascontiguousarray(xmin, dtype=np.float64)
    xmax = np.ascontiguousarray(xmax, dtype=np.float64)
    val = np.zeros(fdim)
    err = np.zeros(fdim)
    getattr(L, which)(fdim, cfn, len(xmin), xmin, xmax, maxeval, abstol, reltol, val, err)
    return val, err, ncalls[0]


def hcubature(f, fdim, xmin, xmax, maxeval=0, abstol=0.0, reltol=1e-8):
    """h-adaptive cubature of a vector integrand f(x) -> fdim values;
    returns (value, error estimate, number of integrand calls)."""
    return _cubature("o_api_hcubature", f, fdim, xmin, xmax, maxeval, abstol, reltol)


def pcubature(f, fdim, xmin, xmax, maxeval=0, abstol=0.0, reltol=1e-8):
    """p-adaptive Clenshaw-Curtis cubature; same conventions as hcubature."""
    return _cubature("o_api_pcubature", f, fdim, xmin, xmax, maxeval, abstol, reltol)


def icosphere(level, radius=1.0, center=None):
    nV, nT = 10 * 4 ** level + 2, 20 * 4 ** level
    V = np.zeros(3 * nV)
    T = np.zeros(3 * nT, dtype=np.int32)
    c = None
    if center is not None:
        cc = np.ascontiguousarray(center, dtype=np.float64)
        c = cc.ctypes.data_as(C.c_void_p)
    lib().o_api_icosphere(level, radius, c, V, T)
    return V.reshape(nV, 3), T.reshape(nT, 3)


def write_gmsh(fn, V, T):
    V = np.ascontiguousarray(V, dtype=np.float64).ravel()
    T = np.ascontiguousarray(T, dtype=np.int32).ravel()
    return lib().o_api_write_gmsh(fn.encode(), V, len(V) // 3, T, len(T) // 3)


def mie(x, m=None, nmax=0):
    """Mie (Qsca, Qext) for size parameter x; m=None selects the PEC sphere."""
    out = np.zeros(2)
    if m is None:
        lib().o_api_mie(x, -1.0, 0.0, nmax, out)
    else:
        lib().o_api_mie(x, complex(m).real, complex(m).imag, nmax, out)
    return out[0], out[1]


class Geometry:
    """Compact objects in vacuum; kind 0=PEC, 1=constant eps, 2=Drude."""

    def __init__(self):
        self.L = lib()
        self.h = self.L.o_api_geom_new()

    def __del__(self):
        if getattr(self, "h", None):
            self.L.o_api_geom_free(self.h)
            self.h = None

    def add_surface(self, V, T, kind=0, eps=1.0, wp=0.0, gamma=0.0):
        V = np.ascontiguousarray(V, dtype=np.float64).ravel()
        T = np.ascontiguousarray(T, dtype=np.int32).ravel()
        eps = complex(eps)
        r = self.L.o_api_add_surface(self.h, V, len(V) // 3, T, len(T) // 3, kind,
                                     eps.real, eps.imag, wp, gamma)
        if r < 0:
            raise RuntimeError("oracle: bad surface")
        return r

    def add_msh(self, fn, kind=0, eps=1.0, wp=0.0, gamma=0.0):
        eps = complex(eps)
        r = self.L.o_api_add_surface_msh(self.h, fn.encode(), kind, eps.real, eps.imag, wp, gamma)
        if r < 0:
            raise RuntimeError("oracle: cannot read " + fn)
        return r

    @property
    def N(self):
        return self.L.o_api_total_bfs(self.h)

    def edges(self, ns=0):
        ne = self.L.o_api_num_edges(self.h, ns)
        out = np.zeros(9 * ne, dtype=np.int32)
        lens = np.zeros(ne)
        self.L.o_api_edges(self.h, ns, out, lens)
        return out.reshape(ne, 9), lens

    def panels(self, ns=0):
        npan = self.L.o_api_num_panels(self.h, ns)
        out = np.zeros(8 * npan)
        self.L.o_api_panels(self.h, ns, out)
        return out.reshape(npan, 8)

    def ppi(self, npa, iQa, npb, iQb, k, ngrad=0, cache=True):
        H = np.zeros(4)
        G = np.zeros(12)
        k = complex(k)
        alg = self.L.o_api_ppi(self.h, npa, iQa, npb, iQb, k.real, k.imag, ngrad,
                               1 if cache else 0, H, G)
        return (H[0::2] + 1j * H[1::2], G[0::2] + 1j * G[1::2], PPI_ALGS[alg])

    def assemble(self, omega):
        N = self.N
        M = np.zeros((N, N), dtype=np.complex128, order="F")
        omega = complex(omega)
        if self.L.o_api_assemble(self.h, omega.real, omega.imag, M, N) != 0:
            raise RuntimeError("oracle: assembly failed")
        return M

    def assemble_rows(self, omega, e0, e1, M=None):
        N = self.N
        if M is None:
            M = np.zeros((N, N), dtype=np.complex128, order="F")
        omega = complex(omega)
        self.L.o_api_assemble_rows(self.h, omega.real, omega.imag, M, N, e0, e1)
        return M

    def rows(self, omega, e0, e1):
        """upper-triangle entries of the rows of edges e0..e1-1 of surface 0
        (single-surface geometries), as a (rows x N) array"""
        N = self.N
        nr = (e1 - e0) * (1 if self.num_bfs(0) == self.L.o_api_num_edges(self.h, 0) else 2)
        R = np.zeros((nr, N), dtype=np.complex128, order="F")
        omega = complex(omega)
        self.L.o_api_assemble_rows_compact(self.h, omega.real, omega.imag, R, nr, e0, e1)
        return R

    def assemble_region(self, omega, region):
        """the matrix with one region's contributions only (region 0: the
        exterior)"""
        N = self.N
        M = np.zeros((N, N), dtype=np.complex128, order="F")
        omega = complex(omega)
        if self.L.o_api_assemble_region(self.h, omega.real, omega.imag, M, N, region) != 0:
            raise RuntimeError("oracle: region assembly failed")
        return M

    def power_emt(self, omega, KN, RHS, Mext):
        """{'PAbs', 'PScat', 'PExt'} by energy/momentum transfer"""
        out = np.zeros(2)
        omega = complex(omega)
        KN = np.asfortranarray(KN, dtype=np.complex128)
        RHS = np.asfortranarray(RHS, dtype=np.complex128)
        if self.L.o_api_power_emt(self.h, omega.real, omega.imag, KN, RHS, np.asfortranarray(Mext),
                                  Mext.shape[0], out) != 0:
            raise RuntimeError("oracle: EMT power needs every surface in the exterior region")
        return {"PAbs": out[0], "PScat": out[1], "PExt": out[0] + out[1]}

    def num_bfs(self, ns):
        return self.L.o_api_num_bfs(self.h, ns)

    def assemble_block_grad(self, omega, nsa, nsb, gammas=None):
        """Block (nsa, nsb) of the BEM matrix for two distinct surfaces, its
        derivatives with respect to a rigid displacement (x, y, z) of surface
        nsa and, for each 3x3 rotation generator in `gammas`, the derivative
        with respect to a rotation of surface nsa about that axis:
        returns (B, [dB/dx, dB/dy, dB/dz], [dB/dtheta_1, ...])."""
        na, nb = self.num_bfs(nsa), self.num_bfs(nsb)
        B = np.zeros((na, nb), dtype=np.complex128, order="F")
        GB = np.zeros((3, nb, na), dtype=np.complex128)   # each GB[m].T is column-major na x nb
        nt = 0 if gammas is None else len(gammas)
        TB = np.zeros((max(nt, 1), nb, na), dtype=np.complex128)
        gam = np.zeros(9 * max(nt, 1))
        for t in range(nt):
            gam[9 * t:9 * t + 9] = np.asarray(gammas[t], dtype=np.float64).ravel(order="F")
        omega = complex(omega)
        r = self.L.o_api_assemble_block_grad(self.h, omega.real, omega.imag, nsa, nsb,
                                             B.ctypes.data_as(C.c_void_p), na,
                                             GB.ctypes.data_as(C.c_void_p), nt,
                                             gam.ctypes.data_as(C.c_void_p),
                                             TB.ctypes.data_as(C.c_void_p))
        if r != 0:
            raise RuntimeError("oracle: bad block")
        return B, [GB[m].T.copy() for m in range(3)], [TB[t].T.copy() for t in range(nt)]

    def set_mate(self, ns, nsm):
        """surface ns is identical to the earlier surface nsm: the full
        assembly copies nsm's diagonal block (AssembleBEMMatrix.cc:728-740)"""
        if self.L.o_api_set_mate(self.h, ns, nsm) != 0:
            raise ValueError("bad mate")

    def fippi_count(self):
        return self.L.o_api_fippi_count(self.h)

    def rhs_planewave(self, omega, E0=(1, 0, 0), nhat=(0, 0, 1)):
        E = np.zeros(6)
        E[0::2] = np.real(E0)
        E[1::2] = np.imag(E0)
        rhs = np.zeros(self.N, dtype=np.complex128)
        omega = complex(omega)
        self.L.o_api_rhs(self.h, omega.real, omega.imag, E, np.asarray(nhat, dtype=np.float64), rhs)
        return rhs

    def power(self, omega, KN, RHS):
        out = np.zeros(2)
        omega = complex(omega)
        self.L.o_api_power(self.h, omega.real, omega.imag,
                           np.ascontiguousarray(KN, dtype=np.complex128),
                           np.ascontiguousarray(RHS, dtype=np.complex128), out)
        return {"PAbs": out[0], "PExt": out[1], "PScat": out[1] - out[0]}


def zgetrf(A):
    A = np.asfortranarray(A, dtype=np.complex128).copy(order="F")
    n = A.shape[0]
    ipiv = np.zeros(n, dtype=np.int32)
    info = lib().o_api_zgetrf(n, A, n, ipiv)
    return A, ipiv, info


def zgetrs(LU, ipiv, B, trans="N"):
    B = np.asfortranarray(B, dtype=np.complex128).copy(order="F")
    if B.ndim == 1:
        B = B.reshape(-1, 1, order="F")
    LU = np.asfortranarray(LU, dtype=np.complex128)
    ipiv = np.ascontiguousarray(ipiv, dtype=np.int32)
    n = LU.shape[0]
    lib().o_api_zgetrs(trans.encode(), n, B.shape[1], LU, n, ipiv, B, n)
    return B


def qifippi(Va, Vb, ncv):
    out = np.zeros(36)
    ne = C.c_size_t(0)
    lib().o_api_qifippi(np.ascontiguousarray(Va, dtype=np.float64).ravel(),
                        np.ascontiguousarray(Vb, dtype=np.float64).ravel(), ncv, out, C.byref(ne))
    return out, ne.value


def taylor_duffy(ncv, PIndex, KIndex, KParam, V, Q=None, maxeval=0):
    """Known-answer interface to the Taylor-Duffy integrator.
    V = [V1, V2, V3, V2P, V3P] (5x3), Q = [Q, QP] (2x3) or None."""
    npk = len(PIndex)
    kp = np.zeros(2 * npk)
    kp[0::2] = np.real(KParam)
    kp[1::2] = np.imag(KParam)
    res = np.zeros(2 * npk)
    nc = C.c_int(0)
    qarr = None
    if Q is not None:
        qarr = np.ascontiguousarray(Q, dtype=np.float64).ravel()
    lib().o_api_taylor_duffy(ncv, npk, np.asarray(PIndex, dtype=np.int32),
                             np.asarray(KIndex, dtype=np.int32), kp,
                             np.ascontiguousarray(V, dtype=np.float64).ravel(),
                             qarr.ctypes.data_as(C.c_void_p) if qarr is not None else None,
                             res, C.byref(nc), maxeval)
    return res[0::2] + 1j * res[1::2], nc.value
