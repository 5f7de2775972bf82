"""ctypes loader of oracle/libcpuref.so (TEST INFRASTRUCTURE ONLY, see oracle/__init__.py)."""
import ctypes as C
import os

import numpy as np

_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libcpuref.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        L = C.CDLL(_PATH)
        d = np.ctypeslib.ndpointer
        L.tet4_elastic_energy_grad.argtypes = [C.c_longlong, d(np.int32, flags="C"), d(np.float64, flags="C"),
                                               d(np.float64, flags="C"), d(np.float64, flags="C"), C.c_double,
                                               C.c_double, C.c_double, d(np.int64, flags="C"), d(np.float64, flags="C"),
                                               d(np.float64, flags="C"), d(np.float64, flags="C")]
        L.tet4_elastic_energy_grad.restype = None
        L.cpuref_num_threads.restype = C.c_int
        L.cpuref_set_threads.argtypes = [C.c_int]
        L.cpuref_set_threads.restype = None
        L.quad4_thermal_resnorm_vjp.argtypes = [C.c_longlong, d(np.int32, flags="C"), d(np.float64, flags="C"),
                                                C.c_longlong, C.c_int, d(np.float64, flags="C"),
                                                d(np.float64, flags="C"), C.c_double, C.c_double,
                                                d(np.uint8, flags="C"), d(np.float64, flags="C"),
                                                d(np.float64, flags="C"), d(np.float64, flags="C")]
        L.quad4_thermal_resnorm_vjp.restype = C.c_double
        _lib = L
    return _lib


def set_threads(n):
    lib().cpuref_set_threads(int(n))


def quad4_thermal_resnorm_vjp(coords, conn, k0, T, dirichlet_nodes, beta=0.5, c=2.0, out=None):
    """Nonlinear thermal QUAD4 residual-norm loss L = 1/2 sum_b |R_b|^2 over a batch (Dirichlet
    entries of R masked) and its exact gradient (g_T, g_k0), FP64, OpenMP.  k0, T: (B, n_nodes)."""
    k0 = np.ascontiguousarray(np.atleast_2d(k0), np.float64)
    T = np.ascontiguousarray(np.atleast_2d(T), np.float64)
    B, Nn = T.shape
    mask = np.zeros(Nn, np.uint8)
    mask[np.asarray(dirichlet_nodes, np.int64)] = 1
    if out is None:
        gT, gk = np.empty_like(T), np.empty_like(T)
    else:
        gT, gk = out
    R = np.empty(Nn)
    L = lib().quad4_thermal_resnorm_vjp(conn.shape[0], np.ascontiguousarray(conn, np.int32),
                                        np.ascontiguousarray(coords, np.float64), Nn, B, k0, T, float(beta),
                                        float(c), mask, R, gT, gk)
    return L, gT, gk


def d_coeffs(E, nu):
    c1 = E / ((1.0 + nu) * (1.0 - 2.0 * nu))
    return c1 * (1.0 - nu), c1 * nu, c1 * 0.5 * (1.0 - 2.0 * nu)


def tet4_elastic_energy_grad(coords, conn, E, u, nu=0.3, young=1.0, seg_of_elem=None, n_seg=1, out=None):
    """Returns (Pi per segment, R, dPi/dE per element) in FP64."""
    c2, c3, c4 = d_coeffs(young, nu)
    Ne = conn.shape[0]
    seg = np.zeros(Ne, np.int64) if seg_of_elem is None else np.ascontiguousarray(seg_of_elem, np.int64)
    if out is None:
        R = np.zeros(u.shape[0])
        gE = np.zeros(Ne)
        Pi = np.zeros(n_seg)
    else:
        R, gE, Pi = out
        R[:] = 0.0
        Pi[:] = 0.0
    lib().tet4_elastic_energy_grad(Ne, np.ascontiguousarray(conn, np.int32), np.ascontiguousarray(coords, np.float64),
                                   np.ascontiguousarray(E, np.float64), np.ascontiguousarray(u, np.float64), c2, c3,
                                   c4, seg, R, gE, Pi)
    return Pi, R, gE


def num_threads():
    return int(lib().cpuref_num_threads())
