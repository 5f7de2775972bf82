"""ctypes binding of libfemhot.so (the C ABI declared in include/femhot.h).

The library is built in-tree (folax_amd/libfemhot.so, see folax_amd/build.py).  There is no CPU
fallback: if the shared library is missing or fails to load, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libfemhot.so")

QUAD4, HEX8, TET4 = 0, 1, 2
THERMAL, MECHANICAL, NEOHOOKE = 0, 1, 2
F32, F64 = 0, 1

MASK_DIRICHLET = 0x1
ELEMENT_FIELD = 0x2
PICARD = 0x4
TRANSPOSE = 0x8
ACCUMULATE = 0x10
SCATTER_COLOUR = 0x20
SCATTER_CHUNK = 0x40
SCATTER_ROWS = 0x80  # tangent: row-owner assembly (each CSR value summed in a fixed order, written once)
CHUNK_DMA = 0x10000  # block | CHUNK_DMA: the pipelined LDS-DMA chunk kernel (FP32)
PCG_CONTEXT_PATTERN = 0x100  # femhot_pcg mode flag: the matrix is on this context's pattern

OK, ERR_ARG, ERR_SHAPE, ERR_INDEX, ERR_NONFINITE, ERR_HIP, ERR_NOT_CONVERGED, ERR_UNSUPPORTED = range(8)

ELEMENT_TYPES = {"quad": QUAD4, "hexahedron": HEX8, "tetra": TET4}

# every symbol include/femhot.h declares
EXPORTED = [
    "femhot_last_error", "femhot_version", "femhot_ctx_create", "femhot_ctx_destroy", "femhot_ctx_info",
    "femhot_pattern", "femhot_colouring", "femhot_residual", "femhot_tangent_csr", "femhot_residual_vjp",
    "femhot_energy_grad", "femhot_pcg", "femhot_spmv", "femhot_orient_elements", "femhot_autotune",
    "femhot_kernel_timing", "femhot_set_chunk_config", "femhot_residual_shape_vjp", "femhot_newton",
    "femhot_gauss_point_data", "femhot_energy_grad_noalloc", "femhot_pcg_async", "femhot_newton_async",
    "femhot_chunk_plan_check",
]

OP_RESIDUAL, OP_VJP, OP_ENERGY_GRAD, OP_TANGENT = 0, 1, 2, 3
PATH_ATOMIC, PATH_COLOUR = 1, 2  # autotune's block when an element-parallel path won (epc = 0)
PATH_ROWS = 3  # ... and for the tangent, when the row-owner assembly won


def gauss_points_per_element(elem_type, num_gp):
    """Points of Gauss rule num_gp (1/2/3) per element (femhot_gauss_point_data's G)."""
    if elem_type == TET4:
        return {1: 1, 2: 4, 3: 8}[num_gp]
    return num_gp ** (2 if elem_type == QUAD4 else 3)


class FemhotError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"femhot error {code}: {msg}")
        self.code = code


class MeshDesc(C.Structure):
    _fields_ = [
        ("elem_type", C.c_int), ("physics", C.c_int), ("num_gp", C.c_int),
        ("young_modulus", C.c_double), ("poisson_ratio", C.c_double), ("beta", C.c_double), ("c", C.c_double),
        ("body_force", C.c_double * 3),
        ("n_nodes", C.c_longlong), ("n_elems", C.c_longlong),
        ("coords", C.POINTER(C.c_double)), ("conn", C.POINTER(C.c_int)),
        ("n_dirichlet", C.c_longlong), ("dirichlet_idx", C.POINTER(C.c_longlong)),
        ("dirichlet_val", C.POINTER(C.c_double)),
        ("n_segments", C.c_int), ("segment_offsets", C.POINTER(C.c_longlong)),
    ]


_lib = None


def lib():
    """Load (once) and return the library handle; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run `python -c 'import folax_amd.build as b; b.build()'`")
    # torch's ROCm wheel carries its own HIP runtime (file libamdhip64.so, SONAME libamdhip64.so.7).
    # Loading torch first makes the library's libamdhip64.so.7 resolve to that same copy; loading the
    # library first would bring in /opt/rocm's runtime and leave two HIP runtimes in the process.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    vp, ll, i32, u32, dbl = C.c_void_p, C.c_longlong, C.c_int, C.c_uint, C.c_double
    L.femhot_last_error.restype = C.c_char_p
    L.femhot_version.restype = C.c_char_p
    L.femhot_ctx_create.argtypes = [C.POINTER(MeshDesc), C.POINTER(vp)]
    L.femhot_ctx_destroy.argtypes = [vp]
    L.femhot_ctx_destroy.restype = None
    L.femhot_ctx_info.argtypes = [vp, C.POINTER(ll), C.POINTER(i32), C.POINTER(i32), C.POINTER(ll), C.POINTER(i32),
                                  C.POINTER(i32)]
    L.femhot_pattern.argtypes = [vp, vp, vp, vp]
    L.femhot_colouring.argtypes = [vp, C.POINTER(i32)]
    L.femhot_residual.argtypes = [vp, i32, i32, vp, ll, vp, ll, vp, ll, vp, u32, vp]
    L.femhot_tangent_csr.argtypes = [vp, i32, vp, vp, vp, vp, u32, vp]
    L.femhot_residual_vjp.argtypes = [vp, i32, i32, vp, ll, vp, ll, vp, ll, vp, ll, vp, ll, u32, vp]
    L.femhot_energy_grad.argtypes = [vp, i32, i32, vp, ll, vp, ll, vp, vp, ll, vp, ll, u32, vp]
    L.femhot_energy_grad_noalloc.argtypes = [vp, i32, i32, vp, ll, vp, ll, vp, vp, ll, vp, ll, u32, vp]
    L.femhot_residual_shape_vjp.argtypes = [vp, i32, vp, vp, vp, vp, u32, vp]
    L.femhot_gauss_point_data.argtypes = [vp, i32, vp, vp, vp, vp]
    L.femhot_pcg.argtypes = [vp, i32, vp, vp, vp, vp, vp, dbl, i32, C.POINTER(i32), C.POINTER(dbl), vp]
    L.femhot_pcg_async.argtypes = [vp, i32, vp, vp, vp, vp, vp, dbl, i32, vp, vp, vp, vp]
    L.femhot_spmv.argtypes = [vp, vp, vp, vp, vp, vp, vp]
    L.femhot_newton.argtypes = [vp, i32, vp, vp, u32, i32, i32, dbl, dbl, dbl, i32, C.POINTER(dbl), C.POINTER(i32), vp]
    L.femhot_newton_async.argtypes = [vp, vp, vp, u32, i32, i32, dbl, dbl, dbl, i32, vp, vp, vp, vp]
    L.femhot_autotune.argtypes = [vp, i32, i32, i32, vp, ll, vp, ll, vp, ll, vp, ll, vp, ll, vp, u32,
                                   C.POINTER(i32), C.POINTER(i32), C.POINTER(i32), C.POINTER(C.c_float), vp]
    L.femhot_orient_elements.argtypes = [i32, ll, C.POINTER(dbl), ll, C.POINTER(i32), C.POINTER(ll)]
    L.femhot_chunk_plan_check.argtypes = [C.POINTER(MeshDesc), i32, C.POINTER(ll)]
    L.femhot_kernel_timing.argtypes = [vp, i32, C.POINTER(dbl), C.POINTER(ll)]
    L.femhot_set_chunk_config.argtypes = [vp, i32, i32, i32, u32, i32, i32, i32]
    for name in EXPORTED:
        if name not in ("femhot_last_error", "femhot_version", "femhot_ctx_destroy"):
            getattr(L, name).restype = C.c_int
    _lib = L
    return L


def check(rc):
    if rc != OK:
        raise FemhotError(rc, lib().femhot_last_error().decode(errors="replace"))


def orient_elements(elem_type, coords, conn):
    """Mesh.CheckAndOrientElements (fol/mesh_input_output/mesh.py:123-142) through the C ABI (host
    arrays, no GPU).  Returns (new_conn, n_swapped)."""
    coords = np.ascontiguousarray(coords, dtype=np.float64)
    conn = np.array(conn, dtype=np.int32, order="C", copy=True)
    n_sw = C.c_longlong(0)
    check(lib().femhot_orient_elements(int(elem_type), coords.shape[0],
                                       coords.ctypes.data_as(C.POINTER(C.c_double)), conn.shape[0],
                                       conn.ctypes.data_as(C.POINTER(C.c_int)), C.byref(n_sw)))
    return conn, int(n_sw.value)


PLAN_STATS = ("n_chunks", "max_nodes", "max_elems", "chunk_nodes", "shared_nodes", "slots", "zero_nodes",
              "caller_order")


def chunk_plan_check(elem_type, coords, conn, epc, segment_offsets=None):
    """Test hook: the chunk plan of `epc` elements per chunk built on the host (no GPU) and checked
    against every invariant the chunk kernels index by; raises FemhotError naming the first violation.
    Returns the plan's figures as a dict (keys PLAN_STATS)."""
    coords = np.ascontiguousarray(coords, dtype=np.float64)
    if coords.shape[1] == 2:
        coords = np.concatenate([coords, np.zeros((coords.shape[0], 1))], axis=1)
    conn = np.ascontiguousarray(conn, dtype=np.int32)
    d = MeshDesc()
    d.elem_type, d.physics, d.num_gp = int(elem_type), THERMAL, 0
    d.n_nodes, d.n_elems = coords.shape[0], conn.shape[0]
    d.coords = coords.ctypes.data_as(C.POINTER(C.c_double))
    d.conn = conn.ctypes.data_as(C.POINTER(C.c_int))
    d.n_dirichlet = 0
    so = None
    if segment_offsets is not None:
        so = np.ascontiguousarray(segment_offsets, dtype=np.int64)
        d.n_segments = so.shape[0] - 1
        d.segment_offsets = so.ctypes.data_as(C.POINTER(C.c_longlong))
    else:
        d.n_segments = 1
    stats = (C.c_longlong * 8)()
    check(lib().femhot_chunk_plan_check(C.byref(d), int(epc), stats))
    return dict(zip(PLAN_STATS, [int(v) for v in stats]))
