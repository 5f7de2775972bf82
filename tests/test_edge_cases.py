"""Edge cases of the boundary: argument validation (host side, no GPU needed) and degenerate but
legal meshes on the GPU (one element, nodes in no element, empty segments, every dof or no dof
Dirichlet), each against the FP64 oracle on both scatter paths."""
import ctypes as C

import numpy as np
import pytest

from oracle import fe
from tests.cases import PHYS_CODE

KINDS_ETYPES = [(k, e) for k in ("thermal", "mechanical", "neohooke") for e in ("quad", "hexahedron", "tetra")]


def rel(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def _desc(elem_type=2, physics=1, coords=None, conn=None, didx=None, dval=None, nseg=1, segoff=None, keep=None):
    from folax_amd import _lib as L
    coords = np.ascontiguousarray(np.zeros((4, 3)) if coords is None else coords, dtype=np.float64)
    conn = np.ascontiguousarray(np.array([[0, 1, 2, 3]]) if conn is None else conn, dtype=np.int32)
    d = L.MeshDesc()
    d.elem_type, d.physics, d.num_gp = elem_type, physics, 0
    d.young_modulus, d.poisson_ratio, d.beta, d.c = 1.0, 0.3, 0.0, 1.0
    d.n_nodes, d.n_elems = coords.shape[0], conn.shape[0]
    d.coords = coords.ctypes.data_as(C.POINTER(C.c_double))
    d.conn = conn.ctypes.data_as(C.POINTER(C.c_int))
    keep.extend([coords, conn])
    if didx is not None:
        didx = np.ascontiguousarray(didx, dtype=np.int64)
        dval = np.ascontiguousarray(dval, dtype=np.float64)
        keep.extend([didx, dval])
        d.n_dirichlet = didx.shape[0]
        d.dirichlet_idx = didx.ctypes.data_as(C.POINTER(C.c_longlong))
        d.dirichlet_val = dval.ctypes.data_as(C.POINTER(C.c_double))
    d.n_segments = nseg
    if segoff is not None:
        so = np.ascontiguousarray(segoff, dtype=np.int64)
        keep.append(so)
        d.segment_offsets = so.ctypes.data_as(C.POINTER(C.c_longlong))
    return d


@pytest.mark.parametrize("case,code", [
    ("elem_type", "ERR_ARG"), ("physics", "ERR_ARG"), ("empty", "ERR_SHAPE"), ("conn_range", "ERR_INDEX"),
    ("nan_coords", "ERR_NONFINITE"), ("dirichlet_range", "ERR_INDEX"), ("seg_no_offsets", "ERR_ARG"),
    ("seg_bad_offsets", "ERR_SHAPE"), ("num_gp", "ERR_ARG")])
def test_ctx_create_rejects_bad_meshes(case, code):
    """femhot_ctx_create validates everything on the host before touching the device and returns a
    distinct status with a message (no context is created)."""
    from folax_amd import _lib as L
    keep = []
    tet = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1.0]])
    kw = dict(coords=tet, keep=keep)
    if case == "elem_type":
        d = _desc(elem_type=7, **kw)
    elif case == "physics":
        d = _desc(physics=9, **kw)
    elif case == "empty":
        d = _desc(conn=np.zeros((0, 4)), **kw)
    elif case == "conn_range":
        d = _desc(conn=np.array([[0, 1, 2, 4]]), **kw)
    elif case == "nan_coords":
        bad = tet.copy()
        bad[2, 1] = np.nan
        d = _desc(coords=bad, keep=keep)
    elif case == "dirichlet_range":
        d = _desc(didx=[0, 12], dval=[0.0, 0.0], **kw)
    elif case == "seg_no_offsets":
        d = _desc(nseg=2, **kw)
    elif case == "seg_bad_offsets":
        d = _desc(nseg=2, segoff=[0, 2, 1], **kw)
    else:
        d = _desc(**kw)
        d.num_gp = 5
    h = C.c_void_p()
    rc = L.lib().femhot_ctx_create(C.byref(d), C.byref(h))
    assert rc == getattr(L, code), (rc, L.lib().femhot_last_error())
    assert not h.value
    assert len(L.lib().femhot_last_error()) > 0


def test_solver_and_tuning_entry_points_validate_on_the_host():
    """The solver, Newton and tuning entry points reject a null context, and the Newton driver its
    settings, before any device work (no GPU needed)."""
    from folax_amd import _lib as L
    lib = L.lib()
    dbl = C.c_double()
    it = C.c_int()
    hist = (C.c_double * 4)()
    steps = (C.c_int * 2)()
    assert lib.femhot_pcg(None, L.F64, None, None, None, None, None, 1e-8, 10, C.byref(it), C.byref(dbl),
                          None) == L.ERR_ARG
    assert lib.femhot_newton(None, L.F64, None, None, 0, 1, 1, 1e-8, 1e-8, 1e-10, 10, hist, steps, None) == L.ERR_ARG
    assert lib.femhot_set_chunk_config(None, 0, L.F32, 1, 0, 256, 256, 0) == L.ERR_ARG
    assert lib.femhot_residual_shape_vjp(None, L.F64, None, None, None, None, 0, None) == L.ERR_ARG
    assert len(lib.femhot_last_error()) > 0


# ------------------------------------------------------------------------------------------------
# GPU: degenerate but legal meshes
# ------------------------------------------------------------------------------------------------
def _single_element(etype):
    if etype == "quad":
        coords, conn, _ = fe.square_quad_mesh(2, 1.0)
    elif etype == "hexahedron":
        coords, conn, _ = fe.box_hex_mesh(1, 1, 1)
    else:
        coords = np.array([[0, 0, 0], [1.1, 0.1, 0], [0.2, 0.9, 0.1], [0.1, 0.2, 1.2]])
        conn = np.array([[0, 1, 2, 3]])
    return np.asarray(coords, dtype=np.float64), np.asarray(conn, dtype=np.int32)


def _ctx(kind, etype, coords, conn, didx=None, dval=None, segoff=None):
    from folax_amd.context import FemhotContext
    ph = fe.Physics(kind, etype, E=1.7, nu=0.3)
    return ph, FemhotContext(etype, PHYS_CODE[kind], coords, conn, num_gp=ph.num_gp, young_modulus=1.7,
                             poisson_ratio=0.3, dirichlet_idx=didx, dirichlet_val=dval, segment_offsets=segoff)


@pytest.mark.gpu
@pytest.mark.parametrize("kind,etype", KINDS_ETYPES)
def test_single_element_mesh(kind, etype):
    """One element: residual, energy and residual VJP on both paths, FP64 and FP32."""
    torch = pytest.importorskip("torch")
    coords, conn = _single_element(etype)
    print(f"[edge] single {kind} {etype}", flush=True)
    ph, ctx = _ctx(kind, etype, coords, conn)
    rng = np.random.default_rng(3)
    Nn = coords.shape[0]
    d = rng.uniform(0.5, 2.0, Nn)
    u = rng.uniform(-0.05, 0.05, ph.ndpn * Nn) if kind != "thermal" else rng.uniform(0, 1, Nn)
    w = rng.standard_normal(ph.ndpn * Nn)
    e_el, r_el, _ = ph.element(coords[conn], d[conn], u[fe.element_dofs(conn, ph.ndpn)], want_K=False)
    R_ref = fe.scatter_nodal(conn, r_el, Nn, ph.ndpn)
    gu_ref, gd_ref = fe.global_residual_vjp(ph, coords, conn, d, u, w)
    for dt, tol in ((torch.float64, 1e-12), (torch.float32, 1e-5)):
        t = lambda x: torch.as_tensor(x, dtype=dt, device="cuda")[None]
        for flags in (0, 0x40):
            R, en = ctx.residual(t(d), t(u), flags=flags)
            gu, gd = ctx.residual_vjp(t(d), t(u), t(w), flags=flags)
            torch.cuda.synchronize()
            assert rel(R.cpu().numpy(), R_ref) <= tol
            assert rel(en.cpu().numpy().sum(), e_el.sum()) <= tol
            assert rel(gu.cpu().numpy(), gu_ref) <= tol
            assert rel(gd.cpu().numpy(), gd_ref) <= tol
    ctx.close()


@pytest.mark.gpu
@pytest.mark.parametrize("flags", [0, 0x40, 0x20])
def test_nodes_in_no_element_and_empty_segment(flags):
    """Two TET4 meshes with an empty segment between them and two nodes referenced by no element:
    those nodes' outputs are exactly zero (also after a call that wrote them), the empty segment's
    energy is zero, and everything else matches the oracle."""
    torch = pytest.importorskip("torch")
    c1, k1, _ = fe.box_tet_mesh(3, jitter=0.1, seed=1)
    c2, k2, _ = fe.box_tet_mesh(2, jitter=0.1, seed=2)
    orphan = np.array([[5.0, 5.0, 5.0], [6.0, 5.0, 5.0]])
    coords = np.concatenate([c1, orphan, c2])
    off2 = c1.shape[0] + 2
    conn = np.concatenate([k1, k2 + off2]).astype(np.int32)
    segoff = [0, k1.shape[0], k1.shape[0], conn.shape[0]]
    ph, ctx = _ctx("mechanical", "tetra", coords, conn, segoff=segoff)
    rng = np.random.default_rng(4)
    Ne, Nn = conn.shape[0], coords.shape[0]
    E = rng.uniform(1, 10, Ne)
    u = rng.uniform(-1e-2, 1e-2, 3 * Nn)
    dev = lambda x: torch.as_tensor(x, dtype=torch.float64, device="cuda")[None]
    # per-element field: oracle with the field taken per element
    ph1 = fe.Physics("mechanical", "tetra", E=1.7, nu=0.3)
    gu_ref = np.zeros(3 * Nn)
    gd_ref = np.zeros(Ne)
    en_ref = np.zeros(3)
    for s, (a, b) in enumerate(zip(segoff[:-1], segoff[1:])):
        if b == a:
            continue
        ks = conn[a:b]
        P, g, gdn = fe.energy_grad(ph1, coords[ks], np.repeat(E[a:b, None], 4, axis=1), u[fe.element_dofs(ks, 3)])
        en_ref[s] = P.sum()
        np.add.at(gu_ref, fe.element_dofs(ks, 3), g)
        gd_ref[a:b] = gdn.sum(1)
    gu0 = torch.full((1, 3 * Nn), 7.0, dtype=torch.float64, device="cuda")  # stale values must be cleared
    out = (torch.empty(3, dtype=torch.float64, device="cuda"), gu0, torch.empty(1, Ne, dtype=torch.float64, device="cuda"))
    en, gu, gd = ctx.energy_grad(dev(E), dev(u), flags=flags | 0x2, out=out)
    torch.cuda.synchronize()
    gu = gu.cpu().numpy().ravel()
    orphan_dofs = np.arange(3 * c1.shape[0], 3 * off2)
    assert np.all(gu[orphan_dofs] == 0.0)
    assert rel(gu, gu_ref) <= 1e-12
    assert rel(gd.cpu().numpy(), gd_ref) <= 1e-12
    en = en.cpu().numpy().ravel()
    assert en[1] == 0.0
    assert rel(en, en_ref) <= 1e-12
    ctx.close()


@pytest.mark.gpu
def test_every_dof_dirichlet_and_no_dirichlet():
    """All dofs Dirichlet: the masked residual is zero, the tangent is its diagonal, and PCG returns
    the initial guess after zero iterations (b = 0).  No Dirichlet dofs: masking changes nothing."""
    torch = pytest.importorskip("torch")
    coords, conn, _ = fe.square_quad_mesh(5, 1.0)
    Nn = coords.shape[0]
    didx = np.arange(Nn)
    dval = np.linspace(0, 1, Nn)
    ph, ctx = _ctx("thermal", "quad", coords, conn, didx=didx, dval=dval)
    rng = np.random.default_rng(5)
    k = torch.as_tensor(rng.uniform(0.5, 2, Nn), device="cuda")
    T = torch.as_tensor(rng.uniform(0, 1, Nn), device="cuda")
    R, _ = ctx.residual(k[None], T[None], flags=0x1)
    vals, R2 = ctx.tangent_csr(k, T)
    rp, ci = ctx.pattern()
    torch.cuda.synchronize()
    assert float(R.abs().max()) == 0.0 and float(R2.abs().max()) == 0.0
    A = torch.sparse_csr_tensor(rp, ci.long(), vals, (Nn, Nn)).to_dense().cpu().numpy()
    assert np.count_nonzero(A - np.diag(np.diag(A))) == 0
    x0 = torch.as_tensor(rng.standard_normal(Nn), device="cuda")
    x, it, rr, ok = ctx.pcg(rp, ci, vals, torch.zeros(Nn, dtype=torch.float64, device="cuda"), x0=x0)
    assert ok and it == 0
    assert torch.equal(x, x0)
    ctx.close()
    ph, ctx = _ctx("thermal", "quad", coords, conn)
    Ra, _ = ctx.residual(k[None], T[None], flags=0)
    Rb, _ = ctx.residual(k[None], T[None], flags=0x1)
    torch.cuda.synchronize()
    assert torch.equal(Ra, Rb)
    ctx.close()
