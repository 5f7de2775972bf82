"""GPU parity: libfemhot (through the C ABI, via folax_amd.context) against the FP64 oracle.

Bars: FP64 kernels within 1e-12 (norm-wise relative error vs the oracle), FP32 kernels within
1e-5 of the FP64 oracle; CSR pattern and colouring bit-exact.
"""
import numpy as np
import pytest

from oracle import fe
from tests.cases import make_case

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

KINDS = ["thermal", "mechanical", "neohooke"]
ETYPES = ["quad", "hexahedron", "tetra"]
ALL = [(k, e) for k in KINDS for e in ETYPES]
TOL = {"f64": 1e-12, "f32": 1e-5}
TDT = {"f64": torch.float64, "f32": torch.float32}


def rel(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def dev(x, dt):
    return torch.as_tensor(np.asarray(x), dtype=TDT[dt], device="cuda")


def ctx_for(case):
    from tests.cases import femhot_context
    return femhot_context(case)


def thermal_params(kind):
    return dict(beta=0.5, c=2.0) if kind == "thermal" else {}


@pytest.mark.parametrize("dt", ["f64", "f32"])
@pytest.mark.parametrize("kind,etype", ALL)
def test_residual_and_energy(kind, etype, dt):
    case = make_case(kind, etype, n=3, seed=1, **thermal_params(kind))
    print(f"[residual] {kind} {etype} {dt}", flush=True)
    ph = case["ph"]
    X = case["coords"][case["conn"]]
    dofs = fe.element_dofs(case["conn"], ph.ndpn)
    e_el, r_el, _ = ph.element(X, case["field"][case["conn"]], case["u"][dofs], want_K=False)
    R_ref = fe.scatter_nodal(case["conn"], r_el, case["coords"].shape[0], ph.ndpn)
    ctx = ctx_for(case)
    R, en = ctx.residual(dev(case["field"], dt)[None], dev(case["u"], dt)[None])
    torch.cuda.synchronize()
    assert rel(R.cpu().numpy(), R_ref) <= TOL[dt]
    assert rel(en.cpu().numpy().sum(), e_el.sum()) <= TOL[dt]
    # BC-masked residual (fe_loss.py:191-207)
    Rm, _ = ctx.residual(dev(case["field"], dt)[None], dev(case["u"], dt)[None], want_energy=False, flags=0x1)
    R_m = R_ref.copy()
    R_m[case["didx"]] = 0.0
    assert rel(Rm.cpu().numpy(), R_m) <= TOL[dt]


@pytest.mark.parametrize("dt", ["f64", "f32"])
@pytest.mark.parametrize("kind,etype", ALL)
def test_residual_vjp(kind, etype, dt):
    case = make_case(kind, etype, n=3, seed=2, **thermal_params(kind))
    print(f"[vjp] {kind} {etype} {dt}", flush=True)
    ph = case["ph"]
    ctx = ctx_for(case)
    for picard in ([False, True] if kind == "thermal" else [False]):
        gu_ref, gd_ref = fe.global_residual_vjp(ph, case["coords"], case["conn"], case["field"], case["u"], case["w"],
                                                picard=picard)
        gu, gd = ctx.residual_vjp(dev(case["field"], dt)[None], dev(case["u"], dt)[None], dev(case["w"], dt)[None],
                                  flags=0x4 if picard else 0)
        torch.cuda.synchronize()
        assert rel(gu.cpu().numpy(), gu_ref) <= TOL[dt], picard
        assert rel(gd.cpu().numpy(), gd_ref) <= TOL[dt], picard


@pytest.mark.parametrize("dt", ["f64", "f32"])
@pytest.mark.parametrize("kind,etype", ALL)
def test_energy_grad(kind, etype, dt):
    case = make_case(kind, etype, n=3, seed=3)
    print(f"[egrad] {kind} {etype} {dt}", flush=True)
    ph = case["ph"]
    Pi_ref, R_ref, gd_ref = fe.global_energy_grad(ph, case["coords"], case["conn"], case["field"], case["u"])
    ctx = ctx_for(case)
    en, gu, gd = ctx.energy_grad(dev(case["field"], dt)[None], dev(case["u"], dt)[None])
    torch.cuda.synchronize()
    assert rel(gu.cpu().numpy(), R_ref) <= TOL[dt]
    assert rel(gd.cpu().numpy(), gd_ref) <= TOL[dt]
    assert rel(en.cpu().numpy().sum(), Pi_ref) <= TOL[dt]


@pytest.mark.parametrize("dt", ["f64", "f32"])
@pytest.mark.parametrize("kind,etype", ALL)
def test_tangent_csr(kind, etype, dt):
    case = make_case(kind, etype, n=3, seed=4, **thermal_params(kind))
    print(f"[tangent] {kind} {etype} {dt}", flush=True)
    ph = case["ph"]
    A, R_ref = fe.assemble(ph, case["coords"], case["conn"], case["field"], case["u"], case["didx"])
    ip, ii = fe.csr_pattern(case["conn"], ph.ndpn, case["coords"].shape[0])
    ctx = ctx_for(case)
    rp, ci = ctx.pattern()
    assert np.array_equal(rp.cpu().numpy(), ip)
    assert np.array_equal(ci.cpu().numpy().astype(np.int64), ii)
    assert np.array_equal(A.indptr, ip) and np.array_equal(A.indices, ii)
    for flags in (0, 0x20, 0x8, 0x80, 0x88):  # atomic, colour-ordered, transposed, row-owner (+ transposed)
        vals, R = ctx.tangent_csr(dev(case["field"], dt), dev(case["u"], dt), flags=flags)
        torch.cuda.synchronize()
        assert rel(vals.cpu().numpy(), A.data) <= TOL[dt], flags
        assert rel(R.cpu().numpy(), R_ref) <= TOL[dt], flags
    # ACCUMULATE adds the assembly into existing values, in every assembly order
    base = torch.linspace(-1.0, 1.0, A.data.size, dtype=torch.float64 if dt == "f64" else torch.float32,
                          device="cuda")
    for flags in (0x10, 0x30, 0x90):
        out = base.clone()
        ctx.tangent_csr(dev(case["field"], dt), dev(case["u"], dt), want_R=False, flags=flags, out_vals=out)
        torch.cuda.synchronize()
        assert rel(out.cpu().numpy(), base.cpu().numpy() + A.data) <= TOL[dt], flags


@pytest.mark.parametrize("kind,etype", [("mechanical", "hexahedron"), ("thermal", "quad"), ("mechanical", "tetra")])
def test_colouring_bit_exact(kind, etype):
    case = make_case(kind, etype, n=4, seed=5)
    ctx = ctx_for(case)
    col = ctx.colouring()
    ref = fe.greedy_colouring(case["conn"], case["coords"].shape[0])
    assert np.array_equal(col, ref)


@pytest.mark.parametrize("dt", ["f64", "f32"])
@pytest.mark.parametrize("kind,etype", [("mechanical", "tetra"), ("neohooke", "hexahedron"), ("thermal", "quad")])
def test_scatter_colour_matches_atomic(kind, etype, dt):
    case = make_case(kind, etype, n=4, seed=6)
    ctx = ctx_for(case)
    f, u, w = dev(case["field"], dt)[None], dev(case["u"], dt)[None], dev(case["w"], dt)[None]
    Ra, ea = ctx.residual(f, u)
    Rc, ec = ctx.residual(f, u, flags=0x20)
    ga, da = ctx.residual_vjp(f, u, w)
    gc, dc = ctx.residual_vjp(f, u, w, flags=0x20)
    torch.cuda.synchronize()
    tol = 1e-14 if dt == "f64" else 1e-6
    assert rel(Rc.cpu().numpy(), Ra.cpu().numpy()) <= tol
    assert rel(gc.cpu().numpy(), ga.cpu().numpy()) <= tol
    assert rel(dc.cpu().numpy(), da.cpu().numpy()) <= tol
    assert rel(ec.cpu().numpy(), ea.cpu().numpy()) <= tol


@pytest.mark.parametrize("dt", ["f64", "f32"])
def test_batched_samples(dt):
    case = make_case("thermal", "quad", n=6, seed=7, beta=0.5, c=2.0)
    ph = case["ph"]
    rng = np.random.default_rng(3)
    B = 5
    F = rng.uniform(0.1, 1.0, size=(B, case["coords"].shape[0]))
    U = rng.uniform(0.0, 1.0, size=(B, ph.ndpn * case["coords"].shape[0]))
    W = rng.standard_normal(U.shape)
    ctx = ctx_for(case)
    R, en = ctx.residual(dev(F, dt), dev(U, dt))
    gu, gd = ctx.residual_vjp(dev(F, dt), dev(U, dt), dev(W, dt))
    torch.cuda.synchronize()
    for b in range(B):
        X = case["coords"][case["conn"]]
        dofs = fe.element_dofs(case["conn"], 1)
        e_el, r_el, _ = ph.element(X, F[b][case["conn"]], U[b][dofs], want_K=False)
        assert rel(R[b].cpu().numpy(), fe.scatter_nodal(case["conn"], r_el, case["coords"].shape[0], 1)) <= TOL[dt]
        assert rel(en[b].cpu().numpy().sum(), e_el.sum()) <= TOL[dt]
        g1, g2 = fe.global_residual_vjp(ph, case["coords"], case["conn"], F[b], U[b], W[b])
        assert rel(gu[b].cpu().numpy(), g1) <= TOL[dt]
        assert rel(gd[b].cpu().numpy(), g2) <= TOL[dt]


@pytest.mark.parametrize("dt", ["f64", "f32"])
def test_element_field_and_segments(dt):
    """Uneven meshes as one disjoint mesh with per-segment energies and a per-element field."""
    from folax_amd.context import FemhotContext
    parts = [fe.box_tet_mesh(n, 1.0, jitter=0.2, seed=10 + n) for n in (2, 3, 4)]
    coords = np.concatenate([p[0] for p in parts])
    noff = np.cumsum([0] + [p[0].shape[0] for p in parts])
    conn = np.concatenate([p[1] + noff[i] for i, p in enumerate(parts)]).astype(np.int32)
    eoff = np.cumsum([0] + [p[1].shape[0] for p in parts])
    rng = np.random.default_rng(11)
    E = np.exp(rng.uniform(np.log(1.0), np.log(100.0), size=conn.shape[0]))
    u = rng.uniform(-0.01, 0.01, size=3 * coords.shape[0])
    ph = fe.Physics("mechanical", "tetra", E=1.0, nu=0.3)
    ctx = FemhotContext("tetra", 1, coords, conn, young_modulus=1.0, poisson_ratio=0.3, segment_offsets=eoff)
    en, gu, gd = ctx.energy_grad(dev(E, dt)[None], dev(u, dt)[None], flags=0x2)
    torch.cuda.synchronize()
    X = coords[conn]
    Pi, r, gde = fe.energy_grad(ph, X, E, u[fe.element_dofs(conn, 3)])
    R_ref = fe.scatter_nodal(conn, r, coords.shape[0], 3)
    assert rel(gu.cpu().numpy(), R_ref) <= TOL[dt]
    assert rel(gd.cpu().numpy(), gde) <= TOL[dt]
    seg_ref = np.array([Pi[eoff[s]:eoff[s + 1]].sum() for s in range(3)])
    assert rel(en.cpu().numpy().ravel(), seg_ref) <= TOL[dt]


@pytest.mark.parametrize("cp", [False, True])
@pytest.mark.parametrize("mode", ["f64", "mixed"])
@pytest.mark.parametrize("kind,etype", [("thermal", "quad"), ("mechanical", "quad"), ("mechanical", "hexahedron"),
                                        ("mechanical", "tetra")])
def test_pcg_against_direct_solve(kind, etype, mode, cp):
    case = make_case(kind, etype, n=5, seed=8)
    ph = case["ph"]
    ub = fe.apply_dirichlet(case["u"], case["didx"], case["dval"])
    A, R = fe.assemble(ph, case["coords"], case["conn"], case["field"], ub, case["didx"])
    import scipy.sparse.linalg as spla
    du_ref = spla.spsolve(A.tocsc(), -R)
    ctx = ctx_for(case)
    rp, ci = ctx.pattern()
    vals, Rg = ctx.tangent_csr(dev(case["field"], "f64"), dev(ub, "f64"))
    x, it, relres, ok = ctx.pcg(rp, ci, vals, -Rg, tol=1e-11, maxiter=5000, mode=mode, context_pattern=cp)
    torch.cuda.synchronize()
    assert ok and relres <= 1e-11
    assert rel(x.cpu().numpy(), du_ref) <= 1e-9


@pytest.mark.parametrize("dt", ["f64", "f32"])
@pytest.mark.parametrize("kind,etype", ALL)
def test_residual_shape_vjp(kind, etype, dt):
    """(dR/dX)^T w on the GPU (forward-mode duals) against the oracle (complex step; central
    differences for Neo-Hookean), and translation invariance of the result."""
    from tests.cases import femhot_context
    case = make_case(kind, etype, n=3, seed=51, **(dict(beta=0.5, c=2.0) if kind == "thermal" else {}))
    print(f"[shape] {kind} {etype} {dt}", flush=True)
    ph = case["ph"]
    ref = fe.global_residual_shape_vjp(ph, case["coords"], case["conn"], case["field"], case["u"], case["w"])
    ctx = femhot_context(case)
    gx = ctx.residual_shape_vjp(dev(case["field"], dt), dev(case["u"], dt), dev(case["w"], dt))
    torch.cuda.synchronize()
    gx = gx.cpu().numpy()
    tol = {"f64": 1e-7 if kind == "neohooke" else 1e-12, "f32": 1e-5}[dt]
    assert rel(gx, ref) <= tol
    assert np.abs(gx.sum(axis=0)).max() <= (1e-12 if dt == "f64" else 1e-5) * np.abs(gx).max() * gx.shape[0]
    if ph.dim == 2:
        assert np.all(gx[:, 2] == 0.0)
