"""GPU parity of the LDS-staged chunk path (FEMHOT_SCATTER_CHUNK) against the FP64 oracle and the
plain element-parallel path: every physics x element type, batched samples, per-element fields
with per-segment energies, Dirichlet masking, accumulation and the autotuner."""
import numpy as np
import pytest

from oracle import fe
from tests.cases import make_case

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

CHUNK = 0x40
KINDS = ["thermal", "mechanical", "neohooke"]
ETYPES = ["quad", "hexahedron", "tetra"]
TOL = {"f64": 1e-12, "f32": 1e-5}
TDT = {"f64": torch.float64, "f32": torch.float32}


def rel(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def dev(x, dt):
    return torch.as_tensor(np.asarray(x), dtype=TDT[dt], device="cuda")


@pytest.mark.parametrize("dt", ["f64", "f32"])
@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("etype", ETYPES)
def test_chunk_residual_vjp_egrad(kind, etype, dt):
    from tests.cases import femhot_context
    case = make_case(kind, etype, n=5, seed=31, **(dict(beta=0.5, c=2.0) if kind == "thermal" else {}))
    print(f"[chunk] {kind} {etype} {dt}", flush=True)
    ph = case["ph"]
    ctx = femhot_context(case)
    for op in (0, 1, 2):  # force the chunk kernel (FP64 Neo-Hookean would otherwise take the plain path)
        ctx.set_chunk_config(op, TDT[dt], 1, 128, 256, 0)
    f, u, w = dev(case["field"], dt)[None], dev(case["u"], dt)[None], dev(case["w"], dt)[None]
    X = case["coords"][case["conn"]]
    dofs = fe.element_dofs(case["conn"], ph.ndpn)
    e_el, r_el, _ = ph.element(X, case["field"][case["conn"]], case["u"][dofs], want_K=False)
    R_ref = fe.scatter_nodal(case["conn"], r_el, case["coords"].shape[0], ph.ndpn)
    R, en = ctx.residual(f, u, flags=CHUNK)
    Rm, _ = ctx.residual(f, u, flags=CHUNK | 0x1)
    gu, gd = ctx.residual_vjp(f, u, w, flags=CHUNK)
    torch.cuda.synchronize()
    assert rel(R.cpu().numpy(), R_ref) <= TOL[dt]
    assert rel(en.cpu().numpy().sum(), e_el.sum()) <= TOL[dt]
    R_m = R_ref.copy()
    R_m[case["didx"]] = 0.0
    assert rel(Rm.cpu().numpy(), R_m) <= TOL[dt]
    gu_ref, gd_ref = fe.global_residual_vjp(ph, case["coords"], case["conn"], case["field"], case["u"], case["w"])
    assert rel(gu.cpu().numpy(), gu_ref) <= TOL[dt]
    assert rel(gd.cpu().numpy(), gd_ref) <= TOL[dt]
    if kind != "thermal":
        Pi, Rr, gdr = fe.global_energy_grad(ph, case["coords"], case["conn"], case["field"], case["u"])
        en2, gu2, gd2 = ctx.energy_grad(f, u, flags=CHUNK)
        torch.cuda.synchronize()
        assert rel(gu2.cpu().numpy(), Rr) <= TOL[dt]
        assert rel(gd2.cpu().numpy(), gdr) <= TOL[dt]
        assert rel(en2.cpu().numpy().sum(), Pi) <= TOL[dt]


@pytest.mark.parametrize("dt", ["f64", "f32"])
def test_chunk_batched_accumulate_and_epc(dt):
    from tests.cases import femhot_context
    case = make_case("thermal", "quad", n=12, seed=32, beta=0.5, c=2.0)
    ph = case["ph"]
    rng = np.random.default_rng(4)
    B = 6
    Nn = case["coords"].shape[0]
    F = rng.uniform(0.1, 1.0, size=(B, Nn))
    U = rng.uniform(0.0, 1.0, size=(B, Nn))
    W = rng.standard_normal((B, Nn))
    ctx = femhot_context(case)
    gu0, gd0 = ctx.residual_vjp(dev(F, dt), dev(U, dt), dev(W, dt))  # plain path
    gu1, gd1 = ctx.residual_vjp(dev(F, dt), dev(U, dt), dev(W, dt), flags=CHUNK)
    # accumulate onto a non-zero start
    import ctypes
    from folax_amd import _lib as L
    from folax_amd.context import _ptr, _stream
    acc_u = dev(np.ones((B, Nn)), dt)
    acc_d = dev(np.full((B, Nn), 2.0), dt)
    Ft, Ut, Wt = dev(F, dt), dev(U, dt), dev(W, dt)  # keep the inputs alive across the raw call
    L.check(ctx._lib.femhot_residual_vjp(ctx._h, L.F32 if dt == "f32" else L.F64, B, _ptr(Ft), Nn, _ptr(Ut), Nn,
                                         _ptr(Wt), Nn, _ptr(acc_u), Nn, _ptr(acc_d), Nn, CHUNK | L.ACCUMULATE,
                                         _stream()))
    torch.cuda.synchronize()
    for b in range(B):
        g1, g2 = fe.global_residual_vjp(ph, case["coords"], case["conn"], F[b], U[b], W[b])
        assert rel(gu1[b].cpu().numpy(), g1) <= TOL[dt]
        assert rel(gd1[b].cpu().numpy(), g2) <= TOL[dt]
        assert rel(gu0[b].cpu().numpy(), g1) <= TOL[dt]
        assert rel(acc_u[b].cpu().numpy() - 1.0, g1) <= 10 * TOL[dt]
        assert rel(acc_d[b].cpu().numpy() - 2.0, g2) <= 10 * TOL[dt]
    # autotuner keeps results valid and the tuned plan gives the same answer
    epc, blk, grid, ms = ctx.autotune(L.OP_VJP, dev(F, dt), dev(U, dt), dev(W, dt))
    assert ms > 0
    assert (epc >= 128 and (blk & ~L.CHUNK_DMA) in (128, 256, 512)) or (epc == 0 and blk in (L.PATH_ATOMIC,
                                                                                                L.PATH_COLOUR))
    gu2, gd2 = ctx.residual_vjp(dev(F, dt), dev(U, dt), dev(W, dt), flags=CHUNK)
    torch.cuda.synchronize()
    assert rel(gu2.cpu().numpy(), gu1.cpu().numpy()) <= TOL[dt]
    assert rel(gd2.cpu().numpy(), gd1.cpu().numpy()) <= TOL[dt]


@pytest.mark.parametrize("dt", ["f64", "f32"])
def test_chunk_uneven_meshes_element_field(dt):
    from folax_amd.context import FemhotContext
    parts = [fe.box_tet_mesh(n, 1.0, jitter=0.2, seed=40 + n) for n in (3, 5, 7, 4)]
    coords = np.concatenate([p[0] for p in parts])
    noff = np.cumsum([0] + [p[0].shape[0] for p in parts])
    conn = np.concatenate([p[1] + noff[i] for i, p in enumerate(parts)]).astype(np.int32)
    eoff = np.cumsum([0] + [p[1].shape[0] for p in parts])
    rng = np.random.default_rng(12)
    E = np.exp(rng.uniform(0.0, np.log(100.0), size=conn.shape[0]))
    u = rng.uniform(-0.01, 0.01, size=3 * coords.shape[0])
    ctx = FemhotContext("tetra", 1, coords, conn, young_modulus=1.0, poisson_ratio=0.3, segment_offsets=eoff)
    en, gu, gd = ctx.energy_grad(dev(E, dt)[None], dev(u, dt)[None], flags=CHUNK | 0x2)
    torch.cuda.synchronize()
    ph = fe.Physics("mechanical", "tetra", E=1.0, nu=0.3)
    Pi, r, gde = fe.energy_grad(ph, coords[conn], E, u[fe.element_dofs(conn, 3)])
    assert rel(gu.cpu().numpy(), fe.scatter_nodal(conn, r, coords.shape[0], 3)) <= TOL[dt]
    assert rel(gd.cpu().numpy(), gde) <= TOL[dt]
    seg_ref = np.array([Pi[eoff[s]:eoff[s + 1]].sum() for s in range(len(parts))])
    assert rel(en.cpu().numpy().ravel(), seg_ref) <= TOL[dt]


@pytest.mark.parametrize("dt", ["f64", "f32"])
def test_tet4_spare_rows_same_bits_across_grids(dt):
    """The TET4 chunk kernel shifts each chunk node's corner rows by its local index (one spare row per
    node) unless the rows would cost a workgroup per CU. On this mesh at 256 threads: the potential
    gradient (per-element field, 512-tet chunks) takes the shift at 2 and 4 workgroups per CU and not at
    5 or 8; the residual VJP (nodal field, so the field-gradient rows too; 256-tet chunks) at 2 and 6 and
    not at 7 or 8. The shift moves no term of any sum, so every grid gives the same bits, and the
    oracle's values."""
    from folax_amd import _lib as L
    from folax_amd.context import FemhotContext
    coords, conn, _ = fe.box_tet_mesh(12, 1.0, jitter=0.2, seed=77)
    conn = conn.astype(np.int32)
    rng = np.random.default_rng(78)
    Ee = np.exp(rng.uniform(0.0, np.log(100.0), size=conn.shape[0]))
    En = np.exp(rng.uniform(0.0, np.log(100.0), size=coords.shape[0]))
    u = rng.uniform(-0.01, 0.01, size=3 * coords.shape[0])
    w = rng.uniform(-1.0, 1.0, size=3 * coords.shape[0])
    ctx = FemhotContext("tetra", L.MECHANICAL, coords, conn, young_modulus=1.0, poisson_ratio=0.3)
    outs = []
    for grid, vgrid in ((2, 2), (4, 6), (5, 7), (8, 8)):
        print(f"[spare rows] {dt} 512 x 256 x {grid}, 256 x 256 x {vgrid}", flush=True)
        ctx.set_chunk_config(L.OP_ENERGY_GRAD, TDT[dt], 1, 512, 256, grid, flags=L.ELEMENT_FIELD)
        ctx.set_chunk_config(L.OP_VJP, TDT[dt], 1, 256, 256, vgrid)
        en, gu, gd = ctx.energy_grad(dev(Ee, dt)[None], dev(u, dt)[None], flags=CHUNK | L.ELEMENT_FIELD)
        vu, vd = ctx.residual_vjp(dev(En, dt)[None], dev(u, dt)[None], dev(w, dt)[None], flags=CHUNK)
        torch.cuda.synchronize()
        outs.append([x.clone() for x in (en, gu, gd, vu, vd)])
    for o in outs[1:]:
        for a, b in zip(outs[0], o):
            assert torch.equal(a, b)
    ph = fe.Physics("mechanical", "tetra", E=1.0, nu=0.3)
    Pi, r, gde = fe.energy_grad(ph, coords[conn], Ee, u[fe.element_dofs(conn, 3)])
    en, gu, gd, vu, vd = outs[0]
    assert rel(gu.cpu().numpy(), fe.scatter_nodal(conn, r, coords.shape[0], 3)) <= TOL[dt]
    assert rel(gd.cpu().numpy(), gde) <= TOL[dt]
    assert rel(en.cpu().numpy().sum(), Pi.sum()) <= TOL[dt]
    vu_ref, vd_ref = fe.global_residual_vjp(ph, coords, conn, En, u, w)
    assert rel(vu.cpu().numpy(), vu_ref) <= TOL[dt]
    assert rel(vd.cpu().numpy(), vd_ref) <= TOL[dt]
    ctx.close()


def test_chunk_flag_takes_plain_path_for_fp64_neohooke():
    """Untuned FEMHOT_SCATTER_CHUNK calls in FP64 Neo-Hookean take the plain path: same numbers as
    an explicit plain call, and as the chunk kernel when it is forced."""
    from tests.cases import femhot_context
    case = make_case("neohooke", "hexahedron", n=4, seed=5)
    ctx = femhot_context(case)
    f, u = dev(case["field"], "f64")[None], dev(case["u"], "f64")[None]
    en_c, gu_c, gd_c = ctx.energy_grad(f, u, flags=CHUNK)
    en_p, gu_p, gd_p = ctx.energy_grad(f, u, flags=0)
    ctx.set_chunk_config(2, torch.float64, 1, 64, 128, 0)
    en_k, gu_k, gd_k = ctx.energy_grad(f, u, flags=CHUNK)
    torch.cuda.synchronize()
    for a, b in ((gu_c, gu_p), (gd_c, gd_p), (en_c, en_p), (gu_k, gu_p), (gd_k, gd_p), (en_k, en_p)):
        assert rel(a.cpu().numpy(), b.cpu().numpy()) <= 1e-12


def test_scatter_path_recorded_and_taken():
    """FEMHOT_SCATTER_CHUNK calls take the path the tuning table holds for their shape: a recorded
    colour-ordered path gives the colour path's bits, and the tangent takes what autotune measured."""
    from folax_amd import _lib as L
    from tests.cases import femhot_context
    case = make_case("mechanical", "hexahedron", n=4, seed=6)
    ctx = femhot_context(case)
    f, u = dev(case["field"], "f64")[None], dev(case["u"], "f64")[None]
    COLOUR = L.SCATTER_COLOUR
    en_c, gu_c, gd_c = ctx.energy_grad(f, u, flags=COLOUR)
    ctx.set_chunk_config(L.OP_ENERGY_GRAD, torch.float64, 1, 0, L.PATH_COLOUR, 0)
    en_k, gu_k, gd_k = ctx.energy_grad(f, u, flags=CHUNK)
    torch.cuda.synchronize()
    assert torch.equal(gu_k, gu_c) and torch.equal(gd_k, gd_c)  # the colour path is deterministic
    Pi, R, gd = fe.global_energy_grad(case["ph"], case["coords"], case["conn"], case["field"], case["u"])
    assert rel(gu_k.cpu().numpy(), R) <= 1e-12 and rel(gd_k.cpu().numpy(), gd) <= 1e-12
    path, ms = ctx.autotune_tangent(f[0], u[0])
    assert path in (L.PATH_ATOMIC, L.PATH_COLOUR, L.PATH_ROWS) and ms > 0
    vk, Rk = ctx.tangent_csr(f, u, flags=CHUNK)
    vc, Rc = ctx.tangent_csr(f, u, flags={L.PATH_COLOUR: COLOUR, L.PATH_ROWS: L.SCATTER_ROWS}.get(path, 0))
    torch.cuda.synchronize()
    if path in (L.PATH_COLOUR, L.PATH_ROWS):  # both orders are fixed: same bits
        assert torch.equal(vk, vc)
    assert rel(vk.cpu().numpy(), vc.cpu().numpy()) <= 1e-12
    # the whole autotuner reports a consistent winner and leaves valid outputs
    epc, blk, grid, ms = ctx.autotune(L.OP_RESIDUAL, f, u)
    assert (epc >= 128 and (blk & ~L.CHUNK_DMA) in (128, 256, 512)) or (epc == 0 and blk in (L.PATH_ATOMIC,
                                                                                                L.PATH_COLOUR))
    R2, _ = ctx.residual(f, u, flags=CHUNK)
    torch.cuda.synchronize()
    assert rel(R2.cpu().numpy(), R) <= 1e-12
