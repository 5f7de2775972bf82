"""GPU parity of the software-pipelined chunk kernel (LDS-DMA staging one step ahead, FP32): it must
give bit-identical outputs to the staged chunk kernel with the same chunks and workgroup size, and
match the FP64 oracle to the FP32 bar, across element types and physics, samples, plan-buffer
alternation (a persistent grid of 1, 2 or 5 workgroups walking many steps), caller-order and
permuted (Morton) chunks, per-element and nodal fields, Dirichlet masking and accumulation."""
import numpy as np
import pytest

from oracle import fe
from tests.cases import make_case

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

CHUNK, MASK, EFIELD, ACC = 0x40, 0x1, 0x2, 0x10
DMA = 0x10000
TOL = 1e-5


def rel(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def dev(x):
    return torch.as_tensor(np.asarray(x), dtype=torch.float32, device="cuda")


def configure(ctx, B, epc, block, grid, dma, flags=0):
    for op in (0, 1, 2):
        ctx.set_chunk_config(op, torch.float32, B, epc, block | (DMA if dma else 0), grid, flags=flags)


def run_ops(ctx, f, u, w, flags, with_egrad):
    R, en = ctx.residual(f, u, flags=CHUNK | flags)
    gu, gd = ctx.residual_vjp(f, u, w, flags=CHUNK | flags)
    out = [R, en, gu, gd]
    if with_egrad:
        out += list(ctx.energy_grad(f, u, flags=CHUNK | flags))
    torch.cuda.synchronize()
    return [t.cpu().numpy().copy() for t in out]


SIZES = {"quad": 20, "hexahedron": 7, "tetra": 5}


@pytest.mark.parametrize("grid", [-1, -2, -5])
@pytest.mark.parametrize("kind", ["thermal", "mechanical", "neohooke"])
@pytest.mark.parametrize("etype", ["quad", "hexahedron", "tetra"])
def test_dma_matches_staged_and_oracle(etype, kind, grid):
    from tests.cases import femhot_context
    extra = dict(beta=0.5, c=2.0) if kind == "thermal" else {}
    if kind == "neohooke":  # nodal displacements well below the element size: no inverted elements
        extra = dict(u_scale=0.1 / SIZES[etype])
    case = make_case(kind, etype, n=SIZES[etype], seed=61, **extra)
    print(f"[dma] {kind} {etype} grid {grid}", flush=True)
    ph = case["ph"]
    ctx = femhot_context(case)
    B = 3
    rng = np.random.default_rng(7)
    F = np.stack([case["field"] * s for s in (1.0, 0.8, 1.3)])
    U = np.stack([case["u"] * s for s in (1.0, -0.7, 0.4)])
    W = rng.standard_normal((B, case["u"].size))
    f, u, w = dev(F), dev(U), dev(W)
    egrad = kind != "thermal"
    for flags in (0, MASK):
        configure(ctx, B, 32, 256, 0, dma=False)
        ref = run_ops(ctx, f, u, w, flags, egrad)
        configure(ctx, B, 32, 256, grid, dma=True)
        got = run_ops(ctx, f, u, w, flags, egrad)
        for r, g in zip(ref, got):
            assert np.isfinite(r).all()
            assert np.array_equal(r, g), f"flags {flags}: DMA kernel differs from the staged kernel"
    # and the FP64 oracle, every sample (unmasked outputs from the last DMA run are masked; redo)
    R, en, gu, gd = run_ops(ctx, f, u, w, 0, False)
    dofs = fe.element_dofs(case["conn"], ph.ndpn)
    X = case["coords"][case["conn"]]
    for b in range(B):
        e_el, r_el, _ = ph.element(X, F[b][case["conn"]], U[b][dofs], want_K=False)
        assert rel(R[b], fe.scatter_nodal(case["conn"], r_el, case["coords"].shape[0], ph.ndpn)) <= TOL
        assert rel(en[b], e_el.sum()) <= TOL
        g1, g2 = fe.global_residual_vjp(ph, case["coords"], case["conn"], F[b], U[b], W[b])
        assert rel(gu[b], g1) <= TOL
        assert rel(gd[b], g2) <= TOL


def test_dma_permuted_elements_element_field_accumulate():
    """Elements in random order: the planner takes spatial (Morton) chunks, so the kernel gathers the
    per-element field through the permutation; per-element field gradient and accumulation."""
    from folax_amd import _lib as L
    from folax_amd.context import FemhotContext, _ptr, _stream
    coords, conn, _ = fe.box_tet_mesh(7, 1.0, jitter=0.2, seed=3)
    rng = np.random.default_rng(5)
    conn = np.ascontiguousarray(conn[rng.permutation(conn.shape[0])]).astype(np.int32)
    Ne, Nn = conn.shape[0], coords.shape[0]
    B = 2
    E = np.exp(rng.uniform(0.0, np.log(100.0), size=(B, Ne)))
    U = rng.uniform(-0.01, 0.01, size=(B, 3 * Nn))
    ctx = FemhotContext("tetra", L.MECHANICAL, coords, conn, young_modulus=1.0, poisson_ratio=0.3)
    Et, Ut = dev(E), dev(U)
    configure(ctx, B, 64, 256, 0, dma=False, flags=EFIELD)
    en0, gu0, gd0 = [t.cpu().numpy().copy() for t in ctx.energy_grad(Et, Ut, flags=CHUNK | EFIELD)]
    configure(ctx, B, 64, 256, -3, dma=True, flags=EFIELD)
    en1, gu1, gd1 = [t.cpu().numpy().copy() for t in ctx.energy_grad(Et, Ut, flags=CHUNK | EFIELD)]
    assert np.array_equal(gu0, gu1) and np.array_equal(gd0, gd1) and np.array_equal(en0, en1)
    ph = fe.Physics("mechanical", "tetra", E=1.0, nu=0.3)
    for b in range(B):
        Pi, r, gde = fe.energy_grad(ph, coords[conn], E[b], U[b][fe.element_dofs(conn, 3)])
        assert rel(gu1[b], fe.scatter_nodal(conn, r, Nn, 3)) <= TOL
        assert rel(gd1[b], gde) <= TOL
        assert rel(en1[b], Pi.sum()) <= TOL
    # accumulate onto non-zero outputs (of the outputs' own magnitude, so FP32 keeps the increment)
    # through the raw entry point
    start_u, start_d = 0.5 * gu1, -0.5 * gd1
    acc_u = dev(start_u)
    acc_d = dev(start_d)
    en = torch.zeros(B, dtype=torch.float64, device="cuda")
    L.check(ctx._lib.femhot_energy_grad(ctx._h, L.F32, B, _ptr(Et), Ne, _ptr(Ut), 3 * Nn, _ptr(en), _ptr(acc_u),
                                        3 * Nn, _ptr(acc_d), Ne, CHUNK | EFIELD | ACC, _stream()))
    torch.cuda.synchronize()
    assert rel(acc_u.cpu().numpy(), start_u + gu1) <= TOL
    assert rel(acc_d.cpu().numpy(), start_d + gd1) <= TOL


def test_dma_config_rejected_for_fp64_and_small_block():
    from tests.cases import femhot_context
    from folax_amd import _lib as L
    case = make_case("thermal", "quad", n=4, seed=1)
    ctx = femhot_context(case)
    with pytest.raises(Exception):
        ctx.set_chunk_config(1, torch.float64, 1, 64, 256 | DMA, 0)
    with pytest.raises(Exception):
        ctx.set_chunk_config(1, torch.float32, 1, 64, 128 | DMA, 0)
    assert L.CHUNK_DMA == DMA


def test_dma_kernel_graph_replays_match_eager():
    """The pipelined kernel captured in a HIP graph: three replays, with eager work and synchronisation
    between them, each equal bit for bit to the eager call."""
    from folax_amd import _lib as L
    from folax_amd.context import FemhotContext
    coords, conn, _ = fe.box_tet_mesh(8, 1.0, jitter=0.2, seed=11)
    conn = np.asarray(conn, dtype=np.int32)
    rng = np.random.default_rng(12)
    E = dev(np.exp(rng.uniform(0.0, np.log(100.0), size=(1, conn.shape[0]))))
    U = dev(rng.uniform(-0.01, 0.01, size=(1, 3 * coords.shape[0])))
    ctx = FemhotContext("tetra", L.MECHANICAL, coords, conn, young_modulus=1.0, poisson_ratio=0.3)
    configure(ctx, 1, 128, 256, -7, dma=True, flags=EFIELD)
    en = torch.empty(1, dtype=torch.float64, device="cuda")
    gu = torch.empty(1, 3 * coords.shape[0], dtype=torch.float32, device="cuda")
    gd = torch.empty(1, conn.shape[0], dtype=torch.float32, device="cuda")

    def step():
        ctx.energy_grad(E, U, flags=CHUNK | EFIELD, out=(en, gu, gd))

    step()
    torch.cuda.synchronize()
    ref = (en.clone(), gu.clone(), gd.clone())
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        step()
    torch.cuda.current_stream().wait_stream(side)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        step()
    for _ in range(3):
        gu.zero_()
        gd.zero_()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(gu, ref[1]) and torch.equal(gd, ref[2]) and torch.equal(en, ref[0])
        _ = float((E * 2).sum())
    ctx.close()
