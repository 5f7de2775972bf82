"""GPU tests of the context's cached chunk plans and of assembly on meshes with collapsed elements."""
import numpy as np
import pytest

from oracle import fe

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

CHUNK, EFIELD = 0x40, 0x2


def rel(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def test_forward_and_vjp_on_different_chunk_sizes_capture_in_one_graph():
    """The residual tuned to 512 elements per chunk and the VJP to 256: both plans stay cached, so
    alternating calls rebuild nothing and both can be captured into one HIP graph (a plan that had
    to be built during capture is refused).  Replays, with eager work between them, equal the eager
    calls bit for bit."""
    from folax_amd import _lib as L
    from folax_amd.context import FemhotContext
    coords, conn, _ = fe.box_tet_mesh(10, 1.0, jitter=0.2, seed=61)
    conn = np.asarray(conn, dtype=np.int32)
    rng = np.random.default_rng(62)
    Nn, Ne = coords.shape[0], conn.shape[0]
    E = torch.as_tensor(np.exp(rng.uniform(0.0, np.log(100.0), size=(1, Ne))), dtype=torch.float32, device="cuda")
    U = torch.as_tensor(rng.uniform(-0.01, 0.01, size=(1, 3 * Nn)), dtype=torch.float32, device="cuda")
    W = torch.as_tensor(rng.standard_normal((1, 3 * Nn)), dtype=torch.float32, device="cuda")
    ctx = FemhotContext("tetra", L.MECHANICAL, coords, conn, young_modulus=1.0, poisson_ratio=0.3)
    ctx.set_chunk_config(L.OP_RESIDUAL, torch.float32, 1, 512, 256, 4, flags=EFIELD)
    ctx.set_chunk_config(L.OP_VJP, torch.float32, 1, 256, 256, 0, flags=EFIELD)
    R = torch.empty(1, 3 * Nn, dtype=torch.float32, device="cuda")
    gu = torch.empty_like(R)
    gd = torch.empty(1, Ne, dtype=torch.float32, device="cuda")
    from folax_amd.context import _ptr, _stream

    def step():
        L.check(ctx._lib.femhot_residual(ctx._h, L.F32, 1, _ptr(E), Ne, _ptr(U), 3 * Nn, _ptr(R), 3 * Nn, None,
                                         CHUNK | EFIELD, _stream()))
        L.check(ctx._lib.femhot_residual_vjp(ctx._h, L.F32, 1, _ptr(E), Ne, _ptr(U), 3 * Nn, _ptr(W), 3 * Nn, _ptr(gu),
                                             3 * Nn, _ptr(gd), Ne, CHUNK | EFIELD, _stream()))

    step()
    torch.cuda.synchronize()
    ref = (R.clone(), gu.clone(), gd.clone())
    ph = fe.Physics("mechanical", "tetra", E=1.0, nu=0.3)
    dofs = fe.element_dofs(conn, 3)
    _, r, _ = ph.element(coords[conn], E.double().cpu().numpy()[0], U.double().cpu().numpy()[0][dofs], want_K=False)
    assert rel(R.cpu().numpy(), fe.scatter_nodal(conn, r, Nn, 3)) <= 1e-5
    g1, g2 = fe.global_residual_vjp(ph, coords, conn, E.double().cpu().numpy()[0], U.double().cpu().numpy()[0],
                                    W.double().cpu().numpy()[0])
    assert rel(gu.cpu().numpy(), g1) <= 1e-5 and rel(gd.cpu().numpy(), g2) <= 1e-5
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        step()
    torch.cuda.current_stream().wait_stream(side)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        step()
    for _ in range(3):
        R.zero_()
        gu.zero_()
        gd.zero_()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(R, ref[0]) and torch.equal(gu, ref[1]) and torch.equal(gd, ref[2])
        step()  # eager calls of both shapes between replays
        torch.cuda.synchronize()
        assert torch.equal(R, ref[0]) and torch.equal(gu, ref[1])
    ctx.close()


@pytest.mark.parametrize("kind", ["thermal", "mechanical"])
def test_tangent_with_a_collapsed_element(kind):
    """A QUAD4 collapsed to a triangle (a node listed twice): the colour-ordered request assembles
    with atomics for such meshes, and the row-owner assembly takes the element twice (once per local
    index of the repeated node); both equal the oracle's scipy assembly (duplicates summed)."""
    from tests.cases import femhot_context, make_case
    case = make_case(kind, "quad", n=3, seed=63)
    conn = case["conn"].copy()
    conn[0, 3] = conn[0, 2]  # element 0 becomes the triangle (n0, n1, n2)
    case["conn"] = conn
    ctx = femhot_context(case)
    ph = case["ph"]
    f = torch.as_tensor(case["field"], dtype=torch.float64, device="cuda")
    u = torch.as_tensor(case["u"], dtype=torch.float64, device="cuda")
    A, R = fe.assemble(ph, case["coords"], conn, case["field"], case["u"], case["didx"])
    rp, ci = ctx.pattern()
    assert np.array_equal(rp.cpu().numpy(), A.indptr) and np.array_equal(ci.cpu().numpy(), A.indices)
    for flags in (0x20, 0, 0x80):
        vals, Rg = ctx.tangent_csr(f, u, flags=flags)
        torch.cuda.synchronize()
        assert rel(vals.cpu().numpy(), A.data) <= 1e-12, flags
        assert rel(Rg.cpu().numpy(), R) <= 1e-12, flags
