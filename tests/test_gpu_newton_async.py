"""GPU tests of femhot_newton_async, the load-stepped Newton driver
(fol/solvers/fe_nonlinear_residual_based_solver.py:76-170) with every decision taken on the device.

Pinned against the oracle's Newton iteration (iteration counts and every norm per load step, the
solution) and against femhot_newton on the same problem; a whole nonlinear solve captured into a HIP
graph replays to the bits of the eager call; a NaN residual norm is reported through the device status
with nothing after it applied."""
import numpy as np
import pytest

from oracle import fe
from tests.cases import femhot_context, make_case
from tests.test_gpu_newton import _thermal_case, oracle_newton, rel, t64

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _thermal_ctx():
    from folax_amd import _lib as L
    from folax_amd.context import FemhotContext
    m, k = _thermal_case()
    coords, conn = m.GetNodesCoordinates(), m.GetElementsNodes("quad")
    idx, val = fe.dirichlet_dofs(m.node_sets, ["T"], {"T": {"left": 1.0, "right": 0.0}})
    ctx = FemhotContext("quad", L.THERMAL, coords, conn, num_gp=2, beta=0.5, c=2.0, dirichlet_idx=idx,
                        dirichlet_val=val)
    return ctx, k, coords, conn, idx, val


def test_async_newton_matches_oracle_and_host_driver():
    """Nonlinear thermal, 3 load steps: per-step iteration counts equal the oracle's and femhot_newton's,
    every norm far above round-off equals the oracle's, and the solution matches both."""
    from folax_amd import _lib as L
    ctx, k, coords, conn, idx, val = _thermal_ctx()
    Nn = coords.shape[0]
    kw = dict(load_steps=3, max_iter=10, abs_tol=1e-9, rel_tol=1e-9, lin_tol=1e-14)
    us, _, ss = ctx.newton(t64(k), t64(np.zeros(Nn)), lin_maxiter=5000, **kw)
    ua = t64(np.zeros(Nn))
    hist, steps, status = ctx.newton_async(t64(k), ua, lin_iterations=600, **kw)
    torch.cuda.synchronize()
    hist, steps = hist.cpu().numpy(), steps.cpu().tolist()
    print(f"[newton_async] thermal steps {steps}, host-driven {ss}", flush=True)
    assert int(status.item()) == L.OK
    ph = fe.Physics("thermal", "quad", num_gp=2, beta=0.5, c=2.0)
    ur, href = oracle_newton(ph, coords, conn, k, idx, val, np.zeros(Nn), 3, 10, 1e-9, 1e-9)
    assert steps == [len(h) for h in href] == ss
    for s, h in enumerate(href):
        ref = np.array(h)
        got = hist[s, :steps[s]]
        big = ref[:, 0] > 1e-7
        assert np.allclose(got[big, 0], ref[big, 0], rtol=1e-8)
        big = ref[:, 1] > 1e-7
        assert np.allclose(got[big, 1], ref[big, 1], rtol=1e-6)
        assert np.isnan(hist[s, steps[s]:]).all()
    assert rel(ua.cpu().numpy(), ur) <= 1e-8
    assert rel(ua.cpu().numpy(), us.cpu().numpy()) <= 1e-10
    ctx.close()


def test_async_newton_replays_as_a_graph():
    """Neo-Hookean HEX8 block stretched 5 % in 2 load steps: the eager call converges to the oracle's
    Newton solution, and a graph of the whole solve (reset of u included) replays to the same bits.
    FEMHOT_SCATTER_CHUNK assembles in a fixed order (colour-ordered tangent, chunk-path residual); the
    default element-parallel atomics would make every run differ in the last bits."""
    from folax_amd import _lib as L
    case = make_case("neohooke", "hexahedron", n=3, seed=9)
    case["dval"] = np.where(case["dval"] != 0.0, 0.05, 0.0)
    ctx = femhot_context(case)
    Nn = case["coords"].shape[0]
    f, u0 = t64(case["field"]), t64(np.zeros(3 * Nn))
    kw = dict(load_steps=2, max_iter=8, abs_tol=1e-11, rel_tol=1e-12, lin_tol=1e-13, lin_iterations=300,
              flags=L.SCATTER_CHUNK)
    ue = u0.clone()
    he, se, ste = ctx.newton_async(f, ue, **kw)
    torch.cuda.synchronize()
    assert int(ste.item()) == L.OK
    he, se = he.clone(), se.clone()
    ur, href = oracle_newton(case["ph"], case["coords"], case["conn"], case["field"], case["didx"], case["dval"],
                             np.zeros(3 * Nn), 2, 8, 1e-11, 1e-12)
    assert se.cpu().tolist() == [len(h) for h in href]
    assert rel(ue.cpu().numpy(), ur) <= 1e-9
    ug = torch.empty_like(u0)
    out = (torch.empty(2, 8, 2, dtype=torch.float64, device="cuda"), torch.empty(2, dtype=torch.int32, device="cuda"),
           torch.empty(1, dtype=torch.int32, device="cuda"))

    def solve():
        ug.copy_(u0)
        ctx.newton_async(f, ug, out=out, **kw)

    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        solve()
    torch.cuda.current_stream().wait_stream(side)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        solve()
    nan_e = torch.isnan(he)
    for rep in range(3):
        ug.fill_(3.0)
        out[0].fill_(-1.0)
        out[1].fill_(-1)
        out[2].fill_(-1)
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(ug, ue), rep
        assert torch.equal(out[1], se) and int(out[2].item()) == L.OK, rep
        assert torch.equal(torch.isnan(out[0]), nan_e) and torch.equal(out[0][~nan_e], he[~nan_e]), rep
        _ = float((f * 2).sum())  # eager work between replays
    ctx.close()


def test_async_newton_nan_guard():
    """A NaN conductivity: the first residual norm is NaN, the status says so, the later load step does
    nothing, and u holds only the first load step's Dirichlet values."""
    from folax_amd import _lib as L
    ctx, k, coords, conn, idx, val = _thermal_ctx()
    kb = np.array(k)
    kb[7] = np.nan
    u = t64(np.zeros(coords.shape[0]))
    hist, steps, status = ctx.newton_async(t64(kb), u, load_steps=2, max_iter=5, lin_iterations=50)
    torch.cuda.synchronize()
    assert int(status.item()) == L.ERR_NONFINITE
    assert steps.cpu().tolist() == [1, 0]
    assert np.isnan(hist.cpu().numpy()).all()
    assert np.array_equal(u.cpu().numpy(), fe.apply_dirichlet(np.zeros(coords.shape[0]), idx, val, 0.5))
    ctx.close()
