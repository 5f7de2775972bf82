"""femhot_pcg_async: FP64 Jacobi-PCG with the convergence test on the device.

Pinned against femhot_pcg on the same system: the device test stops at the first converged iteration
(femhot_pcg checks every 10, so it runs 0..9 iterations more), and femhot_pcg capped at that many
iterations gives the same bits; a HIP graph of the whole solve replays to the same bits; the status
reports an exhausted budget and a non-finite right-hand side.
"""
import numpy as np
import pytest
import torch

from tests.cases import femhot_context, make_case

pytestmark = pytest.mark.gpu


def _system(kind="mechanical", etype="hexahedron", n=10, seed=71, context_pattern=True):
    case = make_case(kind, etype, n=n, seed=seed)
    ctx = femhot_context(case)
    f = torch.as_tensor(case["field"], dtype=torch.float64, device="cuda")
    u = torch.as_tensor(case["u"], dtype=torch.float64, device="cuda")
    vals, R = ctx.tangent_csr(f, u)
    rp, ci = ctx.pattern()
    b = (-R).contiguous()
    return ctx, rp, ci, vals, b


@pytest.mark.parametrize("cp", [True, False])
def test_async_matches_host_checked_pcg(cp):
    ctx, rp, ci, vals, b = _system(context_pattern=cp)
    tol = 1e-10
    xs, its, rels, ok = ctx.pcg(rp, ci, vals, b, tol=tol, maxiter=5000, context_pattern=cp)
    assert ok
    xa = torch.zeros_like(b)
    it, rel, st = ctx.pcg_async(rp, ci, vals, b, xa, tol=tol, iterations=its + 20, context_pattern=cp)
    torch.cuda.synchronize()
    ita, rela, sta = int(it.item()), float(rel.item()), int(st.item())
    print(f"[pcg_async] host-checked {its} iterations, device-checked {ita}", flush=True)
    assert sta == 1 and rela <= tol
    assert 0 <= its - ita < 10
    # the host-checked solve capped at the same iteration count runs the same kernels: same bits
    xc, itc, relc, okc = ctx.pcg(rp, ci, vals, b, tol=1e-300, maxiter=ita, raise_on_fail=False, context_pattern=cp)
    torch.cuda.synchronize()
    assert itc == ita and torch.equal(xc, xa)
    assert float(torch.linalg.norm(xa - xs) / torch.linalg.norm(xs)) <= 1e-8
    ctx.close()


def test_async_solve_replays_as_a_graph():
    ctx, rp, ci, vals, b = _system(n=8, seed=72)
    tol, iters = 1e-10, 400
    x_eager = torch.zeros_like(b)
    it, rel, st = ctx.pcg_async(rp, ci, vals, b, x_eager, tol=tol, iterations=iters, context_pattern=True)
    torch.cuda.synchronize()
    assert int(st.item()) == 1
    ref_it = int(it.item())
    xg = torch.empty_like(b)
    out = (torch.empty(1, dtype=torch.int32, device="cuda"), torch.empty(1, dtype=torch.float64, device="cuda"),
           torch.empty(1, dtype=torch.int32, device="cuda"))

    def solve():
        xg.zero_()
        ctx.pcg_async(rp, ci, vals, b, xg, tol=tol, iterations=iters, context_pattern=True, out=out)

    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        solve()
    torch.cuda.current_stream().wait_stream(side)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        solve()
    for rep in range(3):
        xg.fill_(7.0)
        out[2].fill_(-1)
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(xg, x_eager), rep
        assert int(out[0].item()) == ref_it and int(out[2].item()) == 1, rep
        _ = float((b * 2).sum())  # eager work between replays
    ctx.close()


def test_async_status_budget_and_nonfinite():
    ctx, rp, ci, vals, b = _system(n=6, seed=73)
    x = torch.zeros_like(b)
    it, rel, st = ctx.pcg_async(rp, ci, vals, b, x, tol=1e-14, iterations=5, context_pattern=True)
    torch.cuda.synchronize()
    assert int(st.item()) == 3 and int(it.item()) == 5 and float(rel.item()) > 1e-14
    bn = b.clone()
    bn[3] = float("nan")
    x = torch.zeros_like(b)
    it, rel, st = ctx.pcg_async(rp, ci, vals, bn, x, tol=1e-10, iterations=5, context_pattern=True)
    torch.cuda.synchronize()
    assert int(st.item()) == 2 and int(it.item()) == 0
    assert bool(torch.all(x == 0.0))  # a stopped solve leaves the initial guess untouched
    ctx.close()


def test_async_refuses_without_pattern():
    from folax_amd import _lib as L
    case = make_case("thermal", "quad", n=4, seed=74)
    ctx = femhot_context(case)
    b = torch.zeros(ctx.n_dofs, dtype=torch.float64, device="cuda")
    dummy_rp = torch.zeros(ctx.n_dofs + 1, dtype=torch.int64, device="cuda")
    dummy_ci = torch.zeros(1, dtype=torch.int32, device="cuda")
    dummy_v = torch.zeros(1, dtype=torch.float64, device="cuda")
    with pytest.raises(L.FemhotError):
        ctx.pcg_async(dummy_rp, dummy_ci, dummy_v, b, torch.zeros_like(b), iterations=3)
    ctx.close()
