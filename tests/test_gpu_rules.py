"""GPU parity of the other Gauss rules and of body forces, for every element routine.

The parity suite runs each element at its default rule (QUAD4 / HEX8 2 points per direction, TET4 one
point: thermal.py:61-62,69-70, mechanical.py:127-128,141-142, fe_loss.py:81-84); the library takes rules
1-3 of quadrilateral_2d_4.py:17-66, hexahedra_3d_8.py:17-34 and tetrahedra_3d_4.py:17-30. Here the
residual and energy, the residual VJP, the potential gradient and the CSR tangent run at the rules the
parity suite does not reach, and every routine runs with a body force (the Fe terms of thermal.py:39,44,
mechanical.py:108,113); FP64 within 1e-12 of the FP64 oracle and FP32 within 1e-5, on the same distorted
meshes as the parity suite.
"""
import numpy as np
import pytest

from oracle import fe
from tests.cases import femhot_context, make_case
from tests.test_gpu_parity import ALL, TOL, dev, rel, thermal_params

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

RULES = [("quad", 1), ("quad", 3), ("hexahedron", 1), ("hexahedron", 3), ("tetra", 2), ("tetra", 3)]
CASES = [(k, e, g) for k in ("thermal", "mechanical", "neohooke") for (e, g) in RULES]


def _residual(case, dt):
    ph = case["ph"]
    X = case["coords"][case["conn"]]
    dofs = fe.element_dofs(case["conn"], ph.ndpn)
    e_el, r_el, _ = ph.element(X, case["field"][case["conn"]], case["u"][dofs], want_K=False)
    R_ref = fe.scatter_nodal(case["conn"], r_el, case["coords"].shape[0], ph.ndpn)
    ctx = femhot_context(case)
    R, en = ctx.residual(dev(case["field"], dt)[None], dev(case["u"], dt)[None])
    torch.cuda.synchronize()
    assert rel(R.cpu().numpy(), R_ref) <= TOL[dt]
    assert rel(en.cpu().numpy().sum(), e_el.sum()) <= TOL[dt]
    return ctx


def _energy_grad(case, ctx, dt):
    Pi_ref, R_ref, gd_ref = fe.global_energy_grad(case["ph"], case["coords"], case["conn"], case["field"], case["u"])
    en, gu, gd = ctx.energy_grad(dev(case["field"], dt)[None], dev(case["u"], dt)[None])
    torch.cuda.synchronize()
    assert rel(gu.cpu().numpy(), R_ref) <= TOL[dt]
    assert rel(gd.cpu().numpy(), gd_ref) <= TOL[dt]
    assert rel(en.cpu().numpy().sum(), Pi_ref) <= TOL[dt]


@pytest.mark.parametrize("dt", ["f64", "f32"])
@pytest.mark.parametrize("kind,etype,num_gp", CASES)
def test_rule_residual_vjp_tangent(kind, etype, num_gp, dt):
    case = make_case(kind, etype, n=3, seed=11, num_gp=num_gp, **thermal_params(kind))
    print(f"[rule] {kind} {etype} num_gp={num_gp} {dt}", flush=True)
    ph = case["ph"]
    ctx = _residual(case, dt)
    for picard in ([False, True] if kind == "thermal" else [False]):
        gu_ref, gd_ref = fe.global_residual_vjp(ph, case["coords"], case["conn"], case["field"], case["u"], case["w"],
                                                picard=picard)
        gu, gd = ctx.residual_vjp(dev(case["field"], dt)[None], dev(case["u"], dt)[None], dev(case["w"], dt)[None],
                                  flags=0x4 if picard else 0)
        torch.cuda.synchronize()
        assert rel(gu.cpu().numpy(), gu_ref) <= TOL[dt], picard
        assert rel(gd.cpu().numpy(), gd_ref) <= TOL[dt], picard
    A, R_ref = fe.assemble(ph, case["coords"], case["conn"], case["field"], case["u"], case["didx"])
    for flags in (0, 0x20, 0x80):  # atomic, colour-ordered, row-owner (colour-ordered for Neo-Hookean)
        vals, R = ctx.tangent_csr(dev(case["field"], dt), dev(case["u"], dt), flags=flags)
        torch.cuda.synchronize()
        assert rel(vals.cpu().numpy(), A.data) <= TOL[dt], flags
        assert rel(R.cpu().numpy(), R_ref) <= TOL[dt], flags
    ctx.close()


@pytest.mark.parametrize("dt", ["f64", "f32"])
@pytest.mark.parametrize("kind,etype,num_gp", CASES)
def test_rule_energy_grad(kind, etype, num_gp, dt):
    case = make_case(kind, etype, n=3, seed=12, num_gp=num_gp)
    print(f"[rule egrad] {kind} {etype} num_gp={num_gp} {dt}", flush=True)
    ctx = femhot_context(case)
    _energy_grad(case, ctx, dt)
    ctx.close()


@pytest.mark.parametrize("dt", ["f64", "f32"])
@pytest.mark.parametrize("kind,etype", ALL)
def test_body_force(kind, etype, dt):
    """Residual, energy, VJP (the body term drops out of it) and potential gradient with a body force."""
    dim = 2 if etype == "quad" else 3
    body = 0.7 if kind == "thermal" else (0.3, -0.2, 0.5)[:dim]
    case = make_case(kind, etype, n=3, seed=13, body=body)
    print(f"[body] {kind} {etype} {dt}", flush=True)
    ctx = _residual(case, dt)
    gu_ref, gd_ref = fe.global_residual_vjp(case["ph"], case["coords"], case["conn"], case["field"], case["u"],
                                            case["w"])
    gu, gd = ctx.residual_vjp(dev(case["field"], dt)[None], dev(case["u"], dt)[None], dev(case["w"], dt)[None])
    torch.cuda.synchronize()
    assert rel(gu.cpu().numpy(), gu_ref) <= TOL[dt]
    assert rel(gd.cpu().numpy(), gd_ref) <= TOL[dt]
    _energy_grad(case, ctx, dt)
    ctx.close()
