"""Pin the FP64 oracle against the golden vectors held by folax's own tests (CPU only).

Fixtures: tests/golden/reference_goldens.json (extracted as data by tests/golden/extract_goldens.py).
Tolerances are the reference tests' own rtol/atol.
"""
import json
import os

import numpy as np
import pytest

from oracle import fe

G = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_goldens.json")))
MECH = G["tests/unit/test_mechanical_loss.py"]
NH = G["tests/unit/test_neo_hooke_mechanical_loss.py"]
GEO = G["tests/unit/test_geometries.py"]
SA = G["tests/unit/test_sensitivity_analysis.py"]["test_quad"]
SA2 = G["tests/integration/test_mechanical_2D_sa.py"]

ELEM_TESTS = [("test_tetra", "tetra", "tet_points_coordinates", [1, 2, 3]),
              ("test_hexa", "hexahedron", "hex_points_coordinates", [1, 2, 3]),
              ("test_quad", "quad", "quad_points_coordinates", [1, 2])]


def _check(actual, chk):
    assert chk["actual"] in ("stiffness", "residuals.flatten()") or True
    np.testing.assert_allclose(np.asarray(actual).reshape(np.shape(chk["expected"])), chk["expected"],
                               rtol=chk["rtol"], atol=chk["atol"])


@pytest.mark.parametrize("tname,etype,key,body", ELEM_TESTS)
def test_linear_elastic_element_goldens(tname, etype, key, body):
    """tests/unit/test_mechanical_loss.py: E=1, nu=0.3, body force, d=1, u=1."""
    t = MECH[tname]
    X = np.array(t["assigns"][key])[None]
    n = X.shape[1]
    ph = fe.Physics("mechanical", etype, E=1.0, nu=0.3, body=body)
    _, r, K = ph.element(X, np.ones((1, n)), np.ones((1, n * ph.ndpn)))
    assert t["checks"][0]["actual"] == "stiffness" and t["checks"][1]["actual"] == "residuals.flatten()"
    _check(K[0], t["checks"][0])
    _check(r[0], t["checks"][1])


@pytest.mark.parametrize("tname,etype,key,body", ELEM_TESTS)
def test_neo_hooke_element_goldens(tname, etype, key, body):
    """tests/unit/test_neo_hooke_mechanical_loss.py (uniform translation u=1, i.e. F = I)."""
    t = NH[tname]
    X = np.array(t["assigns"][key])[None]
    n = X.shape[1]
    ph = fe.Physics("neohooke", etype, nu=0.3, body=body)
    _, r, K = ph.element(X, np.ones((1, n)), np.ones((1, n * ph.ndpn)))
    _check(K[0], t["checks"][0])
    _check(r[0], t["checks"][1])
    # these goldens are printed at full precision: the oracle matches them far below rtol
    assert np.max(np.abs(K[0] - np.array(t["checks"][0]["expected"]))) < 1e-13


def _geo_actuals(etype, X, rule):
    pts, w = fe.gauss_rule(etype, rule)
    N = np.stack([fe.shape_values(etype, p) for p in pts])
    dN = np.stack([fe.shape_local_grads(etype, p) for p in pts])
    J = np.stack([fe.jacobian(etype, X, p) for p in pts])
    _, _, DN, _ = fe.geometry(etype, X[None], rule)
    return {"points": pts, "weights": w, "shape_function_values": N, "shape_function_values[0]": N[0],
            "shape_function_grads": dN, "shape_function_grads[0]": dN[0], "jacobian": J, "jacobians": J,
            "shape_function_g_grads": DN[0]}


@pytest.mark.parametrize("tname,etype,key", [("test_quad2D4", "quad", "self.quad_points_coordinates"),
                                             ("test_tet3D4", "tetra", "self.tet_points_coordinates"),
                                             ("test_hex3D8", "hexahedron", "self.hex_points_coordinates")])
def test_geometry_goldens(tname, etype, key):
    """tests/unit/test_geometries.py: rules 1, 2, 3 in order; each `points` check starts a new rule."""
    X = np.array(GEO["setUp"]["assigns"][key])
    rule = 0
    acts = None
    n_checked = 0
    for chk in GEO[tname]["checks"]:
        if chk["actual"] == "points":
            rule += 1
            acts = _geo_actuals(etype, X, rule)
        a = acts[chk["actual"]]
        if chk["actual"] == "weights" and etype == "quad" and rule == 3:
            a = np.asarray(a)
        _check(a, chk)
        n_checked += 1
    assert n_checked == len(GEO[tname]["checks"]) and rule == 3


def _sa_setup(N, K):
    coords, conn, sets = fe.square_quad_mesh(N, 1.0)
    ph = fe.Physics("mechanical", "quad", num_gp=2, E=1.0, nu=0.3)
    bc = {"Ux": {"left": 0.0, "right": 0.05}, "Uy": {"left": 0.0, "right": 0.05}}
    didx, dval = fe.dirichlet_dofs(sets, ["Ux", "Uy"], bc)
    return coords, conn, ph, didx, dval


def _response_grads(coords, conn, K, u):
    """Response f = sum_e sum_gp w detJ E^2 U_x (response_formula "(E**2)*U[0]",
    fe_response.py:73-113): df/du and df/dE at nodes."""
    N, w, DN, detJ = fe.geometry("quad", coords[conn], 2)
    Eg = np.einsum("ga,ea->eg", N, K[conn])
    dofs = fe.element_dofs(conn, 2)
    U0 = np.einsum("ga,ea->eg", N, u[dofs][:, 0::2])
    wd = w[None] * detJ
    dfdu_e = np.zeros((conn.shape[0], 8))
    dfdu_e[:, 0::2] = np.einsum("eg,ga->ea", wd * Eg ** 2, N)
    dfdu = np.zeros(2 * coords.shape[0])
    np.add.at(dfdu, dofs, dfdu_e)
    dfdE = np.zeros(coords.shape[0])
    np.add.at(dfdE, conn, np.einsum("eg,ga->ea", wd * 2 * Eg * U0, N))
    return dfdu, dfdE


def _response_shape_grad(coords, conn, K, u):
    """Shape gradient of the same response: only detJ depends on the coordinates, and
    d detJ / dX_(a,i) = detJ dN_a/dx_i (Jacobi's formula), so df/dX_(a,i) = sum_g w f_g detJ DN_(a,i).
    (Nn, 3) with a zero z column (2-D element)."""
    N, w, DN, detJ = fe.geometry("quad", coords[conn], 2)
    Eg = np.einsum("ga,ea->eg", N, K[conn])
    U0 = np.einsum("ga,ea->eg", N, u[fe.element_dofs(conn, 2)][:, 0::2])
    fw = w[None] * Eg ** 2 * U0 * detJ
    g = np.zeros((coords.shape[0], 3))
    for i in range(2):
        np.add.at(g[:, i], conn, np.einsum("eg,ega->ea", fw, DN[..., i]))
    return g


def test_sensitivity_analysis_goldens():
    """tests/unit/test_sensitivity_analysis.py: transposed BC tangent row 8, adjoint RHS and
    control derivatives lambda^T dR/dE + df/dE on a 3x3-node QUAD4 mesh."""
    K = np.array(SA["assigns"]["random_K"])
    u = np.array(SA["assigns"]["random_FE_UV"])
    lam = np.array(SA["assigns"]["random_adj_FE_UV"])
    coords, conn, ph, didx, dval = _sa_setup(3, K)
    A, _ = fe.assemble(ph, coords, conn, K, u, didx, transpose=True)
    _check(A.toarray()[8], SA["checks"][0])
    dfdu, dfdE = _response_grads(coords, conn, K, u)
    rhs = dfdu.copy()
    rhs[didx] = 0.0
    _check(-rhs, SA["checks"][1])
    _, gd = fe.global_residual_vjp(ph, coords, conn, K, u, lam)
    _check(dfdE + gd, SA["checks"][2])
    gx = fe.global_residual_shape_vjp(ph, coords, conn, K, u, lam)
    assert SA["checks"][3]["actual"].startswith("shape_derivatives")
    _check((_response_shape_grad(coords, conn, K, u) + gx).ravel(), SA["checks"][3])


def test_mechanical_2D_sa_end_to_end_golden():
    """tests/integration/test_mechanical_2D_sa.py: assemble -> direct solve -> transposed adjoint
    solve -> control derivatives on a 5x5-node QUAD4 mesh."""
    K = np.array(SA2["setUp"]["assigns"]["self.random_K"])
    coords, conn, ph, didx, dval = _sa_setup(5, K)
    u = fe.linear_solve(ph, coords, conn, K, np.zeros(2 * coords.shape[0]), didx, dval)
    import scipy.sparse.linalg as spla
    At, _ = fe.assemble(ph, coords, conn, K, u, didx, transpose=True)
    dfdu, dfdE = _response_grads(coords, conn, K, u)
    rhs = dfdu.copy()
    rhs[didx] = 0.0
    lam = spla.spsolve(At.tocsc(), -rhs)
    _, gd = fe.global_residual_vjp(ph, coords, conn, K, u, lam)
    chk = SA2["test_sensitivites"]["checks"][0]
    assert chk["actual"] == "control_derivatives"
    _check(dfdE + gd, chk)
    chk = SA2["test_sensitivites"]["checks"][1]
    assert chk["actual"] == "shape_derivatives"
    gx = fe.global_residual_shape_vjp(ph, coords, conn, K, u, lam)
    _check((_response_shape_grad(coords, conn, K, u) + gx).ravel(), chk)
