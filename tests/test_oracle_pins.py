"""Closed-form pins of the two oracle laws the reference goldens leave open (CPU only), each with a
planted mistake that must fail:

* the nonlinear conductivity k = (N.k0)(1 + beta (N.T)^c) of ThermalLoss (thermal.py:31-33), through
  the Kirchhoff transform Phi(T) = integral of k(s) ds;
* the compressible Neo-Hookean law (neo_hooke.py:64-109, 2-D :15-58) at finite strain: objectivity,
  pure dilation (no isochoric energy or stress) and isochoric motions (no volumetric energy or
  pressure).
"""
import numpy as np
import pytest
from numpy.polynomial import Polynomial

from oracle import fe


# ---------------------------------------------------------------------------------------------
# nonlinear thermal: Kirchhoff transform
# ---------------------------------------------------------------------------------------------
def _rect(x0, hx, y0, hy):
    return np.array([[[x0, y0, 0.0], [x0 + hx, y0, 0.0], [x0 + hx, y0 + hy, 0.0], [x0, y0 + hy, 0.0]]])


def _exact_rect_residual(k_left, k_right, T_left, T_right, hx, hy, beta, c):
    """Element residual of an axis-aligned QUAD4 whose nodal k0 and T vary along x only.

    Along the element k0(t) and T(t) are linear in t = (x - x0)/hx and grad T = (dT/hx, 0), so
    r_a = s_a hy/(2 hx) dT * integral_0^1 k0(t) (1 + beta T(t)^c) dt with s = (-1, +1, +1, -1)
    (dN_a/dx integrates over y to s_a hy/(2 hx)).  The t-integral is a polynomial for integer c and
    is integrated exactly here, independently of any quadrature."""
    t = Polynomial([0.0, 1.0])
    k0 = k_left + (k_right - k_left) * t
    T = T_left + (T_right - T_left) * t
    integrand = k0 * (1.0 + beta * T ** int(c))
    I = integrand.integ()(1.0) - integrand.integ()(0.0)
    dT = T_right - T_left
    return np.array([-1.0, 1.0, 1.0, -1.0]) * (hy / (2.0 * hx)) * dT * I


@pytest.mark.parametrize("beta,c", [(0.5, 2.0), (0.5, 1.0), (1.3, 2.0)])
def test_thermal_element_residual_matches_exact_integral(beta, c):
    """With k0 and T linear in x the integrand of k is a polynomial of degree c+1 <= 3 in x and
    linear in y, which the 2x2 Gauss rule integrates exactly: the oracle must equal the exact
    integral.  A nodally interpolated k = N.(k0 (1 + beta T^c)), or (N.k0)(1 + beta N.(T^c)),
    gives another integral."""
    rng = np.random.default_rng(310)
    for _ in range(5):
        x0, y0 = rng.uniform(-1, 1, 2)
        hx, hy = rng.uniform(0.2, 2.0, 2)
        kl, kr = rng.uniform(0.1, 1.0, 2)
        Tl, Tr = rng.uniform(0.0, 1.0, 2)
        X = _rect(x0, hx, y0, hy)
        d = np.array([[kl, kr, kr, kl]])
        T = np.array([[Tl, Tr, Tr, Tl]])
        e, r, _ = fe.thermal_element("quad", X, d, T, 2, beta, c, want_K=False)
        ref = _exact_rect_residual(kl, kr, Tl, Tr, hx, hy, beta, c)
        np.testing.assert_allclose(r[0], ref, rtol=1e-13, atol=1e-15)
        np.testing.assert_allclose(e[0], T[0] @ ref, rtol=1e-13, atol=1e-15)


def _kirchhoff_strip(nx, ny, beta, c, k0, seed):
    """QUAD4 rectangle [0,1] x [0,0.4], nx x ny elements with randomly graded column widths; T = 1
    on x = 0, T = 0 on x = 1, insulated top and bottom.  With constant k0 the exact *discrete*
    solution has Phi(T_i) = k0 (T_i + beta T_i^(c+1) / (c+1)) linear in the column position x_i:
    each element contributes s_a hy/(2 hx) (Phi(T_right) - Phi(T_left)) (exact integral above), so
    every interior node's residual is the difference of two equal column fluxes."""
    rng = np.random.default_rng(seed)
    w = rng.uniform(0.5, 1.5, nx)
    xs = np.concatenate([[0.0], np.cumsum(w) / w.sum()])
    xs[-1] = 1.0
    ys = np.linspace(0.0, 0.4, ny + 1)
    Y, Xg = np.meshgrid(ys, xs, indexing="ij")
    coords = np.stack([Xg.ravel(), Y.ravel(), np.zeros(Xg.size)], axis=1)
    i, j = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    b = (i * (nx + 1) + j).ravel()
    conn = np.stack([b, b + 1, b + nx + 2, b + nx + 1], axis=1)
    left = np.where(coords[:, 0] == 0.0)[0]
    right = np.where(coords[:, 0] == 1.0)[0]
    didx = np.concatenate([left, right]).astype(np.int64)
    dval = np.concatenate([np.ones(left.size), np.zeros(right.size)])
    phi = lambda T: T + beta * T ** (c + 1) / (c + 1)
    target = phi(1.0) * (1.0 - coords[:, 0])
    T = np.array(target)  # Newton on the scalar cubic Phi(T) = target, started at the target
    for _ in range(60):
        T = T - (phi(T) - target) / (1.0 + beta * T ** c)
    field = np.full(coords.shape[0], k0)
    return coords, conn, field, T, didx, dval


@pytest.mark.parametrize("beta,c", [(0.5, 2.0), (2.0, 2.0), (0.5, 1.0)])
def test_thermal_strip_kirchhoff_solution_is_exact(beta, c):
    """The Kirchhoff nodal values zero every free residual of the oracle (to rounding), at two mesh
    sizes; the oracle's Picard iteration (assembled Se, BC-masked, direct solves) started from the
    linear profile converges to them."""
    ph = fe.Physics("thermal", "quad", num_gp=2, beta=beta, c=c)
    for nx, ny in ((8, 3), (16, 5)):
        coords, conn, field, T, didx, dval = _kirchhoff_strip(nx, ny, beta, c, 0.7, seed=311)
        _, R = fe.assemble(ph, coords, conn, field, T, didx, want_K=False)
        assert np.abs(R).max() <= 1e-14 * max(1.0, beta), np.abs(R).max()
        Tk = 1.0 - coords[:, 0]
        for _ in range(80):
            Tk = fe.linear_solve(ph, coords, conn, field, Tk, didx, dval)
        np.testing.assert_allclose(Tk, T, rtol=0, atol=1e-12)
    # the nonlinear profile is far from linear: the test sees beta
    assert np.abs(T - (1.0 - coords[:, 0])).max() > 0.02


def _wrong_nodal_conductivity(N, d, T, beta, c):
    dg = np.einsum("ga,ea->eg", N, d)
    Tg = np.einsum("ga,ea->eg", N, T)
    return np.einsum("ga,ea->eg", N, d * (1.0 + beta * T ** c)), dg, Tg


def _wrong_interpolated_power(N, d, T, beta, c):
    dg = np.einsum("ga,ea->eg", N, d)
    Tg = np.einsum("ga,ea->eg", N, T)
    return dg * (1.0 + beta * np.einsum("ga,ea->eg", N, T ** c)), dg, Tg


@pytest.mark.parametrize("wrong", [_wrong_nodal_conductivity, _wrong_interpolated_power])
def test_planted_conductivity_law_mistakes_are_caught(monkeypatch, wrong):
    """Both wrong laws converge to the same continuum solution, so only a pin of the discrete law
    sees them: the exact element integral and the exact strip solution both fail."""
    monkeypatch.setattr(fe, "conductivity_at_gp", wrong)
    X = _rect(0.0, 1.0, 0.0, 1.0)
    r = fe.thermal_element("quad", X, np.array([[0.3, 0.9, 0.9, 0.3]]), np.array([[0.2, 0.8, 0.8, 0.2]]), 2,
                           0.5, 2.0, want_K=False)[1]
    ref = _exact_rect_residual(0.3, 0.9, 0.2, 0.8, 1.0, 1.0, 0.5, 2.0)
    assert np.abs(r[0] - ref).max() > 1e-4 * np.abs(ref).max()
    ph = fe.Physics("thermal", "quad", num_gp=2, beta=0.5, c=2.0)
    coords, conn, field, T, didx, _ = _kirchhoff_strip(8, 3, 0.5, 2.0, 0.7, seed=311)
    _, R = fe.assemble(ph, coords, conn, field, T, didx, want_K=False)
    assert np.abs(R).max() > 1e-5


# ---------------------------------------------------------------------------------------------
# Neo-Hookean at finite strain
# ---------------------------------------------------------------------------------------------
def _rotation(rng, dim):
    Q, Rr = np.linalg.qr(rng.standard_normal((dim, dim)))
    Q = Q * np.sign(np.diag(Rr))
    if np.linalg.det(Q) < 0:
        Q[:, 0] = -Q[:, 0]
    return Q


def _finite_F(rng, dim, n):
    F = np.eye(dim) + 0.35 * rng.standard_normal((n, dim, dim))
    F[np.linalg.det(F) <= 0.2] = np.eye(dim)
    return F


@pytest.mark.parametrize("dim", [2, 3])
def test_neo_hooke_objectivity(dim):
    """psi(QF) = psi(F) and S(QF) = S(F) (S is the second Piola-Kirchhoff stress, a function of
    C = F^T F), and so is the material tangent, at strains of order 0.35."""
    rng = np.random.default_rng(320 + dim)
    F = _finite_F(rng, dim, 40)
    kap, mu = rng.uniform(0.5, 3.0, 40), rng.uniform(0.2, 2.0, 40)
    Q = np.stack([_rotation(rng, dim) for _ in range(40)])
    p0, S0, C0 = fe.nh_material(F, kap, mu)
    p1, S1, C1 = fe.nh_material(np.einsum("eij,ejk->eik", Q, F), kap, mu)
    np.testing.assert_allclose(p1, p0, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(S1, S0, rtol=1e-11, atol=1e-12)
    np.testing.assert_allclose(C1, C0, rtol=1e-10, atol=1e-11)
    assert np.abs(S0).max() > 0.1  # finite strain, finite stress


@pytest.mark.parametrize("dim", [2, 3])
def test_neo_hooke_pure_dilation(dim):
    """F = s I: J = s^d, J^(-2/d) tr C = d, so psi_iso = 0 and S_iso = 0; what remains is
    psi = kappa/4 (J^2 - 2 ln J - 1) and S = J p C^-1 = J p s^-2 I with p = kappa/2 (J - 1/J)."""
    s = np.array([0.6, 0.9, 0.999, 1.0, 1.001, 1.2, 1.7])
    kap, mu = 2.3, 0.8
    F = s[:, None, None] * np.eye(dim)
    psi, Sv, _ = fe.nh_material(F, np.full(s.size, kap), np.full(s.size, mu))
    J = s ** dim
    p = 0.5 * kap * (J - 1.0 / J)
    np.testing.assert_allclose(psi, kap / 4 * (J ** 2 - 2 * np.log(J) - 1), rtol=1e-13, atol=1e-16)
    nv = 3 if dim == 2 else 6
    Sref = np.zeros((s.size, nv))
    Sref[:, :dim] = (J * p / s ** 2)[:, None]
    np.testing.assert_allclose(Sv, Sref, rtol=1e-13, atol=1e-15)


@pytest.mark.parametrize("dim", [2, 3])
def test_neo_hooke_isochoric_motion(dim):
    """det F = 1 (simple shear and volume-preserving stretch): psi_vol = 0 and p = 0, so
    psi = mu/2 (tr C - d) and S = mu (I - tr C / d C^-1)."""
    g = 0.7
    Fs = [np.eye(dim), np.eye(dim)]
    Fs[0][0, 1] = g  # simple shear
    a = 1.3
    Fs[1] = np.diag([a, 1.0 / a] if dim == 2 else [a, 1.1, 1.0 / (1.1 * a)])
    F = np.stack(Fs)
    kap, mu = 4.0, 1.1
    psi, Sv, _ = fe.nh_material(F, np.full(2, kap), np.full(2, mu))
    C = np.einsum("eki,ekj->eij", F, F)
    trC = np.trace(C, axis1=1, axis2=2)
    np.testing.assert_allclose(psi, mu / 2 * (trC - dim), rtol=1e-13)
    Sm = mu * (np.eye(dim) - (trC / dim)[:, None, None] * np.linalg.inv(C))
    pairs = [(0, 0), (1, 1), (0, 1)] if dim == 2 else [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]
    np.testing.assert_allclose(Sv, np.stack([Sm[:, i, j] for i, j in pairs], axis=1), rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(psi[0], mu / 2 * g * g, rtol=1e-13)


@pytest.mark.parametrize("etype", ["quad", "hexahedron", "tetra"])
def test_neo_hooke_element_objectivity(etype):
    """Element level: a rigid rotation of the deformed configuration, x' = Q (X + u), leaves the
    element energy unchanged and rotates every nodal force, f'_a = Q f_a."""
    from tests.cases import mesh_for
    coords, conn, _ = mesh_for(etype, 2, seed=330)
    dim = fe.ELEM_DIM[etype]
    rng = np.random.default_rng(331)
    Xe = coords[conn][..., :dim]
    ue = 0.06 * rng.standard_normal(Xe.shape)  # strains of order 0.2 on h = 0.5, J > 0
    Q = _rotation(rng, dim)
    ue2 = np.einsum("ij,eaj->eai", Q, Xe + ue) - Xe
    d = rng.uniform(0.5, 2.0, conn.shape)
    X = coords[conn]
    E0, r0, _ = fe.nh_element(etype, X, d, ue.reshape(conn.shape[0], -1), 2 if etype != "tetra" else 1,
                              want_K=False)
    E1, r1, _ = fe.nh_element(etype, X, d, ue2.reshape(conn.shape[0], -1), 2 if etype != "tetra" else 1,
                              want_K=False)
    np.testing.assert_allclose(E1, E0, rtol=1e-11)
    rot = np.einsum("ij,eaj->eai", Q, r0.reshape(conn.shape[0], -1, dim)).reshape(conn.shape[0], -1)
    np.testing.assert_allclose(r1, rot, rtol=1e-10, atol=1e-12 * np.abs(r0).max())


def test_planted_isochoric_exponent_mistake_is_caught(monkeypatch):
    """J^(-1/3) in place of J^(-2/3) leaves F = I (the goldens) untouched but puts isochoric energy
    into a pure dilation."""
    monkeypatch.setattr(fe, "ISO_J_POWER", -1.0)
    F = np.eye(3)[None]
    psi, Sv, _ = fe.nh_material(F, np.ones(1), np.ones(1))
    assert abs(psi[0]) < 1e-15 and np.abs(Sv).max() < 1e-15  # invisible at F = I
    with pytest.raises(AssertionError):
        test_neo_hooke_pure_dilation(3)
