"""FP64 NumPy restatement of folax's finite-element residual hot path.

TEST INFRASTRUCTURE ONLY -- see oracle/__init__.py.  The product package
never imports this module.

Everything is vectorised over elements: coordinates `X` are (Ne, n, 3), nodal
fields `d` are gathered per element as (Ne, n) (or (Ne,) for a per-element
field), and nodal solutions `u` are gathered as (Ne, n*ndpn) in node-major
interleaved order (local dof = ndpn*a + comp), exactly as
`FiniteElementLoss.ComputeElementEnergyVmapCompatible` gathers them
(fol/loss_functions/fe_loss.py:155-164).

Parity status of each function is recorded next to it ("pinned by ..."); the
checks live in tests/test_oracle_*.py.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

# --------------------------------------------------------------------------
# L1: geometry -- Gauss rules, shape functions, Jacobians
# --------------------------------------------------------------------------

_S3 = 1.0 / np.sqrt(3.0)
_S35 = np.sqrt(3.0 / 5.0)

ELEM_NODES = {"quad": 4, "hexahedron": 8, "tetra": 4}
ELEM_DIM = {"quad": 2, "hexahedron": 3, "tetra": 3}


def gauss_rule(etype: str, num_gp: int):
    """Integration points (ngp, 3) and weights (ngp,).

    QUAD4: fol/geometries/quadrilateral_2d_4.py:17-41; HEX8:
    fol/geometries/hexahedra_3d_8.py:17-70; TET4:
    fol/geometries/tetrahedra_3d_4.py:17-43 (the truncated 4-point constants
    0.58541020/0.13819660 are kept as the reference writes them).
    Pinned by tests/unit/test_geometries.py goldens + polynomial exactness.
    """
    if etype == "quad":
        if num_gp == 1:
            return np.array([[0.0, 0.0, 0.0]]), np.array([4.0])
        if num_gp == 2:
            p = np.array([[-_S3, -_S3, 0.0], [_S3, -_S3, 0.0], [_S3, _S3, 0.0], [-_S3, _S3, 0.0]])
            return p, np.ones(4)
        if num_gp == 3:
            a = _S35
            xs = [-a, 0.0, a]
            w1 = [5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0]
            p = np.array([[xs[i], xs[j], 0.0] for j in range(3) for i in range(3)])
            w = np.array([w1[i] * w1[j] for j in range(3) for i in range(3)])
            return p, w
    elif etype == "hexahedron":
        if num_gp == 1:
            return np.array([[0.0, 0.0, 0.0]]), np.array([8.0])
        if num_gp == 2:
            q = [(-1, -1), (1, -1), (1, 1), (-1, 1)]
            p = np.array([[a * _S3, b * _S3, c * _S3] for c in (-1, 1) for (a, b) in q])
            return p, np.ones(8)
        if num_gp == 3:
            a = _S35
            xs = [-a, 0.0, a]
            w1 = [5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0]
            p = np.array([[xs[i], xs[j], xs[k]] for k in range(3) for j in range(3) for i in range(3)])
            w = np.array([w1[i] * w1[j] * w1[k] for k in range(3) for j in range(3) for i in range(3)])
            return p, w
    elif etype == "tetra":
        if num_gp == 1:
            return np.array([[0.25, 0.25, 0.25]]), np.array([1.0 / 6.0])
        if num_gp == 2:
            a, b = 0.58541020, 0.13819660
            p = np.array([[a, b, b], [b, a, b], [b, b, a], [b, b, b]])
            return p, np.full(4, 1.0 / 24.0)
        if num_gp == 3:
            a, b = 0.015835909865720057993, 0.32805469671142664734
            c, d = 0.67914317820120795168, 0.10695227393293068277
            p = np.array([[a, b, b], [b, a, b], [b, b, a], [b, b, b],
                          [c, d, d], [d, c, d], [d, d, c], [d, d, d]])
            w = np.array([0.02308799441864369039] * 4 + [0.01857867224802297628] * 4)
            return p, w
    raise ValueError(f"unsupported element/rule {etype}/{num_gp}")


def shape_values(etype: str, xi):
    """N(xi) -- quadrilateral_2d_4.py:49-54, hexahedra_3d_8.py:79-88, tetrahedra_3d_4.py:52-56."""
    x, y, z = xi[0], xi[1], xi[2]
    if etype == "quad":
        return 0.25 * np.array([(1 - x) * (1 - y), (1 + x) * (1 - y), (1 + x) * (1 + y), (1 - x) * (1 + y)])
    if etype == "hexahedron":
        s = [(-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1), (-1, -1, 1), (1, -1, 1), (1, 1, 1), (-1, 1, 1)]
        return 0.125 * np.array([(1 + a * x) * (1 + b * y) * (1 + c * z) for a, b, c in s])
    if etype == "tetra":
        return np.array([1.0 - (x + y + z), x, y, z])
    raise ValueError(etype)


def shape_local_grads(etype: str, xi):
    """dN/dxi (n, dim) -- quadrilateral_2d_4.py:56-61, hexahedra_3d_8.py:90-114, tetrahedra_3d_4.py:58-60."""
    x, y, z = xi[0], xi[1], xi[2]
    if etype == "quad":
        s = [(-1, -1), (1, -1), (1, 1), (-1, 1)]
        return 0.25 * np.array([[a * (1 + b * y), b * (1 + a * x)] for a, b in s])
    if etype == "hexahedron":
        s = [(-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1), (-1, -1, 1), (1, -1, 1), (1, 1, 1), (-1, 1, 1)]
        return 0.125 * np.array([[a * (1 + b * y) * (1 + c * z), b * (1 + a * x) * (1 + c * z),
                                  c * (1 + a * x) * (1 + b * y)] for a, b, c in s])
    if etype == "tetra":
        return np.array([[-1.0, -1.0, -1.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])
    raise ValueError(etype)


def jacobian(etype: str, X, xi):
    """J[i,k] = sum_a X[a,i] dN_a/dxi_k  (geometry.py:94-97; QUAD4 uses x,y only, quadrilateral_2d_4.py:63-66)."""
    dim = ELEM_DIM[etype]
    return np.einsum("...ai,ak->...ik", np.asarray(X)[..., :dim], shape_local_grads(etype, xi))


def geometry(etype: str, X, num_gp: int):
    """Per element and Gauss point: N (g,n), weights (g,), global gradients DN (e,g,n,dim), signed detJ (e,g).

    DN = dN/dxi . inv(J) (geometry.py:87-92); detJ = jnp.linalg.det(J), no abs
    (thermal.py:37, mechanical.py:107).
    """
    pts, wts = gauss_rule(etype, num_gp)
    dim = ELEM_DIM[etype]
    X = np.asarray(X)
    Xd = (X if np.iscomplexobj(X) else X.astype(np.float64))[..., :dim]  # complex kept for complex-step
    N = np.stack([shape_values(etype, p) for p in pts])
    dN = np.stack([shape_local_grads(etype, p) for p in pts])
    J = np.einsum("eai,gak->egik", Xd, dN)
    detJ = np.linalg.det(J)
    invJ = np.linalg.inv(J)
    DN = np.einsum("gak,egkj->egaj", dN, invJ)
    return N, wts, DN, detJ


def _field_at_gp(N, d):
    d = np.asarray(d, dtype=np.float64)
    if d.ndim == 1:  # per-element field
        return np.repeat(d[:, None], N.shape[0], axis=1)
    return np.einsum("ga,ea->eg", N, d)


def _field_weights(N, d_ndim, Ne):
    """d(field at gp)/d(field dofs): N for nodal fields, 1 for per-element fields."""
    return N if d_ndim == 2 else np.ones((N.shape[0], 1))


# --------------------------------------------------------------------------
# L2/L3: element physics
# --------------------------------------------------------------------------

def _body_load(N, w, detJ, body, ndpn):
    """Fe[a*ndpn+c] = sum_g w detJ N_a b_c (thermal.py:39 with scalar b; mechanical.py:109 N_mat^T b)."""
    body = np.atleast_1d(np.asarray(body, dtype=np.float64)).reshape(-1)
    if body.size == 1 and ndpn > 1:
        body = np.repeat(body, ndpn)
    Ne = detJ.shape[0]
    fe = np.einsum("g,eg,ga->ea", w, detJ, N)
    return (fe[:, :, None] * body[None, None, :ndpn]).reshape(Ne, -1)


def conductivity_at_gp(N, d, T, beta, c):
    """k at each Gauss point (thermal.py:31-33): the field and the temperature are interpolated
    first, k = (N.d) (1 + beta (N.T)^c).  Returns (k, N.d, N.T), each (e, g).
    Pinned by the Kirchhoff closed forms (tests/test_oracle_pins.py): exact element residuals for
    fields linear along x, and the discrete strip solution Phi(T_i) linear in x_i."""
    dg = _field_at_gp(N, d)
    Tg = np.einsum("ga,ea->eg", N, T)
    return dg * (1.0 + beta * Tg ** c), dg, Tg


def thermal_element(etype, X, d, T, num_gp, beta=0.0, c=1.0, body=0.0, want_K=True):
    """ThermalLoss.ComputeElement (fol/loss_functions/thermal.py:28-49).

    k_gp = (N.d)(1 + beta (N.T)^c) with T under stop_gradient, Se = sum k DN DN^T detJ w,
    Fe = sum w detJ b N.  Returns (energy = T.(Se T - Fe), r = Se T - Fe, Se).
    Pinned by: unit-square closed form, patch test, R = d/dT(1/2 T^T Se T) at beta=0, and the
    Kirchhoff closed forms of beta != 0 (exact element residuals, exact strip solution).
    """
    N, w, DN, detJ = geometry(etype, X, num_gp)
    T = np.asarray(T, dtype=np.float64)
    k, _, _ = conductivity_at_gp(N, d, T, beta, c)
    coef = k * w[None, :] * detJ
    Se = np.einsum("eg,egai,egbi->eab", coef, DN, DN)
    Fe = _body_load(N, w, detJ, body, 1)
    r = np.einsum("eab,eb->ea", Se, T) - Fe
    energy = np.einsum("ea,ea->e", T, r)
    return energy, r, (Se if want_K else None)


def d_matrix(dim, E, nu):
    """Isotropic D: 2-D plane stress (mechanical.py:60-61), 3-D (mechanical.py:63-82)."""
    if dim == 2:
        return np.array([[1, nu, 0], [nu, 1, 0], [0, 0, (1 - nu) / 2]], dtype=np.float64) * (E / (1 - nu ** 2))
    c1 = E / ((1.0 + nu) * (1.0 - 2.0 * nu))
    c2, c3, c4 = c1 * (1.0 - nu), c1 * nu, c1 * 0.5 * (1.0 - 2.0 * nu)
    D = np.zeros((6, 6))
    D[:3, :3] = c3
    D[0, 0] = D[1, 1] = D[2, 2] = c2
    D[3, 3] = D[4, 4] = D[5, 5] = c4
    return D


def b_matrix(DN):
    """Small-strain B (..., nv, n*dim).  2-D rows [xx,yy,xy] (mechanical.py:37-44);
    3-D rows [xx,yy,zz,xy,yz,xz] (mechanical.py:46-58)."""
    n, dim = DN.shape[-2], DN.shape[-1]
    lead = DN.shape[:-2]
    if dim == 2:
        B = np.zeros(lead + (3, 2 * n), dtype=DN.dtype)
        B[..., 0, 0::2] = DN[..., 0]
        B[..., 1, 1::2] = DN[..., 1]
        B[..., 2, 0::2] = DN[..., 1]
        B[..., 2, 1::2] = DN[..., 0]
        return B
    B = np.zeros(lead + (6, 3 * n), dtype=DN.dtype)
    B[..., 0, 0::3] = DN[..., 0]
    B[..., 1, 1::3] = DN[..., 1]
    B[..., 2, 2::3] = DN[..., 2]
    B[..., 3, 0::3] = DN[..., 1]
    B[..., 3, 1::3] = DN[..., 0]
    B[..., 4, 1::3] = DN[..., 2]
    B[..., 4, 2::3] = DN[..., 1]
    B[..., 5, 0::3] = DN[..., 2]
    B[..., 5, 2::3] = DN[..., 0]
    return B


def mech_element(etype, X, d, u, num_gp, E=1.0, nu=0.3, body=None, want_K=True):
    """MechanicalLoss.ComputeElement (fol/loss_functions/mechanical.py:98-117).

    Se = sum w detJ (N.d) B^T D B, Fe = sum w detJ N_mat^T b, returns
    (u.(Se u - Fe), Se u - Fe, Se).  Pinned by tests/unit/test_mechanical_loss.py goldens,
    rigid-body null space and patch tests.
    """
    dim = ELEM_DIM[etype]
    N, w, DN, detJ = geometry(etype, X, num_gp)
    u = np.asarray(u, dtype=np.float64)
    D = d_matrix(dim, E, nu)
    B = b_matrix(DN)
    eg = _field_at_gp(N, d)
    coef = w[None, :] * detJ * eg
    Ke = np.einsum("eg,egvi,vw,egwj->eij", coef, B, D, B)
    Fe = _body_load(N, w, detJ, np.zeros(dim) if body is None else body, dim)
    r = np.einsum("eij,ej->ei", Ke, u) - Fe
    energy = np.einsum("ei,ei->e", u, r)
    return energy, r, (Ke if want_K else None)


_VOIGT = {3: [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)], 2: [(0, 0), (1, 1), (0, 1)]}

# The isochoric part of the Neo-Hookean energy scales trC by J^(ISO_J_POWER / dim)
# (neo_hooke.py:77-80 for 3-D, :29-32 for 2-D).  A module constant so the pins of
# tests/test_oracle_pins.py can plant a wrong exponent and see it caught.
ISO_J_POWER = -2.0


def nh_material(F, kappa, mu, want_C=True):
    """NeoHookianModel.evaluate (3-D, fol/constitutive_material_models/neo_hooke.py:64-109) and
    NeoHookianModel2D.evaluate (2-D, neo_hooke.py:15-58), vectorised over leading axes.

    psi = k/4 (J^2 - 2 lnJ - 1) + mu/2 (J^(-2/d) trC - d);
    S = J p C^-1 + J^(-2/d) mu (I - trC/d C^-1);  p = k/2 (J - 1/J);
    C = (Jp + dp/dJ J^2) C^-1 (x) C^-1 - 2Jp DS(C^-1,C^-1)
        + (2/d) J^(-2/d) mu trC (DS(C^-1,C^-1) - 1/d C^-1 (x) C^-1) - (2/d)(C^-1 (x) S_iso + S_iso (x) C^-1)
    with DS(A,B)_ijkl = (A_ik B_jl + A_il B_jk)/2 (utils.py:191-196) and the Voigt pairs of
    utils.py:103-130.  The P:Cbar:P^T term is identically zero in the reference (C_ = 0).
    Pinned by: equality with linear elasticity at F=I (3-D), FD of psi -> S and of S -> C.
    """
    # psi and S are evaluated in extended precision (x87 long double): the reference formula
    # cancels to first order at small strain, and the oracle must stay accurate there.
    F = np.asarray(F, dtype=np.float64)
    dim = F.shape[-1]
    LD = np.longdouble
    FL = F.astype(LD)
    kL = np.asarray(kappa, dtype=np.float64).astype(LD)
    mL = np.asarray(mu, dtype=np.float64).astype(LD)
    CL = np.einsum("...ki,...kj->...ij", FL, FL)
    JL = _det_ld(FL)
    invCL = _adj_ld(CL) / (JL * JL)[..., None, None]
    trCL = np.trace(CL, axis1=-2, axis2=-1)
    aL = JL ** (LD(ISO_J_POWER) / LD(dim))
    psiL = (kL / 4) * (JL ** 2 - 2 * np.log(JL) - 1) + mL / 2 * (aL * trCL - dim)
    pL = kL / 2 * (JL - 1 / JL)
    SisoL = (aL * mL)[..., None, None] * (np.eye(dim, dtype=LD) - (trCL / dim)[..., None, None] * invCL)
    SL = (JL * pL)[..., None, None] * invCL + SisoL
    psi = psiL.astype(np.float64)
    S = SL.astype(np.float64)
    S_iso = SisoL.astype(np.float64)
    invC = invCL.astype(np.float64)
    J = JL.astype(np.float64)
    trC = trCL.astype(np.float64)
    a = aL.astype(np.float64)
    kappa = kL.astype(np.float64)
    mu = mL.astype(np.float64)
    p = 0.5 * kappa * (J - 1.0 / J)
    dpdJ = 0.5 * kappa * (1.0 + J ** (-2))
    pairs = _VOIGT[dim]
    S_v = np.stack([S[..., i, j] for (i, j) in pairs], axis=-1)
    if not want_C:
        return psi, S_v, None
    oo = np.einsum("...ij,...kl->...ijkl", invC, invC)
    ds = 0.5 * (np.einsum("...ik,...jl->...ijkl", invC, invC) + np.einsum("...il,...jk->...ijkl", invC, invC))
    Jp = (J * p)[..., None, None, None, None]
    Cvol = (J * p + dpdJ * J ** 2)[..., None, None, None, None] * oo - 2.0 * Jp * ds
    Cf_iso = ((2.0 / dim) * a * mu * trC)[..., None, None, None, None] * (ds - oo / dim) \
        - (2.0 / dim) * (np.einsum("...ij,...kl->...ijkl", invC, S_iso) + np.einsum("...ij,...kl->...ijkl", S_iso, invC))
    Cf = Cvol + Cf_iso
    nv = len(pairs)
    Cv = np.empty(F.shape[:-2] + (nv, nv))
    for I_, (i, j) in enumerate(pairs):
        for J_, (k, l) in enumerate(pairs):
            Cv[..., I_, J_] = Cf[..., i, j, k, l]
    if dim == 2:  # FourthTensorToVoigt 2-D fills the lower triangle from the upper (utils.py:107-118)
        for I_ in range(3):
            for J_ in range(I_):
                Cv[..., I_, J_] = Cv[..., J_, I_]
    return psi, S_v, Cv


def _det_ld(A):
    if A.shape[-1] == 2:
        return A[..., 0, 0] * A[..., 1, 1] - A[..., 0, 1] * A[..., 1, 0]
    return (A[..., 0, 0] * (A[..., 1, 1] * A[..., 2, 2] - A[..., 1, 2] * A[..., 2, 1])
            - A[..., 0, 1] * (A[..., 1, 0] * A[..., 2, 2] - A[..., 1, 2] * A[..., 2, 0])
            + A[..., 0, 2] * (A[..., 1, 0] * A[..., 2, 1] - A[..., 1, 1] * A[..., 2, 0]))


def _adj_ld(A):
    """Adjugate (transpose of the cofactor matrix) of 2x2 / 3x3 stacks, dtype-preserving."""
    if A.shape[-1] == 2:
        out = np.empty_like(A)
        out[..., 0, 0] = A[..., 1, 1]
        out[..., 1, 1] = A[..., 0, 0]
        out[..., 0, 1] = -A[..., 0, 1]
        out[..., 1, 0] = -A[..., 1, 0]
        return out
    out = np.empty_like(A)
    for i in range(3):
        for j in range(3):
            r = [k for k in range(3) if k != j]
            c = [k for k in range(3) if k != i]
            m = A[..., r[0], c[0]] * A[..., r[1], c[1]] - A[..., r[0], c[1]] * A[..., r[1], c[0]]
            out[..., i, j] = m if (i + j) % 2 == 0 else -m
    return out


def nh_kinematics(DN, u):
    """H = grad u, F = I + H, nonlinear B -- CalculateKinematics3D/2D (mechanical_neohooke.py:49-91)."""
    n, dim = DN.shape[-2], DN.shape[-1]
    U = u.reshape(u.shape[0], n, dim)  # (e, a, i)
    H = np.einsum("eai,egaj->egij", U, DN)
    F = H + np.eye(dim)
    lead = DN.shape[:-2]
    if dim == 3:
        B = np.zeros(lead + (6, 3 * n))
        for c in range(3):
            B[..., 0, c::3] = F[..., c, 0, None] * DN[..., 0]
            B[..., 1, c::3] = F[..., c, 1, None] * DN[..., 1]
            B[..., 2, c::3] = F[..., c, 2, None] * DN[..., 2]
            B[..., 3, c::3] = F[..., c, 1, None] * DN[..., 2] + F[..., c, 2, None] * DN[..., 1]
            B[..., 4, c::3] = F[..., c, 0, None] * DN[..., 2] + F[..., c, 2, None] * DN[..., 0]
            B[..., 5, c::3] = F[..., c, 0, None] * DN[..., 1] + F[..., c, 1, None] * DN[..., 0]
    else:
        B = np.zeros(lead + (3, 2 * n))
        for c in range(2):
            B[..., 0, c::2] = F[..., c, 0, None] * DN[..., 0]
            B[..., 1, c::2] = F[..., c, 1, None] * DN[..., 1]
            B[..., 2, c::2] = F[..., c, 1, None] * DN[..., 0] + F[..., c, 0, None] * DN[..., 1]
    return H, F, B


def _voigt_to_mat(S_v, dim):
    if dim == 3:
        S0, S1, S2, S3, S4, S5 = [S_v[..., i] for i in range(6)]
        return np.stack([np.stack([S0, S5, S4], -1), np.stack([S5, S1, S3], -1), np.stack([S4, S3, S2], -1)], -2)
    S0, S1, S2 = [S_v[..., i] for i in range(3)]
    return np.stack([np.stack([S0, S2], -1), np.stack([S2, S1], -1)], -2)


def nh_moduli(e, nu):
    """kappa = e/(3(1-2nu)), mu = e/(2(1+nu)) at each Gauss point (mechanical_neohooke.py:253-255)."""
    return e / (3.0 * (1.0 - 2.0 * nu)), e / (2.0 * (1.0 + nu))


def nh_element(etype, X, d, u, num_gp, nu=0.3, body=None, want_K=True):
    """NeoHookeMechanicalLoss.ComputeElement (fol/loss_functions/mechanical_neohooke.py:243-275).

    Ee = sum w detJ psi; Fint = sum w detJ B^T S; K = sum w detJ (B^T C B + K_geo) with
    K_geo[a*d+i, b*d+j] = delta_ij DN_a^T S DN_b (:171-241).  E_dict is not used (only nu).
    Returns (Ee, Fint - Fe, K).  Pinned by test_neo_hooke_mechanical_loss.py goldens (F=I),
    FD of Ee -> Fint and Fint -> K at random F.
    """
    dim = ELEM_DIM[etype]
    N, w, DN, detJ = geometry(etype, X, num_gp)
    u = np.asarray(u, dtype=np.float64)
    eg = _field_at_gp(N, d)
    kap, mu = nh_moduli(eg, nu)
    H, F, B = nh_kinematics(DN, u)
    psi, S_v, Cv = nh_material(F, kap, mu, want_C=want_K)
    wd = w[None, :] * detJ
    Ee = np.einsum("eg,eg->e", wd, psi)
    Fint = np.einsum("eg,egvi,egv->ei", wd, B, S_v)
    Fe = _body_load(N, w, detJ, np.zeros(dim) if body is None else body, dim)
    K = None
    if want_K:
        Smat = _voigt_to_mat(S_v, dim)
        g = np.einsum("egai,egij,egbj->egab", DN, Smat, DN)
        n = DN.shape[2]
        Kgeo = np.einsum("egab,ij->egaibj", g, np.eye(dim)).reshape(g.shape[0], g.shape[1], n * dim, n * dim)
        K = np.einsum("eg,egvi,egvw,egwj->eij", wd, B, Cv, B) + np.einsum("eg,egij->eij", wd, Kgeo)
    return Ee, Fint - Fe, K


# --------------------------------------------------------------------------
# Element dispatch, VJPs and energy gradients
# --------------------------------------------------------------------------

class Physics:
    """Element physics + parameters, mirroring the loss_settings of the reference losses."""

    def __init__(self, kind, etype, num_gp=None, beta=0.0, c=1.0, E=1.0, nu=0.3, body=None):
        assert kind in ("thermal", "mechanical", "neohooke")
        self.kind, self.etype = kind, etype
        if num_gp is None:  # subclass defaults: thermal.py:61-62,69-70; mechanical.py:127-128,141-142;
            num_gp = 1 if etype == "tetra" else 2  # mechanical_neohooke.py:298-300; fe_loss.py:81-84
        self.num_gp, self.beta, self.c, self.E, self.nu, self.body = num_gp, beta, c, E, nu, body
        self.dim = ELEM_DIM[etype]
        self.ndpn = 1 if kind == "thermal" else self.dim
        self.n = ELEM_NODES[etype]

    def element(self, X, d, u, want_K=True):
        if self.kind == "thermal":
            return thermal_element(self.etype, X, d, u, self.num_gp, self.beta, self.c,
                                   0.0 if self.body is None else self.body, want_K)
        if self.kind == "mechanical":
            return mech_element(self.etype, X, d, u, self.num_gp, self.E, self.nu, self.body, want_K)
        return nh_element(self.etype, X, d, u, self.num_gp, self.nu, self.body, want_K)


def residual_vjp(ph: Physics, X, d, u, wv, picard=False):
    """Element VJP of the residual r_e(u, d) with cotangent w_e:
    g_u = (dr/du)^T w, g_d = (dr/dd)^T w  (per element, local ordering).

    This is the quantity fol/responses/fe_response.py:434-453 contracts
    (jacrev(ComputeElement[1]) with the adjoint vector).  For thermal losses
    `picard=True` reproduces the reference's stop_gradient on T inside k(T)
    (thermal.py:31), i.e. g_u = Se^T w; the default is the exact derivative.
    Pinned by test_sensitivity_analysis.py control-derivative golden and FD.
    """
    wv = np.asarray(wv, dtype=np.float64)
    d = np.asarray(d, dtype=np.float64)
    N, w, DN, detJ = geometry(ph.etype, X, ph.num_gp)
    wd = w[None, :] * detJ
    Nd = _field_weights(N, d.ndim, X.shape[0])
    if ph.kind == "thermal":
        T = np.asarray(u, dtype=np.float64)
        k, dg, Tg = conductivity_at_gp(N, d, T, ph.beta, ph.c)
        gw = np.einsum("egai,ea->egi", DN, wv)
        gT = np.einsum("egai,ea->egi", DN, T)
        q = wd * np.einsum("egi,egi->eg", gw, gT)  # w^T G_g T
        Se = np.einsum("eg,egai,egbi->eab", k * wd, DN, DN)
        g_u = np.einsum("eab,ea->eb", Se, wv)
        if not picard:
            dk = dg * ph.beta * ph.c * Tg ** (ph.c - 1.0) if ph.beta != 0.0 else np.zeros_like(dg)
            g_u = g_u + np.einsum("eg,ga->ea", q * dk, N)
        g_d = np.einsum("eg,ga->ea", q * (1.0 + ph.beta * Tg ** ph.c), Nd)
    elif ph.kind == "mechanical":
        D = d_matrix(ph.dim, ph.E, ph.nu)
        B = b_matrix(DN)
        eg = _field_at_gp(N, d)
        u = np.asarray(u, dtype=np.float64)
        Bu = np.einsum("egvi,ei->egv", B, u)
        Bw = np.einsum("egvi,ei->egv", B, wv)
        sig = np.einsum("vw,egw->egv", D, Bu)
        g_u = np.einsum("eg,egvi,vw,egw->ei", wd * eg, B, D, Bw)
        g_d = np.einsum("eg,ga->ea", wd * np.einsum("egv,egv->eg", Bw, sig), Nd)
    else:
        u = np.asarray(u, dtype=np.float64)
        _, _, K = nh_element(ph.etype, X, d, u, ph.num_gp, ph.nu, None, want_K=True)
        g_u = np.einsum("eij,ei->ej", K, wv)
        k1, m1 = nh_moduli(np.ones_like(detJ), ph.nu)
        H, F, B = nh_kinematics(DN, u)
        _, S1, _ = nh_material(F, k1, m1, want_C=False)
        Bw = np.einsum("egvi,ei->egv", B, wv)
        g_d = np.einsum("eg,ga->ea", wd * np.einsum("egv,egv->eg", Bw, S1), Nd)
    if d.ndim == 1:
        g_d = g_d[:, 0]
    return g_u, g_d


def energy_grad(ph: Physics, X, d, u):
    """Element potential Pi_e and its gradient (g_u = dPi/du = r_e, g_d = dPi/dd).

    mechanical: Pi = 1/2 u^T Se u - u^T Fe (the reference's loss u^T stop_grad(Se u - Fe),
      mechanical.py:116-117, has jax.grad_u = r_e as here, and jax.grad_d = 0 by stop_gradient);
    thermal (beta = 0 only): Pi = 1/2 T^T Se T - T^T Fe;
    neohooke: Pi = Ee - u^T Fe with Ee of mechanical_neohooke.py:266-275 (no stop_gradient:
      jax.grad of ComputeTotalEnergy equals this gradient when the body force is zero).
    Pinned by FD (central differences) and by r_e from the element goldens.
    """
    d = np.asarray(d, dtype=np.float64)
    N, w, DN, detJ = geometry(ph.etype, X, ph.num_gp)
    wd = w[None, :] * detJ
    Nd = _field_weights(N, d.ndim, X.shape[0])
    u = np.asarray(u, dtype=np.float64)
    if ph.kind == "thermal":
        if ph.beta != 0.0:
            raise ValueError("thermal potential exists only for beta = 0")
        _, r, Se = thermal_element(ph.etype, X, d, u, ph.num_gp, 0.0, 1.0,
                                   0.0 if ph.body is None else ph.body, want_K=True)
        Fe = np.einsum("eab,eb->ea", Se, u) - r
        Pi = 0.5 * np.einsum("ea,eab,eb->e", u, Se, u) - np.einsum("ea,ea->e", u, Fe)
        gT = np.einsum("egai,ea->egi", DN, u)
        g_d = np.einsum("eg,ga->ea", 0.5 * wd * np.einsum("egi,egi->eg", gT, gT), Nd)
    elif ph.kind == "mechanical":
        _, r, Ke = mech_element(ph.etype, X, d, u, ph.num_gp, ph.E, ph.nu, ph.body, want_K=True)
        Fe = np.einsum("eij,ej->ei", Ke, u) - r
        Pi = 0.5 * np.einsum("ei,eij,ej->e", u, Ke, u) - np.einsum("ei,ei->e", u, Fe)
        D = d_matrix(ph.dim, ph.E, ph.nu)
        Bu = np.einsum("egvi,ei->egv", b_matrix(DN), u)
        g_d = np.einsum("eg,ga->ea", 0.5 * wd * np.einsum("egv,vw,egw->eg", Bu, D, Bu), Nd)
    else:
        Ee, r, _ = nh_element(ph.etype, X, d, u, ph.num_gp, ph.nu, ph.body, want_K=False)
        Fe =_body_load(N, w, detJ, np.zeros(ph.dim) if ph.body is None else ph.body, ph.dim)
        Pi = Ee - np.einsum("ei,ei->e", u, Fe)
        k1, m1 = nh_moduli(np.ones_like(detJ), ph.nu)
        _, F, _ = nh_kinematics(DN, u)
        psi1, _, _ = nh_material(F, k1, m1, want_C=False)
        g_d = np.einsum("eg,ga->ea", wd * psi1, Nd)
    if d.ndim == 1:
        g_d = g_d[:, 0]
    return Pi, r, g_d


# --------------------------------------------------------------------------
# L4: gather / scatter / Dirichlet / assembly / batch loss
# --------------------------------------------------------------------------

def element_dofs(conn, ndpn):
    """Global dof ids per element, node-major interleaved: ndpn*node + comp (fe_loss.py:163, :178-181)."""
    conn = np.asarray(conn, dtype=np.int64)
    return (ndpn * conn[:, :, None] + np.arange(ndpn)[None, None, :]).reshape(conn.shape[0], -1)


def gather(mesh_X, conn, field):
    return np.asarray(field)[np.asarray(conn)]


def dirichlet_dofs(node_sets, dofs_list, bc_dict):
    """FiniteElementLoss.__CreateDofsDict (fe_loss.py:34-55): indices ndpn*node + dof_index,
    concatenated per dof, then per boundary set (dict order); values broadcast; free dofs =
    setdiff1d(all, dirichlet) (sorted, unique)."""
    ndpn = len(dofs_list)
    idx, val = [], []
    for di, dof in enumerate(dofs_list):
        for name, value in bc_dict.get(dof, {}).items():
            nodes = np.asarray(node_sets[name], dtype=np.int64)
            idx.append(ndpn * nodes + di)
            val.append(np.full(nodes.shape, float(value)))
    if idx:
        return np.concatenate(idx), np.concatenate(val)
    return np.zeros(0, np.int64), np.zeros(0)


def free_dofs(n_dofs, dirichlet_idx):
    return np.setdiff1d(np.arange(n_dofs), dirichlet_idx)


def apply_dirichlet(u, dirichlet_idx, dirichlet_val, load_increment=1.0):
    """ApplyDirichletBCOnDofVector (fe_loss.py:186-189); last write wins for repeated indices."""
    out = np.array(u, dtype=np.float64, copy=True)
    out[..., dirichlet_idx] = load_increment * dirichlet_val
    return out


def assemble(ph: Physics, coords, conn, field, u, dirichlet_idx, transpose=False, want_K=True):
    """ComputeJacobianMatrixAndResidualVector (fe_loss.py:264-318) + JaxDirectLinearSolver's
    BCOO -> scipy csr_array conversion (fe_solver.py:71-72).

    Per element: ke optionally transposed (:219-228) *before* masking; r <- diag(bc) r;
    K <- diag(bc) K + diag(mask * diag(K)) (:191-207).  Returns (csr_array K, R).
    Pinned by test_sensitivity_analysis.py (transposed row 8) and test_mechanical_2D_sa.py.
    """
    conn = np.asarray(conn, dtype=np.int64)
    X = np.asarray(coords, dtype=np.float64)[conn]
    f = np.asarray(field, dtype=np.float64)
    d = f[conn] if f.ndim == 1 and f.shape[0] == coords.shape[0] else f
    dofs = element_dofs(conn, ph.ndpn)
    ue = np.asarray(u, dtype=np.float64)[dofs]
    _, r, K = ph.element(X, d, ue, want_K=want_K)
    n_dofs = ph.ndpn * coords.shape[0]
    bc = np.ones(n_dofs)
    bc[dirichlet_idx] = 0.0
    mask = 1.0 - bc
    bce, me = bc[dofs], mask[dofs]
    R = np.zeros(n_dofs)
    np.add.at(R, dofs, bce * r)
    if not want_K:
        return None, R
    if transpose:
        K = np.transpose(K, (0, 2, 1))
    Km = bce[:, :, None] * K
    diag = np.einsum("eii->ei", K)
    idx = np.arange(K.shape[1])
    Km[:, idx, idx] += me * diag
    nl = dofs.shape[1]
    rows = np.repeat(dofs, nl, axis=1).reshape(-1)
    cols = np.tile(dofs, (1, nl)).reshape(-1)
    A = sp.csr_array((Km.reshape(-1), (rows, cols)), shape=(n_dofs, n_dofs))
    return A, R


def csr_pattern(conn, ndpn, n_nodes):
    """Sparsity of the assembled tangent: scipy csr_array semantics on the reference COO
    (duplicates summed, columns sorted, explicit zeros kept).  Returns (indptr, indices) int64.
    Pinned by a brute-force node-adjacency construction and the closed-form nnz (3N-2)^d ndpn^2."""
    dofs = element_dofs(conn, ndpn)
    nl = dofs.shape[1]
    rows = np.repeat(dofs, nl, axis=1).reshape(-1)
    cols = np.tile(dofs, (1, nl)).reshape(-1)
    n = ndpn * n_nodes
    A = sp.csr_array((np.ones(rows.size), (rows, cols)), shape=(n, n))
    return A.indptr.astype(np.int64), A.indices.astype(np.int64)


def total_energy(ph: Physics, coords, conn, field, u):
    """ComputeTotalEnergy (fe_loss.py:166-176): sum of element energies (reference definitions)."""
    conn = np.asarray(conn, dtype=np.int64)
    X = np.asarray(coords, dtype=np.float64)[conn]
    f = np.asarray(field, dtype=np.float64)
    d = f[conn] if f.ndim == 1 and f.shape[0] == coords.shape[0] else f
    ue = np.asarray(u, dtype=np.float64)[element_dofs(conn, ph.ndpn)]
    e, _, _ = ph.element(X, d, ue, want_K=False)
    return e.sum()


def batch_loss(ph: Physics, coords, conn, batch_fields, batch_dofs, dirichlet_idx, dirichlet_val, exponent=1.0):
    """ComputeBatchLoss (fe_loss.py:250-262): overwrite Dirichlet dofs (GetFullDofVector :91-100),
    per-sample (sum_e E_e)^p, returns (mean, (min, max, mean))."""
    bd = apply_dirichlet(np.atleast_2d(batch_dofs), dirichlet_idx, dirichlet_val)
    bf = np.atleast_2d(batch_fields)
    es = np.array([total_energy(ph, coords, conn, bf[b], bd[b]) ** exponent for b in range(bd.shape[0])])
    return es.mean(), (es.min(), es.max(), es.mean())


def scatter_nodal(conn, vals, n_nodes, ndpn):
    out = np.zeros(ndpn * n_nodes)
    np.add.at(out, element_dofs(conn, ndpn), vals)
    return out


def global_residual_vjp(ph, coords, conn, field, u, w, picard=False):
    """Assembled residual VJP: g_u = (dR/du)^T w, g_d = (dR/dd)^T w (no BC masking).  Pinned through
    residual_vjp (control-derivative golden, central differences) and the assembled sensitivity goldens."""
    conn = np.asarray(conn, dtype=np.int64)
    X = np.asarray(coords, dtype=np.float64)[conn]
    f = np.asarray(field, dtype=np.float64)
    nodal = f.ndim == 1 and f.shape[0] == coords.shape[0]
    d = f[conn] if nodal else f
    dofs = element_dofs(conn, ph.ndpn)
    gu_e, gd_e = residual_vjp(ph, X, d, np.asarray(u)[dofs], np.asarray(w)[dofs], picard)
    g_u = np.zeros(ph.ndpn * coords.shape[0])
    np.add.at(g_u, dofs, gu_e)
    if nodal:
        g_d = np.zeros(coords.shape[0])
        np.add.at(g_d, conn, gd_e)
    else:
        g_d = gd_e
    return g_u, g_d


def global_residual_shape_vjp(ph, coords, conn, field, u, w, h=None):
    """(dR/dX)^T w assembled per node component, (Nn, 3): the adjoint shape-derivative term of
    fol/responses/fe_response.py:312-332 (jacrev of ComputeElement[1] w.r.t. the element coordinates,
    contracted with the element adjoint) and :365-406 (assembly).

    Thermal and linear-elastic elements are differentiated by complex step (exact to rounding: the
    geometry keeps complex coordinates).  The Neo-Hookean oracle evaluates psi/S in long double, which
    has no complex form, so it takes central differences (step h, accuracy ~1e-9 relative).
    Pinned by the reference's shape-derivative goldens (test_sensitivity_analysis.py,
    test_mechanical_2D_sa.py), by central differences of the assembled residual and by translation
    invariance (the columns of a rigid shift sum to zero)."""
    coords = np.asarray(coords, dtype=np.float64)
    conn = np.asarray(conn)
    field = np.asarray(field, dtype=np.float64)
    u = np.asarray(u, dtype=np.float64)
    w = np.asarray(w, dtype=np.float64)
    dofs = element_dofs(conn, ph.ndpn)
    X = coords[conn]
    de = field[conn] if field.shape[0] == coords.shape[0] and field.ndim == 1 and field.shape[0] != conn.shape[0] \
        else field
    ue, we = u[dofs], w[dofs]
    n = conn.shape[1]
    dim = ELEM_DIM[ph.etype]
    gx = np.zeros((coords.shape[0], 3))
    for a in range(n):
        for i in range(dim):
            if ph.kind != "neohooke":
                hc = 1e-30
                Xc = X.astype(complex)
                Xc[:, a, i] += 1j * hc
                _, r, _ = ph.element(Xc, de, ue, want_K=False)
                dr = np.imag(r) / hc
            else:
                hh = h if h is not None else 1e-6 * max(1.0, float(np.abs(X).max()))
                Xp, Xm = X.copy(), X.copy()
                Xp[:, a, i] += hh
                Xm[:, a, i] -= hh
                dr = (ph.element(Xp, de, ue, want_K=False)[1] - ph.element(Xm, de, ue, want_K=False)[1]) / (2 * hh)
            np.add.at(gx[:, i], conn[:, a], np.einsum("ek,ek->e", np.real(dr), we))
    return gx


def global_energy_grad(ph, coords, conn, field, u):
    """Assembled potential: (sum Pi_e, R = dPi/du, dPi/dd)."""
    conn = np.asarray(conn, dtype=np.int64)
    X = np.asarray(coords, dtype=np.float64)[conn]
    f = np.asarray(field, dtype=np.float64)
    nodal = f.ndim == 1 and f.shape[0] == coords.shape[0]
    d = f[conn] if nodal else f
    dofs = element_dofs(conn, ph.ndpn)
    Pi, r, gd_e = energy_grad(ph, X, d, np.asarray(u)[dofs])
    R = np.zeros(ph.ndpn * coords.shape[0])
    np.add.at(R, dofs, r)
    if nodal:
        g_d = np.zeros(coords.shape[0])
        np.add.at(g_d, conn, gd_e)
    else:
        g_d = gd_e
    return Pi.sum(), R, g_d


def residual_norm_loss_grad(ph, coords, conn, field, u, dirichlet_idx, picard=False):
    """L = 1/2 ||R_free||^2 with R the BC-masked residual; returns (L, g_u, g_d) with
    g_u zero at Dirichlet dofs (those entries are overwritten by GetFullDofVector)."""
    _, R = assemble(ph, coords, conn, field, u, dirichlet_idx, want_K=False)
    L = 0.5 * float(R @ R)
    g_u, g_d = global_residual_vjp(ph, coords, conn, field, u, R, picard)
    g_u[dirichlet_idx] = 0.0
    return L, g_u, g_d


# --------------------------------------------------------------------------
# L5: linear solve (oracle = direct solve on the same CSR)
# --------------------------------------------------------------------------

def linear_solve(ph, coords, conn, field, u, dirichlet_idx, dirichlet_val, load_increment=1.0):
    """FiniteElementLinearResidualBasedSolver.Solve (fe_linear_residual_based_solver.py:15-22) with
    JaxDirectLinearSolver (fe_solver.py:69-78): u <- BC(u); (K,R) <- assemble; u + K^-1(-R)."""
    ub = apply_dirichlet(u, dirichlet_idx, dirichlet_val, load_increment)
    A, R = assemble(ph, coords, conn, field, ub, dirichlet_idx)
    du = spla.spsolve(A.tocsc(), -R)
    return ub + du


# --------------------------------------------------------------------------
# L0: meshes, node sets, orientation, colouring
# --------------------------------------------------------------------------

def square_quad_mesh(N, L=1.0):
    """create_2D_square_mesh (fol/tools/usefull_functions.py:213-258): N nodes per side, x fastest,
    element e = i*(N-1)+j -> [id, id+1, id+N+1, id+N] (CCW); left/right node sets."""
    Ne = N - 1
    x = np.linspace(0.0, L, N)
    X, Y = np.meshgrid(x, x)
    coords = np.stack([X.ravel(), Y.ravel(), np.zeros(N * N)], axis=1)
    i, j = np.meshgrid(np.arange(Ne), np.arange(Ne), indexing="ij")
    base = (i * N + j).ravel()
    conn = np.stack([base, base + 1, base + N + 1, base + N], axis=1).astype(np.int32)
    sets = {"left": np.arange(0, N * N, N), "right": np.arange(N - 1, N * N, N),
            "bottom": np.arange(0, N), "top": np.arange(N * (N - 1), N * N)}
    return coords, conn, sets


def box_hex_mesh(nx, ny, nz, Lx=1.0, Ly=1.0, Lz=1.0):
    """Structured HEX8 box, lexicographic x-fastest nodes (id = i + (nx+1)(j + (ny+1)k)), element
    local order (---),(+--),(++-),(-+-),(--+),(+-+),(+++),(-++) of hexahedra_3d_8.py:79-88.
    (The reference builds boxes with gmsh, usefull_functions.py:146-211, whose numbering is not
    reproducible offline; this numbering is the one the tests and benchmarks use.)"""
    xs, ys, zs = np.linspace(0, Lx, nx + 1), np.linspace(0, Ly, ny + 1), np.linspace(0, Lz, nz + 1)
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    coords = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    sx, sy = 1, nx + 1
    sz = (nx + 1) * (ny + 1)
    b = (i * sx + j * sy + k * sz).ravel()
    conn = np.stack([b, b + sx, b + sx + sy, b + sy, b + sz, b + sx + sz, b + sx + sy + sz, b + sy + sz], axis=1)
    nid = np.arange(coords.shape[0])
    sets = {"left": nid[coords[:, 0] == 0], "right": nid[coords[:, 0] == Lx],
            "bottom": nid[coords[:, 2] == 0], "top": nid[coords[:, 2] == Lz]}
    return coords, conn.astype(np.int32), sets


_KUHN_PERMS = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]


def box_tet_mesh(n, L=1.0, jitter=0.0, seed=None):
    """Structured cube of n^3 hexes split into 6 Kuhn tets each (path 000 -> 111 along the axis
    permutations), node jitter U(-jitter*h, jitter*h)^3 on interior nodes only, then the
    reference orientation fix (swap local nodes 0<->1 where detJ < 0, fol/mesh_input_output/mesh.py:123-142)."""
    coords, hconn, sets = box_hex_mesh(n, n, n, L, L, L)
    h = L / n
    sx, sy, sz = 1, n + 1, (n + 1) ** 2
    b = hconn[:, 0].astype(np.int64)
    off = [sx, sy, sz]
    tets = []
    for perm in _KUHN_PERMS:
        p1 = b + off[perm[0]]
        p2 = p1 + off[perm[1]]
        p3 = p2 + off[perm[2]]
        tets.append(np.stack([b, p1, p2, p3], axis=1))
    conn = np.stack(tets, axis=1).reshape(-1, 4)
    if jitter > 0.0:
        rng = np.random.default_rng(seed)
        interior = np.all((coords > 1e-12 * L) & (coords < L - 1e-12 * L), axis=1)
        while True:
            c2 = coords.copy()
            c2[interior] += rng.uniform(-jitter * h, jitter * h, size=(int(interior.sum()), 3))
            conn2 = orient_elements("tetra", c2, conn)
            if np.all(_tet_dets(c2, conn2) > 0.0):
                coords, conn = c2, conn2
                break
    else:
        conn = orient_elements("tetra", coords, conn)
    return coords, conn.astype(np.int32), sets


def _tet_dets(coords, conn):
    X = coords[conn]
    return np.linalg.det(np.stack([X[:, 1] - X[:, 0], X[:, 2] - X[:, 0], X[:, 3] - X[:, 0]], axis=2))


def orient_elements(etype, coords, conn):
    """Mesh.CheckAndOrientElements (mesh.py:123-142): detJ at the one-point rule; where negative,
    swap local nodes 0 and 1."""
    conn = np.array(conn, dtype=np.int64, copy=True)
    pts, _ = gauss_rule(etype, 1)
    dim = ELEM_DIM[etype]
    J = np.einsum("eai,ak->eik", coords[conn][..., :dim], shape_local_grads(etype, pts[0]))
    neg = np.linalg.det(J) < 0
    conn[neg, 0], conn[neg, 1] = conn[neg, 1].copy(), conn[neg, 0].copy()
    return conn


def greedy_colouring(conn, n_nodes):
    """Deterministic greedy first-fit element colouring in element order over the shared-node
    graph: an element takes the lowest colour not used by any earlier element sharing a node.
    (Not in the reference; pinned by specification + validity + known counts.)"""
    conn = np.asarray(conn, dtype=np.int64)
    used = [0] * n_nodes
    colours = np.empty(conn.shape[0], dtype=np.int32)
    for e, nodes in enumerate(conn.tolist()):
        m = 0
        for a in nodes:
            m |= used[a]
        c = (~m & (m + 1)).bit_length() - 1
        colours[e] = c
        bit = 1 << c
        for a in nodes:
            used[a] |= bit
    return colours


# --------------------------------------------------------------------------
# Synthetic input fields
# --------------------------------------------------------------------------

def sigmoid(x):
    return 1.0 / (1.0 + np.exp(-x))


def fourier_field(coords, coeffs, x_freqs, y_freqs, z_freqs=(0,), beta=1.0, vmin=0.1, vmax=1.0, c0=1.0):
    """FourierControl.ComputeControlledVariables (fol/controls/fourier_control.py:49-59) with
    variable_vector = [c0, coeffs...]; frequency order = meshgrid(x,y,z, indexing='ij') raveled."""
    fx, fy, fz = np.meshgrid(np.asarray(x_freqs), np.asarray(y_freqs), np.asarray(z_freqs), indexing="ij")
    fr = np.stack([fx.ravel(), fy.ravel(), fz.ravel()], axis=1).astype(np.float64)
    x, y, z = coords[:, 0], coords[:, 1], coords[:, 2]
    K = c0 / 2.0 + sum(cf * np.cos(f[0] * np.pi * x) * np.cos(f[1] * np.pi * y) * np.cos(f[2] * np.pi * z)
                       for f, cf in zip(fr, np.asarray(coeffs, dtype=np.float64)))
    return (vmax - vmin) * sigmoid(beta * (K - 0.5)) + vmin


def grf_two_phase(shape, corr_len, seed, lo=1.0, hi=10.0):
    """Seeded Gaussian random field on a node lattice `shape` (nz, ny, nx order of the flattened
    x-fastest numbering): white noise N(0,1), Gaussian filter of correlation length `corr_len`
    lattice spacings applied in Fourier space (periodic), thresholded at 0 -> {lo, hi}."""
    rng = np.random.default_rng(seed)
    noise = rng.standard_normal(shape)
    ks = [np.fft.fftfreq(s) * 2.0 * np.pi for s in shape]
    K2 = sum(np.meshgrid(*[k ** 2 for k in ks], indexing="ij"))
    field = np.real(np.fft.ifftn(np.fft.fftn(noise) * np.exp(-0.5 * corr_len ** 2 * K2)))
    return np.where(field > 0.0, hi, lo).ravel()
