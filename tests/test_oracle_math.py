"""Pins of the oracle that the reference goldens leave open (CPU only): closed forms, invariants,
finite differences, brute-force assembly, known colouring/pattern counts -- and planted mistakes
in the oracle that these checks must catch."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import fe
from tests.cases import make_case

ALL = [(k, e) for k in ("thermal", "mechanical", "neohooke") for e in ("quad", "hexahedron", "tetra")]


def _R(case, u, f):
    ph = case["ph"]
    X = case["coords"][case["conn"]]
    ff = f[case["conn"]] if f.shape[0] == case["coords"].shape[0] else f
    _, r, _ = ph.element(X, ff, u[fe.element_dofs(case["conn"], ph.ndpn)], want_K=False)
    return fe.scatter_nodal(case["conn"], r, case["coords"].shape[0], ph.ndpn)


@pytest.mark.parametrize("kind,etype", ALL)
def test_residual_vjp_matches_central_differences(kind, etype):
    case = make_case(kind, etype, n=2, seed=21, **(dict(beta=0.5, c=2.0) if kind == "thermal" else {}))
    ph = case["ph"]
    gu, gd = fe.global_residual_vjp(ph, case["coords"], case["conn"], case["field"], case["u"], case["w"])
    rng = np.random.default_rng(5)
    eps = 1e-6
    for _ in range(3):
        v = rng.standard_normal(case["u"].shape)
        fd = case["w"] @ (_R(case, case["u"] + eps * v, case["field"]) - _R(case, case["u"] - eps * v, case["field"]))
        assert abs(fd / (2 * eps) - gu @ v) <= 1e-7 * max(1.0, abs(gu @ v))
        vf = rng.standard_normal(case["field"].shape)
        fd = case["w"] @ (_R(case, case["u"], case["field"] + eps * vf) - _R(case, case["u"], case["field"] - eps * vf))
        assert abs(fd / (2 * eps) - gd @ vf) <= 1e-7 * max(1.0, abs(gd @ vf))


def test_picard_vjp_is_the_frozen_conductivity_transpose():
    """thermal.py:31 stop_gradient: jax.grad through the reference residual gives Se^T w."""
    case = make_case("thermal", "quad", n=3, seed=22, beta=0.5, c=2.0)
    ph = case["ph"]
    A, _ = fe.assemble(ph, case["coords"], case["conn"], case["field"], case["u"], np.zeros(0, np.int64))
    gu, _ = fe.global_residual_vjp(ph, case["coords"], case["conn"], case["field"], case["u"], case["w"], picard=True)
    np.testing.assert_allclose(gu, A.T @ case["w"], rtol=1e-12, atol=1e-12)
    gx, _ = fe.global_residual_vjp(ph, case["coords"], case["conn"], case["field"], case["u"], case["w"])
    assert np.linalg.norm(gx - gu) > 1e-3 * np.linalg.norm(gu)  # beta != 0: exact differs


@pytest.mark.parametrize("kind,etype", ALL)
def test_energy_gradient_matches_central_differences(kind, etype):
    case = make_case(kind, etype, n=2, seed=23)
    ph = case["ph"]
    P = lambda u, f: fe.global_energy_grad(ph, case["coords"], case["conn"], f, u)[0]
    _, R, gd = fe.global_energy_grad(ph, case["coords"], case["conn"], case["field"], case["u"])
    np.testing.assert_allclose(R, _R(case, case["u"], case["field"]), rtol=1e-12, atol=1e-14)
    rng = np.random.default_rng(6)
    eps = 1e-6
    v = rng.standard_normal(case["u"].shape)
    fd = (P(case["u"] + eps * v, case["field"]) - P(case["u"] - eps * v, case["field"])) / (2 * eps)
    assert abs(fd - R @ v) <= 1e-7 * max(1.0, abs(fd))
    vf = rng.standard_normal(case["field"].shape)
    fd = (P(case["u"], case["field"] + eps * vf) - P(case["u"], case["field"] - eps * vf)) / (2 * eps)
    assert abs(fd - gd @ vf) <= 1e-7 * max(1.0, abs(fd))


@pytest.mark.parametrize("kind,etype", [(k, e) for k in ("mechanical", "neohooke") for e in ("quad", "hexahedron", "tetra")])
def test_tangent_is_derivative_of_residual_and_symmetric(kind, etype):
    case = make_case(kind, etype, n=2, seed=24)
    ph = case["ph"]
    A, _ = fe.assemble(ph, case["coords"], case["conn"], case["field"], case["u"], np.zeros(0, np.int64))
    assert abs(A - A.T).max() <= 1e-12 * abs(A).max()
    rng = np.random.default_rng(7)
    v = rng.standard_normal(case["u"].shape)
    eps = 1e-6
    fd = (_R(case, case["u"] + eps * v, case["field"]) - _R(case, case["u"] - eps * v, case["field"])) / (2 * eps)
    assert np.linalg.norm(A @ v - fd) <= 1e-7 * np.linalg.norm(fd)


def test_unit_square_laplacian_closed_form():
    """QUAD4 unit square, k = 1, 2x2 Gauss: Ke = 1/6 [[4,-1,-2,-1],[-1,4,-1,-2],[-2,-1,4,-1],[-1,-2,-1,4]]."""
    X = np.array([[[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0.0]]])
    _, _, K = fe.thermal_element("quad", X, np.ones((1, 4)), np.zeros((1, 4)), 2)
    ref = np.array([[4, -1, -2, -1], [-1, 4, -1, -2], [-2, -1, 4, -1], [-1, -2, -1, 4]]) / 6.0
    np.testing.assert_allclose(K[0], ref, rtol=0, atol=1e-15)


@pytest.mark.parametrize("etype", ["quad", "hexahedron", "tetra"])
def test_patch_tests(etype):
    """Affine T (thermal) and affine u (elasticity): interior residual vanishes on distorted meshes."""
    case = make_case("thermal", etype, n=3, seed=25)
    X = case["coords"]
    interior = np.all((X > 1e-9) & (X < 1 - 1e-9), axis=1) if etype != "quad" else \
        np.all((X[:, :2] > 1e-9) & (X[:, :2] < 1 - 1e-9), axis=1)
    T = 0.3 + X[:, 0] - 2 * X[:, 1] + (0.5 * X[:, 2] if etype != "quad" else 0)
    R = _R(case, T, np.ones(X.shape[0]))
    assert np.max(np.abs(R[interior])) < 1e-12
    case = make_case("mechanical", etype, n=3, seed=25)
    d = case["ph"].dim
    Amat = np.random.default_rng(1).standard_normal((d, d))
    u = (X[:, :d] @ Amat.T).ravel()
    R = _R(case, u, np.ones(X.shape[0])).reshape(-1, d)
    assert np.max(np.abs(R[interior])) < 1e-12


@pytest.mark.parametrize("etype", ["quad", "hexahedron", "tetra"])
def test_rigid_body_modes_in_null_space(etype):
    case = make_case("mechanical", etype, n=2, seed=26)
    ph = case["ph"]
    A, _ = fe.assemble(ph, case["coords"], case["conn"], case["field"], case["u"], np.zeros(0, np.int64))
    X = case["coords"][:, :ph.dim]
    modes = [np.tile(np.eye(ph.dim)[i], X.shape[0]) for i in range(ph.dim)]
    if ph.dim == 2:
        modes.append(np.stack([-X[:, 1], X[:, 0]], 1).ravel())
    else:
        for (i, j) in [(0, 1), (1, 2), (0, 2)]:
            m = np.zeros_like(X)
            m[:, i], m[:, j] = -X[:, j], X[:, i]
            modes.append(m.ravel())
    for m in modes:
        assert np.linalg.norm(A @ m) <= 1e-12 * abs(A).max() * np.linalg.norm(m)


def test_linear_elastic_energy_identity():
    """u^T K u = integral of eps : C : eps (sum of element potentials times two)."""
    case = make_case("mechanical", "hexahedron", n=2, seed=27)
    ph = case["ph"]
    A, _ = fe.assemble(ph, case["coords"], case["conn"], case["field"], case["u"], np.zeros(0, np.int64))
    Pi, _, _ = fe.global_energy_grad(ph, case["coords"], case["conn"], case["field"], case["u"])
    np.testing.assert_allclose(case["u"] @ A @ case["u"], 2 * Pi, rtol=1e-12)


@pytest.mark.parametrize("etype", ["hexahedron", "tetra"])
def test_neo_hooke_at_identity_equals_linear_elasticity(etype):
    """3-D: NH tangent at F = I is linear elasticity with lambda = kappa - 2 mu/3."""
    case = make_case("neohooke", etype, n=2, seed=28)
    X = case["coords"][case["conn"]]
    d = case["field"][case["conn"]]
    zero = np.zeros((X.shape[0], case["ph"].n * 3))
    _, rn, Kn = fe.nh_element(etype, X, d, zero, case["ph"].num_gp, nu=0.3)
    _, _, Kl = fe.mech_element(etype, X, d, zero, case["ph"].num_gp, E=1.0, nu=0.3)
    np.testing.assert_allclose(Kn, Kl, rtol=1e-12, atol=1e-13)
    assert np.abs(rn).max() < 1e-14


def test_csr_pattern_matches_brute_force_and_closed_form():
    """Pattern = scipy csr of all element blocks; structured counts (3N-2)^d per node-block."""
    for N, dim, nd in [(41, 2, 1), (9, 3, 3)]:
        if dim == 2:
            coords, conn, _ = fe.square_quad_mesh(N)
        else:
            coords, conn, _ = fe.box_hex_mesh(N - 1, N - 1, N - 1)
        ip, ii = fe.csr_pattern(conn, nd, coords.shape[0])
        assert ip[-1] == (3 * N - 2) ** dim * nd * nd
        dense = np.zeros((nd * coords.shape[0],) * 2, dtype=bool) if coords.shape[0] < 1000 else None
        if dense is not None:
            dofs = fe.element_dofs(conn, nd)
            for row in dofs:
                dense[np.ix_(row, row)] = True
            A = sp.csr_array(dense.astype(float))
            assert np.array_equal(A.indptr, ip) and np.array_equal(A.indices, ii)


def test_config_pattern_sizes():
    """Closed-form nnz at the benchmark sizes: QUAD4 40^2 thermal 14641; HEX8 64^3 elasticity 64701513."""
    assert (3 * 41 - 2) ** 2 == 14641
    assert (3 * 65 - 2) ** 3 * 9 == 64701513
    coords, conn, _ = fe.square_quad_mesh(41)
    ip, _ = fe.csr_pattern(conn, 1, coords.shape[0])
    assert ip[-1] == 14641


@pytest.mark.parametrize("mesh,expect", [("quad", 4), ("hexahedron", 8)])
def test_colouring_valid_and_known_counts(mesh, expect):
    if mesh == "quad":
        coords, conn, _ = fe.square_quad_mesh(9)
    else:
        coords, conn, _ = fe.box_hex_mesh(5, 4, 3)
    col = fe.greedy_colouring(conn, coords.shape[0])
    assert col.max() + 1 == expect
    for c in range(expect):
        nodes = conn[col == c].ravel()
        assert np.unique(nodes).size == nodes.size  # no two same-colour elements share a node
    # first-fit: each element's colour is the smallest not used by earlier neighbours
    assert col[0] == 0 and col[1] == 1


def test_tet_mesh_orientation_and_volume():
    coords, conn, _ = fe.box_tet_mesh(4, 1.0, jitter=0.2, seed=3)
    X = coords[conn]
    det = fe._tet_dets(coords, conn)
    assert np.all(det > 0)
    np.testing.assert_allclose(det.sum() / 6.0, 1.0, rtol=1e-12)


def test_dirichlet_dofs_semantics():
    """fe_loss.py:34-55: order = per dof, then per set; values broadcast; free = setdiff1d."""
    coords, conn, sets = fe.square_quad_mesh(3)
    idx, val = fe.dirichlet_dofs(sets, ["Ux", "Uy"], {"Ux": {"left": 0.0, "right": 0.05}, "Uy": {"left": 0.1}})
    np.testing.assert_array_equal(idx, [0, 6, 12, 4, 10, 16, 1, 7, 13])
    np.testing.assert_array_equal(val, [0, 0, 0, 0.05, 0.05, 0.05, 0.1, 0.1, 0.1])
    np.testing.assert_array_equal(fe.free_dofs(18, idx), [2, 3, 5, 8, 9, 11, 14, 15, 17])


def test_batch_loss_semantics():
    case = make_case("mechanical", "quad", n=2, seed=29)
    ph = case["ph"]
    rng = np.random.default_rng(2)
    bu = rng.uniform(-1, 1, size=(3, case["u"].size))
    bf = rng.uniform(0.5, 1.5, size=(3, case["coords"].shape[0]))
    m, (lo, hi, mean) = fe.batch_loss(ph, case["coords"], case["conn"], bf, bu, case["didx"], case["dval"], 2.0)
    ub = fe.apply_dirichlet(bu, case["didx"], case["dval"])
    es = np.array([fe.total_energy(ph, case["coords"], case["conn"], bf[b], ub[b]) ** 2 for b in range(3)])
    np.testing.assert_allclose([m, lo, hi], [es.mean(), es.min(), es.max()], rtol=1e-14)


# ---------------------------------------------------------------------------------------------
# planted mistakes: each must be caught by one of the pins above
# ---------------------------------------------------------------------------------------------
def test_planted_shear_order_mistake_is_caught(monkeypatch):
    """Swapping the yz/xz shear rows of the 3-D B matrix breaks the HEX8 golden."""
    import json, os
    G = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_goldens.json")))
    t = G["tests/unit/test_mechanical_loss.py"]["test_hexa"]
    orig = fe.b_matrix

    def wrong(DN):
        B = orig(DN)
        if DN.shape[-1] == 3:
            B = B.copy()
            B[..., [4, 5], :] = B[..., [5, 4], :]
            B[..., 4, 0::3], B[..., 4, 2::3] = DN[..., 1], DN[..., 0]  # xz -> xy-like mistake
        return B

    monkeypatch.setattr(fe, "b_matrix", wrong)
    X = np.array(t["assigns"]["hex_points_coordinates"])[None]
    _, _, K = fe.Physics("mechanical", "hexahedron", E=1.0, nu=0.3, body=[1, 2, 3]).element(X, np.ones((1, 8)),
                                                                                           np.ones((1, 24)))
    assert not np.allclose(K[0], t["checks"][0]["expected"], rtol=1e-5, atol=1e-6)


def test_planted_transposed_jacobian_mistake_is_caught(monkeypatch):
    """Using J^T instead of J changes global gradients on distorted elements (geometry golden)."""
    import json, os
    G = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_goldens.json")))
    X = np.array(G["tests/unit/test_geometries.py"]["setUp"]["assigns"]["self.tet_points_coordinates"])
    chk = [c for c in G["tests/unit/test_geometries.py"]["test_tet3D4"]["checks"] if c["actual"] == "shape_function_g_grads"][0]
    orig_inv = np.linalg.inv
    monkeypatch.setattr(np.linalg, "inv", lambda A: orig_inv(np.swapaxes(A, -1, -2)))
    _, _, DN, _ = fe.geometry("tetra", X[None], 1)
    assert not np.allclose(DN[0], chk["expected"], rtol=1e-5, atol=1e-6)


def test_planted_missing_exact_term_is_caught(monkeypatch):
    """Dropping the dk/dT term of the exact thermal VJP fails the finite-difference pin."""
    case = make_case("thermal", "quad", n=2, seed=30, beta=0.5, c=2.0)
    ph = case["ph"]
    gu, _ = fe.global_residual_vjp(ph, case["coords"], case["conn"], case["field"], case["u"], case["w"], picard=True)
    v = np.random.default_rng(8).standard_normal(case["u"].shape)
    eps = 1e-6
    fd = case["w"] @ (_R(case, case["u"] + eps * v, case["field"]) - _R(case, case["u"] - eps * v, case["field"])) / (2 * eps)
    assert abs(fd - gu @ v) > 1e-4 * abs(fd)


def test_c_restatement_matches_numpy_oracle():
    """oracle/cpu_ref.c (the timed CPU baseline) equals the NumPy oracle on a jittered TET4 mesh."""
    from oracle import cpuref
    coords, conn, _ = fe.box_tet_mesh(4, 1.0, jitter=0.2, seed=9)
    rng = np.random.default_rng(10)
    E = np.exp(rng.uniform(0, np.log(100.0), size=conn.shape[0]))
    u = rng.uniform(-1e-3, 1e-3, size=3 * coords.shape[0])
    seg = (np.arange(conn.shape[0]) >= conn.shape[0] // 2).astype(np.int64)
    Pi, R, gE = cpuref.tet4_elastic_energy_grad(coords, conn, E, u, seg_of_elem=seg, n_seg=2)
    ph = fe.Physics("mechanical", "tetra", E=1.0, nu=0.3)
    Pe, r, gde = fe.energy_grad(ph, coords[conn], E, u[fe.element_dofs(conn, 3)])
    np.testing.assert_allclose(R, fe.scatter_nodal(conn, r, coords.shape[0], 3), rtol=1e-12, atol=1e-18)
    np.testing.assert_allclose(gE, gde, rtol=1e-12)
    np.testing.assert_allclose(Pi, [Pe[seg == 0].sum(), Pe[seg == 1].sum()], rtol=1e-12)


def test_c_thermal_restatement_matches_numpy_oracle():
    """oracle/cpu_ref.c's nonlinear-thermal residual-norm loss and gradient (the CPU baseline of the
    QUAD4 128^2 x 64 workload) equal the NumPy oracle on a distorted mesh, sample by sample, with
    one thread and with several."""
    from oracle import cpuref
    case = make_case("thermal", "quad", n=6, seed=33, beta=0.5, c=2.0)
    ph, coords, conn = case["ph"], case["coords"], case["conn"]
    rng = np.random.default_rng(34)
    B, Nn = 3, coords.shape[0]
    K0 = rng.uniform(0.1, 1.0, size=(B, Nn))
    T = rng.uniform(0.0, 1.0, size=(B, Nn))
    T[:, case["didx"]] = case["dval"]
    ref = [fe.residual_norm_loss_grad(ph, coords, conn, K0[b], T[b], case["didx"]) for b in range(B)]
    default = cpuref.num_threads()
    for threads in (1, 4):
        cpuref.set_threads(threads)
        L, gT, gk = cpuref.quad4_thermal_resnorm_vjp(coords, conn, K0, T, case["didx"], beta=0.5, c=2.0)
        np.testing.assert_allclose(L, sum(r[0] for r in ref), rtol=1e-12)
        np.testing.assert_allclose(gT, np.stack([r[1] for r in ref]), rtol=1e-11, atol=1e-14)
        np.testing.assert_allclose(gk, np.stack([r[2] for r in ref]), rtol=1e-11, atol=1e-14)
    cpuref.set_threads(default)


def _wR_of_coords(case, coords):
    ph = case["ph"]
    dofs = fe.element_dofs(case["conn"], ph.ndpn)
    _, r, _ = ph.element(coords[case["conn"]], case["field"][case["conn"]], case["u"][dofs], want_K=False)
    return float(np.einsum("ek,ek->", r, case["w"][dofs]))


@pytest.mark.parametrize("kind,etype", ALL)
def test_shape_vjp_finite_differences_and_translation(kind, etype):
    """(dR/dX)^T w against central differences of w.R(X) along random coordinate directions, and
    translation invariance (R does not change when every node moves by the same vector, so the
    shape gradient sums to zero per component; 2-D elements have a zero z column)."""
    case = make_case(kind, etype, n=2, seed=41, **(dict(beta=0.5, c=2.0) if kind == "thermal" else {}))
    ph = case["ph"]
    gx = fe.global_residual_shape_vjp(ph, case["coords"], case["conn"], case["field"], case["u"], case["w"])
    rng = np.random.default_rng(42)
    for _ in range(3):
        v = rng.standard_normal(case["coords"].shape)
        if ph.dim == 2:
            v[:, 2] = 0.0
        eps = 1e-6
        fd = (_wR_of_coords(case, case["coords"] + eps * v) - _wR_of_coords(case, case["coords"] - eps * v)) / (2 * eps)
        assert abs(fd - np.sum(gx * v)) <= 1e-6 * max(abs(fd), np.abs(gx).sum() * 1e-3)
    # complex step is exact to rounding; the Neo-Hookean oracle takes central differences
    assert np.abs(gx.sum(axis=0)).max() <= (1e-8 if kind == "neohooke" else 1e-12) * np.abs(gx).max()
    if ph.dim == 2:
        assert np.all(gx[:, 2] == 0.0)


def test_planted_shape_vjp_without_detj_term_is_caught():
    """A shape gradient that ignores how detJ moves with the nodes fails the difference pin."""
    case = make_case("mechanical", "quad", n=2, seed=43)
    ph = case["ph"]
    gx = fe.global_residual_shape_vjp(ph, case["coords"], case["conn"], case["field"], case["u"], case["w"])
    N, w, DN, detJ = fe.geometry("quad", case["coords"][case["conn"]], ph.num_gp)
    # planted mistake: drop d(detJ)/dX, i.e. subtract the part of the gradient that scales the whole
    # element integrand (w.r_e times dN_a/dx_i per Gauss point)
    dofs = fe.element_dofs(case["conn"], 2)
    _, r, _ = ph.element(case["coords"][case["conn"]], case["field"][case["conn"]], case["u"][dofs], want_K=False)
    wr_e = np.einsum("ek,ek->e", r, case["w"][dofs])
    wrong = gx.copy()
    for i in range(2):
        np.add.at(wrong[:, i], case["conn"], -np.einsum("e,eg,ega->ea", wr_e / ph.num_gp ** 2, np.ones_like(detJ), DN[..., i]))
    v = np.random.default_rng(44).standard_normal(case["coords"].shape)
    v[:, 2] = 0.0
    eps = 1e-6
    fd = (_wR_of_coords(case, case["coords"] + eps * v) - _wR_of_coords(case, case["coords"] - eps * v)) / (2 * eps)
    assert abs(fd - np.sum(gx * v)) <= 1e-6 * abs(fd)
    assert abs(fd - np.sum(wrong * v)) > 1e-3 * abs(fd)
