"""Mesh ingestion (Kratos .mdpa) and export, against the reference's own mesh fixture
(tests/meshes/coarse_sphere.mdpa of the reference, copied as data to tests/golden/meshes/) and
tests/unit/test_mesh_io.py:24-32's counts; then the mesh through the loss and solver on the GPU."""
import os

import numpy as np
import pytest

from oracle import fe

MESHES = os.path.join(os.path.dirname(__file__), "golden", "meshes")


def _sphere():
    from folax_amd.mesh import Mesh
    m = Mesh("mdpa_io", file_name="coarse_sphere.mdpa", case_dir=MESHES)
    m.Initialize()
    return m


def test_mdpa_counts_match_reference_test():
    """The counts tests/unit/test_mesh_io.py:24-28 asserts."""
    m = _sphere()
    assert m.GetNumberOfNodes() == 85
    assert m.GetNumberOfElements("tetra") == 249
    assert len(m.GetNodeSet("Skin_Part")) == 66
    assert len(m.GetNodeSet("Partial_Skin_Part")) == 13


def test_mdpa_contents_are_geometrically_consistent():
    """0-based connectivity in range; first node as written in the file; skin nodes on the sphere of
    radius 0.5 and interior nodes inside it; the partial skin a subset of the skin; every tet
    positively oriented and the tets fill a volume below the sphere's (inscribed polyhedron)."""
    m = _sphere()
    c, conn = m.GetNodesCoordinates(), m.GetElementsNodes("tetra")
    assert conn.dtype == np.int32 and conn.min() == 0 and conn.max() == m.GetNumberOfNodes() - 1
    np.testing.assert_array_equal(c[0], [0.16372, 0.18066, -0.43653])
    r = np.linalg.norm(c, axis=1)
    skin = m.GetNodeSet("Skin_Part")
    assert np.all(np.abs(r[skin] - 0.5) < 1e-5)
    assert np.delete(r, skin).max() < 0.45
    assert np.isin(m.GetNodeSet("Partial_Skin_Part"), skin).all()
    X = c[conn]
    det = np.linalg.det(np.stack([X[:, 1] - X[:, 0], X[:, 2] - X[:, 0], X[:, 3] - X[:, 0]], axis=2))
    assert np.all(det > 0)
    assert 0.8 * (4 / 3) * np.pi * 0.125 < det.sum() / 6 < (4 / 3) * np.pi * 0.125
    # scale factor (mesh.py:216)
    from folax_amd.mesh import Mesh
    m2 = Mesh("s", file_name="coarse_sphere.mdpa", case_dir=MESHES, scale_factor=2.0)
    m2.Initialize()
    np.testing.assert_allclose(m2.GetNodesCoordinates(), 2.0 * c)


def test_export_round_trip(tmp_path):
    """Finalize writes a legacy VTK file whose points, cells and point data read back."""
    m = _sphere()
    T = np.zeros(m.GetNumberOfNodes())
    T[m.GetNodeSet("Skin_Part")] = 1.0
    T[m.GetNodeSet("Partial_Skin_Part")] = 2.0
    m["T"] = T
    m.Finalize(export_dir=str(tmp_path))
    text = (tmp_path / "coarse_sphere.vtk").read_text().split("\n")
    i = next(k for k, l in enumerate(text) if l.startswith("POINTS"))
    assert int(text[i].split()[1]) == 85
    pts = np.loadtxt(text[i + 1:i + 86])
    np.testing.assert_allclose(pts, m.GetNodesCoordinates())
    j = next(k for k, l in enumerate(text) if l.startswith("CELLS"))
    assert text[j].split()[1:] == ["249", str(249 * 5)]
    cells = np.loadtxt(text[j + 1:j + 250], dtype=np.int64)
    np.testing.assert_array_equal(cells[:, 1:], m.GetElementsNodes("tetra"))
    k = next(k for k, l in enumerate(text) if l.startswith("SCALARS T"))
    np.testing.assert_allclose(np.loadtxt(text[k + 2:k + 87]), T)


def test_med_needs_meshio():
    from folax_amd.mesh import Mesh
    m = Mesh("med_io", file_name="coarse_sphere.mdpa", case_dir=MESHES)
    m.mesh_format = "med"
    with pytest.raises(NotImplementedError):
        m.Initialize()


@pytest.mark.gpu
@pytest.mark.parametrize("solver", ["PCG", "PCG-mixed"])
def test_sphere_elasticity_solve_against_oracle(solver):
    """The read mesh through MechanicalLoss3DTetra and the linear solver: clamped partial skin,
    body force (0, 0, -1), nodal Young's modulus field; equals the oracle's direct solve."""
    torch = pytest.importorskip("torch")
    from folax_amd import loss_functions as lf
    from folax_amd.solvers import FiniteElementLinearResidualBasedSolver
    m = _sphere()
    bc = {d: {"Partial_Skin_Part": 0.0} for d in ("Ux", "Uy", "Uz")}
    loss = lf.MechanicalLoss3DTetra("sphere", {"dirichlet_bc_dict": bc, "body_foce": [0.0, 0.0, -1.0],
                                               "material_dict": {"young_modulus": 1.0, "poisson_ratio": 0.3}}, m)
    loss.Initialize()
    E = np.random.default_rng(61).uniform(1.0, 3.0, m.GetNumberOfNodes())
    s = FiniteElementLinearResidualBasedSolver("lin", loss, {"linear_solver_settings": {"solver": solver, "tol": 1e-12}})
    s.Initialize()
    u = s.Solve(torch.as_tensor(E, device="cuda"), np.zeros(3 * m.GetNumberOfNodes())).cpu().numpy()
    ph = fe.Physics("mechanical", "tetra", E=1.0, nu=0.3, body=[0.0, 0.0, -1.0])
    idx, val = fe.dirichlet_dofs(m.node_sets, ["Ux", "Uy", "Uz"], bc)
    ref = fe.linear_solve(ph, m.GetNodesCoordinates(), m.GetElementsNodes("tetra"), E,
                          np.zeros(3 * m.GetNumberOfNodes()), idx, val)
    assert np.linalg.norm(u - ref) <= 1e-9 * np.linalg.norm(ref)
