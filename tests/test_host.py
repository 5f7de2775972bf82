"""CPU tests of the host side: the C-ABI library loads and exports every declared symbol, host-only
entry points (orientation), mesh generators and loss-object bookkeeping match the oracle."""
import ctypes
import os
import re

import numpy as np
import pytest

from oracle import fe

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    src = open(os.path.join(ROOT, "include", "femhot.h")).read()
    return sorted(set(re.findall(r"\b(femhot_[a-z_0-9]+)\s*\(", src)))


def test_library_loads_and_exports_every_declared_symbol():
    from folax_amd import _lib
    lib = _lib.lib()
    declared = _declared_symbols()
    assert len(declared) >= 14
    for name in declared:
        assert hasattr(lib, name), name
    assert sorted(_lib.EXPORTED) == declared
    assert lib.femhot_version().decode().startswith("femhot")


def test_error_reporting_without_gpu():
    from folax_amd import _lib
    lib = _lib.lib()
    rc = lib.femhot_ctx_create(None, None)
    assert rc == _lib.ERR_ARG
    assert "null" in lib.femhot_last_error().decode()
    # connectivity out of range is rejected by the host-side orientation routine
    coords = np.zeros((3, 3))
    conn = np.array([[0, 1, 2, 7]], dtype=np.int32)
    with pytest.raises(_lib.FemhotError) as ei:
        _lib.orient_elements(_lib.TET4, coords, conn)
    assert ei.value.code == _lib.ERR_INDEX


@pytest.mark.parametrize("etype", ["quad", "hexahedron", "tetra"])
def test_orientation_matches_oracle(etype):
    from folax_amd import _lib
    rng = np.random.default_rng(4)
    if etype == "quad":
        coords, conn, _ = fe.square_quad_mesh(6)
    elif etype == "hexahedron":
        coords, conn, _ = fe.box_hex_mesh(3, 3, 3)
    else:
        coords, conn, _ = fe.box_tet_mesh(3)
    conn = conn.copy()
    flip = rng.random(conn.shape[0]) < 0.5
    # mirrored local order: negative detJ at the element centre
    mirror = {"quad": [0, 3, 2, 1], "hexahedron": [0, 3, 2, 1, 4, 7, 6, 5], "tetra": [1, 0, 2, 3]}[etype]
    conn[flip] = conn[flip][:, mirror]
    ours, nsw = _lib.orient_elements(_lib.ELEMENT_TYPES[etype], coords, conn)
    ref = fe.orient_elements(etype, coords, conn)
    assert np.array_equal(ours, ref)
    assert nsw == int(flip.sum())


def test_mesh_generators_match_oracle():
    from folax_amd import mesh
    m = mesh.create_2D_square_mesh(1.0, 7)
    c, k, s = fe.square_quad_mesh(7, 1.0)
    assert np.array_equal(m.GetElementsNodes("quad"), k) and np.allclose(m.GetNodesCoordinates(), c)
    assert np.array_equal(m.GetNodeSet("left"), s["left"]) and np.array_equal(m.GetNodeSet("right"), s["right"])
    m = mesh.create_3D_box_mesh(3, 4, 5)
    c, k, s = fe.box_hex_mesh(3, 4, 5)
    assert np.array_equal(m.GetElementsNodes("hexahedron"), k) and np.allclose(m.GetNodesCoordinates(), c)
    m = mesh.create_3D_tet_box_mesh(3)
    c, k, s = fe.box_tet_mesh(3)
    assert np.array_equal(m.GetElementsNodes("tetra"), k) and np.allclose(m.GetNodesCoordinates(), c)
    mj = mesh.create_3D_tet_box_mesh(4, jitter=0.2, seed=9)
    X = mj.GetNodesCoordinates()[mj.GetElementsNodes("tetra")]
    det = np.linalg.det(np.stack([X[:, 1] - X[:, 0], X[:, 2] - X[:, 0], X[:, 3] - X[:, 0]], axis=2))
    assert np.all(det > 0) and abs(det.sum() / 6 - 1.0) < 1e-12


def test_fourier_control_matches_oracle():
    from folax_amd import controls, mesh
    m = mesh.create_2D_square_mesh(1.0, 5)
    fc = controls.FourierControl("K", {"x_freqs": np.array([2, 4, 6]), "y_freqs": np.array([2, 4, 6]),
                                       "z_freqs": np.array([0]), "beta": 20, "min": 0.1, "max": 1.0}, m)
    fc.Initialize()
    v = np.random.default_rng(0).standard_normal(10)
    ref = fe.fourier_field(m.GetNodesCoordinates(), v[1:], [2, 4, 6], [2, 4, 6], [0], beta=20, vmin=0.1, vmax=1.0,
                           c0=v[0])
    np.testing.assert_allclose(fc.ComputeControlledVariables(v), ref, rtol=1e-14)


def test_loss_dirichlet_bookkeeping_matches_oracle():
    from folax_amd import loss_functions as lf, mesh
    m = mesh.create_2D_square_mesh(1.0, 4)
    bc = {"Ux": {"left": 0.0, "right": 0.05}, "Uy": {"left": 0.0, "right": 0.05}}
    loss = lf.MechanicalLoss2DQuad("m", {"dirichlet_bc_dict": bc, "material_dict": {"young_modulus": 1,
                                                                                   "poisson_ratio": 0.3}}, m)
    loss._create_dofs_dict()
    c, k, s = fe.square_quad_mesh(4)
    idx, val = fe.dirichlet_dofs(s, ["Ux", "Uy"], bc)
    assert np.array_equal(loss.dirichlet_indices, idx) and np.array_equal(loss.dirichlet_values, val)
    assert np.array_equal(loss.non_dirichlet_indices, fe.free_dofs(32, idx))
    with pytest.raises(ValueError):
        lf.ThermalLoss2DQuad("t", {}, m)


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus N > 1 exits with status 2 and says why, before touching a GPU."""
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr and r.stdout == ""
