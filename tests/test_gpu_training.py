"""Trainer integration: the operator-learning step of folax (explicit_parametric_operator_learning.py:99-117)
in torch over the drop-in loss, trained like tests/integration/test_mechanical_2D.py:17-86 and held to its
bar: after 200 epochs the network's displacement field is within 10 % (L2) of the finite-element solution."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def test_mechanical_2D_operator_learning_matches_fem():
    from folax_amd import loss_functions as lf
    from folax_amd.controls import FourierControl
    from folax_amd.mesh import create_2D_square_mesh
    from folax_amd.solvers import FiniteElementLinearResidualBasedSolver
    dev = torch.device("cuda")
    m = create_2D_square_mesh(L=1, N=11)
    bc = {"Ux": {"left": 0.0, "right": 0.1}, "Uy": {"left": 0.0, "right": 0.1}}
    loss = lf.MechanicalLoss2DQuad("mechanical_loss_2d", {"dirichlet_bc_dict": bc, "num_gp": 2,
                                                          "material_dict": {"young_modulus": 1, "poisson_ratio": 0.3}}, m)
    loss.Initialize()
    ctrl = FourierControl("fourier_control", {"x_freqs": np.array([1, 2, 3]), "y_freqs": np.array([1, 2, 3]),
                                              "z_freqs": np.array([0]), "beta": 2}, m)
    ctrl.Initialize()
    # create_random_fourier_samples(control, 0) (usefull_functions.py:260-277): the uniform sample c0 = 10
    coeffs = np.zeros((1, ctrl.GetNumberOfVariables()))
    coeffs[0, 0] = 10.0
    K = torch.as_tensor(ctrl.ComputeBatchControlledVariables(coeffs), dtype=torch.float64, device=dev)
    X = torch.as_tensor(coeffs, dtype=torch.float64, device=dev)
    # MLP(in, 1, n_unknowns) with zero-initialised kernels and biases and swish (test_mechanical_2D.py:44-61)
    net = torch.nn.Sequential(torch.nn.Linear(X.shape[1], 1), torch.nn.SiLU(),
                              torch.nn.Linear(1, loss.GetNumberOfUnknowns())).to(dev, torch.float64)
    for p in net.parameters():
        torch.nn.init.zeros_(p)
    opt = torch.optim.Adam(net.parameters(), lr=1e-3)
    free = torch.as_tensor(loss.non_dirichlet_indices, device=dev)

    def full_prediction():
        full = torch.zeros(1, loss.GetTotalNumberOfDOFs(), dtype=torch.float64, device=dev)
        return full.index_copy(1, free, net(X))

    for _ in range(200):
        opt.zero_grad(set_to_none=True)
        L, _ = loss.ComputeBatchLoss(K, full_prediction())
        L.backward()
        # optax.normalize_by_update_norm before adam: the update direction has unit global norm
        gn = torch.sqrt(sum((p.grad ** 2).sum() for p in net.parameters()))
        for p in net.parameters():
            p.grad /= gn
        opt.step()
    with torch.no_grad():
        u_fol = loss.GetFullDofVector(None, full_prediction())[0]
    s = FiniteElementLinearResidualBasedSolver("fe", loss, {"linear_solver_settings": {"solver": "PCG", "tol": 1e-12}})
    s.Initialize()
    u_fem = s.Solve(K[0], torch.zeros_like(u_fol))
    l2 = 100 * float(torch.linalg.norm(u_fol - u_fem) / torch.linalg.norm(u_fem))
    print(f"[training] L2 error {l2:.3f} %", flush=True)
    assert l2 <= 10.0
