"""FE solvers over libfemhot: assembled tangent (CSR) + Jacobi-PCG on the device.

FiniteElementLinearResidualBasedSolver follows fol/solvers/fe_linear_residual_based_solver.py:15-22
(u <- BC(u); (K, R) <- assemble; u + solve(K, -R)); the nonlinear solver follows
fol/solvers/fe_nonlinear_residual_based_solver.py:76-170 (load increments x Newton iterations,
NaN guard raising ValueError).  The linear solve is Jacobi-PCG ("PCG", FP64) or FP32 inner
iterations with FP64 iterative refinement ("PCG-mixed"); the reference's "JAX-bicgstab" and
"JAX-direct" names select "PCG".
"""
from __future__ import annotations

import torch

from . import _lib as L
from .loss_functions import FiniteElementLoss, _as_device_tensor


def _update(default, given):
    """UpdateDefaultDict (fol/tools/usefull_functions.py:569-573): unknown keys are dropped."""
    out = dict(default)
    out.update({k: given[k] for k in default if k in given})
    return out


class FiniteElementSolver:
    def __init__(self, fe_solver_name: str, fe_loss_function: FiniteElementLoss, fe_solver_settings: dict = None):
        self.name = fe_solver_name
        self.fe_loss_function = fe_loss_function
        self.fe_solver_settings = fe_solver_settings or {}
        self.linear_solver_settings = {"solver": "PCG", "tol": 1e-10, "atol": 0.0, "maxiter": 20000}
        self.last_iterations = 0
        self.last_relative_residual = 0.0

    def Initialize(self):
        if "linear_solver_settings" in self.fe_solver_settings:
            self.linear_solver_settings = _update(self.linear_solver_settings,
                                                  self.fe_solver_settings["linear_solver_settings"])
        s = self.linear_solver_settings["solver"]
        if s in ("PCG", "JAX-bicgstab", "JAX-direct"):
            self.mode = "f64"
        elif s == "PCG-mixed":
            self.mode = "mixed"
        else:
            raise ValueError(f"linear solver {s} does exist, available options are ['PCG', 'PCG-mixed']")

    def LinearSolve(self, tangent_matrix, residual_vector, dofs_vector=None):
        """Solve K du = -R (fe_solver.py:60-78); returns du."""
        ctx = self.fe_loss_function.ctx
        b = (-residual_vector).to(torch.float64).contiguous()
        x, it, rel, ok = ctx.pcg(tangent_matrix.row_ptr, tangent_matrix.col_idx,
                                 tangent_matrix.values.to(torch.float64).contiguous(), b,
                                 tol=float(self.linear_solver_settings["tol"]),
                                 maxiter=int(self.linear_solver_settings["maxiter"]), mode=self.mode,
                                 context_pattern=True)
        self.last_iterations, self.last_relative_residual = it, rel
        return x.to(residual_vector.dtype)

    def Finalize(self):
        pass


class FiniteElementLinearResidualBasedSolver(FiniteElementSolver):
    def Solve(self, current_control_vars, current_dofs):
        u = self.fe_loss_function.ApplyDirichletBCOnDofVector(_as_device_tensor(current_dofs))
        K, R = self.fe_loss_function.ComputeJacobianMatrixAndResidualVector(current_control_vars, u)
        return u + self.LinearSolve(K, R, u)


class AdjointFiniteElementSolver(FiniteElementSolver):
    """fol/solvers/adjoint_fe_solver.py:11-23: solves K^T lambda = -df/du with the response's
    BC-applied transposed tangent (device CSR assembly + PCG)."""

    def __init__(self, adj_fe_solver_name, fe_response, adj_fe_solver_settings=None):
        super().__init__(adj_fe_solver_name, fe_response.fe_loss, adj_fe_solver_settings or {})
        self.fe_response = fe_response

    def Solve(self, current_control_vars, current_dofs, current_adjoint_dofs=None):
        jac, rhs = self.fe_response.ComputeAdjointJacobianMatrixAndRHSVector(current_control_vars, current_dofs)
        # LinearSolve solves K x = -r, and rhs already carries the minus sign of -df/du
        return self.LinearSolve(jac, -rhs, current_adjoint_dofs)


class FiniteElementNonLinearResidualBasedSolver(FiniteElementLinearResidualBasedSolver):
    def __init__(self, fe_solver_name, fe_loss_function, fe_solver_settings=None, history_plot_settings=None):
        super().__init__(fe_solver_name, fe_loss_function, fe_solver_settings)
        self.nonlinear_solver_settings = {"rel_tol": 1e-8, "abs_tol": 1e-8, "maxiter": 20, "load_incr": 5}

    def Initialize(self):
        super().Initialize()
        if "nonlinear_solver_settings" in self.fe_solver_settings:
            self.nonlinear_solver_settings = _update(self.nonlinear_solver_settings,
                                                     self.fe_solver_settings["nonlinear_solver_settings"])

    def Solve(self, current_control_vars, current_dofs_np):
        """The load-stepped Newton loop runs inside the library (femhot_newton): assembly, norms, the
        NaN guard, the linear solve and the update stay on the device."""
        loss = self.fe_loss_function
        u = _as_device_tensor(current_dofs_np).to(torch.float64)
        d = _as_device_tensor(current_control_vars, dtype=torch.float64)
        s = self.nonlinear_solver_settings
        # the tangent path of ComputeJacobianMatrixAndResidualVector: the library's tuned choice under
        # "chunk" (colour-ordered untuned), hardware atomics under "atomic"
        flags = loss._flags() & (L.ELEMENT_FIELD | L.SCATTER_CHUNK)
        try:
            u, hist, steps = loss.ctx.newton(d, u, load_steps=s["load_incr"], max_iter=s["maxiter"],
                                             abs_tol=s["abs_tol"], rel_tol=s["rel_tol"],
                                             lin_tol=float(self.linear_solver_settings["tol"]),
                                             lin_maxiter=int(self.linear_solver_settings["maxiter"]),
                                             mode=self.mode, flags=flags)
        except L.FemhotError as e:
            if e.code == L.ERR_NONFINITE:
                raise ValueError("Residual norm contains NaN values.") from e
            raise
        self.convergence_history = {
            step + 1: {"res_norm": [float(v) for v in hist[step, :steps[step], 0]],
                       "delta_dofs_norm": [float(v) for v in hist[step, :steps[step], 1]]}
            for step in range(s["load_incr"])}
        return u.to(_as_device_tensor(current_dofs_np).dtype)
