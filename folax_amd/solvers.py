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
                                 maxiter=int(self.linear_solver_settings["maxiter"]), mode=self.mode)
        self.last_iterations, self.last_relative_residual = it, rel
        return x.to(residual_vector.dtype)

    def Finalize(self):
        pass


class FiniteElementLinearResidualBasedSolver(FiniteElementSolver):
    def Solve(self, current_control_vars, current_dofs):
        u = self.fe_loss_function.ApplyDirichletBCOnDofVector(_as_device_tensor(current_dofs))
        K, R = self.fe_loss_function.ComputeJacobianMatrixAndResidualVector(current_control_vars, u)
        return u + self.LinearSolve(K, R, u)


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
        u = _as_device_tensor(current_dofs_np)
        s = self.nonlinear_solver_settings
        self.convergence_history = {}
        for step in range(1, s["load_incr"] + 1):
            u = self.fe_loss_function.ApplyDirichletBCOnDofVector(u, step / s["load_incr"])
            hist = {"res_norm": [], "delta_dofs_norm": []}
            self.convergence_history[step] = hist
            for i in range(1, s["maxiter"] + 1):
                K, R = self.fe_loss_function.ComputeJacobianMatrixAndResidualVector(current_control_vars, u)
                res_norm = float(torch.linalg.norm(R.double()))
                if res_norm != res_norm:
                    raise ValueError("Residual norm contains NaN values.")
                du = self.LinearSolve(K, R, u)
                dnorm = float(torch.linalg.norm(du.double()))
                hist["res_norm"].append(res_norm)
                hist["delta_dofs_norm"].append(dnorm)
                if res_norm < s["abs_tol"] or dnorm < s["rel_tol"] or i == s["maxiter"]:
                    break
                u = u + du
        return u
