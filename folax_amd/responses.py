"""Finite-element responses and their adjoint sensitivities over libfemhot (PyTorch-ROCm host side).

FiniteElementResponse follows fol/responses/fe_response.py:17-719 with the same method names and
meaning.  A response is f = sum_e sum_g w detJ phi(N.d, N.u) for a user formula phi written in
terms of the control's name and the first letter of the first dof ("(E**2)*U[0]" reads E = N.d
and U[c] = N.u_c at each Gauss point, fe_response.py:61-66, 92-113).

What runs where:
* the Gauss-point geometry (w detJ, N, dN/dx) comes from the library (femhot_gauss_point_data);
* phi is evaluated on the device with torch on all Gauss points at once, and its derivatives with
  respect to the Gauss-point field and state come from torch autograd of that evaluation (the
  formula is arbitrary Python, as the reference's jax.grad of its eval'd lambda);
* the adjoint terms lambda^T dR/dk and lambda^T dR/dX run in the hand-written residual VJP and
  shape-VJP kernels, the transposed tangent in the CSR assembly kernel, and the adjoint solve in
  the device PCG.
"""
from __future__ import annotations

import torch

from .loss_functions import FiniteElementLoss, _as_device_tensor


class _ArrayNamespace:
    """The `jnp` a response formula may call (fe_response.py:64 evaluates it with jnp in scope):
    numpy-named functions resolved to torch's elementwise ones."""

    _alias = {"power": torch.pow, "arctan": torch.atan, "arcsin": torch.asin, "arccos": torch.acos,
              "absolute": torch.abs}

    def __getattr__(self, name):
        if name in self._alias:
            return self._alias[name]
        return getattr(torch, name)


class FiniteElementResponse:
    """fe_response.py:17-45: a named response formula over an FE loss and a control."""

    def __init__(self, name: str, response_formula: str, fe_loss: FiniteElementLoss, control):
        self.name = name
        self.response_formula = response_formula
        self.fe_loss = fe_loss
        self.control = control
        self.initialized = False

    def GetName(self):
        return self.name

    def Initialize(self, reinitialize=False):
        """fe_response.py:47-69: the formula becomes lambda <control name>, <first dof letter>: formula."""
        if self.initialized and not reinitialize:
            return
        self.fe_loss.Initialize()
        self.control.Initialize()
        variables = [self.control.GetName(), self.fe_loss.dofs[0][0]]
        self.response_function = eval(f"lambda {', '.join(variables)}: {self.response_formula}",
                                      {"jnp": _ArrayNamespace()})
        self._gp = {}
        self._conn = None
        self.initialized = True

    # ---- Gauss-point integrand ----
    def _geometry(self, dtype, device):
        key = (dtype, device)
        if key not in self._gp:
            self._gp[key] = self.fe_loss.ctx.gauss_point_data(self.fe_loss.num_gp, dtype=dtype, device=device)
        if self._conn is None or self._conn.device != device:
            self._conn = torch.as_tensor(self.fe_loss.fe_mesh.GetElementsNodes(self.fe_loss.element_type),
                                         dtype=torch.int64, device=device)
        return self._gp[key]

    def _inputs(self, control, dofs):
        u = _as_device_tensor(dofs)
        d = _as_device_tensor(control, dtype=u.dtype, device=u.device)
        return d.reshape(-1), u.reshape(-1)

    def _element_values(self, de, ue, N, wdet):
        """w detJ phi per element and Gauss point: de (Ne, n), ue (Ne, n, ndpn)."""
        gp_d = de @ N.T                                  # (Ne, G)
        gp_u = torch.einsum("ga,eac->ceg", N, ue)        # (ndpn, Ne, G): U[c] is component c
        return wdet * self.response_function(gp_d, gp_u)

    def _gather(self, d, u):
        conn = self._conn
        ndpn = self.fe_loss.number_dofs_per_node
        return d[conn], u.view(-1, ndpn)[conn]

    def _scatter_nodal(self, vals, width):
        """Assembles per-element nodal values (Ne, n, width) into (Nn * width,) (fe_response.py:271-272)."""
        conn = self._conn
        out = torch.zeros(self.fe_loss.fe_mesh.GetNumberOfNodes(), width, dtype=vals.dtype, device=vals.device)
        out.index_add_(0, conn.reshape(-1), vals.reshape(-1, width))
        return out.reshape(-1)

    # ---- reference API ----
    def ComputeValue(self, nodal_control_values, nodal_dof_values):
        """fe_response.py:196-212: sum over elements and Gauss points."""
        d, u = self._inputs(nodal_control_values, nodal_dof_values)
        N, wdet, _ = self._geometry(u.dtype, u.device)
        de, ue = self._gather(d, u)
        return self._element_values(de, ue, N, wdet).sum()

    def _state_and_control_grads(self, d, u):
        """d f / d u (Ndof,) and d f / d control (Nn,) of the response integrand alone."""
        N, wdet, _ = self._geometry(u.dtype, u.device)
        de, ue = self._gather(d, u)
        de = de.detach().requires_grad_(True)
        ue = ue.detach().requires_grad_(True)
        with torch.enable_grad():
            f = self._element_values(de, ue, N, wdet).sum()
            gd, gu = torch.autograd.grad(f, (de, ue))
        ndpn = self.fe_loss.number_dofs_per_node
        return self._scatter_nodal(gu, ndpn), self._scatter_nodal(gd.unsqueeze(-1), 1)

    def ComputeAdjointJacobianMatrixAndRHSVector(self, nodal_control_values, nodal_dof_values):
        """fe_response.py:244-283: rhs = -(df/du) with Dirichlet entries zeroed, and the BC-applied
        transposed tangent of the loss."""
        d, u = self._inputs(nodal_control_values, nodal_dof_values)
        dfdu, _ = self._state_and_control_grads(d, u)
        idx, _ = self.fe_loss._dir_idx(u.device)
        rhs = dfdu.clone()
        rhs[idx] = 0.0
        rhs = -rhs
        K, _ = self.fe_loss.ComputeJacobianMatrixAndResidualVector(d, u, transpose_jacobian=True)
        return K, rhs

    def ComputeAdjointNodalControlDerivatives(self, nodal_control_values, nodal_dof_values, nodal_adj_dof_values):
        """fe_response.py:486-526: df/dk + lambda^T dR/dk (the residual VJP kernel)."""
        d, u = self._inputs(nodal_control_values, nodal_dof_values)
        lam = _as_device_tensor(nodal_adj_dof_values, dtype=u.dtype, device=u.device).reshape(-1)
        _, dfdk = self._state_and_control_grads(d, u)
        _, gk = self.fe_loss.ComputeResidualVJP(d, u, lam)
        return dfdk + gk

    def ComputeAdjointNodalShapeDerivatives(self, nodal_control_values, nodal_dof_values, nodal_adj_dof_values):
        """fe_response.py:365-406: df/dX + lambda^T dR/dX, flattened (3 n_nodes,).  Only detJ of the
        integrand moves with the nodes: d(w detJ)/dX_a,i = w detJ dN_a/dx_i at each Gauss point; the
        residual term runs in the shape-VJP kernel."""
        d, u = self._inputs(nodal_control_values, nodal_dof_values)
        lam = _as_device_tensor(nodal_adj_dof_values, dtype=u.dtype, device=u.device).reshape(-1)
        N, wdet, dndx = self._geometry(u.dtype, u.device)
        de, ue = self._gather(d, u)
        wphi = self._element_values(de, ue, N, wdet)                  # (Ne, G)
        ge = torch.einsum("eg,egaj->eaj", wphi, dndx)                # (Ne, n, dim)
        if ge.shape[-1] < 3:
            ge = torch.nn.functional.pad(ge, (0, 3 - ge.shape[-1]))
        dfdX = self._scatter_nodal(ge, 3)
        return dfdX + self.fe_loss.ComputeResidualShapeVJP(d, u, lam)

    def Finalize(self):
        pass
