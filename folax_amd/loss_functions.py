"""Drop-in FE loss objects over libfemhot (PyTorch-ROCm host side).

Same class names, settings keys and method semantics as folax's loss objects
(fol/loss_functions/fe_loss.py, thermal.py, mechanical.py, mechanical_neohooke.py); the element
loop, scatter and tangent assembly run in the HIP kernels of libfemhot.  Differentiable entry
points are `torch.autograd.Function`s whose backward calls the hand-written VJP kernels.

Arithmetic type follows the input tensors: float64 -> FP64 kernels, float32 -> FP32 kernels.
"""
from __future__ import annotations

from collections import namedtuple

import numpy as np
import torch

from . import _lib as L
from .context import FemhotContext


class CSRMatrix(namedtuple("CSRMatrix", "row_ptr col_idx values shape")):
    """Assembled tangent in CSR form (the pattern scipy builds from the reference COO)."""

    def to_torch(self):
        return torch.sparse_csr_tensor(self.row_ptr, self.col_idx.to(torch.int64), self.values, size=self.shape)

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_array((self.values.cpu().numpy(), self.col_idx.cpu().numpy(), self.row_ptr.cpu().numpy()),
                            shape=self.shape)

    def todense(self):
        return self.to_torch().to_dense()


def _as_device_tensor(x, dtype=None, device=None):
    device = device or torch.device("cuda")
    if isinstance(x, torch.Tensor):
        t = x
        if dtype is not None and t.dtype != dtype:
            t = t.to(dtype)
        if not t.is_cuda:
            t = t.to(device)
        return t.contiguous()
    return torch.as_tensor(np.asarray(x), dtype=dtype or torch.float64, device=device).contiguous()


class FiniteElementLoss:
    """Base FE loss (fe_loss.py:19-318)."""

    physics = None
    default_num_gp = {"quad": 2, "hexahedron": 2, "tetra": 1}

    def __init__(self, name: str, loss_settings: dict, fe_mesh):
        self.name = name
        self.loss_settings = dict(loss_settings)
        self.dofs = self.loss_settings["ordered_dofs"]
        self.element_type = self.loss_settings["element_type"]
        self.fe_mesh = fe_mesh
        self.initialized = False
        if "dirichlet_bc_dict" not in self.loss_settings:
            raise ValueError("dirichlet_bc_dict should provided in the loss settings !")

    def GetName(self):
        return self.name

    # ---- fe_loss.py:34-55 ----
    def _create_dofs_dict(self):
        ndpn = len(self.dofs)
        idx, val = [], []
        for di, dof in enumerate(self.dofs):
            for bname, bval in self.loss_settings["dirichlet_bc_dict"].get(dof, {}).items():
                nodes = np.asarray(self.fe_mesh.GetNodeSet(bname), dtype=np.int64)
                idx.append(ndpn * nodes + di)
                val.append(np.full(nodes.shape, float(bval)))
        self.dirichlet_indices = np.concatenate(idx) if idx else np.zeros(0, np.int64)
        self.dirichlet_values = np.concatenate(val) if val else np.zeros(0)
        self.non_dirichlet_indices = np.setdiff1d(np.arange(ndpn * self.fe_mesh.GetNumberOfNodes()),
                                                  self.dirichlet_indices)

    def _material(self):
        return {}

    def Initialize(self, reinitialize=False):
        """fe_loss.py:57-122: dof bookkeeping, Gauss rule, loss exponent; builds the device context."""
        if self.initialized and not reinitialize:
            return
        self.number_dofs_per_node = len(self.dofs)
        self.total_number_of_dofs = self.number_dofs_per_node * self.fe_mesh.GetNumberOfNodes()
        self._create_dofs_dict()
        self.number_of_unknown_dofs = self.non_dirichlet_indices.size
        num_gp = self.loss_settings.get("num_gp", self.default_num_gp[self.element_type])
        if num_gp not in (1, 2, 3):
            raise ValueError(f" number gauss points {num_gp} is not supported ! ")
        self.num_gp = num_gp
        self.loss_settings["num_gp"] = num_gp
        self.dim = self.loss_settings["compute_dims"]
        # fe_loss.py:91-103: with parametric_boundary_learning the Dirichlet dofs of a sample are
        # its own known_dofs (the batch parameters) instead of the fixed values, and the control
        # field comes from get_param_function (identity in the base class; a constant or mesh
        # point-data heterogeneity field when material_dict names one, as
        # mechanical_saint_venant.py:59-66 does)
        self.parametric_boundary_learning = bool(self.loss_settings.get("parametric_boundary_learning"))
        self.get_param_function = lambda x: x
        md = self.loss_settings.get("material_dict", {})
        if self.parametric_boundary_learning and "heterogeneity_field_name" in md:
            name = md["heterogeneity_field_name"]
            if not self.fe_mesh.HasPointData(name):
                self.fe_mesh[name] = md["heterogeneity_default_value"] * np.ones(self.fe_mesh.GetNumberOfNodes())
            self.heterogeneity_field = np.asarray(self.fe_mesh[name], dtype=np.float64)
            self.get_param_function = self._heterogeneity_params
        self.loss_function_exponent = float(self.loss_settings.get("loss_function_exponent", 1.0))
        self.control_location = self.loss_settings.get("control_location", "node")
        # "chunk": LDS-staged, atomic-free and deterministic (the library still takes the plain path
        # where that is faster: FP64 Neo-Hookean, or where femhot_autotune found it so);
        # "atomic": one lane per element with hardware atomics
        self.scatter = self.loss_settings.get("scatter", "chunk")
        if self.scatter not in ("chunk", "atomic"):
            raise ValueError(f"scatter must be 'chunk' or 'atomic', got {self.scatter!r}")
        # "autotune": time chunk configurations (and the plain path) the first time each call shape
        # (operation, dtype, batch) is seen, and keep the fastest (femhot_autotune)
        self.autotune = bool(self.loss_settings.get("autotune", False))
        self._tuned = set()
        self.ctx = FemhotContext(self.element_type, self.physics, self.fe_mesh.GetNodesCoordinates(),
                                 self.fe_mesh.GetElementsNodes(self.element_type), num_gp=num_gp,
                                 dirichlet_idx=self.dirichlet_indices, dirichlet_val=self.dirichlet_values,
                                 segment_offsets=self.loss_settings.get("segment_offsets"), **self._material())
        self._dir_idx_t = None
        self.initialized = True

    def _tune(self, op, params, dofs, w=None):
        """First call of a shape with autotune on: femhot_autotune picks the fastest configuration."""
        if not (self.autotune and self.scatter == "chunk"):
            return
        key = (op, dofs.dtype, dofs.shape[0] if dofs.dim() > 1 else 1)
        if key not in self._tuned:
            self.ctx.autotune(op, params, dofs, w, flags=self._flags() & ~L.SCATTER_CHUNK)
            self._tuned.add(key)

    def _flags(self):
        f = L.ELEMENT_FIELD if self.control_location == "element" else 0
        return f | (L.SCATTER_CHUNK if self.scatter == "chunk" else 0)

    def _dir_idx(self, device):
        if self._dir_idx_t is None or self._dir_idx_t.device != device:
            self._dir_idx_t = torch.as_tensor(self.dirichlet_indices, dtype=torch.int64, device=device)
            self._dir_val_t = torch.as_tensor(self.dirichlet_values, dtype=torch.float64, device=device)
        return self._dir_idx_t, self._dir_val_t

    def Finalize(self):
        pass

    def GetNumberOfUnknowns(self):
        return self.number_of_unknown_dofs

    def GetTotalNumberOfDOFs(self):
        return self.total_number_of_dofs

    def GetDOFs(self):
        return self.dofs

    # ---- fe_loss.py:91-100, 124-128, 186-189 ----
    def GetFullDofVector(self, known_dofs, unknown_dofs):
        """Dirichlet dofs overwritten with the fixed values, or, with parametric_boundary_learning,
        with known_dofs ((B, n_dirichlet) or (n_dirichlet,), in dirichlet_indices order)."""
        u = _as_device_tensor(unknown_dofs)
        idx, val = self._dir_idx(u.device)
        out = u.clone()
        if self.parametric_boundary_learning:
            k = _as_device_tensor(known_dofs, dtype=u.dtype, device=u.device)
            out[..., idx] = k.reshape(*out.shape[:-1], idx.numel()) if out.dim() > 1 else k.reshape(-1)
        else:
            out[..., idx] = val.to(u.dtype)
        return out

    def _heterogeneity_params(self, param_vector):
        p = torch.as_tensor(self.heterogeneity_field, device=param_vector.device, dtype=param_vector.dtype)
        return p.expand(param_vector.shape[0], -1) if param_vector.dim() > 1 else p

    def ApplyDirichletBCOnDofVector(self, full_dof_vector, load_increment: float = 1.0):
        u = _as_device_tensor(full_dof_vector)
        idx, val = self._dir_idx(u.device)
        out = u.clone()
        out[..., idx] = (load_increment * val).to(u.dtype)
        return out

    def GetParametersVectors(self, param_vector):
        return self.get_param_function(param_vector)

    # ---- energies ----
    def ComputeTotalEnergy(self, total_control_vars, total_primal_vars):
        """fe_loss.py:175-176: sum of the reference element energies (FP64 accumulation)."""
        u = _as_device_tensor(total_primal_vars)
        d = _as_device_tensor(total_control_vars, dtype=u.dtype)
        _, en = self.ctx.residual(d.reshape(1, -1), u.reshape(1, -1), want_R=False, flags=self._flags())
        return en.sum()

    def ComputeBatchLoss(self, batch_params, batch_dofs):
        """fe_loss.py:250-262: mean_b (sum_e E_e)^p with (min, max, mean); differentiable."""
        bd = _as_device_tensor(batch_dofs) if not isinstance(batch_dofs, torch.Tensor) else batch_dofs
        bp = batch_params if isinstance(batch_params, torch.Tensor) else _as_device_tensor(batch_params, dtype=bd.dtype)
        bp = torch.atleast_2d(bp).reshape(bp.shape[0] if bp.dim() > 1 else 1, -1)
        bd = torch.atleast_2d(bd).reshape(bd.shape[0] if bd.dim() > 1 else 1, -1)
        full = _FullDofs.apply(self, bp, bd)
        energies = _BatchEnergy.apply(self, torch.atleast_2d(self.GetParametersVectors(bp)), full)
        be = energies ** self.loss_function_exponent
        m = be.mean()
        return m, (be.min(), be.max(), m)

    def ComputeJacobianMatrixAndResidualVector(self, total_control_vars, total_primal_vars,
                                               transpose_jacobian: bool = False):
        """fe_loss.py:264-318 -> (CSRMatrix K, R), BC-masked as :191-207.  All element matrices of
        the supported physics are symmetric, so transpose_jacobian leaves them unchanged."""
        u = _as_device_tensor(total_primal_vars)
        d = _as_device_tensor(total_control_vars, dtype=u.dtype)
        # "chunk" leaves the assembly order to the library: the path femhot_autotune measured for the
        # tangent (with autotune on), else row-owner assembly (thermal, linear elastic) or colour-ordered
        # adds (Neo-Hookean), both deterministic; "atomic": hardware atomics
        flags = self._flags() | (L.TRANSPOSE if transpose_jacobian else 0)
        if self.scatter == "chunk" and self.autotune and (L.OP_TANGENT, u.dtype, 1) not in self._tuned:
            self.ctx.autotune_tangent(d, u, flags=flags & ~L.SCATTER_CHUNK)
            self._tuned.add((L.OP_TANGENT, u.dtype, 1))
        vals, R = self.ctx.tangent_csr(d, u, want_R=True, flags=flags)
        if not hasattr(self, "_pattern") or self._pattern[0].device != u.device:
            self._pattern = self.ctx.pattern(u.device)
        rp, ci = self._pattern
        n = self.total_number_of_dofs
        return CSRMatrix(rp, ci, vals, (n, n)), R

    # ---- hot-path extensions (same kernels, exposed for solvers, adjoints and training) ----
    def ComputeResidual(self, total_control_vars, total_primal_vars, mask_dirichlet=True):
        u = _as_device_tensor(total_primal_vars)
        d = _as_device_tensor(total_control_vars, dtype=u.dtype)
        f = self._flags() | (L.MASK_DIRICHLET if mask_dirichlet else 0)
        R, _ = self.ctx.residual(d.reshape(1, -1), u.reshape(1, -1), want_energy=False, flags=f)
        return R[0]

    def ComputeResidualVJP(self, total_control_vars, total_primal_vars, cotangent, picard=False):
        """(dR/du)^T w and (dR/dcontrol)^T w of the unmasked residual (fe_response.py:434-453)."""
        u = _as_device_tensor(total_primal_vars)
        d = _as_device_tensor(total_control_vars, dtype=u.dtype)
        w = _as_device_tensor(cotangent, dtype=u.dtype)
        f = self._flags() | (L.PICARD if picard else 0)
        gu, gd = self.ctx.residual_vjp(d.reshape(1, -1), u.reshape(1, -1), w.reshape(1, -1), flags=f)
        return gu[0], gd[0]

    def ComputeResidualShapeVJP(self, total_control_vars, total_primal_vars, cotangent):
        """(dR/dX)^T w, the adjoint shape-derivative term (fe_response.py:312-332, 365-406), flattened
        as (3 * n_nodes,) like the reference's nodal shape derivative vector."""
        u = _as_device_tensor(total_primal_vars)
        d = _as_device_tensor(total_control_vars, dtype=u.dtype)
        w = _as_device_tensor(cotangent, dtype=u.dtype)
        f = L.ELEMENT_FIELD if self.control_location == "element" else 0
        return self.ctx.residual_shape_vjp(d, u, w, flags=f).reshape(-1)

    def ComputePotentialAndGradient(self, total_control_vars, total_primal_vars):
        """Potential Pi with dPi/du (= residual) and dPi/dcontrol."""
        u = _as_device_tensor(total_primal_vars)
        d = _as_device_tensor(total_control_vars, dtype=u.dtype)
        en, gu, gd = self.ctx.energy_grad(d.reshape(1, -1), u.reshape(1, -1), flags=self._flags())
        return en.sum(), gu[0], gd[0]

    def ComputeBatchResidualNormLoss(self, batch_params, batch_dofs, picard=False):
        """L = 1/2 sum_b ||R_free(u_b, k_b)||^2 with Dirichlet dofs overwritten first; differentiable
        with respect to both arguments through the residual VJP kernel (exact unless picard)."""
        bd = batch_dofs if isinstance(batch_dofs, torch.Tensor) else _as_device_tensor(batch_dofs)
        bp = batch_params if isinstance(batch_params, torch.Tensor) else _as_device_tensor(batch_params, dtype=bd.dtype)
        bp2 = torch.atleast_2d(bp)
        full = _FullDofs.apply(self, bp2, torch.atleast_2d(bd))
        return _ResidualNormLoss.apply(self, torch.atleast_2d(self.GetParametersVectors(bp2)), full, picard)


class _FullDofs(torch.autograd.Function):
    """GetFullDofVector: Dirichlet entries overwritten, so their gradient with respect to the
    incoming dofs is zero; with parametric_boundary_learning they are the known dofs, which receive
    the gradient of the entries they set (a gather of g at dirichlet_indices, as the transpose of
    the reference's .at[:, dirichlet_indices].set(known_dofs))."""

    @staticmethod
    def forward(ctx, loss, known, dofs):
        ctx.loss = loss
        return loss.GetFullDofVector(known, dofs)

    @staticmethod
    def backward(ctx, g):
        loss = ctx.loss
        idx, _ = loss._dir_idx(g.device)
        gk = g[..., idx].clone() if loss.parametric_boundary_learning else None
        g = g.clone()
        g[..., idx] = 0
        return None, gk, g


class _BatchEnergy(torch.autograd.Function):
    """Per-sample energies of ComputeBatchLoss.  Backward follows jax.grad of the reference loss.
    * Linear elastic: E = u^T stop_gradient(Se u - Fe) (mechanical.py:116-117), so dE/du = r and
      dE/dcontrol = 0.
    * Thermal: E = T^T (Se(k0, sg(T)) sg(T) - Fe) (thermal.py:31, :44-49).  Only T is under
      stop_gradient, not Se, so dE/dT = r and dE/dk0 = T^T (dSe/dk0) T, which is the residual VJP's
      field part with the cotangent w = T.
    * Neo-Hookean: Ee has no stop_gradient, so dE/du = Fint and dE/dcontrol = dEe/dcontrol
      (energy-gradient kernel)."""

    @staticmethod
    def forward(ctx, loss, params, dofs):
        params = params.to(dofs.dtype).contiguous()
        dofs = dofs.contiguous()
        loss._tune(L.OP_RESIDUAL, params, dofs)
        _, en = loss.ctx.residual(params, dofs, want_R=False, flags=loss._flags())
        ctx.loss = loss
        ctx.save_for_backward(params, dofs)
        return en.sum(dim=1).to(dofs.dtype)

    @staticmethod
    def backward(ctx, gE):
        loss = ctx.loss
        params, dofs = ctx.saved_tensors
        flags = loss._flags()
        if loss.physics == L.NEOHOOKE:
            loss._tune(L.OP_ENERGY_GRAD, params, dofs)
            _, gu, gd = loss.ctx.energy_grad(params, dofs, want_energy=False, flags=flags)
            if np.any(np.asarray(loss.loss_settings.get("body_foce", [0.0]), dtype=np.float64) != 0.0):
                # Ee excludes the body-force work: add Fe = -r(u = 0) back
                r0, _ = loss.ctx.residual(params, torch.zeros_like(dofs), want_energy=False, flags=flags)
                gu = gu - r0
        else:
            gu, _ = loss.ctx.residual(params, dofs, want_energy=False, flags=flags)
            if loss.physics == L.THERMAL:
                _, gd = loss.ctx.residual_vjp(params, dofs, dofs, want_gu=False, flags=flags)
            else:
                gd = torch.zeros_like(params)
        s = gE.to(dofs.dtype).reshape(-1, 1)
        return None, gd * s, gu * s


class _ResidualNormLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, loss, params, dofs, picard):
        params = params.to(dofs.dtype).contiguous()
        dofs = dofs.contiguous()
        flags = loss._flags()
        loss._tune(L.OP_RESIDUAL, params, dofs)
        R, _ = loss.ctx.residual(params, dofs, want_energy=False, flags=flags | L.MASK_DIRICHLET)
        ctx.loss, ctx.picard = loss, picard
        ctx.save_for_backward(params, dofs, R)
        return 0.5 * (R.double() ** 2).sum().to(dofs.dtype)

    @staticmethod
    def backward(ctx, g):
        loss = ctx.loss
        params, dofs, R = ctx.saved_tensors
        flags = loss._flags() | (L.PICARD if ctx.picard else 0)
        loss._tune(L.OP_VJP, params, dofs, R)
        gu, gd = loss.ctx.residual_vjp(params, dofs, R, flags=flags | L.MASK_DIRICHLET)
        s = g.to(dofs.dtype)
        return None, gd * s, gu * s, None


# --------------------------------------------------------------------------------------------
# physics classes (same names and defaults as the reference)
# --------------------------------------------------------------------------------------------
class ThermalLoss(FiniteElementLoss):
    """thermal.py:15-49: k = (N.k0)(1 + beta (N.T)^c); settings beta (default 0), c (default 1)."""

    physics = L.THERMAL

    def _material(self):
        self.thermal_loss_settings = {"beta": self.loss_settings.get("beta", 0), "c": self.loss_settings.get("c", 1)}
        bf = self.loss_settings.get("body_force", 0.0)
        return {"beta": float(self.thermal_loss_settings["beta"]), "c": float(self.thermal_loss_settings["c"]),
                "body_force": (float(bf), 0.0, 0.0)}


class ThermalLoss2DQuad(ThermalLoss):
    def __init__(self, name, loss_settings, fe_mesh):
        super().__init__(name, {**loss_settings, "compute_dims": 2, "ordered_dofs": ["T"], "element_type": "quad"},
                         fe_mesh)


class ThermalLoss3DHexa(ThermalLoss):
    def __init__(self, name, loss_settings, fe_mesh):
        super().__init__(name, {**loss_settings, "compute_dims": 3, "ordered_dofs": ["T"],
                                "element_type": "hexahedron"}, fe_mesh)


class ThermalLoss3DTetra(ThermalLoss):
    def __init__(self, name, loss_settings, fe_mesh):
        super().__init__(name, {**loss_settings, "compute_dims": 3, "ordered_dofs": ["T"], "element_type": "tetra"},
                         fe_mesh)


class MechanicalLoss(FiniteElementLoss):
    """mechanical.py:14-117: small-strain isotropic elasticity, D from material_dict scaled by the
    nodal control field; 2-D is plane stress.  body force key `body_foce` as in the reference."""

    physics = L.MECHANICAL

    def _material(self):
        if "material_dict" not in self.loss_settings:
            raise ValueError("material_dict should provided in the loss settings !")
        md = self.loss_settings["material_dict"]
        bf = np.zeros(3)
        if "body_foce" in self.loss_settings:
            b = np.asarray(self.loss_settings["body_foce"], dtype=np.float64).reshape(-1)
            bf[: b.size] = b
        return {"young_modulus": float(md["young_modulus"]), "poisson_ratio": float(md["poisson_ratio"]),
                "body_force": tuple(bf)}


class MechanicalLoss2DQuad(MechanicalLoss):
    def __init__(self, name, loss_settings, fe_mesh):
        super().__init__(name, {**loss_settings, "compute_dims": 2, "ordered_dofs": ["Ux", "Uy"],
                                "element_type": "quad"}, fe_mesh)


class MechanicalLoss3DHexa(MechanicalLoss):
    def __init__(self, name, loss_settings, fe_mesh):
        super().__init__(name, {**loss_settings, "compute_dims": 3, "ordered_dofs": ["Ux", "Uy", "Uz"],
                                "element_type": "hexahedron"}, fe_mesh)


class MechanicalLoss3DTetra(MechanicalLoss):
    def __init__(self, name, loss_settings, fe_mesh):
        super().__init__(name, {**loss_settings, "compute_dims": 3, "ordered_dofs": ["Ux", "Uy", "Uz"],
                                "element_type": "tetra"}, fe_mesh)


class NeoHookeMechanicalLoss(MechanicalLoss):
    """mechanical_neohooke.py:16-275: compressible Neo-Hookean (hand-derived tangent); kappa and mu
    from the control field and poisson_ratio only (young_modulus is not used)."""

    physics = L.NEOHOOKE


class NeoHookeMechanicalLoss2DQuad(NeoHookeMechanicalLoss):
    def __init__(self, name, loss_settings, fe_mesh):
        super().__init__(name, {**loss_settings, "compute_dims": 2, "ordered_dofs": ["Ux", "Uy"],
                                "element_type": "quad"}, fe_mesh)


class NeoHookeMechanicalLoss3DHexa(NeoHookeMechanicalLoss):
    def __init__(self, name, loss_settings, fe_mesh):
        super().__init__(name, {**loss_settings, "compute_dims": 3, "ordered_dofs": ["Ux", "Uy", "Uz"],
                                "element_type": "hexahedron"}, fe_mesh)


class NeoHookeMechanicalLoss3DTetra(NeoHookeMechanicalLoss):
    def __init__(self, name, loss_settings, fe_mesh):
        super().__init__(name, {**loss_settings, "compute_dims": 3, "ordered_dofs": ["Ux", "Uy", "Uz"],
                                "element_type": "tetra"}, fe_mesh)
