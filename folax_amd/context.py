"""Thin, typed wrapper of one libfemhot context (one mesh + loss settings) over torch device memory.

All buffers are torch tensors on the current CUDA (HIP) device; calls are enqueued on torch's
current stream.  Inputs are validated on the host (device, dtype, shape, contiguity) before any
launch; the library validates sizes again and reports errors through `FemhotError`.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib as L


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else C.c_void_p(0)


def _dtype_code(t):
    if t.dtype == torch.float32:
        return L.F32
    if t.dtype == torch.float64:
        return L.F64
    raise TypeError(f"femhot: unsupported dtype {t.dtype} (float32 or float64)")


class FemhotContext:
    """Owns a `femhot_ctx` built from host mesh arrays (fe_loss.py:57-122 Initialize)."""

    def __init__(self, elem_type: str, physics: int, coords, conn, num_gp=0, young_modulus=1.0, poisson_ratio=0.3,
                 beta=0.0, c=1.0, body_force=(0.0, 0.0, 0.0), dirichlet_idx=None, dirichlet_val=None,
                 segment_offsets=None):
        if not torch.cuda.is_available():
            raise RuntimeError("femhot requires a ROCm GPU (torch.cuda.is_available() is False)")
        lib = L.lib()
        self._lib = lib
        coords = np.ascontiguousarray(coords, dtype=np.float64)
        if coords.ndim != 2 or coords.shape[1] not in (2, 3):
            raise ValueError("coords must be (n_nodes, 2|3)")
        if coords.shape[1] == 2:
            coords = np.concatenate([coords, np.zeros((coords.shape[0], 1))], axis=1)
        conn = np.ascontiguousarray(conn, dtype=np.int32)
        self.elem_type = elem_type
        self.et = L.ELEMENT_TYPES[elem_type]
        self.physics = physics
        self.n_nodes = coords.shape[0]
        self.n_elems = conn.shape[0]
        didx = np.ascontiguousarray(np.zeros(0) if dirichlet_idx is None else dirichlet_idx, dtype=np.int64)
        dval = np.ascontiguousarray(np.zeros(0) if dirichlet_val is None else dirichlet_val, dtype=np.float64)
        if didx.shape != dval.shape:
            raise ValueError("dirichlet_idx and dirichlet_val must have the same length")
        self._keep = (coords, conn, didx, dval)
        d = L.MeshDesc()
        d.elem_type, d.physics, d.num_gp = self.et, int(physics), int(num_gp)
        d.young_modulus, d.poisson_ratio, d.beta, d.c = float(young_modulus), float(poisson_ratio), float(beta), float(c)
        bf = list(np.asarray(body_force, dtype=np.float64).reshape(-1)) + [0.0, 0.0, 0.0]
        d.body_force = (C.c_double * 3)(*bf[:3])
        d.n_nodes, d.n_elems = self.n_nodes, self.n_elems
        d.coords = coords.ctypes.data_as(C.POINTER(C.c_double))
        d.conn = conn.ctypes.data_as(C.POINTER(C.c_int))
        d.n_dirichlet = didx.shape[0]
        d.dirichlet_idx = didx.ctypes.data_as(C.POINTER(C.c_longlong))
        d.dirichlet_val = dval.ctypes.data_as(C.POINTER(C.c_double))
        if segment_offsets is not None:
            so = np.ascontiguousarray(segment_offsets, dtype=np.int64)
            self._keep = self._keep + (so,)
            d.n_segments = so.shape[0] - 1
            d.segment_offsets = so.ctypes.data_as(C.POINTER(C.c_longlong))
        else:
            d.n_segments = 1
        h = C.c_void_p()
        L.check(lib.femhot_ctx_create(C.byref(d), C.byref(h)))
        self._h = h
        n_dofs, npe, dpn = C.c_longlong(), C.c_int(), C.c_int()
        nseg = C.c_int()
        L.check(lib.femhot_ctx_info(h, C.byref(n_dofs), C.byref(npe), C.byref(dpn), None, None, C.byref(nseg)))
        self.n_dofs, self.nodes_per_elem, self.dofs_per_node = n_dofs.value, npe.value, dpn.value
        self.n_segments = nseg.value

    def close(self):
        if getattr(self, "_h", None):
            self._lib.femhot_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------- structure ----------------
    def nnz(self):
        v = C.c_longlong()
        L.check(self._lib.femhot_ctx_info(self._h, None, None, None, C.byref(v), None, None))
        return v.value

    def pattern(self, device=None):
        """CSR pattern (row_ptr int64 (n_dofs+1), col_idx int32 (nnz)) on the device."""
        device = device or torch.device("cuda")
        nnz = self.nnz()
        rp = torch.empty(self.n_dofs + 1, dtype=torch.int64, device=device)
        ci = torch.empty(max(nnz, 1), dtype=torch.int32, device=device)
        L.check(self._lib.femhot_pattern(self._h, _ptr(rp), _ptr(ci), _stream()))
        return rp, ci[:nnz]

    def colouring(self):
        out = np.empty(self.n_elems, dtype=np.int32)
        L.check(self._lib.femhot_colouring(self._h, out.ctypes.data_as(C.POINTER(C.c_int))))
        return out

    # ---------------- checks ----------------
    def _check_in(self, t, name, last, batch, dtype, device):
        if t is None:
            raise ValueError(f"{name} is required")
        if not t.is_cuda:
            raise ValueError(f"{name} must be a device tensor")
        if t.dtype != dtype:
            raise TypeError(f"{name} has dtype {t.dtype}, expected {dtype}")
        if t.device != device:
            raise ValueError(f"{name} is on {t.device}, expected {device}")
        if t.dim() != 2 or t.shape[0] != batch or t.shape[1] != last:
            raise ValueError(f"{name} must have shape ({batch}, {last}), got {tuple(t.shape)}")
        if t.stride(1) != 1:
            raise ValueError(f"{name} must be contiguous in its last dimension")
        return t.stride(0) if batch > 1 else last

    def _field_size(self, flags):
        return self.n_elems if flags & L.ELEMENT_FIELD else self.n_nodes

    # ---------------- compute ----------------
    def residual(self, field, u, want_R=True, want_energy=True, flags=0, out_R=None):
        """R (B, n_dofs) and per-sample (per-segment) energies (float64).  out_R: a preallocated (B, n_dofs)
        residual to write (graph-capture friendly)."""
        u = u if u.dim() == 2 else u.unsqueeze(0)
        field = field if field.dim() == 2 else field.unsqueeze(0)
        B = u.shape[0]
        dt, dev = u.dtype, u.device
        us = self._check_in(u, "u", self.n_dofs, B, dt, dev)
        fs = self._check_in(field, "field", self._field_size(flags), B, dt, dev)
        if out_R is not None:
            self._check_in(out_R, "out_R", self.n_dofs, B, dt, dev)
            if B > 1 and out_R.stride(0) != self.n_dofs:
                raise ValueError("out_R must be contiguous")
        R = out_R if out_R is not None else (torch.empty(B, self.n_dofs, dtype=dt, device=dev) if want_R else None)
        en = torch.empty(B * self.n_segments, dtype=torch.float64, device=dev) if want_energy else None
        L.check(self._lib.femhot_residual(self._h, _dtype_code(u), B, _ptr(field), fs, _ptr(u), us, _ptr(R),
                                          self.n_dofs, _ptr(en), flags, _stream()))
        return R, (en.view(B, self.n_segments) if en is not None else None)

    def residual_vjp(self, field, u, w, want_gu=True, want_gd=True, flags=0, out=None):
        """(dR/du)^T w and (dR/dfield)^T w; out: preallocated contiguous (g_u, g_field) to write."""
        u = u if u.dim() == 2 else u.unsqueeze(0)
        w = w if w.dim() == 2 else w.unsqueeze(0)
        field = field if field.dim() == 2 else field.unsqueeze(0)
        B = u.shape[0]
        dt, dev = u.dtype, u.device
        us = self._check_in(u, "u", self.n_dofs, B, dt, dev)
        ws = self._check_in(w, "w", self.n_dofs, B, dt, dev)
        fsz = self._field_size(flags)
        fs = self._check_in(field, "field", fsz, B, dt, dev)
        if out is not None:
            gu, gd = out
            for t, name, n in ((gu, "g_u", self.n_dofs), (gd, "g_field", fsz)):
                if t is not None:
                    self._check_in(t, name, n, B, dt, dev)
                    if B > 1 and t.stride(0) != n:
                        raise ValueError(f"{name} must be contiguous")
        else:
            gu = torch.empty(B, self.n_dofs, dtype=dt, device=dev) if want_gu else None
            gd = torch.empty(B, fsz, dtype=dt, device=dev) if want_gd else None
        L.check(self._lib.femhot_residual_vjp(self._h, _dtype_code(u), B, _ptr(field), fs, _ptr(u), us, _ptr(w), ws,
                                              _ptr(gu), self.n_dofs, _ptr(gd), fsz, flags, _stream()))
        return gu, gd

    def energy_grad(self, field, u, want_gu=True, want_gd=True, want_energy=True, flags=0, out=None):
        u = u if u.dim() == 2 else u.unsqueeze(0)
        field = field if field.dim() == 2 else field.unsqueeze(0)
        B = u.shape[0]
        dt, dev = u.dtype, u.device
        us = self._check_in(u, "u", self.n_dofs, B, dt, dev)
        fsz = self._field_size(flags)
        fs = self._check_in(field, "field", fsz, B, dt, dev)
        if out is not None:
            en, gu, gd = out
        else:
            gu = torch.empty(B, self.n_dofs, dtype=dt, device=dev) if want_gu else None
            gd = torch.empty(B, fsz, dtype=dt, device=dev) if want_gd else None
            en = torch.empty(B * self.n_segments, dtype=torch.float64, device=dev) if want_energy else None
        L.check(self._lib.femhot_energy_grad(self._h, _dtype_code(u), B, _ptr(field), fs, _ptr(u), us, _ptr(en),
                                             _ptr(gu), self.n_dofs, _ptr(gd), fsz, flags, _stream()))
        return (en.view(B, self.n_segments) if en is not None else None), gu, gd

    def residual_shape_vjp(self, field, u, w, flags=0):
        """(dR/dX)^T w as an (n_nodes, 3) tensor (one sample): the shape sensitivity of the residual."""
        u, w = u.reshape(-1), w.reshape(-1)
        field = field.reshape(-1)
        dt, dev = u.dtype, u.device
        self._check_in(u.view(1, -1), "u", self.n_dofs, 1, dt, dev)
        self._check_in(w.view(1, -1), "w", self.n_dofs, 1, dt, dev)
        self._check_in(field.view(1, -1), "field", self._field_size(flags), 1, dt, dev)
        gx = torch.empty(self.n_nodes, 3, dtype=dt, device=dev)
        L.check(self._lib.femhot_residual_shape_vjp(self._h, _dtype_code(u), _ptr(field), _ptr(u), _ptr(w), _ptr(gx),
                                                   flags, _stream()))
        return gx

    def gauss_point_data(self, num_gp, dtype=torch.float64, device=None):
        """Integration data at the context's Gauss rule: N (G, n), w detJ (n_elems, G) and the global
        shape gradients dN/dx (n_elems, G, n, dim), from the library's element geometry."""
        device = device or torch.device("cuda")
        G = L.gauss_points_per_element(self.et, num_gp)
        dim = 2 if self.et == L.QUAD4 else 3
        n = self.nodes_per_elem
        N = torch.empty(G, n, dtype=dtype, device=device)
        wd = torch.empty(self.n_elems, G, dtype=dtype, device=device)
        dn = torch.empty(self.n_elems, G, n, dim, dtype=dtype, device=device)
        code = L.F32 if dtype == torch.float32 else L.F64
        L.check(self._lib.femhot_gauss_point_data(self._h, code, _ptr(N), _ptr(wd), _ptr(dn), _stream()))
        return N, wd, dn

    def set_chunk_config(self, op, dtype, batch, epc, block, grid_per_cu=0, flags=0):
        code = L.F32 if dtype == torch.float32 else L.F64
        L.check(self._lib.femhot_set_chunk_config(self._h, int(op), code, int(batch), int(flags), int(epc),
                                                  int(block), int(grid_per_cu)))

    def start_kernel_timing(self):
        L.check(self._lib.femhot_kernel_timing(self._h, 1, None, None))

    def stop_kernel_timing(self):
        """(total ms, number of timed calls) of the dominant kernels since start_kernel_timing()."""
        ms, n = C.c_double(), C.c_longlong()
        L.check(self._lib.femhot_kernel_timing(self._h, 0, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def autotune(self, op, field, u, w=None, flags=0):
        """Pick elements-per-chunk / workgroup size / grid shape for the chunk path by timing them on
        these inputs (L.OP_RESIDUAL / OP_VJP / OP_ENERGY_GRAD).  Returns (epc, block, grid_per_cu, ms)."""
        u = u if u.dim() == 2 else u.unsqueeze(0)
        field = field if field.dim() == 2 else field.unsqueeze(0)
        B = u.shape[0]
        dt, dev = u.dtype, u.device
        us = self._check_in(u, "u", self.n_dofs, B, dt, dev)
        fsz = self._field_size(flags)
        fs = self._check_in(field, "field", fsz, B, dt, dev)
        ws = 0
        if w is not None:
            w = w if w.dim() == 2 else w.unsqueeze(0)
            ws = self._check_in(w, "w", self.n_dofs, B, dt, dev)
        ou = torch.empty(B, self.n_dofs, dtype=dt, device=dev)
        od = torch.empty(B, fsz, dtype=dt, device=dev) if op != L.OP_RESIDUAL else None
        en = torch.empty(B * self.n_segments, dtype=torch.float64, device=dev)
        epc, blk, grid, ms = C.c_int(), C.c_int(), C.c_int(), C.c_float()
        L.check(self._lib.femhot_autotune(self._h, int(op), _dtype_code(u), B, _ptr(field), fs, _ptr(u), us, _ptr(w),
                                          ws, _ptr(ou), self.n_dofs, _ptr(od), fsz, _ptr(en), flags, C.byref(epc),
                                          C.byref(blk), C.byref(grid), C.byref(ms), _stream()))
        return epc.value, blk.value, grid.value, ms.value

    def autotune_tangent(self, field, u, flags=0):
        """Times atomic, colour-ordered and (thermal, linear elastic) row-owner tangent assembly on these
        inputs and keeps the fastest for tangent_csr calls with SCATTER_CHUNK.  Returns (L.PATH_ATOMIC,
        L.PATH_COLOUR or L.PATH_ROWS, ms)."""
        field = field.reshape(-1).contiguous()
        u = u.reshape(-1).contiguous()
        vals = torch.empty(max(self.nnz(), 1), dtype=u.dtype, device=u.device)
        epc, blk, grid, ms = C.c_int(), C.c_int(), C.c_int(), C.c_float()
        L.check(self._lib.femhot_autotune(self._h, L.OP_TANGENT, _dtype_code(u), 1, _ptr(field), 0, _ptr(u), 0, None,
                                          0, _ptr(vals), 0, None, 0, None, int(flags), C.byref(epc), C.byref(blk),
                                          C.byref(grid), C.byref(ms), _stream()))
        return blk.value, ms.value

    def tangent_csr(self, field, u, want_R=True, flags=0, out_vals=None):
        """CSR values (nnz) on the pattern of `pattern()`, BC-masked, and the masked residual.  out_vals:
        a contiguous (nnz) tensor to write (or, with ACCUMULATE, to add) the values into."""
        field = field.reshape(-1)
        u = u.reshape(-1)
        dt, dev = u.dtype, u.device
        self._check_in(u.view(1, -1), "u", self.n_dofs, 1, dt, dev)
        self._check_in(field.view(1, -1), "field", self._field_size(flags), 1, dt, dev)
        if out_vals is None:
            if flags & L.ACCUMULATE:
                raise ValueError("tangent_csr: ACCUMULATE adds into out_vals, which is not given")
            vals = torch.empty(max(self.nnz(), 1), dtype=dt, device=dev)
        else:
            if not out_vals.is_contiguous():
                raise ValueError("tangent_csr: out_vals must be contiguous")
            self._check_in(out_vals.view(1, -1), "out_vals", self.nnz(), 1, dt, dev)
            vals = out_vals
        R = torch.empty(self.n_dofs, dtype=dt, device=dev) if want_R else None
        L.check(self._lib.femhot_tangent_csr(self._h, _dtype_code(u), _ptr(field), _ptr(u), _ptr(vals), _ptr(R),
                                             flags, _stream()))
        return vals[: self.nnz()], R

    def spmv(self, row_ptr, col_idx, vals, x):
        y = torch.empty_like(x)
        L.check(self._lib.femhot_spmv(self._h, _ptr(row_ptr), _ptr(col_idx), _ptr(vals), _ptr(x), _ptr(y), _stream()))
        return y

    def newton(self, field, u, load_steps=1, max_iter=20, abs_tol=1e-8, rel_tol=1e-8, lin_tol=1e-10,
               lin_maxiter=10000, mode="f64", flags=0):
        """Load-stepped Newton solve on the device (femhot_newton).  field, u: FP64 device vectors; u is
        the initial guess.  Returns (u, history (load_steps, max_iter, 2) of (||R||, ||du||) with NaN
        after each step's last iteration, iterations per load step).  Raises FemhotError with
        ERR_NONFINITE when the residual norm becomes NaN."""
        if u.dtype != torch.float64 or field.dtype != torch.float64:
            raise TypeError("newton expects FP64 field and dofs")
        field = field.reshape(-1).contiguous()
        x = u.reshape(-1).clone()
        self._check_in(x.view(1, -1), "u", self.n_dofs, 1, x.dtype, x.device)
        self._check_in(field.view(1, -1), "field", self._field_size(flags), 1, x.dtype, x.device)
        hist = (C.c_double * (2 * load_steps * max_iter))()
        steps = (C.c_int * load_steps)()
        m = L.F64 if mode == "f64" else L.F32
        L.check(self._lib.femhot_newton(self._h, m, _ptr(field), _ptr(x), int(flags), int(load_steps), int(max_iter),
                                        float(abs_tol), float(rel_tol), float(lin_tol), int(lin_maxiter), hist, steps,
                                        _stream()))
        return x, np.array(hist[:], dtype=np.float64).reshape(load_steps, max_iter, 2), list(steps)

    def newton_async(self, field, u, load_steps=1, max_iter=20, abs_tol=1e-8, rel_tol=1e-8, lin_tol=1e-10,
                     lin_iterations=1000, flags=0, out=None):
        """femhot_newton with every decision on the device (femhot_newton_async): u (FP64, initial guess)
        is solved in place and nothing waits for the host, so the call can be captured into a graph
        (after one uncaptured call on this context).  Returns device tensors: history (load_steps,
        max_iter, 2) of (||R||, ||du||) with NaN after each step's last iteration, steps (load_steps)
        int32, status (1) int32 (L.OK, or L.ERR_NONFINITE when a residual norm was NaN).  `out` may
        pass those three tensors in."""
        if u.dtype != torch.float64 or field.dtype != torch.float64:
            raise TypeError("newton_async expects FP64 field and dofs")
        if not u.is_contiguous() or u.numel() != self.n_dofs:
            raise ValueError("newton_async: u must be a contiguous vector of n_dofs")
        field = field.reshape(-1)
        if not field.is_contiguous():
            raise ValueError("newton_async: field must be contiguous")
        self._check_in(u.view(1, -1), "u", self.n_dofs, 1, u.dtype, u.device)
        self._check_in(field.view(1, -1), "field", self._field_size(flags), 1, u.dtype, u.device)
        if out is None:
            out = (torch.empty(load_steps, max_iter, 2, dtype=torch.float64, device=u.device),
                   torch.empty(load_steps, dtype=torch.int32, device=u.device),
                   torch.empty(1, dtype=torch.int32, device=u.device))
        hist, steps, status = out
        if hist.numel() != 2 * load_steps * max_iter or steps.numel() != load_steps:
            raise ValueError("newton_async: out tensors do not match load_steps / max_iter")
        L.check(self._lib.femhot_newton_async(self._h, _ptr(field), _ptr(u), int(flags), int(load_steps), int(max_iter),
                                              float(abs_tol), float(rel_tol), float(lin_tol), int(lin_iterations),
                                              _ptr(hist), _ptr(steps), _ptr(status), _stream()))
        return hist, steps, status

    def pcg(self, row_ptr, col_idx, vals, b, x0=None, tol=1e-10, maxiter=10000, mode="f64", raise_on_fail=True,
            context_pattern=False):
        """Jacobi-PCG (FP64), or FP32 inner iterations + FP64 refinement (mode='mixed').  context_pattern:
        row_ptr/col_idx are this context's pattern (`pattern()`), which lets the SpMV read one column
        index per node block."""
        if vals.dtype != torch.float64 or b.dtype != torch.float64:
            raise TypeError("pcg expects the FP64 tangent and right-hand side")
        x = torch.zeros_like(b) if x0 is None else x0.clone()
        it, rel = C.c_int(), C.c_double()
        m = (L.F64 if mode == "f64" else L.F32) | (L.PCG_CONTEXT_PATTERN if context_pattern else 0)
        rc = self._lib.femhot_pcg(self._h, m, _ptr(row_ptr), _ptr(col_idx), _ptr(vals), _ptr(b), _ptr(x), float(tol),
                                  int(maxiter), C.byref(it), C.byref(rel), _stream())
        if rc == L.ERR_NOT_CONVERGED and not raise_on_fail:
            return x, it.value, rel.value, False
        L.check(rc)
        return x, it.value, rel.value, True

    def pcg_async(self, row_ptr, col_idx, vals, b, x, tol=1e-10, iterations=1000, context_pattern=False,
                  out=None):
        """FP64 Jacobi-PCG with the convergence test on the device (femhot_pcg_async): x (initial guess,
        solved in place), `iterations` launched, nothing waits for the host, so the call can be captured
        into a graph.  Returns device tensors (iters int32, rel_res float64, status int32: 1 converged,
        2 non-finite, 3 iterations spent); `out` may pass those three tensors in."""
        if vals.dtype != torch.float64 or b.dtype != torch.float64 or x.dtype != torch.float64:
            raise TypeError("pcg_async expects the FP64 tangent, right-hand side and solution")
        if not x.is_contiguous() or x.numel() != self.n_dofs:
            raise ValueError("pcg_async: x must be a contiguous vector of n_dofs")
        if out is None:
            out = (torch.empty(1, dtype=torch.int32, device=x.device), torch.empty(1, dtype=torch.float64, device=x.device),
                   torch.empty(1, dtype=torch.int32, device=x.device))
        it, rel, st = out
        m = L.F64 | (L.PCG_CONTEXT_PATTERN if context_pattern else 0)
        L.check(self._lib.femhot_pcg_async(self._h, m, _ptr(row_ptr), _ptr(col_idx), _ptr(vals), _ptr(b), _ptr(x),
                                           float(tol), int(iterations), _ptr(it), _ptr(rel), _ptr(st), _stream()))
        return it, rel, st
