#!/usr/bin/env python3
"""Benchmark of the FE residual + VJP hot path on one MI355X.

Headline workload (synthetic, seeded): 32 TET4 linear-elastic meshes of different sizes -- cubes of
n^3 hexes, n = round(linspace(12, 40, 32)), each hex split into 6 Kuhn tetrahedra (10,368 ... 384,000
tets; 4.40 M tets, 0.81 M nodes in total), interior nodes jittered by U(-0.2h, 0.2h)^3, per-element
E ~ LogUniform(1, 100), nu = 0.3, nodal displacements U(-1e-3, 1e-3).

One step = the residual (dPi/du) and the VJP with respect to the per-element E (dPi/dE), plus the
per-mesh potentials, for all 32 meshes:
  * batched FP32: all meshes in ONE launch of the chunk kernel (persistent), captured in a HIP graph
    and replayed; the chunk configuration is the profiled one (profiles/tet4_uneven_chunk_traffic.json)
    unless --autotune or --epc/--block/--grid say otherwise;
  * plain FP64 (the drop-in single-instance kernel): one element-parallel, atomic-scatter call per
    mesh, 32 calls per step.
value = batched FP32 Melem/s; also reported: the plain FP64 rate and the ratio, the FP32 error
against FP64, the roofline of the dominant kernels, the FP64 C/OpenMP CPU baseline (all host cores
and one), and two more workloads: nonlinear thermal QUAD4 128^2 x 64 samples, the residual-norm loss
and its VJP with respect to T and k0 (FP32, one graph replay per step); and HEX8 64^3 linear
elasticity, BC application + Jacobian (CSR tangent) + residual in FP64, the call the reference's own
timing harness times (examples/performance_study/matrix_vector_perf.py).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "batched FP32 residual+VJP Melem/s vs plain FP64 at ≤1e-5 rel err; % HBM peak"
HBM_PEAK_GBS = 8000.0
TRAFFIC_JSON = os.path.join(ROOT, "profiles", "tet4_uneven_chunk_traffic.json")
C3_TRAFFIC_JSON = os.path.join(ROOT, "profiles", "quad128x64_thermal_traffic.json")
C2_TRAFFIC_JSON = os.path.join(ROOT, "profiles", "hex64_tangent_traffic.json")


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def build_workload(n_meshes=32, nmin=12, nmax=40):
    from folax_amd.mesh import create_3D_tet_box_mesh
    ns = np.round(np.linspace(nmin, nmax, n_meshes)).astype(int)
    coords, conns, E, u, noff, eoff = [], [], [], [], [0], [0]
    for m, n in enumerate(ns):
        msh = create_3D_tet_box_mesh(int(n), 1.0, jitter=0.2, seed=5000 + m)
        c = msh.GetNodesCoordinates()
        k = msh.GetElementsNodes("tetra")
        rng_e = np.random.default_rng(5100 + m)
        rng_u = np.random.default_rng(5200 + m)
        coords.append(c)
        conns.append(k.astype(np.int64) + noff[-1])
        E.append(np.exp(rng_e.uniform(0.0, np.log(100.0), size=k.shape[0])))
        u.append(rng_u.uniform(-1e-3, 1e-3, size=3 * c.shape[0]))
        noff.append(noff[-1] + c.shape[0])
        eoff.append(eoff[-1] + k.shape[0])
    return dict(ns=ns, coords=np.concatenate(coords), conn=np.concatenate(conns).astype(np.int32),
                E=np.concatenate(E), u=np.concatenate(u), noff=np.array(noff), eoff=np.array(eoff))


def build_c3(B=64, N=129):
    """QUAD4 128^2: per sample b, k0_b = 0.1 + 0.9 sigmoid(sum of 16 Fourier modes, seed 3000+b) and
    T_b ~ U(0, 1) (seed 3100+b); T = 1 on the left edge, 0 on the right; beta = 0.5, c = 2."""
    from folax_amd.controls import fourier_field
    from folax_amd.mesh import create_2D_square_mesh
    m = create_2D_square_mesh(1.0, N)
    coords = m.GetNodesCoordinates()
    K0 = np.stack([fourier_field(coords, np.random.default_rng(3000 + b).standard_normal(16)) for b in range(B)])
    T = np.stack([np.random.default_rng(3100 + b).uniform(0.0, 1.0, size=coords.shape[0]) for b in range(B)])
    left, right = m.GetNodeSet("left"), m.GetNodeSet("right")
    T[:, left], T[:, right] = 1.0, 0.0
    return dict(m=m, coords=coords, conn=m.GetElementsNodes("quad"), K0=K0, T=T,
                didx=np.concatenate([left, right]).astype(np.int64),
                dval=np.concatenate([np.ones(left.size), np.zeros(right.size)]))


def rel(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _timed_reps(fn, budget_s):
    fn()  # warm-up
    t0 = time.perf_counter()
    fn()
    t1 = time.perf_counter() - t0
    reps = int(max(1, min(5000, budget_s / max(t1, 1e-6))))
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    return reps, time.perf_counter() - t0


def cpu_baseline_c5(W, budget_s=10.0, budget_1t_s=8.0):
    """FP64 C/OpenMP restatement (oracle/libcpuref.so) of the whole workload, on every host core and on
    one, each repeated for about its budget."""
    from oracle import cpuref
    seg = np.repeat(np.arange(len(W["ns"])), np.diff(W["eoff"])).astype(np.int64)
    Ne = W["conn"].shape[0]
    out = (np.zeros(W["u"].shape[0]), np.zeros(Ne), np.zeros(len(W["ns"])))
    fn = lambda: cpuref.tet4_elastic_energy_grad(W["coords"], W["conn"], W["E"], W["u"], seg_of_elem=seg,
                                                 n_seg=len(W["ns"]), out=out)
    cores = cpuref.num_threads()
    reps, t = _timed_reps(fn, budget_s)
    cpuref.set_threads(1)
    reps1, t1 = _timed_reps(fn, budget_1t_s)
    cpuref.set_threads(cores)
    return {"value": round(Ne * reps / t / 1e6, 3), "unit": "Melem/s", "cores": cores, "kind": "port",
            "sample": f"FP64 residual + dPi/dE of all 32 meshes ({Ne} tets): {reps} repetitions after 1 warm-up "
                      f"in {t:.1f} s on {cores} threads; {reps1} in {t1:.1f} s on 1 thread",
            "single_thread": {"value": round(Ne * reps1 / t1 / 1e6, 3), "unit": "Melem/s", "cores": 1},
            "cpu_model": cpu_model()}, out


def cpu_baseline_c3(C, samples=64, budget_s=6.0, budget_1t_s=8.0):
    """FP64 C/OpenMP restatement of the nonlinear-thermal residual-norm loss and its exact VJP over the
    first `samples` samples (all 64 by default), repeated for about the budget on every core and on one."""
    from oracle import cpuref
    Ne = C["conn"].shape[0]
    K0, T = C["K0"][:samples], C["T"][:samples]
    fn = lambda: cpuref.quad4_thermal_resnorm_vjp(C["coords"], C["conn"], K0, T, C["didx"], beta=0.5, c=2.0)
    cores = cpuref.num_threads()
    reps, t = _timed_reps(fn, budget_s)
    cpuref.set_threads(1)
    reps1, t1 = _timed_reps(fn, budget_1t_s)
    cpuref.set_threads(cores)
    per = Ne * samples
    return {"value": round(per * reps / t / 1e6, 3), "unit": "Melem/s", "cores": cores, "kind": "port",
            "sample": f"FP64 residual-norm loss + VJP (g_T, g_k0) of {samples} of the 64 samples ({per} element-"
                      f"samples): {reps} repetitions in {t:.1f} s on {cores} threads; {reps1} in {t1:.1f} s on 1 "
                      f"thread",
            "single_thread": {"value": round(per * reps1 / t1 / 1e6, 3), "unit": "Melem/s", "cores": 1},
            "cpu_model": cpu_model()}


def load_profiled(path):
    try:
        return json.load(open(path))
    except (OSError, ValueError):
        return None


def capture(step, torch):
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        step()
    torch.cuda.current_stream().wait_stream(side)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        step()
    return g


def time_replays(run, steps, torch):
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(steps):
        run()
    ev1.record()
    torch.cuda.synchronize()
    return ev0.elapsed_time(ev1) / steps


def roofline(alg_bytes, kernel_ms, traffic):
    achieved = alg_bytes / (kernel_ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic}


def traffic_if_same(prof, cfg):
    """HBM bytes per launch from the profile, only when it was taken at this very configuration."""
    if prof and prof.get("config") == cfg:
        return prof.get("bytes_per_launch")
    return None


def pick_config(ctx, op, E, u, w, flags, args, prof, torch, L):
    """The chunk configuration of a workload: explicit flags, else the profiled one, else the autotuner's."""
    if args.epc or args.block in (L.PATH_ATOMIC, L.PATH_COLOUR):
        cfg = {"epc": args.epc, "block": args.block, "grid_per_cu": args.grid}
        src = "command line"
    elif prof and prof.get("config") and not args.autotune:
        cfg = dict(prof["config"])
        src = "profiled"
    else:
        epc, blk, grid, ms = ctx.autotune(op, E, u, w, flags=flags)
        cfg = {"epc": epc, "block": blk, "grid_per_cu": grid}
        return cfg, f"autotuned ({ms * 1e3:.1f} us)"
    ctx.set_chunk_config(op, E.dtype, E.shape[0], cfg["epc"], cfg["block"], cfg["grid_per_cu"], flags=flags)
    return cfg, src


def run_c3(args, torch, L):
    """QUAD4 128^2 x 64 nonlinear thermal: R (Dirichlet-masked), L = 1/2 sum |R|^2 and the VJP with w = R
    (g_T, g_k0), FP32, all samples in one chunk launch per kernel, replayed as a HIP graph."""
    from folax_amd.context import FemhotContext
    C = build_c3()
    B, Nn = C["T"].shape
    Ne = C["conn"].shape[0]
    dev = torch.device("cuda")
    ctx = FemhotContext("quad", L.THERMAL, C["coords"], C["conn"], num_gp=2, beta=0.5, c=2.0,
                        dirichlet_idx=C["didx"], dirichlet_val=C["dval"])
    flags = L.SCATTER_CHUNK
    tune = 0
    out = {}
    for dt in (torch.float64, torch.float32):
        K = torch.as_tensor(C["K0"], dtype=dt, device=dev)
        T = torch.as_tensor(C["T"], dtype=dt, device=dev)
        R = torch.empty(B, Nn, dtype=dt, device=dev)
        gT, gk = torch.empty_like(R), torch.empty_like(R)
        loss = torch.empty((), dtype=torch.float64, device=dev)
        cfgs = {}
        if dt == torch.float32:
            prof = load_profiled(C3_TRAFFIC_JSON)
            for op, key in ((L.OP_RESIDUAL, "residual"), (L.OP_VJP, "vjp")):
                p = {"config": prof["config"][key]} if prof and prof.get("config") else None
                cfgs[key], src = pick_config(ctx, op, K, T, T, tune, args, p, torch, L)
                log(f"C3 {key} config: {cfgs[key]} ({src})")

        def step():
            ctx.residual(K, T, want_energy=False, flags=flags | L.MASK_DIRICHLET, out_R=R)
            loss.copy_(torch.linalg.vector_norm(R).square())  # the residual's own precision, one reduction
            ctx.residual_vjp(K, T, R, flags=flags | L.MASK_DIRICHLET, out=(gT, gk))

        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        out[dt] = (float(loss) * 0.5, gT.double().clone(), gk.double().clone())
        if dt == torch.float64:
            continue
        ctx.start_kernel_timing()
        for _ in range(args.steps):
            step()
        kms, kn = ctx.stop_kernel_timing()
        kernel_ms = kms * 2 / max(kn, 1)  # two timed calls per step
        g = None if args.no_graph else capture(step, torch)
        step_ms = time_replays(g.replay if g is not None else step, args.steps, torch)
    L64, gT64, gk64 = out[torch.float64]
    L32, gT32, gk32 = out[torch.float32]
    err = {"loss": float(f"{abs(L32 - L64) / abs(L64):.3e}"), "g_T": float(f"{rel(gT32.cpu(), gT64.cpu()):.3e}"),
           "g_k0": float(f"{rel(gk32.cpu(), gk64.cpu()):.3e}")}
    # compulsory bytes: T and k0 read, g_T and g_k0 written (4 B each per node and sample), the mesh once
    alg = B * Nn * 16 + Ne * 16 + Nn * 8
    prof = load_profiled(C3_TRAFFIC_JSON) or {}
    traffic = prof.get("bytes_per_step") if prof.get("config") == cfgs else None
    res = {"workload": "nonlinear thermal QUAD4 128^2 x 64 samples (k = k0 (1 + 0.5 T^2)): residual-norm loss "
                       "and its VJP wrt T and k0, residual + VJP chunk launches replayed as one HIP graph",
           "value": round(B * Ne / (step_ms * 1e-3) / 1e6, 2), "unit": "Melem/s", "ms_per_step": round(step_ms, 5),
           "kernel_ms": round(kernel_ms, 5), "dtype": "f32", "elements": Ne, "samples": B,
           "chunk": cfgs, "rel_err_vs_fp64": err, "algorithmic_bytes_per_step": int(alg),
           "roofline": roofline(alg, kernel_ms, traffic)}
    log(f"C3: step {step_ms * 1e3:.1f} us, kernels {kernel_ms * 1e3:.1f} us, err {err}")
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline_c3(C)
        log(f"C3 CPU baseline: {res['cpu_baseline']}")
    ctx.close()
    return res


def cpu_baseline_c2(n=20):
    """The oracle's FP64 assembly (numpy, one process) of the reference timing harness's own problem
    (examples/performance_study/matrix_vector_perf.py: HEX8 20^3 box, E = 1, BCs on left/right)."""
    from oracle import fe
    coords, conn, sets = fe.box_hex_mesh(n, n, n)
    bc = {"Ux": {"left": 0.0, "right": 0.1}, "Uy": {"left": 0.0, "right": 0.1}, "Uz": {"left": 0.0, "right": 0.1}}
    didx, dval = fe.dirichlet_dofs(sets, ["Ux", "Uy", "Uz"], bc)
    ph = fe.Physics("mechanical", "hexahedron", num_gp=2, E=1.0, nu=0.3)
    u = fe.apply_dirichlet(np.zeros(3 * coords.shape[0]), didx, dval)
    t0 = time.time()
    fe.assemble(ph, coords, conn, np.ones(coords.shape[0]), u, didx)
    dt = time.time() - t0
    return {"value": round(conn.shape[0] / dt / 1e6, 5), "unit": "Melem/s", "cores": 1, "kind": "port",
            "cpu": cpu_model(), "sample": f"oracle/fe.py assemble (numpy FP64, tangent + residual) of the HEX8 "
                                         f"{n}^3 harness problem, one call, {dt:.1f} s"}


def run_c2(args, torch, L):
    """HEX8 64^3 linear elasticity (E in {1, 10} from the two-phase random field, clamped at z = 0, 1 %
    stretch at z = 1), FP64: BC application + ComputeJacobianMatrixAndResidualVector through
    MechanicalLoss3DHexa, as the reference's timing harness times it (matrix_vector_perf.py:50-69), and
    the tangent assembly alone in each of its three orders."""
    from folax_amd import loss_functions as lf
    from folax_amd.controls import grf_two_phase
    from folax_amd.mesh import create_3D_box_mesh
    t0 = time.time()
    m = create_3D_box_mesh(64, 64, 64)
    bc = {"Ux": {"bottom": 0.0}, "Uy": {"bottom": 0.0}, "Uz": {"bottom": 0.0, "top": 0.01}}
    loss = lf.MechanicalLoss3DHexa("c2", {"dirichlet_bc_dict": bc, "num_gp": 2,
                                          "material_dict": {"young_modulus": 1.0, "poisson_ratio": 0.3}}, m)
    loss.Initialize()
    dev = torch.device("cuda")
    Nn, Ne = m.GetNumberOfNodes(), m.GetNumberOfElements("hexahedron")
    E = torch.as_tensor(grf_two_phase((65, 65, 65), 8.0, 2001), dtype=torch.float64, device=dev)
    u0 = torch.zeros(3 * Nn, dtype=torch.float64, device=dev)

    def call():
        u = loss.ApplyDirichletBCOnDofVector(u0)
        return loss.ComputeJacobianMatrixAndResidualVector(E, u)

    K, R = call()  # pattern, plans and first launch
    torch.cuda.synchronize()
    log(f"C2 set-up {time.time() - t0:.1f}s, nnz {K.values.numel()}")
    reps = max(5, min(args.steps, 20))
    call_ms = time_replays(call, reps, torch)
    ctx, u = loss.ctx, loss.ApplyDirichletBCOnDofVector(u0)
    vals = torch.empty_like(K.values)
    t = {}
    for name, fl in (("rows", L.SCATTER_ROWS), ("colour", L.SCATTER_COLOUR), ("atomic", 0)):
        ctx.tangent_csr(E, u, want_R=False, flags=fl, out_vals=vals)
        t[name] = time_replays(lambda: ctx.tangent_csr(E, u, want_R=False, flags=fl, out_vals=vals),
                               reps if name == "rows" else 3, torch)
    nnz = K.values.numel()
    # compulsory bytes of one tangent assembly: the CSR values written once, the mesh (connectivity,
    # coordinates) and the nodal E read once
    alg = nnz * 8 + Ne * 32 + Nn * (24 + 8)
    prof = load_profiled(C2_TRAFFIC_JSON) or {}
    res = {"workload": "HEX8 64^3 linear elasticity, two-phase E, FP64: ApplyDirichletBCOnDofVector + "
                       "ComputeJacobianMatrixAndResidualVector (CSR tangent + masked residual)",
           "elements": Ne, "nnz": nnz, "ms_per_call": round(call_ms, 4),
           "value": round(Ne / (call_ms * 1e-3) / 1e6, 2), "unit": "Melem/s",
           "tangent_ms": {k: round(v, 4) for k, v in t.items()},
           "algorithmic_bytes_tangent": int(alg),
           "roofline": roofline(alg, t["rows"], prof.get("bytes_per_launch") if prof.get("path") == "rows" else None)}
    log(f"C2: call {call_ms * 1e3:.0f} us, tangent rows/colour/atomic {t['rows'] * 1e3:.0f} / "
        f"{t['colour'] * 1e3:.0f} / {t['atomic'] * 1e3:.0f} us")
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline_c2()
        log(f"C2 CPU baseline: {res['cpu_baseline']}")
    loss.ctx.close()
    return res


def solver_profile(iters=200):
    """HEX8 64^3 linear elasticity (274,625 nodes, 823,875 dofs), E in {1, 10} from a seeded Gaussian random
    field (correlation length 8 voxels), clamped at z = 0 with 1 % stretch at z = 1: FP64 tangent + residual
    (colour-ordered assembly), then `iters` FP64 Jacobi-PCG iterations with the node-blocked SpMV, then
    `iters` iterations of the mixed solve (FP32 inner iterations, FP64 refinement) on the same system."""
    import torch
    from folax_amd import loss_functions as lf
    from folax_amd.controls import grf_two_phase
    from folax_amd.mesh import create_3D_box_mesh
    m = create_3D_box_mesh(64, 64, 64)
    bc = {"Ux": {"bottom": 0.0}, "Uy": {"bottom": 0.0}, "Uz": {"bottom": 0.0, "top": 0.01}}
    loss = lf.MechanicalLoss3DHexa("c2", {"dirichlet_bc_dict": bc, "num_gp": 2,
                                          "material_dict": {"young_modulus": 1.0, "poisson_ratio": 0.3}}, m)
    loss.Initialize()
    dev = torch.device("cuda")
    E = torch.as_tensor(grf_two_phase((65, 65, 65), 8.0, 2001), dtype=torch.float64, device=dev)
    u = loss.ApplyDirichletBCOnDofVector(torch.zeros(3 * m.GetNumberOfNodes(), dtype=torch.float64, device=dev))
    K, R = loss.ComputeJacobianMatrixAndResidualVector(E, u)
    x, it, rr, ok = loss.ctx.pcg(K.row_ptr, K.col_idx, K.values, (-R).contiguous(), tol=1e-14, maxiter=iters,
                                 raise_on_fail=False, context_pattern=True)
    torch.cuda.synchronize()
    log(f"solver profile: HEX8 64^3 tangent nnz {K.values.numel()}, {it} FP64 PCG iterations, rel residual {rr:.3e}")
    x, it, rr, ok = loss.ctx.pcg(K.row_ptr, K.col_idx, K.values, (-R).contiguous(), tol=1e-14, maxiter=iters,
                                 mode="mixed", raise_on_fail=False, context_pattern=True)
    torch.cuda.synchronize()
    log(f"solver profile: {it} mixed PCG iterations, rel residual {rr:.3e}")


def c4_profile(args):
    """HEX8 48^3 x 16 samples compressible Neo-Hookean (mu = 1, lambda = 5 in the E = 1 phase): FP32
    energy + gradient (dPi/du, dPi/dE) of all samples in one chunk launch, `steps` times (for rocprofv3),
    at --epc/--block/--grid when given, else at the autotuner's configuration."""
    steps = args.steps
    import torch
    from folax_amd import _lib as L
    from folax_amd.context import FemhotContext
    from folax_amd.controls import grf_two_phase
    from folax_amd.mesh import create_3D_box_mesh
    m = create_3D_box_mesh(48, 48, 48)
    B = 16
    E = np.stack([(17.0 / 6.0) * grf_two_phase((49, 49, 49), 8.0, 4000 + b) for b in range(B)])
    ctx = FemhotContext("hexahedron", L.NEOHOOKE, m.GetNodesCoordinates(), m.GetElementsNodes("hexahedron"),
                        num_gp=2, poisson_ratio=5.0 / 12.0)
    dev = torch.device("cuda")
    Et = torch.as_tensor(E, dtype=torch.float32, device=dev)
    U = torch.as_tensor(np.random.default_rng(4300).uniform(-1e-3, 1e-3, size=(B, 3 * m.GetNumberOfNodes())),
                        dtype=torch.float32, device=dev)
    if args.epc or args.block in (L.PATH_ATOMIC, L.PATH_COLOUR):
        ctx.set_chunk_config(L.OP_ENERGY_GRAD, torch.float32, B, args.epc, args.block, args.grid)
        log(f"C4 config: {args.epc} elements/chunk, block {args.block}, grid {args.grid}")
    else:
        epc, blk, grid, ms = ctx.autotune(L.OP_ENERGY_GRAD, Et, U)
        log(f"C4 config: --epc {epc} --block {blk} --grid {grid} ({ms * 1e3:.1f} us)")
    for _ in range(3):
        ctx.energy_grad(Et, U, flags=L.SCATTER_CHUNK)
    torch.cuda.synchronize()
    ctx.start_kernel_timing()
    for _ in range(steps):
        ctx.energy_grad(Et, U, flags=L.SCATTER_CHUNK)
    ms, n = ctx.stop_kernel_timing()
    Ne = m.GetNumberOfElements("hexahedron")
    log(f"C4: {ms / max(n, 1) * 1e3:.1f} us of kernels per call, {B * Ne / (ms / max(n, 1) * 1e-3) / 1e6:.1f} "
        f"Melem/s over {B} x {Ne} element-samples")
    log("C4 profile run done")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline legs")
    ap.add_argument("--no-graph", action="store_true", help="time eager launches instead of graph replays")
    ap.add_argument("--no-c3", action="store_true", help="skip the QUAD4 128^2 x 64 thermal line")
    ap.add_argument("--no-c2", action="store_true", help="skip the HEX8 64^3 Jacobian + residual line")
    ap.add_argument("--c2-only", action="store_true", help="run only the HEX8 64^3 Jacobian line (for rocprofv3)")
    ap.add_argument("--autotune", action="store_true", help="autotune instead of the profiled configuration")
    ap.add_argument("--profile", choices=["c5", "c3", "c4"], default=None,
                    help="short run of one workload's kernels for rocprofv3 (c5: + a 256 MiB copy for calibration)")
    ap.add_argument("--epc", type=int, default=0, help="fixed elements per chunk")
    ap.add_argument("--block", type=int, default=256, help="with --epc: threads (| FEMHOT_CHUNK_DMA); with --epc 0: 1 = atomic, 2 = colour-ordered "
                         "element-parallel path")
    ap.add_argument("--grid", type=int, default=0, help="with --epc: workgroups per CU (0: one per chunk)")
    ap.add_argument("--solver-profile", action="store_true",
                    help="short run for rocprofv3 of the linear-solver kernels instead: HEX8 64^3 two-phase "
                         "elastic tangent (colour-ordered FP64 assembly) + 200 FP64 and 200 mixed PCG iterations")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: --gpus {args.gpus} is not supported (the library has no multi-GPU path)",
              file=sys.stderr)
        sys.exit(2)
    import torch
    from folax_amd import _lib as L
    from folax_amd.context import FemhotContext

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    if args.solver_profile:
        return solver_profile()
    if args.profile == "c4":
        return c4_profile(args)
    if args.c2_only:
        args.no_cpu = True
        print(json.dumps(run_c2(args, torch, L)), flush=True)
        return
    if args.profile == "c3":
        args.no_cpu = True
        run_c3(args, torch, L)
        log("C3 profile run done")
        return
    t0 = time.time()
    W = build_workload()
    Ne, Nn = W["conn"].shape[0], W["coords"].shape[0]
    nm = len(W["ns"])
    log(f"workload: {nm} meshes, {Ne} tets, {Nn} nodes ({time.time() - t0:.1f}s)")
    dev = torch.device("cuda")
    flags = L.ELEMENT_FIELD
    # ---------------- batched FP32: one context over the disjoint union, chunk path ----------------
    ctx = FemhotContext("tetra", L.MECHANICAL, W["coords"], W["conn"], young_modulus=1.0, poisson_ratio=0.3,
                        segment_offsets=W["eoff"])
    E32 = torch.as_tensor(W["E"], dtype=torch.float32, device=dev).view(1, -1)
    u32 = torch.as_tensor(W["u"], dtype=torch.float32, device=dev).view(1, -1)
    en32 = torch.empty(nm, dtype=torch.float64, device=dev)
    gu32 = torch.empty(1, 3 * Nn, dtype=torch.float32, device=dev)
    gd32 = torch.empty(1, Ne, dtype=torch.float32, device=dev)
    fl32 = flags | L.SCATTER_CHUNK
    prof = load_profiled(TRAFFIC_JSON)
    t1 = time.time()
    cfg, src = pick_config(ctx, L.OP_ENERGY_GRAD, E32, u32, None, flags, args, prof, torch, L)
    log(f"chunk config: {cfg} ({src}, {time.time() - t1:.1f}s)")

    def step32():
        ctx.energy_grad(E32, u32, flags=fl32, out=(en32, gu32, gd32))

    for _ in range(args.warmup):
        step32()
    torch.cuda.synchronize()
    ref_eager = (en32.clone(), gu32.clone(), gd32.clone())
    if args.profile == "c5":
        for _ in range(args.steps):
            step32()
        # calibration for HBM counters: a plain copy of 256 MiB (read 256 MiB, write 256 MiB)
        src_t = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device=dev)
        dst_t = torch.empty_like(src_t)
        for _ in range(3):
            dst_t.copy_(src_t)
        torch.cuda.synchronize()
        log("profile run done")
        return
    # kernel-only timing (HIP events recorded by the library on the launch stream around the chunk kernel
    # and its reduction)
    ctx.start_kernel_timing()
    for _ in range(args.steps):
        step32()
    kms, kn = ctx.stop_kernel_timing()
    kernel_ms = kms / max(kn, 1)
    graph = None
    graph_ok = None
    if not args.no_graph:
        graph = capture(step32, torch)
        graph_ok = True
        for _ in range(3):  # replays checked against the eager result (bit for bit), with eager work between
            gu32.zero_()
            graph.replay()
            torch.cuda.synchronize()
            graph_ok &= bool(torch.equal(gd32, ref_eager[2])) and bool(torch.equal(gu32, ref_eager[1])) and \
                bool(torch.equal(en32, ref_eager[0]))
            _ = (E32 * 2).sum().item()
    step_ms = time_replays(graph.replay if graph is not None else step32, args.steps, torch)
    log(f"FP32 batched: step {step_ms * 1e3:.1f} us, chunk kernels {kernel_ms * 1e3:.1f} us")
    # ---------------- plain FP64 single-instance kernel, one call per mesh ----------------
    ctxs, ins, outs = [], [], []
    for m in range(nm):
        n0, n1 = W["noff"][m], W["noff"][m + 1]
        e0, e1 = W["eoff"][m], W["eoff"][m + 1]
        cm = FemhotContext("tetra", L.MECHANICAL, W["coords"][n0:n1], W["conn"][e0:e1] - n0, young_modulus=1.0,
                           poisson_ratio=0.3)
        ctxs.append(cm)
        ins.append((torch.as_tensor(W["E"][e0:e1], dtype=torch.float64, device=dev).view(1, -1),
                    torch.as_tensor(W["u"][3 * n0:3 * n1], dtype=torch.float64, device=dev).view(1, -1)))
        outs.append((torch.empty(1, dtype=torch.float64, device=dev),
                     torch.empty(1, 3 * (n1 - n0), dtype=torch.float64, device=dev),
                     torch.empty(1, e1 - e0, dtype=torch.float64, device=dev)))

    def step64():
        for m in range(nm):
            ctxs[m].energy_grad(ins[m][0], ins[m][1], flags=L.ELEMENT_FIELD, out=outs[m])

    for _ in range(args.warmup):
        step64()
    step64_ms = time_replays(step64, args.steps, torch)
    for c in ctxs:
        c.start_kernel_timing()
    step64()
    k64 = sum(c.stop_kernel_timing()[0] for c in ctxs)
    log(f"FP64 plain: step {step64_ms * 1e3:.1f} us (kernels {k64 * 1e3:.1f} us)")
    gu64 = torch.cat([o[1].view(-1) for o in outs]).cpu().numpy()
    gd64 = torch.cat([o[2].view(-1) for o in outs]).cpu().numpy()
    en64 = torch.cat([o[0].view(-1) for o in outs]).cpu().numpy()
    err = {"g_u": rel(gu32.cpu().numpy(), gu64), "g_E": rel(gd32.cpu().numpy(), gd64),
           "energy": rel(en32.cpu().numpy(), en64)}
    log(f"FP32 vs FP64 relative error: {err}")
    # ---------------- roofline of the dominant kernels ----------------
    alg_bytes = Ne * (16 + 4 + 4) + Nn * (12 + 12 + 12)
    roof = roofline(alg_bytes, kernel_ms, traffic_if_same(prof, cfg))
    cpu = None
    if not args.no_cpu:
        cpu, cpu_out = cpu_baseline_c5(W)
        err["cpu_fp64_vs_gpu_fp64_g_u"] = rel(gu64, cpu_out[0])
        log(f"CPU baseline: {cpu}")
    c3 = None if args.no_c3 else run_c3(args, torch, L)
    c2 = None if args.no_c2 else run_c2(args, torch, L)
    value = Ne / (step_ms * 1e-3) / 1e6
    v64 = Ne / (step64_ms * 1e-3) / 1e6
    out = {"metric": METRIC, "value": round(value, 2), "unit": "Melem/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(step_ms, 5), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "32 uneven TET4 elastic meshes (12^3..40^3 hexes x 6 Kuhn tets), residual + "
                                  "VJP wrt per-element E; chunk kernel + one reduction kernel per step, replayed as a "
                                  "HIP graph",
                      "meshes": nm, "tets": int(Ne), "nodes": int(Nn), "parallelism": "single GPU",
                      "chunk": cfg, "chunk_source": src, "graph": graph is not None,
                      "graph_replays_match_eager": graph_ok},
           "fp64_plain": {"value": round(v64, 2), "unit": "Melem/s", "ms_per_step": round(step64_ms, 5),
                          "kernel_ms": round(k64, 5), "launches_per_step": nm},
           "speedup_vs_fp64_plain": round(value / v64, 2),
           "rel_err_vs_fp64": {k: float(f"{v:.3e}") for k, v in err.items()},
           "kernel_ms": round(kernel_ms, 5), "algorithmic_bytes_per_launch": int(alg_bytes),
           "roofline": roof, "cpu_baseline": cpu, "c3_quad128x64_thermal": c3, "c2_hex64_jacobian": c2}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
