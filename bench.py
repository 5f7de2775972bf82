#!/usr/bin/env python3
"""Benchmark of the FE residual + VJP hot path on one MI355X.

Workload (synthetic, seeded): 32 TET4 linear-elastic meshes of different sizes -- cubes of n^3
hexes, n = round(linspace(12, 40, 32)), each hex split into 6 Kuhn tetrahedra (10,368 ... 384,000
tets; 4.40 M tets, 0.81 M nodes in total), interior nodes jittered by U(-0.2h, 0.2h)^3, per-element
E ~ LogUniform(1, 100), nu = 0.3, nodal displacements U(-1e-3, 1e-3).

One step = the residual (dPi/du) and the VJP with respect to the per-element E (dPi/dE), plus the
per-mesh potentials, for all 32 meshes:
  * batched FP32: all meshes in ONE launch of the LDS-staged chunk kernel (persistent, autotuned),
    captured in a HIP graph and replayed;
  * plain FP64 (the drop-in single-instance kernel): one element-parallel, atomic-scatter call per
    mesh, 32 calls per step.
value = batched FP32 Melem/s; also reported: the plain FP64 rate and the ratio, the FP32 error
against FP64, the roofline of the dominant kernel and the FP64 C/OpenMP CPU baseline.
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


def rel(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def cpu_baseline(W, budget_s=10.0):
    """FP64 C/OpenMP restatement (oracle/libcpuref.so) on the whole workload, repeated to ~budget_s."""
    from oracle import cpuref
    seg = np.repeat(np.arange(len(W["ns"])), np.diff(W["eoff"])).astype(np.int64)
    Ne = W["conn"].shape[0]
    out = (np.zeros(W["u"].shape[0]), np.zeros(Ne), np.zeros(len(W["ns"])))
    cpuref.tet4_elastic_energy_grad(W["coords"], W["conn"], W["E"], W["u"], seg_of_elem=seg, n_seg=len(W["ns"]),
                                    out=out)  # warm-up
    t0 = time.perf_counter()
    cpuref.tet4_elastic_energy_grad(W["coords"], W["conn"], W["E"], W["u"], seg_of_elem=seg, n_seg=len(W["ns"]),
                                    out=out)
    t1 = time.perf_counter() - t0
    reps = int(max(1, min(200, budget_s / max(t1, 1e-6))))
    t0 = time.perf_counter()
    for _ in range(reps):
        cpuref.tet4_elastic_energy_grad(W["coords"], W["conn"], W["E"], W["u"], seg_of_elem=seg, n_seg=len(W["ns"]),
                                        out=out)
    t = time.perf_counter() - t0
    return {"value": round(Ne * reps / t / 1e6, 3), "unit": "Melem/s", "cores": cpuref.num_threads(),
            "kind": "port",
            "sample": f"FP64 residual + dPi/dE of all 32 meshes ({Ne} tets), {reps} repetitions after 1 warm-up, "
                      f"{t:.1f} s"}, out


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-graph", action="store_true", help="time eager launches instead of graph replays")
    ap.add_argument("--profile", action="store_true", help="short run for rocprofv3 (FP32 chunk path only)")
    ap.add_argument("--epc", type=int, default=0, help="fixed elements per chunk (skips the autotuner)")
    ap.add_argument("--block", type=int, default=256)
    ap.add_argument("--grid", type=int, default=0, help="workgroups per CU (0: one per chunk)")
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
    t1 = time.time()
    if args.epc:
        epc, blk, grid, tms = args.epc, args.block, args.grid, float("nan")
        ctx.set_chunk_config(L.OP_ENERGY_GRAD, torch.float32, 1, epc, blk, grid, flags=flags)
    else:
        epc, blk, grid, tms = ctx.autotune(L.OP_ENERGY_GRAD, E32, u32, flags=flags)
    log(f"chunk config: {epc} elements/chunk, {blk} threads, grid {'all chunks' if grid == 0 else f'{grid}/CU'}, "
        f"tuned {tms * 1e3:.1f} us ({time.time() - t1:.1f}s)")

    def step32():
        ctx.energy_grad(E32, u32, flags=fl32, out=(en32, gu32, gd32))

    for _ in range(args.warmup):
        step32()
    torch.cuda.synchronize()
    ref_eager = (en32.clone(), gu32.clone(), gd32.clone())
    if args.profile:
        for _ in range(args.steps):
            step32()
        # calibration for HBM counters: a plain copy of 256 MiB (read 256 MiB, write 256 MiB)
        src = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("profile run done")
        return
    # kernel-only timing (HIP events on the launch stream around the chunk kernel)
    ctx.start_kernel_timing()
    for _ in range(args.steps):
        step32()
    kms, kn = ctx.stop_kernel_timing()
    kernel_ms = kms / max(kn, 1)
    # whole-step timing: graph replays (or eager calls)
    graph = None
    graph_ok = None
    if not args.no_graph:
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            step32()
        torch.cuda.current_stream().wait_stream(side)
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            step32()
        graph_ok = True
        for _ in range(3):  # replays checked against the eager result, with eager work in between
            gu32.zero_()
            graph.replay()
            torch.cuda.synchronize()
            graph_ok &= bool(torch.equal(gd32, ref_eager[2])) and rel(gu32.cpu(), ref_eager[1].cpu()) < 1e-6 and \
                rel(en32.cpu(), ref_eager[0].cpu()) < 1e-9
            _ = (E32 * 2).sum().item()
    run = graph.replay if graph is not None else step32
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(args.steps):
        run()
    ev1.record()
    torch.cuda.synchronize()
    step_ms = ev0.elapsed_time(ev1) / args.steps
    log(f"FP32 batched: step {step_ms * 1e3:.1f} us, chunk kernel {kernel_ms * 1e3:.1f} us")
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
            ctxs[m].energy_grad(ins[m][0], ins[m][1], flags=flags, out=outs[m])

    for _ in range(args.warmup):
        step64()
    torch.cuda.synchronize()
    ev0.record()
    for _ in range(args.steps):
        step64()
    ev1.record()
    torch.cuda.synchronize()
    step64_ms = ev0.elapsed_time(ev1) / args.steps
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
    # ---------------- roofline of the dominant kernel ----------------
    alg_bytes = Ne * (16 + 4 + 4) + Nn * (12 + 12 + 12)
    achieved = alg_bytes / (kernel_ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None}
    tp = os.path.join(ROOT, "profiles", "tet4_uneven_chunk_traffic.json")
    if os.path.exists(tp):
        try:
            roof["traffic"] = json.load(open(tp)).get("bytes_per_launch")
        except Exception:
            pass
    cpu = None
    if not args.no_cpu:
        cpu, cpu_out = cpu_baseline(W)
        err["cpu_fp64_vs_gpu_fp64_g_u"] = rel(gu64, cpu_out[0])
        log(f"CPU baseline: {cpu}")
    value = Ne / (step_ms * 1e-3) / 1e6
    v64 = Ne / (step64_ms * 1e-3) / 1e6
    out = {"metric": METRIC, "value": round(value, 2), "unit": "Melem/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(step_ms, 5), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "32 uneven TET4 elastic meshes (12^3..40^3 hexes x 6 Kuhn tets), residual + "
                                  "VJP wrt per-element E; chunk kernel + one reduction kernel per step, replayed as a "
                                  "HIP graph",
                      "meshes": nm, "tets": int(Ne), "nodes": int(Nn), "parallelism": "single GPU",
                      "chunk": {"elements_per_chunk": epc, "block": blk, "grid_per_cu": grid},
                      "graph": graph is not None, "graph_replays_match_eager": graph_ok},
           "fp64_plain": {"value": round(v64, 2), "unit": "Melem/s", "ms_per_step": round(step64_ms, 5),
                          "kernel_ms": round(k64, 5), "launches_per_step": nm},
           "speedup_vs_fp64_plain": round(value / v64, 2),
           "rel_err_vs_fp64": {k: float(f"{v:.3e}") for k, v in err.items()},
           "kernel_ms": round(kernel_ms, 5), "algorithmic_bytes_per_launch": int(alg_bytes),
           "roofline": roof, "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
