"""Benchmark: batched uneven anisotropic multigrid solves to 1e-8 relative residual.

Workload (one step): 512 independent anisotropic resistivity forward solves
div(sigma grad phi) = b, cubic edges drawn from {32, 48, 64, 96, 128},
log-normal conductivity with per-instance sigma_ln in [0.5, 2], 1-4 dipole
sources each; every solve is taken from its conductivities and right-hand side
(resident in HBM) to ||b - A x|| / ||b|| <= 1e-8: face coefficients, Galerkin
hierarchy, coarsest inverses, then BF16-storage V(2,2) cycles (FP32
arithmetic) preconditioning FP64 conjugate gradients, all in libhfm.so
work-table kernels, one launch per phase covering all 512 instances.

Printed: one JSON line with solves/s, roofline of the dominant kernel
(measured live with HIP events on the launch stream), the CPU baseline (the
oracle's restatement of the same solve on the host, bounded sample) and the
FP32 one-at-a-time GPU path on a sample of instances.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "ai-hfm-solver_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0
METRIC = "MG solves/s at 1e-8 rel. residual vs FP64"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def load_batch(mg, specs, dev, chunk=64):
    """Stage every instance's fields: drawn on the host from numpy's PCG64
    (hfm.workloads.host_fields, the oracle's draw order), 16 threads, ``chunk``
    instances at a time, then copied to the device."""
    import torch
    from hfm.workloads import host_fields_many
    for c0 in range(0, len(specs), chunk):
        part = specs[c0:c0 + chunk]
        for i, arrs in enumerate(host_fields_many(part, workers=16)):
            sx, sy, sz, b = (torch.from_numpy(a).to(dev) for a in arrs)
            mg.stage_sigma(c0 + i, sx, sy, sz, b)


def one_step(mg, persistent=False):
    """faces from sigma -> hierarchy -> solve, for every instance."""
    mg.build_faces()
    mg.setup()
    return mg.solve(persistent=persistent)


def kernel_time_ms(fn, reps, torch):
    """Average duration of the launches issued by fn() on torch's current stream."""
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    fn()
    e0.record(s)
    for _ in range(reps):
        fn()
    e1.record(s)
    e1.synchronize()
    return e0.elapsed_time(e1) / reps


def _cpu_solve(spec):
    """One oracle solve (worker process): the instance regenerated from its seed
    (the same PCG64 draw as the GPU's inputs), numpy FCG with FP32-storage V(2,2)."""
    from oracle import hfm_oracle as O
    shape, sig, nd, seed = spec
    sx, sy, sz, b = O.make_anisotropic_fields(shape, seed, sig, nd)
    op = O.harmonic_faces(sx, sy, sz)
    t0 = time.perf_counter()
    res = O.fcg_solve(op, b, tol=1e-8, storage="f32")
    return {"t": time.perf_counter() - t0, "iters": res.iters, "status": res.status,
            "x": res.x.astype(np.float64)}


def cpu_baseline_torch(specs, iters, nsample, threads, gpu_x):
    """The baseline plan's CPU restatement: the solve as unfused PyTorch-CPU FP32
    array operations (the operator as six shifted face products, Jacobi sweeps
    as separate passes, avg_pool3d restriction, repeat_interleave prolongation;
    oracle/torch_cpu.py) with ``threads`` torch threads, on the first
    ``nsample`` instances of the workload; the rate in fine-cell-iterations per
    second of solve time is scaled to the workload's sum(cells x iterations).
    The GPU solutions of those instances (``gpu_x``) are compared with it."""
    import torch
    from oracle import torch_cpu as TC
    from hfm.workloads import host_fields
    sample = specs[:nsample]
    prev = torch.get_num_threads()
    try:
        TC.cell_iterations_per_s(sample[:1], threads, host_fields)     # warm
        res = TC.cell_iterations_per_s(sample, threads, host_fields)
    finally:
        torch.set_num_threads(prev)
    if not all(o["converged"] for o in res["solves"]):
        raise SystemExit("bench: a CPU-restatement solve did not converge")
    total = sum(math.prod(s[0]) * int(k) for s, k in zip(specs, iters))
    diffs, it_diff = [], []
    for q, o in enumerate(res["solves"]):
        if q in gpu_x:
            xg, kg = gpu_x[q]
            diffs.append(float(np.linalg.norm(xg - o["x"]) / np.linalg.norm(o["x"])))
            it_diff.append(int(kg) - int(o["iters"]))
    edges = sorted({sp[0][0] for sp in sample})
    return {"value": round(len(specs) * res["rate"] / total, 3), "unit": "solves/s",
            "cores": threads, "kind": "port",
            "sample": f"CPU restatement of the solve path as unfused PyTorch-CPU FP32 array ops "
                      f"(FP64 FCG around an FP32 V(2,2)) with {threads} torch threads, on the "
                      f"first {len(sample)} instances of the workload (edges {edges}), "
                      f"{res['seconds']:.1f} s of solves, {res['rate'] / 1e6:.2f} M "
                      f"fine-cell-iterations/s, scaled to the workload's {total / 1e9:.2f} G "
                      f"cell-iterations",
            "gpu_vs_cpu_max_rel_diff": max(diffs) if diffs else None,
            "iters_gpu_minus_cpu": sorted(set(it_diff))}


def cpu_baseline(specs, iters, nsample, workers, gpu_x):
    """The oracle's solve (numpy FCG + FP32-storage V(2,2), unfused array ops) on
    ``workers`` host processes: the first ``nsample`` instances of the workload,
    all to completion; the rate in fine-cell-iterations/s (wall clock) is
    scaled to the whole workload's sum(cells x iterations) to give solves/s
    for the same 512-instance job.  The GPU solutions of those instances
    (``gpu_x``) are compared with the oracle's."""
    import multiprocessing as mp
    sample = specs[:nsample]
    saved = {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS")}
    for k in saved:
        os.environ[k] = "1"          # one thread per worker process
    try:
        with mp.get_context("spawn").Pool(workers) as pool:
            t0 = time.perf_counter()
            outs = pool.map(_cpu_solve, sample, chunksize=1)
            wall = time.perf_counter() - t0
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    if any(o["status"] != "converged" for o in outs):
        raise SystemExit("bench: an oracle solve of the CPU baseline did not converge")
    work = sum(math.prod(sp[0]) * o["iters"] for sp, o in zip(sample, outs))
    rate = work / wall
    total = sum(math.prod(s[0]) * int(k) for s, k in zip(specs, iters))
    diffs, it_diff = [], []
    for q, o in enumerate(outs):
        if q in gpu_x:
            xg, kg = gpu_x[q]
            diffs.append(float(np.linalg.norm(xg - o["x"]) / np.linalg.norm(o["x"])))
            it_diff.append(int(kg) - int(o["iters"]))
    edges = sorted({sp[0][0] for sp in sample})
    return {"value": round(len(specs) * rate / total, 3), "unit": "solves/s", "cores": workers,
            "kind": "port",
            "sample": f"the oracle's numpy FCG (FP32-storage V(2,2)) on the first {len(sample)} "
                      f"instances of the workload (edges {edges}), {workers} worker processes, "
                      f"{wall:.1f} s wall, {rate / 1e6:.2f} M fine-cell-iterations/s, scaled to "
                      f"the workload's {total / 1e9:.2f} G cell-iterations",
            # the GPU solved these instances with the bench's storage, the oracle
            # here with FP32 storage: both to 1e-8 on the same FP32 operator
            "gpu_vs_oracle_f32_max_rel_diff": max(diffs) if diffs else None,
            "iters_gpu_minus_oracle_f32": sorted(set(it_diff))}


def fp32_one_at_a_time(specs, dev, torch):
    """The plain FP32 path, one instance at a time, on a sample: (solves/s, mean cells)."""
    from hfm.mg import MGBatch
    from hfm.workloads import host_fields
    t_tot = 0.0
    for (shape, sig, nd, seed) in specs:
        sx, sy, sz, b = (torch.from_numpy(a).to(dev) for a in host_fields(shape, sig, nd, seed))
        mg = MGBatch([shape], storage="f32", tol=1e-8, device=dev)
        mg.stage_sigma(0, sx, sy, sz, b)
        one_step(mg)   # warm: graph capture
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        info = one_step(mg)
        torch.cuda.synchronize()
        t_tot += time.perf_counter() - t0
        assert int(info["status"][0]) == 1
    return len(specs) / t_tot


def true_rel_residual(inst, b, x, torch):
    """||b - A x||_2 / ||b||_2 in FP64 with A the instance's stored FP32 face
    operator (low faces cx/cy/cz, high boundary faces bxh/byh/bzh), with torch
    array ops apart from the solver's kernels."""
    D, H, W = inst.shape
    N = D * H * W
    f, o = inst.face32[0], inst.face_off[0]
    cx, cy, cz = (f[o[q]:o[q] + N].view(D, H, W).double() for q in range(3))
    bxh = f[o[3]:o[3] + D * H].view(D, H).double()
    byh = f[o[4]:o[4] + D * W].view(D, W).double()
    bzh = f[o[5]:o[5] + H * W].view(H, W).double()
    ex = torch.cat([cx[:, :, 1:], bxh[:, :, None]], 2)
    ey = torch.cat([cy[:, 1:, :], byh[:, None, :]], 1)
    ez = torch.cat([cz[1:], bzh[None]], 0)
    x = x.reshape(D, H, W).double()
    b = b.reshape(-1)[:N].view(D, H, W).double()
    ax = (cx + ex + cy + ey + cz + ez) * x
    ax[:, :, 1:] -= cx[:, :, 1:] * x[:, :, :-1]
    ax[:, :, :-1] -= ex[:, :, :-1] * x[:, :, 1:]
    ax[:, 1:, :] -= cy[:, 1:, :] * x[:, :-1, :]
    ax[:, :-1, :] -= ey[:, :-1, :] * x[:, 1:, :]
    ax[1:] -= cz[1:] * x[:-1]
    ax[:-1] -= ez[:-1] * x[1:]
    return float((b - ax).norm() / b.norm())


def true_residuals(mg, torch):
    """Per instance, ||b - A x||_2 / ||b||_2 in FP64 with A the stored FP32 face
    operator, computed here with torch array ops (independent of the solver's
    kernels and of its recursive residual), and the largest relative
    difference between the stored faces and an FP64 harmonic mean of the staged
    conductivities (boundary face = the cell's own value) rounded to FP32."""
    rels, face_dev = [], 0.0
    for inst in mg.insts:
        D, H, W = inst.shape
        N = D * H * W
        rels.append(true_rel_residual(inst, inst.b32, inst.solution(), torch))
        f, o = inst.face32[0], inst.face_off[0]
        cx, cy, cz = (f[o[q]:o[q] + N].view(D, H, W).double() for q in range(3))
        N8 = inst.sigma.numel() // 3
        sg = [inst.sigma[q * N8:q * N8 + N].view(D, H, W).double() for q in range(3)]
        hm = lambda a, c: 2.0 * a * c / (a + c)  # noqa: E731
        ref = (torch.cat([sg[0][:, :, :1], hm(sg[0][:, :, 1:], sg[0][:, :, :-1])], 2),
               torch.cat([sg[1][:, :1, :], hm(sg[1][:, 1:, :], sg[1][:, :-1, :])], 1),
               torch.cat([sg[2][:1], hm(sg[2][1:], sg[2][:-1])], 0))
        for got, want in zip((cx, cy, cz), ref):
            want32 = want.float().double()
            face_dev = max(face_dev, float(((got - want32).abs() / want32.abs()).max()))
    return np.array(rels), face_dev


def serving(args, dev, torch):
    """Poisson stream of anisotropic solves, edges 32..256 log-uniform (multiples of
    16), sigma_ln = 1, one dipole: continuous batching (BF16 storage, 200 GB budget)
    vs the plain FP32 path serving one request at a time.  Two offered loads: the
    stated arrival rate (latency under that load) and a saturating rate (every
    request arrives within ~0.1 s: each server's capacity in solves/s)."""
    from hfm.mg import Params
    from hfm.server import SolveServer, latency_stats, poisson_stream, run_stream
    from hfm.workloads import device_fields
    mk = lambda shape, seed: device_fields(shape, 1.0, 1, seed, dev)  # noqa: E731
    out = {}
    servers = (("batched_bf16", Params(storage="bf16", tol=1e-8), 4096),
               ("fp32_one_at_a_time", Params(storage="f32", tol=1e-8), 1))
    for load, rate in (("at_rate", args.serve_rate), ("saturated", args.serve_sat_rate)):
        stream = poisson_stream(args.serve_requests, rate, args.seed + 1)
        res_load = {}
        for name, params, cap in servers:
            srv = SolveServer(params, hbm_budget=200e9, capacity=cap, device=dev)
            warm = SolveServer(params, hbm_budget=200e9, capacity=cap, device=dev)
            warm.submit(-1, *mk((64, 64, 64), 1))
            warm.drain()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            res = run_stream(srv, stream, mk)
            torch.cuda.synchronize()
            wall = time.perf_counter() - t0
            bad = [r for r in res if r.status != "converged" or r.rel_residual > 1e-8]
            if bad:
                raise SystemExit(f"serving {name}: {len(bad)} requests failed")
            st = latency_stats(res)
            res_load[name] = {"solves_per_s": round(len(res) / wall, 2),
                              "p50_ms": round(st["p50_ms"], 2), "p99_ms": round(st["p99_ms"], 2),
                              "requests": len(res), "peak_resident_GB": round(srv.peak_used / 1e9, 2),
                              "epochs": srv.epochs}
            log(f"bench: serving {load} {name}: {res_load[name]}")
        res_load["arrival_rate_per_s"] = rate
        res_load["speedup_solves_per_s"] = round(res_load["batched_bf16"]["solves_per_s"] /
                                                 res_load["fp32_one_at_a_time"]["solves_per_s"], 2)
        out[load] = res_load
    return out


def cpu_baseline_config0(sx, sy, sz, b):
    """The oracle's stationary cycle on the same 64^3 problem, one host core."""
    from oracle import hfm_oracle as O
    t0 = time.perf_counter()
    ref = O.solve_refined(O.harmonic_faces(sx, sy, sz).astype("f32"), b, tol=1e-6, storage="f32",
                          inner_cycles=1, alpha=1.0, max_outer=400)
    return ref, time.perf_counter() - t0


def config0(args, dev, torch):
    """The plain method of the CPU-runnable configuration: one 64^3 solve
    (sigma_ln 1, one dipole, seed 0), stationary V(2,2) cycles (omega 2/3,
    nearest prolongation, alpha 1) with an FP32 V-cycle to 1e-6, as FP64 defect
    correction (method refine); GPU wall time from staged fields (faces,
    hierarchy, solve; median of 5 after a warm run) and the oracle's restatement
    of the same cycle on one host core (numpy, its cycle count checked: this
    leg's CPU baseline)."""
    from hfm.mg import MGBatch
    from hfm.workloads import host_fields
    shape = (64, 64, 64)
    sx, sy, sz, b = host_fields(shape, 1.0, 1, 0)
    mg = MGBatch([shape], storage="f32", tol=1e-6, maxit=400, alpha=1.0, method="refine",
                 device=dev)
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    mg.stage_sigma(0, T(sx), T(sy), T(sz), T(b))
    one_step(mg)
    torch.cuda.synchronize()
    ts = []
    for _ in range(5):
        t0 = time.perf_counter()
        info = one_step(mg)
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    rel, _ = true_residuals(mg, torch)
    ref, tcpu = cpu_baseline_config0(sx, sy, sz, b)
    out = {"grid": list(shape), "method": "FP64 defect correction, FP32 V(2,2), alpha 1",
           "tol": 1e-6, "ms_per_solve": round(float(np.median(ts)) * 1e3, 3),
           "vcycles": int(info["iter"][0]), "oracle_vcycles": int(ref.vcycles),
           "status": int(info["status"][0]), "true_rel_residual_fp64": float(rel[0]),
           "gpu_vs_oracle_rel_diff": float(np.linalg.norm(mg.solution(0).cpu().numpy() - ref.x)
                                           / np.linalg.norm(ref.x)),
           "cpu_oracle_s": round(tcpu, 3), "cpu_cores": 1,
           "speedup_vs_cpu_oracle": round(tcpu / float(np.median(ts)), 1)}
    log(f"bench: config0 {out}")
    return out


def small_batches(args, dev, torch, sizes=(1, 8, 32)):
    """Launch-bound batches: n instances of edges 32..64 solved with graph-replayed
    launches and with the persistent kernel, the batch's wall time from staged
    fields (faces, hierarchy, solve) per mode, median of ``reps`` after a warm
    run.  The two modes give bitwise the same solutions."""
    from hfm.mg import MGBatch
    from hfm.workloads import batched_specs
    out = {}
    for n in sizes:
        specs = batched_specs(n, args.seed + 7, edges=(32, 48, 64))
        mg = MGBatch([s[0] for s in specs], storage=args.storage, tol=1e-8, maxit=300,
                     device=dev, iters_per_graph=args.iters_per_graph, shrink=args.shrink,
                     outer=args.outer, fuse=not args.no_fuse, streams=args.streams,
                     **_fuse_kw(args))
        load_batch(mg, specs, dev)
        row = {"cells": int(sum(math.prod(s[0]) for s in specs))}
        xs = {}
        for mode in (False, True):
            one_step(mg, persistent=mode)
            torch.cuda.synchronize()
            ts = []
            for _ in range(7):
                t0 = time.perf_counter()
                info = one_step(mg, persistent=mode)
                torch.cuda.synchronize()
                ts.append(time.perf_counter() - t0)
            xs[mode] = [mg.solution(q).clone() for q in range(n)]
            row["persistent_ms" if mode else "graphs_ms"] = round(float(np.median(ts)) * 1e3, 3)
            row["iters_max"] = int(info["iter"].max())
            if not np.all(info["status"] == 1):
                raise SystemExit(f"small batch {n}: not converged")
        row["bitwise_equal"] = all(bool((a == b).all()) for a, b in zip(xs[False], xs[True]))
        row["speedup_persistent"] = round(row["graphs_ms"] / row["persistent_ms"], 2)
        out[str(n)] = row
        log(f"bench: small batch {n}: {row}")
        del mg
    return out


def batched_fp32(specs, dev, torch, storage="f32"):
    """The same 512-instance batch with FP32 storage (the configuration's plain
    run), or another storage type."""
    from hfm.mg import MGBatch
    mg = MGBatch([s[0] for s in specs], storage=storage, tol=1e-8, maxit=300, device=dev)
    load_batch(mg, specs, dev)
    one_step(mg)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    info = one_step(mg)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    out = {"solves_per_s": round(len(specs) / t, 2), "ms_per_step": round(t * 1e3, 2),
           "iters_mean": round(float(info["iter"].mean()), 2), "iters_max": int(info["iter"].max()),
           "resident_GiB": round(mg.nbytes() / 2 ** 30, 2),
           "converged": int(np.sum(info["status"] == 1))}
    if storage != "f32":
        rel, _ = true_residuals(mg, torch)
        out["max_true_rel_residual_fp64"] = float(rel.max())
    elif not np.all(info["status"] == 1):
        raise SystemExit("bench: FP32 batch did not converge")
    log(f"bench: batched {storage} {out}")
    return out


L_STATUS = {0: "running", 1: "converged", 2: "maxit", 3: "breakdown"}


def urban(args, dev, torch):
    """Urban flow on 256 x 256 x 64 (D = 64 heights): seeded blocks covering 20 % of
    the ground, log-law inflow, the reference's explicit 27-point momentum and
    scalar transport, and a 7-point MG pressure projection to 1e-8 per step;
    BF16-storage V-cycles (z and p in BF16, or in FP32) vs FP32, with the final
    pressure of each compared with
    an FP64-converged (1e-12) solve of the same system."""
    from hfm.timestep import UrbanProjection
    from hfm.workloads import log_law_profile, urban_blocks
    sh = (64, 256, 256)
    dt = 0.05       # CFL ~0.4 for the log-law inflow (up to 7.9 at the top)
    mask = urban_blocks(sh, args.seed)
    sig = torch.from_numpy(np.where(mask, 1e5, 0.0).astype(np.float32)).to(dev)
    prof = log_law_profile(sh[0])
    out = {"grid": list(sh), "steps": args.urban_steps, "dt": dt,
           "ground_cover": round(float(mask[0].mean()), 3)}
    finals = {}
    variants = {"bf16": ("bf16", None), "bf16_fp32_zp": ("bf16", "f32"), "f32": ("f32", None),
                "f16": ("f16", None), "f16_fp32_zp": ("f16", "f32")}
    for name, (storage, outer) in variants.items():
        flow = UrbanProjection(sh, sigma=sig, inflow=prof, dt=dt, storage=storage, outer=outer,
                               tol=1e-8, device=dev)
        u0 = torch.from_numpy(np.broadcast_to(prof[:, None, None], sh).copy()).to(dev)
        z = torch.zeros(sh, device=dev)
        state = (u0, z, z.clone(), z.clone(), z.clone())
        state = flow.step(*state)          # warm (graph capture)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        its = []
        failed = None
        for _ in range(args.urban_steps):
            state = flow.step(*state, donate=True)
            its.append(int(flow.last_info["iter"][0]))
            if int(flow.last_info["status"][0]) != 1:
                failed = L_STATUS[int(flow.last_info["status"][0])]
                break
        torch.cuda.synchronize()
        if failed is not None:
            if storage != "f16":
                raise SystemExit(f"bench: urban {name} pressure solve failed: {flow.last_info}")
            out[name] = {"failed": failed, "after_steps": len(its)}
            log(f"bench: urban {name}: the pressure solve stopped ({failed})")
            continue
        tstep = (time.perf_counter() - t0) / args.urban_steps
        out[name] = {"ms_per_step": round(tstep * 1e3, 3),
                        "cell_steps_per_s": round(float(np.prod(sh)) / tstep, 1),
                        "mg_iters_mean": round(float(np.mean(its)), 2),
                        "rel_residual": float(flow.last_info["rel"][0])}
        inst = flow.solver.insts[0]
        finals[name] = (inst.b32.clone(), inst.solution().clone())
        out[name]["true_rel_residual_fp64"] = true_rel_residual(inst, inst.b32,
                                                                inst.solution(), torch)
        if name == "f32":
            ref_solver = UrbanProjection(sh, sigma=sig, inflow=prof, dt=dt, storage="f32",
                                         tol=1e-12, device=dev).solver
    for name, (b, x) in finals.items():
        ref_solver.insts[0].b32.copy_(b)
        ref_solver.solve()
        xr = ref_solver.insts[0].solution()
        out[name]["pressure_rel_err_vs_fp64"] = float((x - xr).norm() / xr.norm())
    log(f"bench: urban {out}")
    return out


def kolmogorov(args, dev, torch):
    """2D Kolmogorov flow, 256^2 periodic, nu = 1/100, f_x = sin 4y, dt = 0.005:
    one flow for ``--flow-steps`` steps (the configuration's run is 1000), an
    ensemble of ``--flow-ensemble`` flows in one launch, and the oracle's step
    on one host core over a bounded sample."""
    from hfm.flow2d import KolmogorovFlow
    from hfm.workloads import random_modes_ic
    N = 256
    out = {"grid": [N, N], "nu": 0.01, "dt": 0.005, "tol": 1e-6, "levels": 6}
    fl = KolmogorovFlow(N=N, nflows=1, team=args.flow_team, device=dev)
    u0, v0 = random_modes_ic(N, args.seed)
    fl.set_state(0, torch.from_numpy(u0), torch.from_numpy(v0))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    done = 0
    while done < args.flow_steps:            # progress every 250 steps
        k = min(250, args.flow_steps - done)
        fl.step(k)
        torch.cuda.synchronize()
        done += k
        log(f"bench: kolmogorov step {done}")
    t1 = time.perf_counter() - t0
    st = fl.stats()[0]
    if int(st["status"]) != 1 or int(st["steps"]) != args.flow_steps:
        raise SystemExit(f"bench: kolmogorov flow stopped: {st}")
    div = fl.divergence(0)
    u, v, _ = fl.state(0)
    out["single_flow"] = {
        "steps": args.flow_steps, "workgroups": fl.team, "wall_s": round(t1, 3),
        "ms_per_step": round(t1 / args.flow_steps * 1e3, 4),
        "steps_per_s": round(args.flow_steps / t1, 1),
        "cell_steps_per_s": round(N * N * args.flow_steps / t1, 1),
        "substeps": int(st["substeps"]),
        "cg_iters_per_substep": round(float(st["iters_sum"] / st["substeps"]), 3),
        "max_rel_residual": float(st["rel_max"]),
        "max_abs_divergence": float(div.abs().max()),
        "kinetic_energy": float(0.5 * (u.double() ** 2 + v.double() ** 2).mean())}
    if args.flow_ensemble > 0:
        ens = KolmogorovFlow(N=N, nflows=args.flow_ensemble, device=dev)
        for q in range(args.flow_ensemble):
            u, v = random_modes_ic(N, args.seed + q)
            ens.set_state(q, torch.from_numpy(u), torch.from_numpy(v))
        ens.step(1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ens.step(args.flow_ensemble_steps)
        torch.cuda.synchronize()
        t2 = time.perf_counter() - t0
        es = ens.stats()
        if not np.all(es["status"] == 1):
            raise SystemExit("bench: kolmogorov ensemble flow stopped")
        fs = args.flow_ensemble * args.flow_ensemble_steps / t2
        out["ensemble"] = {"flows": args.flow_ensemble, "steps": args.flow_ensemble_steps,
                           "flow_steps_per_s": round(fs, 1),
                           "cell_steps_per_s": round(fs * N * N, 1),
                           "resident_GiB": round(ens.nbytes() / 2 ** 30, 2)}
        del ens
    if not args.no_cpu_baseline:
        # the oracle's substeps from the run's first and last states (the early
        # flow is slower and needs fewer substeps), scaled to the run's substeps
        from oracle import hfm_flow2d as F
        mg = F.build_mg(N)
        fu = F.forcing_u(N)
        uf, vf, pf = (t.cpu().numpy() for t in fl.state(0))
        nsub, tc = 0, 0.0
        for (us_, vs_, ps_) in ((u0, v0, np.zeros((N, N))), (uf, vf, pf)):
            stc = F.FlowState(us_.copy(), vs_.copy(), ps_.copy())
            t0 = time.perf_counter()
            k = 0
            while k < 5 or time.perf_counter() - t0 < args.cpu_budget / 6:
                nsub += F.step(stc, mg, 0.005, 0.01, fu)[0]
                k += 1
            tc += time.perf_counter() - t0
        cpu_sps = args.flow_steps / (int(st["substeps"]) * tc / nsub)
        out["cpu_baseline"] = {"value": round(cpu_sps, 2), "unit": "steps/s", "cores": 1,
                               "kind": "port",
                               "sample": f"oracle (numpy) steps from the run's first and last "
                                         f"states: {nsub} substeps in {tc:.1f} s, scaled to the "
                                         f"run's {int(st['substeps'])} substeps"}
        out["speedup_vs_cpu"] = round(out["single_flow"]["steps_per_s"] / cpu_sps, 1)
    log(f"bench: kolmogorov {out}")
    return out


def reference_context(args, dev, torch):
    """The reference timings taken at this build's algorithm and size
    (hardware differs; context, not the headline metric):

    * 1000 reference sawtooth cycles on 128^3, constant-coefficient 7-point
      (Dx, Dy, Dz) = (200, 100, 50), 8 levels to 1^3:
      295.16 s (ART/Anisotropic_3d_Multigrid_7point.ipynb:395, hardware not
      recorded) -- here layers.SawtoothMG, the same cycle;
    * the urban step on 512^3 with 5 sawtooth cycles of the 27-point Poisson
      layer per step: 5.06 s/step (UM/AI_HFM_v3.ipynb:2508, RTX A5000) --
      here timestep.UrbanFlow, the same step (seeded blocks: the reference's
      mesh file is not in its repository);
    * the INHALE demo step (Petrov-Galerkin dissipation) on 256 x 4096 x 4096:
      4000 steps in 22580.29 s (UM/INHALE_demo.ipynb cell 24 output, hardware
      not recorded) -- here timestep.InhaleFlow on a 128 x 2048 x 2048 box,
      compared per cell-step."""
    from hfm import layers as Ly
    from hfm.timestep import UrbanFlow
    from hfm.workloads import urban_blocks
    out = {}
    if args.ref_cycles > 0:
        K = np.zeros((3, 3, 3), np.float32)
        K[0, 1, 1] = K[2, 1, 1] = -50.0
        K[1, 0, 1] = K[1, 2, 1] = -100.0
        K[1, 1, 0] = K[1, 1, 2] = -200.0
        K[1, 1, 1] = 2.0 * (200.0 + 100.0 + 50.0)
        mg = Ly.SawtoothMG([Ly.FixedConv(K, (n, n, n)) for n in (128, 64, 32, 16, 8, 4, 2, 1)])
        b = torch.zeros((1, 128, 128, 128, 1), device=dev)
        b[0, 40, 64, 64, 0], b[0, 88, 64, 64, 0] = 1.0, -1.0
        x = torch.zeros_like(b)
        for _ in range(3):
            x = mg.cycle(x, b)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.ref_cycles):
            x = mg.cycle(x, b)
        torch.cuda.synchronize()
        per = (time.perf_counter() - t0) / args.ref_cycles
        out["sawtooth_128"] = {"cycles": args.ref_cycles, "ms_per_cycle": round(per * 1e3, 4),
                               "s_per_1000_cycles": round(per * 1e3, 3),
                               "fine_cell_cycles_per_s": round(128 ** 3 / per, 1),
                               "reference_s_per_1000_cycles": 295.16,
                               "ratio_vs_reference": round(295.16 / (per * 1e3), 1)}
    if args.ref_urban_steps > 0:
        sh = (512, 512, 512)
        mask = urban_blocks(sh, args.seed)
        sig = torch.from_numpy(np.where(mask, 1e5, 0.0).astype(np.float32)).to(dev)
        del mask
        flow = UrbanFlow(sh, sigma=sig, multi_itr=5, device=dev)
        z = torch.zeros(sh, device=dev)
        state = (torch.ones(sh, device=dev), z, z.clone(), z.clone(), z.clone())
        state = flow.step(*state)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.ref_urban_steps):
            state = flow.step(*state, donate=True)
        torch.cuda.synchronize()
        per = (time.perf_counter() - t0) / args.ref_urban_steps
        out["urban_512"] = {"steps": args.ref_urban_steps, "mg_cycles_per_step": 5,
                            "ms_per_step": round(per * 1e3, 2),
                            "cell_steps_per_s": round(512 ** 3 / per, 1),
                            "reference_s_per_step": 5.06, "ratio_vs_reference": round(5.06 / per, 1),
                            "finite": bool(all(torch.isfinite(t).all() for t in state))}
        del flow, state, sig, z
        torch.cuda.empty_cache()
    if args.ref_inhale_steps > 0:
        # the INHALE demo's step (Petrov-Galerkin dissipation) on a quarter-size
        # box of its 256 x 4096 x 4096 grid (same 1:16 aspect ratio; the full
        # grid's fields and temporaries exceed one GPU's memory)
        from hfm.timestep import InhaleFlow
        sh = (128, 2048, 2048)
        mask = urban_blocks(sh, args.seed, foot=(16, 48), height=(10, 60))
        sig = torch.from_numpy(np.where(mask, 1e5, 0.0).astype(np.float32)).to(dev)
        del mask
        flow = InhaleFlow(sh, sigma=sig, multi_itr=5, device=dev)
        z = torch.zeros(sh, device=dev)
        state = (torch.ones(sh, device=dev), z, z.clone(), z.clone(), z.clone())
        state = flow.step(*state)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.ref_inhale_steps):
            state = flow.step(*state, donate=True)
        torch.cuda.synchronize()
        per = (time.perf_counter() - t0) / args.ref_inhale_steps
        cells = float(np.prod(sh))
        ref_rate = 256 * 4096 * 4096 / (22580.29 / 4000)
        out["inhale_128x2048x2048"] = {
            "steps": args.ref_inhale_steps, "mg_cycles_per_step": 5,
            "ms_per_step": round(per * 1e3, 2), "cell_steps_per_s": round(cells / per, 1),
            "reference_cell_steps_per_s": round(ref_rate, 1),
            "ratio_vs_reference_per_cell": round(cells / per / ref_rate, 2),
            "finite": bool(all(torch.isfinite(t).all() for t in state[:4]))}
        del flow, state, sig, z
        torch.cuda.empty_cache()
    log(f"bench: reference context {out}")
    return out


def profile_iteration(mg, cells, torch):
    """Time every launch of one FCG iteration separately (all instances active)."""
    from hfm import _lib as L
    plan, pool = mg.plan, mg.pool
    s = L.stream_handle(mg.device)
    names = {0: "zero", 1: "plain", 2: "correct"}
    rows = []
    ops = [("down", op) for op in plan.ops_down] + [("up", o) for ups in plan.ops_up for o in ups]
    for leg, op in ops:
        kind, mode, lev = op[0], op[1], op[5]
        label = f"{leg} L{lev} " + {"restrict": "restrict",
                                    "smooth_restrict": "sweep plain + restrict (fused)",
                                    "smooth2": "sweep correct + plain (fused)"}.get(
                                        kind, "sweep " + names.get(mode, ""))
        label += " + r.z partials" if op[4] else ""
        rows.append((label, kernel_time_ms(lambda op=op: plan._launch(op, s), 5, torch)))
    rows.append((f"coarse cycle (levels >= {plan.chain_levels})",
                 kernel_time_ms(lambda: plan.launch_coarse_cycle(s), 5, torch)))
    rows.append(("fcg dir", kernel_time_ms(lambda: plan.launch_dir(0, s), 5, torch)))
    upd = []
    for lag in (1, 0):
        mg.set_xlag(lag)
        upd.append(kernel_time_ms(lambda: plan.launch_update(0, s), 5, torch))
    mg.set_xlag(0)
    rows.append(("fcg update, mean of the two variants", 0.5 * (upd[0] + upd[1])))
    rows.append(("fcg flush", kernel_time_ms(lambda: plan.launch_flush(0, s), 5, torch)))
    rows.append(("finalize", kernel_time_ms(lambda: pool.finalize("check", s), 5, torch)))
    total = sum(t for _, t in rows)
    rows += [("  fcg update with x (odd iterations)", upd[0]),
             ("  fcg update without x (even iterations)", upd[1])]
    print(json.dumps({"fine_cells": cells, "iteration_ms": round(total, 3),
                      "launches": {k: round(v, 4) for k, v in rows}}), flush=True)


def profile_flow2d(args, dev, torch):
    """The Kolmogorov flow alone (for kernel traces): one 256^2 flow for
    ``--flow-steps`` steps, then ``--flow-ensemble`` flows for
    ``--flow-ensemble-steps`` steps; prints one JSON line."""
    from hfm.flow2d import KolmogorovFlow
    from hfm.workloads import random_modes_ic
    N, out = 256, {"workload": "kolmogorov_2d"}
    for name, nflows, steps in (("single_flow", 1, args.flow_steps),
                                ("ensemble", args.flow_ensemble, args.flow_ensemble_steps)):
        if nflows < 1 or steps < 1:
            continue
        fl = KolmogorovFlow(N=N, nflows=nflows, team=args.flow_team if nflows == 1 else None,
                            device=dev)
        for q in range(nflows):
            u, v = random_modes_ic(N, args.seed + q)
            fl.set_state(q, torch.from_numpy(u), torch.from_numpy(v))
        s = torch.cuda.current_stream()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(s)
        fl.step(steps)
        e1.record(s)
        e1.synchronize()
        st = fl.stats()
        out[name] = {"flows": nflows, "steps": steps, "launch_ms": round(e0.elapsed_time(e1), 3),
                     "ms_per_step": round(e0.elapsed_time(e1) / steps, 4),
                     "substeps": int(st["substeps"].sum()),
                     "cg_iters": int(st["iters_sum"].sum())}
        del fl
    print(json.dumps(out), flush=True)


def profile_urban(args, dev, torch):
    """The urban projection step alone (for kernel traces): ``--urban-steps``
    steps of the FP16 (the default), FP32 and BF16 + FP32-z/p variants;
    prints one JSON line."""
    from hfm.timestep import UrbanProjection
    from hfm.workloads import log_law_profile, urban_blocks
    sh = (64, 256, 256)
    mask = urban_blocks(sh, args.seed)
    sig = torch.from_numpy(np.where(mask, 1e5, 0.0).astype(np.float32)).to(dev)
    prof = log_law_profile(sh[0])
    out = {"workload": "urban_projection", "grid": list(sh)}
    for name, (storage, outer) in {"f16": ("f16", None), "f32": ("f32", None),
                                   "bf16_fp32_zp": ("bf16", "f32")}.items():
        flow = UrbanProjection(sh, sigma=sig, inflow=prof, dt=0.05, storage=storage, outer=outer,
                               tol=1e-8, device=dev)
        u0 = torch.from_numpy(np.broadcast_to(prof[:, None, None], sh).copy()).to(dev)
        z = torch.zeros(sh, device=dev)
        state = (u0, z, z.clone(), z.clone(), z.clone())
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        its = []
        for _ in range(args.urban_steps):
            state = flow.step(*state, donate=True)
            its.append(int(flow.last_info["iter"][0]))
        torch.cuda.synchronize()
        out[name] = {"steps": args.urban_steps,
                     "ms_per_step": round((time.perf_counter() - t0) / args.urban_steps * 1e3, 3),
                     "mg_iters": its}
    print(json.dumps(out), flush=True)


def profile_kernel(args, dev, torch):
    """Launch only one kernel (all instances active) and a copy of known size,
    for rocprofv3 kernel-trace / PMC runs; prints one JSON line."""
    import numpy as np
    from hfm.mg import MGBatch
    from hfm.workloads import batched_specs
    if args.profile_kernel == "flow2d":
        return profile_flow2d(args, dev, torch)
    if args.profile_kernel == "urban":
        return profile_urban(args, dev, torch)
    if args.profile_kernel == "reference":
        # the reference-context runs selected by --ref-cycles / --ref-urban-steps
        # / --ref-inhale-steps, alone
        print(json.dumps(reference_context(args, dev, torch)), flush=True)
        return None
    if args.single_shape:
        shape = tuple(int(v) for v in args.single_shape.split(","))
        specs = [(shape, 1.0, 1, args.seed)]
    else:
        specs = batched_specs(args.instances, args.seed)
    cells = sum(math.prod(s[0]) for s in specs)
    mg = MGBatch([s[0] for s in specs], storage=args.storage, tol=1e-8, device=dev,
                 outer=args.outer, fuse=not args.no_fuse, **_fuse_kw(args))
    load_batch(mg, specs, dev)
    mg.build_faces()
    mg.setup()
    mg.restart()
    if args.profile_kernel == "all":
        profile_iteration(mg, cells, torch)
        if not args.no_tune:
            from hfm.tune import tune
            choice = tune(mg)
            print(json.dumps({"coarse_cycle_ms": {str(k): round(v, 4) for k, v in mg.coarse_times.items()},
                              "coarse_cells": choice["coarse_cells"],
                              "tile_planes": choice["tile_planes"],
                              "vcycle_ms_by_tile_planes": {str(k): round(v, 4)
                                                           for k, v in mg.tile_times.items()},
                              "tune_ms": {"/".join(map(str, k)): {str(b): round(t, 4) for b, t in v.items()}
                                          for k, v in sorted(mg.tune_times.items())},
                              "blocks": {"/".join(map(str, k)): v for k, v in sorted(choice["blocks"].items())}}),
                  flush=True)
            mg.restart()
            profile_iteration(mg, cells, torch)
        return None
    if args.profile_kernel == "update":
        # the two variants alternate: with the deferred solution update, without
        times, per = [], []
        for lag in (1, 0):
            mg.set_xlag(lag)
            times.append(kernel_time_ms(lambda: mg.launch_update(0), args.reps, torch))
            per.append(mg.update_bytes_per_cell(lag))
        ms, per_cell = 0.5 * sum(times), {"update": 0.5 * sum(per)}
        variants = {"with_x_ms": round(times[0], 4), "without_x_ms": round(times[1], 4),
                    "bytes_per_cell": per}
    else:
        variants = {}
        per_cell = {"sweep": mg.fine_pre_bytes_per_cell()}
        variants = {"launch": mg.fine_pre_kernel()}
        ms = kernel_time_ms(mg.launch_fine_pre, args.reps, torch)
    nbytes = 1 << 30
    src = torch.empty(nbytes // 4, dtype=torch.float32, device=dev).fill_(1.0)
    dst = torch.empty_like(src)
    cp_ms = kernel_time_ms(lambda: dst.copy_(src), 5, torch)
    alg = cells * per_cell[args.profile_kernel]
    print(json.dumps({"kernel": args.profile_kernel, "instances": len(specs), "fine_cells": cells,
                      "alg_bytes_per_launch": alg, "launch_ms": round(ms, 4), **variants,
                      "achieved_GBs": round(alg / (ms * 1e-3) / 1e9, 1), "reps": args.reps + 1,
                      "copy_bytes_each_way": nbytes, "copy_ms": round(cp_ms, 4),
                      "copy_GBs_read_plus_write": round(2 * nbytes / (cp_ms * 1e-3) / 1e9, 1)}),
          flush=True)


# HBM bytes per k_fcg_update launch of the headline batch (512 instances, seed
# 2024, BF16 storage and search directions), by variant (with / without the
# deferred solution update), from rocprofv3 --pmc runs of ``bench.py
# --traffic --traffic-lag {1,0}`` (profiles/k_fcg_update_traffic_batched512.md):
# FETCH_SIZE x 2.000 + WRITE_SIZE x 1.000 in KiB, the factors being what a
# 1 GiB device copy reads on each counter (524299.5 KiB fetched, 1048576.0 KiB
# written).
UPDATE_TRAFFIC = {"workload": (512, 2024, "bf16", "bf16"),
                  "kib": {1: (5524986.7, 6763025.9), 0: (3833670.3, 4058333.7)}}


def update_traffic():
    """Measured HBM bytes of the mean update launch (both variants), or None."""
    kib = UPDATE_TRAFFIC["kib"]
    if any(v is None for v in kib.values()):
        return None
    return int(round(sum(f * 2.0 + w * 1.0 for f, w in kib.values()) * 1024 / len(kib)))


def traffic_driver(args):
    """Torch-free driver for hardware-counter runs of the dominant kernel,
    k_fcg_update (the _sweep form the solver launches, BF16 storage), or with
    ``--traffic-kernel sr`` / ``s2`` of the fine level's fused pre-smoothing
    launch k_smooth_pair<T, false> (Jacobi sweep + residual restriction) /
    post-smoothing launch k_smooth_pair<T, true> (CORRECT + PLAIN sweeps, r.z
    partials), T = ``--traffic-storage``, or with ``--traffic-kernel dir`` of
    the direction pass k_fcg_dir<bf16>.

    The work tables and buffers of the headline batch (``--instances`` shapes
    from ``--seed``, the same draw as the solve workload) or, with
    ``--traffic-edge E``, of one E^3 instance come from hipMalloc; every
    instance is active.  The kernel's memory traffic does not depend on the
    buffer contents.  ``--reps`` launches are timed with HIP events, then the
    same number of device-to-device copies of 1 GiB (the counter calibration).
    Prints one JSON line."""
    import ctypes as C
    os.environ["HFM_NO_TORCH"] = "1"
    from hfm import _lib as L
    from hfm.workloads import batched_specs
    CHUNK = 8 * 256           # cells per k_fcg_update block (hfm.mg.CHUNK; mg imports torch)
    hip = C.CDLL("libamdhip64.so")
    vp = C.c_void_p

    def malloc(nbytes):
        p = vp()
        if hip.hipMalloc(C.byref(p), C.c_size_t(nbytes)) != 0:
            raise SystemExit(f"hipMalloc({nbytes}) failed")
        if hip.hipMemset(p, 0, C.c_size_t(nbytes)) != 0:
            raise SystemExit("hipMemset failed")
        return p.value

    def upload(arr):
        p = malloc(arr.nbytes)
        if hip.hipMemcpy(vp(p), arr.ctypes.data_as(vp), C.c_size_t(arr.nbytes), 1) != 0:
            raise SystemExit("hipMemcpy failed")
        return p

    if args.traffic_edge:
        shapes = [(args.traffic_edge,) * 3]
        what = f"one {args.traffic_edge}^3 instance"
    else:
        shapes = [sp[0] for sp in batched_specs(args.instances, args.seed)]
        what = f"{args.instances} instances, edges from seed {args.seed}"
    n = len(shapes)
    cells = np.array([D * H * W for (D, H, W) in shapes], dtype=np.int64)
    if args.traffic_kernel in ("sr", "s2"):
        return _traffic_pair(args, shapes, cells, what, malloc, upload, hip)
    nblk = ((cells + CHUNK - 1) // CHUNK + 7) // 8 * 8
    blk0 = np.concatenate([[0], np.cumsum(nblk)[:-1]])
    total = int(nblk.sum())
    t = np.zeros(n, dtype=L.TASK_DTYPE)
    for q, (D, H, W) in enumerate(shapes):
        N = D * H * W
        r = t[q]
        r["D"], r["H"], r["W"], r["inst"] = D, H, W, q
        r["blk0"], r["nblk"], r["slot0"] = blk0[q], nblk[q], blk0[q]
        r["magW"], r["magH"] = L.magic(W), L.magic(H)
        r["c"] = [malloc(4 * N) for _ in range(3)] + [malloc(4 * D * H), malloc(4 * D * W),
                                                      malloc(4 * H * W)]
        r["xin"], r["xout"], r["rlow"] = malloc(2 * N), malloc(2 * N), malloc(2 * N)
        r["ec"] = malloc(2 * N)
        r["x64"], r["r64"] = malloc(8 * ((N + 511) // 512 * 512)), malloc(8 * ((N + 511) // 512 * 512))
        if args.traffic_kernel == "dir":
            r["b"] = malloc(2 * N)   # z; the old direction is xin, the new one xout
    tasks = upload(t.view(np.uint8))
    blk = upload(np.repeat(np.arange(n, dtype=np.uint16), nblk))
    sc = np.zeros(n, dtype=L.SCALARS_DTYPE)
    sc["active"], sc["alpha"], sc["slot0"], sc["nslot"] = 1, 0.5, blk0, nblk
    sc["xlag"], sc["alpha_prev"] = args.traffic_lag, 0.25
    sc["beta"] = 0.5
    scp = upload(sc.view(np.uint8))
    dots = malloc(8 * total)
    hip.hipDeviceSynchronize()

    timed = lambda fn, reps: _hip_timed(hip, fn, reps)  # noqa: E731
    if args.traffic_kernel == "dir":
        fdir = lambda: L.call("hfm_fcg_dir", tasks, blk, total, scp, L.STORE["bf16"],  # noqa: E731
                              dots, None)
        ms = timed(fdir, args.reps)
        nbytes, cp_ms = _copy_reference(hip, malloc, args.reps)
        N = int(cells.sum())
        rd, wr = 16 * N, 2 * N    # z, p_old, three FP32 faces; p written
        print(json.dumps({"kernel": "k_fcg_dir<bf16>", "workload": what, "cells": N,
                          "blocks": total, "launches": args.reps + 1,
                          "alg_bytes_per_launch": rd + wr, "alg_read_bytes": rd,
                          "alg_write_bytes": wr, "launch_ms": round(ms, 4),
                          "achieved_GBs": round((rd + wr) / (ms * 1e-3) / 1e9, 1),
                          "copy_bytes_each_way": nbytes, "copy_launches": args.reps + 1,
                          "copy_ms": round(cp_ms, 4)}), flush=True)
        return
    upd = lambda: L.call("hfm_fcg_update_sweep", tasks, blk, total, scp, L.STORE["bf16"],  # noqa: E731
                         L.STORE["bf16"], C.c_float(2.0 / 3.0), dots, None)
    ms = timed(upd, args.reps)
    nbytes, cp_ms = _copy_reference(hip, malloc, args.reps)
    N = int(cells.sum())
    # bf16 MGBatch.update_bytes_per_cell(lag): p 2, faces 12, r64 16, rlow 2,
    # first sweep 2; with the solution update also x64 16 and p_prev 2
    rd, wr = (30 + 2, 20) if args.traffic_lag else (22, 12)
    alg = N * (rd + wr)
    print(json.dumps({"kernel": "k_fcg_update<bf16, true, bf16>", "workload": what, "cells": N,
                      "xlag": args.traffic_lag, "blocks": total, "launches": args.reps + 1,
                      "alg_bytes_per_launch": alg,
                      "alg_read_bytes": N * rd, "alg_write_bytes": N * wr,
                      "launch_ms": round(ms, 4), "achieved_GBs": round(alg / (ms * 1e-3) / 1e9, 1),
                      "copy_bytes_each_way": nbytes, "copy_launches": args.reps + 1,
                      "copy_ms": round(cp_ms, 4)}), flush=True)


def _fuse_kw(args):
    return {} if args.fuse_levels is None else {"fuse_levels": args.fuse_levels}


def _hip_timed(hip, fn, reps):
    """Mean ms of ``reps`` calls of fn after one warm call (HIP events on the
    null stream)."""
    import ctypes as C
    ev = [C.c_void_p(), C.c_void_p()]
    for x in ev:
        hip.hipEventCreate(C.byref(x))
    fn()
    hip.hipEventRecord(ev[0], None)
    for _ in range(reps):
        fn()
    hip.hipEventRecord(ev[1], None)
    hip.hipEventSynchronize(ev[1])
    ms = C.c_float()
    hip.hipEventElapsedTime(C.byref(ms), ev[0], ev[1])
    return ms.value / reps


def _copy_reference(hip, malloc, reps):
    """(bytes, mean ms) of a 1 GiB device-to-device copy: the counters'
    calibration."""
    import ctypes as C
    nbytes = 1 << 30
    src, dst = malloc(nbytes), malloc(nbytes)
    cp = lambda: hip.hipMemcpy(C.c_void_p(dst), C.c_void_p(src), C.c_size_t(nbytes), 3)  # noqa: E731
    return nbytes, _hip_timed(hip, cp, reps)


def _traffic_pair(args, shapes, cells, what, malloc, upload, hip):
    """A fused pass pair of the fine level over the batch, one tile per
    512-thread block (tile geometry from hfm_mg_tile), in ``--traffic-storage``:
    sr, x1 = Jacobi(x0, r) and the coarse right-hand side of r - A x1; s2, the
    CORRECT sweep (x0 + alpha P ec) followed by a PLAIN sweep, z stored in the
    same type, with the r.z partials."""
    import ctypes as C
    from hfm import _lib as L
    s2 = args.traffic_kernel == "s2"
    st = args.traffic_storage
    T = {"bf16": 2, "f16": 2, "f32": 4}[st]
    tiles = [L.tile(s) for s in shapes]
    if any(tl is None for tl in tiles):
        raise SystemExit("a level of the workload does not tile")
    n = len(shapes)
    ntile = np.array([tl[2] for tl in tiles], dtype=np.int64)
    blk0 = np.concatenate([[0], np.cumsum(ntile)[:-1]])
    total = int(ntile.sum())
    padded = (total + 7) // 8 * 8
    t = np.zeros(n, dtype=L.TASK_DTYPE)
    for q, ((D, H, W), (R, Z, nt)) in enumerate(zip(shapes, tiles)):
        N = D * H * W
        Dc, Hc, Wc = (D + 1) // 2, (H + 1) // 2, (W + 1) // 2
        r = t[q]
        r["D"], r["H"], r["W"], r["inst"] = D, H, W, q
        r["blk0"], r["nblk"], r["tile_rows"], r["tile_planes"] = blk0[q], nt, R, Z
        r["slot0"] = blk0[q]
        r["magW"], r["magH"] = L.magic(W), L.magic(H)
        r["Dc"], r["Hc"], r["Wc"] = Dc, Hc, Wc
        r["magWc"], r["magHc"] = L.magic(Wc), L.magic(Hc)
        r["c"] = [malloc(T * N) for _ in range(3)] + [malloc(T * D * H), malloc(T * D * W),
                                                      malloc(T * H * W)]
        r["xin"], r["xout"], r["b"] = malloc(T * N), malloc(T * N), malloc(T * N)
        if s2:
            r["ec"] = malloc(T * Dc * Hc * Wc)
        else:
            r["rlow"] = malloc(T * Dc * Hc * Wc)
    tasks = upload(t.view(np.uint8))
    bt = np.full(padded, 0xFFFF, dtype=np.uint16)
    bt[:total] = np.repeat(np.arange(n, dtype=np.uint16), ntile)
    blk = upload(bt)
    sc = np.zeros(n, dtype=L.SCALARS_DTYPE)
    sc["active"], sc["slot0"], sc["nslot"] = 1, blk0, ntile
    scp = upload(sc.view(np.uint8))
    dots = malloc(8 * padded)
    hip.hipDeviceSynchronize()
    om = C.c_float(2.0 / 3.0)
    if s2:
        fn = lambda: L.call("hfm_mg_smooth2", tasks, blk, padded, scp, L.STORE[st],  # noqa: E731
                            L.STORE[st], om, C.c_float(1.6), dots, None)
    else:
        fn = lambda: L.call("hfm_mg_smooth_restrict", tasks, blk, padded, scp,  # noqa: E731
                            L.STORE[st], om, None)
    ms = _hip_timed(hip, fn, args.reps)
    nbytes, cp_ms = _copy_reference(hip, malloc, args.reps)
    N = int(cells.sum())
    # x0, r, three faces read and x1 written (T bytes each); sr: the coarse
    # right-hand side written, s2: the coarse correction read (T per 8 cells)
    rd, wr = 5 * T * N, T * N
    if s2:
        rd += T * N // 8
    else:
        wr += T * N // 8
    print(json.dumps({"kernel": f"k_smooth_pair<{st}, {str(s2).lower()}, {st}, {str(s2).lower()}>",
                      "workload": what,
                      "cells": N, "tiles": total, "blocks": padded, "launches": args.reps + 1,
                      "alg_bytes_per_launch": rd + wr, "alg_read_bytes": rd, "alg_write_bytes": wr,
                      "launch_ms": round(ms, 4),
                      "achieved_GBs": round((rd + wr) / (ms * 1e-3) / 1e9, 1),
                      "copy_bytes_each_way": nbytes, "copy_launches": args.reps + 1,
                      "copy_ms": round(cp_ms, 4)}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--instances", type=int, default=512)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--storage", default="bf16", choices=["bf16", "f16", "f32"])
    ap.add_argument("--outer", default=None, choices=["bf16", "f16", "f32"],
                    help="storage of z and the search direction (default: --storage)")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--fp32-sample", type=int, default=64)
    ap.add_argument("--cpu-sample", type=int, default=32,
                    help="instances (workload order) the CPU baseline solves")
    ap.add_argument("--cpu-workers", type=int, default=16,
                    help="host processes of the CPU baseline (at most os.cpu_count())")
    ap.add_argument("--serve-requests", type=int, default=300)
    ap.add_argument("--serve-rate", type=float, default=300.0)
    ap.add_argument("--serve-sat-rate", type=float, default=3000.0)
    ap.add_argument("--no-fp32-batch", action="store_true")
    ap.add_argument("--iters-per-graph", type=int, default=4,
                    help="FCG iterations per HIP-graph replay (active flags read between)")
    ap.add_argument("--shrink", type=float, default=0.5,
                    help="re-plan over the active instances once at most this fraction remain")
    ap.add_argument("--streams", type=int, default=2, choices=[1, 2],
                    help="solve the batch as this many instance groups on their own streams")
    ap.add_argument("--no-fuse", action="store_true",
                    help="every V-cycle pass as its own launch (no fused pass pairs)")
    ap.add_argument("--fuse-levels", type=int, default=None,
                    help="finest levels that run fused pass pairs (default: the solver's)")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep 256-thread blocks (skip the launch-shape tuner)")
    ap.add_argument("--persistent-steps", type=int, default=3,
                    help="also time the batch with one persistent launch per iters_per_graph "
                         "iterations (0: skip)")
    ap.add_argument("--config0", type=int, default=1,
                    help="also time the CPU-runnable configuration's plain 64^3 solve (and "
                         "the oracle's on one host core)")
    ap.add_argument("--tune-schedule", type=int, default=1,
                    help="1: also rank sweep schedules by time to 1e-8 (reported beside "
                         "the V(2,2) headline)")
    ap.add_argument("--profile-kernel",
                    choices=["update", "sweep", "all", "flow2d", "urban", "reference"],
                    default=None,
                    help="launch only this kernel (after setup) or workload, for profiler runs")
    ap.add_argument("--single-shape", default=None,
                    help="D,H,W: profile one instance of this shape (sigma_ln 1, one dipole) "
                         "instead of the batch")
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--urban-steps", type=int, default=10)
    ap.add_argument("--ref-cycles", type=int, default=200,
                    help="reference sawtooth cycles timed on 128^3 (0: skip)")
    ap.add_argument("--ref-inhale-steps", type=int, default=2,
                    help="INHALE demo steps on 128 x 2048 x 2048 (0: skip)")
    ap.add_argument("--ref-urban-steps", type=int, default=3,
                    help="reference urban steps timed on 512^3 (0: skip)")
    ap.add_argument("--flow-steps", type=int, default=1000,
                    help="Kolmogorov 2D flow steps of the single-flow run (0: skip)")
    ap.add_argument("--flow-ensemble", type=int, default=256)
    ap.add_argument("--flow-ensemble-steps", type=int, default=20)
    ap.add_argument("--flow-team", type=int, default=None,
                    help="workgroups sharing the single flow (default: 16 at 256^2)")
    ap.add_argument("--traffic", action="store_true",
                    help="torch-free k_fcg_update driver for hardware-counter runs")
    ap.add_argument("--traffic-edge", type=int, default=0)
    ap.add_argument("--traffic-lag", type=int, choices=(0, 1), default=1,
                    help="update variant: 1 with the deferred solution update, 0 without")
    ap.add_argument("--traffic-kernel", choices=("update", "sr", "s2", "dir"), default="update",
                    help="update: k_fcg_update; sr / s2: the fine level's fused pre- / "
                         "post-smoothing launch; dir: k_fcg_dir")
    ap.add_argument("--traffic-storage", choices=("bf16", "f16", "f32"), default="bf16",
                    help="storage type of the fused pass pairs (sr, s2)")
    args = ap.parse_args()
    if args.traffic:
        return traffic_driver(args)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the solves are independent and this build runs "
              "on a single MI355X; multi-GPU runs are not part of it", file=sys.stderr)
        sys.exit(2)

    import torch

    from hfm.mg import MGBatch
    from hfm.workloads import batched_specs

    dev = torch.device("cuda:0")
    if args.profile_kernel:
        return profile_kernel(args, dev, torch)
    specs = batched_specs(args.instances, args.seed)
    shapes = [s[0] for s in specs]
    cells = sum(math.prod(s) for s in shapes)
    log(f"bench: {len(specs)} instances, {cells / 1e6:.1f} M fine cells, storage {args.storage}")
    mg = MGBatch(shapes, storage=args.storage, tol=1e-8, maxit=300, device=dev,
                 iters_per_graph=args.iters_per_graph, shrink=args.shrink, outer=args.outer,
                 fuse=not args.no_fuse, streams=args.streams, **_fuse_kw(args))
    load_batch(mg, specs, dev)
    tuning = None
    if not args.no_tune:
        from hfm.tune import tune
        mg.build_faces()
        mg.setup()
        t0 = time.perf_counter()
        choice = tune(mg)
        tuning = {"tune_s": round(time.perf_counter() - t0, 2),
                  "coarse_cells": choice["coarse_cells"],
                  "coarse_cycle_ms": {str(k): round(v, 4) for k, v in mg.coarse_times.items()},
                  "tile_planes": choice["tile_planes"],
                  "vcycle_ms_by_tile_planes": {str(k): round(v, 4)
                                               for k, v in mg.tile_times.items()},
                  "blocks": {"/".join(map(str, k)): v for k, v in sorted(choice["blocks"].items())
                             if v != 256}}
        log(f"bench: launch tuning {tuning}")
    torch.cuda.synchronize()
    log(f"bench: resident {mg.nbytes() / 2**30:.1f} GiB; warmup {args.warmup}")
    for _ in range(args.warmup):
        info = one_step(mg)
    torch.cuda.synchronize()
    times = []
    for k in range(args.steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        info = one_step(mg)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        log(f"bench: step {k} {times[-1] * 1e3:.1f} ms, iters mean {info['iter'].mean():.1f} "
            f"max {info['iter'].max()}")
    status = info["status"]
    if not np.all(status == 1):
        raise SystemExit(f"bench: {int(np.sum(status != 1))} instances did not converge")
    # independent FP64 check of every result (after the timed region)
    true_rel, face_dev = true_residuals(mg, torch)
    log(f"bench: true FP64 residual max {true_rel.max():.3e}, stored faces vs FP64 harmonic "
        f"{face_dev:.2e}")
    gpu_x = {q: (mg.solution(q).cpu().numpy(), int(info["iter"][q]))
             for q in range(min(args.cpu_sample, len(specs)))} if not args.no_cpu_baseline else {}
    total = sum(times)
    value = args.steps * len(specs) / total
    # Sigma over instances of iterations x (nu1 + nu2) x cells of the swept levels
    its = info["iter"].astype(np.int64)
    swept = np.array([sum(math.prod(lv) for lv in inst.levels[:-1]) for inst in mg.insts])
    fine = np.array([inst.N for inst in mg.insts])
    rates = {"cell_updates_per_s": float(np.sum(its * 4 * swept)) * args.steps / total,
             "fine_cell_vcycles_per_s": float(np.sum(its * fine)) * args.steps / total}

    probe_q = (0, len(specs) // 2, len(specs) - 1)
    xs0 = [mg.solution(q).clone() for q in probe_q] if args.persistent_steps > 0 else []
    # dominant kernel: the fine-level FCG update (x += a p, r -= a A p), timed live
    mg.restart()
    fine_cells = cells
    # its two variants alternate (the solution update rides on every other
    # iteration): a launch here is the mean of one of each
    upd_t = []
    for lag in (1, 0):
        mg.set_xlag(lag)
        upd_t.append(kernel_time_ms(lambda: mg.launch_update(0), 20, torch))
    upd_ms = 0.5 * sum(upd_t)
    upd_bytes = fine_cells * 0.5 * (mg.update_bytes_per_cell(1) + mg.update_bytes_per_cell(0))
    achieved = upd_bytes / (upd_ms * 1e-3) / 1e9
    sweep_ms = kernel_time_ms(lambda: mg.launch_fine_pre(), 20, torch)
    sweep_bytes = fine_cells * mg.fine_pre_bytes_per_cell()
    traffic = None
    if (args.instances, args.seed, mg.params.storage, mg.params.outer) == UPDATE_TRAFFIC["workload"]:
        traffic = update_traffic()
    roofline = {"bound": "hbm", "kernel": "k_fcg_update", "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic, "launch_ms": round(upd_ms, 4),
                "alg_bytes_per_launch": int(upd_bytes),
                "variants": {"with_x": {"launch_ms": round(upd_t[0], 4),
                                        "bytes_per_cell": mg.update_bytes_per_cell(1)},
                             "without_x": {"launch_ms": round(upd_t[1], 4),
                                           "bytes_per_cell": mg.update_bytes_per_cell(0)}},
                "fine_pre": {"kernel": mg.fine_pre_kernel(), "launch_ms": round(sweep_ms, 4),
                               "achieved": round(sweep_bytes / (sweep_ms * 1e-3) / 1e9, 1)}}
    extra = {}
    outer_storage = mg.params.outer
    if args.persistent_steps > 0:
        # the same batch, each group of iters_per_graph iterations one
        # persistent launch (per-instance dependency counters) instead of
        # graph-replayed launches; bitwise the same results
        from hfm.mg import persistent_workgroups
        info_p = one_step(mg, persistent=True)          # warm: programs built
        torch.cuda.synchronize()
        pt = []
        for _ in range(args.persistent_steps):
            t0 = time.perf_counter()
            info_p = one_step(mg, persistent=True)
            torch.cuda.synchronize()
            pt.append(time.perf_counter() - t0)
        same = all(bool((mg.solution(q) == x).all())
                   for q, x in zip(probe_q, xs0))
        extra["persistent"] = {
            "solves_per_s": round(len(specs) * len(pt) / sum(pt), 2),
            "ms_per_step": round(sum(pt) / len(pt) * 1e3, 2),
            "latency_ms": {"p50": round(float(np.percentile(pt, 50)) * 1e3, 2),
                           "p99": round(float(np.percentile(pt, 99)) * 1e3, 2)},
            "workgroups": persistent_workgroups(mg.pool),
            "iters_equal": bool(np.array_equal(info_p["iter"], info["iter"])),
            "solutions_bitwise_equal": same}
        log(f"bench: persistent {extra['persistent']}")
        extra["small_batches"] = small_batches(args, dev, torch)
    if args.config0:
        extra["config0_plain_64"] = config0(args, dev, torch)
    if args.tune_schedule:
        # the sweep schedule (nu1, nu2, alpha) ranked by the batch's time to 1e-8;
        # the headline keeps the configuration's V(2,2)
        from hfm.tune import tune_schedule
        t0 = time.perf_counter()
        best = tune_schedule(mg)
        tsec = time.perf_counter() - t0
        one_step(mg)                    # warm: the new plans' graphs
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        inf2 = one_step(mg)
        torch.cuda.synchronize()
        t1 = time.perf_counter() - t0
        extra["schedule_tuning"] = {
            "best": list(best), "tune_s": round(tsec, 2),
            "solve_ms_by_schedule": {f"V({a},{b}) alpha={c}": (round(v, 2) if math.isfinite(v)
                                                               else None)
                                     for (a, b, c), v in mg.schedule_times.items()},
            "best_solves_per_s": round(len(specs) / t1, 2),
            "best_iters_mean": round(float(inf2["iter"].mean()), 2),
            "converged": bool((inf2["status"] == 1).all())}
        log(f"bench: schedule tuning {extra['schedule_tuning']}")
    if not args.no_fp32_batch:
        del mg
        torch.cuda.empty_cache()
        extra["batched_fp32"] = batched_fp32(specs, dev, torch)
        extra["batched_f16"] = batched_fp32(specs, dev, torch, storage="f16")
        extra["speedup_mixed_vs_fp32_batched"] = round(value / extra["batched_fp32"]["solves_per_s"], 2)
    if args.fp32_sample > 0:
        # a systematic sample (every k-th instance of the randomly ordered
        # workload), so its size mix follows the batch's
        k = max(1, len(specs) // args.fp32_sample)
        sub = specs[::k][:args.fp32_sample]
        sps = fp32_one_at_a_time(sub, dev, torch)
        sub_cells = float(np.mean([math.prod(sp[0]) for sp in sub]))
        bat_cells = cells / len(specs)
        # per-cell normalisation to the batch's mean size
        sps_norm = sps * sub_cells / bat_cells
        extra["fp32_one_at_a_time"] = {
            "sample": f"every {k}th instance ({len(sub)})", "solves_per_s": round(sps, 2),
            "sample_mean_cells": round(sub_cells), "batch_mean_cells": round(bat_cells),
            "solves_per_s_at_batch_mean_size": round(sps_norm, 2)}
        extra["speedup_vs_fp32_one_at_a_time"] = round(value / sps_norm, 2)
    if args.serve_requests > 0:
        extra["serving"] = serving(args, dev, torch)
    if args.urban_steps > 0:
        extra["urban_flow"] = urban(args, dev, torch)
    if args.flow_steps > 0:
        extra["kolmogorov_2d"] = kolmogorov(args, dev, torch)
    if args.ref_cycles > 0 or args.ref_urban_steps > 0 or args.ref_inhale_steps > 0:
        extra["reference_context"] = reference_context(args, dev, torch)
    cpu = None
    if not args.no_cpu_baseline:
        workers = max(1, min(os.cpu_count() or 1, args.cpu_workers))
        # the baseline plan's per-op PyTorch-CPU restatement on all host cores
        # (the line's cpu_baseline), and the oracle's numpy solve in one worker
        # process per core beside it
        cpu = cpu_baseline_torch(specs, info["iter"], args.cpu_sample, workers, gpu_x)
        log(f"bench: cpu baseline (torch per-op) {cpu}")
        cpu["numpy_oracle_per_core_processes"] = cpu_baseline(specs, info["iter"],
                                                              args.cpu_sample, workers, gpu_x)
        log(f"bench: cpu baseline (numpy oracle) {cpu['numpy_oracle_per_core_processes']}")
    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "solves/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(total / args.steps * 1e3, 2), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64", "storage": args.storage,
        "outer_storage": outer_storage,
        "data": "synthetic",
        "config": {"workload": f"batched uneven anisotropic solves: {len(specs)} instances, edges "
                               "{32,48,64,96,128}^3, log-normal sigma_ln U[0.5,2], 1-4 dipoles, "
                               f"{args.storage}-storage V(2,2) preconditioned FP64 CG to 1e-8",
                   "instances": len(specs), "fine_cells": cells, "seed": args.seed},
        "iters_mean": round(float(info["iter"].mean()), 2), "iters_max": int(info["iter"].max()),
        "cell_updates_per_s": round(rates["cell_updates_per_s"], 1),
        "fine_cell_vcycles_per_s": round(rates["fine_cell_vcycles_per_s"], 1),
        "max_rel_residual": float(info["rel"].max()),
        "max_true_rel_residual_fp64": float(true_rel.max()),
        "faces_max_rel_diff_vs_fp64_harmonic": face_dev,
        "latency_ms": {"p50": round(float(np.percentile(times, 50)) * 1e3, 2),
                       "p99": round(float(np.percentile(times, 99)) * 1e3, 2)},
        "roofline": roofline, "cpu_baseline": cpu, "launch_tuning": tuning, **extra,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
