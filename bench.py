#!/usr/bin/env python3
"""PD steps/s on the 64^3-hex cube (1,572,864 tets, 274,625 vertices), one MI355X.

Workload (one step = one PdSolver::Update of the whole body): 64x64x64 hex
grid, 0.1 m spacing, 6 Kuhn tets per hex, E = 1e5 Pa, nu = 0.3, rho = 1000,
top face pinned, gravity -9.8 m/s^2, seeded +-1 mm jitter, dt = 1/60 s;
20 local/global iterations per step, each a Jacobi-PCG global solve
(||D^-1 r|| reduced 1e-6, floor 1e-8 ||D^-1 b||) followed by the Chebyshev
recurrence (rho = 0.9992); FP32 arithmetic.  Inputs are resident in HBM
before the timed region.

Prints one JSON line: steps/s, ms/step, the roofline of the dominant kernel
(timed live with HIP events on the solver's stream), the FP32 run against the
FP32 oracle after 10 steps (parity) and against an FP64 run of the same steps
(accuracy), the CPU baseline (the oracle's C++ restatement of the same step
with OpenMP) on a bounded sample, and under "configs" short runs of the other
workloads: the 8^3 cube on the direct solve (100 steps), 256 small bodies on
the dense MFMA solve, 1,024 uneven bodies on the batched per-body solve
against per-body solvers, and the FP64 IPC contact scene of two 16^3
Neo-Hookean cubes (100 steps).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "soft-body-simulation-cuda_amd"))

from sbs_amd import mesh  # noqa: E402
from sbs_amd.capi import Stream  # noqa: E402
from sbs_amd.pd import PdSolver, SolverData, SolverParams, SolverType  # noqa: E402

METRIC = "PD steps/sec at 64^3-hex cube vs plain FP32 at stated error; HBM GB/s"
HBM_PEAK_GBS = 8000.0


def workload_params(fp64: bool = False, halo: bool = True) -> SolverParams:
    return SolverParams(numIterations=20, dt=1.0 / 60.0, gravity=9.8, tol=1e-6, rho=0.9992, pcg_tol=1e-6,
                        pcg_floor=0.0 if fp64 else 1e-8, pcg_max_iter=200, pcg_chebyshev=True,
                        pcg_adaptive=True, pcg_halo=halo)


def local_step_bytes(nt: int, nv: int) -> int:
    # per tet: 4 u32 indices + 9 f32 Dm^-1 + 1 f32 w; per vertex: read q, read
    # sn_old, read m/h^2, write b (the unit of the reference's computeLocal +
    # addM_h2Sn pass)
    return 56 * nt + 40 * nv


def pcg_iteration_bytes(nnz_off: int, nv: int) -> int:
    # the work unit of one Jacobi-PCG iteration as a CSR implementation moves
    # it: matrix values + u32 columns of the scalar matrix (nnz_s = N_v +
    # off-diagonal entries), row pointers, D^-1, and 11 passes over 12-byte
    # vertex vectors (SpMV p -> q: 2; x += alpha p: 3; r -= alpha q: 3;
    # p = D^-1 r + beta p: 3): 8 nnz_s + 4 (N_v + 1) + 4 N_v + 132 N_v
    nnz_s = nnz_off + nv
    return 8 * nnz_s + 4 * (nv + 1) + 4 * nv + 132 * nv


TRAFFIC_FILE = "profiles/pd_step_64cube_fp32_kernels.json"
_TRAFFIC_KERNELS = {"local_step": ["local_chunk_kernel<float, true>", "local_chunk_kernel<float, false>"],
                    "pcg_iteration": ["pcg_iter_halo_kernel<float, false>", "pcg_iter_halo_kernel<float, true>"],
                    "pcg_iteration_gather": ["pcg_iter_kernel<float, false>", "pcg_iter_kernel<float, true>"]}


def measured_traffic(kernel: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (scaled by the copy calibration), averaged over the
    profiled launches of all its instantiations; None when not profiled."""
    path = os.path.join(REPO, TRAFFIC_FILE)
    if not os.path.exists(path):
        return None
    k = json.load(open(path))["kernels"]
    parts = [k[n] for n in _TRAFFIC_KERNELS.get(kernel, []) if n in k]
    if not parts:
        return None
    n = sum(p["dispatches"] for p in parts)
    return int(sum(p["mean_mb"] * p["dispatches"] for p in parts) / n * 1e6)


def run_gpu(scene, steps: int, warmup: int, precision: int, timing_steps: int, halo: bool = True,
            graph: bool = False, polar: bool = False):
    data = SolverData(scene, precision)
    pd = PdSolver(128, data)
    pd.SetGlobalSolver(SolverType.PCG)
    stream = Stream()            # the solver's own stream (graph replay needs a non-null one)
    pd.set_stream(stream)
    p = workload_params(precision == 64, halo)
    p.use_graph = graph
    p.polar_rotation = polar
    for _ in range(warmup):
        pd.Update(data, p)
    pd.synchronize()
    st0 = pd.stats()
    t0 = time.perf_counter()
    for _ in range(steps):
        pd.Update(data, p)
    pd.synchronize()
    t1 = time.perf_counter()
    st1 = pd.stats()
    X = data.get("X")
    kt, tst = {}, None
    if timing_steps > 0:
        # kernel durations from HIP events on the solver's stream, in extra
        # steps after the timed region (events between launches add gaps: the
        # PCG iterations of a solve are timed as one run of launches)
        names = ("local_step", "pcg_iteration", "pcg_solve")
        pd.set_kernel_timing(names)
        for _ in range(timing_steps):
            pd.Update(data, p)
        pd.synchronize()
        st2 = pd.stats()
        kt = {k: pd.kernel_time(k) for k in names}
        tst = dict(steps=timing_steps, pcg_iters=int(st2.pcg_iterations_total - st1.pcg_iterations_total),
                   pcg_solves=int(st2.pcg_solves_total - st1.pcg_solves_total))
    out = dict(seconds=t1 - t0, X=X, stats=st1, kernel_times=kt, timing=tst, nnz_off=int(st1.nnz_offdiag),
               pcg_iters=int(st1.pcg_iterations_total - st0.pcg_iterations_total),
               pcg_solves=int(st1.pcg_solves_total - st0.pcg_solves_total),
               unconverged=int(st1.pcg_cap_hits - st0.pcg_cap_hits),
               graph_replays=int(st1.graph_replays - st0.graph_replays),
               graph_captures=int(st1.graph_captures - st0.graph_captures))
    pd.close()
    data.free()
    stream.close()
    return out


def cpu_baseline(scene, budget_s: float, max_steps: int):
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O
    p = O.make_params(solver=O.SOLVER_PCG, num_iterations=20, dt=1.0 / 60.0, gravity=9.8, tol=1e-6, rho=0.9992,
                      pcg_tol=1e-6, pcg_floor=1e-8, pcg_max_iter=200, pcg_chebyshev=1)
    o = O.PdOracle(scene, p, fp64=False, bench=True)
    o.step(1)  # warm-up step
    n, t0 = 0, time.perf_counter()
    while n < max_steps:
        o.step(1)
        n += 1
        if time.perf_counter() - t0 > budget_s:
            break
    el = time.perf_counter() - t0
    threads = O.num_threads(bench=True)
    o.close()
    model = next((l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo") if l.startswith("model name")), None)
    return dict(value=n / el, unit="steps/s", cores=threads, kind="port",
                cpu_model=model, affinity_cores=len(os.sched_getaffinity(0)), nproc=os.cpu_count(),
                sample=f"{n} step(s) of the same 64^3 workload after 1 warm-up step, FP32, "
                       f"oracle C++ restatement with OpenMP ({threads} threads; every loop of the step, the PCG "
                       f"dot products and updates included, runs in parallel), {el:.1f} s")


def parity_vs_oracle(scene, steps: int, polar: bool):
    """The FP32 GPU step against the FP32 oracle (the reference's step in its
    precision, the SVD rotation) after `steps` steps from rest."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O
    g = run_gpu(scene, steps, 0, 32, timing_steps=0, polar=polar)
    o = O.PdOracle(scene, O.make_params(solver=O.SOLVER_PCG, num_iterations=20, dt=1.0 / 60.0, gravity=9.8,
                                        tol=1e-6, rho=0.9992, pcg_tol=1e-6, pcg_floor=1e-8, pcg_max_iter=200,
                                        pcg_chebyshev=1), fp64=False)
    o.step(steps)
    ref = o.get("X")
    o.close()
    return dict(steps=steps, rel_l2=float(np.linalg.norm(g["X"] - ref) / np.linalg.norm(ref)),
                max_rel=float(np.abs(g["X"] - ref).max() / np.abs(ref).max()),
                reference="FP32 oracle (oracle/pd_oracle.cpp), same scene and parameters")


def dense_scene(nbodies: int = 256, n: int = 6, seed: int = 0):
    return mesh.small_bodies_scene(nbodies, n, seed)


def dense_measure(steps: int, warmup: int):
    """256 bodies of 6^3 hexes, 10 local/global iterations per step with the
    dense small-body solver (per-body inverse on MFMA + FP64 refinement to
    1e-10), against the FP64 oracle's direct mode on the same steps."""
    scene = dense_scene()
    print(f"# dense workload: {scene.nv} vertices, {scene.nt} tets, 256 bodies", file=sys.stderr, flush=True)
    out = {}
    for name, prec in (("fp32_mfma", 0), ("bf16_mfma", 1)):
        data = SolverData(scene, 32)
        pd = PdSolver(128, data)
        pd.SetGlobalSolver(SolverType.DenseCholesky)
        # all 10 iterations every step (no early exit), so FP32 and the FP64
        # oracle take the same discrete path and the error is the arithmetic's
        p = SolverParams(numIterations=10, dt=1.0 / 60.0, gravity=9.8, tol=0.0, rho=0.9992, dense_precision=prec,
                         dense_tol=1e-10, dense_max_iter=16)
        for _ in range(warmup):
            pd.Update(data, p)
        pd.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            pd.Update(data, p)
        pd.synchronize()
        el = time.perf_counter() - t0
        info = pd.dense_info()
        out[name] = dict(steps_per_s=round(steps / el, 3), ms_per_step=round(1e3 * el / steps, 4),
                         refinement_corrections_last_solve=info["iterations"], residual_ratio=info["ratio"],
                         inverse_bytes=info["inverse_bytes"], X=data.get("X"))
        pd.close()
        data.free()
        print(f"# {name}: {out[name]['steps_per_s']} steps/s", file=sys.stderr, flush=True)
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O
    o = O.PdOracle(scene, O.make_params(solver=O.SOLVER_CHOLESKY, num_iterations=10, dt=1.0 / 60.0, gravity=9.8,
                                        tol=0.0, rho=0.9992), fp64=True)
    o.step(warmup + steps)
    ref = o.get("X")
    o.close()
    disp = np.abs(ref - scene.X).max()
    for name in out:
        X = out[name].pop("X")
        out[name]["max_abs_err_vs_fp64_oracle_m"] = float(np.abs(X - ref).max())
        out[name]["rel_l2_vs_fp64_oracle"] = float(np.linalg.norm(X - ref) / np.linalg.norm(ref))
        out[name]["max_err_over_max_displacement"] = float(np.abs(X - ref).max() / disp)
    return dict(modes=out, verts=scene.nv, tets=scene.nt, bodies=256, steps=steps + warmup)


def run_dense_workload(args):
    r = dense_measure(args.steps, args.warmup)
    out = r["modes"]
    best = max(out, key=lambda k: out[k]["steps_per_s"])
    line = {"metric": "PD steps/sec, 256 bodies of 6^3 hexes, dense MFMA solves + FP64 refinement",
            "value": out[best]["steps_per_s"], "unit": "steps/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": out[best]["ms_per_step"], "higher_is_better": True,
            "scaling": "strong", "vs_baseline": None, "dtype": "f32 (FP64 refinement)", "data": "synthetic",
            "config": {"workload": "256 x 6^3-hex bodies, 10 it/step, dense solver", "verts": r["verts"],
                       "tets": r["tets"], "bodies": 256},
            "modes": out, "reference": "FP64 CPU oracle, direct (Cholesky) mode, same steps"}
    print(json.dumps(line), flush=True)


def batched_scene(nbodies: int = 1024, seed: int = 0):
    """The batched workload: nbodies cubes of (nx, ny, nz) hexes drawn per
    axis uniformly from 3..20, each with a rigid offset on a grid (no
    overlap) and E uniform in [5e4, 5e5] Pa; top faces pinned."""
    idx = np.arange(4 * nbodies, dtype=np.uint64)
    u = mesh.uniform01(seed + 101, idx)
    dims = (3 + np.floor(u[: 3 * nbodies] * 18)).astype(int).reshape(nbodies, 3)
    E = u[3 * nbodies:] * 4.5e5 + 5e4
    side = int(np.ceil(np.sqrt(nbodies)))
    bodies = [dict(nx=int(d[0]), ny=int(d[1]), nz=int(d[2]), spacing=0.1,
                   origin=(2.5 * (i % side), 0.0, 2.5 * (i // side)), jitter=1e-3, seed=seed + 7 + i,
                   E=float(E[i]), nu=0.3, density=1000.0, pin_top=True) for i, d in enumerate(dims)]
    return mesh.make_scene(bodies), bodies


def batched_measure(steps: int, warmup: int, sample: int, fp64_check: bool = True, polar: bool = False):
    """The 1024 uneven bodies on the batched solver (every body its own
    Jacobi-PCG + Chebyshev, one workgroup per body, one launch per size class,
    workgroup sizes from the autotuner at the first update), against one
    PdSolver per body (the plain FP32 per-body form) on a sample of the bodies,
    extrapolated to all of them; accuracy against an FP64 batched run of the
    same steps (FP64 batched is parity-tested against the FP64 oracle per body)."""
    scene, bodies = batched_scene()
    print(f"# batched workload: {scene.nv} vertices, {scene.nt} tets, {len(bodies)} bodies", file=sys.stderr,
          flush=True)
    p = SolverParams(numIterations=10, dt=1.0 / 60.0, gravity=9.8, tol=1e-6, rho=0.9992, pcg_tol=1e-6,
                     pcg_floor=1e-8, pcg_max_iter=200, pcg_chebyshev=True, pcg_adaptive=True, polar_rotation=polar)
    out = {}
    runs = [(32, p)]
    if fp64_check:
        runs.append((64, SolverParams(**{**p.__dict__, "pcg_floor": 0.0})))
    for prec, pp in runs:
        data = SolverData(scene, prec)
        pd = PdSolver(128, data)
        pd.SetGlobalSolver(SolverType.BatchedPCG)
        stream = Stream()
        pd.set_stream(stream)
        t0 = time.perf_counter()
        pd.Update(data, pp)      # prepares: bodies, systems, the autotuner
        pd.synchronize()
        t_prep = time.perf_counter() - t0
        for _ in range(max(warmup - 1, 0)):
            pd.Update(data, pp)
        pd.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            pd.Update(data, pp)
        pd.synchronize()
        el = time.perf_counter() - t0
        info = pd.batch_info()
        out[prec] = dict(seconds=el, X=data.get("X"), info=info, prepare_s=t_prep)
        pd.close()
        data.free()
        stream.close()
        print(f"# batched fp{prec}: {steps / el:.2f} steps/s", file=sys.stderr, flush=True)
    # per-body engines on a sample, same steps
    idx = list(range(0, len(bodies), max(1, len(bodies) // sample)))[:sample]
    t_sample, errs = 0.0, []
    pb = SolverParams(**{**p.__dict__, "pcg_adaptive": True})
    Xb = out[32]["X"]
    for i in idx:
        sb = mesh.make_scene([bodies[i]])
        d = SolverData(sb, 32)
        e = PdSolver(128, d)
        e.SetGlobalSolver(SolverType.PCG)
        for _ in range(warmup):
            e.Update(d, pb)
        e.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            e.Update(d, pb)
        e.synchronize()
        t_sample += time.perf_counter() - t0
        v0, v1 = scene.body_vert_start[i], scene.body_vert_start[i + 1]
        xb = Xb[v0:v1]
        errs.append(float(np.linalg.norm(d.get("X") - xb) / np.linalg.norm(xb)))
        e.close()
        d.free()
    t_per_body_all = t_sample / len(idx) * len(bodies)
    sps_b = steps / out[32]["seconds"]
    sps_p = steps / t_per_body_all
    info = out[32]["info"]
    st = info["body_stats"]
    res = {"steps_per_s": round(sps_b, 3), "ms_per_step": round(1e3 * out[32]["seconds"] / steps, 3),
           "local_step_rotation": "Newton polar iteration, TR1690 SVD fallback" if polar else "TR1690 SVD",
           "bodies": len(bodies), "verts": scene.nv, "tets": scene.nt,
           "classes": [{k: c[k] for k in ("min_tets", "max_tets", "bodies", "threads")} |
                       {"tuner_ms": [(t, round(ms, 3)) for t, ms in c["tried"]]} for c in info["classes"]],
           "pcg_iterations_per_solve": round(float(st[:, 2].sum()) / max(float(st[:, 3].sum()), 1), 2),
           "max_pcg_iterations_one_solve": int(st[:, 4].max()), "solves_cut_by_max_iter": int(st[:, 5].sum()),
           "per_body_solvers": {"steps_per_s_extrapolated": round(sps_p, 4), "sample_bodies": len(idx),
                                "speedup_batched": round(sps_b / sps_p, 2),
                                "max_rel_l2_vs_batched": max(errs)},
           "prepare_s_incl_autotune": round(out[32]["prepare_s"], 2)}
    if 64 in out:
        X64 = out[64]["X"]
        worst = 0.0
        for i in range(len(bodies)):
            v0, v1 = scene.body_vert_start[i], scene.body_vert_start[i + 1]
            d = np.linalg.norm(X64[v0:v1] - scene.X[v0:v1])
            worst = max(worst, float(np.linalg.norm(Xb[v0:v1] - X64[v0:v1]) / max(d, 1e-12)))
        res["error_vs_fp64"] = {"rel_l2": float(np.linalg.norm(Xb - X64) / np.linalg.norm(X64)),
                                "max_rel": float(np.abs(Xb - X64).max() / np.abs(X64).max()),
                                "max_abs_m": float(np.abs(Xb - X64).max()),
                                "worst_body_err_over_displacement": worst,
                                "reference": "FP64 batched run of the same steps (parity-tested against the FP64 "
                                             "oracle per body)"}
    return res


def run_batched_workload(args):
    r = batched_measure(args.steps, args.warmup, args.per_body_sample, polar=args.rotation == "polar")
    line = {"metric": "PD steps/sec, 1024 uneven bodies (3..20 hexes per axis), batched per-body solves vs "
                      "per-body solvers", "value": r["steps_per_s"], "unit": "steps/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": r["ms_per_step"], "higher_is_better": True,
            "scaling": "strong", "vs_baseline": None, "dtype": "f32 (FP64 reductions)", "data": "synthetic",
            "config": {"workload": "1024 bodies, Jacobi-PCG + Chebyshev per body, 10 it/step", "verts": r["verts"],
                       "tets": r["tets"], "bodies": r["bodies"]}, "batched": r}
    print(json.dumps(line), flush=True)


def ipc_measure(nsteps: int = 100, cpu_steps: int = 3, cpu_budget: float = 20.0):
    """The FP64 IPC contact scene: two 16^3-hex Neo-Hookean cubes (E = 1e5 Pa,
    nu = 0.4), the upper one 5 mm above the lower (pinned on its bottom face)
    with +-0.5 mm jitter; h = 1/100, dhat = 1e-3, Newton tol 5e-2 (the
    reference context's), Jacobi-PCG inner solve; `nsteps` steps from rest.
    Latency-bound (a host-driven Newton loop), so it reports time per step.
    The oracle's single-thread restatement runs the first `cpu_steps` steps."""
    from sbs_amd.ipc import IpcData, IpcParams, IpcSolver
    scene = mesh.stack_scene(16, spacing=0.1, gap=5e-3, jitter=5e-4, seed=0, E=1e5, nu=0.4)
    print(f"# ipc workload: {scene.nv} vertices, {scene.nt} tets, {scene.ntri} surface triangles", file=sys.stderr,
          flush=True)
    p = IpcParams(dt=0.01, damp=0.999, dhat=1e-3, tol=5e-2, maxIterations=100, lin_tol=1e-10, lin_max_iter=5000)
    data = IpcData(scene)
    solver = IpcSolver(data, elastic="neohookean")
    newton, lin, queries, kappa, first_contact = 0, 0, 0, 0.0, None
    times = []
    for k in range(nsteps):
        t0 = time.perf_counter()
        assert solver.Update(data, p), solver.stats()
        times.append(time.perf_counter() - t0)
        st = solver.stats()
        newton += st["newton_iterations"]
        lin += st["linear_iterations"]
        queries = max(queries, st["queries"])
        kappa = max(kappa, st["kappa"])
        if first_contact is None and st["kappa"] > 100.0:
            first_contact = k + 1
    up = np.arange(scene.nv) >= scene.body_vert_start[1]
    X = data.get("X")
    gap = float(X[up, 1].min() - X[~up, 1].max())
    solver.close()
    data.free()
    el = sum(times[1:]) if nsteps > 1 else times[0]
    cpu = None
    if cpu_steps > 0:
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import oracle as O
        o = O.IpcOracle(scene.X, np.zeros_like(scene.X), scene.tet, scene.tri, scene.tri_fathers, scene.mass,
                        scene.mu, scene.lam, scene.dbc > 0, dt=p.dt, damp=p.damp, dhat=p.dhat, tol=p.tol,
                        lin_tol=p.lin_tol, max_iter=p.maxIterations, elastic="neohookean")
        n, c0 = 0, time.perf_counter()
        while n < cpu_steps and time.perf_counter() - c0 < cpu_budget:
            o.step()
            n += 1
        cel = time.perf_counter() - c0
        cpu = dict(value=round(n / cel, 4), unit="steps/s", cores=1, kind="port",
                   sample=f"the first {n} step(s) of the same scene, FP64, oracle C++ restatement (one thread), "
                          f"{cel:.1f} s", gpu_steps_per_s_same_steps=round(n / sum(times[:n]), 3))
        o.close()
    return {"steps": nsteps, "steps_per_s": round((len(times) - 1) / el, 3) if nsteps > 1 else None,
            "ms_per_step": round(1e3 * el / max(len(times) - 1, 1), 3), "first_step_ms_incl_prepare":
            round(1e3 * times[0], 1), "newton_iterations_per_step": round(newton / nsteps, 2),
            "pcg_iterations_per_solve": round(lin / max(newton + nsteps, 1), 1), "max_queries": queries,
            "max_kappa": kappa, "first_step_with_barrier": first_contact, "gap_after_m": gap,
            "verts": scene.nv, "tets": scene.nt, "tris": scene.ntri, "elastic": "neohookean", "cpu_baseline": cpu}


def run_ipc_workload(args):
    r = ipc_measure(args.ipc_steps, 0 if args.no_cpu else 3, args.cpu_budget)
    line = {"metric": "IPC steps/sec, two 16^3-hex Neo-Hookean cubes dropped 5 mm into contact (FP64 Newton + "
                      "barrier)", "value": r["steps_per_s"], "unit": "steps/s", "n_gpus": 1, "steps": r["steps"],
            "warmup": 0, "ms_per_step": r["ms_per_step"], "higher_is_better": True, "scaling": "strong",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "2 x 16^3-hex Neo-Hookean cubes, IPC step (h 1/100, dhat 1e-3, tol 5e-2)",
                       "verts": r["verts"], "tets": r["tets"], "tris": r["tris"]},
            "ipc": {k: v for k, v in r.items() if k not in ("cpu_baseline",)}, "cpu_baseline": r["cpu_baseline"]}
    print(json.dumps(line), flush=True)


def config0_measure(steps: int = 100):
    """The 8^3-hex cube (729 vertices, 3,072 tets), 10 local/global iterations
    per step on the prefactored sparse Cholesky (the band factor), 100 steps,
    FP32: time, and positions against the FP32 oracle's Cholesky mode."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O
    scene = mesh.cube_scene(8)
    p = SolverParams(numIterations=10, dt=1.0 / 60.0, gravity=9.8, tol=1e-6, rho=0.9992)
    data = SolverData(scene, 32)
    pd = PdSolver(128, data)
    pd.SetGlobalSolver(SolverType.CuSolverCholesky)
    pd.Update(data, p)
    pd.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps - 1):
        pd.Update(data, p)
    pd.synchronize()
    el = time.perf_counter() - t0
    X = data.get("X")
    pd.close()
    data.free()
    o = O.PdOracle(scene, O.make_params(solver=O.SOLVER_CHOLESKY, num_iterations=10), fp64=False)
    c0 = time.perf_counter()
    o.step(steps)
    cel = time.perf_counter() - c0
    Xo = o.get("X")
    o.close()
    return {"steps": steps, "steps_per_s": round((steps - 1) / el, 2), "verts": scene.nv, "tets": scene.nt,
            "rel_l2_vs_fp32_oracle": float(np.linalg.norm(X - Xo) / np.linalg.norm(Xo)),
            "max_rel_vs_fp32_oracle": float(np.abs(X - Xo).max() / np.abs(Xo).max()),
            "cpu_oracle_steps_per_s": round(steps / cel, 2)}


def run_solver_workload(args, solver, n: int, kernel: str, label: str):
    """The 20-iteration PD step of an n^3 cube (the headline's body and
    parameters) on another global solver: Jacobi + Chebyshev (the reference's
    default) or the direct band Cholesky.  The global-step kernel is timed
    with HIP events in as many steps again after the timed region."""
    scene = mesh.cube_scene(n, spacing=0.1, jitter=1e-3, seed=0, E=1e5, nu=0.3, density=1000.0, pin_top=True)
    print(f"# {label}: {scene.nv} vertices, {scene.nt} tets", file=sys.stderr, flush=True)
    p = workload_params()
    data = SolverData(scene, 32)
    pd = PdSolver(128, data)
    pd.SetGlobalSolver(solver)
    for _ in range(args.warmup):
        pd.Update(data, p)
    pd.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        pd.Update(data, p)
    pd.synchronize()
    el = time.perf_counter() - t0
    pd.set_kernel_timing(("local_step", kernel))
    for _ in range(args.steps):
        pd.Update(data, p)
    pd.synchronize()
    st = pd.stats()
    kinfo = {}
    for k in ("local_step", kernel):
        nl, ms = pd.kernel_time(k)
        kinfo[k] = dict(launches=nl, avg_us=round(ms / max(nl, 1) * 1e3, 2))
    kinfo["local_step"]["gbs"] = round(local_step_bytes(scene.nt, scene.nv) / (kinfo["local_step"]["avg_us"] * 1e3), 1)
    roof = None
    if solver == SolverType.Jacobi:
        # per vertex: pin flag, partial range, x~, m/h^2, the diagonal, sn and prev
        # read and written; per (chunk, vertex) partial: 12 bytes
        jb = 76 * scene.nv + 12 * st.partial_entries
        kinfo[kernel]["bytes_per_launch"] = jb
        kinfo[kernel]["gbs"] = round(jb / (kinfo[kernel]["avg_us"] * 1e3), 1)
        roof = dict(bound="hbm", kernel=kernel, achieved=kinfo[kernel]["gbs"], peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(kinfo[kernel]["gbs"] / HBM_PEAK_GBS, 4))
    pd.close()
    data.free()
    line = {"metric": f"PD steps/sec, {n}^3-hex cube, {label}", "value": round(args.steps / el, 3),
            "unit": "steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(1e3 * el / args.steps, 4), "higher_is_better": True, "scaling": "strong",
            "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": f"{n}^3-hex cube PD step ({label}, 20 it/step)", "verts": scene.nv,
                       "tets": scene.nt}, "kernels": kinfo, "roofline": roof}
    print(json.dumps(line), flush=True)


def run_collide_workload(args):
    """Two 24^3-hex cubes, the upper one falling onto the lower one (pinned on
    its bottom face), with the PD update's collision handling (LBVH broad
    phase over the step's motion, CCD narrow phase, the CCD response) and two
    fixed bodies (a floor plane and a sphere beside the upper cube): the
    collision rows of the step, timed per phase by the solver's own events in
    as many steps again after the timed region."""
    from sbs_amd import fixed as F
    n = 24
    scene = mesh.stack_scene(n, spacing=0.1, gap=2e-3, jitter=5e-4, seed=0)
    print(f"# collide workload: {scene.nv} vertices, {scene.nt} tets, {scene.ntri} surface triangles",
          file=sys.stderr, flush=True)
    p = workload_params()
    p.handleCollision = True
    data = SolverData(scene, 32)
    bodies = F.FixedBodyData([F.Plane(pos=(0.0, -0.05, 0.0)), F.Sphere(pos=(n * 0.1 + 0.35, n * 0.15, n * 0.05),
                                                                       radius=0.3)])
    data.set_fixed_bodies(bodies)
    pd = PdSolver(128, data)
    pd.SetGlobalSolver(SolverType.PCG)
    for _ in range(args.warmup):
        pd.Update(data, p)
    pd.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        pd.Update(data, p)
    pd.synchronize()
    el = time.perf_counter() - t0
    pd.SetPerf(True)
    before = dict(pd.GetPerformanceData())
    for _ in range(args.steps):
        pd.Update(data, p)
    pd.synchronize()
    after = dict(pd.GetPerformanceData())
    phases = {k: round((after[k] - before.get(k, 0.0)) / args.steps, 4) for k in after}
    pd.close()
    bodies.close()
    data.free()
    line = {"metric": "PD steps/sec, two 24^3-hex cubes colliding, with collision handling and fixed bodies",
            "value": round(args.steps / el, 3), "unit": "steps/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(1e3 * el / args.steps, 4), "higher_is_better": True,
            "scaling": "strong", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "2 x 24^3-hex cubes, PD (Jacobi-PCG + Chebyshev, 20 it/step), collisions, "
                                   "a plane and a sphere", "verts": scene.nv, "tets": scene.nt,
                       "tris": scene.ntri},
            "phases_ms_per_step": phases}
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=64, help="hexes per axis (64 = the measured workload)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 error reference run")
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of CPU baseline work")
    ap.add_argument("--timing-steps", type=int, default=None,
                    help="steps timed kernel by kernel after the timed region (default: as many as --steps)")
    ap.add_argument("--workload", default="cube64",
                    choices=["cube64", "dense256", "batched", "ipc", "jacobi64", "direct14", "direct27", "collide"],
                    help="cube64: the headline 64^3 cube; dense256: 256 small bodies on the dense solver; "
                         "batched: 1024 uneven bodies in one engine vs per-body engines; "
                         "ipc: the FP64 IPC step on two 16^3 cubes in contact; jacobi64: the 64^3 cube on the "
                         "Jacobi + Chebyshev global step; direct14, direct27: a 14^3 or 27^3 cube on the band Cholesky; "
                         "collide: "
                         "two 24^3 cubes colliding, with fixed bodies")
    ap.add_argument("--per-body-sample", type=int, default=16, help="bodies timed as separate engines (batched)")
    ap.add_argument("--ipc-steps", type=int, default=100, help="steps of the IPC contact scene")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the other workloads' summaries in the default line (8^3 direct, 256 dense "
                         "bodies, 1024 batched bodies, the IPC contact scene)")
    ap.add_argument("--graph", action="store_true", help="replay each step as a captured HIP graph")
    ap.add_argument("--rotation", default="polar", choices=["polar", "svd"],
                    help="FP32 local step rotation: Newton's polar iteration (SVD fallback) or the TR1690 SVD "
                         "throughout (the reference's arithmetic)")
    ap.add_argument("--pcg-kernels", default="halo", choices=["halo", "gather"],
                    help="PCG kernel family: halo-staged (LDS) or per-row neighbour gathers")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU (the PD step of one body on one MI355X); "
              f"--gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)
    if args.workload == "dense256":
        run_dense_workload(args)
        return
    if args.workload == "batched":
        run_batched_workload(args)
        return
    if args.workload == "ipc":
        run_ipc_workload(args)
        return
    if args.workload == "jacobi64":
        run_solver_workload(args, SolverType.Jacobi, 64, "jacobi_chebyshev", "Jacobi + Chebyshev global step")
        return
    if args.workload in ("direct14", "direct27"):
        # band factor (host) + ring-buffer band triangular solves (GPU)
        run_solver_workload(args, SolverType.CuSolverCholesky, int(args.workload[6:]), "direct_solve",
                            "band Cholesky global step")
        return
    if args.workload == "collide":
        run_collide_workload(args)
        return

    scene = mesh.cube_scene(args.n, spacing=0.1, jitter=1e-3, seed=0, E=1e5, nu=0.3, density=1000.0, pin_top=True)
    print(f"# scene: {scene.nv} vertices, {scene.nt} tets", file=sys.stderr, flush=True)
    halo = args.pcg_kernels == "halo"
    timing_steps = args.steps if args.timing_steps is None else args.timing_steps
    polar = args.rotation == "polar"
    g = run_gpu(scene, args.steps, args.warmup, 32, timing_steps=timing_steps, halo=halo, polar=polar,
                graph=args.graph)
    print(f"# gpu fp32: {args.steps} steps in {g['seconds']:.3f} s", file=sys.stderr, flush=True)
    sps = args.steps / g["seconds"]

    err = None
    if not args.no_fp64:
        r = run_gpu(scene, args.steps, args.warmup, 64, timing_steps=0, halo=halo)
        print(f"# gpu fp64 reference run: {r['seconds']:.3f} s", file=sys.stderr, flush=True)
        disp = np.abs(r["X"] - scene.X).max()
        err = dict(rel_l2=float(np.linalg.norm(g["X"] - r["X"]) / np.linalg.norm(r["X"])),
                   max_abs_m=float(np.abs(g["X"] - r["X"]).max()),
                   max_abs_over_max_displacement=float(np.abs(g["X"] - r["X"]).max() / disp),
                   reference="FP64 run of the same steps on the GPU (FP64 path parity-tested vs the FP64 oracle)",
                   steps_simulated=args.warmup + args.steps,
                   note="FP32 arithmetic of the reference's algorithm: the FP32 oracle is as far from FP64 "
                        "(tests/test_gpu_cube64.py); the distance grows with the displacement, about 1.4e-4 of it",
                   fp64_steps_per_s=round(args.steps / r["seconds"], 3))

    roof = None
    kt, tm = g["kernel_times"], g["timing"]
    if tm is not None:
        iters = tm["pcg_iters"]
        per_launch = {
            "local_step": local_step_bytes(scene.nt, scene.nv),
            # algorithmic bytes of the effective iterations spread over every
            # launch (the stop probe and launches after convergence move nothing)
            "pcg_iteration": iters * pcg_iteration_bytes(g["nnz_off"], scene.nv) / max(kt["pcg_iteration"][0], 1),
        }
        kinfo = {}
        for k, (n, ms) in kt.items():
            d = dict(launches=n, total_ms=round(ms, 3), avg_us=round(ms / max(n, 1) * 1e3, 2))
            if k in per_launch:
                d["bytes_per_launch"] = int(per_launch[k])
                d["gbs"] = round(per_launch[k] / (ms / max(n, 1) * 1e-3) / 1e9, 1)
            kinfo[k] = d
        dom = max(per_launch, key=lambda k: kt[k][1])
        ach = kinfo[dom]["gbs"]
        traffic = measured_traffic(dom if halo or dom != "pcg_iteration" else "pcg_iteration_gather")
        roof = dict(bound="hbm", kernel=dom, achieved=ach, peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(ach / HBM_PEAK_GBS, 4), traffic=traffic,
                    traffic_source=TRAFFIC_FILE if traffic is not None else None,
                    algorithmic_bytes_per_launch=kinfo[dom]["bytes_per_launch"], kernels=kinfo,
                    timing_sample=f"{tm['steps']} steps after the timed region (as many as it has), HIP events "
                                  f"around every local-step launch and around each solve's run of PCG iteration "
                                  f"launches, {iters} PCG iterations in {tm['pcg_solves']} solves")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(scene, args.cpu_budget, 10)
        print(f"# cpu baseline: {cpu['value']:.4f} steps/s", file=sys.stderr, flush=True)

    parity = None
    if not args.no_cpu and args.n == 64:
        parity = parity_vs_oracle(scene, 10, polar)
        print(f"# parity vs FP32 oracle after 10 steps: {parity['rel_l2']:.2e}", file=sys.stderr, flush=True)

    configs = None
    if not args.no_configs and args.n == 64:
        configs = {}
        for name, fn in (("config0_8cube_direct", lambda: config0_measure(100)),
                         ("config2_256_dense_bodies", lambda: dense_measure(8, 2)),
                         ("config3_1024_batched_bodies", lambda: batched_measure(10, 2, args.per_body_sample,
                                                                                 polar=args.rotation == "polar")),
                         ("config4_ipc_two_16cubes", lambda: ipc_measure(args.ipc_steps,
                                                                           0 if args.no_cpu else 3,
                                                                           args.cpu_budget))):
            try:
                configs[name] = fn()
            except Exception as e:  # a failed side workload is reported, the headline still prints
                configs[name] = {"error": f"{type(e).__name__}: {e}"}
            print(f"# {name} done", file=sys.stderr, flush=True)

    st = g["stats"]
    line = {
        "metric": METRIC, "value": round(sps, 3), "unit": "steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * g["seconds"] / args.steps, 4),
        "higher_is_better": True, "scaling": "strong", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{args.n}^3-hex cube PD step (Jacobi-PCG + Chebyshev, 20 it/step)",
                   "verts": scene.nv, "tets": scene.nt, "nnz_scalar": g["nnz_off"] + scene.nv,
                   "iterations_per_step": 20, "dt": 1.0 / 60.0, "pcg_tol": 1e-6, "pcg_floor": 1e-8,
                   "pcg_kernels": args.pcg_kernels, "halo_groups": int(g["stats"].halo_groups),
                   "local_step_rotation": "Newton polar iteration, TR1690 SVD fallback" if polar else "TR1690 SVD",
                   "graph_replays": g["graph_replays"], "graph_captures": g["graph_captures"]},
        "roofline": roof,
        "cpu_baseline": cpu,
        "parity_vs_fp32_oracle": parity,
        "error_vs_fp64": err,
        "pcg": {"iterations_per_solve": round(g["pcg_iters"] / max(g["pcg_solves"], 1), 3),
                "solves": g["pcg_solves"], "largest_launch_cap": st.pcg_cap,
                "unconverged_solves": g["unconverged"]},
        "configs": configs,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
