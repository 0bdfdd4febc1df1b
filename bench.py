#!/usr/bin/env python3
"""Benchmark of the sparse two-phase PF-FLIP step on one MI355X.

Workload (the headline metric "step ms @256^3 two-phase"): Rayleigh-Taylor on a
256^3 MAC grid of the unit box, 8 jittered particles per interior cell
(131,096,512 FP32 particles), heavy rho = 1000 above
y = 0.5 + 0.05 cos(2 pi x) cos(2 pi z), light rho = 1 below, zero initial
velocity, g = -9.81, dt = 1e-3, FLIP 0.95, pressure solved until
||b - A x||_2 / ||b||_2 <= 1e-6 measured in FP64 (--precond: aggregation-multigrid
PCG by default, or the reference's Jacobi PCG; both stop on that same FP64
residual).  Synthetic data from the seeded scene generator; inputs are resident in
HBM before the timed region.

One "step" is SparseSim3DTP::step (bin, mark, P2G, save/gravity/walls, assembly,
solve, velocity update, G2P + RK2 advection).  W untimed warm-up steps, then K
timed steps bracketed by device synchronisation.  Prints one JSON line.

The roofline object is for the step's dominant kernel (the one with the largest
summed event time in the per-kernel table: the P2G splat), timed live with HIP events
on the library stream inside the timed steps.

The CPU baseline times the repository's C oracle in float-faithful mode (the same
results as the reference's SparseSim3DTP::step, bit for bit, on identical state;
8 P2G threads as the reference's 8 colour tasks, the other loops on one thread) on a
bounded sample: the same scene construction at 64^3 (1,906,624 particles) with the
reference's Jacobi PCG and stop rule, next to the GPU running the same 64^3 step with
the same solver and stop rule; the 256^3 figure scales the sample by the particle
count and is an estimate.  It is a reported baseline, not the target.

Secondary lines: the 256^3 step with the reference's Jacobi preconditioner, the
bubble tank at 128^3 and the dam-break at 256^3 (200 CFL-1 steps are the test; the
bench times a few), each with per-stage ms and algorithmic GB/s, and the multires
multi-bubble tank at 512^3 effective (three levels) with per-stage ms.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "adaptive-pf-flip_amd"))

HBM_PEAK_GBS = 8000.0
SPMV_BYTES_PER_CELL = 32  # z, p_old, diag, 3 couplings read; p_new, Ap written (FP32)
# algorithmic bytes per stage (FP32 = 4 B): per particle and per FLUID cell
P2G_BYTES = (25, 24)        # particle x, y, z, u, v, w, type read; u, v, w, mu, mv, mw written
STAGE_BYTES = {             # (per particle, per fluid cell, per grid cell)
    "bin": (58, 0, 0),      # particle read + sorted write (rank and counts not counted)
    "mark": (0, 0, 5),      # counts read, marker written
    "p2g": (25, 24, 0),
    "grid": (0, 60, 0),     # 3 axes: velocity + mass read, velocity + saved written
    "assemble": (0, 45, 0),
    "velocity": (0, 40, 0),
    "g2p_advect": (49, 24, 0),
}


def read_traffic(path):
    """HBM bytes per launch from a counter summary line 'hbm_bytes_per_launch: N'."""
    for line in open(path):
        if line.strip().startswith("hbm_bytes_per_launch:"):
            return int(float(line.split(":", 1)[1].strip().strip("`")))
    return None


def read_table_value(path, label):
    """The value column of the summary table row whose first cell starts with label."""
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) >= 2 and cells[0].startswith(label):
            try:
                return float(cells[-1])
            except ValueError:
                return None
    return None


def _ensure_built():
    import pfflip_amd as G
    if not os.path.exists(G.LIB_PATH):
        import importlib.util
        spec = importlib.util.spec_from_file_location("pf_build", os.path.join(ROOT, "adaptive-pf-flip_amd", "build.py"))
        m = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(m)
        m.build_lib()
    return G


# per preconditioner: (pf_precond, FP32 reduction per FP64 refinement pass, iterations
# between host checks) -- a multigrid iteration costs about four Jacobi ones, so its
# passes go deeper and the host looks more often.  Measured at 256^3 (multigrid): a
# 1e-4 reduction per pass takes 2 replacements and 20 iterations (18.4 ms), 1e-2 four
# and 18 (19.2 ms), 1e-3 three and 26 (24.1 ms: its restart falls mid-convergence)
SOLVERS = {"jacobi": (0, 0.1, 16), "multigrid": (2, 1e-4, 4), "learned": (1, 1e-4, 4)}
SOLVER_NAMES = {"jacobi": "plain Jacobi PCG", "multigrid": "aggregation-multigrid PCG",
                "learned": "multigrid PCG with learned correction weights"}


INNER_OVERRIDE = None  # --inner-rel-tol


def solve_cfg(G, precond, tol, max_iters, warm=0):
    pre, inner, check = SOLVERS[precond]
    if INNER_OVERRIDE is not None and precond != "jacobi":
        inner = INNER_OVERRIDE
    return G.SolveCfg(max_iters=max_iters, stop_rule=G.STOP_REL2, tol=tol, inner_rel_tol=inner,
                      max_outer=10000, warm_start=warm, check_every=check, precond=pre)


def step_cfg(G, n, tol, max_iters, warm, precond="jacobi"):
    ph = G.Phase(1000.0, 1.0, 1.0, 1.0)
    ph.rho_tilde_0 = G.calibrate_rho_tilde0(ph, 1.0)
    solve = solve_cfg(G, precond, tol, max_iters, warm)
    return G.StepCfg(dt=1e-3, gravity=-9.81, Vp=1.0, alpha_liquid=0.95, alpha_gas=0.95,
                     adaptive_cfl=0.5, adaptive_min_dt=1e-5, c_div=0.0, adaptive_timestep=0,
                     advection_order=2, timing=1, phase=ph, solve=solve)


def pressure_comparison(G, ctx, cfg, tol, learned=True):
    """Pressure solve to ||b - Ax||_2/||b||_2 <= tol (FP64) on the current system:
    plain Jacobi PCG, multigrid PCG and the learned preconditioner (the multigrid cycle
    with the network's correction weights) inside flexible PCG, same inputs.  The
    network has the package's trained weights when there are any, else the untrained
    network with its zero head (exactly the plain cycle's weights, plus the network's
    cost).  The pre-projection state (mark, P2G, save, gravity, walls) is rebuilt for
    each solve."""
    from pfflip_amd import lpc_train
    net, trained = None, False
    if learned:
        net = G.LearnedPreconditioner(lpc_train.DEFAULT_SEED)
        if os.path.exists(lpc_train.DEFAULT_WEIGHTS):
            net.set_weights(lpc_train.load_weights())
            trained = True
        net.attach(ctx)
    dt, ph = cfg.dt, cfg.phase

    def prepare():
        ctx.mark_cells()
        ctx.p2g(ph, cfg.Vp)
        ctx.save_velocity()
        ctx.apply_gravity(dt, cfg.gravity)
        ctx.apply_walls()

    runs = [("plain_pcg", "jacobi"), ("multigrid_pcg", "multigrid"), ("multigrid_pcg", "multigrid")]
    if net is not None:
        # the first learned solve also allocates the network's buffers; keep the second
        runs += [("learned_pc", "learned"), ("learned_pc", "learned")]
    out = {"learned_pc": None}
    for name, pre in runs:
        prepare()
        solve = solve_cfg(G, pre, tol, 50000)
        ctx.synchronize()
        t0 = time.perf_counter()
        st = ctx.project(ph, dt, solve)
        ctx.synchronize()
        ms = (time.perf_counter() - t0) * 1e3
        out[name] = {"ms": round(ms, 2), "iterations": st.iterations, "outer": st.outer_iterations,
                     "true_rel_residual": st.true_rel_residual, "converged": st.converged,
                     "fallback_to_jacobi": st.fallback_used}
        print(f"[bench] pressure solve {name}: {ms:.1f} ms, {st.iterations} it, fallback {st.fallback_used}",
              file=sys.stderr, flush=True)
    out["speedup_multigrid_vs_plain"] = round(out["plain_pcg"]["ms"] / out["multigrid_pcg"]["ms"], 3)
    if net is not None:
        out["speedup_learned_vs_plain"] = round(out["plain_pcg"]["ms"] / out["learned_pc"]["ms"], 3)
        out["speedup_learned_vs_multigrid"] = round(out["multigrid_pcg"]["ms"] / out["learned_pc"]["ms"], 3)
        out["learned_weights"] = "trained" if trained else (
            "untrained (zero head): training with validation selection kept no weights better than the plain cycle")
        G.lib().pf_lpc_attach(ctx.h, None)
        net.close()
    else:
        out["note"] = "learned preconditioner skipped"
    return out


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    if not os.path.exists(O.ORACLE_SO):
        O.build(ref=False)
    return O


def cpu_baseline(n_sample, n_full, seed):
    """The float-faithful C oracle's SparseSim3DTP::step on the n_sample^3 RT scene
    (reference stop rule |r|_inf < 1e-7, 600 Jacobi PCG iterations, P2G on 8 threads,
    the rest on one), scaled to n_full^3 by the particle count."""
    O = _oracle()
    dx = 1.0 / n_sample
    x, y, z, t = O.scene(O.SCENE_RT, n_sample, n_sample, n_sample, dx, seed)
    ph = O.Phase(1000.0, 1.0, 1.0, 1.0)
    ph.rho_tilde_0 = O.calibrate_rho_tilde0(ph, 1.0)
    cfg = O.StepCfg(dt=1e-3, gravity=-9.81, Vp=1.0, alpha_liquid=0.95, alpha_gas=0.95,
                    adaptive_timestep=0, advection_order=2, cg_iters=600, stop_rule=O.STOP_REF_MAXABS,
                    adaptive_cfl=0.5, adaptive_min_dt=1e-5, cg_tol=1e-7, c_div_volume_correction=0,
                    c_div_strength=1.0, liquid_volume_target=0.0, phase=ph)
    ps = O.Particles(x, y, z, type=t)
    scale = O.lib().orc_scene_count(O.SCENE_RT, n_full, n_full, n_full) / len(x)
    O.set_threads_split(8, 1)
    try:
        g = O.Grid(n_sample, n_sample, n_sample, dx, faithful=True)
        t0 = time.perf_counter()
        st = g.step(ps, cfg)
        ms = (time.perf_counter() - t0) * 1e3
    finally:
        O.set_threads(1)
    stages = {"mark": st.ms_mark, "p2g": st.ms_p2g, "save_gravity_walls": st.ms_save_gravity_walls,
              "project": st.ms_project, "g2p": st.ms_g2p, "advect": st.ms_advect}
    return {
        "value": round(ms * scale, 1),
        "unit": f"ms/step (estimated at {n_full}^3 from the {n_sample}^3 sample)",
        "cores": 8,
        "kind": "port",
        "sample": (f"one SparseSim3DTP::step of the {n_sample}^3 RT scene ({len(x)} particles) by the "
                   f"float-faithful C oracle (bit-exact to the reference step on identical state): "
                   f"{ms:.0f} ms on the host, Jacobi PCG {st.solve.iterations} it, reference stop rule "
                   f"|r|_inf < 1e-7 / 600 iterations; P2G on 8 threads (the reference's 8 colour tasks), "
                   f"the rest on one; scaled x{scale:.2f} by particle count"),
        "sample_ms": round(ms, 1),
        "sample_stage_ms": {k: round(v, 1) for k, v in stages.items()},
        "sample_solve_iterations": st.solve.iterations,
        "nproc": os.cpu_count(),
    }


def gpu_sample_pair(G, n, seed, steps=3):
    """the same n^3 RT step on the GPU with the reference's solver and stop rule
    (Jacobi PCG, |r|_inf < 1e-7, 600 iterations): the GPU side of the CPU pair"""
    ctx = G.Context(n, n, n, 1.0 / n)
    try:
        ctx.scene(G.SCENE_RT, seed)
        ph = G.Phase(1000.0, 1.0, 1.0, 1.0)
        ph.rho_tilde_0 = G.calibrate_rho_tilde0(ph, 1.0)
        solve = G.SolveCfg(max_iters=600, stop_rule=G.STOP_REF_MAXABS, tol=1e-7, inner_rel_tol=0.1,
                           max_outer=10000, warm_start=0, check_every=16, precond=0)
        cfg = G.StepCfg(dt=1e-3, gravity=-9.81, Vp=1.0, alpha_liquid=0.95, alpha_gas=0.95, adaptive_cfl=0.5,
                        adaptive_min_dt=1e-5, c_div=0.0, adaptive_timestep=0, advection_order=2, timing=0,
                        phase=ph, solve=solve)
        first = ctx.step(cfg)  # the state the CPU sample steps from
        ctx.synchronize()
        ctx2 = G.Context(n, n, n, 1.0 / n)
        try:
            ctx2.scene(G.SCENE_RT, seed)
            ctx2.synchronize()
            t0 = time.perf_counter()
            st = ctx2.step(cfg)
            ctx2.synchronize()
            ms = (time.perf_counter() - t0) * 1e3
        finally:
            ctx2.close()
        return {"ms": round(ms, 3), "solve_iterations": st.solve.iterations,
                "first_step_iterations": first.solve.iterations, "converged": st.solve.converged}
    finally:
        ctx.close()


def stage_gbs(stage_ms, np_, nc, ncell_total):
    out = {}
    for k, (bp, bc, bg) in STAGE_BYTES.items():
        ms = stage_ms.get(k, 0.0)
        if ms > 0:
            out[k] = round((bp * np_ + bc * nc + bg * ncell_total) / (ms * 1e-3) / 1e9, 1)
    return out


def run_workload(G, kind, n, seed, dt, steps, warmup, precond, adaptive=False, cap=None, label=""):
    """ms/step, per-stage ms and algorithmic GB/s of a scene at n^3"""
    ctx = G.Context(n, n, n, 1.0 / n)
    try:
        ctx.scene(kind, seed)
        cfg = step_cfg(G, n, 1e-6, 50000, 0, precond)
        cfg.dt = cap if cap else dt
        if adaptive:
            cfg.adaptive_timestep, cfg.adaptive_cfl = 1, 1.0
        for _ in range(warmup):
            ctx.step(cfg)
        ctx.synchronize()
        stats = []
        t0 = time.perf_counter()
        for _ in range(steps):
            stats.append(ctx.step(cfg))
        ctx.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        stage = {name: round(sum(s.stage_ms[i] for s in stats) / len(stats), 3) for i, name in enumerate(G.STAGES)}
        nc = stats[-1].solve.fluid_cells
        print(f"[bench] {label}: {ms:.2f} ms/step", file=sys.stderr, flush=True)
        return {"workload": label, "ms_per_step": round(ms, 3), "steps": steps, "warmup": warmup,
                "particles": ctx.count(), "fluid_cells": nc, "stage_ms": stage,
                "stage_gbs": stage_gbs(stage, ctx.count(), nc, n ** 3),
                "solve_iterations": [s.solve.iterations for s in stats],
                "true_rel_residual_max": max(s.solve.true_rel_residual for s in stats),
                "effective_dt": [round(s.effective_dt, 8) for s in stats]}
    finally:
        ctx.close()


def run_multires(G, n, seed, dt, steps, warmup, mode, label=""):
    """ms/step and per-stage ms of the multires multi-bubble tank at n^3 effective: three
    levels (level 0 within 6 cells of the free surface and of 16 bubbles, level 1 within
    22, level 2 beyond), static layout, the multires solve in `mode` (multigrid cycle or
    the plain Jacobi PCG) to the reference's stop rule |r|_inf < 1e-6 |r0|_inf"""
    sim = G.MRSim3DTP(n, n, n, 1.0 / n)
    try:
        sim.ctx.scene_bubbles(seed, 16, 6)
        levels = np.bincount(sim.ctx.layout()[:, 0], minlength=3)
        ph = G.Phase(1000.0, 1.0, 1.0, 1.0)
        ph.rho_tilde_0 = G.calibrate_rho_tilde0(ph, 1.0)
        sim.phase, sim.dt, sim.dynamic_refinement, sim.timing = ph, dt, False, True
        sim.cg_iters, sim.cg_tol, sim.cg_rel_tol = 50000, 0.0, 1e-6
        sim.solve_options = {"mode": mode}
        for _ in range(warmup):
            sim.step()
        sim.ctx.synchronize()
        stats = []
        t0 = time.perf_counter()
        for _ in range(steps):
            stats.append(sim.step())
        sim.ctx.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        stage = {name: round(sum(s.stage_ms[i] for s in stats) / len(stats), 3) for i, name in enumerate(G.STAGES)}
        print(f"[bench] {label}: {ms:.2f} ms/step", file=sys.stderr, flush=True)
        return {"workload": label, "ms_per_step": round(ms, 3), "steps": steps, "warmup": warmup,
                "leaves_per_level": [int(v) for v in levels], "particles": sim.ctx.count(),
                "fluid_cells": stats[-1].solve.fluid_cells, "stage_ms": stage,
                "solve_iterations": [s.solve.iterations for s in stats],
                "converged": [s.solve.converged for s in stats],
                "true_max_residual_rel": max(s.solve.true_max_residual / max(s.solve.initial_max_residual, 1e-300)
                                             for s in stats),
                "true_rel_residual_max": max(s.solve.true_rel_residual for s in stats)}
    finally:
        sim.ctx.close()


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=256, help="grid cells per axis")
    ap.add_argument("--seed", type=int, default=20250601)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--max-iters", type=int, default=50000)
    ap.add_argument("--warm-start", type=int, default=0)
    ap.add_argument("--cpu-sample-n", type=int, default=64)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-kernel-timing", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned preconditioner in the comparison")
    ap.add_argument("--no-pressure", action="store_true", help="skip the pressure-solver comparison")
    ap.add_argument("--multires-only", action="store_true",
                    help="run only the multires 512^3-effective multigrid line and print its JSON")
    ap.add_argument("--no-workloads", action="store_true",
                    help="skip the secondary lines (Jacobi step, bubble 128^3, dam-break 256^3)")
    ap.add_argument("--inner-rel-tol", type=float, default=None,
                    help="FP32 residual reduction per FP64 refinement pass of the cycle solvers")
    ap.add_argument("--precond", choices=sorted(SOLVERS), default="multigrid",
                    help="preconditioner of the timed steps' pressure solve")
    args = ap.parse_args()
    global INNER_OVERRIDE
    INNER_OVERRIDE = args.inner_rel_tol
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (this library has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)

    G = _ensure_built()
    if args.multires_only:
        # the multires multigrid line alone (for profiling it): --steps timed steps
        r = run_multires(G, 512, args.seed, 5e-4, args.steps, args.warmup, G.MR_SOLVE_MULTIGRID,
                         label="multires multi-bubble tank 512^3 effective, 3 levels, multigrid cycle on V A")
        print(json.dumps(r), flush=True)
        return
    n = args.n
    ctx = G.Context(n, n, n, 1.0 / n)
    ctx.scene(G.SCENE_RT, args.seed)
    cfg = step_cfg(G, n, args.tol, args.max_iters, args.warm_start, args.precond)
    step_net = None
    if args.precond == "learned":
        from pfflip_amd import lpc_train
        step_net = G.LearnedPreconditioner(lpc_train.DEFAULT_SEED)
        if os.path.exists(lpc_train.DEFAULT_WEIGHTS):
            step_net.set_weights(lpc_train.load_weights())
        step_net.attach(ctx)
    print(f"[bench] {n}^3 RT, {ctx.count()} particles; warmup {args.warmup}", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        st = ctx.step(cfg)
        print(f"[bench] warmup step: {st.total_ms:.1f} ms, solve {st.solve.iterations} it", file=sys.stderr,
              flush=True)
    ctx.synchronize()
    # inside the timed steps only the roofline kernels carry events (a pair per
    # launch: the P2G colour launches and the SpMV); the per-kernel table comes from
    # one more step afterwards
    if not args.no_kernel_timing:
        ctx.kernel_timing(True, ["p2g", "pcg_spmv"])
    stats = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        stats.append(ctx.step(cfg))
        print(f"[bench] step {i + 1}/{args.steps} issued", file=sys.stderr, flush=True)
    ctx.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    print(f"[bench] timed {args.steps} steps: {ms_per_step:.2f} ms/step", file=sys.stderr, flush=True)

    stage = {name: round(sum(s.stage_ms[i] for s in stats) / len(stats), 3) for i, name in enumerate(G.STAGES)}
    iters = [s.solve.iterations for s in stats]
    fluid = stats[-1].solve.fluid_cells
    npart = ctx.count()
    kernels = {}
    roofline = roofline_spmv = None
    if not args.no_kernel_timing:
        live = {}
        for name in ("p2g", "pcg_spmv"):
            nl, ms = ctx.kernel_time(G.KERNELS.index(name))
            live[name] = (nl, ms)
        print("[bench] per-kernel event times over one more step", file=sys.stderr, flush=True)
        ctx.kernel_timing(True)
        ctx.step(cfg)
        for k, name in enumerate(G.KERNELS):
            nl, ms = ctx.kernel_time(k)
            if nl:
                kernels[name] = {"launches": nl, "avg_ms": round(ms / nl, 5), "total_ms": round(ms, 3)}
        ctx.kernel_timing(False)
        # (mg_apply, lpc_apply, interface and surface bracket a whole cycle or stage of
        # several kernels: not single kernels)
        single = {k: v for k, v in kernels.items() if k not in ("mg_apply", "lpc_apply", "interface", "surface")}
        dominant = max(single, key=lambda k: single[k]["total_ms"]) if single else None
        nl, ms = live["p2g"]
        if nl:
            # one launch per colour: each carries an eighth of the step's algorithmic P2G bytes
            avg = ms / nl
            per_launch = (P2G_BYTES[0] * npart + P2G_BYTES[1] * fluid) / 8
            achieved = per_launch / (avg * 1e-3) / 1e9
            tf = os.path.join(ROOT, "profiles", "p2g_color_rt256_traffic.md")
            roofline = {"bound": "hbm", "kernel": "k_p2g_color<0, true>", "achieved": round(achieved, 1),
                        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                        "traffic": read_traffic(tf) if (os.path.exists(tf) and n == 256) else None,
                        "bytes_per_launch": int(per_launch), "avg_launch_ms": round(avg, 5),
                        "launches_timed": nl, "dominant_kernel_in_step_table": dominant,
                        "note": "bytes = (25 B/particle + 24 B/fluid cell) / 8 colour launches; the kernel is "
                                "VALU-bound: valu_busy and waves_per_simd are the SQ counters of the committed "
                                "256^3 counter summary"}
            if os.path.exists(tf) and n == 256:
                roofline["valu_busy"] = read_table_value(tf, "VALU busy")
                roofline["waves_per_simd"] = read_table_value(tf, "resident waves per SIMD")
        nl, ms = live["pcg_spmv"]
        if nl:
            bytes_per_launch = SPMV_BYTES_PER_CELL * fluid
            avg = ms / nl
            achieved = bytes_per_launch / (avg * 1e-3) / 1e9
            tf = os.path.join(ROOT, "profiles", "pcg_spmv_rt256_traffic.md")
            roofline_spmv = {"bound": "hbm", "kernel": "k_pcg_spmv", "achieved": round(achieved, 1),
                             "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                             "traffic": read_traffic(tf) if (os.path.exists(tf) and n == 256) else None,
                             "bytes_per_launch": bytes_per_launch, "avg_launch_ms": round(avg, 5)}
    # achievable HBM bandwidth of a plain copy kernel on this box (reported beside
    # the 8 TB/s peak; also the known-size kernel for calibrating HBM counters)
    print("[bench] copy probe", file=sys.stderr, flush=True)
    ncopy = 256 << 20  # 1 GiB read + 1 GiB written: past every cache
    copy_ms = min(ctx.copy_probe(ncopy) for _ in range(3))
    copy_gbs = round(2 * 4 * ncopy / (copy_ms * 1e-3) / 1e9, 1)
    pressure = None
    if not args.no_pressure:
        pressure = pressure_comparison(G, ctx, cfg, args.tol, not args.no_learned)
    ctx.close()
    if step_net is not None:
        step_net.close()
    workloads = {}
    if not args.no_workloads:
        workloads["rt256_jacobi"] = run_workload(
            G, G.SCENE_RT, n, args.seed, 1e-3, 2, 1, "jacobi",
            label=f"RT {n}^3 with the reference's Jacobi preconditioner, FP64 residual <= 1e-6")
        workloads["bubble128"] = run_workload(
            G, G.SCENE_BUBBLE, 128, args.seed, 5e-4, 10, 2, args.precond,
            label="bubble tank 128^3, 1000:1, dt 5e-4")
        workloads["dambreak256"] = run_workload(
            G, G.SCENE_DAMBREAK, 256, args.seed, None, 20, 5, args.precond, adaptive=True, cap=5e-3,
            label="dam-break 256^3 (liquid only), CFL 1, requested dt cap 5e-3")
        mr_label = "multires multi-bubble tank 512^3 effective, 3 levels, 16 bubbles, 1000:1, dt 5e-4"
        workloads["multires512"] = run_multires(G, 512, args.seed, 5e-4, 3, 1, G.MR_SOLVE_MULTIGRID,
                                                label=mr_label + ", multigrid cycle on V A")
        workloads["multires512_jacobi"] = run_multires(G, 512, args.seed, 5e-4, 1, 1, G.MR_SOLVE_FAST,
                                                       label=mr_label + ", plain Jacobi PCG")
    cpu = pair = None
    if not args.no_cpu_baseline:
        print(f"[bench] gpu side of the {args.cpu_sample_n}^3 pair", file=sys.stderr, flush=True)
        pair = {"workload": f"RT {args.cpu_sample_n}^3, Jacobi PCG, |r|_inf < 1e-7, 600 iterations (the "
                            f"reference's solver and stop rule on both sides)",
                "gpu": gpu_sample_pair(G, args.cpu_sample_n, args.seed)}
        print("[bench] cpu baseline (float-faithful oracle on the 64^3 sample) ...", file=sys.stderr, flush=True)
        cpu = cpu_baseline(args.cpu_sample_n, n, args.seed)
        print(f"[bench] cpu sample {cpu['sample_ms']} ms", file=sys.stderr, flush=True)
        pair["cpu"] = {"ms": cpu["sample_ms"], "solve_iterations": cpu["sample_solve_iterations"],
                       "threads": "P2G 8, rest 1"}
        pair["speedup"] = round(cpu["sample_ms"] / pair["gpu"]["ms"], 1)

    line = {
        "metric": "step ms @256\u00b3 two-phase; pressure ms to 1e-6 resid, learned-PC vs plain PCG",
        "value": round(ms_per_step, 3),
        "unit": "ms/step",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (FP64 dots, scalars and refinement)",
        "data": "synthetic (seeded RT scene)",
        "config": {"workload": f"sparse two-phase Rayleigh-Taylor {n}^3, rho 1000:1, dt 1e-3, "
                               f"{npart} particles, {SOLVER_NAMES[args.precond]} + FP64 refinement to "
                               f"{args.tol:g} rel. residual",
                   "grid": n, "particles": npart, "fluid_cells": fluid,
                   "solver": f"{args.precond}-pcg-fp32+ir-fp64", "warm_start": bool(args.warm_start)},
        "stage_ms": stage,
        "stage_gbs": stage_gbs(stage, npart, fluid, n ** 3),
        "pressure_ms_to_1e-6": pressure,
        "solve_iterations": iters,
        "true_rel_residual": [s.solve.true_rel_residual for s in stats],
        "kernels": kernels,
        "kernels_note": "event times of every kernel over one step run after the timed steps; inside the "
                        "timed steps only the P2G and SpMV launches carry events (for the rooflines)",
        "roofline": roofline,
        "roofline_spmv": roofline_spmv,
        "copy_kernel_gbs": copy_gbs,
        "workloads": workloads,
        "cpu_gpu_pair": pair,
        "cpu_baseline": cpu,
    }
    if cpu:
        line["speedup_vs_cpu_estimate"] = round(cpu["value"] / ms_per_step, 1)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
