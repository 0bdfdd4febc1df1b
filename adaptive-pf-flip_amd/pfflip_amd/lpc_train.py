"""Train the learned preconditioner on the GPU.

Training data are residuals recorded by the solver itself.  A system is one pressure
projection of a two-phase scene on an n^3 grid (default 128^3, 8 jittered particles
per interior cell): Rayleigh-Taylor, or (one system in four) the bubble tank with its
empty headspace, with the light phase's density drawn log-uniform so that
rho_l / rho_g lies in [1, 1000] (rho_l = 1000).  Each scene is stepped with the
multigrid-preconditioned solve; from its second step on, every projection records its
FP32 residuals (one per host check, up to `snapshots`) and the network takes `steps`
Adam steps on them (pf_lpc_train, the library's own forward and backward kernels),
where M is the multigrid cycle with the network's per-aggregate correction weights.
The default objective is the energy-norm error reduction e1^T A e1 / e0^T A e0,
e1 = e0 - M A e0, over the cycle's slowest modes: each recorded residual after
`power_steps` preconditioned Richardson steps with the current M.  PCG's iteration
count follows the spread of the spectrum of M A, which these modes bound; the
residual objective ||r - A M(r)||^2 / ||r||^2 (--objective residual) weighs the
typical residual instead.  The default 40 scenes x 8 recorded
steps give about 2000 residual/system pairs.  After every `val_every` scenes the
network's solve iterations on three validation systems are counted, and the weights
with the fewest (the untrained network, which is the plain cycle, included) are the
ones written.  Held-out systems then compare the solve with those weights against the
plain cycle.  Seeds, ratios and results are written beside the weights.

Usage (on a GPU):
  python -m pfflip_amd.lpc_train --out weights.npy [--n 128 --scenes 40 ...]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
if os.path.dirname(HERE) not in sys.path:
    sys.path.insert(0, os.path.dirname(HERE))

import pfflip_amd as G  # noqa: E402

DEFAULT_WEIGHTS = os.path.join(HERE, "lpc_weights.npy")
DEFAULT_SEED = 20251


def load_weights(path=DEFAULT_WEIGHTS):
    """Trained weights as float32 (numpy format, no pickled objects)."""
    w = np.load(path, allow_pickle=False)
    if w.dtype != np.float32 or w.shape != (G.LearnedPreconditioner.num_weights(),):
        raise ValueError(f"{path}: expected {G.LearnedPreconditioner.num_weights()} float32 weights")
    return w


def system_phase(ratio):
    ph = G.Phase(1000.0, 1000.0 / ratio, 1.0, 1.0)
    ph.rho_tilde_0 = G.calibrate_rho_tilde0(ph, 1.0)
    return ph


def solve_cfg(precond, tol=1e-6, max_iters=20000):
    return G.SolveCfg(max_iters=max_iters, stop_rule=G.STOP_REL2, tol=tol, inner_rel_tol=0.01, max_outer=10000,
                      warm_start=0, check_every=4, precond=precond)


def step_cfg(ph, dt, precond=G.PRECOND_MULTIGRID):
    return G.StepCfg(dt=dt, gravity=-9.81, Vp=1.0, alpha_liquid=0.95, alpha_gas=0.95, adaptive_cfl=0.5,
                     adaptive_min_dt=1e-5, c_div=0.0, adaptive_timestep=0, advection_order=2, timing=0,
                     model=G.MODEL_TWO_PHASE, phase=ph, solve=solve_cfg(precond))


def compare(net, n, kind, seed, ratio, steps=2, repeats=2):
    """iterations and solve ms of the plain cycle and of the learned one on the same
    systems (each projection solved by both from the same pre-projection state)"""
    ctx = G.Context(n, n, n, 1.0 / n)
    try:
        ctx.scene(kind, seed)
        ph = system_phase(ratio)
        dt = 1e-3 if kind == G.SCENE_RT else 5e-4
        net.attach(ctx)
        out = []
        for _ in range(steps):
            ctx.step(step_cfg(ph, dt))
            row = {}
            for name, pre in (("plain", G.PRECOND_MULTIGRID), ("learned", G.PRECOND_LEARNED)):
                best = None
                for _ in range(repeats):
                    ctx.mark_cells(); ctx.p2g(ph, 1.0); ctx.save_velocity(); ctx.apply_gravity(dt, -9.81)
                    ctx.apply_walls()
                    ctx.synchronize()
                    t0 = time.perf_counter()
                    st = ctx.project(ph, dt, solve_cfg(pre))
                    ctx.synchronize()
                    ms = (time.perf_counter() - t0) * 1e3
                    best = ms if best is None else min(best, ms)
                row[name] = {"iterations": st.iterations, "ms": round(best, 3), "converged": st.converged,
                             "fallback": st.fallback_used, "residual": st.true_rel_residual}
            out.append(row)
        G.lib().pf_lpc_attach(ctx.h, None)
        return out
    finally:
        ctx.close()


# validation systems (kind, seed, ratio): seeds unused by training and by the held-out set
VALIDATION = ((G.SCENE_RT, 5, 300.0), (G.SCENE_RT, 9, 10.0), (G.SCENE_BUBBLE, 15, 500.0))


def validation_iterations(net, n):
    """summed iterations of the learned solve on the validation systems (one projection
    each, after one multigrid step of the scene)"""
    total = 0
    for kind, seed, ratio in VALIDATION:
        ctx = G.Context(n, n, n, 1.0 / n)
        try:
            ctx.scene(kind, seed)
            ph = system_phase(ratio)
            dt = 1e-3 if kind == G.SCENE_RT else 5e-4
            ctx.step(step_cfg(ph, dt))
            net.attach(ctx)
            ctx.mark_cells(); ctx.p2g(ph, 1.0); ctx.save_velocity(); ctx.apply_gravity(dt, -9.81)
            ctx.apply_walls()
            st = ctx.project(ph, dt, solve_cfg(G.PRECOND_LEARNED))
            total += st.iterations if st.converged and not st.fallback_used else 10 ** 6
            G.lib().pf_lpc_attach(ctx.h, None)
        finally:
            ctx.close()
    return total


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--scenes", type=int, default=40)
    ap.add_argument("--steps-per-scene", type=int, default=8, help="recorded projections per scene")
    ap.add_argument("--snapshots", type=int, default=16)
    ap.add_argument("--steps", type=int, default=24, help="Adam steps per recorded system")
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=DEFAULT_SEED)
    ap.add_argument("--init", default=None, help="start from these weights")
    ap.add_argument("--objective", choices=("energy", "residual"), default="energy")
    ap.add_argument("--power-steps", type=int, default=6)
    ap.add_argument("--val-every", type=int, default=4, help="scenes between validation counts")
    args = ap.parse_args(argv)
    rng = np.random.default_rng(args.seed)
    net = G.LearnedPreconditioner(seed=args.seed)
    net.set_objective(G.LPC_OBJECTIVE_ENERGY if args.objective == "energy" else G.LPC_OBJECTIVE_RESIDUAL,
                      args.power_steps if args.objective == "energy" else 0)
    if args.init:
        net.set_weights(load_weights(args.init))
    log = {"n": args.n, "scenes": [], "args": vars(args), "validation": []}
    t0 = time.time()
    pairs = 0
    best_it = validation_iterations(net, args.n)
    best_w, best_at = net.weights(), 0
    log["validation"].append({"scenes": 0, "iterations": best_it})
    print(f"[lpc_train] validation before training: {best_it} it", flush=True)
    for s in range(args.scenes):
        ratio = float(10.0 ** rng.uniform(0.0, 3.0))
        scene_seed = int(rng.integers(1, 2 ** 31))
        kind = G.SCENE_BUBBLE if s % 4 == 3 else G.SCENE_RT
        dt = 1e-3 if kind == G.SCENE_RT else 5e-4
        ph = system_phase(ratio)
        ctx = G.Context(args.n, args.n, args.n, 1.0 / args.n)
        try:
            ctx.scene(kind, scene_seed)
            ctx.step(step_cfg(ph, dt))
            net.attach(ctx)
            lr = args.lr * (0.5 ** (s // 10))
            rec = {"kind": kind, "ratio": ratio, "scene_seed": scene_seed, "systems": []}
            for k in range(args.steps_per_scene):
                G.lpc_record(ctx, args.snapshots, 1)
                st = ctx.step(step_cfg(ph, dt))
                ns = G.lpc_snapshot_count(ctx)
                pairs += ns
                loss = net.train(ctx, args.steps, lr=lr, seed=scene_seed + k)
                probe = [net.loss(ctx, q) for q in range(ns)]
                rec["systems"].append({"snapshots": ns, "iterations": st.solve.iterations, "train_loss": loss,
                                       "loss_learned": float(np.mean([a for a, _ in probe])),
                                       "loss_plain": float(np.mean([b for _, b in probe]))})
            G.lpc_record(ctx, 0, 1)
            G.lib().pf_lpc_attach(ctx.h, None)
        finally:
            ctx.close()
        last = rec["systems"][-1]
        print(f"[lpc_train] scene {s} ({'bubble' if kind else 'RT'}, ratio {ratio:.1f}): {pairs} pairs so far; "
              f"last system {last['iterations']} it, loss learned {last['loss_learned']:.5f} vs plain "
              f"{last['loss_plain']:.5f} ({time.time() - t0:.0f} s)", flush=True)
        log["scenes"].append(rec)
        if (s + 1) % args.val_every == 0 or s + 1 == args.scenes:
            it = validation_iterations(net, args.n)
            log["validation"].append({"scenes": s + 1, "iterations": it})
            print(f"[lpc_train] validation after {s + 1} scenes: {it} it (best {best_it})", flush=True)
            if it < best_it:
                best_it, best_w, best_at = it, net.weights(), s + 1
    log["pairs"] = pairs
    log["selected"] = {"scenes": best_at, "validation_iterations": best_it}
    net.set_weights(best_w)
    np.save(args.out, best_w)
    print(f"[lpc_train] selected the weights after {best_at} scenes ({best_it} validation it)", flush=True)
    # held-out systems: seeds and ratios not used above
    evals = []
    for kind, seed, ratio in ((G.SCENE_RT, 7, 1000.0), (G.SCENE_RT, 11, 30.0), (G.SCENE_BUBBLE, 13, 1000.0),
                              (G.SCENE_BUBBLE, 17, 100.0)):
        res = compare(net, args.n, kind, seed, ratio)
        evals.append({"kind": kind, "seed": seed, "ratio": ratio, "steps": res})
        for row in res:
            print(f"[lpc_train] held-out {'bubble' if kind else 'RT'} ratio {ratio:g}: plain "
                  f"{row['plain']['iterations']} it {row['plain']['ms']} ms, learned {row['learned']['iterations']} it "
                  f"{row['learned']['ms']} ms", flush=True)
    log["held_out"] = evals
    with open(os.path.splitext(args.out)[0] + "_train.json", "w") as f:
        json.dump(log, f, indent=1)
    print(f"[lpc_train] wrote {args.out} ({pairs} pairs, {time.time() - t0:.0f} s)", flush=True)


if __name__ == "__main__":
    main()
