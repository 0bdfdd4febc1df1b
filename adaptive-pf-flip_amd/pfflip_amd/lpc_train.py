"""Train the learned block preconditioner on the GPU.

Training data are residuals recorded by the solver itself: for each system, a
two-phase Rayleigh-Taylor scene on an n^3 grid (default 128^3, 8 jittered particles
per interior cell) with the light phase's density drawn log-uniform so that
rho_l / rho_g lies in [1, 1000] (rho_l = 1000), one step is taken with the network
attached and recording, and the plain Jacobi PCG's FP32 residuals (one per `every`
iterations, up to `snapshots`) become that system's training set.  The network is
trained with Adam on loss ||r - A M(r)||^2 / ||r||^2 using the library's own forward
and backward kernels (pf_lpc_train).  With 10 systems x 200 snapshots this is the
~2000 residual/system pairs the workload calls for.  The seeds and ratios are
printed and written beside the weights.

Usage (on a GPU):
  python -m pfflip_amd.lpc_train --out weights.npy [--n 128 --systems 10 ...]
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


def step_cfg(ph, max_iters=20000, tol=1e-6, precond=G.PRECOND_JACOBI):
    solve = G.SolveCfg(max_iters=max_iters, stop_rule=G.STOP_REL2, tol=tol, inner_rel_tol=0.1, max_outer=10000,
                       warm_start=0, check_every=8, precond=precond)
    return G.StepCfg(dt=1e-3, gravity=-9.81, Vp=1.0, alpha_liquid=0.95, alpha_gas=0.95, adaptive_cfl=0.5,
                     adaptive_min_dt=1e-5, c_div=0.0, adaptive_timestep=0, advection_order=2, timing=0,
                     model=G.MODEL_TWO_PHASE, phase=ph, solve=solve)


def harvest(net, n, seed, ratio, snapshots, every, warm_steps=1):
    """A context whose last projection recorded `snapshots` residuals."""
    ctx = G.Context(n, n, n, 1.0 / n)
    ctx.scene(G.SCENE_RT, seed)
    ph = system_phase(ratio)
    cfg = step_cfg(ph)
    for _ in range(warm_steps):
        ctx.step(cfg)
    net.attach(ctx)
    G.lpc_record(ctx, snapshots, every)
    st = ctx.step(cfg)
    return ctx, st


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--systems", type=int, default=10)
    ap.add_argument("--snapshots", type=int, default=200)
    ap.add_argument("--every", type=int, default=4)
    ap.add_argument("--steps", type=int, default=300, help="Adam steps per system")
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=20251)
    args = ap.parse_args(argv)
    rng = np.random.default_rng(args.seed)
    net = G.LearnedPreconditioner(seed=args.seed)
    log = {"n": args.n, "systems": [], "args": vars(args)}
    t0 = time.time()
    for s in range(args.systems):
        ratio = float(10.0 ** rng.uniform(0.0, 3.0))
        scene_seed = int(rng.integers(1, 2 ** 31))
        ctx, st = harvest(net, args.n, scene_seed, ratio, args.snapshots, args.every)
        ns = G.lpc_snapshot_count(ctx)
        lr = args.lr * (0.5 ** (s // 4))
        loss = net.train(ctx, args.steps, lr=lr, seed=scene_seed)
        probe = [net.loss(ctx, k) for k in range(0, ns, max(1, ns // 8))]
        rec = {"ratio": ratio, "scene_seed": scene_seed, "snapshots": ns, "jacobi_iterations": st.solve.iterations,
               "train_loss": loss, "loss_learned": float(np.mean([a for a, _ in probe])),
               "loss_jacobi": float(np.mean([b for _, b in probe]))}
        log["systems"].append(rec)
        print(f"[lpc_train] system {s}: ratio {ratio:.1f}, {ns} snapshots, Jacobi {st.solve.iterations} it, "
              f"train loss {loss:.4f}, eval learned {rec['loss_learned']:.4f} vs Jacobi {rec['loss_jacobi']:.4f} "
              f"({time.time() - t0:.0f} s)", flush=True)
        ctx.close()
    w = net.weights()
    np.save(args.out, w)
    with open(os.path.splitext(args.out)[0] + "_train.json", "w") as f:
        json.dump(log, f, indent=1)
    print(f"[lpc_train] wrote {args.out}", flush=True)


if __name__ == "__main__":
    main()
