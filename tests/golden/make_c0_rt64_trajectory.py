#!/usr/bin/env python3
"""Writes c0_rt64_trajectory.npz: 100 steps of the 64^3 two-phase Rayleigh-Taylor
workload (1,906,624 particles, rho 1000 above the perturbed interface, 1 below,
dt 1e-3, g -9.81, FLIP 0.95) run by the FP64 oracle (oracle/pf_oracle.c,
faithful=False), with the pressure solved to ||b - Ax||_2/||b||_2 <= 1e-9.

Stored per step: mean y of each phase, kinetic energy sum |v|^2 per phase, maximum
particle speed, fluid cells and pin; after steps 1, 10 and 100: every 997th particle
(caller order) as FP32 x, y, z, u, v, w.  The GPU test replays the same scene with
its FP32 step and compares.

Usage: python tests/golden/make_c0_rt64_trajectory.py [--threads N] (about 15 min on 8 cores)
"""
import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pyoracle as O  # noqa: E402

N, SEED, DT, STEPS, TOL = 64, 20250601, 1e-3, 100, 1e-9
SNAP_STEPS = (1, 10, 100)
STRIDE = 997


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    ap.add_argument("--out", default=os.path.join(HERE, "c0_rt64_trajectory.npz"))
    a = ap.parse_args()
    O.set_threads(a.threads)
    dx = 1.0 / N
    x, y, z, t = O.scene(O.SCENE_RT, N, N, N, dx, SEED)
    ph = O.Phase(1000.0, 1.0, 1.0, 1.0)
    ph.rho_tilde_0 = O.calibrate_rho_tilde0(ph, 1.0)
    ps = O.Particles(x, y, z, type=t)
    g = O.Grid(N, N, N, dx, faithful=False)
    cfg = O.StepCfg(dt=DT, gravity=-9.81, Vp=1.0, alpha_liquid=0.95, alpha_gas=0.95, adaptive_timestep=0,
                    advection_order=2, cg_iters=100000, stop_rule=O.STOP_REL2, adaptive_cfl=0.5,
                    adaptive_min_dt=1e-5, cg_tol=TOL, c_div_volume_correction=0, c_div_strength=1.0,
                    liquid_volume_target=0.0, phase=ph)
    heavy, light = t == 0, t == 1
    rows, snaps = [], {}
    t0 = time.time()
    for s in range(1, STEPS + 1):
        st = g.step(ps, cfg)
        sp2 = ps.u ** 2 + ps.v ** 2 + ps.w ** 2
        rows.append([ps.y[heavy].mean(), ps.y[light].mean(), sp2[heavy].sum(), sp2[light].sum(),
                     np.sqrt(sp2.max()), st.solve.fluid_cells, st.solve.pinned, st.solve.iterations,
                     st.solve.true_rel_residual])
        if s in SNAP_STEPS:
            snaps[s] = np.stack([a[::STRIDE] for a in (ps.x, ps.y, ps.z, ps.u, ps.v, ps.w)]).astype(np.float32)
        print(f"step {s}: {time.time() - t0:.0f} s, {st.solve.iterations} it, resid {st.solve.true_rel_residual:.2e}",
              flush=True)
    np.savez_compressed(a.out, summary=np.array(rows), snap_steps=np.array(SNAP_STEPS),
                        snaps=np.stack([snaps[s] for s in SNAP_STEPS]), stride=STRIDE, seed=SEED, n=N, dt=DT,
                        tol=TOL, np=len(x))
    print("wrote", a.out)


if __name__ == "__main__":
    main()
