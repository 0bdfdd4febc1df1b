#!/usr/bin/env python3
"""Writes ref_rt_8x12x8.npz from the reference's own SparseSim3DTP (oracle/_ref,
built in this container from the reference sources by oracle/Makefile): the
Rayleigh-Taylor test scene of tests/test_sparse_sim3d_tp.cpp:110-148 (8x12x8 cells,
dx 1, dt 0.03, the class defaults rho_l 1 / rho_g 0.01, initRayleighTaylor's
lattice), its initial particles and phase, and the particle state after steps 1 and
12 plus the mean y of each phase after every step.  The GPU test replays the scene
from the stored initial state; the reference binary itself never leaves this
container.

Usage: python tests/golden/make_ref_rt_scene.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pyoracle as O  # noqa: E402


def main():
    if not O.ref_available():
        sys.exit("oracle/_ref is not built (python -c 'import pyoracle; pyoracle.build()')")
    rs = O.RefSim(8, 12, 8, 1.0)
    ps, ph = rs.init_scene(0)
    cfg = O.StepCfg(dt=0.03, gravity=-9.81, Vp=1.0, alpha_liquid=0.95, alpha_gas=0.95, adaptive_timestep=0,
                    advection_order=2, cg_iters=600, stop_rule=O.STOP_REF_MAXABS, adaptive_cfl=0.5,
                    adaptive_min_dt=1e-5, cg_tol=1e-7, c_div_volume_correction=0, c_div_strength=1.0,
                    liquid_volume_target=0.0, phase=ph)
    rs.config(cfg)
    init = np.stack([ps.x, ps.y, ps.z, ps.u, ps.v, ps.w])
    means, after = [], {}
    for s in range(1, 13):
        rs.step()
        q = rs.get_particles(ps)
        means.append([q.y[ps.type == 0].mean(), q.y[ps.type == 1].mean()])
        if s in (1, 12):
            after[s] = np.stack([q.x, q.y, q.z, q.u, q.v, q.w])
    out = os.path.join(HERE, "ref_rt_8x12x8.npz")
    np.savez_compressed(out, init=init, type=ps.type, phase=ph.as_array(), dt=0.03, steps=12,
                        after1=after[1], after12=after[12], mean_y=np.array(means))
    print("wrote", out, len(ps), "particles")


if __name__ == "__main__":
    main()
