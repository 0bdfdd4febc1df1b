"""Shared set-up for the parity tests: seeded scenes, oracle runs, error norms."""
import numpy as np

import pyoracle as O


def rel_err(a, b):
    """Norm-wise relative error ||a - b||_2 / ||b||_2 (pointwise relative error is
    meaningless where momentum cancels)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / nb) if nb > 0 else float(np.linalg.norm(a))


def phase(rho_l=1000.0, rho_g=1.0, alpha=1.0):
    ph = O.Phase(rho_l, rho_g, alpha, 1.0)
    ph.rho_tilde_0 = O.calibrate_rho_tilde0(ph, 1.0)
    return ph


def scene(kind, n, seed=11, ny=None, nz=None):
    ny = ny or n
    nz = nz or n
    dx = 1.0 / n
    x, y, z, t = O.scene(kind, n, ny, nz, dx, seed)
    return dx, x, y, z, t


def step_cfg(ph, dt=1e-3, cg_iters=600, cg_tol=1e-7, stop_rule=O.STOP_REF_MAXABS, order=2):
    return O.StepCfg(dt=dt, gravity=-9.81, Vp=1.0, alpha_liquid=0.95, alpha_gas=0.95,
                     adaptive_timestep=0, advection_order=order, cg_iters=cg_iters,
                     stop_rule=stop_rule, adaptive_cfl=0.5, adaptive_min_dt=1e-5, cg_tol=cg_tol,
                     c_div_volume_correction=0, c_div_strength=1.0, liquid_volume_target=0.0,
                     phase=ph)


# option sets of MRPressureSolveConfig3D's coarse machinery (multires_pressure3d.h:148-165)
# exercised by the multires oracle and GPU tests; "flexible" selects use_flexible_cg_beta
COARSE_CASES = [
    dict(use_coarse_correction=1),
    dict(use_coarse_correction=1, coarse_correction_sweeps=3, coarse_correction_iterations=6),
    dict(use_coarse_correction=1, coarse_correction_sweeps=2, coarse_correction_min_scale=0.25,
         coarse_correction_relative_tolerance=0.0, coarse_correction_iterations=200),
    dict(use_coarse_preconditioner=1),
    dict(use_coarse_preconditioner=1, coarse_preconditioner_min_rz_gain=0.5, coarse_preconditioner_auto_disable=1,
         coarse_preconditioner_auto_disable_after=2),
    dict(use_coarse_preconditioner=1, coarse_preconditioner_max_work_ratio=1.5, coarse_preconditioner_iterations=4),
    dict(use_coarse_preconditioner=1, coarse_preconditioner_scale=0.5, use_coarse_correction=1, flexible=1),
]


def bubble_lattice(n, seed=3, vel=True):
    """the reference's bubble-tank lattice (multires_sim3d_tp.cpp:483-496), FP32 values"""
    water = n // 2
    cx, cy, cz = n * 0.5, water * 0.375, n * 0.5
    r = max(1.25, n * 0.1875)
    pts = []
    for k in range(1, n - 1):
        for j in range(1, water):
            for i in range(1, n - 1):
                gas = (i + 0.5 - cx) ** 2 + (j + 0.5 - cy) ** 2 + (k + 0.5 - cz) ** 2 < r * r
                for s in range(8):
                    pts.append((i + 0.25 + 0.5 * (s & 1), j + 0.25 + 0.5 * ((s >> 1) & 1),
                                k + 0.25 + 0.5 * ((s >> 2) & 1), 1 if gas else 0))
    p = np.array(pts)
    rng = np.random.default_rng(seed)
    x, y, z = (p[:, q].astype(np.float32) for q in range(3))
    u, v, w = ((rng.normal(scale=0.3, size=len(p)) if vel else np.zeros(len(p))).astype(np.float32) for _ in range(3))
    t = p[:, 3].astype(np.uint8)
    return x, y, z, u, v, w, t


def refine_box(nx, ny, nz, leaves, box):
    """refineFineCellBox (multires_layout3d.h:62-97) on an (n, 4) leaf array of (level,
    bx, by, bz): leaves above level 0 that overlap the box split into their children
    inside the domain; sorted and unique"""
    x0, y0, z0, x1, y1, z1 = box
    out = set()
    for L, bx, by, bz in np.asarray(leaves).tolist():
        s = 4 << L
        ov = (bx * s < x1 and min(nx, bx * s + s) > x0 and by * s < y1 and min(ny, by * s + s) > y0
              and bz * s < z1 and min(nz, bz * s + s) > z0)
        if L == 0 or not ov:
            out.add((L, bx, by, bz))
            continue
        h = 4 << (L - 1)
        for c in range(8):
            ch = (L - 1, 2 * bx + (c & 1), 2 * by + ((c >> 1) & 1), 2 * bz + (c >> 2))
            if ch[1] * h < nx and ch[2] * h < ny and ch[3] * h < nz:
                out.add(ch)
    return np.array(sorted(out), np.int32).reshape(-1, 4)


def three_level_bubble_layout(n):
    """levels 2, 1 and 0 around the bubble tank of bubble_lattice(n) (n a multiple of 16,
    at least 64): level 2 in the bulk, level 1 in a slab around the free surface, level
    0 in a box around the bubble (refined twice from level 2) and in a band around the
    free surface"""
    water = n // 2
    lv = O.mr_layout_fine_box(n, n, n, [0, water - water // 4 + 4, 0, n, water + water // 4 + 4, n], level=2)
    cy, r = water * 0.375, max(1.25, n * 0.1875)
    bub = [int(n / 2 - r - 2), 0, int(n / 2 - r - 2), int(n / 2 + r + 2), int(cy + r + 2), int(n / 2 + r + 2)]
    for box in (bub, bub, [0, water - 4, 0, n, water + 4, n]):
        lv = refine_box(n, n, n, lv, box)
    return lv
