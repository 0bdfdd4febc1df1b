"""Three resolution levels on the HIP multires path.

The reference's layout stops at level 1 (multires_layout3d.h:25-26); the GPU
representation takes leaf levels 0-6.  On a level 2-1-0 layout around the bubble tank:
layout, enumerated faces and markers bit-exact against the oracle; face mass and
momentum within 1e-5 of the FP64 oracle; the fast projection (FP32 PCG with FP64
refinement on V A) against the exact FP64 solution of the oracle's rows
(multires_pressure3d.cpp:783-812) followed by the oracle's face update.  The
reference's own iteration does not converge on such layouts, so the exact solution is
the comparison.  The reference mode's level-1 coarse space refuses level-2 leaves, as
the reference's level1ParentTuple does.
"""
import numpy as np
import pytest

import pfflip_amd as G
import pyoracle as O
from _helpers import bubble_lattice, refine_box, rel_err, three_level_bubble_layout
from test_multires3_oracle import _system, exact_solution

pytestmark = pytest.mark.gpu

FIELD_TOL = 1e-5


def _phase():
    ph = O.Phase(1.0, 0.01, 1.0, 1.0)
    ph.rho_tilde_0 = O.calibrate_rho_tilde0(ph, 1.0)
    return ph, G.Phase(ph.rho_l, ph.rho_g, ph.alpha_phi, ph.rho_tilde_0)


def _setup(n, dt):
    ph, gph = _phase()
    leaves = three_level_bubble_layout(n)
    parts = bubble_lattice(n)
    ctx = G.MRContext(n, n, n, 1.0)
    ctx.set_layout(leaves)
    ctx.upload(*parts)
    g = O.MRGrid(n, n, n, 1.0, faithful=False)
    g.set_leaves(leaves)
    x, y, z, u, v, w, t = parts
    ps = O.Particles(x, y, z, u, v, w, type=t)
    return ctx, g, ps, ph, gph


def test_layout_refine_with_level_two_leaves():
    n = 64
    water = n // 2
    base = G.mr_layout_fine_box(n, n, n, [0, water - water // 4 + 4, 0, n, water + water // 4 + 4, n], level=2)
    assert np.array_equal(base, O.mr_layout_fine_box(n, n, n, [0, water - water // 4 + 4, 0, n,
                                                                 water + water // 4 + 4, n], level=2))
    box = [14, 0, 14, 50, 26, 50]
    once = G.mr_layout_refine(n, n, n, base, box)
    assert np.array_equal(once, refine_box(n, n, n, base, box))
    assert np.array_equal(G.mr_layout_refine(n, n, n, once, box), refine_box(n, n, n, once, box))


@pytest.mark.parametrize("mode", [G.MR_SOLVE_FAST, G.MR_SOLVE_MULTIGRID])
def test_three_levels_match_oracle_and_exact_projection(mode):
    n, dt = 64, 0.02
    ctx, g, ps, ph, gph = _setup(n, dt)
    assert np.array_equal(ctx.layout(), g.leaves())
    for a in range(3):
        assert np.array_equal(ctx.face_valid(a), g.face_valid(a)), a
    ctx.mr_mark_cells()
    g.mark_cells(ps)
    assert np.array_equal(ctx.field(G.FIELD_MARKER), g.dense(O.MK))
    ctx.mr_p2g(gph, 1.0)
    g.p2g(ps, ph, 1.0)
    for f in range(6):
        assert rel_err(ctx.field(f), g.dense(f)) < FIELD_TOL, f
    ctx.save_velocity(); ctx.apply_gravity(dt, -9.81); ctx.apply_walls()
    g.apply_gravity(dt, -9.81); g.apply_walls()
    for f in range(3):
        assert rel_err(ctx.field(f), g.dense(f)) < FIELD_TOL, f
    cells, A, rhs, pin, vol = _system(g, ph, dt)
    x = exact_solution(A, rhs, vol)
    g.apply_pressure(ph, dt, x)
    st = ctx.mr_project(gph, dt, G.mr_solve_cfg(20000, 1e-11, 0.0, mode=mode))
    assert st.converged == 1 and st.true_max_residual < 1e-11
    assert st.fluid_cells == len(cells) and st.pinned_cell == pin
    assert rel_err(ctx.field(G.FIELD_P), g.dense(O.P)) < 1e-6
    for f in range(3):
        assert rel_err(ctx.field(f), g.dense(f)) < 1e-5, f


def test_reference_mode_coarse_space_refuses_level_two():
    n, dt = 64, 0.02
    ctx, g, ps, ph, gph = _setup(n, dt)
    ctx.mr_mark_cells(); ctx.mr_p2g(gph, 1.0); ctx.apply_gravity(dt, -9.81); ctx.apply_walls()
    cfg = G.mr_solve_cfg(200, 1e-7, 1e-6, mode=G.MR_SOLVE_REFERENCE, use_coarse_correction=1)
    with pytest.raises(G.PfError, match="level 0 and 1"):
        ctx.mr_project(gph, dt, cfg)
    # without the coarse space the reference iteration runs (and need not converge)
    st = ctx.mr_project(gph, dt, G.mr_solve_cfg(50, 1e-7, 1e-6, mode=G.MR_SOLVE_REFERENCE))
    assert st.iterations > 0


def _box_distance(bub, water, lo, hi):
    """distance in cells from the cell centres of the fine box [lo, hi) to the free
    surface y = water and the bubble spheres (domain units = cells here)"""
    a, b = np.asarray(lo) + 0.5, np.asarray(hi) - 0.5
    d = max(0.0, a[1] - water, water - b[1])
    for cx, cy, cz, r in np.asarray(bub, np.float64):
        c = np.array([cx, cy, cz])
        near = np.sqrt((np.maximum(0, np.maximum(a - c, c - b)) ** 2).sum())
        far = np.sqrt((np.maximum(np.abs(a - c), np.abs(b - c)) ** 2).sum())
        d = min(d, max(0.0, near - r, r - far))
    return d


def test_multi_bubble_scene():
    # the multires multi-bubble tank at 96^3 (dx = 1): the level rule, 2:1 balance,
    # tiling, the seeded particle count and types; then P2G against the oracle and the
    # fast projection against the exact solution of the oracle's rows
    n, dt, band = 96, 0.02, 6
    ph, gph = _phase()
    ctx = G.MRContext(n, n, n, 1.0)
    bub = ctx.scene_bubbles(7, 4, band)
    lv = ctx.layout()
    assert set(np.unique(lv[:, 0]).tolist()) == {0, 1, 2}
    assert ((lv[:, 0] == 2) & (lv[:, 2] * 16 < n // 2)).any()   # level-2 leaves in the water
    level = np.full((n, n, n), -1, np.int64)
    for L, bx, by, bz in lv.tolist():
        s = 4 << L
        assert (level[bz * s:bz * s + s, by * s:by * s + s, bx * s:bx * s + s] == -1).all()
        level[bz * s:bz * s + s, by * s:by * s + s, bx * s:bx * s + s] = L
        lim = {0: None, 1: band, 2: band + 16}[L]
        if lim is not None:
            assert _box_distance(bub, n // 2, (bx * s, by * s, bz * s), (bx * s + s, by * s + s, bz * s + s)) > lim
        if L < 2:
            p = 8 << L    # the parent that split: within its threshold
            plim = {0: band, 1: band + 16}[L]
            lo = ((bx * s) // p * p, (by * s) // p * p, (bz * s) // p * p)
            assert _box_distance(bub, n // 2, lo, tuple(q + p for q in lo)) <= plim
    assert (level >= 0).all()
    for ax in range(3):
        d = np.abs(np.diff(level, axis=ax))
        assert d.max() <= 1, ax
    x, y, z, u, v, w = ctx.download()
    t, _ = ctx.download_ext()
    seeded = 0
    for L, bx, by, bz in lv.tolist():
        h = 1 << L
        for c in range(64):
            x0, y0, z0 = bx * 4 * h + (c & 3) * h, by * 4 * h + ((c >> 2) & 3) * h, bz * 4 * h + (c >> 4) * h
            if max(x0, 1) < min(x0 + h, n - 1) and max(y0, 1) < min(y0 + h, n - 1, n // 2) \
                    and max(z0, 1) < min(z0 + h, n - 1):
                seeded += 1
    assert len(x) == 8 * seeded
    assert y.max() < n // 2 and min(x.min(), y.min(), z.min()) >= 1 and max(x.max(), z.max()) < n - 1
    gas = np.zeros(len(x), bool)
    for cx, cy, cz, r in np.asarray(bub, np.float64):
        gas |= (x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2 < r * r
    assert np.array_equal(t == 1, gas) and gas.any()
    g = O.MRGrid(n, n, n, 1.0, faithful=False)
    g.set_leaves(lv)
    ps = O.Particles(x, y, z, u, v, w, type=t)
    ctx.mr_mark_cells(); g.mark_cells(ps)
    assert np.array_equal(ctx.field(G.FIELD_MARKER), g.dense(O.MK))
    ctx.mr_p2g(gph, 1.0); g.p2g(ps, ph, 1.0)
    for f in range(6):
        assert rel_err(ctx.field(f), g.dense(f)) < FIELD_TOL, f
    ctx.save_velocity(); ctx.apply_gravity(dt, -9.81); ctx.apply_walls()
    g.apply_gravity(dt, -9.81); g.apply_walls()
    cells, A, rhs, pin, vol = _system(g, ph, dt)
    g.apply_pressure(ph, dt, exact_solution(A, rhs, vol))
    st = ctx.mr_project(gph, dt, G.mr_solve_cfg(20000, 1e-11, 0.0))
    assert st.converged == 1 and st.fluid_cells == len(cells)
    assert rel_err(ctx.field(G.FIELD_P), g.dense(O.P)) < 1e-6
    for f in range(3):
        assert rel_err(ctx.field(f), g.dense(f)) < 1e-5, f


def test_multigrid_cycle_cuts_iterations():
    # the V-cycle on V A (block grid, then the uniform levels) against Jacobi on the
    # same three-level system: the same solution to the solve tolerance, several times
    # fewer iterations
    n, dt = 64, 0.02
    runs = {}
    for mode in (G.MR_SOLVE_FAST, G.MR_SOLVE_MULTIGRID):
        ctx, g, ps, ph, gph = _setup(n, dt)
        ctx.mr_mark_cells(); ctx.mr_p2g(gph, 1.0); ctx.apply_gravity(dt, -9.81); ctx.apply_walls()
        st = ctx.mr_project(gph, dt, G.mr_solve_cfg(20000, 0.0, 1e-8, mode=mode))
        assert st.converged == 1
        runs[mode] = (st.iterations, ctx.field(G.FIELD_P))
    assert rel_err(runs[G.MR_SOLVE_MULTIGRID][1], runs[G.MR_SOLVE_FAST][1]) < 1e-6
    assert 4 * runs[G.MR_SOLVE_MULTIGRID][0] <= runs[G.MR_SOLVE_FAST][0], runs


def test_multigrid_refuses_level_three():
    n = 64
    ctx = G.MRContext(n, n, n, 1.0)
    ctx.set_layout(O.mr_layout_fine_box(n, n, n, [0, 0, 0, n, 8, n], level=3))
    ctx.upload(*bubble_lattice(n))
    ph, gph = _phase()
    ctx.mr_mark_cells(); ctx.mr_p2g(gph, 1.0)
    with pytest.raises(G.PfError, match="level 0 to 2"):
        ctx.mr_project(gph, 0.02, G.mr_solve_cfg(100, 0.0, 1e-6, mode=G.MR_SOLVE_MULTIGRID))
