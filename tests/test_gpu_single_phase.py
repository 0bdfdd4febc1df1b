"""Single-phase sparse path (SparseSim3D, src/driver/sparse_sim3d.cpp; spP2G3D and
spProjectStep3D, src/driver/sparse_ops3d.cpp) on the HIP path against the oracle.

Bars: markers and active-block sets bit-exact; grid mass and momentum within 1e-5
norm-wise relative error of the FP64 oracle; pressure to the FP64 residual asked
for; particle state after whole steps within the stated tolerances.
"""
import numpy as np
import pytest

import pfflip_amd as G
import pyoracle as O
from _helpers import rel_err, scene

pytestmark = pytest.mark.gpu

FIELD_TOL = 1e-5


def _setup(n, seed=5, ny=None, vel=True):
    ny = ny or n
    dx, x, y, z, t = scene(O.SCENE_DAMBREAK, n, seed=seed, ny=ny)
    rng = np.random.default_rng(seed)
    u, v, w = (rng.normal(scale=0.3, size=len(x)).astype(np.float32) if vel else np.zeros(len(x), np.float32)
               for _ in range(3))
    ctx = G.Context(n, ny, n, dx)
    ctx.upload(x, y, z, u, v, w)
    return dx, ctx, O.Particles(x, y, z, u, v, w)


def test_single_phase_p2g_active_sets_and_fields():
    n = 24
    dx, ctx, ps = _setup(n)
    g = O.Grid(n, n, n, dx, faithful=False)
    g.mark_cells(ps)
    g.p2g_single(ps)
    ctx.mark_cells()
    ctx.p2g_single()
    assert np.array_equal(ctx.field(G.FIELD_MARKER), g.dense(O.MK))
    for f in range(6):
        assert np.array_equal(np.sort(ctx.active_blocks(f)), np.sort(g.active_blocks(f))), f
        assert rel_err(ctx.field(f), g.dense(f)) < FIELD_TOL, f
    for fv, fm in ((0, 3), (1, 4), (2, 5)):
        mom = ctx.field(fv).astype(np.float64) * ctx.field(fm)
        assert rel_err(mom, g.dense(fv) * g.dense(fm)) < FIELD_TOL
    # unit masses: the raw mass sums to the particle count wherever no stencil node
    # leaves the grid (the scene keeps particles in interior cells)
    assert abs(ctx.field(G.FIELD_MV).astype(np.float64).sum() - len(ps)) < 1e-3 * len(ps)


@pytest.mark.parametrize("precond", [0, 2])
def test_single_phase_projection_matches_fp64_oracle(precond):
    # precond 2: the multigrid preconditioner (same FP64 stop, same bars)
    n, dt = 24, 2e-3
    dx, ctx, ps = _setup(n)
    g = O.Grid(n, n, n, dx, faithful=False)
    ctx.mark_cells(); ctx.p2g_single(); ctx.save_velocity(); ctx.apply_gravity(dt, -9.81); ctx.apply_walls()
    g.mark_cells(ps); g.p2g_single(ps); g.apply_gravity(dt, -9.81); g.apply_walls()
    st = ctx.project_single(dt, G.default_solve_cfg(G.STOP_REL2, 1e-9, 20000, precond))
    ost = g.project_single(dt, 20000, 1e-12, O.STOP_REL2)
    assert st.converged == 1 and st.true_rel_residual <= 1e-9 and st.fallback_used == 0
    assert st.pinned == 0 and ost.pinned == 0
    assert st.fluid_cells == ost.fluid_cells
    assert np.array_equal(np.sort(ctx.active_blocks(G.FIELD_P)), np.sort(g.active_blocks(O.P)))
    assert rel_err(ctx.field(G.FIELD_P), g.dense(O.P)) < 1e-6
    for f in range(3):
        assert rel_err(ctx.field(f), g.dense(f)) < 1e-5
    # free-surface projection: divergence on FLUID cells removed to the solve's level
    mk = g.dense(O.MK)
    U, V, W = (ctx.field(f).astype(np.float64) for f in range(3))
    d = ((U[:, :, 1:] - U[:, :, :-1]) + (V[:, 1:, :] - V[:, :-1, :]) + (W[1:, :, :] - W[:-1, :, :]))[mk == 1]
    Us, Vs, Ws = (ctx.field(f).astype(np.float64) for f in (8, 9, 10))
    d0 = ((Us[:, :, 1:] - Us[:, :, :-1]) + (Vs[:, 1:, :] - Vs[:, :-1, :]) + (Ws[1:, :, :] - Ws[:-1, :, :]))[mk == 1]
    assert np.abs(d).max() < 1e-4 * (np.abs(d0).max() + 1.0)


def test_single_phase_steps_match_fp64_oracle():
    n, dt = 24, 2e-3
    dx, ctx, ps = _setup(n, vel=False)
    sim = G.SparseSim3D(n, n, n, dx)
    sim.ctx = ctx
    sim.dt, sim.alpha = dt, 0.95
    sim.stop_rule, sim.cg_tol, sim.cg_iters = G.STOP_REL2, 1e-9, 20000
    g = O.Grid(n, n, n, dx, faithful=False)
    pa = ps.copy()
    for _ in range(3):
        st = sim.step()
        ost = g.step_single(pa, dt, 0.95, -9.81, 20000, 1e-12, O.STOP_REL2)
        assert st.solve.converged == 1 and st.solve.pinned == 0
        assert st.solve.fluid_cells == ost.solve.fluid_cells
    x, y, z, u, v, w = ctx.download()
    assert max(np.abs(x - pa.x).max(), np.abs(y - pa.y).max(), np.abs(z - pa.z).max()) < 1e-6
    assert rel_err(np.concatenate([u, v, w]), np.concatenate([pa.u, pa.v, pa.w])) < 1e-4


def test_reference_dambreak_test_on_the_gpu():
    # tests/test_sparse_sim3d.cpp:6-36 through the Python mirror: 16^3, initDamBreak's
    # lattice, 20 steps with the reference defaults (dt 0.05, cg 300 / 1e-6)
    n = 16
    xs, ys, zs = [], [], []
    for k in range(1, n - 1):
        for j in range(1, n * 7 // 10):
            for i in range(1, n * 4 // 10):
                for s in range(8):
                    xs.append(i + 0.25 + 0.5 * (s & 1))
                    ys.append(j + 0.25 + 0.5 * ((s >> 1) & 1))
                    zs.append(k + 0.25 + 0.5 * ((s >> 2) & 1))
    sim = G.SparseSim3D(n, n, n, 1.0)
    sim.set_particles(xs, ys, zs)
    my0 = float(np.mean(ys))
    max_active = 0
    for _ in range(20):
        sim.step()
        max_active = max(max_active, len(sim.ctx.active_blocks(G.FIELD_P)))
    x, y, z, *_ = sim.ctx.download()
    assert sim.ctx.count() == len(xs)
    assert np.isfinite(x).all() and np.isfinite(y).all() and np.isfinite(z).all()
    assert float(y.mean()) < my0
    assert 0 < max_active < 4 * 4 * 4
