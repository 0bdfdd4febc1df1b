"""Multigrid-preconditioned pressure solve (PF_PRECOND_MULTIGRID) on the GPU.

The preconditioner changes how fast the solve gets there, not where: the solve still
stops on the FP64 residual ||b - Ax||_2/||b||_2, so the bars are those of the Jacobi
solve against the FP64 oracle (pressure within 1e-6 norm-wise relative, faces within
1e-5, same FLUID count and pinned cell), plus far fewer iterations than Jacobi PCG on
the same system.
"""
import numpy as np
import pytest

import pfflip_amd as G
import pyoracle as O
from _helpers import phase, rel_err, scene

pytestmark = pytest.mark.gpu


def _gphase(oph):
    return G.Phase(oph.rho_l, oph.rho_g, oph.alpha_phi, oph.rho_tilde_0)


def _prepare(kind, n, seed=11, ny=None, nz=None, dt=1e-3, rho_g=None):
    dx, x, y, z, t = scene(kind, n, seed=seed, ny=ny, nz=nz)
    rng = np.random.default_rng(seed)
    u, v, w = (rng.normal(scale=0.3, size=len(x)).astype(np.float32) for _ in range(3))
    oph = phase() if rho_g is None else phase(1000.0, rho_g)
    ctx = G.Context(n, ny or n, nz or n, dx)
    ctx.upload(x, y, z, u, v, w, t)
    ps = O.Particles(x, y, z, u, v, w, t)
    return dx, ctx, ps, oph


def _to_projection(ctx, oph, dt):
    ctx.mark_cells()
    ctx.p2g(_gphase(oph), 1.0)
    ctx.save_velocity()
    ctx.apply_gravity(dt, -9.81)
    ctx.apply_walls()


def _faces(ctx):
    return [ctx.field(f) for f in (G.FIELD_U, G.FIELD_V, G.FIELD_W)]


@pytest.mark.parametrize("kind", [O.SCENE_RT, O.SCENE_BUBBLE, O.SCENE_DAMBREAK])
def test_multigrid_projection_matches_fp64_oracle(kind):
    n, dt = 24, 1e-3
    print(f"[mg] kind={kind} n={n}", flush=True)
    dx, ctx, ps, oph = _prepare(kind, n)
    g = O.Grid(n, n, n, dx, faithful=False)
    g.mark_cells(ps)
    g.p2g(ps, oph, 1.0)
    g.apply_gravity(dt, -9.81)
    g.apply_walls()
    _to_projection(ctx, oph, dt)
    st = ctx.project(_gphase(oph), dt, G.default_solve_cfg(G.STOP_REL2, 1e-9, 20000, G.PRECOND_MULTIGRID))
    ost = g.project(oph, dt, 20000, 1e-12, 0.0, O.STOP_REL2)
    assert st.converged == 1 and st.true_rel_residual <= 1e-9 and st.fallback_used == 0
    assert st.fluid_cells == ost.fluid_cells
    assert st.pinned == ost.pinned and (st.pinned_cell == ost.pinned_cell or not st.pinned)
    assert rel_err(ctx.field(G.FIELD_P), g.dense(O.P)) < 1e-6
    for f in range(3):
        assert rel_err(ctx.field(f), g.dense(f)) < 1e-5


@pytest.mark.parametrize("shape", [(40, None, None), (20, 13, 29), (48, 48, 12)])
def test_multigrid_needs_far_fewer_iterations(shape):
    n, ny, nz = shape
    dt = 1e-3
    print(f"[mg-iters] shape={shape}", flush=True)
    dx, ctx, ps, oph = _prepare(O.SCENE_RT, n, seed=5, ny=ny, nz=nz)
    _to_projection(ctx, oph, dt)
    faces = _faces(ctx)
    st_j = ctx.project(_gphase(oph), dt, G.default_solve_cfg(G.STOP_REL2, 1e-8, 50000))
    p_j = ctx.field(G.FIELD_P).astype(np.float64)
    after_j = _faces(ctx)
    for f, a in zip((G.FIELD_U, G.FIELD_V, G.FIELD_W), faces):
        ctx.set_field(f, a)
    st_m = ctx.project(_gphase(oph), dt, G.default_solve_cfg(G.STOP_REL2, 1e-8, 50000, G.PRECOND_MULTIGRID))
    print(f"iterations: jacobi {st_j.iterations}, multigrid {st_m.iterations}")
    assert st_j.converged == 1 and st_m.converged == 1
    assert st_m.true_rel_residual <= 1e-8 and st_m.fallback_used == 0
    assert st_m.iterations * 3 <= st_j.iterations
    # both are the same FP64-converged solution
    assert rel_err(ctx.field(G.FIELD_P), p_j) < 1e-5
    for a, b in zip(_faces(ctx), after_j):
        assert rel_err(a, b) < 1e-5


def test_multigrid_steps_track_jacobi_steps():
    # the coarse operators are rebuilt per step as the FLUID set moves
    n = 24
    dx, x, y, z, t = scene(O.SCENE_DAMBREAK, n, seed=3)
    oph = phase()
    runs = []
    for pre in (G.PRECOND_JACOBI, G.PRECOND_MULTIGRID):
        sim = G.SparseSim3DTP(n, n, n, dx)
        sim.phase = _gphase(oph)
        sim.dt = 2e-3
        sim.stop_rule, sim.cg_tol, sim.cg_iters = G.STOP_REL2, 1e-9, 20000
        sim.precond = pre
        sim.set_particles(x, y, z, type=t)
        its = []
        for _ in range(3):
            sim.step()
            its.append(sim.stats_last.solve.iterations)
            assert sim.stats_last.solve.converged == 1
        runs.append((np.stack(sim.ctx.download()).astype(np.float64), its))
    (a, it_j), (b, it_m) = runs
    print(f"iterations per step: jacobi {it_j}, multigrid {it_m}")
    assert sum(it_m) < sum(it_j)
    pos_a, pos_b = a[:3], b[:3]
    assert np.abs(pos_a - pos_b).max() < 1e-5
    assert rel_err(b[3:], a[3:]) < 1e-4


@pytest.mark.parametrize("kind", [O.SCENE_RT, O.SCENE_BUBBLE])
def test_multigrid_steps_stay_converged(kind):
    # ten adaptive-dt steps of an evolving interface: every solve reaches the FP64
    # tolerance with the cycle (no Jacobi fallback) in a steady number of iterations
    n = 48
    print(f"[mg-steps] kind={kind}", flush=True)
    ctx = G.Context(n, n, n, 1.0 / n)
    ctx.scene(kind, 7)
    oph = phase()
    solve = G.default_solve_cfg(G.STOP_REL2, 1e-6, 50000, G.PRECOND_MULTIGRID)
    solve.inner_rel_tol = 0.01
    cfg = G.StepCfg(dt=2e-3, gravity=-9.81, Vp=1.0, alpha_liquid=0.95, alpha_gas=0.95, adaptive_cfl=0.5,
                    adaptive_min_dt=1e-5, c_div=0.0, adaptive_timestep=1, advection_order=2, timing=0,
                    phase=_gphase(oph), solve=solve)
    its = []
    for _ in range(10):
        st = ctx.step(cfg)
        assert st.solve.converged == 1 and st.solve.fallback_used == 0
        assert st.solve.true_rel_residual <= 1e-6
        its.append(st.solve.iterations)
    print(f"iterations {its}")
    assert max(its) <= 40


def test_multigrid_with_tiny_and_empty_systems():
    # a grid of at most 4 cells a side has no coarse level; a grid with no FLUID cell
    # has nothing to solve
    for n in (4, 7):
        dx, ctx, ps, oph = _prepare(O.SCENE_BUBBLE, n, seed=2)
        _to_projection(ctx, oph, 1e-3)
        st = ctx.project(_gphase(oph), 1e-3, G.default_solve_cfg(G.STOP_REL2, 1e-9, 5000, G.PRECOND_MULTIGRID))
        assert st.converged == 1 and st.true_rel_residual <= 1e-9
    ctx = G.Context(8, 8, 8, 1.0)
    st = ctx.project(_gphase(phase()), 1e-3, G.default_solve_cfg(G.STOP_REL2, 1e-9, 5000, G.PRECOND_MULTIGRID))
    assert st.converged == 1 and st.iterations == 0


@pytest.mark.parametrize("cycle", [dict(sweeps=1), dict(sweeps=3), dict(sweeps=2, coarsest_sweeps=1),
                                   dict(omega=0.6, correction=1.0), dict(sweeps=4, omega=0.7, correction=1.5),
                                   dict(chebyshev_ratio=0.0), dict(chebyshev_ratio=20.0, sweeps=3)])
def test_every_cycle_shape_reaches_the_jacobi_solution(cycle):
    # one sweep (first sweep + fused prolongation), the fused pair of first sweeps,
    # more sweeps, a one-sweep coarsest level, no over-correction
    n, dt = 32, 1e-3
    print(f"[mg-cycle] {cycle}", flush=True)
    dx, ctx, ps, oph = _prepare(O.SCENE_BUBBLE, n, seed=9)
    _to_projection(ctx, oph, dt)
    faces = _faces(ctx)
    st_j = ctx.project(_gphase(oph), dt, G.default_solve_cfg(G.STOP_REL2, 1e-9, 50000))
    p_j = ctx.field(G.FIELD_P).astype(np.float64)
    for f, a in zip((G.FIELD_U, G.FIELD_V, G.FIELD_W), faces):
        ctx.set_field(f, a)
    ctx.set_multigrid(**cycle)
    got = ctx.multigrid()
    for k, v in cycle.items():
        assert getattr(got, k) == pytest.approx(v)
    st_m = ctx.project(_gphase(oph), dt, G.default_solve_cfg(G.STOP_REL2, 1e-9, 50000, G.PRECOND_MULTIGRID))
    print(f"iterations: jacobi {st_j.iterations}, multigrid {st_m.iterations}")
    assert st_m.converged == 1 and st_m.true_rel_residual <= 1e-9 and st_m.fallback_used == 0
    assert st_m.iterations < st_j.iterations
    assert rel_err(ctx.field(G.FIELD_P), p_j) < 1e-6


@pytest.mark.parametrize("rho_g", [0.1, 1000.0])
def test_multigrid_at_extreme_density_ratios(rho_g):
    # the cycle reads an FP16 copy of the level-0 operator scaled by its maximum: a
    # liquid/gas ratio of 10^4 (and none) must still reach the FP64 tolerance and the
    # Jacobi solution
    n, dt = 32, 1e-3
    print(f"[mg-ratio] rho_g={rho_g}", flush=True)
    dx, ctx, ps, oph = _prepare(O.SCENE_RT, n, seed=13, rho_g=rho_g)
    _to_projection(ctx, oph, dt)
    faces = _faces(ctx)
    st_j = ctx.project(_gphase(oph), dt, G.default_solve_cfg(G.STOP_REL2, 1e-9, 50000))
    p_j = ctx.field(G.FIELD_P).astype(np.float64)
    for f, a in zip((G.FIELD_U, G.FIELD_V, G.FIELD_W), faces):
        ctx.set_field(f, a)
    st_m = ctx.project(_gphase(oph), dt, G.default_solve_cfg(G.STOP_REL2, 1e-9, 50000, G.PRECOND_MULTIGRID))
    print(f"iterations: jacobi {st_j.iterations}, multigrid {st_m.iterations}")
    assert st_j.converged == 1 and st_m.converged == 1
    assert st_m.true_rel_residual <= 1e-9 and st_m.fallback_used == 0
    assert st_m.iterations * 2 <= st_j.iterations
    assert rel_err(ctx.field(G.FIELD_P), p_j) < 1e-6


def test_invalid_cycles_are_refused():
    ctx = G.Context(8, 8, 8, 1.0)
    for bad in (dict(sweeps=0), dict(sweeps=17), dict(coarsest_sweeps=0), dict(omega=1.0), dict(omega=0.0),
                dict(correction=2.0), dict(correction=0.0), dict(chebyshev_ratio=1.0),
                dict(chebyshev_ratio=-3.0), dict(chebyshev_ratio=2000.0)):
        with pytest.raises(G.PfError, match="invalid multigrid cycle"):
            ctx.set_multigrid(**bad)
    assert ctx.multigrid().sweeps == G.mg_defaults().sweeps


def test_unknown_preconditioner_is_refused():
    dx, ctx, ps, oph = _prepare(O.SCENE_RT, 8, seed=1)
    _to_projection(ctx, oph, 1e-3)
    with pytest.raises(G.PfError, match="unknown preconditioner"):
        ctx.project(_gphase(oph), 1e-3, G.default_solve_cfg(G.STOP_REL2, 1e-9, 100, 7))
