"""Interface diagnostics and CSF surface tension on the GPU against the oracle
(src/driver/interface_diagnostics3d.cpp:362-440, sparse_sim3d_tp.cpp:865-874).

Both sides start from the same FP32 face masses, velocities and markers (the GPU's
own P2G output copied into the oracle grid), so every per-cell value is computed
from identical inputs in the same FP64 operation order.  Bars: counts, minima and
maxima exact; sums and means within 1e-12 relative (the GPU adds in another order);
face velocities after the CSF increments bit-exact.  The multires variants sample
each FLUID leaf cell once and follow the reference's leaf order; they are held to the
same bars on two-level layouts with the interface in fine and in coarse cells.
Through a whole step the GPU's FP32 P2G differs from the FP64 oracle's, so particle
state is held to the same 1e-4 norm-wise bar as the step without surface tension,
and the step's statistics are recomputed by the oracle from the GPU's own post-step
masses.
"""
import numpy as np
import pytest

import pfflip_amd as G
import pyoracle as O
from _helpers import bubble_lattice, phase, rel_err, scene, step_cfg

pytestmark = pytest.mark.gpu

IFACE_EXACT = ("sample_cells", "interface_cells", "finite", "surface_tension_candidate", "phi_min", "phi_max",
               "grad_max", "curvature_abs_max")
IFACE_SUMS = ("phi_mean", "grad_mean", "curvature_abs_mean")
ST_EXACT = ("enabled", "applied_cells", "finite", "curvature_smoothing_radius", "capillary_stable", "strength",
            "max_delta_speed_limit", "max_delta_speed", "raw_curvature_abs_max", "smoothed_curvature_abs_max",
            "capillary_dt_limit")
ST_SUMS = ("mean_delta_speed", "raw_curvature_abs_mean", "smoothed_curvature_abs_mean")
SUM_TOL = 1e-12


def _gphase(oph):
    return G.Phase(oph.rho_l, oph.rho_g, oph.alpha_phi, oph.rho_tilde_0)


def _close(a, b, tol=SUM_TOL):
    return abs(a - b) <= tol * max(abs(a), abs(b), 1e-300)


def _check_iface(gs, os_):
    for f in IFACE_EXACT:
        assert getattr(gs, f) == getattr(os_, f), f
    for f in IFACE_SUMS:
        assert _close(getattr(gs, f), getattr(os_, f)), (f, getattr(gs, f), getattr(os_, f))


def _check_surface(gs, os_):
    for f in ST_EXACT:
        assert getattr(gs, f) == getattr(os_, f), f
    for f in ST_SUMS:
        assert _close(getattr(gs, f), getattr(os_, f)), (f, getattr(gs, f), getattr(os_, f))


def _after_p2g(kind, n, seed=11):
    """GPU context after markCells + P2G (velocities normalised) of a seeded scene"""
    print(f"[iface] kind={kind} n={n} seed={seed}", flush=True)
    dx, x, y, z, t = scene(kind, n, seed=seed)
    rng = np.random.default_rng(seed)
    u, v, w = (rng.normal(scale=0.3, size=len(x)).astype(np.float32) for _ in range(3))
    ctx = G.Context(n, n, n, dx)
    ctx.upload(x, y, z, u, v, w, t)
    oph = phase()
    ctx.mark_cells()
    ctx.p2g(_gphase(oph), 1.0)
    return dx, ctx, oph


def _oracle_copy(ctx, n, dx, fields=(0, 1, 2, 3, 4, 5)):
    """an oracle grid holding the GPU's face fields and markers"""
    g = O.Grid(n, n, n, dx, faithful=True)
    for f in fields:
        g.load_dense(f, ctx.field(f).astype(np.float64))
    g.load_dense(O.MK, ctx.field(G.FIELD_MARKER).astype(np.float64))
    return g


@pytest.mark.parametrize("kind", [O.SCENE_RT, O.SCENE_BUBBLE, O.SCENE_DAMBREAK])
def test_interface_diagnostics_match_oracle(kind):
    n = 24
    dx, ctx, oph = _after_p2g(kind, n)
    g = _oracle_copy(ctx, n, dx, fields=(3, 4, 5))
    gs = ctx.interface_diagnostics(_gphase(oph))
    os_ = g.interface_diagnostics(oph)
    assert gs.sample_cells > 0 and gs.interface_cells > 0
    _check_iface(gs, os_)
    # the diagnostics read the grid and change nothing
    gs2 = ctx.interface_diagnostics(_gphase(oph))
    assert bytes(gs2) == bytes(gs)


@pytest.mark.parametrize("strength,maxd,radius", [(0.02, 0.25, 0), (0.5, 0.01, 1), (2.0, 0.0, 2), (0.1, 0.2, 3)])
def test_surface_tension_matches_oracle(strength, maxd, radius):
    n = 24
    dt = 0.01
    dx, ctx, oph = _after_p2g(O.SCENE_BUBBLE, n, seed=5)
    g = _oracle_copy(ctx, n, dx)
    gs = ctx.apply_surface_tension(_gphase(oph), dt, strength, maxd, radius)
    os_ = g.surface_tension(oph, dt, strength, maxd, radius)
    assert gs.applied_cells > 0
    _check_surface(gs, os_)
    for f in range(3):
        assert np.array_equal(ctx.field(f).astype(np.float64), g.dense(f)), f


def test_surface_tension_disabled_and_capillary_limit():
    n = 16
    dx, ctx, oph = _after_p2g(O.SCENE_BUBBLE, n, seed=2)
    u0 = [ctx.field(f).copy() for f in range(3)]
    g = _oracle_copy(ctx, n, dx)
    # strength 0: initSurfaceStats only (enabled 0, nothing applied)
    gs = ctx.apply_surface_tension(_gphase(oph), 0.01, 0.0, 0.25, 5)
    _check_surface(gs, g.surface_tension(oph, 0.01, 0.0, 0.25, 5))
    assert gs.enabled == 0 and gs.curvature_smoothing_radius == 3
    for f in range(3):
        assert np.array_equal(ctx.field(f), u0[f])
    # a dt above sqrt(rho_min dx^3 / (2 pi strength)) is flagged unstable
    big = ctx.apply_surface_tension(_gphase(oph), 10.0, 0.5, 0.25, 0)
    lim = np.sqrt(min(oph.rho_l, oph.rho_g) * dx ** 3 / (2 * np.pi * 0.5))
    assert big.capillary_stable == 0 and abs(big.capillary_dt_limit - lim) <= 1e-15 * lim


def test_nan_mass_matches_oracle():
    # a NaN face mass gives a NaN phase in the two cells that share the face
    # (std::min(NaN, 1.0) is NaN): the diagnostics report finite = 0 and skip the
    # cells; the CSF turns the NaN gradients next to them into zero increments
    n = 20
    dx, ctx, oph = _after_p2g(O.SCENE_BUBBLE, n, seed=7)
    mk = ctx.field(G.FIELD_MARKER)
    k, j, i = (int(q[len(q) // 2]) for q in np.nonzero(mk == 1))
    mv = ctx.field(G.FIELD_MV)
    mv[k, j, i] = np.nan
    ctx.set_field(G.FIELD_MV, mv)
    g = _oracle_copy(ctx, n, dx)
    gs = ctx.interface_diagnostics(_gphase(oph))
    os_ = g.interface_diagnostics(oph)
    assert os_.finite == 0
    _check_iface(gs, os_)
    gt = ctx.apply_surface_tension(_gphase(oph), 0.01, 0.5, 0.2, 1)
    ot = g.surface_tension(oph, 0.01, 0.5, 0.2, 1)
    _check_surface(gt, ot)
    for f in (0, 2):
        assert np.array_equal(ctx.field(f).astype(np.float64), g.dense(f)), f
    gv, ov = ctx.field(1).astype(np.float64), g.dense(1)
    assert np.array_equal(np.isnan(gv), np.isnan(ov)) and np.array_equal(gv[~np.isnan(gv)], ov[~np.isnan(ov)])


def _sim(n, dx, ctx, oph, csf=True, diag=True):
    sim = G.SparseSim3DTP(n, n, n, dx)
    sim.ctx = ctx
    sim.phase = _gphase(oph)
    sim.dt = 2e-3
    sim.stop_rule, sim.cg_tol, sim.cg_iters = G.STOP_REL2, 1e-9, 20000
    sim.surface_tension = csf
    sim.surface_tension_strength = 0.3
    sim.surface_tension_max_delta_speed = 0.1
    sim.surface_tension_curvature_smoothing_radius = 1
    sim.interface_diagnostics = diag
    return sim


def test_step_with_surface_tension_matches_oracle():
    n = 20
    print(f"[iface-step] n={n}", flush=True)
    dx, x, y, z, t = scene(O.SCENE_BUBBLE, n, seed=2)
    oph = phase()
    ctx = G.Context(n, n, n, dx)
    ctx.upload(x, y, z, type=t)
    sim = _sim(n, dx, ctx, oph)
    cfg = step_cfg(oph, dt=2e-3, cg_iters=20000, cg_tol=1e-12, stop_rule=O.STOP_REL2)
    cfg.surface_tension, cfg.surface_tension_strength = 1, 0.3
    cfg.surface_tension_max_delta_speed, cfg.surface_tension_smoothing_radius = 0.1, 1
    go = O.Grid(n, n, n, dx, faithful=False)
    pa = O.Particles(x, y, z, type=t)
    for _ in range(2):
        st = sim.step()
        ost = go.step(pa, cfg)
        assert st.solve.converged == 1
        assert st.surface.enabled == 1 and st.surface.applied_cells > 0
        assert st.iface.interface_cells > 0
        # the step's statistics against the oracle on the GPU's own masses and markers
        g = _oracle_copy(ctx, n, dx, fields=(3, 4, 5))
        _check_iface(st.iface, g.interface_diagnostics(oph))
        _check_surface(st.surface, g.surface_tension(oph, st.effective_dt, 0.3, 0.1, 1))
        # and, with FP32 P2G on the GPU, close to the FP64 oracle step's
        assert abs(st.surface.applied_cells - ost.surface.applied_cells) <= 0.02 * ost.surface.applied_cells
        assert sim.surface_tension_stats_last.applied_cells == st.surface.applied_cells
        assert sim.interface_diagnostics_last.interface_cells == st.iface.interface_cells
    gx, gy, gz, gu, gv, gw = ctx.download()
    assert np.abs(gx - pa.x).max() < 1e-6 and np.abs(gy - pa.y).max() < 1e-6 and np.abs(gz - pa.z).max() < 1e-6
    assert rel_err(np.concatenate([gu, gv, gw]), np.concatenate([pa.u, pa.v, pa.w])) < 1e-4


def test_step_diagnostics_flag_changes_only_the_statistics():
    n = 16
    dx, x, y, z, t = scene(O.SCENE_RT, n, seed=9)
    oph = phase()
    out = []
    for diag in (True, False):
        ctx = G.Context(n, n, n, dx)
        ctx.upload(x, y, z, type=t)
        sim = _sim(n, dx, ctx, oph, csf=False, diag=diag)
        st = sim.step()
        out.append((st, ctx.download()))
    (s1, p1), (s2, p2) = out
    assert s1.iface.sample_cells > 0 and s1.iface.finite == 1
    assert s2.iface.sample_cells == 0 and s2.iface.finite == 1
    # surface tension off: SurfaceTensionStats3D() defaults
    assert s1.surface.enabled == 0 and s1.surface.finite == 1 and s1.surface.capillary_stable == 1
    for a, b in zip(p1, p2):
        assert np.array_equal(a, b)


# ---- multires variants (diagnoseMRInterface3D, applyMRSurfaceTension3D)

# fine boxes on a level-1 base for a 32^3 bubble tank (water below y = 16): the
# interface band fine, liquid below y = 8 coarse, the water surface in coarse cells
MR_BOXES = {"band": None, "mid": [0, 8, 0, 32, 20, 32], "low": [0, 0, 0, 32, 8, 32]}


def _mr_leaves(n, box):
    return O.mr_layout_bubble_band(n, n, n) if MR_BOXES[box] is None else O.mr_layout_fine_box(n, n, n, MR_BOXES[box])


def _mr_after_p2g(box, n=32, seed=3):
    print(f"[mr-iface] box={box} n={n} seed={seed}", flush=True)
    leaves = _mr_leaves(n, box)
    ctx = G.MRContext(n, n, n, 1.0)
    ctx.set_layout(leaves)
    ctx.upload(*bubble_lattice(n, seed))
    oph = phase(1.0, 0.01)
    ctx.mr_mark_cells()
    ctx.mr_p2g(_gphase(oph), 1.0)
    return ctx, oph, leaves


def _mr_oracle_copy(ctx, n, leaves, fields=(0, 1, 2, 3, 4, 5)):
    g = O.MRGrid(n, n, n, 1.0, faithful=True)
    g.set_leaves(leaves)
    for f in fields:
        g.load_dense(f, ctx.field(f).astype(np.float64))
    g.load_dense(O.MK, ctx.field(G.FIELD_MARKER).astype(np.float64))
    return g


@pytest.mark.parametrize("box", sorted(MR_BOXES))
def test_mr_interface_diagnostics_match_oracle(box):
    n = 32
    ctx, oph, leaves = _mr_after_p2g(box, n)
    g = _mr_oracle_copy(ctx, n, leaves, fields=(3, 4, 5))
    gs = ctx.mr_interface_diagnostics(_gphase(oph))
    assert gs.interface_cells > 0
    _check_iface(gs, g.interface_diagnostics(oph))


@pytest.mark.parametrize("box,strength,maxd,radius", [("band", 0.02, 0.25, 0), ("low", 0.5, 0.01, 1),
                                                      ("mid", 2.0, 0.0, 2), ("low", 0.1, 0.2, 3)])
def test_mr_surface_tension_matches_oracle(box, strength, maxd, radius):
    n = 32
    ctx, oph, leaves = _mr_after_p2g(box, n, seed=5)
    g = _mr_oracle_copy(ctx, n, leaves)
    gs = ctx.mr_apply_surface_tension(_gphase(oph), 0.01, strength, maxd, radius)
    assert gs.applied_cells > 0
    _check_surface(gs, g.surface_tension(oph, 0.01, strength, maxd, radius))
    for f in range(3):
        assert np.array_equal(ctx.field(f).astype(np.float64), g.dense(f)), f


def test_interface_entry_points_check_the_grid_kind():
    n = 16
    ctx, oph, leaves = _mr_after_p2g("band", n)
    with pytest.raises(G.PfError):
        ctx.interface_diagnostics(_gphase(oph))
    sp = G.MRContext(n, n, n, 1.0)      # no layout set: a sparse context
    with pytest.raises(G.PfError):
        sp.mr_interface_diagnostics(_gphase(oph))


def test_mr_step_with_surface_tension_matches_oracle():
    n, dt = 32, 0.02
    print(f"[mr-iface-step] n={n}", flush=True)
    leaves = _mr_leaves(n, "low")
    parts = bubble_lattice(n, vel=False)
    oph = phase(1.0, 0.01)
    ctx = G.MRContext(n, n, n, 1.0)
    ctx.set_layout(leaves)
    ctx.upload(*parts)
    sim = G.MRSim3DTP(n, n, n, 1.0)
    sim.ctx = ctx
    sim.phase = _gphase(oph)
    sim.dt = dt
    sim.dynamic_refinement = False
    sim.cg_iters, sim.cg_tol, sim.cg_rel_tol = 20000, 1e-11, 0.0
    sim.surface_tension, sim.surface_tension_strength = True, 0.3
    sim.surface_tension_max_delta_speed, sim.surface_tension_curvature_smoothing_radius = 0.1, 1
    osim = O.mr_sim(n, n, n, 1.0, phase=oph, dt=dt, dynamic=0,
                    solve=O.mr_solve_cfg(max_iterations=20000, abs_tol=1e-12, rel_tol=0.0))
    osim.surface_tension, osim.surface_tension_strength = 1, 0.3
    osim.surface_tension_max_delta_speed, osim.surface_tension_smoothing_radius = 0.1, 1
    go = O.MRGrid(n, n, n, 1.0, faithful=False)
    go.set_leaves(leaves)
    x, y, z, u, v, w, t = parts
    pa = O.Particles(x, y, z, u, v, w, type=t)
    for _ in range(2):
        st = sim.step()
        ost = O.mr_step(go, osim, pa)
        assert st.solve.converged == 1
        assert st.surface.enabled == 1 and st.surface.applied_cells > 0 and st.iface.interface_cells > 0
        g = _mr_oracle_copy(ctx, n, leaves, fields=(3, 4, 5))
        _check_iface(st.iface, g.interface_diagnostics(oph))
        _check_surface(st.surface, g.surface_tension(oph, st.effective_dt, 0.3, 0.1, 1))
        assert abs(st.surface.applied_cells - ost.surface.applied_cells) <= 0.02 * ost.surface.applied_cells
        assert sim.surface_tension_stats_last.applied_cells == st.surface.applied_cells
    gx, gy, gz, gu, gv, gw = ctx.download()
    assert max(np.abs(gx - pa.x).max(), np.abs(gy - pa.y).max(), np.abs(gz - pa.z).max()) < 1e-5
    assert rel_err(np.concatenate([gu, gv, gw]), np.concatenate([pa.u, pa.v, pa.w])) < 1e-4
