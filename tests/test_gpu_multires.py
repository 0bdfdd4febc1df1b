"""Multires two-phase path (MRSim3DTP) on the HIP path against the multires oracle.

Bars: layouts, enumerated faces, markers, refinement boxes and fine-leaf counts
bit-exact; face mass and momentum within 1e-5 norm-wise relative error of the FP64
oracle; pressure solved to the reference's max-norm rule, compared with a tight
tolerance; particle state after whole steps within the stated tolerances.
"""
import numpy as np
import pytest

import pfflip_amd as G
import pyoracle as O
from _helpers import COARSE_CASES, bubble_lattice as _bubble_particles, rel_err

pytestmark = pytest.mark.gpu

FIELD_TOL = 1e-5


def _phase(rho_l=1.0, rho_g=0.01):
    ph = O.Phase(rho_l, rho_g, 1.0, 1.0)
    ph.rho_tilde_0 = O.calibrate_rho_tilde0(ph, 1.0)
    return ph, G.Phase(ph.rho_l, ph.rho_g, ph.alpha_phi, ph.rho_tilde_0)


def _ctx(n, leaves, parts):
    ctx = G.MRContext(n, n, n, 1.0)
    ctx.set_layout(leaves)
    ctx.upload(*parts)
    g = O.MRGrid(n, n, n, 1.0, faithful=False)
    g.set_leaves(leaves)
    x, y, z, u, v, w, t = parts
    return ctx, g, O.Particles(x, y, z, u, v, w, type=t)


@pytest.mark.parametrize("dims,box", [((16, 16, 16), [4, 4, 4, 8, 12, 12]), ((17, 10, 9), [4, 3, 2, 13, 9, 8])])
def test_layout_and_enumerated_faces(dims, box):
    nx, ny, nz = dims
    leaves = O.mr_layout_fine_box(nx, ny, nz, box)
    ctx = G.MRContext(nx, ny, nz, 1.0)
    ctx.set_layout(leaves)
    g = O.MRGrid(nx, ny, nz, 1.0)
    g.set_leaves(leaves)
    assert np.array_equal(ctx.layout(), g.leaves())
    for a in range(3):
        assert np.array_equal(ctx.face_valid(a), g.face_valid(a)), a


def test_markers_and_p2g_match_fp64_oracle():
    n = 16
    ph, gph = _phase()
    leaves = O.mr_layout_bubble_band(n, n, n)
    ctx, g, ps = _ctx(n, leaves, _bubble_particles(n))
    ctx.mr_mark_cells()
    g.mark_cells(ps)
    assert np.array_equal(ctx.field(G.FIELD_MARKER), g.dense(O.MK))
    ctx.mr_p2g(gph, 1.0)
    g.p2g(ps, ph, 1.0)
    for f in range(6):
        assert rel_err(ctx.field(f), g.dense(f)) < FIELD_TOL, f
        valid = g.face_valid(f % 3)
        assert not ((ctx.field(f) != 0) & (valid == 0)).any(), f
    for fv in range(3):
        mom = ctx.field(fv).astype(np.float64) * ctx.field(fv + 3)
        assert rel_err(mom, g.dense(fv) * g.dense(fv + 3)) < FIELD_TOL


def test_multires_entry_points_need_a_layout():
    # every entry point that launches a multires kernel refuses a context without a
    # layout (PF_ERR_STATE) rather than reading unset buffers; the state queries answer
    n = 16
    ph, gph = _phase()
    ctx = G.MRContext(n, n, n, 1.0)
    try:
        ctx.upload(*_bubble_particles(n))
        calls = (ctx.mr_mark_cells, lambda: ctx.mr_p2g(gph, 1.0), lambda: ctx.mr_project(gph, 0.02),
                 lambda: ctx.face_valid(0), lambda: ctx.mr_interface_diagnostics(gph),
                 lambda: ctx.mr_apply_surface_tension(gph, 0.02))
        for call in calls:
            with pytest.raises(G.PfError):
                call()
        assert len(ctx.layout()) == 0
        assert ctx.mr_update_refinement(G.MRRefineCfg(0, 1, 2, 1, 0, 0)) == (0, 0)  # refinement off: no change
        assert ctx.mr_retained_box()[0] is False
        assert len(ctx.layout()) == 0
        sim = G.MRSim3DTP(n, n, n, 1.0)
        sim.ctx.close()
        sim.ctx, sim.phase, sim.dynamic_refinement = ctx, gph, False
        with pytest.raises(G.PfError):
            sim.step()
        # and the same context works once a layout is set
        ctx.set_layout(O.mr_layout_bubble_band(n, n, n))
        ctx.mr_mark_cells()
        ctx.mr_p2g(gph, 1.0)
    finally:
        ctx.close()


def _explain(ctx, g, ps, f):
    """what differs, for the failure message: the mismatching faces, the markers and
    the binned particle counts per cell"""
    got, want = ctx.field(f), g.dense(f)
    bad = np.argwhere(np.abs(got - want) > 1e-5 * max(np.abs(want).max(), 1e-30))
    counts = np.zeros((ctx.nz, ctx.ny, ctx.nx), np.int64)
    cell = np.floor(np.stack([ps.z, ps.y, ps.x], 1)).astype(np.int64)
    np.add.at(counts, tuple(cell.T), 1)
    return (f"field {f}: {len(bad)} faces differ, first (k,j,i) {bad[:8].tolist()}; "
            f"markers equal {np.array_equal(ctx.field(G.FIELD_MARKER), g.dense(O.MK))}; "
            f"binned counts equal {np.array_equal(ctx.bin_counts(), counts)}; "
            f"particles {ctx.count()} of {len(ps.x)}")


def test_projection_matches_fp64_oracle():
    n, dt = 16, 0.02
    ph, gph = _phase()
    leaves = O.mr_layout_bubble_band(n, n, n)
    ctx, g, ps = _ctx(n, leaves, _bubble_particles(n))
    ctx.mr_mark_cells(); ctx.mr_p2g(gph, 1.0)
    g.mark_cells(ps); g.p2g(ps, ph, 1.0)
    for f in range(6):
        assert rel_err(ctx.field(f), g.dense(f)) < FIELD_TOL, _explain(ctx, g, ps, f)
    ctx.save_velocity(); ctx.apply_gravity(dt, -9.81); ctx.apply_walls()
    g.apply_gravity(dt, -9.81); g.apply_walls()
    for f in range(3):
        assert rel_err(ctx.field(f), g.dense(f)) < FIELD_TOL, _explain(ctx, g, ps, f)
    st = ctx.mr_project(gph, dt, G.mr_solve_cfg(20000, 1e-11, 0.0))
    ost = g.project(ph, dt, O.mr_solve_cfg(max_iterations=20000, abs_tol=1e-12, rel_tol=0.0))
    assert st.converged == 1 and st.true_max_residual < 1e-11
    assert st.fluid_cells == ost.active_cells and st.pinned_cell == ost.pinned_cell
    assert rel_err(ctx.field(G.FIELD_P), g.dense(O.P)) < 1e-6
    for f in range(3):
        assert rel_err(ctx.field(f), g.dense(f)) < 1e-5, f


def test_reference_stop_rule_and_defaults():
    # MRSim3DTP's defaults (cg_iters 160, cg_tol 1e-7, cg_rel_tol 1e-6): the solve
    # stops on |rhs - A x|_inf < max(1e-7, 1e-6 |rhs|_inf)
    n, dt = 16, 0.02
    ph, gph = _phase()
    ctx, g, ps = _ctx(n, O.mr_layout_bubble_band(n, n, n), _bubble_particles(n))
    ctx.mr_mark_cells(); ctx.mr_p2g(gph, 1.0); ctx.apply_gravity(dt, -9.81); ctx.apply_walls()
    st = ctx.mr_project(gph, dt, G.mr_solve_cfg())
    tol = max(1e-7, 1e-6 * st.initial_max_residual)
    assert st.converged == 1 and st.true_max_residual < tol


def test_steps_match_fp64_oracle_with_dynamic_refinement():
    n, dt = 16, 0.02
    ph, gph = _phase()
    parts = _bubble_particles(n, vel=False)
    leaves = O.mr_layout_fine_box(n, n, n)          # coarse everywhere: the first step refines
    ctx, g, ps = _ctx(n, leaves, parts)
    sim = G.MRSim3DTP(n, n, n, 1.0)
    sim.ctx = ctx
    sim.phase = gph
    sim.dt = dt
    sim.cg_iters, sim.cg_tol, sim.cg_rel_tol = 20000, 1e-11, 0.0
    osim = O.mr_sim(n, n, n, 1.0, phase=ph, dt=dt,
                    solve=O.mr_solve_cfg(max_iterations=20000, abs_tol=1e-12, rel_tol=0.0))
    pa = ps.copy()
    for s in range(2):
        st = sim.step()
        ost = O.mr_step(g, osim, pa)
        assert st.solve.converged == 1
        if s == 0:
            assert np.array_equal(ctx.layout(), g.leaves())
    x, y, z, u, v, w = ctx.download()
    assert max(np.abs(x - pa.x).max(), np.abs(y - pa.y).max(), np.abs(z - pa.z).max()) < 1e-5
    assert rel_err(np.concatenate([u, v, w]), np.concatenate([pa.u, pa.v, pa.w])) < 1e-4


def test_refinement_box_hysteresis_and_budget():
    # updateDynamicRefinement on a corner column with a gas pocket: the same box,
    # fine-leaf count and budget flag as the oracle
    n = 24
    pts = []
    for k in range(1, 10):
        for j in range(1, 12):
            for i in range(1, 9):
                for s in range(8):
                    pts.append((i + 0.25 + 0.5 * (s & 1), j + 0.25 + 0.5 * ((s >> 1) & 1),
                                k + 0.25 + 0.5 * ((s >> 2) & 1), 1 if (i - 4) ** 2 + (j - 4) ** 2 < 4 else 0))
    p = np.array(pts)
    x, y, z = (p[:, q].astype(np.float32) for q in range(3))
    t = p[:, 3].astype(np.uint8)
    ps = O.Particles(x, y, z, type=t)
    for max_fine in (0, 20):
        ctx = G.MRContext(n, n, n, 1.0)
        ctx.set_layout(O.mr_layout_fine_box(n, n, n))
        ctx.upload(x, y, z, type=t)
        fl, bl = ctx.mr_update_refinement(G.MRRefineCfg(1, 1, 2, 1, max_fine, 0))
        g = O.MRGrid(n, n, n, 1.0)
        g.set_leaves(O.mr_layout_fine_box(n, n, n))
        osim = O.mr_sim(n, n, n, 1.0, max_fine=max_fine)
        O.mr_dynamic_refinement(g, osim, ps)
        assert np.array_equal(ctx.layout(), g.leaves())
        assert (fl, bl) == (osim.dynamic_last_fine_leaves, osim.budget_limited)
        assert bl == (1 if max_fine else 0)


def _shared_system(n, filled):
    """the same multires system on the GPU and in the FP64 oracle: the GPU's FP32
    face fields after P2G, gravity and walls are copied into the oracle grid"""
    ph, gph = _phase()
    leaves = O.mr_layout_bubble_band(n, n, n)
    parts = list(_bubble_particles(n))
    if filled:
        pts = np.array([(i + 0.5, j + 0.5, k + 0.5) for k in range(1, n - 1) for j in range(1, n - 1)
                        for i in range(1, n - 1)])
        rng = np.random.default_rng(5)
        parts = [pts[:, 0].astype(np.float32), pts[:, 1].astype(np.float32), pts[:, 2].astype(np.float32)] + \
            [rng.normal(scale=0.3, size=len(pts)).astype(np.float32) for _ in range(3)] + \
            [np.zeros(len(pts), np.uint8)]
    ctx, g, ps = _ctx(n, leaves, parts)
    ctx.mr_mark_cells(); ctx.mr_p2g(gph, 1.0); ctx.apply_gravity(0.02, -9.81); ctx.apply_walls()
    g.mark_cells(ps)
    for f in range(6):
        a = ctx.field(f)
        for k, j, i in zip(*np.nonzero(a)):
            g.set(f, int(i), int(j), int(k), float(a[k, j, i]))
    return ctx, g, ph, gph


@pytest.mark.parametrize("filled", [False, True])
@pytest.mark.parametrize("case", range(7))
def test_reference_order_solve_with_options_matches_oracle(filled, case):
    # PF_MR_SOLVE_REFERENCE: the reference's FP64 iteration with its coarse
    # correction, coarse preconditioner, flexible beta and restarts on the GPU
    # against the oracle (itself bit-exact with the reference build) on the same
    # system; reductions run in another order, so counts may differ by a step
    opts = dict(COARSE_CASES[case])
    flexible = opts.pop("flexible", 0)
    n = 16
    ctx, g, ph, gph = _shared_system(n, filled)
    ocfg = O.mr_solve_cfg(max_iterations=400, flexible=flexible, **opts)
    ost = g.project(ph, 0.02, ocfg)
    gcfg = G.mr_solve_cfg(400, 1e-7, 1e-6, mode=G.MR_SOLVE_REFERENCE, use_flexible_cg_beta=flexible, **opts)
    st = ctx.mr_project(gph, 0.02, gcfg)
    d = ctx.mr_solve_detail()
    print(case, filled, st.iterations, ost.iterations, d.coarse_correction_sweeps, ost.cc_sweeps,
          d.coarse_preconditioner_applications, ost.cp_applications, flush=True)
    assert d.reference_order == 1
    # (with the work budget and a pin the reference itself stops at the cap)
    assert st.converged == ost.converged
    assert abs(st.iterations - ost.iterations) <= 2
    if ost.converged:
        assert st.true_max_residual < ost.effective_tolerance
    assert d.coarse_correction_cells == ost.cc_cells and d.coarse_preconditioner_cells == ost.cp_cells
    assert abs(d.coarse_correction_sweeps - ost.cc_sweeps) <= 1
    assert abs(d.coarse_preconditioner_applications - ost.cp_applications) <= 2
    assert abs(d.coarse_preconditioner_skipped_applications - ost.cp_skipped) <= 2
    assert d.coarse_preconditioner_auto_disabled == ost.cp_auto_disabled
    assert rel_err(ctx.field(G.FIELD_P), g.dense(O.P)) < 1e-6
    for f in range(3):
        assert rel_err(ctx.field(f), g.dense(f)) < 1e-5, f


def test_reference_order_rejects_invalid_options():
    # validateSolveConfig (multires_pressure3d.cpp:307-333)
    n = 16
    ctx, g, ph, gph = _shared_system(n, False)
    for bad in (dict(coarse_correction_min_scale=0.0), dict(coarse_preconditioner_scale=-1.0),
                dict(relaxation_sweeps=-1)):
        with pytest.raises(G.PfError):
            ctx.mr_project(gph, 0.02, G.mr_solve_cfg(mode=G.MR_SOLVE_REFERENCE, **bad))
    # the fast mode ignores the option fields
    st = ctx.mr_project(gph, 0.02, G.mr_solve_cfg(coarse_correction_min_scale=0.0))
    assert st.converged == 1 and ctx.mr_solve_detail().reference_order == 0
