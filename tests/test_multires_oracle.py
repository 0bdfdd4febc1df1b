"""The multires oracle (oracle/pf_oracle_mr.c) pinned by closed forms, an independent
dense assembly of the finite-volume system, the reference's own multires doctests,
and -- where oracle/_ref was built -- bit-exactness against the reference's
MRSim3DTP and stage functions compiled from its sources."""
import numpy as np
import pytest

import pyoracle as O
from _helpers import COARSE_CASES

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built")


def _phase(rho_l=1.0, rho_g=0.01):
    ph = O.Phase(rho_l, rho_g, 1.0, 1.0)
    ph.rho_tilde_0 = O.calibrate_rho_tilde0(ph, 1.0)
    return ph


def _levels(g):
    """per fine cell: the level of its leaf (from the leaf list)"""
    lv = np.full((g.nz, g.ny, g.nx), -1, np.int64)
    for L, bx, by, bz in g.leaves():
        s = 4 * (1 << L)
        lv[bz * s:(bz + 1) * s, by * s:(by + 1) * s, bx * s:(bx + 1) * s] = L
    return lv


def test_layout_counts_closed_form():
    # setCoarseEverywhere(1) on 16^3: 2^3 blocks of 8^3 cells; refining the box
    # [4,8)x[4,12)x[4,12) splits the level-1 blocks it touches into 8 children each
    a = O.mr_layout_fine_box(16, 16, 16)
    assert len(a) == 8 and set(a[:, 0]) == {1}
    b = O.mr_layout_fine_box(16, 16, 16, [4, 4, 4, 8, 12, 12])
    # blocks touched: bx in {0}, by in {0, 1}, bz in {0, 1} -> 4 split into 32 fine blocks
    assert (b[:, 0] == 0).sum() == 32 and (b[:, 0] == 1).sum() == 4
    # sorted by (level, bx, by, bz) as std::tie orders them
    assert [tuple(r) for r in b] == sorted(tuple(r) for r in b)


def test_face_enumeration_matches_the_level_rule():
    # a fine face is enumerated iff it is the low face of the leaf cell on its high side
    # (coordinate divisible by that cell's step) or the high face (clipped to the domain)
    # of the leaf cell on its low side
    g = O.MRGrid(17, 10, 9, 1.0)
    g.set_leaves(O.mr_layout_fine_box(17, 10, 9, [4, 3, 2, 13, 9, 8]))
    lv = _levels(g)
    n = (g.nx, g.ny, g.nz)
    for axis in range(3):
        v = g.face_valid(axis)
        want = np.zeros_like(v)
        fz, fy, fx = v.shape
        for k in range(fz):
            for j in range(fy):
                for i in range(fx):
                    c = [i, j, k]
                    hi = c.copy()
                    lo = c.copy()
                    lo[axis] -= 1
                    ok = False
                    if hi[axis] < n[axis]:
                        s = 1 << lv[hi[2], hi[1], hi[0]]
                        ok |= hi[axis] % s == 0
                    if lo[axis] >= 0:
                        s = 1 << lv[lo[2], lo[1], lo[0]]
                        ok |= c[axis] % s == 0 or c[axis] == n[axis]
                    want[k, j, i] = ok
        assert np.array_equal(v, want), axis


def test_p2g_momentum_conserved_across_coarse_fine_boundary():
    # tests/test_multires_transfer3d_tp.cpp:38-55
    g = O.MRGrid(16, 16, 16, 1.0, faithful=False)
    g.set_leaves(O.mr_layout_fine_box(16, 16, 16, [4, 4, 4, 8, 12, 12]))
    pp = O.Phase(1.0, 0.01, 1.0, 1.0)
    g.p2g(O.Particles([7.75], [6.5], [6.5], [4.0], [1.0], [-2.0], type=[0]), pp, 1.0)
    for f, want in ((O.U, 4.0), (O.V, 1.0), (O.W, -2.0)):
        assert (g.dense(f) * g.dense(f + 3)).sum() == pytest.approx(want, rel=1e-6)


def test_p2g_gas_density_and_odd_boundary_faces():
    # tests/test_multires_transfer3d_tp.cpp:57-72 and :134-165
    g = O.MRGrid(16, 16, 16, 1.0, faithful=False)
    g.set_leaves(O.mr_layout_fine_box(16, 16, 16, level=0))
    pp = O.Phase(1.0, 0.01, 1.0, 1.0)
    g.p2g(O.Particles([5.0], [4.5], [6.5], [-6.0], [3.0], [2.0], type=[1]), pp, 2.0)
    assert (g.dense(O.U) * g.dense(O.MU)).sum() == pytest.approx(0.01 * 2.0 * -6.0, rel=1e-6)
    g = O.MRGrid(17, 10, 9, 1.0, faithful=False)
    g.set_leaves(O.mr_layout_fine_box(17, 10, 9, [4, 3, 2, 13, 9, 8]))
    g.p2g(O.Particles([12.75, 4.25], [8.25, 3.5], [7.25, 2.5], [2.0, -1.0], [-3.0, 4.0], [1.0, -2.0], type=[0, 1]),
          pp, 1.0)
    for axis in range(3):
        touched = g.dense(axis + 3) != 0
        assert not (touched & (g.face_valid(axis) == 0)).any()


def _closed_box_layout(g):
    return O.mr_layout_fine_box(g.nx, g.ny, g.nz, [4, 4, 4, 8, 12, 12])


def test_projection_matches_an_independent_dense_assembly():
    # finite-volume system assembled here by leaf-cell pairs: each shared fine face
    # patch carries dt beta dx^2 / (h_a/2 + h_b/2); row a is divided by its volume
    # h_a^3 (multires_pressure3d.cpp:214-244); AIR neighbours are Dirichlet 0
    n = 16
    g = O.MRGrid(n, n, n, 1.0, faithful=False)
    g.set_leaves(_closed_box_layout(g))
    rng = np.random.default_rng(3)
    pts = rng.uniform(1.0, 15.0, size=(400, 3))
    pts = pts[pts[:, 1] < 9.0]
    ps = O.Particles(pts[:, 0], pts[:, 1], pts[:, 2], rng.normal(size=len(pts)), rng.normal(size=len(pts)),
                     rng.normal(size=len(pts)), type=np.zeros(len(pts), np.uint8))
    pp = _phase(1.0, 1.0)  # one phase: beta = 1 on every face
    g.mark_cells(ps)
    g.p2g(ps, pp, 1.0)
    g.apply_walls()
    U, V, W = g.dense(O.U), g.dense(O.V), g.dense(O.W)
    mk = g.dense(O.MK)
    lv = _levels(g)
    dt = 0.05
    # leaf cells: anchor = fine cell rounded down to the leaf step
    anchors = {}
    for k in range(n):
        for j in range(n):
            for i in range(n):
                s = 1 << lv[k, j, i]
                anchors[(i, j, k)] = (i - i % s, j - j % s, k - k % s, s)
    fluid = sorted({a for c, a in anchors.items() if mk[c[2], c[1], c[0]] == 1})
    index = {a: t for t, a in enumerate(fluid)}
    N = len(fluid)
    A = np.zeros((N, N))
    b = np.zeros(N)
    vel = (U, V, W)
    for t, (x0, y0, z0, s) in enumerate(fluid):
        Vol = float(s ** 3)
        for axis in range(3):
            for side, sign in ((0, -1.0), (1, 1.0)):
                for u1 in range(s):
                    for u2 in range(s):
                        c = [x0, y0, z0]
                        ta, tb = [a for a in range(3) if a != axis]
                        c[ta] += u1
                        c[tb] += u2
                        f = list(c)
                        nb = list(c)
                        if side == 0:
                            nb[axis] -= 1
                        else:
                            f[axis] += s
                            nb[axis] += s
                        inside = all(0 <= nb[q] < n for q in range(3))
                        m = mk[nb[2], nb[1], nb[0]] if inside else 2
                        if m == 2:
                            continue
                        h_b = float(1 << lv[nb[2], nb[1], nb[0]])
                        coeff = dt * 1.0 / (0.5 * s + 0.5 * h_b) / Vol
                        A[t, t] += coeff
                        if m == 1:
                            A[t, index[anchors[tuple(nb)]]] -= coeff
                        b[t] -= sign * vel[axis][f[2], f[1], f[0]] / Vol
    x = np.linalg.solve(A, b)
    st = g.project(pp, dt, O.mr_solve_cfg(max_iterations=5000, abs_tol=1e-13, rel_tol=0.0))
    assert st.converged == 1 and st.pinned_cell == -1 and st.active_cells == N
    P = g.dense(O.P)
    got = np.array([P[z0, y0, x0] for (x0, y0, z0, s) in fluid])
    assert np.allclose(got, x, rtol=1e-8, atol=1e-10)


def _bubble_ref(n=16):
    r = O.RefMR(n, n, n, 1.0)
    r.init_bubble()
    return r


@needs_ref
@pytest.mark.parametrize("box", [None, [3, 2, 1, 9, 7, 6], [0, 0, 0, 24, 10, 20], [5, 5, 5, 6, 6, 6]])
def test_layout_bit_exact(box):
    assert np.array_equal(O.mr_layout_fine_box(24, 20, 20, box), O.ref_mr_layout_fine_box(24, 20, 20, box))


@needs_ref
def test_stages_bit_exact_on_the_bubble_band():
    n = 16
    r = _bubble_ref(n)
    ph = _phase()
    r.config(dt=0.02, phase=ph)
    ps = r.particles()
    rng = np.random.default_rng(1)
    for a in (ps.u, ps.v, ps.w):
        a[:] = rng.normal(size=len(ps))
    r.set_particles(ps)
    g = O.MRGrid(n, n, n, 1.0, faithful=True)
    g.set_leaves(r.leaves())
    assert np.array_equal(g.leaves(), O.mr_layout_bubble_band(n, n, n))
    for a in range(3):
        assert np.array_equal(g.face_valid(a), r.face_valid(a))
    r.p2g()
    g.p2g(ps, ph, 1.0)
    for f in range(6):
        assert np.array_equal(g.dense(f), r.dense(f)), f
    g.mark_cells(ps)
    mk = g.dense(O.MK)
    for k in range(n):
        for j in range(n):
            for i in range(n):
                r.set(O.MK, i, j, k, mk[k, j, i])
    saved_r, saved_g = r.clone(), g.clone()
    for cfg in (O.mr_solve_cfg(), O.mr_solve_cfg(flexible=1), O.mr_solve_cfg(sweeps=3, restart=0),
                O.mr_solve_cfg(jacobi=0, max_iterations=40)):
        rr, gg = saved_r.clone(), saved_g.clone()
        st = gg.project(ph, 0.02, cfg)
        out = rr.project(0.02, cfg)
        assert (st.iterations, st.restarts, st.beta_resets, st.relaxation_accepted, st.converged) == \
            tuple(int(v) for v in (out[0], out[1], out[2], out[3], out[5]))
        assert st.final_residual == out[8] and st.pinned_cell == out[11]
        for f in (O.U, O.V, O.W, O.P):
            assert np.array_equal(gg.dense(f), rr.dense(f)), f
    # G2P (against the pre-projection copy) and advection through the projected field
    r = saved_r.clone()
    r.project(0.02, O.mr_solve_cfg())
    r.g2p(saved_r, 0.95, 0.9)
    gg = saved_g.clone()
    gg.project(ph, 0.02, O.mr_solve_cfg())
    pa = ps.copy()
    gg.g2p(saved_g, pa, 0.95, 0.9)
    rp = r.particles()
    assert np.array_equal(rp.u, pa.u) and np.array_equal(rp.v, pa.v) and np.array_equal(rp.w, pa.w)
    for order in (2, 3):
        rc = r.clone()
        out = rc.advect(0.02, order)
        pb = pa.copy()
        es = gg.advect(pb, 0.02, order)
        rp = rc.particles()
        assert np.array_equal(rp.x, pb.x) and np.array_equal(rp.y, pb.y) and np.array_equal(rp.z, pb.z)
        assert [es.clamped_liquid, es.clamped_gas, es.x_lo, es.x_hi, es.y_lo, es.y_hi, es.z_lo, es.z_hi] == \
            [int(v) for v in out]


def _compare_step(g, r, ps):
    rp = r.particles()
    for a in ("x", "y", "z", "u", "v", "w"):
        assert np.array_equal(getattr(ps, a), getattr(rp, a)), a
    for f in range(8):
        assert np.array_equal(g.dense(f), r.dense(f)), f
    assert np.array_equal(g.leaves(), r.leaves())


@needs_ref
def test_step_bit_exact_bubble_band():
    n = 16
    r = _bubble_ref(n)
    ph = _phase()
    r.config(dt=0.02, phase=ph)
    ps = r.particles()
    g = O.MRGrid(n, n, n, 1.0, faithful=True)
    g.set_leaves(r.leaves())
    sim = O.mr_sim(n, n, n, 1.0, phase=ph)
    for _ in range(4):
        out = r.step()
        st = O.mr_step(g, sim, ps)
        assert (st.solve.iterations, st.solve.converged) == (int(out[2]), int(out[3]))
        _compare_step(g, r, ps)


@needs_ref
def test_step_bit_exact_with_moving_refinement_and_budget():
    # a liquid column in one corner with a gas pocket: the fine box follows the
    # particles (hysteresis 1) and a fine-leaf budget shrinks it
    n = 24
    pts = []
    for k in range(1, 10):
        for j in range(1, 12):
            for i in range(1, 9):
                for s in range(8):
                    pts.append((i + 0.25 + 0.5 * (s & 1), j + 0.25 + 0.5 * ((s >> 1) & 1),
                                k + 0.25 + 0.5 * ((s >> 2) & 1), 1 if (i - 4) ** 2 + (j - 4) ** 2 < 4 else 0))
    pts = np.array(pts)
    ps = O.Particles(pts[:, 0], pts[:, 1], pts[:, 2], type=pts[:, 3].astype(np.uint8))
    ph = _phase(1000.0, 1.0)
    for max_fine in (0, 20):
        r = O.RefMR(n, n, n, 1.0)
        r.set_leaves(O.mr_layout_fine_box(n, n, n))
        r.config(dt=0.01, phase=ph, max_fine=max_fine)
        r.set_particles(ps)
        g = O.MRGrid(n, n, n, 1.0, faithful=True)
        g.set_leaves(r.leaves())
        sim = O.mr_sim(n, n, n, 1.0, phase=ph, dt=0.01, max_fine=max_fine)
        pa = ps.copy()
        for _ in range(3):
            out = r.step()
            O.mr_step(g, sim, pa)
            assert sim.dynamic_last_fine_leaves == int(out[11]) and sim.budget_limited == int(out[12])
            _compare_step(g, r, pa)
        if max_fine:
            assert sim.budget_limited == 1


def _coarse_case(n, filled):
    """a bubble-band system on the reference build and the oracle, markers shared;
    `filled`: liquid in every interior cell (no AIR, so the solve pins a cell)"""
    r = _bubble_ref(n)
    ph = _phase()
    r.config(dt=0.02, phase=ph)
    ps = r.particles()
    if filled:
        pts = [(i + 0.5, j + 0.5, k + 0.5) for k in range(1, n - 1) for j in range(1, n - 1) for i in range(1, n - 1)]
        p = np.array(pts)
        ps = O.Particles(p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(), type=np.zeros(len(p), np.uint8))
    rng = np.random.default_rng(7)
    for a in (ps.u, ps.v, ps.w):
        a[:] = rng.normal(size=len(ps))
    r.set_particles(ps)
    g = O.MRGrid(n, n, n, 1.0, faithful=True)
    g.set_leaves(r.leaves())
    r.p2g()
    g.p2g(ps, ph, 1.0)
    g.mark_cells(ps)
    mk = g.dense(O.MK)
    for k in range(n):
        for j in range(n):
            for i in range(n):
                r.set(O.MK, i, j, k, mk[k, j, i])
    return r, g, ph




@needs_ref
@pytest.mark.parametrize("filled", [False, True])
@pytest.mark.parametrize("opts", COARSE_CASES)
def test_coarse_machinery_bit_exact(filled, opts):
    # level-1 aggregation, volume-weighted restriction, piecewise-constant
    # prolongation and the pinned R A P coarse CG (multires_pressure3d.cpp:39-70,
    # 470-700), used as the pre-solve correction with its scale search (:886-1060) and
    # as the additive preconditioner with rz-gain guard, work budget and auto-disable
    # (:1119-1341): statistics and fields equal the reference build's
    n = 16
    r, g, ph = _coarse_case(n, filled)
    opts = dict(opts)
    flexible = opts.pop("flexible", 0)
    cfg = O.mr_solve_cfg(max_iterations=400, flexible=flexible, **opts)
    st = g.project(ph, 0.02, cfg)
    out = r.project(0.02, cfg)
    assert st.pinned_cell == out[11] and (st.pinned_cell >= 0) == filled
    assert (st.iterations, st.restarts, st.beta_resets, st.converged, st.breakdown) == \
        tuple(int(v) for v in (out[0], out[1], out[2], out[5], out[6]))
    assert st.final_residual == out[8]
    cc = (st.cc_cells, st.cc_iterations, st.cc_sweeps, st.cc_accepted_sweeps, st.cc_rejected_sweeps, st.cc_accepted)
    assert cc == tuple(int(v) for v in out[12:18])
    assert st.cc_last_scale == out[18] and st.cc_final_residual == out[19]
    cp = (st.cp_cells, st.cp_applications, st.cp_accepted, st.cp_rejected, st.cp_skipped, st.cp_budget_limited,
          st.cp_iterations, st.cp_auto_disabled, st.cp_auto_disable_iteration)
    assert cp == tuple(int(v) for v in out[20:29])
    assert st.cp_last_rz_gain == out[29]
    if opts.get("use_coarse_correction"):
        assert st.cc_cells > 0 and st.cc_sweeps > 0
    if opts.get("use_coarse_preconditioner"):
        assert st.cp_applications + st.cp_skipped > 0
    for f in (O.U, O.V, O.W, O.P):
        assert np.array_equal(g.dense(f), r.dense(f)), f


def test_coarse_correction_lowers_the_residual_and_invalid_configs_are_rejected():
    # without the reference build: an accepted coarse sweep never raises the max-norm
    # residual (the scale search keeps the best of 1, 1/2, ... 1/64 against the start);
    # validateSolveConfig's cases (multires_pressure3d.cpp:307-333) stop the solve
    n = 16
    ph = _phase()
    g = O.MRGrid(n, n, n, 1.0)
    g.set_leaves(O.mr_layout_bubble_band(n, n, n))
    pts = [(i + 0.5, j + 0.5, k + 0.5) for k in range(1, n - 1) for j in range(1, n // 2) for i in range(1, n - 1)]
    p = np.array(pts)
    rng = np.random.default_rng(3)
    ps = O.Particles(p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(), rng.normal(size=len(p)),
                     rng.normal(size=len(p)), rng.normal(size=len(p)), type=np.zeros(len(p), np.uint8))
    g.mark_cells(ps)
    g.p2g(ps, ph, 1.0)
    base = g.clone()
    st = g.project(ph, 0.02, O.mr_solve_cfg(max_iterations=0, use_coarse_correction=1, coarse_correction_sweeps=4))
    assert st.cc_sweeps >= 1 and st.cc_accepted == 1
    assert st.final_residual < st.initial_residual and st.cc_final_residual == st.final_residual
    for bad in (dict(coarse_correction_min_scale=0.0), dict(coarse_correction_min_scale=1.5),
                dict(coarse_preconditioner_scale=-1.0), dict(coarse_preconditioner_auto_disable_after=-1)):
        st = base.clone().project(ph, 0.02, O.mr_solve_cfg(**bad))
        assert st.breakdown == 1 and st.iterations == 0, bad
