"""The oracle against the reference's own known answers (its doctest expectations)
and against closed forms / brute force that a plausible mistake would break.

Each test names the reference test it restates (paths relative to the reference).
"""
import math

import numpy as np
import pytest

import pyoracle as O
from _helpers import phase, rel_err, scene


def _seed_cell(i, j, k, t, xs, ys, zs, ts):
    # the reference's lattice seeding, sparse_sim3d_tp.cpp:12-20
    for s in range(8):
        xs.append(i + 0.25 + 0.5 * (s & 1))
        ys.append(j + 0.25 + 0.5 * ((s >> 1) & 1))
        zs.append(k + 0.25 + 0.5 * ((s >> 2) & 1))
        ts.append(t)


def test_p2g_conserves_normalized_cubic_momentum():
    # tests/test_sparse_transfer3d_tp.cpp:11-29
    g = O.Grid(8, 8, 8, 1.0)
    ps = O.Particles([3.0], [2.5], [2.5], [4.0], [3.0], [2.0], [0])
    g.p2g(ps, O.Phase(1.0, 0.01, 1.0, 1.0), 1.0)
    momU = float((g.dense(O.U) * g.dense(O.MU)).sum())
    assert momU == pytest.approx(4.0, rel=1e-6)
    nb = len(g.active_blocks(O.MU))
    assert 0 < nb < 3 * 2 * 2  # muf has ceil(9/4)*2*2 = 12 blocks
    # the same sum for v and w (a dropped axis offset would break one of them)
    assert float((g.dense(O.V) * g.dense(O.MV)).sum()) == pytest.approx(3.0, rel=1e-6)
    assert float((g.dense(O.W) * g.dense(O.MW)).sum()) == pytest.approx(2.0, rel=1e-6)


def test_p2g_separates_liquid_and_gas():
    # tests/test_sparse_transfer3d_tp.cpp:31-58
    xs, ys, zs, ts = [], [], [], []
    for k in range(1, 7):
        for j in range(1, 4):
            for i in range(1, 7):
                _seed_cell(i, j, k, 0, xs, ys, zs, ts)
        for j in range(4, 7):
            for i in range(1, 7):
                _seed_cell(i, j, k, 1, xs, ys, zs, ts)
    ph = O.Phase(1.0, 0.01, 1.0, 1.0)
    ph.rho_tilde_0 = O.calibrate_rho_tilde0(ph, 1.0)
    g = O.Grid(8, 8, 8, 1.0)
    g.p2g(O.Particles(xs, ys, zs, type=ts), ph, 1.0)
    assert O.phi(g.get(O.MU, 4, 2, 3), ph) > 0.8
    assert O.phi(g.get(O.MU, 4, 5, 3), ph) < 0.2


def test_calibrate_rho_tilde0_brute_force():
    # transfer3d_tp.cpp:29-36: the raw u-face mass at (4,4,4) of an 8^3 lattice,
    # recomputed here by brute force over the particles that reach that node
    ph = O.Phase(1.0, 0.01, 1.0, 1.0)
    got = O.calibrate_rho_tilde0(ph, 1.0)
    total = 0.0
    for k in range(8):
        for j in range(8):
            for i in range(8):
                for s in range(8):
                    gx = i + 0.25 + 0.5 * (s & 1)
                    gy = j + 0.25 + 0.5 * ((s >> 1) & 1) - 0.5
                    gz = k + 0.25 + 0.5 * ((s >> 2) & 1) - 0.5
                    ws = {}
                    for kk in range(0, 8):
                        for jj in range(0, 8):
                            for ii in range(0, 9):
                                d2 = (gx - ii) ** 2 + (gy - jj) ** 2 + (gz - kk) ** 2
                                t = 1.0 - d2 / 2.25
                                if t > 0:
                                    ws[(ii, jj, kk)] = t ** 3
                    if (4, 4, 4) in ws:
                        total += ws[(4, 4, 4)] / sum(ws.values())
    assert got == pytest.approx(total, rel=1e-12)
    assert 7.0 < got < 9.0  # about 8 particles per cell share each node


def test_g2p_uses_phase_specific_flip_alpha():
    # tests/test_sparse_transfer3d_tp.cpp:60-78
    g = O.Grid(8, 8, 8, 1.0)
    sv = O.Grid(8, 8, 8, 1.0)
    for k in (1, 2):
        for j in (1, 2):
            for i in (2, 3):
                g.set(O.U, i, j, k, 5.0)
                sv.set(O.U, i, j, k, 2.0)
    ps = O.Particles([2.0, 2.0], [2.0, 2.0], [2.0, 2.0], [10.0, 10.0], [0, 0], [0, 0], [0, 1])
    g.g2p(sv, ps, 1.0, 0.0)
    assert ps.u[0] == pytest.approx(13.0)
    assert ps.u[1] == pytest.approx(5.0)


def test_dense_and_sparse_p2g_agree():
    # tests/test_sparse_transfer3d_tp.cpp:80-121
    ph = O.Phase(1.0, 0.01, 1.0, 1.0)
    ps = O.Particles([3.0, 4.25], [2.5, 4.0], [2.5, 3.75], [4.0, -1.0], [3.0, 0.5], [2.0, -2.0],
                     [0, 1])
    U, V, W, MU, MV, MW = O.dense_p2g(8, 8, 8, 1.0, ps, ph)
    g = O.Grid(8, 8, 8, 1.0)
    g.p2g(ps, ph, 1.0)
    for f, D in ((O.U, U), (O.V, V), (O.W, W), (O.MU, MU), (O.MV, MV), (O.MW, MW)):
        assert np.abs(g.dense(f) - D).max() < 1e-6


def test_advect_reports_phase_boundary_clamps():
    # tests/test_sparse_transfer3d_tp.cpp:123-161
    g = O.Grid(8, 8, 8, 1.0)
    for k in range(8):
        for j in range(8):
            for i in range(9):
                g.set(O.U, i, j, k, 10.0)
    for k in range(8):
        for j in range(9):
            for i in range(8):
                g.set(O.V, i, j, k, -10.0)
    ps = O.Particles([7.45, 4.0], [4.0, 0.55], [4.0, 4.0], type=[0, 1])
    es, idx = g.advect(ps, 0.1, 2)
    assert ps.x[0] == pytest.approx(7.5, rel=1e-12)
    assert ps.y[1] == pytest.approx(0.5, rel=1e-12)
    assert (es.clamped_liquid, es.clamped_gas, es.x_hi, es.y_lo) == (1, 1, 1, 1)
    assert list(idx) == [0, 1]


def test_advect_rk3_follows_linear_field():
    # tests/test_sparse_transfer3d_tp.cpp:163-180
    g = O.Grid(8, 8, 8, 1.0)
    for k in range(8):
        for j in range(8):
            for i in range(9):
                g.set(O.U, i, j, k, float(i))
    ps = O.Particles([2.0], [4.0], [4.0], type=[0])
    g.advect(ps, 0.1, 3)
    expected = 2.0 + 0.1 / 6.0 * (2.0 + 4.0 * 2.1 + 2.22)
    assert ps.x[0] == pytest.approx(expected, rel=1e-6)
    assert ps.y[0] == pytest.approx(4.0, rel=1e-12) and ps.z[0] == pytest.approx(4.0, rel=1e-12)
    # RK2 midpoint in u = x: x + dt*(x + dt/2*x) closed form
    ps2 = O.Particles([2.0], [4.0], [4.0], type=[0])
    g.advect(ps2, 0.1, 2)
    assert ps2.x[0] == pytest.approx(2.0 + 0.1 * (2.0 + 0.05 * 2.0), rel=1e-6)


def _vc_box_grid():
    # tests/test_sparse_vc3d.cpp:8-38: walls, an interior 4^3 fluid box, unit masses
    g = O.Grid(8, 8, 8, 1.0)
    for k in range(8):
        for j in range(8):
            for i in range(8):
                m = 0
                if i in (0, 7) or j in (0, 7) or k in (0, 7):
                    m = 2
                elif 2 <= i < 6 and 2 <= j < 6 and 2 <= k < 6:
                    m = 1
                if m:
                    g.set(O.MK, i, j, k, m)
    return g


def _vc_fill(g, with_u):
    for k in range(2, 6):
        for j in range(2, 6):
            for i in range(2, 7):
                g.set(O.MU, i, j, k, 1.0)
                if with_u:
                    g.set(O.U, i, j, k, float(i))
            for i in range(2, 6):
                g.set(O.MV, i, j, k, 1.0)
                g.set(O.MW, i, j, k, 1.0)
    for k in range(2, 7):
        for j in range(2, 6):
            for i in range(2, 6):
                g.set(O.MW, i, j, k, 1.0)


def _box_div(g):
    U, V, W = g.dense(O.U), g.dense(O.V), g.dense(O.W)
    d = ((U[2:6, 2:6, 3:7] - U[2:6, 2:6, 2:6]) + (V[2:6, 3:7, 2:6] - V[2:6, 2:6, 2:6]) +
         (W[3:7, 2:6, 2:6] - W[2:6, 2:6, 2:6]))
    return d


def test_vc_projection_removes_divergence():
    # tests/test_sparse_vc3d.cpp:8-54
    g = _vc_box_grid()
    _vc_fill(g, True)
    st = g.project(O.Phase(1.0, 0.01, 1.0, 1.0), 1.0, 1000, 1e-9)
    assert np.abs(_box_div(g)).max() < 1e-4
    assert len(g.active_blocks(O.P)) == 8
    assert st.pinned == 0 and st.fluid_cells == 64


def test_vc_projection_preserves_target_cdiv():
    # tests/test_sparse_vc3d.cpp:56-100
    g = _vc_box_grid()
    _vc_fill(g, False)
    g.project(O.Phase(1.0, 0.01, 1.0, 1.0), 1.0, 1000, 1e-9, 0.125)
    assert _box_div(g).mean() == pytest.approx(0.125, rel=1e-4)


def test_projection_matches_direct_solve_on_tiny_grid():
    # A dense direct solve of the same linear system pins the PCG's answer.  The
    # matrix is assembled here from the face masses by the operator's definition
    # (sparse_ops3d_tp.cpp:285-313): rows = FLUID cells, beta from Eq. 7.
    n = 6
    dx, x, y, z, t = scene(O.SCENE_BUBBLE, n)
    ph = phase()
    ps = O.Particles(x, y, z, type=t)
    g = O.Grid(n, n, n, dx, faithful=False)
    g.mark_cells(ps)
    g.p2g(ps, ph, 1.0)
    g.apply_gravity(1e-3, -9.81)
    g.apply_walls()
    pre = g.clone()
    st = g.project(ph, 1e-3, 5000, 1e-14, 0.0, O.STOP_REL2)
    mk = pre.dense(O.MK)
    cells = [(i, j, k) for k in range(n) for j in range(n) for i in range(n) if mk[k, j, i] == 1]
    idx = {c: r for r, c in enumerate(cells)}
    N = len(cells)
    A = np.zeros((N, N))
    b = np.zeros(N)
    U, V, W = pre.dense(O.U), pre.dense(O.V), pre.dense(O.W)
    MU, MV, MW = pre.dense(O.MU), pre.dense(O.MV), pre.dense(O.MW)

    def beta(raw):
        return O.beta(O.phi(raw, ph), ph)

    def solid(i, j, k):
        return not (0 <= i < n and 0 <= j < n and 0 <= k < n) or mk[k, j, i] == 2

    scale = 1e-3 / dx ** 2
    for (i, j, k), r in idx.items():
        faces = [((i + 1, j, k), MU[k, j, i + 1], U[k, j, i + 1], +1),
                 ((i - 1, j, k), MU[k, j, i], U[k, j, i], -1),
                 ((i, j + 1, k), MV[k, j + 1, i], V[k, j + 1, i], +1),
                 ((i, j - 1, k), MV[k, j, i], V[k, j, i], -1),
                 ((i, j, k + 1), MW[k + 1, j, i], W[k + 1, j, i], +1),
                 ((i, j, k - 1), MW[k, j, i], W[k, j, i], -1)]
        div = 0.0
        for nb, raw, vel, sgn in faces:
            if solid(*nb):
                continue
            div += sgn * vel
            A[r, r] += scale * beta(raw)
            if nb in idx:
                A[r, idx[nb]] -= scale * beta(raw)
        b[r] = -div / dx
    assert st.pinned == 0  # the bubble scene has AIR headspace: no pin
    p = np.linalg.solve(A, b)
    got = np.array([g.get(O.P, *c) for c in cells])
    assert rel_err(got, p) < 1e-9
    assert np.allclose(A, A.T)


def test_timestep_clamps_by_particle_speed():
    # tests/test_sparse_sim3d_tp.cpp:85-107
    st = O.timestep([10.0, 0.0], [0.0, 3.0], [0.0, 4.0], 1.0, 0.2, True, 0.5, 1e-5)
    assert st.max_particle_speed == pytest.approx(10.0)
    assert st.effective_dt == pytest.approx(0.05) and st.limited == 1
    st = O.timestep([10.0], [0.0], [0.0], 1.0, 0.2, False, 0.5, 1e-5)
    assert st.effective_dt == pytest.approx(0.2) and st.limited == 0


def test_phase_field_curve():
    # tests/test_phasefield.cpp:4-13 (Eq. 7, Fig. 7)
    ph = O.Phase(1.0, 0.01, 1.0, 1.0)
    assert O.phi(ph.rho_g, ph) == pytest.approx(0.0)
    # doctest's Approx(1.0).epsilon(0.02) accepts |a - b| < 0.02 * (1 + max(|a|, |b|))
    f = O.phi(ph.rho_l, ph)
    assert abs(f - 1.0) < 0.02 * (1.0 + max(abs(f), 1.0))
    # Eq. 7 at the liquid density: sqrt(1 - ln(rho_l/rho_g) rho_g / rho_l)
    assert f == pytest.approx(math.sqrt(1.0 - math.log(100.0) * 0.01), rel=1e-15)
    prev = -1.0
    for q in range(21):
        f = O.phi(q * 0.05, ph)
        assert f >= prev - 1e-12 and 0.0 <= f <= 1.0
        prev = f
    rmin = math.log(ph.rho_l / ph.rho_g) * ph.rho_g * ph.rho_tilde_0
    assert O.phi(rmin * 0.99, ph) == pytest.approx(0.0)
    assert O.phi(rmin * 1.01 + 0.05, ph) > 0.0
    # closed form at one interior point and beta endpoints
    rt = 0.5
    assert O.phi(rt, ph) == pytest.approx(math.sqrt((rt - rmin) / (1.0 * 1.0 * 1.0)), rel=1e-15)
    assert O.beta(1.0, ph) == pytest.approx(1.0 / ph.rho_l)
    assert O.beta(0.0, ph) == pytest.approx(1.0 / ph.rho_g)


def test_color8_never_colours_6_neighbours_alike():
    # tests/test_block_color.cpp:19-31 (colour = (bx&1) + 2(by&1) + 4(bz&1))
    def color8(bx, by, bz):
        return (bx & 1) + 2 * (by & 1) + 4 * (bz & 1)
    for bz in range(2):
        for by in range(3):
            for bx in range(4):
                c = color8(bx, by, bz)
                for d in ((1, 0, 0), (0, 1, 0), (0, 0, 1)):
                    assert color8(bx + d[0], by + d[1], bz + d[2]) != c


def test_active_blocks_match_brute_force():
    # preactivateKernel, sparse_ops3d_tp.cpp:23-40: a u-face block is active iff some
    # particle has an in-bounds node of it with d^2 < r^2 = 2.25 (brute force over all nodes)
    n = 12
    dx, x, y, z, t = scene(O.SCENE_BUBBLE, n, seed=3)
    rng = np.random.default_rng(5)
    keep = rng.random(len(x)) < 0.05
    ps = O.Particles(x[keep], y[keep], z[keep], type=t[keep])
    g = O.Grid(n, n, n, dx)
    g.p2g(ps, phase(), 1.0)
    for f, off, dims in ((O.MU, (0.0, 0.5, 0.5), (n + 1, n, n)), (O.MV, (0.5, 0.0, 0.5), (n, n + 1, n)),
                         (O.MW, (0.5, 0.5, 0.0), (n, n, n + 1))):
        want = set()
        gx, gy, gz = ps.x / dx - off[0], ps.y / dx - off[1], ps.z / dx - off[2]
        I, J, K = np.meshgrid(np.arange(dims[0]), np.arange(dims[1]), np.arange(dims[2]), indexing="ij")
        I, J, K = I.ravel(), J.ravel(), K.ravel()
        nbx, nby = (dims[0] + 3) // 4, (dims[1] + 3) // 4
        for p in range(len(ps)):
            d2 = (gx[p] - I) ** 2 + (gy[p] - J) ** 2 + (gz[p] - K) ** 2
            hit = d2 < 2.25
            for b in np.unique((I[hit] // 4) + nbx * ((J[hit] // 4) + nby * (K[hit] // 4))):
                want.add(int(b))
        assert set(int(b) for b in g.active_blocks(f)) == want


def test_partition_of_unity_mass_and_momentum():
    # Sum of raw mass = sum of particle masses and sum of momentum = sum m_p v_p for
    # every axis (normalised weights): a wrong normalisation or dropped node fails.
    n = 10
    dx, x, y, z, t = scene(O.SCENE_RT, n, seed=9)
    rng = np.random.default_rng(1)
    u, v, w = rng.normal(size=(3, len(x)))
    ph = phase()
    ps = O.Particles(x, y, z, u, v, w, t)
    g = O.Grid(n, n, n, dx, faithful=False)
    g.p2g(ps, ph, 1.0)
    mp = np.where(t == 0, ph.rho_l, ph.rho_g)
    for fv, fm, vel in ((O.U, O.MU, u), (O.V, O.MV, v), (O.W, O.MW, w)):
        M = g.dense(fm)
        assert M.sum() == pytest.approx(mp.sum(), rel=1e-12)
        assert (g.dense(fv) * M).sum() == pytest.approx((mp * vel).sum(), rel=1e-9)


def test_mark_cells_by_truncation():
    # markCells, sparse_sim3d_tp.cpp:22-50, brute force
    n = 9
    dx, x, y, z, t = scene(O.SCENE_DAMBREAK, n, seed=2)
    ps = O.Particles(x, y, z, type=t)
    g = O.Grid(n, n, n, dx)
    g.mark_cells(ps)
    want = np.zeros((n, n, n))
    want[[0, -1], :, :] = 2
    want[:, [0, -1], :] = 2
    want[:, :, [0, -1]] = 2
    for p in range(len(x)):
        i, j, k = int(ps.x[p] / dx), int(ps.y[p] / dx), int(ps.z[p] / dx)
        if want[k, j, i] != 2:
            want[k, j, i] = 1
    assert np.array_equal(g.dense(O.MK), want)
