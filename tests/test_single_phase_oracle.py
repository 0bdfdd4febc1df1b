"""The single-phase sparse path of the oracle (SparseSim3D, src/driver/sparse_sim3d.cpp,
src/driver/sparse_ops3d.cpp) pinned by closed forms, a dense direct solve, the
reference's own dam-break test, and -- where oracle/_ref was built -- bit-exactness
against the reference compiled from its sources."""
import numpy as np
import pytest

import pyoracle as O
from _helpers import scene

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built")


def test_trilinear_single_particle_known_weights():
    # spP2G3D (sparse_ops3d.cpp:42-68): particle at (3.25, 2.5, 2.75), dx = 1.  The
    # u-face stencil starts at floor((3.25, 2.0, 2.25)) = (3, 2, 2) with fractions
    # (0.25, 0, 0.25): weights 0.75/0.25 along x and z, 1/0 along y.
    g = O.Grid(8, 8, 8, 1.0, faithful=False)
    ps = O.Particles([3.25], [2.5], [2.75], u=[2.0], v=[-1.0], w=[0.5])
    g.p2g_single(ps)
    mu = g.dense(O.MU)
    want = {(3, 2, 2): 0.5625, (4, 2, 2): 0.1875, (3, 2, 3): 0.1875, (4, 2, 3): 0.0625}
    nz = {tuple(int(a) for a in idx[::-1]): float(mu[idx]) for idx in zip(*np.nonzero(mu))}
    assert nz == want
    # normalised velocity equals the particle's at every node it reached
    U = g.dense(O.U)
    for (i, j, k) in want:
        assert U[k, j, i] == 2.0
    # v-face: floor((2.75, 2.5, 2.25)) = (2, 2, 2), fractions (0.75, 0.5, 0.25)
    mv = g.dense(O.MV)
    assert mv[2, 2, 2] == 0.25 * 0.5 * 0.75 and mv[3, 3, 3] == 0.75 * 0.5 * 0.25
    assert np.isclose(mv.sum(), 1.0, rtol=0, atol=1e-15)
    # zero-weight stencil nodes still activate their blocks (preactivateField, :22-40):
    # the u-face stencil reaches j = 3, whose weight is 0, inside block (0, 0, 0)
    assert 0 in set(g.active_blocks(O.MU))


def test_trilinear_partition_of_unity():
    # interior particles: the 8 trilinear weights sum to 1, so the raw mass of each
    # face field is the particle count and the momentum is the velocity sum
    rng = np.random.default_rng(3)
    n = 400
    x, y, z = (rng.uniform(2.0, 10.0, n) for _ in range(3))
    u, v, w = (rng.normal(size=n) for _ in range(3))
    g = O.Grid(12, 12, 12, 1.0, faithful=False)
    g.p2g_single(O.Particles(x, y, z, u, v, w))
    for vf, mf, vel in ((O.U, O.MU, u), (O.V, O.MV, v), (O.W, O.MW, w)):
        m = g.dense(mf)
        assert np.isclose(m.sum(), n, rtol=1e-12)
        assert np.isclose((g.dense(vf) * m).sum(), vel.sum(), rtol=1e-9, atol=1e-9)


def test_trilinear_active_sets_brute_force():
    # a face block is active iff some particle's 2x2x2 stencil has an in-bounds node
    # in it; brute force over every node of every particle
    rng = np.random.default_rng(8)
    n = 60
    nx, ny, nz = 12, 9, 10
    x = rng.uniform(0.0, nx, n); y = rng.uniform(0.0, ny, n); z = rng.uniform(0.0, nz, n)
    g = O.Grid(nx, ny, nz, 1.0, faithful=False)
    g.p2g_single(O.Particles(x, y, z))
    shifts = {O.U: (0.0, -0.5, -0.5), O.V: (-0.5, 0.0, -0.5), O.W: (-0.5, -0.5, 0.0)}
    for f, (sx, sy, sz) in shifts.items():
        fx, fy, fz = g.dims(f)
        nbx, nby = (fx + 3) // 4, (fy + 3) // 4
        want = set()
        for p in range(n):
            i0, j0, k0 = int(np.floor(x[p] + sx)), int(np.floor(y[p] + sy)), int(np.floor(z[p] + sz))
            for dk in (0, 1):
                for dj in (0, 1):
                    for di in (0, 1):
                        i, j, k = i0 + di, j0 + dj, k0 + dk
                        if 0 <= i < fx and 0 <= j < fy and 0 <= k < fz:
                            want.add(i // 4 + nbx * (j // 4 + nby * (k // 4)))
        assert set(g.active_blocks(f)) == want


def _box_grid(n=8):
    # walls SOLID, an interior fluid region with AIR above it (free surface)
    g = O.Grid(n, n, n, 1.0, faithful=False)
    for k in range(n):
        for j in range(n):
            for i in range(n):
                wall = i in (0, n - 1) or j in (0, n - 1) or k in (0, n - 1)
                if wall:
                    g.set(O.MK, i, j, k, 2)
                elif j <= 4:
                    g.set(O.MK, i, j, k, 1)
    return g


def test_single_phase_projection_matches_direct_solve():
    # spProjectStep3D (sparse_ops3d.cpp:170-319): unit coefficients, SOLID faces
    # dropped, AIR neighbours Dirichlet p = 0, no pin.  Assemble the same system
    # densely here and solve it directly.
    n = 8
    g = _box_grid(n)
    rng = np.random.default_rng(4)
    for f in (O.U, O.V, O.W):
        fx, fy, fz = g.dims(f)
        for k in range(fz):
            for j in range(fy):
                for i in range(fx):
                    g.set(f, i, j, k, rng.normal())
    U, V, W = g.dense(O.U), g.dense(O.V), g.dense(O.W)
    mk = g.dense(O.MK)
    fluid = [(i, j, k) for k in range(n) for j in range(n) for i in range(n) if mk[k, j, i] == 1]
    idx = {c: t for t, c in enumerate(fluid)}
    N = len(fluid)
    dt = 0.1
    A = np.zeros((N, N))
    b = np.zeros(N)
    for t, (i, j, k) in enumerate(fluid):
        b[t] = -((U[k, j, i + 1] - U[k, j, i]) + (V[k, j + 1, i] - V[k, j, i]) + (W[k + 1, j, i] - W[k, j, i]))
        for di, dj, dk in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
            q = (i + di, j + dj, k + dk)
            m = mk[q[2], q[1], q[0]]
            if m == 2:
                continue
            A[t, t] += dt
            if m == 1:
                A[t, idx[q]] -= dt
    p = np.linalg.solve(A, b)
    st = g.project_single(dt, iters=2000, tol=1e-13, stop_rule=O.STOP_REL2)
    assert st.converged == 1 and st.pinned == 0
    P = g.dense(O.P)
    got = np.array([P[k, j, i] for (i, j, k) in fluid])
    assert np.allclose(got, p, rtol=1e-10, atol=1e-10)


def test_reference_dambreak_test_on_the_oracle():
    # tests/test_sparse_sim3d.cpp:6-36: 16^3 dam-break, 20 steps: stable, count kept,
    # the column falls and the grid stays sparse
    n = 16
    xs, ys, zs = [], [], []
    for k in range(1, n - 1):
        for j in range(1, n * 7 // 10):
            for i in range(1, n * 4 // 10):
                for s in range(8):
                    xs.append(i + 0.25 + 0.5 * (s & 1))
                    ys.append(j + 0.25 + 0.5 * ((s >> 1) & 1))
                    zs.append(k + 0.25 + 0.5 * ((s >> 2) & 1))
    ps = O.Particles(xs, ys, zs)
    g = O.Grid(n, n, n, 1.0, faithful=True)
    my0 = ps.y.mean()
    max_active = 0
    for _ in range(20):
        g.step_single(ps, 0.05, 0.95, -9.81, 300, 1e-6)
        max_active = max(max_active, len(g.active_blocks(O.P)))
    assert np.isfinite(ps.x).all() and np.isfinite(ps.y).all() and np.isfinite(ps.z).all()
    assert ps.y.mean() < my0
    assert 0 < max_active < 4 * 4 * 4


def _compare1(g, r, pa):
    pr = r.get_particles()
    for f in range(8):
        assert np.array_equal(g.dense(f), r.dense(f)), f"field {f}"
        assert np.array_equal(g.active_blocks(f), r.active_blocks(f)), f"active order {f}"
    for a in ("x", "y", "z", "u", "v", "w"):
        assert np.array_equal(getattr(pa, a), getattr(pr, a)), a


@needs_ref
def test_single_phase_step_bit_exact_reference_scene():
    # the reference's own initDamBreak and defaults (dt 0.05, alpha 0.95, cg 300 / 1e-6)
    r = O.RefSim1(16, 16, 16, 1.0)
    r.config(0.05, 0.95, -9.81, 300, 1e-6)
    ps = r.init_dambreak()
    g = O.Grid(16, 16, 16, 1.0, faithful=True)
    for _ in range(6):
        r.step()
        g.step_single(ps, 0.05, 0.95, -9.81, 300, 1e-6)
        _compare1(g, r, ps)


@needs_ref
def test_single_phase_step_bit_exact_jittered_with_velocities():
    # jittered dam-break column with random initial velocities (non-trivial momentum
    # sums from the first splat), dx = 1/n
    n, ny = 12, 16
    dx, x, y, z, t = scene(O.SCENE_DAMBREAK, n, seed=9, ny=ny)
    rng = np.random.default_rng(2)
    u, v, w = (rng.normal(scale=0.3, size=len(x)) for _ in range(3))
    ps = O.Particles(x, y, z, u, v, w)
    r = O.RefSim1(n, ny, n, dx)
    r.config(2e-3, 0.9, -9.81, 400, 1e-7)
    r.set_particles(ps)
    g = O.Grid(n, ny, n, dx, faithful=True)
    pa = ps.copy()
    for _ in range(3):
        r.step()
        g.step_single(pa, 2e-3, 0.9, -9.81, 400, 1e-7)
        _compare1(g, r, pa)
