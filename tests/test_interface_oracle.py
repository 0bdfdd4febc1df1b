"""Interface diagnostics and CSF surface tension in the oracle
(oracle/pf_oracle.c, restating src/driver/interface_diagnostics3d.cpp): pinned by
closed forms (planar interfaces have zero curvature; a linear phase profile has the
finite-difference gradient exactly; a sphere's curvature approaches 2/R), and -- where
oracle/_ref was built -- bit-exact against the reference's own functions and its
whole SparseSim3DTP::step with surface tension on."""
import numpy as np
import pytest

import pyoracle as O
from _helpers import phase, scene, step_cfg

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built")


def _phi_inverse(ph, phi):
    """raw face mass with phiFromRawDensity(raw) = phi (Eq. 7 inverted)"""
    rmin = np.log(ph.rho_l / ph.rho_g) * ph.rho_g * ph.rho_tilde_0
    return rmin + phi * phi * ph.alpha_phi * ph.rho_tilde_0 * ph.rho_l


def _field_grid(n, ph, phi_of, fluid=lambda i, j, k: True):
    """a grid whose face masses give phi_of(x, y, z) at each face centre; FLUID
    marker where fluid(i, j, k), SOLID walls"""
    g = O.Grid(n, n, n, 1.0, faithful=False)
    for f, off in ((O.MU, (0.0, 0.5, 0.5)), (O.MV, (0.5, 0.0, 0.5)), (O.MW, (0.5, 0.5, 0.0))):
        a = f - O.MU
        for k in range(n + (a == 2)):
            for j in range(n + (a == 1)):
                for i in range(n + (a == 0)):
                    g.set(f, i, j, k, _phi_inverse(ph, phi_of(i + off[0], j + off[1], k + off[2])))
    for k in range(n):
        for j in range(n):
            for i in range(n):
                wall = min(i, j, k) == 0 or max(i, j, k) == n - 1
                g.set(O.MK, i, j, k, 2 if wall else (1 if fluid(i, j, k) else 0))
    return g


def test_planar_interface_has_zero_curvature_and_no_csf_force():
    ph = phase()
    n = 12
    g = _field_grid(n, ph, lambda x, y, z: min(1.0, max(0.0, 0.1 * (y - 1.0))))
    phi, grad, kap = g.interface_fields(ph)
    inner = (slice(2, n - 2),) * 3
    assert np.all(kap[inner] == 0.0)
    st = g.interface_diagnostics(ph)
    assert st.interface_cells > 0 and st.curvature_abs_max == 0.0 and st.finite == 1
    u0 = [g.dense(f).copy() for f in range(3)]
    s2 = g.surface_tension(ph, 0.01, strength=0.5, max_delta_speed=0.0, radius=1)
    assert s2.applied_cells == st.interface_cells and s2.max_delta_speed == 0.0
    for f in range(3):
        assert np.array_equal(g.dense(f), u0[f])


def test_linear_profile_gradient_is_the_central_difference():
    # phi = 0.05 y + 0.02 z (faces carry exact phi): the cell phi is the mean of six
    # face values, so the central difference of cell phi gives the slopes
    ph = phase()
    n = 10
    g = _field_grid(n, ph, lambda x, y, z: 0.05 * y + 0.02 * z)
    phi, grad, kap = g.interface_fields(ph)
    want = np.hypot(0.05, 0.02)
    assert np.allclose(grad[2:-2, 2:-2, 2:-2], want, rtol=1e-9)
    assert np.allclose(phi[3, 4, 5], 0.05 * 4.5 + 0.02 * 3.5, rtol=1e-9)


def test_sphere_curvature_approaches_two_over_radius():
    ph = phase()
    n, R = 32, 9.0
    c = n / 2

    def prof(x, y, z):
        r = np.sqrt((x - c) ** 2 + (y - c) ** 2 + (z - c) ** 2)
        return float(np.clip(0.5 - 0.25 * (r - R), 0.0, 1.0))

    g = _field_grid(n, ph, prof)
    phi, grad, kap = g.interface_fields(ph)
    band = (phi > 0.3) & (phi < 0.7)
    # div(grad phi / |grad phi|) of a field decreasing outwards: -2/r at radius r
    assert abs(np.median(kap[band]) + 2.0 / R) < 0.05 * 2.0 / R


def _ref_case(n=16, seed=4, kind=O.SCENE_BUBBLE):
    dx, x, y, z, t = scene(kind, n, seed=seed, ny=n)
    ph = phase()
    cfg = step_cfg(ph, dt=2e-3)
    ps = O.Particles(x, y, z, type=t)
    rng = np.random.default_rng(seed)
    for a in (ps.u, ps.v, ps.w):
        a[:] = rng.normal(scale=0.2, size=len(ps))
    g = O.Grid(n, n, n, dx, faithful=True)
    rs = O.RefSim(n, n, n, dx)
    rs.config(cfg)
    rs.set_particles(ps)
    g.mark_cells(ps)
    g.p2g(ps, ph, 1.0)
    mk = g.dense(O.MK)
    for k, j, i in zip(*np.nonzero(mk)):
        rs.set(O.MK, int(i), int(j), int(k), float(mk[k, j, i]))
    rs.p2g()
    return g, rs, ph, cfg, ps


@needs_ref
@pytest.mark.parametrize("kind", [O.SCENE_BUBBLE, O.SCENE_RT])
def test_interface_diagnostics_bit_exact(kind):
    g, rs, ph, cfg, ps = _ref_case(kind=kind)
    for f in range(6):
        assert np.array_equal(g.dense(f), rs.dense(f))
    st = g.interface_diagnostics(ph)
    assert st.interface_cells > 0
    assert st.values() == list(rs.interface_diagnostics())


@needs_ref
@pytest.mark.parametrize("strength,maxd,radius", [(0.02, 0.25, 0), (0.5, 0.01, 1), (2.0, 0.0, 2), (0.1, 0.2, 3)])
def test_surface_tension_bit_exact(strength, maxd, radius):
    g, rs, ph, cfg, ps = _ref_case()
    st = g.surface_tension(ph, 0.01, strength, maxd, radius)
    out = rs.surface_tension(0.01, strength, maxd, radius)
    assert st.applied_cells > 0
    assert st.values() == list(out)
    for f in range(3):
        assert np.array_equal(g.dense(f), rs.dense(f)), f


@needs_ref
def test_step_with_surface_tension_bit_exact():
    # SparseSim3DTP::step with surface_tension on (sparse_sim3d_tp.cpp:865-873):
    # diagnostics after P2G, CSF after gravity, then the rest of the step
    dx, x, y, z, t = scene(O.SCENE_BUBBLE, 12, seed=2, ny=12)
    ph = phase()
    cfg = step_cfg(ph, dt=2e-3)
    cfg.surface_tension, cfg.surface_tension_strength = 1, 0.3
    cfg.surface_tension_max_delta_speed, cfg.surface_tension_smoothing_radius = 0.1, 1
    ps = O.Particles(x, y, z, type=t)
    g = O.Grid(12, 12, 12, dx, faithful=True)
    rs = O.RefSim(12, 12, 12, dx)
    rs.config(cfg)
    rs.set_particles(ps)
    pa = ps.copy()
    for _ in range(2):
        st = g.step(pa, cfg)
        rs.step()
        di, sr = rs.last_interface()
        assert st.iface.values() == list(di)
        assert st.surface.values() == list(sr) and st.surface.applied_cells > 0
        pr = rs.get_particles(pa)
        for a in ("x", "y", "z", "u", "v", "w"):
            assert np.array_equal(getattr(pa, a), getattr(pr, a)), a


# ---- multires variants (diagnoseMRInterface3D, applyMRSurfaceTension3D)

# fine boxes on a level-1 base for the reference's bubble tank at 32^3: its own
# interface band (FLUID leaves all fine), liquid below y = 8 coarse, and the water
# surface and bubble in coarse cells
MR_BOXES = {"band": None, "mid": [0, 8, 0, 32, 20, 32], "low": [0, 0, 0, 32, 8, 32]}


def _mr_case(n=32, seed=3, box="band"):
    """a bubble-tank scene on a two-level layout after P2G on the reference build
    and the oracle, markers shared"""
    r = O.RefMR(n, n, n, 1.0)
    r.init_bubble()
    if MR_BOXES[box] is not None:
        r.set_leaves(O.mr_layout_fine_box(n, n, n, MR_BOXES[box]))
    ph = phase(1.0, 0.01)
    r.config(dt=0.02, phase=ph)
    ps = r.particles()
    rng = np.random.default_rng(seed)
    for a in (ps.u, ps.v, ps.w):
        a[:] = rng.normal(scale=0.5, size=len(ps))
    r.set_particles(ps)
    g = O.MRGrid(n, n, n, 1.0, faithful=True)
    g.set_leaves(r.leaves())
    r.p2g()
    g.p2g(ps, ph, 1.0)
    g.mark_cells(ps)
    mk = g.dense(O.MK)
    for k, j, i in zip(*np.nonzero(mk)):
        r.set(O.MK, int(i), int(j), int(k), float(mk[k, j, i]))
    return g, r, ph, ps


def test_mr_on_a_uniform_fine_layout_is_the_sparse_diagnostic():
    # with every leaf at level 0 the multires grid is the sparse grid: the same cells
    # are sampled at the same points, only the summation order differs
    n = 16
    dx, x, y, z, t = scene(O.SCENE_BUBBLE, n, seed=6, ny=n)
    ph = phase()
    ps = O.Particles(np.asarray(x) / dx, np.asarray(y) / dx, np.asarray(z) / dx, type=t)
    g = O.Grid(n, n, n, 1.0, faithful=True)
    g.mark_cells(ps)
    g.p2g(ps, ph, 1.0)
    m = O.MRGrid(n, n, n, 1.0, faithful=True)
    m.set_leaves(O.mr_layout_fine_box(n, n, n, [0, 0, 0, n, n, n]))
    assert set(m.leaves()[:, 0]) == {0}
    for f in (O.MU, O.MV, O.MW, O.U, O.V, O.W):
        a = g.dense(f)
        for k, j, i in zip(*np.nonzero(a)):
            m.set(f, int(i), int(j), int(k), a[k, j, i])
    mk = g.dense(O.MK)
    for k, j, i in zip(*np.nonzero(mk)):
        m.set(O.MK, int(i), int(j), int(k), mk[k, j, i])
    s, q = g.interface_diagnostics(ph), m.interface_diagnostics(ph)
    assert s.interface_cells > 0
    for f in ("sample_cells", "interface_cells", "finite", "phi_min", "phi_max", "grad_max", "curvature_abs_max"):
        assert getattr(s, f) == getattr(q, f), f
    for f in ("phi_mean", "grad_mean", "curvature_abs_mean"):
        assert getattr(q, f) == pytest.approx(getattr(s, f), rel=1e-12)
    a, b = g.surface_tension(ph, 0.01, 0.5, 0.1, 1), m.surface_tension(ph, 0.01, 0.5, 0.1, 1)
    assert a.applied_cells == b.applied_cells > 0 and a.max_delta_speed == b.max_delta_speed
    for f in (O.U, O.V, O.W):
        assert np.array_equal(g.dense(f), m.dense(f)), f


@needs_ref
@pytest.mark.parametrize("box", ["mid", "low"])
def test_mr_samples_each_fluid_leaf_cell_once(box):
    g, r, ph, ps = _mr_case(box=box)
    lv = g.leaves()
    assert set(lv[:, 0]) == {0, 1}
    # FLUID leaf cells: markers are stored per leaf cell, so count distinct leaf cells
    mk = g.dense(O.MK)
    n = g.nx
    level = np.full((n, n, n), -1)
    for L, bx, by, bz in lv:
        s = 4 << L
        level[bz * s:(bz + 1) * s, by * s:(by + 1) * s, bx * s:(bx + 1) * s] = L
    kk, jj, ii = np.nonzero(mk == 1)
    leaf_cells = {(int(i) >> level[k, j, i], int(j) >> level[k, j, i], int(k) >> level[k, j, i], level[k, j, i])
                  for k, j, i in zip(kk, jj, ii)}
    st = g.interface_diagnostics(ph)
    assert st.finite == 1 and st.sample_cells == len(leaf_cells)
    assert any(c[3] == 1 for c in leaf_cells) and any(c[3] == 0 for c in leaf_cells)


@needs_ref
@pytest.mark.parametrize("box", sorted(MR_BOXES))
def test_mr_interface_diagnostics_bit_exact(box):
    g, r, ph, ps = _mr_case(box=box)
    for f in range(6):
        assert np.array_equal(g.dense(f), r.dense(f))
    st = g.interface_diagnostics(ph)
    assert st.interface_cells > 0
    assert st.values() == list(r.interface_diagnostics())


@needs_ref
@pytest.mark.parametrize("box,strength,maxd,radius", [("band", 0.02, 0.25, 0), ("low", 0.5, 0.01, 1),
                                                      ("mid", 2.0, 0.0, 2), ("low", 0.1, 0.2, 3)])
def test_mr_surface_tension_bit_exact(box, strength, maxd, radius):
    g, r, ph, ps = _mr_case(box=box)
    st = g.surface_tension(ph, 0.01, strength, maxd, radius)
    out = r.surface_tension(0.01, strength, maxd, radius)
    assert st.applied_cells > 0
    assert st.values() == list(out)
    for f in range(3):
        assert np.array_equal(g.dense(f), r.dense(f)), f


@needs_ref
def test_mr_step_with_surface_tension_bit_exact():
    n = 32
    r = O.RefMR(n, n, n, 1.0)
    r.init_bubble()
    # a fixed layout with the water surface in coarse cells (no dynamic refinement)
    r.set_leaves(O.mr_layout_fine_box(n, n, n, MR_BOXES["low"]))
    ph = phase(1.0, 0.01)
    r.config(dt=0.02, phase=ph, dynamic=0)
    r.surface_flags(1, 0.3, 0.1, 1)
    ps = r.particles()
    g = O.MRGrid(n, n, n, 1.0, faithful=True)
    g.set_leaves(r.leaves())
    sim = O.mr_sim(n, n, n, 1.0, phase=ph, dynamic=0)
    sim.surface_tension, sim.surface_tension_strength = 1, 0.3
    sim.surface_tension_max_delta_speed, sim.surface_tension_smoothing_radius = 0.1, 1
    for _ in range(3):
        r.step()
        st = O.mr_step(g, sim, ps)
        di, sr = r.last_interface()
        assert st.iface.values() == list(di)
        assert st.surface.values() == list(sr) and st.surface.applied_cells > 0
        rp = r.particles()
        for a in ("x", "y", "z", "u", "v", "w"):
            assert np.array_equal(getattr(ps, a), getattr(rp, a)), a
    # surface tension off: SurfaceTensionStats3D() defaults in the step's stats
    sim.surface_tension = 0
    r.surface_flags(0)
    r.step()
    st = O.mr_step(g, sim, ps)
    assert st.surface.values() == list(r.last_interface()[1])
