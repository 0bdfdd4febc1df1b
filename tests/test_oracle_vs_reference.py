"""The oracle against the reference itself, compiled from its own sources into
oracle/_ref (skipped where that checker was not built).  In float-faithful mode the
oracle must reproduce every field, the first-touch order of every active-block
list, the particle state and the clamp counters of SparseSim3DTP::step bit for bit.
"""
import numpy as np
import pytest

import pyoracle as O
from _helpers import phase, scene, step_cfg

pytestmark = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built")


def _compare(g, rs, pa, pr):
    for f in range(8):
        assert np.array_equal(g.dense(f), rs.dense(f)), f"field {f}"
        assert np.array_equal(g.active_blocks(f), rs.active_blocks(f)), f"active order {f}"
    for a in ("x", "y", "z", "u", "v", "w"):
        assert np.array_equal(getattr(pa, a), getattr(pr, a)), a


@pytest.mark.parametrize("kind,n,ny,steps", [(O.SCENE_RT, 12, 16, 3), (O.SCENE_BUBBLE, 12, 12, 3),
                                             (O.SCENE_DAMBREAK, 10, 12, 2)])
def test_step_bit_exact_on_jittered_scenes(kind, n, ny, steps):
    dx, x, y, z, t = scene(kind, n, seed=5, ny=ny)
    ph = phase()
    cfg = step_cfg(ph, dt=2e-3, cg_iters=400, cg_tol=1e-7)
    ps = O.Particles(x, y, z, type=t)
    g = O.Grid(n, ny, n, dx, faithful=True)
    rs = O.RefSim(n, ny, n, dx)
    rs.config(cfg)
    rs.set_particles(ps)
    pa = ps.copy()
    for _ in range(steps):
        st = g.step(pa, cfg)
        out = rs.step()
        pr = rs.get_particles(pa)
        _compare(g, rs, pa, pr)
        es = st.escape
        assert [es.clamped_liquid, es.clamped_gas, es.x_lo, es.x_hi, es.y_lo, es.y_hi, es.z_lo,
                es.z_hi] == [int(v) for v in out[4:12]]
        assert st.ts.effective_dt == out[0] and st.ts.max_particle_speed == out[1]


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_step_bit_exact_on_reference_scenes(kind):
    # the reference's own initialisers and defaults (8x12x8, dt 0.03 as in
    # tests/test_sparse_sim3d_tp.cpp:110-114 and :749-753), RK3 on one of them
    rs = O.RefSim(8, 12, 8, 1.0)
    ps, ph = rs.init_scene(kind)
    cfg = step_cfg(ph, dt=0.03, cg_iters=600, cg_tol=1e-7, order=3 if kind == 2 else 2)
    rs.config(cfg)
    g = O.Grid(8, 12, 8, 1.0, faithful=True)
    pa = ps.copy()
    for _ in range(4):
        g.step(pa, cfg)
        rs.step()
        _compare(g, rs, pa, rs.get_particles(pa))


def test_adaptive_timestep_and_cdiv_bit_exact():
    # tests/test_sparse_sim3d_tp.cpp:150-185: CFL-limited dt and c_div from the
    # liquid-volume error, through the whole step
    rs = O.RefSim(8, 12, 8, 1.0)
    ps, ph = rs.init_scene(1)
    ps.u[0] = 100.0
    cfg = step_cfg(ph, dt=0.2)
    cfg.adaptive_timestep = 1
    cfg.c_div_volume_correction = 1
    cfg.liquid_volume_target = float((ps.type == 0).sum()) + 1.0
    rs.config(cfg)
    rs.set_particles(ps)
    g = O.Grid(8, 12, 8, 1.0, faithful=True)
    pa = ps.copy()
    st = g.step(pa, cfg)
    out = rs.step()
    assert st.ts.effective_dt == out[0] == pytest.approx(0.005)
    assert st.c_div == out[2] == pytest.approx(1.0 / (0.005 * cfg.liquid_volume_target))
    _compare(g, rs, pa, rs.get_particles(pa))


def test_helpers_bit_exact():
    for ph in (O.Phase(1.0, 0.01, 1.0, 1.0), O.Phase(1000.0, 1.0, 1.3, 2.0)):
        a = ph.as_array()
        assert O.calibrate_rho_tilde0(ph, 1.0) == O.reflib().ref_calibrate_rho_tilde0(a, 1.0)
        for raw in np.linspace(0.0, 3000.0, 97):
            assert O.phi(raw, ph) == O.reflib().ref_phi(raw, a)
        for p in np.linspace(0.0, 1.0, 11):
            assert O.beta(p, ph) == O.reflib().ref_beta(p, a)
