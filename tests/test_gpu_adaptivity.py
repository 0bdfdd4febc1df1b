"""Particle adaptivity on the GPU against the oracle (src/driver/particle_adaptivity3d.h,
applied at the start of SparseSim3DTP::step and MRSim3DTP::step).

Both sides start from the same FP32 particles.  Bars: every integer result -- the
counters, which particles survive, their order, their types, the refilled particles'
count and order -- exact; positions and velocities exact (survivors are copied;
refilled positions and mean velocities are the FP32 roundings of the same FP64
values); volumes within 1e-6 relative (the GPU stores FP32 volumes, so a refill that
follows a coarsening sums rounded volumes).  Through a whole step the particle state
is held to the 1e-4 bar of the steps without adaptivity.
"""
import numpy as np
import pytest

import pfflip_amd as G
import pyoracle as O
from _helpers import bubble_lattice, phase, rel_err, scene, step_cfg

pytestmark = pytest.mark.gpu

VOL_TOL = 1e-6


def _mix(n, seed, nper=6, dx=1.0, origin=(0.0, 0.0, 0.0)):
    """liquid in the lower half and a column, gas above; 0..nper particles per cell,
    FP32 positions"""
    rng = np.random.default_rng(seed)
    pts = []
    for k in range(1, n - 1):
        for j in range(1, n - 1):
            for i in range(1, n - 1):
                liquid = j < n // 2 or (i - n / 2) ** 2 + (k - n / 2) ** 2 < 2
                for _ in range(int(rng.integers(0, nper + 1))):
                    pts.append((i + rng.random(), j + rng.random(), k + rng.random(), 0 if liquid else 1))
    p = np.array(pts)
    x, y, z = ((p[:, q] * dx + origin[q]).astype(np.float32) for q in range(3))
    u, v, w = (rng.normal(size=len(p)).astype(np.float32) for _ in range(3))
    vol = rng.uniform(0.5, 1.5, len(p)).astype(np.float32)
    return x, y, z, u, v, w, p[:, 3].astype(np.uint8), vol


CASES = [
    dict(narrow_band_air=1),
    dict(gas_particle_coarsening=1, gas_particles_per_cell_target=2),
    dict(liquid_particle_coarsening=1, liquid_particles_per_cell_target=3),
    dict(liquid_particle_refill=1),
    dict(narrow_band_air=1, gas_particle_coarsening=1, liquid_particle_coarsening=1, liquid_particle_refill=1),
    dict(narrow_band_air=1, narrow_band_air_radius=0, gas_particle_coarsening=1, liquid_particle_coarsening=1,
         liquid_particles_per_cell_target=5, liquid_particle_refill=1, liquid_refill_particles_per_cell_target=3),
    dict(liquid_particle_refill=1, liquid_particle_refill_interface_only=1, liquid_particle_refill_interface_radius=2,
         liquid_particle_refill_max_added_per_step=50),
    dict(liquid_particle_refill=1, liquid_refill_particles_per_cell_target=9, liquid_particle_refill_max_added_per_step=1),
]


def _compare(ctx, out, ost, gst):
    assert gst.values() == ost.values()
    assert ctx.count() == len(out)
    x, y, z, u, v, w = ctx.download()
    t, vol = ctx.download_ext()
    assert np.array_equal(t, out.type)
    for a, b in ((x, out.x), (y, out.y), (z, out.z), (u, out.u), (v, out.v), (w, out.w)):
        assert np.array_equal(a, b.astype(np.float32))
    assert np.allclose(vol, out.volume, rtol=VOL_TOL, atol=0)


@pytest.mark.parametrize("case", range(len(CASES)))
def test_adaptivity_matches_oracle(case):
    n = 16
    dx, origin = 0.125, (0.5, -1.0, 2.0)
    print(f"[adapt] case={case} n={n}", flush=True)
    x, y, z, u, v, w, t, vol = _mix(n, 20 + case, dx=dx, origin=origin)
    ctx = G.Context(n, n, n, dx, origin)
    ctx.upload(x, y, z, u, v, w, t, vol)
    cfg = G.adapt_cfg(**CASES[case])
    gst = ctx.particle_adaptivity(cfg)
    d = O.PaDomain(n, n, n, 0, dx, *origin)
    out, ost = O.particle_adaptivity(d, O.adapt_cfg(**CASES[case]), O.Particles(x, y, z, u, v, w, type=t, volume=vol))
    assert sum(ost.values()) > 0
    _compare(ctx, out, ost, gst)
    assert ctx.liquid_volume() == pytest.approx(out.volume[out.type == 0].sum(), rel=1e-6)


def test_adaptivity_without_volumes_and_twice_in_a_row():
    # no volumes uploaded (all 1.0); a second pass on the result, with the particle
    # set grown past the first allocation by refill
    n = 12
    x, y, z, u, v, w, t, _ = _mix(n, 5, nper=2)
    ctx = G.Context(n, n, n, 1.0)
    ctx.upload(x, y, z, u, v, w, t)
    ps = O.Particles(x, y, z, u, v, w, type=t)
    d = O.PaDomain(n, n, n, 0, 1.0, 0.0, 0.0, 0.0)
    for cfgd in (dict(liquid_particle_refill=1, liquid_refill_particles_per_cell_target=8),
                 dict(narrow_band_air=1, gas_particle_coarsening=1, liquid_particle_coarsening=1)):
        gst = ctx.particle_adaptivity(G.adapt_cfg(**cfgd))
        ps, ost = O.particle_adaptivity(d, O.adapt_cfg(**cfgd), ps)
        _compare(ctx, ps, ost, gst)
        gx, gy, gz, gu, gv, gw = ctx.download()
        t2, v2 = ctx.download_ext()
        ps = O.Particles(gx, gy, gz, gu, gv, gw, type=t2, volume=v2)


def test_all_stages_off_changes_nothing():
    n = 10
    x, y, z, u, v, w, t, vol = _mix(n, 9)
    ctx = G.Context(n, n, n, 1.0)
    ctx.upload(x, y, z, u, v, w, t, vol)
    st = ctx.particle_adaptivity(G.adapt_cfg())
    assert st.values() == [0] * 21 and ctx.count() == len(x)
    gx = ctx.download()[0]
    assert np.array_equal(gx, x)


def test_step_with_adaptivity_matches_oracle():
    n = 20
    print(f"[adapt-step] n={n}", flush=True)
    dx, x, y, z, t = scene(O.SCENE_BUBBLE, n, seed=2)
    oph = phase()
    acfg = dict(narrow_band_air=1, gas_particle_coarsening=1, liquid_particle_coarsening=1,
                liquid_particle_refill=1, liquid_particles_per_cell_target=6)
    ctx = G.Context(n, n, n, dx)
    ctx.upload(x, y, z, type=t)
    sim = G.SparseSim3DTP(n, n, n, dx)
    sim.ctx = ctx
    sim.phase = G.Phase(oph.rho_l, oph.rho_g, oph.alpha_phi, oph.rho_tilde_0)
    sim.dt = 2e-3
    sim.stop_rule, sim.cg_tol, sim.cg_iters = G.STOP_REL2, 1e-9, 20000
    for k, v in acfg.items():
        setattr(sim, k, v)
    sim.c_div_volume_correction, sim.c_div_strength = True, 0.5
    sim.liquid_volume_target = float((t == 0).sum()) * 1.02
    cfg = step_cfg(oph, dt=2e-3, cg_iters=20000, cg_tol=1e-12, stop_rule=O.STOP_REL2)
    cfg.c_div_volume_correction, cfg.c_div_strength, cfg.liquid_volume_target = 1, 0.5, sim.liquid_volume_target
    go = O.Grid(n, n, n, dx, faithful=False)
    pa = O.Particles(x, y, z, type=t)
    d = O.PaDomain(n, n, n, 0, dx, 0.0, 0.0, 0.0)
    pa, ost = O.particle_adaptivity(d, O.adapt_cfg(**acfg), pa)
    ostep = go.step(pa, cfg)
    st = sim.step()
    # the adaptivity ran on identical particles: exact
    assert st.adapt.values() == ost.values() and sum(ost.values()) > 0
    assert sim.liquid_particle_refill_added_total == ost.refill_added
    assert st.liquid_volume == pytest.approx(ostep.liquid_volume, rel=1e-6)
    assert st.c_div == pytest.approx(ostep.c_div, rel=1e-5, abs=1e-9) and st.c_div != 0.0
    assert st.solve.converged == 1
    gx, gy, gz, gu, gv, gw = ctx.download()
    assert len(gx) == len(pa)
    # positions: the 1e-6 bar of the dt = 1e-3 steps, scaled to this step's dt = 2e-3
    assert max(np.abs(gx - pa.x).max(), np.abs(gy - pa.y).max(), np.abs(gz - pa.z).max()) < 2e-6
    assert rel_err(np.concatenate([gu, gv, gw]), np.concatenate([pa.u, pa.v, pa.w])) < 1e-4


def test_mr_step_with_adaptivity_matches_oracle():
    n, dt = 16, 0.02
    print(f"[adapt-mr-step] n={n}", flush=True)
    parts = bubble_lattice(n, vel=False)
    oph = phase(1.0, 0.01)
    leaves = O.mr_layout_bubble_band(n, n, n)
    ctx = G.MRContext(n, n, n, 1.0)
    ctx.set_layout(leaves)
    ctx.upload(*parts)
    sim = G.MRSim3DTP(n, n, n, 1.0)
    sim.ctx = ctx
    sim.phase = G.Phase(oph.rho_l, oph.rho_g, oph.alpha_phi, oph.rho_tilde_0)
    sim.dt = dt
    sim.dynamic_refinement = False
    sim.cg_iters, sim.cg_tol, sim.cg_rel_tol = 20000, 1e-11, 0.0
    acfg = dict(gas_particle_coarsening=1, liquid_particle_coarsening=1, liquid_particles_per_cell_target=5,
                liquid_particle_refill=1, liquid_refill_particles_per_cell_target=6)
    for k, v in acfg.items():
        setattr(sim, k, v)
    osim = O.mr_sim(n, n, n, 1.0, phase=oph, dt=dt, dynamic=0,
                    solve=O.mr_solve_cfg(max_iterations=20000, abs_tol=1e-12, rel_tol=0.0))
    go = O.MRGrid(n, n, n, 1.0, faithful=False)
    go.set_leaves(leaves)
    x, y, z, u, v, w, t = parts
    pa = O.Particles(x, y, z, u, v, w, type=t)
    d = O.PaDomain(n, n, n, 0, 1.0, 0.0, 0.0, 0.0)
    pa, ost = O.particle_adaptivity(d, O.adapt_cfg(**acfg), pa)
    O.mr_step(go, osim, pa)
    st = sim.step()
    assert st.adapt.values() == ost.values() and ost.liquid_removed > 0 and ost.gas_removed > 0
    assert st.solve.converged == 1
    gx, gy, gz, gu, gv, gw = ctx.download()
    assert len(gx) == len(pa)
    assert max(np.abs(gx - pa.x).max(), np.abs(gy - pa.y).max(), np.abs(gz - pa.z).max()) < 1e-5
    assert rel_err(np.concatenate([gu, gv, gw]), np.concatenate([pa.u, pa.v, pa.w])) < 1e-4
