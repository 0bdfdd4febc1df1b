"""Particle adaptivity in the oracle (oracle/pf_oracle_adapt.c, restating
src/driver/particle_adaptivity3d.h): pinned by the hash's invertibility, closed-form
selections and conservation laws, and -- where oracle/_ref was built -- bit-exact
against the reference's own pa3d functions and through SparseSim3DTP::step and
MRSim3DTP::step with the adaptivity flags on."""
import numpy as np
import pytest

import pyoracle as O
from _helpers import phase, scene, step_cfg

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built")

M32 = (1 << 32) - 1


def _unmix32(h):
    """the inverse of mix32: undo each xorshift and multiply by the modular inverse"""
    def unxorshift(x, s):
        y = x
        for _ in range(32 // s + 1):
            y = x ^ (y >> s)
        return y & M32
    h = unxorshift(h, 16)
    h = (h * pow(0x846CA68B, -1, 1 << 32)) & M32
    h = unxorshift(h, 15)
    h = (h * pow(0x7FEB352D, -1, 1 << 32)) & M32
    return unxorshift(h, 16)


def test_mix32_is_the_invertible_lowbias_hash():
    rng = np.random.default_rng(0)
    xs = [0, 1, 2, 0xFFFFFFFF, 0x9E3779B9] + [int(v) for v in rng.integers(0, 1 << 32, 200, dtype=np.uint64)]
    outs = [O.lib().orc_mix32(x) for x in xs]
    assert all(_unmix32(h) == x for h, x in zip(outs, xs))
    assert O.lib().orc_mix32(0) == 0 and len(set(outs)) == len(xs)


def _domain(n, dx=1.0, origin=(0.0, 0.0, 0.0)):
    return O.PaDomain(n, n, n, 0, dx, *origin)


def _lattice(cells, per_cell, rng, types):
    """per_cell jittered particles in each listed cell (unit cells), types per cell"""
    pts = []
    for (i, j, k), t in zip(cells, types):
        for _ in range(per_cell):
            pts.append((i + rng.random(), j + rng.random(), k + rng.random(), t))
    p = np.array(pts)
    rng.shuffle(p)
    u, v, w = (rng.normal(size=len(p)) for _ in range(3))
    return O.Particles(p[:, 0], p[:, 1], p[:, 2], u, v, w, type=p[:, 3].astype(np.uint8))


def test_narrow_band_keeps_gas_within_the_radius():
    # liquid fills the layer j = 0..1; gas in columns at j = 2..9: with radius 2 the
    # gas cells j <= 3 stay, the rest go
    n = 10
    rng = np.random.default_rng(1)
    cells, types = [], []
    for i in range(n):
        for k in range(n):
            for j in range(n):
                cells.append((i, j, k))
                types.append(0 if j < 2 else 1)
    ps = _lattice(cells, 2, rng, types)
    out, st = O.particle_adaptivity(_domain(n), O.adapt_cfg(narrow_band_air=1, narrow_band_air_radius=2), ps)
    gas = ps.type == 1
    assert st.narrow_band_air_liquid_cells == 2 * n * n
    assert st.narrow_band_air_gas_before == gas.sum()
    keep = (ps.type == 0) | (np.floor(ps.y) <= 3)
    assert st.narrow_band_air_gas_after == (gas & keep).sum() and st.narrow_band_air_removed == (~keep).sum()
    # survivors keep their order (eraseIf)
    assert np.array_equal(out.x, ps.x[keep]) and np.array_equal(out.type, ps.type[keep])
    # no liquid at all: nothing happens, nothing is counted
    none, st2 = O.particle_adaptivity(_domain(n), O.adapt_cfg(narrow_band_air=1), O.Particles(
        ps.x[gas], ps.y[gas], ps.z[gas], type=ps.type[gas]))
    assert len(none) == gas.sum() and st2.values() == [0] * 21


@pytest.mark.parametrize("ptype,target", [(1, 4), (0, 3), (1, 1)])
def test_coarsening_keeps_the_lowest_scores_and_conserves_volume(ptype, target):
    n = 6
    rng = np.random.default_rng(2 + target)
    cells = [(i, j, k) for i in range(n) for j in range(n) for k in range(n)]
    counts = rng.integers(1, 9, len(cells))
    pts = []
    for (i, j, k), c in zip(cells, counts):
        for _ in range(c):
            pts.append((i + rng.random(), j + rng.random(), k + rng.random(), rng.integers(0, 2)))
    p = np.array(pts)
    vol = rng.uniform(0.5, 2.0, len(p))
    ps = O.Particles(p[:, 0], p[:, 1], p[:, 2], type=p[:, 3].astype(np.uint8), volume=vol)
    d = _domain(n)
    key = "gas" if ptype == 1 else "liquid"
    seed = 0x1234567
    cfg = O.adapt_cfg(**{f"{key}_particle_coarsening": 1, f"{key}_particles_per_cell_target": target,
                         f"{key}_particle_coarsening_seed": seed})
    out, st = O.particle_adaptivity(d, cfg, ps)
    # independent selection: per cell sort by (score, index), keep the first target
    ci = np.floor(ps.x).astype(int), np.floor(ps.y).astype(int), np.floor(ps.z).astype(int)
    cell = ci[0] + n * (ci[1] + n * ci[2])
    keep = np.ones(len(ps), bool)
    want_vol = ps.volume.copy()
    groups = {}
    for q in np.nonzero(ps.type == ptype)[0]:
        groups.setdefault(cell[q], []).append(q)
    for c, members in groups.items():
        sc = [(O.lib().orc_particle_score(O.C.byref(d), ps.x[q], ps.y[q], ps.z[q], ci[0][q], ci[1][q], ci[2][q],
                                          seed), q) for q in members]
        sc.sort()
        kept = [q for _, q in sc[:target]]
        for q in members:
            keep[q] = q in kept
        for q in kept:
            want_vol[q] = sum(ps.volume[m] for _, m in sc) / len(kept)
    assert getattr(st, f"{key}_cells") == len(groups)
    assert getattr(st, f"{key}_after") == sum(min(len(m), target) for m in groups.values())
    assert getattr(st, f"{key}_overfull_cells") == sum(len(m) > target for m in groups.values())
    assert np.array_equal(out.x, ps.x[keep]) and np.allclose(out.volume, want_vol[keep], rtol=1e-15)
    # the type's volume per cell is conserved; the other type is untouched
    other = ps.type != ptype
    assert np.array_equal(out.volume[out.type != ptype], ps.volume[other])
    for c, members in groups.items():
        oc = np.floor(out.x).astype(int) + n * (np.floor(out.y).astype(int) + n * np.floor(out.z).astype(int))
        assert np.isclose(out.volume[(oc == c) & (out.type == ptype)].sum(), ps.volume[members].sum(), rtol=1e-13)


def test_refill_fills_underfull_cells_at_the_slots():
    n = 5
    rng = np.random.default_rng(7)
    cells = [(i, j, k) for i in range(n) for j in range(n) for k in range(n) if (i + j + k) % 3]
    pts = []
    for (i, j, k) in cells:
        for _ in range(rng.integers(1, 6)):
            pts.append((i + rng.random(), j + rng.random(), k + rng.random()))
    p = np.array(pts)
    u, v, w = (rng.normal(size=len(p)) for _ in range(3))
    ps = O.Particles(p[:, 0], p[:, 1], p[:, 2], u, v, w, type=np.zeros(len(p), np.uint8))
    target = 4
    out, st = O.particle_adaptivity(_domain(n), O.adapt_cfg(liquid_particle_refill=1,
                                                             liquid_refill_particles_per_cell_target=target), ps)
    cell = (np.floor(ps.x) + n * (np.floor(ps.y) + n * np.floor(ps.z))).astype(int)
    counts = np.bincount(cell, minlength=n ** 3)
    under = (counts > 0) & (counts < target)
    assert st.refill_cells == (counts > 0).sum() and st.refill_underfull_cells == under.sum()
    assert st.refill_added == (target - counts[under]).sum() == len(out) - len(ps)
    assert st.refill_budget_limited == 0 and st.refill_interface_cells == 0
    added = slice(len(ps), len(out))
    # added particles sit at the 2x2x2 sub-cell centres of their cell, with the
    # cell's mean velocity, and the cell's volume is split evenly
    fx = out.x[added] - np.floor(out.x[added])
    assert set(np.round(fx, 12)) <= {0.25, 0.75}
    ac = (np.floor(out.x[added]) + n * (np.floor(out.y[added]) + n * np.floor(out.z[added]))).astype(int)
    for c in np.unique(ac):
        m = cell == c
        assert np.allclose(out.u[added][ac == c], ps.u[m].mean(), rtol=1e-12)
        allc = (np.floor(out.x) + n * (np.floor(out.y) + n * np.floor(out.z))).astype(int) == c
        assert allc.sum() == target and np.allclose(out.volume[allc], counts[c] / target, rtol=1e-15)
    # a budget stops refill in linear cell order and flags it
    b = 7
    out2, st2 = O.particle_adaptivity(_domain(n), O.adapt_cfg(liquid_particle_refill=1,
                                                               liquid_particle_refill_max_added_per_step=b), ps)
    assert st2.refill_added == b and st2.refill_budget_limited == 1
    assert np.array_equal(out2.x[len(ps):], out.x[len(ps):len(ps) + b])


def _random_mix(n, seed, nper=6):
    rng = np.random.default_rng(seed)
    pts = []
    for k in range(1, n - 1):
        for j in range(1, n - 1):
            for i in range(1, n - 1):
                liquid = j < n // 2 or (i - n / 2) ** 2 + (k - n / 2) ** 2 < 2
                c = int(rng.integers(0, nper + 1))
                for _ in range(c):
                    pts.append((i + rng.random(), j + rng.random(), k + rng.random(), 0 if liquid else 1))
    p = np.array(pts)
    u, v, w = (rng.normal(size=len(p)) for _ in range(3))
    return O.Particles(p[:, 0], p[:, 1], p[:, 2], u, v, w, type=p[:, 3].astype(np.uint8),
                       volume=rng.uniform(0.5, 1.5, len(p)))


ALL_ON = dict(narrow_band_air=1, gas_particle_coarsening=1, liquid_particle_coarsening=1, liquid_particle_refill=1)
CASES = [
    ALL_ON,
    dict(ALL_ON, narrow_band_air_radius=0, gas_particles_per_cell_target=2, liquid_particles_per_cell_target=5,
         liquid_refill_particles_per_cell_target=3),
    dict(ALL_ON, liquid_particle_refill_interface_only=1, liquid_particle_refill_interface_radius=2,
         liquid_particle_refill_max_added_per_step=50),
    dict(liquid_particle_refill=1, liquid_refill_particles_per_cell_target=9, liquid_particle_refill_max_added_per_step=1),
]


@needs_ref
@pytest.mark.parametrize("case", range(len(CASES)))
def test_bit_exact_against_the_reference(case):
    n = 12
    ps = _random_mix(n, 10 + case)
    dx, origin = 0.125, (0.5, -1.0, 2.0)
    ps.x, ps.y, ps.z = ps.x * dx + origin[0], ps.y * dx + origin[1], ps.z * dx + origin[2]
    d = _domain(n, dx, origin)
    cfg = O.adapt_cfg(**CASES[case])
    out, st = O.particle_adaptivity(d, cfg, ps)
    rout, rst = O.ref_particle_adaptivity(d, cfg, ps)
    assert st.values() == rst
    for a in ("x", "y", "z", "u", "v", "w", "type", "volume"):
        assert np.array_equal(getattr(out, a), getattr(rout, a)), a
    assert sum(st.values()) > 0


@needs_ref
def test_step_with_adaptivity_bit_exact():
    # SparseSim3DTP::step with every adaptivity stage on (sparse_sim3d_tp.cpp:847-856)
    n = 12
    dx, x, y, z, t = scene(O.SCENE_BUBBLE, n, seed=4, ny=n)
    ph = phase()
    cfg = step_cfg(ph, dt=2e-3)
    acfg = O.adapt_cfg(**CASES[1])
    ps = O.Particles(x, y, z, type=t)
    g = O.Grid(n, n, n, dx, faithful=True)
    rs = O.RefSim(n, n, n, dx)
    rs.config(cfg)
    rs.set_adapt(acfg)
    rs.set_particles(ps)
    d = O.PaDomain(n, n, n, 0, dx, 0.0, 0.0, 0.0)
    changed = 0
    for _ in range(3):
        ps, st = O.particle_adaptivity(d, acfg, ps)
        changed += st.refill_added + st.gas_removed + st.liquid_removed + st.narrow_band_air_removed
        g.step(ps, cfg)
        rs.step()
        assert st.values() == rs.adapt_stats()
        rp = rs.particles_full()
        for a in ("x", "y", "z", "u", "v", "w", "type", "volume"):
            assert np.array_equal(getattr(ps, a), getattr(rp, a)), a
    assert changed > 0


@needs_ref
def test_mr_step_with_adaptivity_bit_exact():
    # MRSim3DTP::step: adaptivity, then dynamic refinement (multires_sim3d_tp.cpp:671-674)
    n = 16
    r = O.RefMR(n, n, n, 1.0)
    r.init_bubble()
    ph = phase(1.0, 0.01)
    r.config(dt=0.02, phase=ph)
    acfg = O.adapt_cfg(**CASES[0])
    r.set_adapt(acfg)
    ps = r.particles_full()
    g = O.MRGrid(n, n, n, 1.0, faithful=True)
    g.set_leaves(r.leaves())
    sim = O.mr_sim(n, n, n, 1.0, phase=ph)
    d = O.PaDomain(n, n, n, 0, 1.0, 0.0, 0.0, 0.0)
    changed = 0
    for _ in range(3):
        ps, st = O.particle_adaptivity(d, acfg, ps)
        changed += st.refill_added + st.gas_removed + st.liquid_removed + st.narrow_band_air_removed
        O.mr_step(g, sim, ps)
        r.step()
        assert st.values() == r.adapt_stats()
        rp = r.particles_full()
        for a in ("x", "y", "z", "u", "v", "w", "type", "volume"):
            assert np.array_equal(getattr(ps, a), getattr(rp, a)), a
        assert np.array_equal(g.leaves(), r.leaves())
    assert changed > 0
