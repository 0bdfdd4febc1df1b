"""Direct device-side checks of the integer parts of P2G and of the phase field.

* Binning: particles per cell from the GPU counting sort equal a brute-force count
  of markCells' cell index (int)((x - o)/dx) evaluated in FP64
  (src/driver/sparse_sim3d_tp.cpp:43-46), bit for bit.
* Colour assignment: the eight P2G source-block lists are exactly the particle
  blocks of colour (bx&1) + 2(by&1) + 4(bz&1) (src/grid/block_color.h:7), and the
  node regions that two blocks of one colour write never overlap (the property
  that makes one launch per colour race-free, as runColor8's tasks are,
  src/driver/sparse_ops3d_common.h:17-28).
* Phase field: the FP64 phi the device computes from its face masses matches the
  FP64 oracle's phi of the oracle's masses to 1e-5 norm-wise, and equals Eq. 7
  (src/physics/phasefield.h:10-14) evaluated on the host for the same masses bit for bit.
* Divergence after projection: the GPU's and the FP64 oracle's post-projection
  divergence differ by at most 1e-5 of the pre-projection divergence, norm-wise.
"""
import math
import os

import numpy as np
import pytest

import pfflip_amd as G
import pyoracle as O
from _helpers import phase, rel_err, scene

pytestmark = pytest.mark.gpu

SIZES = [(O.SCENE_RT, 24), (O.SCENE_BUBBLE, 24), (O.SCENE_DAMBREAK, 28), (O.SCENE_RT, 64), (O.SCENE_BUBBLE, 128)]


def _gphase(oph):
    return G.Phase(oph.rho_l, oph.rho_g, oph.alpha_phi, oph.rho_tilde_0)


def _host_threads():
    return max(1, min(16, os.cpu_count() or 1))


def _brute_counts(x, y, z, n, dx):
    # markCells' cell of a particle: (int)((x - o) / dx) in FP64, truncation toward 0
    i = np.trunc((x.astype(np.float64) - 0.0) / dx).astype(np.int64)
    j = np.trunc((y.astype(np.float64) - 0.0) / dx).astype(np.int64)
    k = np.trunc((z.astype(np.float64) - 0.0) / dx).astype(np.int64)
    cnt = np.bincount(i + n * (j + n * k), minlength=n ** 3)
    return cnt.reshape(n, n, n)


def _color8(bx, by, bz):
    return (bx & 1) + 2 * (by & 1) + 4 * (bz & 1)


@pytest.mark.parametrize("kind,n", SIZES)
def test_bin_counts_and_colour_lists_are_exact(kind, n):
    dx, x, y, z, t = scene(kind, n, seed=17)
    rng = np.random.default_rng(n)
    perm = rng.permutation(len(x))  # caller order must not matter
    x, y, z, t = x[perm], y[perm], z[perm], t[perm]
    ctx = G.Context(n, n, n, dx)
    ctx.upload(x, y, z, type=t)
    print(f"[bitexact] scene {kind} n {n}: {len(x)} particles", flush=True)
    got = ctx.bin_counts()
    want = _brute_counts(x, y, z, n, dx)
    assert got.dtype == np.uint32 and np.array_equal(got.astype(np.int64), want)
    assert int(got.sum()) == len(x)

    nb = (n + 3) // 4
    blk = want.reshape(nb, 4, nb, 4, nb, 4).sum(axis=(1, 3, 5))  # [bz, by, bx]
    bz, by, bx = np.nonzero(blk)
    bids = bx + nb * (by + nb * bz)
    col = _color8(bx, by, bz)
    lists = ctx.color_lists()
    assert sum(len(lst) for lst in lists) == len(bids)
    for c in range(8):
        assert np.array_equal(lists[c], np.sort(bids[col == c]).astype(np.int32)), c
        # a source block writes face nodes [4b - 1, 4b + 5] per axis (kernel support
        # [c - 1, c + 2] of its cells); same-colour blocks are 2 apart on some axis
        b = lists[c].astype(np.int64)
        cx, cy, cz = b % nb, (b // nb) % nb, b // (nb * nb)
        assert np.all(_color8(cx, cy, cz) == c)
        # node intervals [4a - 1, 4a + 5] and [4a' - 1, 4a' + 5] overlap iff |a - a'| <= 1,
        # so two source regions overlap iff the blocks are 26-neighbours: none may be
        # of the same colour
        members = set(b.tolist())
        for ox in (-1, 0, 1):
            for oy in (-1, 0, 1):
                for oz in (-1, 0, 1):
                    if ox == oy == oz == 0:
                        continue
                    qx, qy, qz = cx + ox, cy + oy, cz + oz
                    ok = (qx >= 0) & (qx < nb) & (qy >= 0) & (qy < nb) & (qz >= 0) & (qz < nb)
                    q = (qx + nb * (qy + nb * qz))[ok]
                    assert not any(int(v) in members for v in q)


@pytest.mark.parametrize("kind,n", [(O.SCENE_RT, 24), (O.SCENE_BUBBLE, 24), (O.SCENE_RT, 64)])
def test_device_phase_field(kind, n):
    dx, x, y, z, t = scene(kind, n, seed=21)
    oph = phase()
    ctx = G.Context(n, n, n, dx)
    ctx.upload(x, y, z, type=t)
    ctx.mark_cells()
    ctx.p2g(_gphase(oph), 1.0)
    O.set_threads(_host_threads())
    try:
        g = O.Grid(n, n, n, dx, faithful=False)
        ps = O.Particles(x, y, z, type=t)
        g.mark_cells(ps)
        g.p2g(ps, oph, 1.0)
    finally:
        O.set_threads(1)
    rmin = math.log(oph.rho_l / oph.rho_g) * oph.rho_g * oph.rho_tilde_0
    den = oph.alpha_phi * oph.rho_tilde_0 * oph.rho_l
    for ax in range(3):
        dev = ctx.face_phi(_gphase(oph), ax)
        # Eq. 7 of the oracle's FP64 masses
        mo = g.dense(O.MU + ax)
        phi_o = np.where(mo < rmin, 0.0, np.minimum(np.sqrt(np.maximum(mo - rmin, 0.0) / den), 1.0))
        assert rel_err(dev, phi_o) < 1e-5, ax
        # the same formula on the device's own FP32 masses, evaluated on the host
        mg = ctx.field(G.FIELD_MU + ax).astype(np.float64)
        phi_h = np.where(mg < rmin, 0.0, np.minimum(np.sqrt(np.maximum(mg - rmin, 0.0) / den), 1.0))
        assert np.array_equal(dev, phi_h), ax
        # spot-check the host formula against the oracle's own phi
        for q in np.random.default_rng(ax).integers(0, mo.size, 200):
            assert O.phi(float(mo.flat[q]), oph) == phi_o.flat[q]


@pytest.mark.parametrize("kind,n", [(O.SCENE_RT, 24), (O.SCENE_BUBBLE, 24), (O.SCENE_RT, 64)])
def test_post_projection_divergence_matches_oracle(kind, n):
    dt = 1e-3
    dx, x, y, z, t = scene(kind, n, seed=11)
    rng = np.random.default_rng(11)
    u, v, w = (rng.normal(scale=0.3, size=len(x)).astype(np.float32) for _ in range(3))
    oph = phase()
    ctx = G.Context(n, n, n, dx)
    ctx.upload(x, y, z, u, v, w, t)
    ps = O.Particles(x, y, z, u, v, w, t)
    ctx.mark_cells()
    ctx.p2g(_gphase(oph), 1.0)
    ctx.save_velocity()
    ctx.apply_gravity(dt, -9.81)
    ctx.apply_walls()
    O.set_threads(_host_threads())
    try:
        g = O.Grid(n, n, n, dx, faithful=False)
        g.mark_cells(ps)
        g.p2g(ps, oph, 1.0)
        g.apply_gravity(dt, -9.81)
        g.apply_walls()
    finally:
        O.set_threads(1)
    mk = g.dense(O.MK)

    def div(U, V, W):
        # the reference's RHS divergence: faces next to a SOLID cell count as 0
        # (sparse_ops3d_tp.cpp:270-281); walls are zeroed already, so the plain
        # difference is the same on FLUID cells
        d = (U[:, :, 1:] - U[:, :, :-1]) + (V[:, 1:, :] - V[:, :-1, :]) + (W[1:, :, :] - W[:-1, :, :])
        return d[mk == 1] / dx

    pre = div(*(g.dense(f) for f in (O.U, O.V, O.W)))
    cfg = G.default_solve_cfg(G.STOP_REL2, 1e-9, 20000)
    st = ctx.project(_gphase(oph), dt, cfg)
    ost = g.project(oph, dt, 20000, 1e-12, 0.0, O.STOP_REL2)
    assert st.converged == 1 and ost.converged == 1
    dg = div(*(ctx.field(f).astype(np.float64) for f in range(3)))
    do = div(g.dense(O.U), g.dense(O.V), g.dense(O.W))
    npre = np.linalg.norm(pre)
    assert npre > 0
    assert np.linalg.norm(dg - do) <= 1e-5 * npre
    assert np.linalg.norm(do) <= 1e-6 * npre  # the oracle itself projects to ~0
