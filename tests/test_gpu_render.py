"""Render-cache export from the GPU context against the oracle's writer
(src/driver/render_cache3d.h; the oracle's files are byte-identical to the
reference's, tests/test_render_oracle.py).

Both writers get the same state: the GPU's face masses, markers and particles (FP32
values), and on the sparse grid the oracle's own markCells over the particles the
GPU marked, which fixes the marker blocks' allocation order.  Bar: the files are
byte-identical -- every number is formatted from the same double.
"""
import json

import numpy as np
import pytest

import pfflip_amd as G
import pyoracle as O
from _helpers import bubble_lattice, phase, scene

pytestmark = pytest.mark.gpu


def _gphase(oph):
    return G.Phase(oph.rho_l, oph.rho_g, oph.alpha_phi, oph.rho_tilde_0)


def _info(n, dx, oph, frame=3, time=0.125, dt=2e-3, Vp=1.0):
    return G.RenderFrame(frame, 0, time, dt, Vp, _gphase(oph), G.default_render_camera(n, n, n, dx))


def _oracle_particles(ctx):
    x, y, z, u, v, w = ctx.download()
    t, vol = ctx.download_ext()
    return O.Particles(x, y, z, u, v, w, type=t, volume=vol)


def _sparse_oracle(ctx, n, dx, marked):
    """an oracle grid marked by markCells over `marked` (allocation order) holding the
    GPU's face masses"""
    g = O.Grid(n, n, n, dx, faithful=True)
    g.mark_cells(marked)
    assert np.array_equal(g.dense(O.MK), ctx.field(G.FIELD_MARKER).astype(np.float64))
    for f in (O.MU, O.MV, O.MW):
        g.load_dense(f, ctx.field(f).astype(np.float64))
    return g


@pytest.mark.parametrize("kind", [O.SCENE_BUBBLE, O.SCENE_RT, O.SCENE_DAMBREAK])
def test_sparse_frame_matches_oracle_after_marking(kind, tmp_path):
    n = 16
    print(f"[render] kind={kind} n={n}", flush=True)
    dx, x, y, z, t = scene(kind, n, seed=4)
    rng = np.random.default_rng(kind)
    u, v, w = (rng.normal(size=len(x)).astype(np.float32) for _ in range(3))
    vol = rng.uniform(0.5, 1.5, len(x)).astype(np.float32)
    # a shuffled particle order changes the allocation order of the marker blocks
    perm = rng.permutation(len(x))
    x, y, z, u, v, w, t, vol = (a[perm] for a in (x, y, z, u, v, w, t, vol))
    oph = phase()
    ctx = G.Context(n, n, n, dx)
    ctx.upload(x, y, z, u, v, w, t, vol)
    ctx.mark_cells()
    ctx.p2g(_gphase(oph), 1.0)
    a, b = tmp_path / "gpu.jsonl", tmp_path / "oracle.jsonl"
    nbytes = ctx.write_render_cache(a, _info(n, dx, oph, Vp=0.5))
    ps = _oracle_particles(ctx)
    g = _sparse_oracle(ctx, n, dx, ps)
    assert g.render_cache(ps, oph, O.render_meta(n, n, n, dx, oph, frame=3, time=0.125, dt=2e-3, Vp=0.5), b) == 0
    assert a.read_bytes() == b.read_bytes() and nbytes == a.stat().st_size


def test_sparse_frame_after_a_step(tmp_path):
    n = 16
    print(f"[render-step] n={n}", flush=True)
    dx, x, y, z, t = scene(O.SCENE_BUBBLE, n, seed=6)
    oph = phase()
    sim = G.SparseSim3DTP(n, n, n, dx)
    sim.phase = _gphase(oph)
    sim.dt = 2e-3
    sim.set_particles(x, y, z, type=t)
    for _ in range(2):
        before = _oracle_particles(sim.ctx)
        sim.step()
    a, b = tmp_path / "gpu.jsonl", tmp_path / "oracle.jsonl"
    sim.write_render_cache(a, 2, 0.004)
    # the markers and their block order come from the positions the step marked
    g = _sparse_oracle(sim.ctx, n, dx, before)
    ps = _oracle_particles(sim.ctx)
    meta = O.render_meta(n, n, n, dx, oph, frame=2, time=0.004, dt=sim.effective_dt_last, Vp=1.0)
    assert g.render_cache(ps, oph, meta, b) == 0
    assert a.read_bytes() == b.read_bytes()
    recs = [json.loads(line) for line in a.read_text().splitlines()]
    assert recs[0]["sim_kind"] == "sparse3d_tp" and sum(r["section"] == "particle" for r in recs) == len(ps)


def test_sparse_frame_without_marking_lists_the_same_cells(tmp_path):
    # new particles uploaded after the marking: the allocation order is gone, the
    # blocks follow in id order, and the same cells are written
    n = 12
    dx, x, y, z, t = scene(O.SCENE_RT, n, seed=8)
    oph = phase()
    ctx = G.Context(n, n, n, dx)
    ctx.upload(x, y, z, type=t)
    ctx.mark_cells()
    ctx.p2g(_gphase(oph), 1.0)
    marked = _oracle_particles(ctx)
    ctx.upload(x[::-1].copy(), y[::-1].copy(), z[::-1].copy(), type=t[::-1].copy())
    a, b = tmp_path / "gpu.jsonl", tmp_path / "oracle.jsonl"
    ctx.write_render_cache(a, _info(n, dx, oph))
    g = _sparse_oracle(ctx, n, dx, marked)
    ps = _oracle_particles(ctx)
    assert g.render_cache(ps, oph, O.render_meta(n, n, n, dx, oph, frame=3, time=0.125, dt=2e-3), b) == 0
    la, lb = a.read_text().splitlines(), b.read_text().splitlines()
    cells = lambda ls: [s for s in ls if '"phase_cell"' in s]
    rest = lambda ls: [s for s in ls if '"phase_cell"' not in s]
    assert sorted(cells(la)) == sorted(cells(lb)) and rest(la) == rest(lb)


@pytest.mark.parametrize("box", [None, [0, 0, 0, 32, 8, 32]])
def test_mr_frame_matches_oracle(box, tmp_path):
    n = 32
    print(f"[render-mr] box={box}", flush=True)
    leaves = O.mr_layout_bubble_band(n, n, n) if box is None else O.mr_layout_fine_box(n, n, n, box)
    oph = phase(1.0, 0.01)
    ctx = G.MRContext(n, n, n, 1.0)
    ctx.set_layout(leaves)
    ctx.upload(*bubble_lattice(n, seed=2))
    ctx.mr_mark_cells()
    ctx.mr_p2g(_gphase(oph), 1.0)
    a, b = tmp_path / "gpu.jsonl", tmp_path / "oracle.jsonl"
    ctx.write_render_cache(a, _info(n, 1.0, oph, frame=1, time=0.02, dt=0.02))
    g = O.MRGrid(n, n, n, 1.0, faithful=True)
    g.set_leaves(leaves)
    for f in (O.MU, O.MV, O.MW, O.MK):
        g.load_dense(f, ctx.field(f if f != O.MK else G.FIELD_MARKER).astype(np.float64))
    ps = _oracle_particles(ctx)
    assert g.render_cache(ps, oph, O.render_meta(n, n, n, 1.0, oph, frame=1, time=0.02, dt=0.02), b) == 0
    assert a.read_bytes() == b.read_bytes()
    # both layouts keep coarse leaves (the top of the tank, walls included)
    levels = {json.loads(s)["level"] for s in a.read_text().splitlines() if '"phase_cell"' in s}
    assert levels == {0, 1}


def test_invalid_metadata_is_refused(tmp_path):
    n = 8
    oph = phase()
    ctx = G.Context(n, n, n, 1.0)
    info = _info(n, 1.0, oph)
    info.camera.far_clip = info.camera.near_clip
    with pytest.raises(G.PfError, match="invalid render cache metadata"):
        ctx.write_render_cache(tmp_path / "x.jsonl", info)
