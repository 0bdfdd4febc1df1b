"""Render-cache export in the oracle (oracle/pf_oracle_render.c, restating
src/driver/render_cache3d.h): every line parses as JSON and the sections agree with
each other (counts, volume sums, bounds), and -- where oracle/_ref was built -- the
frame files and the manifest are byte-identical to the reference's own writers after
identical sparse and multires steps."""
import json

import numpy as np
import pytest

import pyoracle as O
from _helpers import phase, scene, step_cfg

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built")


def _lines(path):
    with open(path) as f:
        return [json.loads(line) for line in f]


def _check_structure(recs, n_particles):
    sec = [r["section"] for r in recs]
    assert sec[:5] == ["header", "camera", "water_volume", "secondary_channels", "cinematic_metadata"]
    cells = [r for r in recs if r["section"] == "phase_cell"]
    pf = recs[5]
    assert pf["section"] == "phase_field" and pf["count"] == len(cells)
    wv = recs[2]
    assert wv["phase_field_cells"] == len(cells)
    assert wv["phase_field_liquid_volume"] == pytest.approx(sum(c["liquid_volume"] for c in cells), rel=1e-12)
    parts = [r for r in recs if r["section"] == "particle"]
    assert len(parts) == n_particles and [p["index"] for p in parts] == list(range(n_particles))
    liquid = [p for p in parts if p["phase"] == "liquid"]
    assert wv["primary_liquid_count"] == len(liquid) and recs[4]["primary_particle_count"] == n_particles
    assert all(p["render_channel"] == ("water" if p["phase"] == "liquid" else "air") for p in parts)
    cm = recs[4]
    if liquid:
        pos = np.array([p["position"] for p in liquid])
        assert np.all(np.array(cm["water_bounds_min"]) <= pos.min(axis=0))
        assert np.all(np.array(cm["water_bounds_max"]) >= pos.max(axis=0))
    kinds = [r["kind"] for r in recs if r["section"] == "particles"]
    assert kinds == ["primary", "secondary_droplet", "secondary_bubble"]
    return cells, parts


@needs_ref
def test_default_camera_is_the_references():
    for dims, dx in (((8, 12, 8), 1.0), ((64, 32, 16), 1.0 / 64)):
        ref = np.zeros(13)
        O.reflib().ref_render_default_camera(*dims, dx, ref.ctypes.data_as(O.C.c_void_p))
        assert list(O.default_camera(*dims, dx)) == list(ref)
    cam = list(O.default_camera(10, 20, 30, 0.5))
    assert cam[:6] == [2.5, 5.0 + 0.2 * 15.0, 7.5 + 30.0, 2.5, 5.0, 7.5] and cam[9:] == [45.0, 0.025, 75.0, 50.0]


@needs_ref
@pytest.mark.parametrize("kind", [O.SCENE_BUBBLE, O.SCENE_RT])
def test_sparse_frame_bit_exact(kind, tmp_path):
    n = 12
    dx, x, y, z, t = scene(kind, n, seed=3, ny=n)
    ph = phase()
    cfg = step_cfg(ph, dt=2e-3)
    ps = O.Particles(x, y, z, type=t)
    g = O.Grid(n, n, n, dx, faithful=True)
    rs = O.RefSim(n, n, n, dx)
    rs.config(cfg)
    rs.set_particles(ps)
    for _ in range(2):
        st = g.step(ps, cfg)
        rs.step()
    meta = O.render_meta(n, n, n, dx, ph, frame=7, time=0.25, dt=st.ts.effective_dt, Vp=1.0)
    a, b = tmp_path / "oracle.jsonl", tmp_path / "ref.jsonl"
    assert g.render_cache(ps, ph, meta, a) == 0
    assert rs.render_cache(b, 7, 0.25) == 0
    assert a.read_bytes() == b.read_bytes()
    cells, parts = _check_structure(_lines(a), len(ps))
    assert len(cells) > 0 and any(c["marker"] == 2 for c in cells)


@needs_ref
def test_mr_frame_bit_exact(tmp_path):
    n = 16
    r = O.RefMR(n, n, n, 1.0)
    r.init_bubble()
    ph = phase(1.0, 0.01)
    r.config(dt=0.02, phase=ph)
    ps = r.particles_full()
    g = O.MRGrid(n, n, n, 1.0, faithful=True)
    g.set_leaves(r.leaves())
    sim = O.mr_sim(n, n, n, 1.0, phase=ph)
    st = None
    for _ in range(2):
        st = O.mr_step(g, sim, ps)
        r.step()
    meta = O.render_meta(n, n, n, 1.0, ph, frame=2, time=0.04, dt=st.ts.effective_dt, Vp=1.0)
    a, b = tmp_path / "oracle.jsonl", tmp_path / "ref.jsonl"
    assert g.render_cache(ps, ph, meta, a) == 0
    assert r.render_cache(b, 2, 0.04) == 0
    assert a.read_bytes() == b.read_bytes()
    cells, _ = _check_structure(_lines(a), len(ps))
    assert {c["level"] for c in cells} >= {0}


@needs_ref
def test_manifest_bit_exact_with_escapes(tmp_path):
    frames = [(0, 0, 0.0, "frame_0000.jsonl", 1234), (1, 10, 0.2, 'odd "name"\\dir\n\t\x01.jsonl', 99),
              (2, 20, 0.4000000000000001, "f2.jsonl", 0)]
    a, b = tmp_path / "m_oracle.json", tmp_path / "m_ref.json"
    assert O.render_manifest(str(a), "sparse3d_tp", 8, 12, 8, 0.125, frames) == 0
    assert O.render_manifest(str(b), "sparse3d_tp", 8, 12, 8, 0.125, frames, ref=True) == 0
    assert a.read_bytes() == b.read_bytes()
    m = json.loads(a.read_text())
    assert m["frame_count"] == 3 and m["frames"][1]["path"] == frames[1][3]
    # invalid input is refused by both writers
    for bad in ([(0, -1, 0.0, "a", 0)], [(0, 0, float("nan"), "a", 0)], [(0, 0, 0.0, "", 0)]):
        assert O.render_manifest(str(a), "sparse3d_tp", 8, 12, 8, 0.125, bad) == -1
        assert O.render_manifest(str(b), "sparse3d_tp", 8, 12, 8, 0.125, bad, ref=True) == -1


def test_invalid_metadata_is_refused(tmp_path):
    n = 8
    ph = phase()
    g = O.Grid(n, n, n, 1.0)
    ps = O.Particles([2.5], [2.5], [2.5], type=[0])
    for kw in (dict(dt=float("inf")), dict(time=float("nan"))):
        assert g.render_cache(ps, ph, O.render_meta(n, n, n, 1.0, ph, **kw), tmp_path / "x.jsonl") == -1
    cam = O.default_camera(n, n, n, 1.0)
    cam[11] = cam[10]          # far clip not beyond near clip
    assert g.render_cache(ps, ph, O.render_meta(n, n, n, 1.0, ph, camera=cam), tmp_path / "x.jsonl") == -1
