"""The render-cache manifest writer and default camera of the C ABI (host code, no
device): byte-identical to the oracle's, which the reference's own writer pins
(tests/test_render_oracle.py)."""
import pytest

import pfflip_amd as G
import pyoracle as O


def test_default_camera_matches_the_oracle():
    for dims, dx in (((8, 12, 8), 1.0), ((64, 32, 16), 1.0 / 64), ((3, 3, 200), 0.37)):
        g = G.default_render_camera(*dims, dx)
        o = list(O.default_camera(*dims, dx))
        assert list(g.position) + list(g.target) + list(g.up) + [g.fov_degrees, g.near_clip, g.far_clip,
                                                                  g.focal_length_mm] == o


def test_manifest_bytes_match_the_oracle(tmp_path):
    frames = [(0, 0, 0.0, "frame_0000.jsonl", 1234), (1, 10, 0.2, 'odd "name"\\dir\n\t\x01.jsonl', 99),
              (2, 20, 0.4000000000000001, "f2.jsonl", 0), (3, 30, 1e-300, "f3", 2 ** 40)]
    a, b = tmp_path / "gpu.json", tmp_path / "oracle.json"
    for kind, dims, dx in (("sparse3d_tp", (8, 12, 8), 0.125), ("multires3d_tp", (17, 10, 9), 1.0 / 3)):
        G.write_render_manifest(a, kind, *dims, dx, frames)
        assert O.render_manifest(str(b), kind, *dims, dx, frames) == 0
        assert a.read_bytes() == b.read_bytes()
    G.write_render_manifest(a, "sparse3d_tp", 8, 8, 8, 1.0, [])
    assert O.render_manifest(str(b), "sparse3d_tp", 8, 8, 8, 1.0, []) == 0
    assert a.read_bytes() == b.read_bytes()


@pytest.mark.parametrize("bad", [[(0, -1, 0.0, "a", 0)], [(0, 0, float("nan"), "a", 0)], [(0, 0, 0.0, "", 0)],
                                 [(0, 0, 0.0, "a", -1)]])
def test_manifest_rejects_what_the_reference_rejects(bad, tmp_path):
    with pytest.raises(G.PfError):
        G.write_render_manifest(tmp_path / "m.json", "sparse3d_tp", 8, 8, 8, 1.0, bad)
    assert O.render_manifest(str(tmp_path / "o.json"), "sparse3d_tp", 8, 8, 8, 1.0, bad) == -1
