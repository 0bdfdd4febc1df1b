"""The oracle's threaded particle loops (P2G slabs, G2P and advection ranges) give
the serial loops' results bit for bit: face values, active-block sets in their
first-touch order, particle state and clamp counters, in FP64 and float-faithful
mode.  The large-size GPU tests rely on the threaded oracle."""
import numpy as np
import pytest

import pyoracle as O
from _helpers import phase, scene, step_cfg


@pytest.mark.parametrize("kind,n", [(O.SCENE_RT, 20), (O.SCENE_BUBBLE, 24), (O.SCENE_DAMBREAK, 28)])
@pytest.mark.parametrize("faithful", [False, True])
def test_threaded_oracle_matches_serial(kind, n, faithful):
    dx, x, y, z, t = scene(kind, n, seed=5)
    rng = np.random.default_rng(1)
    u, v, w = (rng.normal(size=len(x)).astype(np.float32) for _ in range(3))
    oph = phase()
    out = []
    try:
        for th in (1, 7):
            O.set_threads(th)
            ps = O.Particles(x, y, z, u, v, w, t)
            g = O.Grid(n, n, n, dx, faithful=faithful)
            cfg = step_cfg(oph)
            g.step(ps, cfg)
            st = g.step(ps, cfg)
            es = st.escape
            out.append(([g.dense(f) for f in range(7)], [g.active_blocks(f) for f in range(7)],
                        [ps.x.copy(), ps.y.copy(), ps.z.copy(), ps.u.copy(), ps.v.copy(), ps.w.copy()],
                        (es.clamped_liquid, es.clamped_gas, es.x_lo, es.x_hi, es.y_lo, es.y_hi, es.z_lo, es.z_hi)))
    finally:
        O.set_threads(1)
    a, b = out
    for p, q in zip(a[0], b[0]):
        assert np.array_equal(p, q)
    for p, q in zip(a[1], b[1]):
        assert np.array_equal(p, q)  # first-touch order, not only the set
    for p, q in zip(a[2], b[2]):
        assert np.array_equal(p, q)
    assert a[3] == b[3]
