"""The reference's own doctest cases for the sparse two-phase path, run through
the HIP C-ABI on the GPU (same inputs, same expected answers)."""
import numpy as np
import pytest

import pfflip_amd as G

pytestmark = pytest.mark.gpu


def _ctx8():
    return G.Context(8, 8, 8, 1.0)


def test_p2g_conserves_normalized_cubic_momentum():
    # tests/test_sparse_transfer3d_tp.cpp:11-29
    c = _ctx8()
    c.upload([3.0], [2.5], [2.5], [4.0], [3.0], [2.0], [0])
    c.p2g(G.Phase(1.0, 0.01, 1.0, 1.0), 1.0)
    momU = float((c.field(G.FIELD_U).astype(np.float64) * c.field(G.FIELD_MU)).sum())
    assert momU == pytest.approx(4.0, rel=1e-6)
    nb = len(c.active_blocks(G.FIELD_MU))
    assert 0 < nb < 12
    assert float((c.field(G.FIELD_V).astype(np.float64) * c.field(G.FIELD_MV)).sum()) == pytest.approx(3.0, rel=1e-6)
    assert float((c.field(G.FIELD_W).astype(np.float64) * c.field(G.FIELD_MW)).sum()) == pytest.approx(2.0, rel=1e-6)


def test_p2g_separates_liquid_and_gas():
    # tests/test_sparse_transfer3d_tp.cpp:31-58
    xs, ys, zs, ts = [], [], [], []
    for k in range(1, 7):
        for j in range(1, 7):
            for i in range(1, 7):
                for s in range(8):
                    xs.append(i + 0.25 + 0.5 * (s & 1))
                    ys.append(j + 0.25 + 0.5 * ((s >> 1) & 1))
                    zs.append(k + 0.25 + 0.5 * ((s >> 2) & 1))
                    ts.append(0 if j < 4 else 1)
    ph = G.Phase(1.0, 0.01, 1.0, 1.0)
    ph.rho_tilde_0 = G.calibrate_rho_tilde0(ph)
    c = _ctx8()
    c.upload(xs, ys, zs, type=ts)
    c.p2g(ph, 1.0)
    mu = c.field(G.FIELD_MU)
    L = G.lib()
    import ctypes as C
    assert L.pf_phi_from_raw(float(mu[3, 2, 4]), C.byref(ph)) > 0.8
    assert L.pf_phi_from_raw(float(mu[3, 5, 4]), C.byref(ph)) < 0.2


def test_g2p_uses_phase_specific_flip_alpha():
    # tests/test_sparse_transfer3d_tp.cpp:60-78
    c = _ctx8()
    U = np.zeros((8, 8, 9), np.float32)
    S = np.zeros((8, 8, 9), np.float32)
    U[1:3, 1:3, 2:4] = 5.0
    S[1:3, 1:3, 2:4] = 2.0
    c.set_field(G.FIELD_U, U)
    c.set_field(G.FIELD_SAVED_U, S)
    c.upload([2.0, 2.0], [2.0, 2.0], [2.0, 2.0], [10.0, 10.0], [0, 0], [0, 0], [0, 1])
    c.g2p(1.0, 0.0)
    x, y, z, u, v, w = c.download()
    assert u[0] == pytest.approx(13.0)
    assert u[1] == pytest.approx(5.0)


def test_advect_reports_phase_boundary_clamps():
    # tests/test_sparse_transfer3d_tp.cpp:123-161
    c = _ctx8()
    c.set_field(G.FIELD_U, np.full((8, 8, 9), 10.0, np.float32))
    c.set_field(G.FIELD_V, np.full((8, 9, 8), -10.0, np.float32))
    c.upload([7.45, 4.0], [4.0, 0.55], [4.0, 4.0], type=[0, 1])
    es = c.advect(0.1, 2)
    x, y, z, *_ = c.download()
    assert x[0] == pytest.approx(7.5, rel=1e-7)
    assert y[1] == pytest.approx(0.5, rel=1e-7)
    assert (es.clamped_liquid, es.clamped_gas, es.x_hi, es.y_lo) == (1, 1, 1, 1)
    assert (es.x_lo, es.y_hi, es.z_lo, es.z_hi) == (0, 0, 0, 0)


def test_advect_rk3_follows_linear_field():
    # tests/test_sparse_transfer3d_tp.cpp:163-180
    c = _ctx8()
    U = np.broadcast_to(np.arange(9, dtype=np.float32), (8, 8, 9)).copy()
    c.set_field(G.FIELD_U, U)
    c.upload([2.0], [4.0], [4.0], type=[0])
    c.advect(0.1, 3)
    x, y, z, *_ = c.download()
    assert x[0] == pytest.approx(2.0 + 0.1 / 6.0 * (2.0 + 4.0 * 2.1 + 2.22), rel=1e-6)
    assert y[0] == pytest.approx(4.0, rel=1e-7) and z[0] == pytest.approx(4.0, rel=1e-7)


def _vc_setup(c, with_u):
    # tests/test_sparse_vc3d.cpp:8-38
    mk = np.zeros((8, 8, 8), np.float32)
    mk[:, :, [0, 7]] = 2
    mk[:, [0, 7], :] = 2
    mk[[0, 7], :, :] = 2
    mk[2:6, 2:6, 2:6] = 1
    c.set_field(G.FIELD_MARKER, mk)
    MU = np.zeros((8, 8, 9), np.float32)
    MU[2:6, 2:6, 2:7] = 1.0
    MV = np.zeros((8, 9, 8), np.float32)
    MV[2:6, 2:6, 2:6] = 1.0
    MW = np.zeros((9, 8, 8), np.float32)
    MW[2:7, 2:6, 2:6] = 1.0
    c.set_field(G.FIELD_MU, MU)
    c.set_field(G.FIELD_MV, MV)
    c.set_field(G.FIELD_MW, MW)
    if with_u:
        U = np.zeros((8, 8, 9), np.float32)
        U[2:6, 2:6, 2:7] = np.arange(2, 7, dtype=np.float32)
        c.set_field(G.FIELD_U, U)


def _box_div(c):
    U, V, W = (c.field(f).astype(np.float64) for f in (G.FIELD_U, G.FIELD_V, G.FIELD_W))
    return ((U[2:6, 2:6, 3:7] - U[2:6, 2:6, 2:6]) + (V[2:6, 3:7, 2:6] - V[2:6, 2:6, 2:6]) +
            (W[3:7, 2:6, 2:6] - W[2:6, 2:6, 2:6]))


def test_vc_projection_removes_divergence():
    # tests/test_sparse_vc3d.cpp:8-54 (cg_iters 1000, cg_tol 1e-9, reference stop rule)
    c = _ctx8()
    _vc_setup(c, True)
    st = c.project(G.Phase(1.0, 0.01, 1.0, 1.0), 1.0, G.default_solve_cfg(G.STOP_REF_MAXABS, 1e-9, 1000))
    assert np.abs(_box_div(c)).max() < 1e-4
    assert len(c.active_blocks(G.FIELD_P)) == 8
    assert st.converged == 1 and st.pinned == 0 and st.fluid_cells == 64


def test_vc_projection_preserves_target_cdiv():
    # tests/test_sparse_vc3d.cpp:56-100
    c = _ctx8()
    _vc_setup(c, False)
    c.project(G.Phase(1.0, 0.01, 1.0, 1.0), 1.0, G.default_solve_cfg(G.STOP_REF_MAXABS, 1e-9, 1000), 0.125)
    assert _box_div(c).mean() == pytest.approx(0.125, rel=1e-4)


def test_adaptive_timestep_clamps_by_particle_speed():
    # tests/test_sparse_sim3d_tp.cpp:85-107
    c = G.Context(4, 4, 4, 1.0)
    c.upload([1.0, 2.0], [1.0, 1.0], [1.0, 1.0], [10.0, 0.0], [0.0, 3.0], [0.0, 4.0], [0, 1])
    eff, vmax, cfl_dt, lim = c.timestep(0.2, True, 0.5, 1e-5)
    assert vmax == pytest.approx(10.0) and eff == pytest.approx(0.05) and lim == 1
    eff, vmax, cfl_dt, lim = c.timestep(0.2, False, 0.5, 1e-5)
    assert eff == pytest.approx(0.2) and lim == 0


def test_rejects_particles_outside_the_grid():
    c = _ctx8()
    with pytest.raises(G.PfError):
        c.upload([8.5], [1.0], [1.0])
    with pytest.raises(G.PfError):
        c.set_field(G.FIELD_MARKER, np.ones((8, 8, 8), np.float32))  # FLUID on the boundary


def test_cpp_drop_in_class_runs_the_reference_integration_tests():
    # adaptive-pf-flip_amd/host/sim_demo.cpp: the reference's SparseSim3DTP tests
    # (tests/test_sparse_sim3d_tp.cpp:110-169, 749-786), SparseSim3D dam-break
    # (tests/test_sparse_sim3d.cpp:6-36) and MRSim3DTP tests
    # (tests/test_multires_sim3d_tp.cpp:136-148, 345-439, 558-594) against the C++
    # drop-in classes
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = os.path.join(root, "adaptive-pf-flip_amd", "host", "sim_demo")
    assert os.path.exists(exe), "build the demo with adaptive-pf-flip_amd/build.py"
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr)
    assert out.returncode == 0
    assert "all checks passed" in out.stdout
