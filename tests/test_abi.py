"""The C-ABI library loads and exports every symbol include/pf_gpu.h declares, and
the Python structures have the C layout (no device work: runs without a GPU)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import pfflip_amd as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pf_gpu.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pf_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(G.LIB_PATH)
    names = _declared()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_layouts_match_the_header(tmp_path):
    src = tmp_path / "sizes.c"
    src.write_text('#include "pf_gpu.h"\n#include <stdio.h>\n#include <stddef.h>\n'
                   "int main(void){printf(\"%zu %zu %zu %zu %zu %zu %zu %zu %zu\\n\","
                   "sizeof(pf_grid_desc),sizeof(pf_phase),sizeof(pf_solve_cfg),"
                   "sizeof(pf_solve_stats),sizeof(pf_escape_stats),sizeof(pf_step_cfg),"
                   "sizeof(pf_step_stats),offsetof(pf_step_cfg,phase),offsetof(pf_step_stats,solve));"
                   "printf(\"%zu %zu %zu %zu %zu %zu\\n\",sizeof(pf_iface_stats),sizeof(pf_surface_stats),"
                   "offsetof(pf_step_cfg,surface_tension_strength),offsetof(pf_step_cfg,skip_interface_diagnostics),"
                   "offsetof(pf_step_stats,iface),offsetof(pf_step_stats,surface));"
                   "printf(\"%zu %zu %zu %zu %zu %zu\\n\",sizeof(pf_adapt_cfg),sizeof(pf_adapt_stats),"
                   "offsetof(pf_step_cfg,liquid_volume_target),offsetof(pf_step_cfg,adapt),"
                   "offsetof(pf_step_stats,c_div),offsetof(pf_step_stats,adapt));"
                   "printf(\"%zu %zu\\n\",sizeof(pf_mg_cfg),offsetof(pf_mg_cfg,chebyshev_ratio));"
                   "return 0;}\n")
    exe = tmp_path / "sizes"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True,
                                          check=True).stdout.split()]
    want = [C.sizeof(G.GridDesc), C.sizeof(G.Phase), C.sizeof(G.SolveCfg), C.sizeof(G.SolveStats),
            C.sizeof(G.EscapeStats), C.sizeof(G.StepCfg), C.sizeof(G.StepStats),
            G.StepCfg.phase.offset, G.StepStats.solve.offset,
            C.sizeof(G.IfaceStats), C.sizeof(G.SurfaceStats), G.StepCfg.surface_tension_strength.offset,
            G.StepCfg.skip_interface_diagnostics.offset, G.StepStats.iface.offset, G.StepStats.surface.offset,
            C.sizeof(G.AdaptCfg), C.sizeof(G.AdaptStats), G.StepCfg.liquid_volume_target.offset,
            G.StepCfg.adapt.offset, G.StepStats.c_div.offset, G.StepStats.adapt.offset,
            C.sizeof(G.MgCfg), G.MgCfg.chebyshev_ratio.offset]
    assert got == want


def test_host_helpers_without_device():
    # host-only entry points: Eq. 7 / beta and the rho_tilde_0 calibration
    import pyoracle as O
    ph = G.Phase(1000.0, 1.0, 1.0, 1.0)
    oph = O.Phase(1000.0, 1.0, 1.0, 1.0)
    assert G.calibrate_rho_tilde0(ph) == O.calibrate_rho_tilde0(oph, 1.0)
    L = G.lib()
    for raw in (0.0, 5.0, 7.0, 100.0, 900.0, 2000.0):
        assert L.pf_phi_from_raw(raw, C.byref(ph)) == O.phi(raw, oph)
    assert L.pf_beta_from_phi(0.25, C.byref(ph)) == O.beta(0.25, oph)
    assert L.pf_abi_version() == 5
    # the multigrid cycle's defaults are in range
    m = G.mg_defaults()
    assert 1 <= m.sweeps <= 16 and 1 <= m.coarsest_sweeps <= 1024 and 0 < m.omega < 1 and 0 < m.correction < 2
    assert m.chebyshev_ratio == 0 or 1 < m.chebyshev_ratio <= 1000


def test_null_arguments_are_rejected_without_device():
    L = G.lib()
    assert L.pf_create(None, None) == 1  # PF_ERR_ARG
    d = G.GridDesc(2, 8, 8, 0, 1.0, 0.0, 0.0, 0.0)
    h = C.c_void_p()
    assert L.pf_create(C.byref(h), C.byref(d)) == 1  # nx < 3
    assert L.pf_mark_cells(None) == 1


def test_multires_struct_layouts_match_the_header(tmp_path):
    src = tmp_path / "mr_sizes.c"
    src.write_text('#include "pf_gpu.h"\n#include <stdio.h>\n#include <stddef.h>\n'
                   "int main(void){printf(\"%zu %zu %zu %zu %zu\\n\","
                   "sizeof(pf_mr_solve_cfg),sizeof(pf_mr_refine_cfg),sizeof(pf_mr_step_cfg),"
                   "offsetof(pf_mr_step_cfg,solve),offsetof(pf_mr_step_cfg,refine));"
                   "printf(\"%zu %zu %zu\\n\",sizeof(pf_mr_solve_detail),offsetof(pf_mr_solve_cfg,mode),"
                   "offsetof(pf_mr_solve_cfg,coarse_preconditioner_max_work_ratio));return 0;}\n")
    exe = tmp_path / "mr_sizes"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True,
                                          check=True).stdout.split()]
    want = [C.sizeof(G.MRSolveCfg), C.sizeof(G.MRRefineCfg), C.sizeof(G.MRStepCfg),
            G.MRStepCfg.solve.offset, G.MRStepCfg.refine.offset, C.sizeof(G.MRSolveDetail),
            G.MRSolveCfg.mode.offset, G.MRSolveCfg.coarse_preconditioner_max_work_ratio.offset]
    assert got == want


@pytest.mark.parametrize("dims", [(16, 16, 16), (8, 12, 8), (17, 10, 9), (24, 40, 20)])
def test_multires_layout_helpers_match_the_oracle(dims):
    # host-only layout builders: layoutWithFineBox for several boxes, and the bubble
    # band (two refineFineCellBox calls) composed from pf_mr_layout_fine_box and
    # pf_mr_layout_refine, leaf for leaf against the oracle's layouts
    import pyoracle as O
    nx, ny, nz = dims
    for box in (None, [0, 0, 0, nx, ny, nz], [1, 2, 3, nx // 2 + 1, ny - 2, nz // 3 + 2], [5, 5, 5, 6, 6, 6]):
        assert np.array_equal(G.mr_layout_fine_box(nx, ny, nz, box), O.mr_layout_fine_box(nx, ny, nz, box))
    water = ny // 2
    band = G.mr_layout_fine_box(nx, ny, nz, [0, 0, 0, nx, min(water + 2, ny), nz])
    bubble = [nx // 2 - nx // 4, water // 4, nz // 2 - nz // 4, nx // 2 + nx // 4, water, nz // 2 + nz // 4]
    got = G.mr_layout_refine(nx, ny, nz, band, bubble)
    assert np.array_equal(got, O.mr_layout_bubble_band(nx, ny, nz))
    # an empty or out-of-domain box leaves the layout as it was
    assert np.array_equal(G.mr_layout_refine(nx, ny, nz, band, [nx, 0, 0, nx + 4, 4, 4]), band)


def test_multires_solve_defaults_are_the_reference_defaults():
    # MRSim3DTP (multires_sim3d_tp.h:34-60) and MRPressureSolveConfig3D
    # (multires_pressure3d.h:137-166) defaults, host-only
    c = G.mr_solve_cfg()
    assert (c.max_iterations, c.absolute_tolerance, c.relative_tolerance) == (160, 1e-7, 1e-6)
    assert (c.mode, c.use_jacobi_preconditioner, c.use_flexible_cg_beta, c.adaptive_restart) == (0, 1, 0, 1)
    assert (c.restart_growth_threshold, c.relaxation_sweeps, c.relaxation_omega, c.relaxation_min_omega) == \
        (10.0, 0, 0.67, 0.05)
    assert (c.use_coarse_correction, c.coarse_correction_iterations, c.coarse_correction_sweeps) == (0, 32, 1)
    assert (c.coarse_correction_relative_tolerance, c.coarse_correction_min_scale) == (1e-3, 1.0 / 64.0)
    assert (c.use_coarse_preconditioner, c.coarse_preconditioner_iterations,
            c.coarse_preconditioner_relative_tolerance, c.coarse_preconditioner_scale) == (0, 8, 1e-2, 1.0)
    assert (c.coarse_preconditioner_auto_disable, c.coarse_preconditioner_auto_disable_after) == (0, 4)
    with pytest.raises(TypeError):
        G.mr_solve_cfg(no_such_option=1)


def test_adaptivity_defaults_are_the_reference_defaults():
    # sparse_sim3d_tp.h:29-60: every stage off, radius 2, targets 4, interface radius 1
    import pyoracle as O
    g = G.adapt_cfg()
    o = O.adapt_cfg()
    assert bytes(g) == bytes(o)
    assert g.gas_particle_coarsening_seed == 0x9E3779B9 and g.liquid_particle_refill_seed == 0x6A09E667
    assert not (g.narrow_band_air or g.gas_particle_coarsening or g.liquid_particle_coarsening
                or g.liquid_particle_refill)
