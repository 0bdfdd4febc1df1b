"""Python binding of the MI355X sparse two-phase PF-FLIP step (include/pf_gpu.h).

Thin ctypes layer over ``libpfflip_hip.so`` (built in-tree by
``adaptive-pf-flip_amd/build.py``).  There is no CPU fallback: if the HIP library
is missing or fails to load, importing ``lib()`` raises.

Naming follows the reference (lsfs): ``SparseSim3DTP`` keeps the fields of
src/driver/sparse_sim3d_tp.h:8-177 that ``step()`` reads, and the field ids are the
members of SparseMacGrid3D<4> (src/grid/sparse_mac_grid3d.h:9-11).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpfflip_hip.so")

FIELD_U, FIELD_V, FIELD_W, FIELD_MU, FIELD_MV, FIELD_MW, FIELD_P, FIELD_MARKER = range(8)
FIELD_SAVED_U, FIELD_SAVED_V, FIELD_SAVED_W = 8, 9, 10
STOP_REF_MAXABS, STOP_REL2 = 0, 1
SCENE_RT, SCENE_BUBBLE, SCENE_DAMBREAK = 0, 1, 2
STAGES = ("bin", "mark", "p2g", "grid", "assemble", "solve", "velocity", "g2p_advect")
KERNELS = ("bin_count", "bin_scatter", "p2g", "face_ops", "assemble", "pcg_spmv", "pcg_update",
           "residual64", "velocity", "g2p_advect", "lpc_apply", "interface", "surface", "mg_apply")
(K_BIN_COUNT, K_BIN_SCATTER, K_P2G, K_FACE_OPS, K_ASSEMBLE, K_PCG_SPMV, K_PCG_UPDATE,
 K_RESIDUAL64, K_VELOCITY, K_G2P_ADVECT, K_LPC_APPLY) = range(11)


class GridDesc(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("reserved", C.c_int32),
                ("dx", C.c_double), ("ox", C.c_double), ("oy", C.c_double), ("oz", C.c_double)]


class Phase(C.Structure):
    _fields_ = [("rho_l", C.c_double), ("rho_g", C.c_double), ("alpha_phi", C.c_double),
                ("rho_tilde_0", C.c_double)]


class SolveCfg(C.Structure):
    _fields_ = [("max_iters", C.c_int32), ("stop_rule", C.c_int32), ("tol", C.c_double),
                ("inner_rel_tol", C.c_double), ("max_outer", C.c_int32), ("warm_start", C.c_int32),
                ("check_every", C.c_int32), ("precond", C.c_int32)]


class SolveStats(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("outer_iterations", C.c_int32),
                ("converged", C.c_int32), ("breakdown", C.c_int32), ("pinned", C.c_int32),
                ("fallback_used", C.c_int32), ("reserved0", C.c_int32), ("reserved1", C.c_int32),
                ("fluid_cells", C.c_int64), ("pinned_cell", C.c_int64),
                ("initial_max_residual", C.c_double), ("final_max_residual", C.c_double),
                ("true_rel_residual", C.c_double), ("true_max_residual", C.c_double),
                ("rhs_norm2", C.c_double), ("ms", C.c_double)]


class EscapeStats(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("clamped_liquid", "clamped_gas", "x_lo", "x_hi", "y_lo",
                                          "y_hi", "z_lo", "z_hi")]


MODEL_TWO_PHASE, MODEL_SINGLE_PHASE = 0, 1


MR_SOLVE_FAST, MR_SOLVE_REFERENCE, MR_SOLVE_MULTIGRID = 0, 1, 2


class MRSolveCfg(C.Structure):
    _fields_ = [("max_iterations", C.c_int32), ("check_every", C.c_int32), ("absolute_tolerance", C.c_double),
                ("relative_tolerance", C.c_double), ("inner_rel_tol", C.c_double), ("max_outer", C.c_int32),
                ("mode", C.c_int32), ("use_jacobi_preconditioner", C.c_int32), ("use_flexible_cg_beta", C.c_int32),
                ("adaptive_restart", C.c_int32), ("relaxation_sweeps", C.c_int32),
                ("restart_growth_threshold", C.c_double), ("relaxation_omega", C.c_double),
                ("relaxation_min_omega", C.c_double), ("use_coarse_correction", C.c_int32),
                ("coarse_correction_iterations", C.c_int32), ("coarse_correction_sweeps", C.c_int32),
                ("use_coarse_preconditioner", C.c_int32), ("coarse_correction_absolute_tolerance", C.c_double),
                ("coarse_correction_relative_tolerance", C.c_double), ("coarse_correction_min_scale", C.c_double),
                ("coarse_preconditioner_iterations", C.c_int32), ("coarse_preconditioner_auto_disable", C.c_int32),
                ("coarse_preconditioner_auto_disable_after", C.c_int32), ("reserved", C.c_int32),
                ("coarse_preconditioner_absolute_tolerance", C.c_double),
                ("coarse_preconditioner_relative_tolerance", C.c_double),
                ("coarse_preconditioner_scale", C.c_double), ("coarse_preconditioner_min_rz_gain", C.c_double),
                ("coarse_preconditioner_max_work_ratio", C.c_double)]


class MRSolveDetail(C.Structure):
    _fields_ = [(n, C.c_int32) for n in (
        "reference_order", "restarts", "beta_resets", "relaxation_accepted", "relaxation_rejected",
        "coarse_correction_cells", "coarse_correction_iterations", "coarse_correction_sweeps",
        "coarse_correction_accepted_sweeps", "coarse_correction_rejected_sweeps", "coarse_correction_accepted",
        "coarse_correction_converged", "coarse_correction_breakdown", "coarse_preconditioner_cells",
        "coarse_preconditioner_applications", "coarse_preconditioner_accepted_applications",
        "coarse_preconditioner_rejected_applications", "coarse_preconditioner_skipped_applications",
        "coarse_preconditioner_budget_limited_applications", "coarse_preconditioner_iterations",
        "coarse_preconditioner_auto_disabled", "coarse_preconditioner_auto_disable_iteration",
        "coarse_preconditioner_breakdown", "coarse_preconditioner_budget_exhausted")] + [
        (n, C.c_double) for n in (
            "initial_residual", "final_residual", "effective_tolerance", "relaxation_final_omega",
            "coarse_correction_effective_tolerance", "coarse_correction_initial_residual",
            "coarse_correction_final_residual", "coarse_correction_last_scale",
            "coarse_preconditioner_effective_tolerance", "coarse_preconditioner_last_rz_gain")]


def mr_solve_cfg(max_iterations=160, abs_tol=1e-7, rel_tol=1e-6, **options):
    """MRSim3DTP's solver defaults (multires_sim3d_tp.h:34-60: cg_iters 160, cg_tol 1e-7,
    cg_rel_tol 1e-6, ...) from pf_mr_solve_defaults; `options` set fields by name (for
    example mode=MR_SOLVE_REFERENCE, use_coarse_preconditioner=1)."""
    c = MRSolveCfg()
    lib().pf_mr_solve_defaults(C.byref(c))
    c.max_iterations, c.absolute_tolerance, c.relative_tolerance = max_iterations, abs_tol, rel_tol
    for k, v in options.items():
        if k not in dict(MRSolveCfg._fields_):
            raise TypeError(f"unknown multires solve option {k}")
        setattr(c, k, v)
    return c


class MRRefineCfg(C.Structure):
    _fields_ = [("dynamic_refinement", C.c_int32), ("particle_padding", C.c_int32), ("gas_padding", C.c_int32),
                ("hysteresis_cells", C.c_int32), ("max_fine_leaves", C.c_int32), ("reserved", C.c_int32)]


class AdaptCfg(C.Structure):
    """pf_adapt_cfg: the particle adaptivity fields of SparseSim3DTP / MRSim3DTP
    (sparse_sim3d_tp.h:29-60), under their names"""
    _fields_ = [("narrow_band_air", C.c_int32), ("narrow_band_air_radius", C.c_int32),
                ("gas_particle_coarsening", C.c_int32), ("gas_particles_per_cell_target", C.c_int32),
                ("gas_particle_coarsening_seed", C.c_uint32),
                ("liquid_particle_coarsening", C.c_int32), ("liquid_particles_per_cell_target", C.c_int32),
                ("liquid_particle_coarsening_seed", C.c_uint32),
                ("liquid_particle_refill", C.c_int32), ("liquid_particle_refill_interface_only", C.c_int32),
                ("liquid_particle_refill_interface_radius", C.c_int32),
                ("liquid_particle_refill_max_added_per_step", C.c_int32),
                ("liquid_refill_particles_per_cell_target", C.c_int32), ("liquid_particle_refill_seed", C.c_uint32)]


ADAPT_FIELDS = [n for n, _ in AdaptCfg._fields_]


class MgCfg(C.Structure):
    """pf_mg_cfg: the cycle of the multigrid preconditioner (PRECOND_MULTIGRID)"""
    _fields_ = [("sweeps", C.c_int32), ("coarsest_sweeps", C.c_int32), ("omega", C.c_float),
                ("correction", C.c_float), ("chebyshev_ratio", C.c_float)]


def mg_defaults() -> MgCfg:
    """the built-in cycle, which every context starts with"""
    c = MgCfg()
    lib().pf_mg_defaults(C.byref(c))
    return c


def adapt_cfg(**kw) -> "AdaptCfg":
    """pf_adapt_defaults (the reference's defaults: every stage off), then kw"""
    c = AdaptCfg()
    lib().pf_adapt_defaults(C.byref(c))
    for k, v in kw.items():
        if k not in ADAPT_FIELDS:
            raise KeyError(k)
        setattr(c, k, v)
    return c


class AdaptStats(C.Structure):
    """pf_adapt_stats: the adaptivity results (particle_adaptivity3d.h:50-73)"""
    _fields_ = [(n, C.c_int32) for n in (
        "narrow_band_air_removed", "narrow_band_air_liquid_cells", "narrow_band_air_gas_before",
        "narrow_band_air_gas_after", "gas_removed", "gas_cells", "gas_overfull_cells", "gas_before", "gas_after",
        "liquid_removed", "liquid_cells", "liquid_overfull_cells", "liquid_before", "liquid_after",
        "refill_added", "refill_cells", "refill_interface_cells", "refill_underfull_cells",
        "refill_budget_limited", "refill_before", "refill_after", "reserved")]

    def values(self):
        return [getattr(self, n) for n, _ in self._fields_ if n != "reserved"]


class RenderCamera(C.Structure):
    """pf_render_camera: RenderCacheCamera3D (render_cache3d.h:19-27)"""
    _fields_ = [("position", C.c_double * 3), ("target", C.c_double * 3), ("up", C.c_double * 3),
                ("fov_degrees", C.c_double), ("near_clip", C.c_double), ("far_clip", C.c_double),
                ("focal_length_mm", C.c_double)]


class RenderFrame(C.Structure):
    """pf_render_frame: frame, time, dt (effective_dt_last), Vp, phase, camera"""
    _fields_ = [("frame", C.c_int32), ("reserved", C.c_int32), ("time", C.c_double), ("dt", C.c_double),
                ("Vp", C.c_double), ("phase", Phase), ("camera", RenderCamera)]


class RenderManifestFrame(C.Structure):
    """pf_render_manifest_frame: RenderCacheManifestFrame3D (render_cache3d.h:39-45)"""
    _fields_ = [("frame", C.c_int32), ("step", C.c_int32), ("time", C.c_double), ("path", C.c_char_p),
                ("bytes", C.c_int64)]


def default_render_camera(nx, ny, nz, dx) -> RenderCamera:
    """defaultRenderCacheCamera3D (render_cache3d.h:48-62)"""
    cam = RenderCamera()
    lib().pf_render_default_camera(nx, ny, nz, dx, C.byref(cam))
    return cam


def write_render_manifest(path, sim_kind, nx, ny, nz, dx, frames):
    """writeRenderCacheManifest3D; frames: (frame, step, time, path, bytes) tuples"""
    arr = (RenderManifestFrame * max(len(frames), 1))(
        *[RenderManifestFrame(f[0], f[1], f[2], str(f[3]).encode(), f[4]) for f in frames])
    st = lib().pf_write_render_manifest(str(path).encode(), sim_kind.encode(), nx, ny, nz, dx, arr, len(frames))
    if st != 0:
        raise PfError(f"pf_write_render_manifest: status {st}: invalid render cache manifest")


class StepCfg(C.Structure):
    _fields_ = [("dt", C.c_double), ("gravity", C.c_double), ("Vp", C.c_double),
                ("alpha_liquid", C.c_double), ("alpha_gas", C.c_double),
                ("adaptive_cfl", C.c_double), ("adaptive_min_dt", C.c_double),
                ("c_div", C.c_double), ("adaptive_timestep", C.c_int32),
                ("advection_order", C.c_int32), ("timing", C.c_int32), ("model", C.c_int32),
                ("phase", Phase), ("solve", SolveCfg),
                ("surface_tension", C.c_int32), ("surface_tension_smoothing_radius", C.c_int32),
                ("surface_tension_strength", C.c_double), ("surface_tension_max_delta_speed", C.c_double),
                ("skip_interface_diagnostics", C.c_int32), ("reserved", C.c_int32),
                ("c_div_volume_correction", C.c_int32), ("reserved2", C.c_int32),
                ("c_div_strength", C.c_double), ("liquid_volume_target", C.c_double), ("adapt", AdaptCfg)]


class MRStepCfg(C.Structure):
    _fields_ = [("dt", C.c_double), ("gravity", C.c_double), ("Vp", C.c_double),
                ("alpha_liquid", C.c_double), ("alpha_gas", C.c_double),
                ("adaptive_cfl", C.c_double), ("adaptive_min_dt", C.c_double),
                ("c_div", C.c_double), ("adaptive_timestep", C.c_int32),
                ("advection_order", C.c_int32), ("timing", C.c_int32), ("reserved", C.c_int32),
                ("phase", Phase), ("solve", MRSolveCfg), ("refine", MRRefineCfg),
                ("surface_tension", C.c_int32), ("surface_tension_smoothing_radius", C.c_int32),
                ("surface_tension_strength", C.c_double), ("surface_tension_max_delta_speed", C.c_double),
                ("skip_interface_diagnostics", C.c_int32), ("reserved2", C.c_int32),
                ("c_div_volume_correction", C.c_int32), ("reserved3", C.c_int32),
                ("c_div_strength", C.c_double), ("liquid_volume_target", C.c_double), ("adapt", AdaptCfg)]


class IfaceStats(C.Structure):
    """InterfaceDiagnostics3D, src/driver/interface_diagnostics3d.h:7-19"""
    _fields_ = [(n, C.c_int32) for n in ("sample_cells", "interface_cells", "finite",
                                         "surface_tension_candidate")] + \
               [(n, C.c_double) for n in ("phi_min", "phi_max", "phi_mean", "grad_mean", "grad_max",
                                          "curvature_abs_mean", "curvature_abs_max")]


class SurfaceStats(C.Structure):
    """SurfaceTensionStats3D, src/driver/interface_diagnostics3d.h:21-36"""
    _fields_ = [(n, C.c_int32) for n in ("enabled", "applied_cells", "finite", "curvature_smoothing_radius",
                                         "capillary_stable", "reserved")] + \
               [(n, C.c_double) for n in ("strength", "max_delta_speed_limit", "mean_delta_speed",
                                          "max_delta_speed", "raw_curvature_abs_mean", "raw_curvature_abs_max",
                                          "smoothed_curvature_abs_mean", "smoothed_curvature_abs_max",
                                          "capillary_dt_limit")]


class StepStats(C.Structure):
    _fields_ = [("effective_dt", C.c_double), ("max_particle_speed", C.c_double),
                ("cfl_limit_dt", C.c_double), ("limited", C.c_int32), ("reserved", C.c_int32),
                ("solve", SolveStats), ("escape", EscapeStats), ("stage_ms", C.c_double * 8),
                ("total_ms", C.c_double), ("iface", IfaceStats), ("surface", SurfaceStats),
                ("liquid_volume", C.c_double), ("c_div", C.c_double), ("adapt", AdaptStats)]


class PfError(RuntimeError):
    pass


_lib = None


def lib():
    """Load libpfflip_hip.so (raises if it is missing: there is no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PfError(f"{LIB_PATH} is missing; run adaptive-pf-flip_amd/build.py")
        L = C.CDLL(LIB_PATH)
        vp, i32, i64, dbl = C.c_void_p, C.c_int32, C.c_int64, C.c_double
        sig = {
            "pf_create": (C.c_int, [C.POINTER(vp), C.POINTER(GridDesc)]),
            "pf_destroy": (C.c_int, [vp]),
            "pf_last_error": (C.c_char_p, [vp]),
            "pf_abi_version": (C.c_int, []),
            "pf_synchronize": (C.c_int, [vp]),
            "pf_particles_upload": (C.c_int, [vp, i64] + [vp] * 8),
            "pf_particles_download": (C.c_int, [vp] + [vp] * 6),
            "pf_particle_count": (i64, [vp]),
            "pf_scene_generate": (C.c_int, [vp, i32, C.c_uint64]),
            "pf_mark_cells": (C.c_int, [vp]),
            "pf_bin_counts": (C.c_int, [vp, vp]),
            "pf_color_lists": (C.c_int, [vp, vp, i64, vp]),
            "pf_face_phi": (C.c_int, [vp, C.POINTER(Phase), i32, vp]),
            "pf_p2g": (C.c_int, [vp, C.POINTER(Phase), dbl]),
            "pf_save_velocity": (C.c_int, [vp]),
            "pf_apply_gravity": (C.c_int, [vp, dbl, dbl]),
            "pf_apply_walls": (C.c_int, [vp]),
            "pf_project": (C.c_int, [vp, C.POINTER(Phase), dbl, C.POINTER(SolveCfg), dbl,
                                     C.POINTER(SolveStats)]),
            "pf_g2p": (C.c_int, [vp, dbl, dbl]),
            "pf_advect": (C.c_int, [vp, dbl, i32, C.POINTER(EscapeStats)]),
            "pf_timestep": (C.c_int, [vp, dbl, i32, dbl, dbl, C.POINTER(dbl), C.POINTER(dbl),
                                      C.POINTER(dbl), C.POINTER(i32)]),
            "pf_step": (C.c_int, [vp, C.POINTER(StepCfg), C.POINTER(StepStats)]),
            "pf_interface_diagnostics": (C.c_int, [vp, C.POINTER(Phase), C.POINTER(IfaceStats)]),
            "pf_apply_surface_tension": (C.c_int, [vp, C.POINTER(Phase), dbl, dbl, dbl, i32,
                                                   C.POINTER(SurfaceStats)]),
            "pf_mr_interface_diagnostics": (C.c_int, [vp, C.POINTER(Phase), C.POINTER(IfaceStats)]),
            "pf_particle_adaptivity": (C.c_int, [vp, C.POINTER(AdaptCfg), C.POINTER(AdaptStats)]),
            "pf_render_default_camera": (None, [i32, i32, i32, dbl, C.POINTER(RenderCamera)]),
            "pf_write_render_cache": (C.c_int, [vp, C.c_char_p, C.POINTER(RenderFrame), C.POINTER(i64)]),
            "pf_write_render_manifest": (C.c_int, [C.c_char_p, C.c_char_p, i32, i32, i32, dbl, vp, i64]),
            "pf_adapt_defaults": (None, [C.POINTER(AdaptCfg)]),
            "pf_mg_defaults": (None, [C.POINTER(MgCfg)]),
            "pf_set_multigrid": (C.c_int, [vp, C.POINTER(MgCfg)]),
            "pf_mg_set_correction_weights": (C.c_int, [vp, vp, i64]),
            "pf_get_multigrid": (C.c_int, [vp, C.POINTER(MgCfg)]),
            "pf_particles_download_ext": (C.c_int, [vp, vp, vp]),
            "pf_liquid_volume": (dbl, [vp]),
            "pf_mr_apply_surface_tension": (C.c_int, [vp, C.POINTER(Phase), dbl, dbl, dbl, i32,
                                                      C.POINTER(SurfaceStats)]),
            "pf_field_dims": (C.c_int, [vp, i32, vp]),
            "pf_field_download": (C.c_int, [vp, i32, vp]),
            "pf_field_upload": (C.c_int, [vp, i32, vp, vp]),
            "pf_active_blocks": (C.c_int, [vp, i32, vp, i64, C.POINTER(i64)]),
            "pf_calibrate_rho_tilde0": (dbl, [C.POINTER(Phase), dbl]),
            "pf_phi_from_raw": (dbl, [dbl, C.POINTER(Phase)]),
            "pf_beta_from_phi": (dbl, [dbl, C.POINTER(Phase)]),
            "pf_p2g_single": (C.c_int, [vp]),
            "pf_project_single": (C.c_int, [vp, dbl, C.POINTER(SolveCfg), C.POINTER(SolveStats)]),
            "pf_stream": (vp, [vp]),
            "pf_mr_set_layout": (C.c_int, [vp, i64, vp]),
            "pf_mr_get_layout": (C.c_int, [vp, vp, i64, C.POINTER(i64)]),
            "pf_mr_layout_fine_box": (i64, [i32, i32, i32, i32, i32, vp, vp, i64]),
            "pf_mr_face_valid": (C.c_int, [vp, i32, vp]),
            "pf_mr_mark_cells": (C.c_int, [vp]),
            "pf_mr_p2g": (C.c_int, [vp, C.POINTER(Phase), dbl]),
            "pf_mr_project": (C.c_int, [vp, C.POINTER(Phase), dbl, C.POINTER(MRSolveCfg), dbl,
                                         C.POINTER(SolveStats)]),
            "pf_mr_update_refinement": (C.c_int, [vp, C.POINTER(MRRefineCfg), C.POINTER(i32), C.POINTER(i32)]),
            "pf_mr_layout_refine": (i64, [i32, i32, i32, vp, i64, vp, vp, i64]),
            "pf_mr_scene_bubbles": (C.c_int, [vp, C.c_uint64, i32, i32, vp]),
            "pf_mr_retained_box": (C.c_int, [vp, vp, C.POINTER(i32)]),
            "pf_mr_step": (C.c_int, [vp, C.POINTER(MRStepCfg), C.POINTER(StepStats)]),
            "pf_mr_solve_defaults": (None, [C.POINTER(MRSolveCfg)]),
            "pf_mr_last_solve_detail": (C.c_int, [vp, C.POINTER(MRSolveDetail)]),
            "pf_lpc_create": (C.c_int, [C.POINTER(vp), C.c_uint64]),
            "pf_lpc_destroy": (C.c_int, [vp]),
            "pf_lpc_last_error": (C.c_char_p, [vp]),
            "pf_lpc_num_weights": (i64, []),
            "pf_lpc_get_weights": (C.c_int, [vp, vp, i64]),
            "pf_lpc_set_weights": (C.c_int, [vp, vp, i64]),
            "pf_lpc_forward_tiles": (C.c_int, [vp, i64, vp, vp]),
            "pf_lpc_grad_tiles": (C.c_int, [vp, i64, vp, vp, vp]),
            "pf_lpc_attach": (C.c_int, [vp, vp]),
            "pf_lpc_record": (C.c_int, [vp, i32, i32]),
            "pf_lpc_snapshot_count": (i32, [vp]),
            "pf_lpc_train": (C.c_int, [vp, vp, i32, dbl, C.c_uint64, C.POINTER(dbl)]),
            "pf_lpc_loss": (C.c_int, [vp, vp, i32, C.POINTER(dbl), C.POINTER(dbl)]),
            "pf_lpc_grad": (C.c_int, [vp, vp, i32, vp, C.POINTER(dbl)]),
            "pf_lpc_set_objective": (C.c_int, [vp, i32, i32]),
            "pf_kernel_timing": (C.c_int, [vp, i32]),
            "pf_kernel_timing_mask": (C.c_int, [vp, C.c_uint32]),
            "pf_kernel_time": (C.c_int, [vp, i32, C.POINTER(i64), C.POINTER(dbl)]),
            "pf_copy_probe": (C.c_int, [vp, i64, C.POINTER(dbl)]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def calibrate_rho_tilde0(phase: Phase, Vp: float = 1.0) -> float:
    return lib().pf_calibrate_rho_tilde0(C.byref(phase), Vp)


PRECOND_JACOBI, PRECOND_LEARNED, PRECOND_MULTIGRID = 0, 1, 2
LPC_OBJECTIVE_RESIDUAL, LPC_OBJECTIVE_ENERGY = 0, 1


def default_solve_cfg(stop_rule=STOP_REF_MAXABS, tol=1e-7, max_iters=600, precond=PRECOND_JACOBI) -> SolveCfg:
    return SolveCfg(max_iters=max_iters, stop_rule=stop_rule, tol=tol, inner_rel_tol=0.1,
                    max_outer=10000, warm_start=0, check_every=8, precond=precond)


class Context:
    """One grid + one particle set on the current HIP device (a pf_ctx)."""

    def __init__(self, nx, ny, nz, dx=1.0, origin=(0.0, 0.0, 0.0)):
        self.nx, self.ny, self.nz, self.dx = nx, ny, nz, dx
        self.h = C.c_void_p()
        d = GridDesc(nx, ny, nz, 0, dx, *origin)
        st = lib().pf_create(C.byref(self.h), C.byref(d))
        if st != 0:
            raise PfError(f"pf_create failed with status {st}")

    def close(self):
        if self.h:
            lib().pf_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st, what):
        if st != 0:
            msg = lib().pf_last_error(self.h)
            raise PfError(f"{what}: status {st}: {msg.decode() if msg else ''}")

    # particles
    def upload(self, x, y, z, u=None, v=None, w=None, type=None, volume=None):
        f = lambda a: np.ascontiguousarray(a, np.float32)
        x, y, z = f(x), f(y), f(z)
        n = len(x)
        u = f(u if u is not None else np.zeros(n)); v = f(v if v is not None else np.zeros(n))
        w = f(w if w is not None else np.zeros(n))
        t = np.ascontiguousarray(type if type is not None else np.zeros(n), np.uint8)
        vol = None if volume is None else f(volume)
        self._keep = (x, y, z, u, v, w, t, vol)
        self._check(lib().pf_particles_upload(self.h, n, _p(x), _p(y), _p(z), _p(u), _p(v), _p(w),
                                              _p(t), _p(vol)), "pf_particles_upload")

    def download(self):
        n = self.count()
        a = [np.zeros(n, np.float32) for _ in range(6)]
        self._check(lib().pf_particles_download(self.h, *[_p(q) for q in a]), "pf_particles_download")
        return a

    def download_ext(self):
        """(types, volumes) in the caller's order"""
        n = self.count()
        t = np.zeros(n, np.uint8)
        vol = np.zeros(n, np.float32)
        self._check(lib().pf_particles_download_ext(self.h, _p(t), _p(vol)), "pf_particles_download_ext")
        return t, vol

    def count(self):
        return lib().pf_particle_count(self.h)

    def liquid_volume(self):
        """the summed volume of the liquid particles (liquidVolume with Vp = 1)"""
        return lib().pf_liquid_volume(self.h)

    def write_render_cache(self, path, info: RenderFrame) -> int:
        """writeSparseRenderCache3D / writeMRRenderCache3D of this context; returns the bytes"""
        nb = C.c_int64()
        self._check(lib().pf_write_render_cache(self.h, str(path).encode(), C.byref(info), C.byref(nb)),
                    "pf_write_render_cache")
        return nb.value

    def set_multigrid(self, cfg: MgCfg = None, **kw):
        """the multigrid cycle of the following PRECOND_MULTIGRID solves (cfg, or the
        current one with the fields in kw changed)"""
        c = cfg if cfg is not None else self.multigrid()
        for k, v in kw.items():
            if k not in [n for n, _ in MgCfg._fields_]:
                raise KeyError(k)
            setattr(c, k, v)
        self._check(lib().pf_set_multigrid(self.h, C.byref(c)), "pf_set_multigrid")

    def set_correction_weights(self, w):
        """per-aggregate weights of the level-1 coarse correction, dense (az, ay, ax)
        over ceil(n/2) aggregates per axis, or None for the plain cycle"""
        if w is None:
            self._check(lib().pf_mg_set_correction_weights(self.h, None, 0), "pf_mg_set_correction_weights")
            return
        a = np.ascontiguousarray(w, np.float32).reshape(-1)
        self._check(lib().pf_mg_set_correction_weights(self.h, _p(a), len(a)), "pf_mg_set_correction_weights")

    def multigrid(self) -> MgCfg:
        c = MgCfg()
        self._check(lib().pf_get_multigrid(self.h, C.byref(c)), "pf_get_multigrid")
        return c

    def particle_adaptivity(self, cfg: AdaptCfg) -> AdaptStats:
        """applyParticleAdaptivity on the context's particles"""
        st = AdaptStats()
        self._check(lib().pf_particle_adaptivity(self.h, C.byref(cfg), C.byref(st)), "pf_particle_adaptivity")
        return st

    def scene(self, kind, seed):
        self._check(lib().pf_scene_generate(self.h, kind, seed), "pf_scene_generate")

    # stages
    def mark_cells(self):
        self._check(lib().pf_mark_cells(self.h), "pf_mark_cells")

    def p2g(self, phase, Vp=1.0):
        self._check(lib().pf_p2g(self.h, C.byref(phase), Vp), "pf_p2g")

    def p2g_single(self):
        """spP2G3D (single-phase trilinear splat, unit mass)."""
        self._check(lib().pf_p2g_single(self.h), "pf_p2g_single")

    def project_single(self, dt, cfg=None):
        """spProjectStep3D (single-phase, rho = 1, no pinned cell)."""
        cfg = cfg or default_solve_cfg(STOP_REF_MAXABS, 1e-6, 300)
        st = SolveStats()
        self._check(lib().pf_project_single(self.h, dt, C.byref(cfg), C.byref(st)), "pf_project_single")
        return st

    def save_velocity(self):
        self._check(lib().pf_save_velocity(self.h), "pf_save_velocity")

    def apply_gravity(self, dt, g):
        self._check(lib().pf_apply_gravity(self.h, dt, g), "pf_apply_gravity")

    def apply_walls(self):
        self._check(lib().pf_apply_walls(self.h), "pf_apply_walls")

    def project(self, phase, dt, cfg=None, div_correction=0.0):
        cfg = cfg or default_solve_cfg()
        st = SolveStats()
        self._check(lib().pf_project(self.h, C.byref(phase), dt, C.byref(cfg), div_correction,
                                     C.byref(st)), "pf_project")
        return st

    def g2p(self, aL, aG):
        self._check(lib().pf_g2p(self.h, aL, aG), "pf_g2p")

    def advect(self, dt, order=2):
        es = EscapeStats()
        self._check(lib().pf_advect(self.h, dt, order, C.byref(es)), "pf_advect")
        return es

    def timestep(self, dt, enabled, cfl, min_dt):
        e, v, c = C.c_double(), C.c_double(), C.c_double()
        lim = C.c_int32()
        self._check(lib().pf_timestep(self.h, dt, int(enabled), cfl, min_dt, C.byref(e), C.byref(v),
                                      C.byref(c), C.byref(lim)), "pf_timestep")
        return e.value, v.value, c.value, lim.value

    def step(self, cfg: StepCfg):
        st = StepStats()
        self._check(lib().pf_step(self.h, C.byref(cfg), C.byref(st)), "pf_step")
        return st

    def interface_diagnostics(self, phase: Phase) -> IfaceStats:
        """diagnoseSparseInterface3D over the current masses and markers"""
        out = IfaceStats()
        self._check(lib().pf_interface_diagnostics(self.h, C.byref(phase), C.byref(out)),
                    "pf_interface_diagnostics")
        return out

    def apply_surface_tension(self, phase: Phase, dt, strength=0.02, max_delta_speed=0.25,
                              smoothing_radius=0) -> SurfaceStats:
        """applySparseSurfaceTension3D on the current face velocities"""
        out = SurfaceStats()
        self._check(lib().pf_apply_surface_tension(self.h, C.byref(phase), dt, strength, max_delta_speed,
                                                   int(smoothing_radius), C.byref(out)),
                    "pf_apply_surface_tension")
        return out

    def synchronize(self):
        self._check(lib().pf_synchronize(self.h), "pf_synchronize")

    # measurement hooks
    def kernel_timing(self, enable: bool, kernels=None):
        """event-time every kernel (enable) or only the named ones (KERNELS names)"""
        if kernels is None:
            self._check(lib().pf_kernel_timing(self.h, int(enable)), "pf_kernel_timing")
            return
        mask = 0
        for k in kernels:
            mask |= 1 << KERNELS.index(k)
        self._check(lib().pf_kernel_timing_mask(self.h, mask if enable else 0), "pf_kernel_timing_mask")

    def kernel_time(self, kernel: int):
        """(launches, summed ms) of one PF_K_* kernel since timing was enabled."""
        n, ms = C.c_int64(), C.c_double()
        self._check(lib().pf_kernel_time(self.h, kernel, C.byref(n), C.byref(ms)), "pf_kernel_time")
        return n.value, ms.value

    def copy_probe(self, n: int) -> float:
        ms = C.c_double()
        self._check(lib().pf_copy_probe(self.h, n, C.byref(ms)), "pf_copy_probe")
        return ms.value

    # fields
    def dims(self, field):
        d = np.zeros(3, np.int32)
        self._check(lib().pf_field_dims(self.h, field, _p(d)), "pf_field_dims")
        return tuple(int(q) for q in d)

    def field(self, field):
        nx, ny, nz = self.dims(field)
        out = np.zeros(nx * ny * nz, np.float32)
        self._check(lib().pf_field_download(self.h, field, _p(out)), "pf_field_download")
        return out.reshape(nz, ny, nx)

    def set_field(self, field, dense, block_active=None):
        d = np.ascontiguousarray(dense, np.float32).reshape(-1)
        ba = None if block_active is None else np.ascontiguousarray(block_active, np.uint8)
        self._check(lib().pf_field_upload(self.h, field, _p(d), _p(ba)), "pf_field_upload")

    def bin_counts(self):
        """particles per cell from the binning, (nz, ny, nx)"""
        out = np.zeros(self.nx * self.ny * self.nz, np.uint32)
        self._check(lib().pf_bin_counts(self.h, _p(out)), "pf_bin_counts")
        return out.reshape(self.nz, self.ny, self.nx)

    def color_lists(self):
        """the eight P2G colour lists of source blocks, each ascending"""
        cnt = np.zeros(8, np.int64)
        self._check(lib().pf_color_lists(self.h, None, 0, _p(cnt)), "pf_color_lists")
        flat = np.zeros(max(int(cnt.sum()), 1), np.int32)
        self._check(lib().pf_color_lists(self.h, _p(flat), len(flat), _p(cnt)), "pf_color_lists")
        off = np.concatenate([[0], np.cumsum(cnt)])
        return [flat[off[k]:off[k + 1]] for k in range(8)]

    def face_phi(self, phase, axis):
        """FP64 phi of the current face masses of one axis, computed on the device"""
        nx, ny, nz = self.dims(FIELD_MU + axis)
        out = np.zeros(nx * ny * nz, np.float64)
        self._check(lib().pf_face_phi(self.h, C.byref(phase), axis, _p(out)), "pf_face_phi")
        return out.reshape(nz, ny, nx)

    def active_blocks(self, field):
        n = C.c_int64()
        self._check(lib().pf_active_blocks(self.h, field, None, 0, C.byref(n)), "pf_active_blocks")
        out = np.zeros(max(n.value, 1), np.int32)
        self._check(lib().pf_active_blocks(self.h, field, _p(out), n.value, C.byref(n)),
                    "pf_active_blocks")
        return out[: n.value]


def _init_adaptivity(sim):
    """the particle adaptivity fields of SparseSim3DTP / MRSim3DTP (sparse_sim3d_tp.h:29-66)
    with the reference's defaults, and their *_last / *_total results"""
    d = adapt_cfg()
    for n in ADAPT_FIELDS:
        v = getattr(d, n)
        setattr(sim, n, bool(v) if n in ("narrow_band_air", "gas_particle_coarsening", "liquid_particle_coarsening",
                                          "liquid_particle_refill", "liquid_particle_refill_interface_only") else v)
    sim.adapt_stats_last = AdaptStats()
    for n in ("narrow_band_air_removed_total", "gas_particle_coarsening_removed_total",
              "liquid_particle_coarsening_removed_total", "liquid_particle_refill_added_total"):
        setattr(sim, n, 0)


def _adapt_of(sim):
    return AdaptCfg(**{n: int(getattr(sim, n)) for n in ADAPT_FIELDS})


def _store_step_stats(sim, st):
    """the step's *_last fields: volume correction and adaptivity (the store*Stats
    helpers, sparse_sim3d_tp.cpp:368-379, 778-845)"""
    sim.liquid_volume_current_last = st.liquid_volume
    sim.liquid_volume_error_last = sim.liquid_volume_target - st.liquid_volume
    sim.c_div_last = st.c_div
    a = st.adapt
    sim.adapt_stats_last = a
    sim.narrow_band_air_removed_total += a.narrow_band_air_removed
    sim.gas_particle_coarsening_removed_total += a.gas_removed
    sim.liquid_particle_coarsening_removed_total += a.liquid_removed
    sim.liquid_particle_refill_added_total += a.refill_added


class SparseSim3DTP:
    """GPU mirror of the reference's SparseSim3DTP (src/driver/sparse_sim3d_tp.h:8-177).

    Keeps the fields step() reads under the reference's names; particles live on the
    device between steps.  Stats of the last step land in the *_last fields.
    """

    def __init__(self, nx, ny, nz, dx):
        self.ctx = Context(nx, ny, nz, dx)
        self.nx, self.ny, self.nz, self.dx = nx, ny, nz, dx
        self.phase = Phase(1.0, 0.01, 1.0, 1.0)
        self.dt, self.gravity, self.Vp = 0.02, -9.81, 1.0
        self.alpha_liquid, self.alpha_gas = 0.95, 0.95
        self.adaptive_timestep, self.adaptive_cfl, self.adaptive_min_dt = False, 0.5, 1e-5
        self.advection_order = 2
        self.c_div_volume_correction, self.c_div_strength = False, 1.0
        self.liquid_volume_target = 0.0
        self.cg_iters, self.cg_tol = 600, 1e-7
        self.stop_rule = STOP_REF_MAXABS
        self.inner_rel_tol, self.warm_start, self.check_every = 0.1, 0, 8
        # PRECOND_JACOBI (the reference's), PRECOND_LEARNED or PRECOND_MULTIGRID
        self.precond = PRECOND_JACOBI
        self.timing = False
        # CSF surface tension (sparse_sim3d_tp.h:156-159)
        self.surface_tension = False
        self.surface_tension_strength = 0.02
        self.surface_tension_max_delta_speed = 0.25
        self.surface_tension_curvature_smoothing_radius = 0
        # False skips the per-step interface statistics (the reference always computes them)
        self.interface_diagnostics = True
        _init_adaptivity(self)
        self.effective_dt_last = self.dt
        self.max_particle_speed_last = 0.0
        self.c_div_last = 0.0
        self.liquid_volume_current_last = 0.0
        self.liquid_volume_error_last = 0.0
        self.interface_diagnostics_last = IfaceStats(finite=1)
        self.surface_tension_stats_last = SurfaceStats(finite=1, capillary_stable=1)
        self.stats_last = None

    def set_particles(self, x, y, z, u=None, v=None, w=None, type=None, volume=None):
        self.ctx.upload(x, y, z, u, v, w, type, volume)

    def calibrate(self):
        """resetSceneState's rho_tilde_0 = calibrateRhoTilde0(phase, Vp), sparse_sim3d_tp.cpp:396"""
        self.phase.rho_tilde_0 = calibrate_rho_tilde0(self.phase, self.Vp)

    def step_cfg(self) -> StepCfg:
        solve = SolveCfg(max_iters=self.cg_iters, stop_rule=self.stop_rule, tol=self.cg_tol,
                         inner_rel_tol=self.inner_rel_tol, max_outer=10000,
                         warm_start=self.warm_start, check_every=self.check_every, precond=self.precond)
        return StepCfg(dt=self.dt, gravity=self.gravity, Vp=self.Vp,
                       alpha_liquid=self.alpha_liquid, alpha_gas=self.alpha_gas,
                       adaptive_cfl=self.adaptive_cfl, adaptive_min_dt=self.adaptive_min_dt,
                       c_div=0.0, adaptive_timestep=int(self.adaptive_timestep),
                       advection_order=self.advection_order, timing=int(self.timing),
                       phase=self.phase, solve=solve,
                       surface_tension=int(self.surface_tension),
                       surface_tension_smoothing_radius=self.surface_tension_curvature_smoothing_radius,
                       surface_tension_strength=self.surface_tension_strength,
                       surface_tension_max_delta_speed=self.surface_tension_max_delta_speed,
                       skip_interface_diagnostics=int(not self.interface_diagnostics),
                       c_div_volume_correction=1,
                       c_div_strength=self.c_div_strength if self.c_div_volume_correction else 0.0,
                       liquid_volume_target=self.liquid_volume_target, adapt=_adapt_of(self))

    def write_render_cache(self, path, frame, time, camera=None):
        """writeSparseRenderCache3D(sim, path, frame, time, camera) (render_cache3d.h:617-658)"""
        cam = camera or default_render_camera(self.nx, self.ny, self.nz, self.dx)
        return self.ctx.write_render_cache(path, RenderFrame(frame, 0, time, self.effective_dt_last, self.Vp,
                                                             self.phase, cam))

    def step(self):
        """SparseSim3DTP::step, src/driver/sparse_sim3d_tp.cpp:854-884 (particle
        adaptivity and updateVolumeCorrectionStats, :368-379, inside the step)."""
        st = self.ctx.step(self.step_cfg())
        _store_step_stats(self, st)
        self.effective_dt_last = st.effective_dt
        self.max_particle_speed_last = st.max_particle_speed
        self.interface_diagnostics_last = st.iface
        self.surface_tension_stats_last = st.surface
        self.stats_last = st
        return st


class SparseSim3D:
    """GPU mirror of the reference's single-phase SparseSim3D (src/driver/sparse_sim3d.h:5-15):
    trilinear unit-mass P2G, rho = 1 Jacobi PCG projection, FLIP blend with one alpha,
    RK2 advection."""

    def __init__(self, nx, ny, nz, dx):
        self.ctx = Context(nx, ny, nz, dx)
        self.nx, self.ny, self.nz, self.dx = nx, ny, nz, dx
        self.dt, self.alpha, self.gravity = 0.05, 0.95, -9.81
        self.cg_iters, self.cg_tol = 300, 1e-6
        self.stop_rule = STOP_REF_MAXABS
        self.inner_rel_tol, self.warm_start, self.check_every = 0.1, 0, 8
        # CFL-limited dt (timestep3d.h:16-50); off in the reference class
        self.adaptive_timestep, self.adaptive_cfl, self.adaptive_min_dt = False, 1.0, 1e-5
        self.precond = PRECOND_JACOBI
        self.timing = False
        self.stats_last = None

    def set_particles(self, x, y, z, u=None, v=None, w=None):
        self.ctx.upload(x, y, z, u, v, w)

    def step_cfg(self) -> StepCfg:
        solve = SolveCfg(max_iters=self.cg_iters, stop_rule=self.stop_rule, tol=self.cg_tol,
                         inner_rel_tol=self.inner_rel_tol, max_outer=10000,
                         warm_start=self.warm_start, check_every=self.check_every, precond=self.precond)
        return StepCfg(dt=self.dt, gravity=self.gravity, Vp=1.0, alpha_liquid=self.alpha,
                       alpha_gas=self.alpha, adaptive_cfl=self.adaptive_cfl,
                       adaptive_min_dt=self.adaptive_min_dt, c_div=0.0,
                       adaptive_timestep=int(self.adaptive_timestep), advection_order=2,
                       timing=int(self.timing), model=MODEL_SINGLE_PHASE,
                       phase=Phase(1.0, 1.0, 1.0, 1.0), solve=solve)

    def step(self):
        """SparseSim3D::step, src/driver/sparse_sim3d.cpp:106-115."""
        st = self.ctx.step(self.step_cfg())
        self.stats_last = st
        return st


class LearnedPreconditioner:
    """The learned preconditioner (a pf_lpc): conv3^3 network 4-16-32-16 plus a 1x1x1
    head, BF16 on MFMA, over 16^3 tiles of the multigrid cycle's aggregates; its
    output weights each aggregate's coarse correction (PRECOND_LEARNED).  Trained on
    the GPU."""

    def __init__(self, seed=0):
        self.h = C.c_void_p()
        st = lib().pf_lpc_create(C.byref(self.h), seed)
        if st != 0:
            raise PfError(f"pf_lpc_create failed with status {st}")

    def close(self):
        if self.h:
            lib().pf_lpc_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st, what, ctx=None):
        if st != 0:
            msg = lib().pf_last_error(ctx.h) if ctx is not None else lib().pf_lpc_last_error(self.h)
            raise PfError(f"{what}: status {st}: {msg.decode() if msg else ''}")

    @staticmethod
    def num_weights():
        return int(lib().pf_lpc_num_weights())

    def weights(self):
        w = np.zeros(self.num_weights(), np.float32)
        self._check(lib().pf_lpc_get_weights(self.h, _p(w), len(w)), "pf_lpc_get_weights")
        return w

    def set_weights(self, w):
        w = np.ascontiguousarray(w, np.float32)
        self._check(lib().pf_lpc_set_weights(self.h, _p(w), len(w)), "pf_lpc_set_weights")

    def forward_tiles(self, X):
        X = np.ascontiguousarray(X, np.float32)
        nt = X.shape[0]
        O = np.zeros((nt, 16, 16, 16), np.float32)
        self._check(lib().pf_lpc_forward_tiles(self.h, nt, _p(X), _p(O)), "pf_lpc_forward_tiles")
        return O

    def grad_tiles(self, X, gO):
        X = np.ascontiguousarray(X, np.float32)
        gO = np.ascontiguousarray(gO, np.float32)
        gW = np.zeros(self.num_weights(), np.float32)
        self._check(lib().pf_lpc_grad_tiles(self.h, X.shape[0], _p(X), _p(gO), _p(gW)), "pf_lpc_grad_tiles")
        return gW

    def attach(self, ctx):
        ctx._check(lib().pf_lpc_attach(ctx.h, self.h), "pf_lpc_attach")

    def set_objective(self, objective, power_steps=0):
        """LPC_OBJECTIVE_RESIDUAL, or LPC_OBJECTIVE_ENERGY over the cycle's slowest modes
        (the snapshot after power_steps preconditioned Richardson steps)"""
        self._check(lib().pf_lpc_set_objective(self.h, objective, power_steps), "pf_lpc_set_objective")

    def train(self, ctx, steps, lr=1e-3, seed=0):
        loss = C.c_double()
        self._check(lib().pf_lpc_train(self.h, ctx.h, steps, lr, seed, C.byref(loss)), "pf_lpc_train", ctx)
        return loss.value

    def loss(self, ctx, k):
        """(loss with the network's weights, loss of the plain cycle) on snapshot k"""
        a, b = C.c_double(), C.c_double()
        self._check(lib().pf_lpc_loss(self.h, ctx.h, k, C.byref(a), C.byref(b)), "pf_lpc_loss", ctx)
        return a.value, b.value

    def grad(self, ctx, k):
        """(gradient of the loss w.r.t. the flat weights, loss) on snapshot k"""
        g = np.zeros(self.num_weights(), np.float32)
        loss = C.c_double()
        self._check(lib().pf_lpc_grad(self.h, ctx.h, k, _p(g), C.byref(loss)), "pf_lpc_grad", ctx)
        return g, loss.value


def lpc_record(ctx, max_snapshots, every=1):
    ctx._check(lib().pf_lpc_record(ctx.h, max_snapshots, every), "pf_lpc_record")


def lpc_snapshot_count(ctx):
    return int(lib().pf_lpc_snapshot_count(ctx.h))


def mr_layout_fine_box(nx, ny, nz, box=None, level=1):
    """layoutWithFineBox: setCoarseEverywhere(level) + refineFineCellBox(box)."""
    b = np.zeros(6, np.int32) if box is None else np.ascontiguousarray(box, np.int32)
    n = lib().pf_mr_layout_fine_box(nx, ny, nz, level, int(box is not None), _p(b), None, 0)
    out = np.zeros((max(n, 1), 4), np.int32)
    lib().pf_mr_layout_fine_box(nx, ny, nz, level, int(box is not None), _p(b), _p(out), n)
    return out[:n]


def mr_layout_refine(nx, ny, nz, leaves, box):
    """refineFineCellBox (multires_layout3d.h:62-97) of `box` = (x0, y0, z0, x1, y1, z1)
    applied to the (n, 4) leaves; the sorted result"""
    lv = np.ascontiguousarray(leaves, np.int32).reshape(-1, 4)
    b = np.ascontiguousarray(box, np.int32)
    n = lib().pf_mr_layout_refine(nx, ny, nz, _p(lv), len(lv), _p(b), None, 0)
    if n < 0:
        raise PfError("pf_mr_layout_refine: bad arguments")
    out = np.zeros((max(n, 1), 4), np.int32)
    lib().pf_mr_layout_refine(nx, ny, nz, _p(lv), len(lv), _p(b), _p(out), n)
    return out[:n]


class MRContext(Context):
    """A context with a multires layout (MRSim3DTP's grid on the GPU)."""

    def set_layout(self, leaves):
        lv = np.ascontiguousarray(np.asarray(leaves, np.int32).reshape(-1, 4))
        self._check(lib().pf_mr_set_layout(self.h, len(lv), _p(lv)), "pf_mr_set_layout")

    def scene_bubbles(self, seed, nbubbles=16, band=6):
        """the multires multi-bubble tank: three-level layout and particles; returns the
        bubbles as (nbubbles, 4) centre and radius in domain units"""
        out = np.zeros((max(nbubbles, 1), 4), np.float32)
        self._check(lib().pf_mr_scene_bubbles(self.h, seed, nbubbles, band, _p(out)), "pf_mr_scene_bubbles")
        return out[:nbubbles]

    def layout(self):
        n = C.c_int64()
        self._check(lib().pf_mr_get_layout(self.h, None, 0, C.byref(n)), "pf_mr_get_layout")
        out = np.zeros((max(n.value, 1), 4), np.int32)
        self._check(lib().pf_mr_get_layout(self.h, _p(out), n.value, C.byref(n)), "pf_mr_get_layout")
        return out[:n.value]

    def face_valid(self, axis):
        d = np.zeros(3, np.int32)
        self._check(lib().pf_field_dims(self.h, axis, _p(d)), "pf_field_dims")
        out = np.zeros(int(d.prod()), np.uint8)
        self._check(lib().pf_mr_face_valid(self.h, axis, _p(out)), "pf_mr_face_valid")
        return out.reshape(d[2], d[1], d[0])

    def mr_mark_cells(self):
        self._check(lib().pf_mr_mark_cells(self.h), "pf_mr_mark_cells")

    def mr_p2g(self, phase, Vp=1.0):
        self._check(lib().pf_mr_p2g(self.h, C.byref(phase), Vp), "pf_mr_p2g")

    def mr_project(self, phase, dt, cfg=None, c_div=0.0):
        cfg = cfg or mr_solve_cfg()
        st = SolveStats()
        self._check(lib().pf_mr_project(self.h, C.byref(phase), dt, C.byref(cfg), c_div, C.byref(st)),
                    "pf_mr_project")
        return st

    def mr_update_refinement(self, refine):
        fl, bl = C.c_int32(), C.c_int32()
        self._check(lib().pf_mr_update_refinement(self.h, C.byref(refine), C.byref(fl), C.byref(bl)),
                    "pf_mr_update_refinement")
        return fl.value, bl.value

    def mr_retained_box(self):
        """(valid, [x0, y0, z0, x1, y1, z1]) of the last refinement update's kept box"""
        box, valid = np.zeros(6, np.int32), C.c_int32()
        self._check(lib().pf_mr_retained_box(self.h, _p(box), C.byref(valid)), "pf_mr_retained_box")
        return bool(valid.value), box

    def mr_solve_detail(self):
        """MRPressureSolveStats3D counters of the last multires projection"""
        d = MRSolveDetail()
        self._check(lib().pf_mr_last_solve_detail(self.h, C.byref(d)), "pf_mr_last_solve_detail")
        return d

    def mr_interface_diagnostics(self, phase: Phase) -> IfaceStats:
        """diagnoseMRInterface3D over the current masses and leaf markers"""
        out = IfaceStats()
        self._check(lib().pf_mr_interface_diagnostics(self.h, C.byref(phase), C.byref(out)),
                    "pf_mr_interface_diagnostics")
        return out

    def mr_apply_surface_tension(self, phase: Phase, dt, strength=0.02, max_delta_speed=0.25,
                                 smoothing_radius=0) -> SurfaceStats:
        """applyMRSurfaceTension3D on the current face velocities"""
        out = SurfaceStats()
        self._check(lib().pf_mr_apply_surface_tension(self.h, C.byref(phase), dt, strength, max_delta_speed,
                                                      int(smoothing_radius), C.byref(out)),
                    "pf_mr_apply_surface_tension")
        return out

    def mr_step(self, cfg):
        st = StepStats()
        self._check(lib().pf_mr_step(self.h, C.byref(cfg), C.byref(st)), "pf_mr_step")
        return st


class MRSim3DTP:
    """GPU mirror of the reference's MRSim3DTP (src/driver/multires_sim3d_tp.h:10-191):
    the fields step() reads, dynamic refinement, step()."""

    def __init__(self, nx, ny, nz, dx):
        self.ctx = MRContext(nx, ny, nz, dx)
        self.nx, self.ny, self.nz, self.dx = nx, ny, nz, dx
        self.phase = Phase(1.0, 0.01, 1.0, 1.0)
        self.dt, self.gravity, self.Vp = 0.02, -9.81, 1.0
        self.alpha_liquid, self.alpha_gas = 0.95, 0.95
        self.adaptive_timestep, self.adaptive_cfl, self.adaptive_min_dt = False, 0.5, 1e-5
        self.advection_order = 2
        self.cg_iters, self.cg_tol, self.cg_rel_tol = 160, 1e-7, 1e-6
        self.dynamic_refinement, self.dynamic_particle_padding, self.dynamic_gas_padding = True, 1, 2
        self.dynamic_hysteresis_cells, self.dynamic_max_fine_leaves = 1, 0
        # the reference's other MRPressureSolveConfig3D options by field name, e.g.
        # {"mode": MR_SOLVE_REFERENCE, "use_coarse_correction": 1}
        self.solve_options = {}
        # CSF surface tension (multires_sim3d_tp.h:170-175)
        self.surface_tension = False
        self.surface_tension_strength = 0.02
        self.surface_tension_max_delta_speed = 0.25
        self.surface_tension_curvature_smoothing_radius = 0
        # False skips the per-step interface statistics (the reference always computes them)
        self.interface_diagnostics = True
        self.interface_diagnostics_last = IfaceStats(finite=1)
        self.surface_tension_stats_last = SurfaceStats(finite=1, capillary_stable=1)
        _init_adaptivity(self)
        # updateVolumeCorrectionStats (multires_sim3d_tp.cpp:439-450)
        self.c_div_volume_correction, self.c_div_strength, self.liquid_volume_target = False, 1.0, 0.0
        self.liquid_volume_current_last = self.liquid_volume_error_last = 0.0
        self.effective_dt_last = self.dt
        self.c_div_last = 0.0
        self.timing = False
        self.stats_last = None

    def set_layout(self, leaves):
        self.ctx.set_layout(leaves)

    def set_particles(self, x, y, z, u=None, v=None, w=None, type=None, volume=None):
        self.ctx.upload(x, y, z, u, v, w, type, volume)

    def step_cfg(self):
        solve = mr_solve_cfg(self.cg_iters, self.cg_tol, self.cg_rel_tol, **self.solve_options)
        refine = MRRefineCfg(int(self.dynamic_refinement), self.dynamic_particle_padding, self.dynamic_gas_padding,
                             self.dynamic_hysteresis_cells, self.dynamic_max_fine_leaves, 0)
        return MRStepCfg(dt=self.dt, gravity=self.gravity, Vp=self.Vp, alpha_liquid=self.alpha_liquid,
                         alpha_gas=self.alpha_gas, adaptive_cfl=self.adaptive_cfl,
                         adaptive_min_dt=self.adaptive_min_dt, c_div=0.0,
                         adaptive_timestep=int(self.adaptive_timestep), advection_order=self.advection_order,
                         timing=int(self.timing), reserved=0, phase=self.phase, solve=solve, refine=refine,
                         surface_tension=int(self.surface_tension),
                         surface_tension_smoothing_radius=self.surface_tension_curvature_smoothing_radius,
                         surface_tension_strength=self.surface_tension_strength,
                         surface_tension_max_delta_speed=self.surface_tension_max_delta_speed,
                         skip_interface_diagnostics=int(not self.interface_diagnostics),
                         c_div_volume_correction=1,
                         c_div_strength=self.c_div_strength if self.c_div_volume_correction else 0.0,
                         liquid_volume_target=self.liquid_volume_target, adapt=_adapt_of(self))

    def write_render_cache(self, path, frame, time, camera=None):
        """writeMRRenderCache3D(sim, path, frame, time, camera) (render_cache3d.h:660-711)"""
        cam = camera or default_render_camera(self.nx, self.ny, self.nz, self.dx)
        return self.ctx.write_render_cache(path, RenderFrame(frame, 0, time, self.effective_dt_last, self.Vp,
                                                             self.phase, cam))

    def step(self):
        """MRSim3DTP::step, src/driver/multires_sim3d_tp.cpp:671-734."""
        st = self.ctx.mr_step(self.step_cfg())
        _store_step_stats(self, st)
        self.effective_dt_last = st.effective_dt
        self.interface_diagnostics_last = st.iface
        self.surface_tension_stats_last = st.surface
        self.stats_last = st
        return st
