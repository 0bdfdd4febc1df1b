"""ctypes bindings for the parity checkers (TEST INFRASTRUCTURE ONLY).

* ``Oracle``: the C restatement in ``oracle/pf_oracle.c`` (liboracle.so).
* ``RefSim``: the reference's own SparseSim3DTP step compiled from its sources by
  ``oracle/Makefile`` into ``oracle/_ref/libpfref.so`` (present only where the
  reference was available at build time).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libpfref.so")

U, V, W, MU, MV, MW, P, MK = range(8)
AIR, FLUID, SOLID = 0, 1, 2
STOP_REF_MAXABS, STOP_REL2 = 0, 1
SCENE_RT, SCENE_BUBBLE, SCENE_DAMBREAK = 0, 1, 2

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_u8 = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_i32 = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class Phase(C.Structure):
    _fields_ = [("rho_l", C.c_double), ("rho_g", C.c_double), ("alpha_phi", C.c_double),
                ("rho_tilde_0", C.c_double)]

    def as_array(self):
        return np.array([self.rho_l, self.rho_g, self.alpha_phi, self.rho_tilde_0], np.float64)


class SolveStats(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("pinned", C.c_int32), ("converged", C.c_int32),
                ("breakdown", C.c_int32), ("fluid_cells", C.c_int64), ("pinned_cell", C.c_int64),
                ("initial_max_residual", C.c_double), ("final_max_residual", C.c_double),
                ("true_rel_residual", C.c_double), ("true_max_residual", C.c_double)]


class EscapeStats(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("clamped_liquid", "clamped_gas", "x_lo", "x_hi",
                                          "y_lo", "y_hi", "z_lo", "z_hi")]


class TimestepStats(C.Structure):
    _fields_ = [("requested_dt", C.c_double), ("effective_dt", C.c_double),
                ("max_particle_speed", C.c_double), ("cfl_limit_dt", C.c_double),
                ("limited", C.c_int32)]


class StepCfg(C.Structure):
    _fields_ = [("dt", C.c_double), ("gravity", C.c_double), ("Vp", C.c_double),
                ("alpha_liquid", C.c_double), ("alpha_gas", C.c_double),
                ("adaptive_timestep", C.c_int32), ("advection_order", C.c_int32),
                ("cg_iters", C.c_int32), ("stop_rule", C.c_int32),
                ("adaptive_cfl", C.c_double), ("adaptive_min_dt", C.c_double),
                ("cg_tol", C.c_double), ("c_div_volume_correction", C.c_int32),
                ("pad0", C.c_int32), ("c_div_strength", C.c_double),
                ("liquid_volume_target", C.c_double), ("phase", Phase),
                ("surface_tension", C.c_int32), ("surface_tension_smoothing_radius", C.c_int32),
                ("surface_tension_strength", C.c_double), ("surface_tension_max_delta_speed", C.c_double)]


class IfaceStats(C.Structure):
    """InterfaceDiagnostics3D (src/driver/interface_diagnostics3d.h:7-19)"""
    _fields_ = [("sample_cells", C.c_int32), ("interface_cells", C.c_int32), ("finite", C.c_int32),
                ("surface_tension_candidate", C.c_int32)] + [
        (n, C.c_double) for n in ("phi_min", "phi_max", "phi_mean", "grad_mean", "grad_max", "curvature_abs_mean",
                                  "curvature_abs_max")]

    def values(self):
        return [getattr(self, n) for n, _ in self._fields_]


class StStats(C.Structure):
    """SurfaceTensionStats3D (src/driver/interface_diagnostics3d.h:21-36)"""
    _fields_ = [(n, C.c_int32) for n in ("enabled", "applied_cells", "finite", "curvature_smoothing_radius",
                                          "capillary_stable", "pad0")] + [
        (n, C.c_double) for n in ("strength", "max_delta_speed_limit", "mean_delta_speed", "max_delta_speed",
                                  "raw_curvature_abs_mean", "raw_curvature_abs_max", "smoothed_curvature_abs_mean",
                                  "smoothed_curvature_abs_max", "capillary_dt_limit")]

    def values(self):
        return [getattr(self, n) for n, _ in self._fields_ if n != "pad0"]


class PaDomain(C.Structure):
    """ParticleCellDomain3D (src/driver/particle_adaptivity3d.h:13-48)"""
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("pad0", C.c_int32),
                ("dx", C.c_double), ("ox", C.c_double), ("oy", C.c_double), ("oz", C.c_double)]


class AdaptCfg(C.Structure):
    """the particle adaptivity fields of SparseSim3DTP (sparse_sim3d_tp.h:29-60)"""
    _fields_ = [("narrow_band_air", C.c_int32), ("narrow_band_air_radius", C.c_int32),
                ("gas_particle_coarsening", C.c_int32), ("gas_particles_per_cell_target", C.c_int32),
                ("gas_particle_coarsening_seed", C.c_uint32),
                ("liquid_particle_coarsening", C.c_int32), ("liquid_particles_per_cell_target", C.c_int32),
                ("liquid_particle_coarsening_seed", C.c_uint32),
                ("liquid_particle_refill", C.c_int32), ("liquid_particle_refill_interface_only", C.c_int32),
                ("liquid_particle_refill_interface_radius", C.c_int32),
                ("liquid_particle_refill_max_added_per_step", C.c_int32),
                ("liquid_refill_particles_per_cell_target", C.c_int32), ("liquid_particle_refill_seed", C.c_uint32)]

    def words(self):
        return np.frombuffer(bytes(self), np.uint32).copy()


def adapt_cfg(**kw) -> AdaptCfg:
    """the reference's defaults (all stages off, radius 2, targets 4, its seeds), then kw"""
    c = AdaptCfg(narrow_band_air_radius=2, gas_particles_per_cell_target=4, gas_particle_coarsening_seed=0x9e3779b9,
                 liquid_particles_per_cell_target=4, liquid_particle_coarsening_seed=0x7f4a7c15,
                 liquid_particle_refill_interface_radius=1, liquid_refill_particles_per_cell_target=4,
                 liquid_particle_refill_seed=0x6a09e667)
    for k, v in kw.items():
        if k not in dict(AdaptCfg._fields_):
            raise KeyError(k)
        setattr(c, k, v)
    return c


class AdaptStats(C.Structure):
    """NarrowBandAirResult3D, GasParticleCoarseningResult3D (gas, then liquid) and
    LiquidParticleRefillResult3D (particle_adaptivity3d.h:50-73)"""
    _fields_ = [(n, C.c_int32) for n in (
        "narrow_band_air_removed", "narrow_band_air_liquid_cells", "narrow_band_air_gas_before",
        "narrow_band_air_gas_after", "gas_removed", "gas_cells", "gas_overfull_cells", "gas_before", "gas_after",
        "liquid_removed", "liquid_cells", "liquid_overfull_cells", "liquid_before", "liquid_after",
        "refill_added", "refill_cells", "refill_interface_cells", "refill_underfull_cells",
        "refill_budget_limited", "refill_before", "refill_after", "pad0")]

    def values(self):
        return [getattr(self, n) for n, _ in self._fields_ if n != "pad0"]


def _adapt_capacity(domain, cfg, n):
    extra = domain.nx * domain.ny * domain.nz * max(1, cfg.liquid_refill_particles_per_cell_target) \
        if cfg.liquid_particle_refill else 0
    return n + extra


def _adapt_arrays(ps, cap):
    n = len(ps)
    out = [np.zeros(cap) for _ in range(6)]
    for a, src in zip(out, (ps.x, ps.y, ps.z, ps.u, ps.v, ps.w)):
        a[:n] = src
    t = np.zeros(cap, np.uint8)
    t[:n] = ps.type
    vol = np.ones(cap)
    if ps.volume is not None:
        vol[:n] = ps.volume
    return out, t, vol


def particle_adaptivity(domain: PaDomain, cfg: AdaptCfg, ps):
    """applyParticleAdaptivity on a copy of ps: (new Particles with volumes, AdaptStats)"""
    cap = _adapt_capacity(domain, cfg, len(ps))
    a, t, vol = _adapt_arrays(ps, cap)
    st = AdaptStats()
    m = lib().orc_particle_adaptivity(C.byref(domain), C.byref(cfg), len(ps), cap, *a, t, vol, C.byref(st))
    assert m >= 0
    return Particles(*(q[:m] for q in a), type=t[:m], volume=vol[:m]), st


def ref_particle_adaptivity(domain: PaDomain, cfg: AdaptCfg, ps):
    """the reference's pa3d functions on a copy of ps: (new Particles, 21 counters)"""
    cap = _adapt_capacity(domain, cfg, len(ps))
    a, t, vol = _adapt_arrays(ps, cap)
    st = np.zeros(22, np.int32)
    m = reflib().ref_pa_apply(np.array([domain.nx, domain.ny, domain.nz], np.int32),
                              np.array([domain.dx, domain.ox, domain.oy, domain.oz]), cfg.words(), len(ps), cap,
                              *a, t, vol, st)
    assert m >= 0
    return Particles(*(q[:m] for q in a), type=t[:m], volume=vol[:m]), list(st[:21])


class RenderMeta(C.Structure):
    """frame, time, dt, Vp, phase and camera of one render-cache frame"""
    _fields_ = [("frame", C.c_int32), ("pad0", C.c_int32), ("time", C.c_double), ("dt", C.c_double),
                ("Vp", C.c_double), ("phase", Phase), ("camera", C.c_double * 13)]


def default_camera(nx, ny, nz, dx):
    """defaultRenderCacheCamera3D (render_cache3d.h:48-62): position, target, up, fov, near, far, focal"""
    cam = (C.c_double * 13)()
    lib().orc_render_default_camera(nx, ny, nz, dx, cam)
    return cam


def render_meta(nx, ny, nz, dx, phase, frame=0, time=0.0, dt=0.0, Vp=1.0, camera=None):
    return RenderMeta(frame, 0, time, dt, Vp, phase, camera if camera is not None else default_camera(nx, ny, nz, dx))


def _render_particles(ps):
    vol = ps.volume if ps.volume is not None else np.ones(len(ps))
    return (len(ps), ps.x, ps.y, ps.z, ps.u, ps.v, ps.w, ps.type, np.ascontiguousarray(vol, np.float64))


def render_manifest(path, kind, nx, ny, nz, dx, frames, ref=False):
    """writeRenderCacheManifest3D; frames: (frame, step, time, path, bytes) tuples"""
    n = len(frames)
    fr = np.array([f[0] for f in frames], np.int32)
    st = np.array([f[1] for f in frames], np.int32)
    tm = np.array([f[2] for f in frames], np.float64)
    by = np.array([f[4] for f in frames], np.int64)
    paths = (C.c_char_p * max(n, 1))(*[f[3].encode() for f in frames])
    L = reflib() if ref else lib()
    fn = L.ref_render_manifest if ref else L.orc_render_manifest
    return fn(path.encode(), kind.encode(), nx, ny, nz, dx, n, fr, st, tm, paths, by)


class MRSolveCfg(C.Structure):
    _fields_ = [("max_iterations", C.c_int32), ("use_jacobi", C.c_int32), ("use_flexible_beta", C.c_int32),
                ("adaptive_restart", C.c_int32), ("absolute_tolerance", C.c_double),
                ("relative_tolerance", C.c_double), ("restart_growth_threshold", C.c_double),
                ("relaxation_sweeps", C.c_int32), ("pad0", C.c_int32), ("relaxation_omega", C.c_double),
                ("relaxation_min_omega", C.c_double), ("divergence_correction", C.c_double),
                ("use_coarse_correction", C.c_int32), ("coarse_correction_iterations", C.c_int32),
                ("coarse_correction_sweeps", C.c_int32), ("use_coarse_preconditioner", C.c_int32),
                ("coarse_preconditioner_iterations", C.c_int32), ("coarse_preconditioner_auto_disable", C.c_int32),
                ("coarse_preconditioner_auto_disable_after", C.c_int32), ("pad1", C.c_int32),
                ("coarse_correction_absolute_tolerance", C.c_double),
                ("coarse_correction_relative_tolerance", C.c_double), ("coarse_correction_min_scale", C.c_double),
                ("coarse_preconditioner_absolute_tolerance", C.c_double),
                ("coarse_preconditioner_relative_tolerance", C.c_double),
                ("coarse_preconditioner_scale", C.c_double), ("coarse_preconditioner_min_rz_gain", C.c_double),
                ("coarse_preconditioner_max_work_ratio", C.c_double)]


def mr_solve_cfg(max_iterations=160, abs_tol=1e-7, rel_tol=1e-6, jacobi=1, flexible=0, restart=1, growth=10.0,
                 sweeps=0, omega=0.67, min_omega=0.05, cdiv=0.0, **coarse):
    """MRPressureSolveConfig3D with MRSim3DTP's defaults (multires_sim3d_tp.h:36-60); the
    coarse options are keyword arguments named as the structure's fields."""
    c = MRSolveCfg(max_iterations, jacobi, flexible, restart, abs_tol, rel_tol, growth, sweeps, 0, omega,
                   min_omega, cdiv)
    c.coarse_correction_iterations, c.coarse_correction_sweeps = 32, 1
    c.coarse_correction_relative_tolerance, c.coarse_correction_min_scale = 1e-3, 1.0 / 64.0
    c.coarse_preconditioner_iterations, c.coarse_preconditioner_relative_tolerance = 8, 1e-2
    c.coarse_preconditioner_scale, c.coarse_preconditioner_auto_disable_after = 1.0, 4
    for k, v in coarse.items():
        if not hasattr(c, k):
            raise TypeError(f"unknown solve option {k}")
        setattr(c, k, v)
    return c


class MRSolveStats(C.Structure):
    _fields_ = [("active_cells", C.c_int64), ("pinned_cell", C.c_int64), ("iterations", C.c_int32),
                ("restarts", C.c_int32), ("beta_resets", C.c_int32), ("relaxation_accepted", C.c_int32),
                ("relaxation_rejected", C.c_int32), ("converged", C.c_int32), ("breakdown", C.c_int32),
                ("pad0", C.c_int32), ("initial_residual", C.c_double), ("final_residual", C.c_double),
                ("effective_tolerance", C.c_double), ("true_max_residual", C.c_double),
                ("true_rel_residual", C.c_double)] + [
        (n, C.c_int32) for n in ("cc_cells", "cc_iterations", "cc_sweeps", "cc_accepted_sweeps",
                                 "cc_rejected_sweeps", "cc_accepted", "cc_converged", "cc_breakdown", "cp_cells",
                                 "cp_applications", "cp_accepted", "cp_rejected", "cp_skipped",
                                 "cp_budget_limited", "cp_iterations", "cp_auto_disabled",
                                 "cp_auto_disable_iteration", "cp_wasted_streak", "cp_breakdown",
                                 "cp_budget_exhausted")] + [
        (n, C.c_double) for n in ("cc_initial_residual", "cc_final_residual", "cc_last_scale",
                                  "cc_effective_tolerance", "cp_effective_tolerance", "cp_last_rz_gain")]


class MRSim(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("pad0", C.c_int32), ("dx", C.c_double),
                ("phase", Phase), ("dt", C.c_double), ("gravity", C.c_double), ("Vp", C.c_double),
                ("alpha_liquid", C.c_double), ("alpha_gas", C.c_double), ("adaptive_timestep", C.c_int32),
                ("advection_order", C.c_int32), ("adaptive_cfl", C.c_double), ("adaptive_min_dt", C.c_double),
                ("c_div_volume_correction", C.c_int32), ("pad1", C.c_int32), ("c_div_strength", C.c_double),
                ("liquid_volume_target", C.c_double), ("solve", MRSolveCfg), ("dynamic_refinement", C.c_int32),
                ("particle_padding", C.c_int32), ("gas_padding", C.c_int32), ("hysteresis", C.c_int32),
                ("max_fine_leaves", C.c_int32), ("retained_valid", C.c_int32), ("retained", C.c_int32 * 6),
                ("budget_limited", C.c_int32), ("dynamic_last_fine_leaves", C.c_int32),
                ("surface_tension", C.c_int32), ("surface_tension_smoothing_radius", C.c_int32),
                ("surface_tension_strength", C.c_double), ("surface_tension_max_delta_speed", C.c_double)]


def mr_sim(nx, ny, nz, dx=1.0, phase=None, dt=0.02, order=2, dynamic=1, ppad=1, gpad=2, hyst=1, max_fine=0,
           adaptive=0, cfl=0.5, mindt=1e-5, solve=None, aL=0.95, aG=0.95, gravity=-9.81):
    """MRSim3DTP's defaults (src/driver/multires_sim3d_tp.h:11-70)."""
    return MRSim(nx, ny, nz, 0, dx, phase or Phase(1.0, 0.01, 1.0, 1.0), dt, gravity, 1.0, aL, aG, adaptive, order,
                 cfl, mindt, 0, 0, 1.0, 0.0, solve or mr_solve_cfg(), dynamic, ppad, gpad, hyst, max_fine, 0,
                 (C.c_int32 * 6)(), 0, 0)


class StepStats(C.Structure):
    _fields_ = [("ts", TimestepStats), ("solve", SolveStats), ("escape", EscapeStats),
                ("liquid_volume", C.c_double), ("c_div", C.c_double),
                ("ms_mark", C.c_double), ("ms_p2g", C.c_double),
                ("ms_save_gravity_walls", C.c_double), ("ms_project", C.c_double),
                ("ms_g2p", C.c_double), ("ms_advect", C.c_double), ("ms_total", C.c_double),
                ("iface", IfaceStats), ("surface", StStats)]


def build(ref: bool = True) -> None:
    """Compile liboracle.so (and _ref/libpfref.so when the reference is present)."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    if ref and os.path.isdir("/root/reference/src"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


_lib = None


def set_threads(n: int) -> None:
    """Host threads for the oracle's particle loops (P2G, G2P, advection); results are
    bit-identical to one thread."""
    lib().orc_set_threads(int(n))


def set_threads_split(p2g: int, particles: int) -> None:
    """P2G threads and G2P/advection threads separately"""
    lib().orc_set_threads_split(int(p2g), int(particles))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build(ref=False)
        L = C.CDLL(ORACLE_SO)
        L.orc_grid_create.restype = C.c_void_p
        L.orc_grid_create.argtypes = [C.c_int] * 3 + [C.c_double] * 4 + [C.c_int]
        L.orc_grid_destroy.argtypes = [C.c_void_p]
        L.orc_grid_set_faithful.argtypes = [C.c_void_p, C.c_int]
        L.orc_grid_clone.restype = C.c_void_p
        L.orc_grid_clone.argtypes = [C.c_void_p]
        L.orc_field_dims.argtypes = [C.c_void_p, C.c_int, _i32]
        L.orc_get.restype = C.c_double
        L.orc_get.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_set.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double]
        L.orc_field_dense.argtypes = [C.c_void_p, C.c_int, _dp]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_set_threads_split.argtypes = [C.c_int, C.c_int]
        L.orc_get_threads.restype = C.c_int
        L.orc_active_blocks.restype = C.c_int64
        L.orc_active_blocks.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.orc_clear_field.argtypes = [C.c_void_p, C.c_int]
        L.orc_phi.restype = C.c_double
        L.orc_phi.argtypes = [C.c_double, C.POINTER(Phase)]
        L.orc_beta.restype = C.c_double
        L.orc_beta.argtypes = [C.c_double, C.POINTER(Phase)]
        L.orc_calibrate_rho_tilde0.restype = C.c_double
        L.orc_calibrate_rho_tilde0.argtypes = [C.POINTER(Phase), C.c_double]
        L.orc_mark_cells.argtypes = [C.c_void_p, C.c_int64, _dp, _dp, _dp]
        L.orc_p2g.argtypes = [C.c_void_p, C.c_int64, _dp, _dp, _dp, _dp, _dp, _dp, _u8,
                              C.c_void_p, C.POINTER(Phase), C.c_double]
        L.orc_apply_gravity.argtypes = [C.c_void_p, C.c_double, C.c_double]
        L.orc_apply_walls.argtypes = [C.c_void_p]
        L.orc_project.argtypes = [C.c_void_p, C.POINTER(Phase), C.c_double, C.c_int,
                                  C.c_double, C.c_double, C.c_int, C.POINTER(SolveStats)]
        L.orc_g2p.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, _dp, _dp, _dp, _dp, _dp, _dp,
                              _u8, C.c_double, C.c_double]
        L.orc_advect.argtypes = [C.c_void_p, C.c_int64, _dp, _dp, _dp, _u8, C.c_double, C.c_int,
                                 C.POINTER(EscapeStats), C.c_void_p, C.c_void_p]
        L.orc_timestep.argtypes = [C.c_int64, _dp, _dp, _dp, C.c_double, C.c_double, C.c_int,
                                   C.c_double, C.c_double, C.POINTER(TimestepStats)]
        L.orc_step.argtypes = [C.c_void_p, C.c_int64, _dp, _dp, _dp, _dp, _dp, _dp, _u8,
                               C.c_void_p, C.POINTER(StepCfg), C.POINTER(StepStats)]
        L.orc_interface_diagnostics.argtypes = [C.c_void_p, C.POINTER(Phase), C.POINTER(IfaceStats)]
        L.orc_surface_tension.argtypes = [C.c_void_p, C.POINTER(Phase), C.c_double, C.c_double, C.c_double,
                                          C.c_int, C.POINTER(StStats)]
        L.orc_interface_fields.argtypes = [C.c_void_p, C.POINTER(Phase), _dp, _dp, _dp]
        L.orc_p2g_single.argtypes = [C.c_void_p, C.c_int64, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_project_single.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_double, C.c_int,
                                         C.POINTER(SolveStats)]
        L.orc_step_single.argtypes = ([C.c_void_p, C.c_int64] + [_dp] * 6 +
                                      [C.c_double, C.c_double, C.c_double, C.c_int, C.c_double,
                                       C.c_int, C.POINTER(StepStats)])
        L.orc_mr_create.restype = C.c_void_p
        L.orc_mr_create.argtypes = [C.c_int] * 3 + [C.c_double, C.c_int]
        L.orc_mr_destroy.argtypes = [C.c_void_p]
        L.orc_mr_clone.restype = C.c_void_p
        L.orc_mr_clone.argtypes = [C.c_void_p]
        L.orc_mr_set_leaves.argtypes = [C.c_void_p, C.c_int, _i32]
        L.orc_mr_leaf_count.restype = C.c_int64
        L.orc_mr_leaf_count.argtypes = [C.c_void_p]
        L.orc_mr_get_leaves.argtypes = [C.c_void_p, _i32]
        L.orc_mr_layout_fine_box.restype = C.c_int64
        L.orc_mr_layout_fine_box.argtypes = [C.c_int] * 5 + [_i32, C.c_void_p, C.c_int64]
        L.orc_mr_layout_bubble_band.restype = C.c_int64
        L.orc_mr_layout_bubble_band.argtypes = [C.c_int] * 3 + [C.c_void_p, C.c_int64]
        L.orc_mr_face_valid.argtypes = [C.c_void_p, C.c_int, _u8]
        L.orc_mr_field_dense.argtypes = [C.c_void_p, C.c_int, _dp]
        L.orc_mr_set.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double]
        L.orc_mr_mark_cells.argtypes = [C.c_void_p, C.c_int64, _dp, _dp, _dp]
        L.orc_mr_p2g.argtypes = [C.c_void_p, C.c_int64] + [_dp] * 6 + [_u8, C.c_void_p, C.POINTER(Phase),
                                                                          C.c_double]
        L.orc_mr_apply_gravity.argtypes = [C.c_void_p, C.c_double, C.c_double]
        L.orc_mr_apply_walls.argtypes = [C.c_void_p]
        L.orc_mr_project.argtypes = [C.c_void_p, C.POINTER(Phase), C.c_double, C.POINTER(MRSolveCfg),
                                     C.POINTER(MRSolveStats)]
        L.orc_mr_system.restype = C.c_int64
        L.orc_mr_system.argtypes = [C.c_void_p, C.POINTER(Phase), C.c_double, C.c_double, C.c_int64, C.c_int64,
                                    C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.orc_mr_apply_pressure.argtypes = [C.c_void_p, C.POINTER(Phase), C.c_double, _dp]
        L.orc_mr_g2p.argtypes = [C.c_void_p, C.c_void_p, C.c_int64] + [_dp] * 6 + [_u8, C.c_double, C.c_double]
        L.orc_mr_advect.argtypes = [C.c_void_p, C.c_int64, _dp, _dp, _dp, _u8, C.c_double, C.c_int,
                                    C.POINTER(EscapeStats)]
        L.orc_mr_step.argtypes = [C.c_void_p, C.c_void_p, C.c_int64] + [_dp] * 6 + [_u8, C.c_void_p, C.c_void_p]
        L.orc_particle_adaptivity.restype = C.c_int64
        L.orc_particle_adaptivity.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_int64] + [_dp] * 6 + \
            [_u8, _dp, C.c_void_p]
        L.orc_mix32.restype = C.c_uint32
        L.orc_mix32.argtypes = [C.c_uint32]
        L.orc_particle_score.restype = C.c_uint32
        L.orc_particle_score.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int,
                                         C.c_uint32]
        L.orc_slot_score.restype = C.c_uint32
        L.orc_slot_score.argtypes = [C.c_uint64, C.c_int, C.c_uint32]
        L.orc_render_default_camera.argtypes = [C.c_int] * 3 + [C.c_double, C.c_void_p]
        _rp = [C.c_int64] + [_dp] * 6 + [_u8, _dp]
        L.orc_render_cache_sparse.argtypes = [C.c_void_p, C.c_void_p] + _rp + [C.c_void_p, C.c_char_p]
        L.orc_render_cache_mr.argtypes = [C.c_void_p, C.c_void_p] + _rp + [C.c_void_p, C.c_char_p]
        L.orc_render_manifest.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int64,
                                          _i32, _i32, _dp, C.c_void_p,
                                          np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")]
        L.orc_mr_interface_diagnostics.argtypes = [C.c_void_p, C.POINTER(Phase), C.POINTER(IfaceStats)]
        L.orc_mr_surface_tension.argtypes = [C.c_void_p, C.POINTER(Phase), C.c_double, C.c_double, C.c_double,
                                             C.c_int, C.POINTER(StStats)]
        L.orc_mr_dynamic_refinement.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, _dp, _dp, _dp, _u8]
        L.orc_scene_count.restype = C.c_int64
        L.orc_scene_count.argtypes = [C.c_int] * 4
        L.orc_scene_fill.restype = C.c_int64
        L.orc_scene_fill.argtypes = [C.c_int] * 4 + [C.c_double, C.c_uint64, _fp, _fp, _fp, _u8]
        L.orc_dense_p2g.argtypes = ([C.c_int] * 3 + [C.c_double, C.c_int64] + [_dp] * 6 +
                                    [_u8, C.c_void_p, C.POINTER(Phase), C.c_double] + [_dp] * 6)
        _lib = L
    return _lib


def _vptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Particles:
    """Double-precision particle state (src/particles/particles3d_tp.h:4-35)."""

    def __init__(self, x, y, z, u=None, v=None, w=None, type=None, volume=None):
        f = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        self.x, self.y, self.z = f(x), f(y), f(z)
        n = len(self.x)
        self.u = f(u if u is not None else np.zeros(n))
        self.v = f(v if v is not None else np.zeros(n))
        self.w = f(w if w is not None else np.zeros(n))
        self.type = np.ascontiguousarray(type if type is not None else np.zeros(n), np.uint8)
        self.volume = None if volume is None else f(volume)

    def __len__(self):
        return len(self.x)

    def copy(self):
        return Particles(self.x.copy(), self.y.copy(), self.z.copy(), self.u.copy(),
                         self.v.copy(), self.w.copy(), self.type.copy(),
                         None if self.volume is None else self.volume.copy())


class Grid:
    """Oracle SparseMacGrid3D<4> (faithful=True: float pool as the reference)."""

    def __init__(self, nx, ny, nz, dx=1.0, origin=(0.0, 0.0, 0.0), faithful=True, _h=None):
        self.nx, self.ny, self.nz, self.dx = nx, ny, nz, dx
        self.origin = origin
        self.faithful = faithful
        self.h = _h if _h is not None else lib().orc_grid_create(nx, ny, nz, dx, *origin,
                                                                  int(faithful))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_grid_destroy(self.h)
            self.h = None

    def clone(self):
        return Grid(self.nx, self.ny, self.nz, self.dx, self.origin, self.faithful,
                    _h=lib().orc_grid_clone(self.h))

    def set_faithful(self, faithful: bool):
        """round later stores as the reference's float pool (True) or keep FP64"""
        lib().orc_grid_set_faithful(self.h, int(faithful))
        self.faithful = bool(faithful)

    def dims(self, f):
        d = np.zeros(3, np.int32)
        lib().orc_field_dims(self.h, f, d)
        return tuple(int(a) for a in d)

    def dense(self, f):
        nx, ny, nz = self.dims(f)
        out = np.zeros(nx * ny * nz, np.float64)
        lib().orc_field_dense(self.h, f, out)
        return out.reshape(nz, ny, nx)

    def active_blocks(self, f):
        n = lib().orc_active_blocks(self.h, f, None)
        out = np.zeros(max(n, 1), np.int32)
        lib().orc_active_blocks(self.h, f, _vptr(out))
        return out[:n]

    def get(self, f, i, j, k):
        return lib().orc_get(self.h, f, i, j, k)

    def set(self, f, i, j, k, v):
        lib().orc_set(self.h, f, i, j, k, float(v))

    def load_dense(self, f, a):
        """write the non-zero entries of a dense [nz, ny, nx] array (activating their
        blocks, as ref() does); entries that are zero are left alone"""
        a = np.asarray(a)
        for k, j, i in zip(*np.nonzero(a)):
            self.set(f, int(i), int(j), int(k), a[k, j, i])

    def mark_cells(self, ps: Particles):
        lib().orc_mark_cells(self.h, len(ps), ps.x, ps.y, ps.z)

    def p2g(self, ps: Particles, phase: Phase, Vp=1.0):
        lib().orc_p2g(self.h, len(ps), ps.x, ps.y, ps.z, ps.u, ps.v, ps.w, ps.type,
                      _vptr(ps.volume), C.byref(phase), Vp)

    def apply_gravity(self, dt, g):
        lib().orc_apply_gravity(self.h, dt, g)

    def apply_walls(self):
        lib().orc_apply_walls(self.h)

    def project(self, phase, dt, iters=600, tol=1e-7, cdiv=0.0, stop_rule=STOP_REF_MAXABS):
        st = SolveStats()
        lib().orc_project(self.h, C.byref(phase), dt, iters, tol, cdiv, stop_rule, C.byref(st))
        return st

    def g2p(self, saved, ps: Particles, aL, aG):
        lib().orc_g2p(self.h, saved.h, len(ps), ps.x, ps.y, ps.z, ps.u, ps.v, ps.w, ps.type,
                      aL, aG)

    def advect(self, ps: Particles, dt, order=2):
        es = EscapeStats()
        n = len(ps)
        idx = np.zeros(max(n, 1), np.int32)
        ne = C.c_int64(0)
        lib().orc_advect(self.h, n, ps.x, ps.y, ps.z, ps.type, dt, order, C.byref(es),
                         _vptr(idx), C.byref(ne))
        return es, idx[: ne.value]

    # single-phase path (src/driver/sparse_ops3d.cpp, sparse_sim3d.cpp)
    def p2g_single(self, ps: Particles):
        lib().orc_p2g_single(self.h, len(ps), ps.x, ps.y, ps.z, ps.u, ps.v, ps.w)

    def project_single(self, dt, iters=300, tol=1e-6, stop_rule=STOP_REF_MAXABS):
        st = SolveStats()
        lib().orc_project_single(self.h, dt, iters, tol, stop_rule, C.byref(st))
        return st

    def step_single(self, ps: Particles, dt, alpha=0.95, gravity=-9.81, iters=300, tol=1e-6,
                    stop_rule=STOP_REF_MAXABS):
        st = StepStats()
        lib().orc_step_single(self.h, len(ps), ps.x, ps.y, ps.z, ps.u, ps.v, ps.w, dt, alpha,
                              gravity, iters, tol, stop_rule, C.byref(st))
        return st

    def step(self, ps: Particles, cfg: StepCfg):
        st = StepStats()
        lib().orc_step(self.h, len(ps), ps.x, ps.y, ps.z, ps.u, ps.v, ps.w, ps.type,
                       _vptr(ps.volume), C.byref(cfg), C.byref(st))
        return st

    def interface_diagnostics(self, phase):
        st = IfaceStats()
        lib().orc_interface_diagnostics(self.h, C.byref(phase), C.byref(st))
        return st

    def surface_tension(self, phase, dt, strength=0.02, max_delta_speed=0.25, radius=0):
        st = StStats()
        lib().orc_surface_tension(self.h, C.byref(phase), dt, strength, max_delta_speed, radius, C.byref(st))
        return st

    def render_cache(self, ps, phase, meta, path):
        """writeSparseRenderCache3D of this grid and the particles ps"""
        return lib().orc_render_cache_sparse(self.h, C.byref(phase), *_render_particles(ps), C.byref(meta),
                                             str(path).encode())

    def interface_fields(self, phase):
        """(phi, |grad phi|, curvature) per cell, arrays [nz, ny, nx]"""
        n = self.nx * self.ny * self.nz
        out = [np.zeros(n) for _ in range(3)]
        lib().orc_interface_fields(self.h, C.byref(phase), *out)
        return [a.reshape(self.nz, self.ny, self.nx) for a in out]


def phi(raw, phase):
    return lib().orc_phi(raw, C.byref(phase))


def beta(p, phase):
    return lib().orc_beta(p, C.byref(phase))


def calibrate_rho_tilde0(phase, Vp=1.0):
    return lib().orc_calibrate_rho_tilde0(C.byref(phase), Vp)


def timestep(u, v, w, dx, dt, enabled, cfl, mindt):
    st = TimestepStats()
    f = lambda a: np.ascontiguousarray(a, np.float64)
    lib().orc_timestep(len(u), f(u), f(v), f(w), dx, dt, int(enabled), cfl, mindt, C.byref(st))
    return st


def scene(kind, nx, ny, nz, dx, seed):
    """Seeded synthetic scene -> float32 positions, uint8 type."""
    n = lib().orc_scene_count(kind, nx, ny, nz)
    x = np.zeros(max(n, 1), np.float32); y = np.zeros_like(x); z = np.zeros_like(x)
    t = np.zeros(max(n, 1), np.uint8)
    lib().orc_scene_fill(kind, nx, ny, nz, dx, seed, x, y, z, t)
    return x[:n], y[:n], z[:n], t[:n]


def dense_p2g(nx, ny, nz, dx, ps: Particles, phase, Vp=1.0):
    nu, nv, nw = (nx + 1) * ny * nz, nx * (ny + 1) * nz, nx * ny * (nz + 1)
    U, MU = np.zeros(nu), np.zeros(nu)
    V, MV = np.zeros(nv), np.zeros(nv)
    W, MW = np.zeros(nw), np.zeros(nw)
    lib().orc_dense_p2g(nx, ny, nz, dx, len(ps), ps.x, ps.y, ps.z, ps.u, ps.v, ps.w, ps.type,
                        _vptr(ps.volume), C.byref(phase), Vp, U, V, W, MU, MV, MW)
    return (U.reshape(nz, ny, nx + 1), V.reshape(nz, ny + 1, nx), W.reshape(nz + 1, ny, nx),
            MU.reshape(nz, ny, nx + 1), MV.reshape(nz, ny + 1, nx), MW.reshape(nz + 1, ny, nx))


# ----------------------------------------------------------------- reference

_ref = None


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def reflib():
    global _ref
    if _ref is None:
        L = C.CDLL(REF_SO)
        L.ref_sim_create.restype = C.c_void_p
        L.ref_sim_create.argtypes = [C.c_int] * 3 + [C.c_double]
        L.ref_sim_destroy.argtypes = [C.c_void_p]
        L.ref_sim_clone.restype = C.c_void_p
        L.ref_sim_clone.argtypes = [C.c_void_p]
        L.ref_sim_config.argtypes = [C.c_void_p, _dp, _i32, _dp]
        L.ref_sim_set_particles.argtypes = [C.c_void_p, C.c_int64, _dp, _dp, _dp, _dp, _dp, _dp,
                                            _u8, C.c_void_p]
        L.ref_sim_get_particles.argtypes = [C.c_void_p] + [_dp] * 6
        L.ref_sim_step.argtypes = [C.c_void_p, _dp]
        L.ref_interface_diagnostics.argtypes = [C.c_void_p, _dp]
        L.ref_surface_tension.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_int, _dp]
        L.ref_sim_surface_tension.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_int]
        L.ref_sim_last_interface.argtypes = [C.c_void_p, _dp]
        L.ref_field_dims.argtypes = [C.c_void_p, C.c_int, _i32]
        L.ref_field_dense.argtypes = [C.c_void_p, C.c_int, _dp]
        L.ref_active_blocks.restype = C.c_int64
        L.ref_active_blocks.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.ref_set.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double]
        L.ref_p2g.argtypes = [C.c_void_p]
        L.ref_project.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_double, C.c_double,
                                  C.c_void_p]
        L.ref_g2p.argtypes = [C.c_void_p, C.c_void_p, C.c_double, C.c_double]
        L.ref_advect.argtypes = [C.c_void_p, C.c_double, C.c_int, _i32]
        L.ref_sim_init.restype = C.c_int64
        L.ref_sim_init.argtypes = [C.c_void_p, C.c_int]
        L.ref_sim_get_types.argtypes = [C.c_void_p, _u8, _dp]
        L.ref_calibrate_rho_tilde0.restype = C.c_double
        L.ref_calibrate_rho_tilde0.argtypes = [_dp, C.c_double]
        L.ref_phi.restype = C.c_double
        L.ref_phi.argtypes = [C.c_double, _dp]
        L.ref_beta.restype = C.c_double
        L.ref_beta.argtypes = [C.c_double, _dp]
        L.ref_timestep.argtypes = [C.c_int64, _dp, _dp, _dp, C.c_double, C.c_double, C.c_int,
                                   C.c_double, C.c_double, _dp]
        L.ref_mr_create.restype = C.c_void_p
        L.ref_mr_create.argtypes = [C.c_int] * 3 + [C.c_double]
        L.ref_mr_destroy.argtypes = [C.c_void_p]
        L.ref_mr_clone.restype = C.c_void_p
        L.ref_mr_clone.argtypes = [C.c_void_p]
        L.ref_mr_set_leaves.argtypes = [C.c_void_p, C.c_int, _i32]
        L.ref_mr_get_leaves.restype = C.c_int64
        L.ref_mr_get_leaves.argtypes = [C.c_void_p, C.c_void_p]
        L.ref_mr_layout_fine_box.restype = C.c_int64
        L.ref_mr_layout_fine_box.argtypes = [C.c_int] * 3 + [C.c_double, C.c_int, _i32, C.c_void_p]
        L.ref_mr_config.argtypes = [C.c_void_p, _dp, _i32, _dp]
        L.ref_mr_set_particles.argtypes = [C.c_void_p, C.c_int64] + [_dp] * 6 + [_u8, C.c_void_p]
        L.ref_mr_get_particles.restype = C.c_int64
        L.ref_mr_get_particles.argtypes = [C.c_void_p] + [C.c_void_p] * 7
        L.ref_mr_init_bubble.restype = C.c_int64
        L.ref_mr_init_bubble.argtypes = [C.c_void_p]
        L.ref_mr_step.argtypes = [C.c_void_p, _dp]
        L.ref_mr_dynamic.argtypes = [C.c_void_p]
        L.ref_mr_p2g.argtypes = [C.c_void_p]
        L.ref_mr_g2p.argtypes = [C.c_void_p, C.c_void_p, C.c_double, C.c_double]
        L.ref_mr_advect.argtypes = [C.c_void_p, C.c_double, C.c_int, _i32]
        L.ref_mr_project.argtypes = [C.c_void_p, C.c_double, _i32, _dp, _dp]
        L.ref_mr_field_dense.argtypes = [C.c_void_p, C.c_int, _dp]
        L.ref_mr_face_valid.argtypes = [C.c_void_p, C.c_int, _u8]
        L.ref_mr_set.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double]
        L.ref_mr_interface_diagnostics.argtypes = [C.c_void_p, _dp]
        L.ref_render_sparse.argtypes = [C.c_void_p, C.c_char_p, C.c_int, C.c_double]
        L.ref_render_mr.argtypes = [C.c_void_p, C.c_char_p, C.c_int, C.c_double]
        L.ref_render_manifest.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int64,
                                          _i32, _i32, _dp, C.c_void_p,
                                          np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")]
        L.ref_render_default_camera.argtypes = [C.c_int] * 3 + [C.c_double, C.c_void_p]
        _u32 = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
        L.ref_pa_apply.restype = C.c_int64
        L.ref_pa_apply.argtypes = [_i32, _dp, _u32, C.c_int64, C.c_int64] + [_dp] * 6 + [_u8, _dp, _i32]
        L.ref_sim_set_adapt.argtypes = [C.c_void_p, _u32]
        L.ref_sim_particles_full.restype = C.c_int64
        L.ref_sim_particles_full.argtypes = [C.c_void_p] + [C.c_void_p] * 8
        L.ref_sim_get_adapt.argtypes = [C.c_void_p, _i32]
        L.ref_mr_set_adapt.argtypes = [C.c_void_p, _u32]
        L.ref_mr_particles_full.restype = C.c_int64
        L.ref_mr_particles_full.argtypes = [C.c_void_p] + [C.c_void_p] * 8
        L.ref_mr_get_adapt.argtypes = [C.c_void_p, _i32]
        L.ref_mr_surface_tension.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_int, _dp]
        L.ref_mr_surface_flags.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_int]
        L.ref_mr_last_interface.argtypes = [C.c_void_p, _dp]
        L.ref1_create.restype = C.c_void_p
        L.ref1_create.argtypes = [C.c_int] * 3 + [C.c_double]
        L.ref1_destroy.argtypes = [C.c_void_p]
        L.ref1_config.argtypes = [C.c_void_p, _dp, C.c_int]
        L.ref1_init_dambreak.restype = C.c_int64
        L.ref1_init_dambreak.argtypes = [C.c_void_p]
        L.ref1_set_particles.argtypes = [C.c_void_p, C.c_int64] + [_dp] * 6
        L.ref1_get_particles.argtypes = [C.c_void_p] + [_dp] * 6
        L.ref1_step.restype = C.c_double
        L.ref1_step.argtypes = [C.c_void_p]
        L.ref1_field_dense.argtypes = [C.c_void_p, C.c_int, _dp]
        L.ref1_active_blocks.restype = C.c_int64
        L.ref1_active_blocks.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.ref1_set.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double]
        L.ref1_p2g.argtypes = [C.c_void_p]
        L.ref1_project.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_double]
        _ref = L
    return _ref


class MRGrid:
    """Oracle multires grid (pf_oracle_mr.c): 2-level layout, fine face patches."""

    def __init__(self, nx, ny, nz, dx=1.0, faithful=True, _h=None):
        self.nx, self.ny, self.nz, self.dx = nx, ny, nz, dx
        self.h = _h if _h is not None else lib().orc_mr_create(nx, ny, nz, dx, int(faithful))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_mr_destroy(self.h)
            self.h = None

    def clone(self):
        return MRGrid(self.nx, self.ny, self.nz, self.dx, _h=lib().orc_mr_clone(self.h))

    def set_leaves(self, leaves):
        lv = np.ascontiguousarray(np.asarray(leaves, np.int32).reshape(-1, 4))
        lib().orc_mr_set_leaves(self.h, len(lv), lv)

    def leaves(self):
        n = lib().orc_mr_leaf_count(self.h)
        out = np.zeros((max(n, 1), 4), np.int32)
        lib().orc_mr_get_leaves(self.h, out)
        return out[:n]

    def dims(self, f):
        nx, ny, nz = self.nx, self.ny, self.nz
        if f <= MW:
            a = f % 3
            return (nx + (a == 0), ny + (a == 1), nz + (a == 2))
        return (nx, ny, nz)

    def dense(self, f):
        fx, fy, fz = self.dims(f)
        out = np.zeros(fx * fy * fz)
        lib().orc_mr_field_dense(self.h, f, out)
        return out.reshape(fz, fy, fx)

    def face_valid(self, axis):
        fx, fy, fz = self.dims(axis)
        out = np.zeros(fx * fy * fz, np.uint8)
        lib().orc_mr_face_valid(self.h, axis, out)
        return out.reshape(fz, fy, fx)

    def set(self, f, i, j, k, v):
        lib().orc_mr_set(self.h, f, i, j, k, float(v))

    def mark_cells(self, ps):
        lib().orc_mr_mark_cells(self.h, len(ps), ps.x, ps.y, ps.z)

    def p2g(self, ps, phase, Vp=1.0):
        lib().orc_mr_p2g(self.h, len(ps), ps.x, ps.y, ps.z, ps.u, ps.v, ps.w, ps.type, _vptr(ps.volume),
                         C.byref(phase), Vp)

    def apply_gravity(self, dt, g):
        lib().orc_mr_apply_gravity(self.h, dt, g)

    def apply_walls(self):
        lib().orc_mr_apply_walls(self.h)

    def project(self, phase, dt, cfg=None):
        cfg = cfg or mr_solve_cfg()
        st = MRSolveStats()
        lib().orc_mr_project(self.h, C.byref(phase), dt, C.byref(cfg), C.byref(st))
        return st

    def system(self, phase, dt, cdiv=0.0):
        """the projection system as (cells, CSR (indptr, indices, data), rhs, pin row):
        rows in leafCells order, A's entries as the reference assembles them"""
        nnz, pin = C.c_int64(), C.c_int64()
        n = lib().orc_mr_system(self.h, C.byref(phase), dt, cdiv, 0, 0, None, None, None, None, None,
                                C.byref(nnz), C.byref(pin))
        cells = np.zeros(max(n, 1), np.int64)
        ptr = np.zeros(n + 1, np.int64)
        col = np.zeros(max(nnz.value, 1), np.int64)
        val = np.zeros(max(nnz.value, 1))
        rhs = np.zeros(max(n, 1))
        lib().orc_mr_system(self.h, C.byref(phase), dt, cdiv, n, nnz.value, _vptr(cells), _vptr(ptr), _vptr(col),
                            _vptr(val), _vptr(rhs), C.byref(nnz), C.byref(pin))
        return cells[:n], (ptr, col[:nnz.value], val[:nnz.value]), rhs[:n], pin.value

    def apply_pressure(self, phase, dt, x):
        """finish a projection from the row pressures x: leaf pressures and face update"""
        lib().orc_mr_apply_pressure(self.h, C.byref(phase), dt, np.ascontiguousarray(x, np.float64))

    def g2p(self, saved, ps, aL, aG):
        lib().orc_mr_g2p(self.h, saved.h, len(ps), ps.x, ps.y, ps.z, ps.u, ps.v, ps.w, ps.type, aL, aG)

    def advect(self, ps, dt, order=2):
        es = EscapeStats()
        lib().orc_mr_advect(self.h, len(ps), ps.x, ps.y, ps.z, ps.type, dt, order, C.byref(es))
        return es

    def render_cache(self, ps, phase, meta, path):
        """writeMRRenderCache3D of this grid and the particles ps"""
        return lib().orc_render_cache_mr(self.h, C.byref(phase), *_render_particles(ps), C.byref(meta),
                                         str(path).encode())

    def load_dense(self, f, a):
        """write the non-zero entries of a dense [z, y, x] array: fine faces, or for
        ORC_P / ORC_MK the leaf cell holding each fine cell"""
        a = np.asarray(a)
        for k, j, i in zip(*np.nonzero(a)):
            self.set(f, int(i), int(j), int(k), a[k, j, i])

    def interface_diagnostics(self, phase):
        st = IfaceStats()
        lib().orc_mr_interface_diagnostics(self.h, C.byref(phase), C.byref(st))
        return st

    def surface_tension(self, phase, dt, strength=0.02, max_delta_speed=0.25, radius=0):
        st = StStats()
        lib().orc_mr_surface_tension(self.h, C.byref(phase), dt, strength, max_delta_speed, radius, C.byref(st))
        return st


class MRStepStats(C.Structure):
    _fields_ = [("ts", TimestepStats), ("solve", MRSolveStats), ("escape", EscapeStats), ("c_div", C.c_double),
                ("iface", IfaceStats), ("surface", StStats)]


def mr_step(g, sim, ps):
    st = MRStepStats()
    lib().orc_mr_step(g.h, C.byref(sim), len(ps), ps.x, ps.y, ps.z, ps.u, ps.v, ps.w, ps.type, _vptr(ps.volume),
                      C.byref(st))
    return st


def mr_dynamic_refinement(g, sim, ps):
    lib().orc_mr_dynamic_refinement(g.h, C.byref(sim), len(ps), ps.x, ps.y, ps.z, ps.type)


def mr_layout_fine_box(nx, ny, nz, box=None, level=1):
    b = np.zeros(6, np.int32) if box is None else np.asarray(box, np.int32)
    n = lib().orc_mr_layout_fine_box(nx, ny, nz, level, int(box is not None), b, None, 0)
    out = np.zeros((max(n, 1), 4), np.int32)
    lib().orc_mr_layout_fine_box(nx, ny, nz, level, int(box is not None), b, _vptr(out), n)
    return out[:n]


def mr_layout_bubble_band(nx, ny, nz):
    n = lib().orc_mr_layout_bubble_band(nx, ny, nz, None, 0)
    out = np.zeros((max(n, 1), 4), np.int32)
    lib().orc_mr_layout_bubble_band(nx, ny, nz, _vptr(out), n)
    return out[:n]


class RefSim1:
    """The reference's single-phase SparseSim3D (src/driver/sparse_sim3d.h:5-15)."""

    def __init__(self, nx, ny, nz, dx=1.0):
        self.nx, self.ny, self.nz, self.dx = nx, ny, nz, dx
        self.h = reflib().ref1_create(nx, ny, nz, dx)
        self.n = 0

    def __del__(self):
        if getattr(self, "h", None):
            reflib().ref1_destroy(self.h)
            self.h = None

    def config(self, dt, alpha=0.95, gravity=-9.81, iters=300, tol=1e-6):
        reflib().ref1_config(self.h, np.array([dt, alpha, gravity, tol], np.float64), iters)

    def init_dambreak(self) -> Particles:
        self.n = reflib().ref1_init_dambreak(self.h)
        a = [np.zeros(self.n) for _ in range(6)]
        reflib().ref1_get_particles(self.h, *a)
        return Particles(*a)

    def set_particles(self, ps: Particles):
        self.n = len(ps)
        reflib().ref1_set_particles(self.h, len(ps), ps.x, ps.y, ps.z, ps.u, ps.v, ps.w)

    def get_particles(self) -> Particles:
        a = [np.zeros(self.n) for _ in range(6)]
        reflib().ref1_get_particles(self.h, *a)
        return Particles(*a)

    def step(self):
        return reflib().ref1_step(self.h)

    def dims(self, f):
        nx, ny, nz = self.nx, self.ny, self.nz
        return {0: (nx + 1, ny, nz), 1: (nx, ny + 1, nz), 2: (nx, ny, nz + 1),
                3: (nx + 1, ny, nz), 4: (nx, ny + 1, nz), 5: (nx, ny, nz + 1)}.get(f, (nx, ny, nz))

    def dense(self, f):
        nx, ny, nz = self.dims(f)
        out = np.zeros(nx * ny * nz)
        reflib().ref1_field_dense(self.h, f, out)
        return out.reshape(nz, ny, nx)

    def active_blocks(self, f):
        n = reflib().ref1_active_blocks(self.h, f, None)
        out = np.zeros(max(n, 1), np.int32)
        reflib().ref1_active_blocks(self.h, f, _vptr(out))
        return out[:n]

    def set(self, f, i, j, k, v):
        reflib().ref1_set(self.h, f, i, j, k, float(v))

    def p2g(self):
        reflib().ref1_p2g(self.h)

    def project(self, dt, iters=300, tol=1e-6):
        reflib().ref1_project(self.h, dt, iters, tol)


class RefSim:
    """The reference's SparseSim3DTP (src/driver/sparse_sim3d_tp.h:8-177)."""

    def __init__(self, nx, ny, nz, dx=1.0, _h=None):
        self.nx, self.ny, self.nz, self.dx = nx, ny, nz, dx
        self.h = _h if _h is not None else reflib().ref_sim_create(nx, ny, nz, dx)

    def __del__(self):
        if getattr(self, "h", None):
            reflib().ref_sim_destroy(self.h)
            self.h = None

    def clone(self):
        return RefSim(self.nx, self.ny, self.nz, self.dx, _h=reflib().ref_sim_clone(self.h))

    def config(self, cfg: StepCfg):
        c = np.array([cfg.dt, cfg.gravity, cfg.Vp, cfg.alpha_liquid, cfg.alpha_gas,
                      cfg.adaptive_cfl, cfg.adaptive_min_dt, cfg.cg_tol, cfg.c_div_strength,
                      cfg.liquid_volume_target], np.float64)
        ic = np.array([cfg.adaptive_timestep, cfg.advection_order, cfg.cg_iters,
                       cfg.c_div_volume_correction], np.int32)
        reflib().ref_sim_config(self.h, c, ic, cfg.phase.as_array())
        reflib().ref_sim_surface_tension(self.h, cfg.surface_tension, cfg.surface_tension_strength,
                                         cfg.surface_tension_max_delta_speed, cfg.surface_tension_smoothing_radius)

    def interface_diagnostics(self):
        out = np.zeros(11)
        reflib().ref_interface_diagnostics(self.h, out)
        return out

    def surface_tension(self, dt, strength=0.02, max_delta_speed=0.25, radius=0):
        out = np.zeros(14)
        reflib().ref_surface_tension(self.h, dt, strength, max_delta_speed, radius, out)
        return out

    def last_interface(self):
        """(interface_diagnostics_last, surface_tension_stats_last) after a step"""
        out = np.zeros(25)
        reflib().ref_sim_last_interface(self.h, out)
        return out[:11], out[11:]

    def render_cache(self, path, frame=0, time=0.0):
        return reflib().ref_render_sparse(self.h, str(path).encode(), frame, time)

    def set_adapt(self, cfg: AdaptCfg):
        reflib().ref_sim_set_adapt(self.h, cfg.words())

    def adapt_stats(self):
        out = np.zeros(22, np.int32)
        reflib().ref_sim_get_adapt(self.h, out)
        return list(out[:21])

    def particles_full(self) -> Particles:
        n = reflib().ref_sim_particles_full(self.h, *([None] * 8))
        a = [np.zeros(max(n, 1)) for _ in range(7)]
        t = np.zeros(max(n, 1), np.uint8)
        reflib().ref_sim_particles_full(self.h, *[_vptr(q) for q in a[:6]], _vptr(t), _vptr(a[6]))
        return Particles(*(q[:n] for q in a[:6]), type=t[:n], volume=a[6][:n])

    def set_particles(self, ps: Particles):
        self.n = len(ps)
        reflib().ref_sim_set_particles(self.h, len(ps), ps.x, ps.y, ps.z, ps.u, ps.v, ps.w,
                                       ps.type, _vptr(ps.volume))

    def get_particles(self, like: Particles) -> Particles:
        n = len(like)
        a = [np.zeros(n) for _ in range(6)]
        reflib().ref_sim_get_particles(self.h, *a)
        return Particles(*a, type=like.type.copy(),
                         volume=None if like.volume is None else like.volume.copy())

    def init_scene(self, kind):
        """Run the reference's own initialiser; returns (Particles, Phase)."""
        n = reflib().ref_sim_init(self.h, kind)
        t = np.zeros(max(n, 1), np.uint8)
        ph = np.zeros(4)
        reflib().ref_sim_get_types(self.h, t, ph)
        a = [np.zeros(n) for _ in range(6)]
        reflib().ref_sim_get_particles(self.h, *a)
        self.n = n
        return Particles(*a, type=t[:n]), Phase(*ph)

    def step(self):
        out = np.zeros(13)
        reflib().ref_sim_step(self.h, out)
        return out

    def dims(self, f):
        d = np.zeros(3, np.int32)
        reflib().ref_field_dims(self.h, f, d)
        return tuple(int(a) for a in d)

    def dense(self, f):
        nx, ny, nz = self.dims(f)
        out = np.zeros(nx * ny * nz)
        reflib().ref_field_dense(self.h, f, out)
        return out.reshape(nz, ny, nx)

    def active_blocks(self, f):
        n = reflib().ref_active_blocks(self.h, f, None)
        out = np.zeros(max(n, 1), np.int32)
        reflib().ref_active_blocks(self.h, f, _vptr(out))
        return out[:n]

    def set(self, f, i, j, k, v):
        reflib().ref_set(self.h, f, i, j, k, float(v))

    def p2g(self):
        reflib().ref_p2g(self.h)

    def project(self, dt, iters=600, tol=1e-7, cdiv=0.0):
        ms = C.c_double(0)
        reflib().ref_project(self.h, dt, iters, tol, cdiv, C.byref(ms))
        return ms.value

    def g2p(self, saved, aL, aG):
        reflib().ref_g2p(self.h, saved.h, aL, aG)

    def advect(self, dt, order=2):
        st = np.zeros(9, np.int32)
        reflib().ref_advect(self.h, dt, order, st)
        return st


class RefMR:
    """The reference's MRSim3DTP (src/driver/multires_sim3d_tp.h) and its stage functions."""

    def __init__(self, nx, ny, nz, dx=1.0, _h=None):
        self.nx, self.ny, self.nz, self.dx = nx, ny, nz, dx
        self.h = _h if _h is not None else reflib().ref_mr_create(nx, ny, nz, dx)

    def __del__(self):
        if getattr(self, "h", None):
            reflib().ref_mr_destroy(self.h)
            self.h = None

    def clone(self):
        return RefMR(self.nx, self.ny, self.nz, self.dx, _h=reflib().ref_mr_clone(self.h))

    def set_leaves(self, leaves):
        lv = np.ascontiguousarray(np.asarray(leaves, np.int32).reshape(-1, 4))
        reflib().ref_mr_set_leaves(self.h, len(lv), lv)

    def leaves(self):
        n = reflib().ref_mr_get_leaves(self.h, None)
        out = np.zeros((max(n, 1), 4), np.int32)
        reflib().ref_mr_get_leaves(self.h, _vptr(out))
        return out[:n]

    def config(self, dt=0.02, phase=None, gravity=-9.81, Vp=1.0, aL=0.95, aG=0.95, adaptive=0, cfl=0.5,
               mindt=1e-5, order=2, cg_iters=160, cg_tol=1e-7, cg_rel_tol=1e-6, jacobi=1, flexible=0, restart=1,
               growth=10.0, sweeps=0, omega=0.67, min_omega=0.05, dynamic=1, ppad=1, gpad=2, hyst=1,
               max_fine=0, cdiv_corr=0, cdiv_strength=1.0, lv_target=0.0):
        phase = phase or Phase(1.0, 0.01, 1.0, 1.0)
        c = np.array([dt, gravity, Vp, aL, aG, cfl, mindt, cg_tol, cg_rel_tol, growth, omega, min_omega,
                      cdiv_strength, lv_target], np.float64)
        ic = np.array([adaptive, order, cg_iters, jacobi, flexible, restart, sweeps, dynamic, ppad, gpad, hyst,
                       max_fine, cdiv_corr], np.int32)
        reflib().ref_mr_config(self.h, c, ic, phase.as_array())

    def set_particles(self, ps):
        reflib().ref_mr_set_particles(self.h, len(ps), ps.x, ps.y, ps.z, ps.u, ps.v, ps.w, ps.type,
                                      _vptr(ps.volume))

    def particles(self):
        n = reflib().ref_mr_get_particles(self.h, *([None] * 7))
        a = [np.zeros(n) for _ in range(6)]
        t = np.zeros(max(n, 1), np.uint8)
        reflib().ref_mr_get_particles(self.h, *[_vptr(q) for q in a], _vptr(t))
        return Particles(*a, type=t[:n])

    def init_bubble(self):
        return reflib().ref_mr_init_bubble(self.h)

    def step(self):
        out = np.zeros(13)
        reflib().ref_mr_step(self.h, out)
        return out

    def dynamic(self):
        reflib().ref_mr_dynamic(self.h)

    def interface_diagnostics(self):
        out = np.zeros(11)
        reflib().ref_mr_interface_diagnostics(self.h, out)
        return out

    def surface_tension(self, dt, strength=0.02, max_delta_speed=0.25, radius=0):
        out = np.zeros(14)
        reflib().ref_mr_surface_tension(self.h, dt, strength, max_delta_speed, radius, out)
        return out

    def render_cache(self, path, frame=0, time=0.0):
        return reflib().ref_render_mr(self.h, str(path).encode(), frame, time)

    def set_adapt(self, cfg: AdaptCfg):
        reflib().ref_mr_set_adapt(self.h, cfg.words())

    def adapt_stats(self):
        out = np.zeros(22, np.int32)
        reflib().ref_mr_get_adapt(self.h, out)
        return list(out[:21])

    def particles_full(self) -> Particles:
        n = reflib().ref_mr_particles_full(self.h, *([None] * 8))
        a = [np.zeros(max(n, 1)) for _ in range(7)]
        t = np.zeros(max(n, 1), np.uint8)
        reflib().ref_mr_particles_full(self.h, *[_vptr(q) for q in a[:6]], _vptr(t), _vptr(a[6]))
        return Particles(*(q[:n] for q in a[:6]), type=t[:n], volume=a[6][:n])

    def surface_flags(self, on, strength=0.02, max_delta_speed=0.25, radius=0):
        reflib().ref_mr_surface_flags(self.h, int(on), strength, max_delta_speed, radius)

    def last_interface(self):
        """(interface_diagnostics_last, surface_tension_stats_last) of the last step"""
        out = np.zeros(25)
        reflib().ref_mr_last_interface(self.h, out)
        return out[:11], out[11:]

    def p2g(self):
        reflib().ref_mr_p2g(self.h)

    def g2p(self, saved, aL, aG):
        reflib().ref_mr_g2p(self.h, saved.h, aL, aG)

    def advect(self, dt, order=2):
        st = np.zeros(8, np.int32)
        reflib().ref_mr_advect(self.h, dt, order, st)
        return st

    def project(self, dt, cfg: MRSolveCfg):
        ic = np.array([cfg.max_iterations, cfg.use_jacobi, cfg.use_flexible_beta, cfg.adaptive_restart,
                       cfg.relaxation_sweeps, cfg.use_coarse_correction, cfg.coarse_correction_iterations,
                       cfg.coarse_correction_sweeps, cfg.use_coarse_preconditioner,
                       cfg.coarse_preconditioner_iterations, cfg.coarse_preconditioner_auto_disable,
                       cfg.coarse_preconditioner_auto_disable_after], np.int32)
        c = np.array([cfg.absolute_tolerance, cfg.relative_tolerance, cfg.restart_growth_threshold,
                      cfg.relaxation_omega, cfg.relaxation_min_omega, cfg.divergence_correction,
                      cfg.coarse_correction_absolute_tolerance, cfg.coarse_correction_relative_tolerance,
                      cfg.coarse_correction_min_scale, cfg.coarse_preconditioner_absolute_tolerance,
                      cfg.coarse_preconditioner_relative_tolerance, cfg.coarse_preconditioner_scale,
                      cfg.coarse_preconditioner_min_rz_gain, cfg.coarse_preconditioner_max_work_ratio], np.float64)
        out = np.zeros(30)
        reflib().ref_mr_project(self.h, dt, ic, c, out)
        return out

    def dims(self, f):
        nx, ny, nz = self.nx, self.ny, self.nz
        if f <= MW:
            a = f % 3
            return (nx + (a == 0), ny + (a == 1), nz + (a == 2))
        return (nx, ny, nz)

    def dense(self, f):
        fx, fy, fz = self.dims(f)
        out = np.zeros(fx * fy * fz)
        reflib().ref_mr_field_dense(self.h, f, out)
        return out.reshape(fz, fy, fx)

    def face_valid(self, axis):
        fx, fy, fz = self.dims(axis)
        out = np.zeros(fx * fy * fz, np.uint8)
        reflib().ref_mr_face_valid(self.h, axis, out)
        return out.reshape(fz, fy, fx)

    def set(self, f, i, j, k, v):
        reflib().ref_mr_set(self.h, f, i, j, k, float(v))


def ref_mr_layout_fine_box(nx, ny, nz, box=None, dx=1.0):
    b = np.zeros(6, np.int32) if box is None else np.asarray(box, np.int32)
    n = reflib().ref_mr_layout_fine_box(nx, ny, nz, dx, int(box is not None), b, None)
    out = np.zeros((max(n, 1), 4), np.int32)
    reflib().ref_mr_layout_fine_box(nx, ny, nz, dx, int(box is not None), b, _vptr(out))
    return out[:n]
