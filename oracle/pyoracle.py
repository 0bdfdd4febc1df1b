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
                ("liquid_volume_target", C.c_double), ("phase", Phase)]


class StepStats(C.Structure):
    _fields_ = [("ts", TimestepStats), ("solve", SolveStats), ("escape", EscapeStats),
                ("liquid_volume", C.c_double), ("c_div", C.c_double),
                ("ms_mark", C.c_double), ("ms_p2g", C.c_double),
                ("ms_save_gravity_walls", C.c_double), ("ms_project", C.c_double),
                ("ms_g2p", C.c_double), ("ms_advect", C.c_double), ("ms_total", C.c_double)]


def build(ref: bool = True) -> None:
    """Compile liboracle.so (and _ref/libpfref.so when the reference is present)."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    if ref and os.path.isdir("/root/reference/src"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build(ref=False)
        L = C.CDLL(ORACLE_SO)
        L.orc_grid_create.restype = C.c_void_p
        L.orc_grid_create.argtypes = [C.c_int] * 3 + [C.c_double] * 4 + [C.c_int]
        L.orc_grid_destroy.argtypes = [C.c_void_p]
        L.orc_grid_clone.restype = C.c_void_p
        L.orc_grid_clone.argtypes = [C.c_void_p]
        L.orc_field_dims.argtypes = [C.c_void_p, C.c_int, _i32]
        L.orc_get.restype = C.c_double
        L.orc_get.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_set.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double]
        L.orc_field_dense.argtypes = [C.c_void_p, C.c_int, _dp]
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
        L.orc_p2g_single.argtypes = [C.c_void_p, C.c_int64, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_project_single.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_double, C.c_int,
                                         C.POINTER(SolveStats)]
        L.orc_step_single.argtypes = ([C.c_void_p, C.c_int64] + [_dp] * 6 +
                                      [C.c_double, C.c_double, C.c_double, C.c_int, C.c_double,
                                       C.c_int, C.POINTER(StepStats)])
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
