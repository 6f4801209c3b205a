"""ctypes binding of the dsph C ABI (include/dsph.h).

``System`` mirrors the reference ``struct System`` (hydrocodes/sph/sph.hpp:72-380):
``add`` / ``seed_damage`` / ``init`` / ``timestep`` / ``step`` and the phase
methods, with whole-field ``get`` / ``set`` in place of direct vector access.
All arrays cross the boundary in IC (insertion) order.  There is no CPU
fallback: if ``libdsph.so`` is missing or no HIP device is visible the
constructor raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdsph.so")

MIXED, FP64, MIXED_FAST = 0, 1, 2
(F_POS, F_VEL, F_ACC, F_MASS, F_RHO, F_U, F_DUDT, F_P, F_CS, F_CSIG, F_HH, F_EPS_WORK, F_D,
 F_EPS_ACT, F_SIGMAX, F_PF, F_RART, F_DSCALE, F_S, F_DSDT, F_MAT, F_FIXED, F_SEG, F_DUDT_OLD,
 F_DSDT_OLD) = range(25)
FIELD = {"pos": F_POS, "vel": F_VEL, "acc": F_ACC, "mass": F_MASS, "rho": F_RHO, "u": F_U,
         "dudt": F_DUDT, "P": F_P, "cs": F_CS, "csig": F_CSIG, "hh": F_HH, "eps_work": F_EPS_WORK,
         "D": F_D, "eps_act": F_EPS_ACT, "sigmax": F_SIGMAX, "Pf": F_PF, "Rart": F_RART,
         "dscale": F_DSCALE, "S": F_S, "dSdt": F_DSDT, "mat": F_MAT, "fixed": F_FIXED, "seg": F_SEG,
         "dudt_old": F_DUDT_OLD, "dSdt_old": F_DSDT_OLD}
WIDTH = {F_POS: 3, F_VEL: 3, F_ACC: 3, F_S: 6, F_DSDT: 6, F_DSDT_OLD: 6}
DTYPE = {F_MAT: np.int32, F_SEG: np.int32, F_FIXED: np.uint8}

# C entry points declared in include/dsph.h (checked by the export test).
EXPORTS = ["dsph_version", "dsph_last_error", "dsph_create", "dsph_batch_create", "dsph_destroy",
           "dsph_upload", "dsph_count", "dsph_seed_damage", "dsph_get_field", "dsph_set_field",
           "dsph_init", "dsph_compute_density_h", "dsph_compute_forces", "dsph_grow_damage",
           "dsph_timestep", "dsph_step", "dsph_step_dts", "dsph_run", "dsph_synchronize", "dsph_set_clock",
           "dsph_diagnostics", "dsph_set_profiling", "dsph_get_profile", "dsph_reset_profile",
           "dsph_get_tuning", "dsph_set_tuning", "dsph_autotune", "dsph_energy_sums_noalloc"]


class Material(C.Structure):
    _fields_ = [("type", C.c_int32), ("gamma", C.c_double), ("rho0", C.c_double),
                ("A", C.c_double), ("B", C.c_double), ("a", C.c_double), ("b", C.c_double),
                ("alpha", C.c_double), ("beta", C.c_double), ("u0", C.c_double),
                ("uiv", C.c_double), ("ucv", C.c_double), ("G", C.c_double), ("Y", C.c_double),
                ("wm", C.c_double), ("wk", C.c_double), ("Emod", C.c_double)]


class Params(C.Structure):
    _fields_ = [("eta", C.c_double), ("h_init", C.c_double), ("alpha", C.c_double),
                ("beta", C.c_double), ("eps", C.c_double), ("alpha_u", C.c_double),
                ("cfl", C.c_double), ("eps_as", C.c_double), ("damage_on", C.c_int32),
                ("periodic", C.c_int32 * 3), ("lo", C.c_double * 3), ("hi", C.c_double * 3),
                ("g", C.c_double), ("damp", C.c_double), ("t_end", C.c_double)]


class Tuning(C.Structure):
    _fields_ = [("block_pair", C.c_int32), ("block_point", C.c_int32), ("cell_factor", C.c_double),
                ("xcd_remap", C.c_int32), ("cell_capacity_per_particle", C.c_int32),
                ("neighbor_capacity", C.c_int32), ("list_skin", C.c_double)]


class Diag(C.Structure):
    _fields_ = [("n", C.c_int64), ("mass", C.c_double), ("e_kin", C.c_double),
                ("e_int", C.c_double), ("e_tot", C.c_double), ("mom", C.c_double * 3),
                ("vmax", C.c_double), ("umax", C.c_double), ("hmin", C.c_double),
                ("hmax", C.c_double), ("rho_min", C.c_double), ("rho_max", C.c_double),
                ("dmax", C.c_double), ("t", C.c_double), ("dt_next", C.c_double),
                ("steps", C.c_int64), ("density_rounds", C.c_int32), ("nonfinite", C.c_int32),
                ("list_walk_steps", C.c_int64), ("list_overflow", C.c_int64),
                ("status", C.c_int32), ("pad_", C.c_int32),
                ("density_lanes", C.c_int64), ("density_rewalks", C.c_int64),
                ("density_waves", C.c_int64), ("density_rewalk_waves", C.c_int64)]


class Profile(C.Structure):
    _fields_ = [("nphases", C.c_int32), ("name", (C.c_char * 24) * 24),
                ("total_ms", C.c_double * 24), ("launches", C.c_int64 * 24)]


class DsphError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"dsph error {code}: {msg}")
        self.code = code


_LIB = None


def load_library(path: str = LIB_PATH) -> C.CDLL:
    """Load libdsph.so; raises if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is missing: build it with `python -m deorbit_amd.build`")
    lib = C.CDLL(path)
    vp, dp = C.c_void_p, C.POINTER(C.c_double)
    lib.dsph_version.restype = C.c_char_p
    lib.dsph_last_error.restype = C.c_char_p
    lib.dsph_create.argtypes = [C.POINTER(Params), C.POINTER(Material), C.c_int32, C.c_int32, C.POINTER(vp)]
    lib.dsph_batch_create.argtypes = [C.POINTER(Params), C.c_int32, C.POINTER(Material), C.c_int32,
                                      C.c_int32, C.POINTER(vp)]
    lib.dsph_destroy.argtypes = [vp]
    lib.dsph_destroy.restype = None
    lib.dsph_upload.argtypes = [vp, C.c_int64, dp, dp, dp, dp, C.POINTER(C.c_int32),
                                C.POINTER(C.c_uint8), C.POINTER(C.c_int32)]
    lib.dsph_count.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int32)]
    lib.dsph_seed_damage.argtypes = [vp, C.c_int32, C.c_double, C.c_uint32]
    lib.dsph_get_field.argtypes = [vp, C.c_int32, vp]
    lib.dsph_set_field.argtypes = [vp, C.c_int32, vp]
    for fn in ("dsph_init", "dsph_compute_density_h", "dsph_compute_forces", "dsph_synchronize",
               "dsph_reset_profile"):
        getattr(lib, fn).argtypes = [vp]
    lib.dsph_grow_damage.argtypes = [vp, C.c_double]
    lib.dsph_timestep.argtypes = [vp, dp]
    lib.dsph_step.argtypes = [vp, C.c_double]
    lib.dsph_step_dts.argtypes = [vp, dp]
    lib.dsph_run.argtypes = [vp, C.c_int32, C.c_double, C.c_int32]
    lib.dsph_set_clock.argtypes = [vp, C.c_int32, C.c_double, C.c_int64, C.c_double]
    lib.dsph_diagnostics.argtypes = [vp, C.c_int32, C.POINTER(Diag)]
    lib.dsph_set_profiling.argtypes = [vp, C.c_int32]
    lib.dsph_get_profile.argtypes = [vp, C.POINTER(Profile)]
    lib.dsph_get_tuning.argtypes = [vp, C.POINTER(Tuning)]
    lib.dsph_set_tuning.argtypes = [vp, C.POINTER(Tuning)]
    lib.dsph_autotune.argtypes = [vp, C.c_int32, C.POINTER(Tuning)]
    lib.dsph_energy_sums_noalloc.argtypes = [vp, vp, vp, vp, vp, C.c_int64, vp, vp]
    _LIB = lib
    return lib


def material_struct(m: dict) -> Material:
    s = Material()
    s.type = 1 if m.get("type", "tillotson") == "tillotson" else 0
    s.gamma = float(m.get("gamma", 1.4))
    for k in ("rho0", "A", "B", "a", "b", "alpha", "beta", "u0", "uiv", "ucv", "G", "Y", "wm", "wk"):
        setattr(s, k, float(m.get(k, 0.0)))
    s.Emod = float(m.get("Emod", 0.0))
    return s


def params_struct(p: dict) -> Params:
    s = Params()
    for k in ("eta", "h_init", "alpha", "beta", "eps", "alpha_u", "cfl", "eps_as"):
        setattr(s, k, float(p[k]))
    s.damage_on = int(bool(p.get("damage_on", False)))
    per = p.get("periodic", (0, 0, 0))
    for a in range(3):
        s.periodic[a] = int(bool(per[a]))
        s.lo[a] = float(p["lo"][a])
        s.hi[a] = float(p["hi"][a])
    s.g = float(p.get("g", 0.0))
    s.damp = float(p.get("damp", 1.0))
    s.t_end = float(p.get("t_end", 0.0))
    return s


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


class System:
    """Device-resident SPH system behind the C ABI (one segment, or a batch)."""

    def __init__(self, params, materials: list[dict], precision: int = MIXED):
        self.lib = load_library()
        self.h = C.c_void_p()
        mats = (Material * len(materials))(*[material_struct(m) for m in materials])
        if isinstance(params, dict):
            prm = params_struct(params)
            self._check(self.lib.dsph_create(C.byref(prm), mats, len(materials), precision,
                                              C.byref(self.h)))
            self.nseg = 1
        else:
            plist = (Params * len(params))(*[params_struct(p) for p in params])
            self._check(self.lib.dsph_batch_create(plist, len(params), mats, len(materials),
                                                    precision, C.byref(self.h)))
            self.nseg = len(params)
        self.precision = precision

    def _check(self, rc: int):
        if rc != 0:
            raise DsphError(rc, self.lib.dsph_last_error().decode())

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            self.lib.dsph_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- state -----------------------------------------------------------
    @property
    def n(self) -> int:
        n = C.c_int64()
        self._check(self.lib.dsph_count(self.h, C.byref(n), None))
        return n.value

    def add(self, pos, vel, mass, u, mat, fixed=None, seg=None):
        pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
        vel = np.ascontiguousarray(vel, np.float64).reshape(-1, 3)
        mass = np.ascontiguousarray(mass, np.float64)
        u = np.ascontiguousarray(u, np.float64)
        mat = np.ascontiguousarray(mat, np.int32)
        n = len(mass)
        fx = np.zeros(n, np.uint8) if fixed is None else np.ascontiguousarray(fixed, np.uint8)
        sg = None if seg is None else np.ascontiguousarray(seg, np.int32)
        self._check(self.lib.dsph_upload(self.h, n, _ptr(pos, C.c_double), _ptr(vel, C.c_double),
                                         _ptr(mass, C.c_double), _ptr(u, C.c_double),
                                         _ptr(mat, C.c_int32), _ptr(fx, C.c_uint8),
                                         None if sg is None else _ptr(sg, C.c_int32)))

    def add_ic(self, ic: dict, seg=None):
        self.add(ic["pos"], ic["vel"], ic["mass"], ic["u"], ic["mat"], ic["fixed"], seg)

    def seed_damage(self, dx: float, seed: int = 12345, seg: int = -1):
        self._check(self.lib.dsph_seed_damage(self.h, seg, dx, seed))

    def get(self, field) -> np.ndarray:
        f = FIELD[field] if isinstance(field, str) else field
        w = WIDTH.get(f, 1)
        out = np.empty(self.n * w, DTYPE.get(f, np.float64))
        self._check(self.lib.dsph_get_field(self.h, f, out.ctypes.data_as(C.c_void_p)))
        return out.reshape(-1, w) if w > 1 else out

    def set(self, field, values):
        f = FIELD[field] if isinstance(field, str) else field
        w = WIDTH.get(f, 1)
        a = np.ascontiguousarray(values, DTYPE.get(f, np.float64)).reshape(-1)
        if a.size != self.n * w:
            raise ValueError(f"field {field}: expected {self.n * w} values, got {a.size}")
        self._check(self.lib.dsph_set_field(self.h, f, a.ctypes.data_as(C.c_void_p)))

    def state(self, names=None) -> dict:
        names = names or [k for k in FIELD if k not in ("dudt_old", "dSdt_old")]
        return {k: self.get(k) for k in names}

    # --- the step ---------------------------------------------------------
    def init(self):
        self._check(self.lib.dsph_init(self.h))

    def compute_density_h(self):
        self._check(self.lib.dsph_compute_density_h(self.h))

    def compute_forces(self):
        self._check(self.lib.dsph_compute_forces(self.h))

    def grow_damage(self, dt: float):
        self._check(self.lib.dsph_grow_damage(self.h, dt))

    def timestep(self):
        out = np.empty(self.nseg, np.float64)
        self._check(self.lib.dsph_timestep(self.h, _ptr(out, C.c_double)))
        return float(out[0]) if self.nseg == 1 else out

    def step(self, dt: float = 0.0):
        self._check(self.lib.dsph_step(self.h, float(dt)))

    def step_dts(self, dts):
        a = np.ascontiguousarray(dts, np.float64)
        self._check(self.lib.dsph_step_dts(self.h, _ptr(a, C.c_double)))

    def run(self, nsteps: int, dt: float = 0.0, graph: bool = True):
        self._check(self.lib.dsph_run(self.h, nsteps, float(dt), int(graph)))

    def synchronize(self):
        self._check(self.lib.dsph_synchronize(self.h))

    def set_clock(self, t: float, steps: int, dt_next: float, seg: int = -1):
        self._check(self.lib.dsph_set_clock(self.h, seg, t, steps, dt_next))

    # --- checkpoint / restart (run_ic format, lossless extension) ---------------
    CHECKPOINT_FIELDS = ("pos", "vel", "mass", "u", "mat", "fixed", "D", "eps_act", "S", "hh", "acc", "dudt",
                         "dSdt")

    def checkpoint_state(self) -> dict:
        """Every field a bit-exact restart needs, IC order, plus the clock of segment 0."""
        st = self.state(list(self.CHECKPOINT_FIELDS))
        d = self.diagnostics(0)
        st.update(t=d["t"], steps=d["steps"], dt_next=d["dt_next"])
        return st

    def restore_state(self, st: dict):
        """Load a checkpoint_state / read_checkpoint dict into a system that holds the
        same particles (same count, IC order)."""
        for f in ("D", "eps_act", "S", "hh", "acc", "dudt", "dSdt"):
            self.set(f, st[f])
        self.set_clock(st["t"], int(st.get("steps", 0)), float(st.get("dt_next", 0.0)))

    # --- diagnostics / profiling / tuning -----------------------------------
    def diagnostics(self, seg: int = 0) -> dict:
        d = Diag()
        self._check(self.lib.dsph_diagnostics(self.h, seg, C.byref(d)))
        out = {k: getattr(d, k) for k, _ in Diag._fields_}
        out["mom"] = tuple(d.mom)
        return out

    def set_profiling(self, on: bool):
        self._check(self.lib.dsph_set_profiling(self.h, int(on)))

    def reset_profile(self):
        self._check(self.lib.dsph_reset_profile(self.h))

    def profile(self) -> dict:
        p = Profile()
        self._check(self.lib.dsph_get_profile(self.h, C.byref(p)))
        return {bytes(p.name[k]).split(b"\0")[0].decode(): (p.total_ms[k], p.launches[k])
                for k in range(p.nphases)}

    def tuning(self) -> dict:
        t = Tuning()
        self._check(self.lib.dsph_get_tuning(self.h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in Tuning._fields_}

    def set_tuning(self, **kw):
        t = Tuning()
        cur = self.tuning()
        cur.update(kw)
        for k, v in cur.items():
            setattr(t, k, v)
        self._check(self.lib.dsph_set_tuning(self.h, C.byref(t)))

    def autotune(self, steps: int = 3) -> dict:
        t = Tuning()
        self._check(self.lib.dsph_autotune(self.h, steps, C.byref(t)))
        return {k: getattr(t, k) for k, _ in Tuning._fields_}


def make_system(ic: dict, precision: int = MIXED, seed_damage: int | None = 12345,
                params: dict | None = None) -> System:
    """System for one IC dict (deorbit_amd.ics), seeded like the reference drivers."""
    p = params or ic["params"]
    s = System(p, ic["materials"], precision)
    s.add_ic(ic)
    if p.get("damage_on") and seed_damage is not None:
        s.seed_damage(ic["dx"], seed_damage)
    return s


def make_batch(ic_list: list, precision: int = MIXED, seed_damage: int | None = 12345) -> System:
    """One batched context holding every IC of `ic_list` as its own segment (its own
    parameters, domain, time step and density rounds), uploaded in one call and
    seeded per segment like separate runs.  All ICs must share one material table."""
    mats = ic_list[0]["materials"]
    s = System([ic["params"] for ic in ic_list], mats, precision)
    cat = lambda k: np.concatenate([np.asarray(ic[k]) for ic in ic_list])
    seg = np.concatenate([np.full(len(ic["mass"]), k, np.int32) for k, ic in enumerate(ic_list)])
    s.add(cat("pos"), cat("vel"), cat("mass"), cat("u"), cat("mat"), cat("fixed"), seg)
    if seed_damage is not None and any(ic["params"].get("damage_on") for ic in ic_list):
        # every segment draws its flaws from its own mt19937(seed), as separate runs do
        dxs = {float(ic["dx"]) for ic in ic_list}
        if len(dxs) == 1:
            s.seed_damage(dxs.pop(), seed_damage, seg=-1)
        else:
            for k, ic in enumerate(ic_list):
                s.seed_damage(ic["dx"], seed_damage, seg=k)
    return s

