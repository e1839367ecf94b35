"""ctypes view of the C ABI in include/disco_hip.h, and the Python driver of
libdisco_hip.so.

`DiscoHip` is a thin handle over the dh_* entry points; it is what the tests
and bench.py call.  The structures below mirror dh_grid / dh_params field for
field.  The library is loaded from the package directory (built in-tree by
`python -m disco_amd.build`); a missing library is an error, never a fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Dict, Optional, Sequence

import numpy as np

from . import configs as cfgmod
from .state_io import MeshState

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdisco_hip.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class DhGrid(C.Structure):
    _fields_ = [("Nr", C.c_int), ("Nz", C.c_int), ("NgRa", C.c_int), ("NgRb", C.c_int),
                ("NgZa", C.c_int), ("NgZb", C.c_int), ("Np", _ip), ("r_jph", _dp),
                ("z_kph", _dp), ("phi_max", C.c_double)]


_PARAM_INTS = ["num_c", "num_n", "ct_mode", "hydro", "riemann_solver", "mesh_motion",
               "exact_mesh_omega", "energy_omega", "timestep_mode", "ct", "isothermal",
               "cs2_profile", "visc_flag", "alpha_flag", "rot_frame", "grav2d",
               "nobc_rmin", "nobc_rmax", "nobc_zmin", "nobc_zmax", "r_periodic",
               "z_periodic", "bc_rmin", "bc_rmax", "bc_zmin", "bc_zmax", "nozzle_type",
               "metric", "c2p_guess", "c2p_stop"]
_PARAM_DOUBLES = ["gamma_law", "plm", "cfl", "max_dt", "rho_floor", "p_floor",
                  "viscosity", "mach", "metric_par2", "rot_omega", "rot_d"]


class DhParams(C.Structure):
    _fields_ = ([(n, C.c_int) for n in _PARAM_INTS]
                + [(n, C.c_double) for n in _PARAM_DOUBLES]
                + [("init_par", C.c_double * 9), ("phi_order", C.c_double),
                   ("metric_par3", C.c_double), ("nozzle_par", C.c_double * 4)])


PLANET_STRIDE = 8
MAX_PLANETS = 8

# in.par key -> dh_params field (readpar.c:70-155 names)
_PAR_INT_KEYS = {
    "Riemann_Solver": "riemann_solver", "Mesh_Motion": "mesh_motion",
    "Exact_Mesh_Omega": "exact_mesh_omega", "Energy_Omega": "energy_omega",
    "Timestep": "timestep_mode", "Constrained_Transport": "ct",
    "Isothermal": "isothermal", "Cs2_Profile": "cs2_profile",
    "Use_Viscosity": "visc_flag", "Use_As_Alpha": "alpha_flag", "RotFrame": "rot_frame",
    "Grav2D": "grav2d", "NoBC_Rmin": "nobc_rmin", "NoBC_Rmax": "nobc_rmax",
    "NoBC_Zmin": "nobc_zmin", "NoBC_Zmax": "nobc_zmax", "R_Periodic": "r_periodic",
    "Z_Periodic": "z_periodic", "Nozzle_Type": "nozzle_type",
}
_PAR_DBL_KEYS = {
    "Adiabatic_Index": "gamma_law", "PLM": "plm", "CFL": "cfl", "Max_DT": "max_dt",
    "Density_Floor": "rho_floor", "Pressure_Floor": "p_floor", "Viscosity": "viscosity",
    "Mach_Number": "mach", "Metric_Par2": "metric_par2", "RotOmega": "rot_omega",
    "RotD": "rot_d", "Metric_Par3": "metric_par3",
}


def make_params(cfg: "cfgmod.Config", par: Dict[str, object]) -> DhParams:
    from .state_io import _DEFAULTS
    full = dict(_DEFAULTS)
    full.update(par)
    p = DhParams()
    p.num_c, p.num_n, p.ct_mode = cfg.num_c, cfg.num_n, cfg.ct_mode
    p.hydro = {"euler": cfgmod.HYDRO_EULER, "mhd": cfgmod.HYDRO_MHD,
               "grmhd2": cfgmod.HYDRO_GRMHD}[cfg.hydro]
    p.metric = cfgmod.METRICS[cfg.metric_module]
    for k, f in _PAR_INT_KEYS.items():
        # readpar.c:31-36 parses integers through %lf and a cast.
        setattr(p, f, int(float(full[k])))
    for k, f in _PAR_DBL_KEYS.items():
        setattr(p, f, float(full[k]))
    for i in range(9):
        p.init_par[i] = float(full[f"Init_Par{i}"])
    for i in range(4):
        p.nozzle_par[i] = float(full.get(f"Nozzle_Par{i + 1}", 0.0))
    p.bc_rmin, p.bc_rmax, p.bc_zmin, p.bc_zmax = cfg.dh_bc(full)
    p.phi_order = 2.0  # PHI_ORDER, planet.c:4
    # grmhd2 cons2prim's guess and stopping rule (DH_C2P_*; not reference
    # in.par keys, so absent from reference runs)
    p.c2p_guess = int(full.get("C2P_Guess", 0))
    p.c2p_stop = int(full.get("C2P_Stop", 0))
    return p


class GridView:
    """dh_grid plus the numpy arrays it points into (kept alive here)."""

    def __init__(self, s: MeshState):
        self.np_ = np.ascontiguousarray(s.np_, dtype=np.int32)
        self.r = np.ascontiguousarray(s.r_jph, dtype=np.float64)
        self.z = np.ascontiguousarray(s.z_kph, dtype=np.float64)
        self.g = DhGrid(s.nr, s.nz, *s.ng, self.np_.ctypes.data_as(_ip),
                        self.r.ctypes.data_as(_dp), self.z.ctypes.data_as(_dp), s.phi_max)


def _ptr(a: Optional[np.ndarray]):
    if a is None:
        return None
    if a.dtype != np.float64 or not a.flags["C_CONTIGUOUS"]:
        raise ValueError("state arrays must be C-contiguous float64")
    return a.ctypes.data_as(_dp)


class DhError(RuntimeError):
    pass


def load_library(path: str = LIB_PATH) -> C.CDLL:
    if not os.path.exists(path):
        raise DhError(f"{path} is missing: build it with `python -m disco_amd.build`")
    lib = C.CDLL(path)
    vp = C.c_void_p
    sig = {
        "dh_create": ([C.POINTER(DhGrid), C.POINTER(DhParams), C.POINTER(vp)], C.c_int),
        "dh_destroy": ([vp], C.c_int),
        "dh_sizes": ([vp, C.POINTER(C.c_long), _ip, _ip], C.c_int),
        "dh_upload_state": ([vp, _dp, _dp, _dp, _dp, _dp, _dp], C.c_int),
        "dh_download_state": ([vp, _dp, _dp, _dp, _dp, _dp, _dp], C.c_int),
        "dh_init_cells": ([vp], C.c_int),
        "dh_set_planets": ([vp, C.c_int, _dp, C.c_int], C.c_int),
        "dh_getmindt": ([vp, _dp], C.c_int),
        "dh_timestep": ([vp, C.c_double], C.c_int),
        "dh_sync": ([vp], C.c_int),
        "dh_enable_timers": ([vp, C.c_int], C.c_int),
        "dh_enable_graphs": ([vp, C.c_int], C.c_int),
        "dh_read_timers": ([vp, C.c_int, C.c_void_p, _dp, C.POINTER(C.c_long), _ip], C.c_int),
        "dh_face_count": ([vp, C.c_int, C.POINTER(C.c_long)], C.c_int),
        "dh_get_faces": ([vp, C.c_int, _ip, _ip, _ip, _dp, _dp], C.c_int),
        "dh_add_diagnostics": ([vp, C.c_int, C.c_double, _dp, C.c_int], C.c_int),
        "dh_get_diagnostics": ([vp, _dp], C.c_int),
        "dh_zero_diagnostics": ([vp], C.c_int),
        "dh_avg_diagnostics": ([vp, C.c_double], C.c_int),
        "dh_report": ([vp, _dp, C.c_int, _dp], C.c_int),
        "dh_set_c2p_model": ([vp, C.POINTER(C.c_float), C.c_long], C.c_int),
        "dh_c2p_harvest": ([vp, C.POINTER(C.c_float), C.POINTER(C.c_float)], C.c_int),
        "dh_c2p_predict": ([vp, C.POINTER(C.c_float)], C.c_int),
        "dh_c2p_stats": ([vp, C.POINTER(C.c_longlong)], C.c_int),
        "dh_c2p_info": ([vp, vp], C.c_int),
        "dh_c2p_train": ([C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_long, C.c_int,
                          C.c_int, C.c_float, C.c_ulonglong, C.POINTER(C.c_float),
                          C.POINTER(C.c_float)], C.c_int),
        "dh_c2p_grad": ([C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float),
                         C.c_long, C.POINTER(C.c_float), C.POINTER(C.c_float)], C.c_int),
        "dh_c2p_grad_noalloc": ([vp, vp, vp, C.c_long, vp, vp, vp], C.c_int),
        "dh_autotune": ([vp, C.c_int], C.c_int),
        "dh_ens_create": ([C.POINTER(vp), C.c_int, C.POINTER(vp)], C.c_int),
        "dh_ens_destroy": ([vp], C.c_int),
        "dh_ens_getmindt": ([vp, _dp], C.c_int),
        "dh_ens_timestep": ([vp, _dp], C.c_int),
        "dh_ens_add_diagnostics": ([vp, C.c_int, _dp, _dp, C.c_int], C.c_int),
        "dh_ens_enable_graphs": ([vp, C.c_int], C.c_int),
        "dh_ens_sync": ([vp], C.c_int),
        "dh_read_tuning": ([vp, C.c_int, C.c_void_p, _ip, _ip, _dp, _dp, _ip], C.c_int),
        "dh_read_tuning_sweep": ([vp, C.c_char_p, C.c_int, _ip, _dp, _ip], C.c_int),
        "dh_last_error": ([], C.c_char_p),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    return lib


EXPORTED_SYMBOLS = ["dh_create", "dh_destroy", "dh_sizes", "dh_upload_state",
                    "dh_download_state", "dh_init_cells", "dh_set_planets", "dh_getmindt", "dh_timestep",
                    "dh_face_count", "dh_get_faces", "dh_sync", "dh_enable_timers", "dh_enable_graphs",
                    "dh_read_timers", "dh_add_diagnostics", "dh_get_diagnostics",
                    "dh_zero_diagnostics", "dh_avg_diagnostics", "dh_report", "dh_set_c2p_model",
                    "dh_c2p_harvest", "dh_c2p_predict", "dh_c2p_stats", "dh_c2p_info", "dh_c2p_train",
                    "dh_c2p_grad", "dh_c2p_grad_noalloc", "dh_autotune", "dh_read_tuning", "dh_read_tuning_sweep",
                    "dh_ens_create", "dh_ens_destroy", "dh_ens_getmindt", "dh_ens_timestep",
                    "dh_ens_add_diagnostics", "dh_ens_enable_graphs", "dh_ens_sync",
                    "dh_last_error"]

VARIANT_NAMES = {0: "storage order, round-robin XCDs", 1: "storage order, XCD-contiguous"}


def read_tuning(lib, ctx):
    """{kernel: {variant, order, ms_chosen, ms_default, ms_variants}} of a
    context's tuner (dh_read_tuning, dh_read_tuning_sweep)."""
    cap = 64
    names = (C.c_char * 48 * cap)()
    var = (C.c_int * cap)()
    strip = (C.c_int * cap)()
    msc = (C.c_double * cap)()
    msd = (C.c_double * cap)()
    n = C.c_int()
    if lib.dh_read_tuning(ctx, cap, names, var, strip, msc, msd, C.byref(n)):
        raise DhError(f"dh_read_tuning: {lib.dh_last_error().decode()}")
    out = {}
    sw_strip = (C.c_int * 16)()
    sw_ms = (C.c_double * 16)()
    nvar = C.c_int()
    for i in range(n.value):
        v = var[i]
        name = bytes(names[i]).split(b"\0")[0].decode()
        if lib.dh_read_tuning_sweep(ctx, name.encode(), 16, sw_strip, sw_ms, C.byref(nvar)):
            raise DhError(f"dh_read_tuning_sweep: {lib.dh_last_error().decode()}")
        sweep = {VARIANT_NAMES.get(u, f"strips of {sw_strip[u]} layers"): sw_ms[u]
                 for u in range(min(nvar.value, 16))}
        out[name] = {
            "variant": v,
            "order": VARIANT_NAMES.get(v, f"strips of {strip[i]} layers"),
            "ms_chosen": msc[i], "ms_default": msd[i], "ms_variants": sweep}
    return out


_LIB: Optional[C.CDLL] = None


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        _LIB = load_library()
    return _LIB


class DiscoHip:
    """One mesh on the GPU, driven through the dh_* C ABI."""

    def __init__(self, cfg: "cfgmod.Config", par: Dict[str, object], state: MeshState,
                 planets: Optional[np.ndarray] = None, library: Optional[C.CDLL] = None):
        self.lib = library if library is not None else lib()
        self.cfg = cfg
        self.grid = GridView(state)
        self.params = make_params(cfg, par)
        self.ctx = C.c_void_p()
        self._check(self.lib.dh_create(C.byref(self.grid.g), C.byref(self.params),
                                       C.byref(self.ctx)))
        self.ntot = state.ntot
        self.nq = cfg.nq
        self.nf = cfg.nf
        self.upload(state)
        pl = state.planets if planets is None else planets
        self.set_planets(0, pl)
        self.set_planets(1, pl)

    def _check(self, rc: int):
        if rc != 0:
            msg = self.lib.dh_last_error()
            raise DhError(f"dh call failed ({rc}): {msg.decode() if msg else ''}")

    def upload(self, s: MeshState):
        phi = s.phi_flux if self.nf else None
        self._check(self.lib.dh_upload_state(
            self.ctx, _ptr(np.ascontiguousarray(s.prim)), _ptr(np.ascontiguousarray(s.cons)),
            _ptr(s.piph), _ptr(s.dphi), _ptr(s.wiph),
            _ptr(np.ascontiguousarray(phi)) if phi is not None else None))

    def download(self, s: MeshState) -> MeshState:
        out = s.copy()
        self._check(self.lib.dh_download_state(
            self.ctx, _ptr(out.prim), _ptr(out.cons), _ptr(out.piph), _ptr(out.dphi),
            _ptr(out.wiph), _ptr(out.phi_flux) if self.nf else None))
        return out

    def set_planets(self, substep: int, planets: np.ndarray):
        pl = np.ascontiguousarray(planets, dtype=np.float64).reshape(-1, PLANET_STRIDE)
        self._check(self.lib.dh_set_planets(self.ctx, substep, _ptr(pl), pl.shape[0]))

    def faces(self, dim: int):
        """(L, R, LRtype, cm[n,3], dA) of the r (1) or z (2) faces of the mesh."""
        n = C.c_long()
        self._check(self.lib.dh_face_count(self.ctx, dim, C.byref(n)))
        L = np.zeros(n.value, np.int32)
        R = np.zeros(n.value, np.int32)
        LR = np.zeros(n.value, np.int32)
        cm = np.zeros((n.value, 3))
        dA = np.zeros(n.value)
        ip = lambda a: a.ctypes.data_as(_ip)  # noqa: E731
        self._check(self.lib.dh_get_faces(self.ctx, dim, ip(L), ip(R), ip(LR), _ptr(cm), _ptr(dA)))
        return L, R, LR, cm, dA

    def add_diagnostics(self, analysis: str, dt: float, planets: np.ndarray):
        """add_diagnostics (analysis.c:52-106) of the named module with the
        given planets (the state after the step)."""
        kind, _ = cfgmod.DIAG_KINDS[analysis]
        pl = np.ascontiguousarray(planets, dtype=np.float64).reshape(-1, PLANET_STRIDE)
        self._check(self.lib.dh_add_diagnostics(self.ctx, kind, dt, _ptr(pl), pl.shape[0]))
        self._diag = analysis

    def diagnostics(self) -> np.ndarray:
        """The accumulated Qrz as [Nr*Nz, nq]."""
        nq = cfgmod.DIAG_KINDS[self._diag][1]
        out = np.zeros((len(self.grid.np_), nq))
        self._check(self.lib.dh_get_diagnostics(self.ctx, _ptr(out)))
        return out

    def avg_diagnostics(self, t_avg: float):
        self._check(self.lib.dh_avg_diagnostics(self.ctx, t_avg))

    def zero_diagnostics(self):
        self._check(self.lib.dh_zero_diagnostics(self.ctx))

    REPORT_COLUMNS = ("Torque", "Power", "Fr", "rho_min", "rhoavg_min", "PsiR", "PsiI", "Mass",
                      "Mdot", "S_R", "L1_rho", "L1_isen", "L1_B", "Br2", "aM", "bM")

    def report(self, planets: Optional[np.ndarray] = None) -> np.ndarray:
        """report.c's 16 columns after t for the current state."""
        out = np.zeros(16)
        pl = None if planets is None else np.ascontiguousarray(planets, np.float64)
        npl = 0 if pl is None else len(pl)
        self._check(self.lib.dh_report(self.ctx, _ptr(pl), npl,
                                       _ptr(out)))
        return out

    def set_c2p_model(self, params: np.ndarray):
        """Loads the learned cons2prim guess's parameters (float32, the layout
        of disco_amd/csrc/dh_c2pnet.h)."""
        p = np.ascontiguousarray(params, dtype=np.float32).ravel()
        self._check(self.lib.dh_set_c2p_model(self.ctx, p.ctypes.data_as(C.POINTER(C.c_float)),
                                              p.size))

    def c2p_harvest(self):
        """(features [ntot, 8], targets [ntot, 2]) of the current state; NaN
        rows outside the real cells."""
        f = np.zeros((self.ntot, 8), np.float32)
        t = np.zeros((self.ntot, 2), np.float32)
        fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))  # noqa: E731
        self._check(self.lib.dh_c2p_harvest(self.ctx, fp(f), fp(t)))
        return f, t

    def c2p_predict(self) -> np.ndarray:
        """The model's guesses (v^2, eta) [ntot, 2] for the current state."""
        g = np.zeros((self.ntot, 2), np.float32)
        self._check(self.lib.dh_c2p_predict(self.ctx, g.ctypes.data_as(C.POINTER(C.c_float))))
        return g

    def c2p_stats(self) -> Dict[str, int]:
        """Totals of the last cons2prim over the real cells."""
        out = (C.c_longlong * 4)()
        self._check(self.lib.dh_c2p_stats(self.ctx, out))
        return dict(zip(("cells", "iterations", "learned", "fallbacks"), list(out)))

    def c2p_info(self) -> np.ndarray:
        """The last cons2prim's record of every cell (dh_c2p_info): bits 0-5
        Newton iterations (63: the stopping rule was not met), 64 learned
        guess, 128 fall-back; 0 outside the real cells."""
        out = np.zeros(self.ntot, dtype=np.uint8)
        self._check(self.lib.dh_c2p_info(self.ctx, out.ctypes.data_as(C.c_void_p)))
        return out

    def init_cells(self):
        self._check(self.lib.dh_init_cells(self.ctx))

    def getmindt(self) -> float:
        v = C.c_double()
        self._check(self.lib.dh_getmindt(self.ctx, C.byref(v)))
        return v.value

    def timestep(self, dt: float):
        self._check(self.lib.dh_timestep(self.ctx, dt))

    def sync(self):
        self._check(self.lib.dh_sync(self.ctx))

    def enable_timers(self, on: bool = True):
        self._check(self.lib.dh_enable_timers(self.ctx, 1 if on else 0))

    def enable_graphs(self, on: bool = True):
        """Record an RK2 timestep once as a HIP graph and replay it."""
        self._check(self.lib.dh_enable_graphs(self.ctx, 1 if on else 0))

    def autotune(self, steps: int):
        """Tune the launch variant of every kernel over the next `steps`
        timesteps (dh_autotune)."""
        self._check(self.lib.dh_autotune(self.ctx, steps))

    def read_tuning(self):
        return read_tuning(self.lib, self.ctx)

    def read_timers(self):
        cap = 64
        names = (C.c_char * 48 * cap)()
        ms = (C.c_double * cap)()
        calls = (C.c_long * cap)()
        n = C.c_int()
        self._check(self.lib.dh_read_timers(self.ctx, cap, names, ms, calls, C.byref(n)))
        return [(bytes(names[i]).split(b"\0")[0].decode(), ms[i], calls[i]) for i in range(n.value)]

    def close(self):
        if self.ctx:
            self.lib.dh_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
