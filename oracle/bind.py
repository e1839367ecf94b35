"""ctypes binding of oracle/liboracle.so (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, as the checker of the HIP path and as the timed CPU baseline.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from typing import Dict, Optional

import numpy as np

from disco_amd.capi import DhGrid, DhParams, GridView, make_params, PLANET_STRIDE
from disco_amd.state_io import MeshState

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_LIB = None


def build() -> None:
    subprocess.run(["make", "-s", "-f", os.path.join(HERE, "Makefile")], check=True)


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.or_create.argtypes = [C.POINTER(DhGrid), C.POINTER(DhParams)]
        L.or_create.restype = vp
        L.or_destroy.argtypes = [vp]
        L.or_ntot.argtypes = [vp]
        L.or_ntot.restype = C.c_long
        L.or_set_state.argtypes = [vp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.or_get_state.argtypes = [vp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.or_set_planets.argtypes = [vp, C.c_int, _dp, C.c_int]
        L.or_getmindt.argtypes = [vp]
        L.or_getmindt.restype = C.c_double
        L.or_timestep.argtypes = [vp, C.c_double]
        L.or_plm_phi.argtypes = [vp, _dp]
        L.or_faces.argtypes = [vp, C.c_int, _ip, _ip, _ip, _dp, _dp, _dp]
        L.or_faces.restype = C.c_long
        L.or_plm_trans.argtypes = [vp, C.c_int, _dp]
        L.or_riemann_flux.argtypes = [vp, _dp, _dp, _dp, _dp, C.c_double, _dp]
        L.or_prim2cons.argtypes = [vp, _dp, _dp, _dp, C.c_double]
        L.or_cons2prim.argtypes = [vp, _dp, _dp, _dp, C.c_double]
        _LIB = L
    return _LIB


def _p(a):
    return None if a is None else a.ctypes.data_as(_dp)


class Oracle:
    def __init__(self, cfg, par: Dict[str, object], state: MeshState,
                 planets: Optional[np.ndarray] = None):
        self.L = lib()
        self.cfg = cfg
        self.grid = GridView(state)
        self.params = make_params(cfg, par)
        self.o = self.L.or_create(C.byref(self.grid.g), C.byref(self.params))
        self.nq, self.nf = cfg.nq, cfg.nf
        self.state = state.copy()
        self.set_state(state)
        pl = state.planets if planets is None else planets
        self.set_planets(0, pl)
        self.set_planets(1, pl)

    def set_state(self, s: MeshState):
        self.L.or_set_state(self.o, _p(np.ascontiguousarray(s.prim)),
                            _p(np.ascontiguousarray(s.cons)), _p(s.piph), _p(s.dphi),
                            _p(s.wiph), _p(np.ascontiguousarray(s.phi_flux)) if self.nf else None)

    def get_state(self) -> MeshState:
        out = self.state.copy()
        self.L.or_get_state(self.o, _p(out.prim), _p(out.cons), _p(out.piph), _p(out.dphi),
                            _p(out.wiph), _p(out.phi_flux) if self.nf else None)
        return out

    def set_planets(self, substep: int, planets: np.ndarray):
        pl = np.ascontiguousarray(planets, dtype=np.float64).reshape(-1, PLANET_STRIDE)
        self.L.or_set_planets(self.o, substep, _p(pl), pl.shape[0])

    def getmindt(self) -> float:
        return self.L.or_getmindt(self.o)

    def timestep(self, dt: float):
        self.L.or_timestep(self.o, dt)

    def plm_phi(self) -> np.ndarray:
        g = np.zeros((self.nq, self.state.ntot))
        self.L.or_plm_phi(self.o, _p(g))
        return g

    def plm_trans(self, dim: int) -> np.ndarray:
        g = np.zeros((self.nq, self.state.ntot))
        self.L.or_plm_trans(self.o, dim, _p(g))
        return g

    def faces(self, dim: int):
        n = self.L.or_faces(self.o, dim, None, None, None, None, None, None)
        L = np.zeros(n, np.int32); R = np.zeros(n, np.int32); LR = np.zeros(n, np.int32)
        dA = np.zeros(n); phic = np.zeros(n); dphi = np.zeros(n)
        self.L.or_faces(self.o, dim, L.ctypes.data_as(_ip), R.ctypes.data_as(_ip),
                        LR.ctypes.data_as(_ip), _p(dA), _p(phic), _p(dphi))
        return dict(L=L, R=R, LRtype=LR, dA=dA, phic=phic, dphi=dphi)

    def riemann_flux(self, primL, primR, x, n, w) -> np.ndarray:
        F = np.zeros(self.nq)
        a = [np.ascontiguousarray(v, dtype=np.float64) for v in (primL, primR, x, n)]
        self.L.or_riemann_flux(self.o, _p(a[0]), _p(a[1]), _p(a[2]), _p(a[3]), w, _p(F))
        return F

    def prim2cons(self, prim, x, dV) -> np.ndarray:
        out = np.zeros(self.nq)
        a = [np.ascontiguousarray(v, dtype=np.float64) for v in (prim, x)]
        self.L.or_prim2cons(self.o, _p(a[0]), _p(out), _p(a[1]), dV)
        return out

    def cons2prim(self, cons, x, dV) -> np.ndarray:
        out = np.zeros(self.nq)
        a = [np.ascontiguousarray(v, dtype=np.float64) for v in (cons, x)]
        self.L.or_cons2prim(self.o, _p(a[0]), _p(out), _p(a[1]), dV)
        return out

    def close(self):
        if self.o:
            self.L.or_destroy(self.o)
            self.o = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
