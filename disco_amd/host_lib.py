"""ctypes handle on a host library libdisco_host_<config>.so: the drop-in's
in.par reader, mesh set-up and the reference-signature timestep/getmindt."""
from __future__ import annotations

import ctypes as C
import os

from .capi import lib as _hip_lib  # loads libdisco_hip.so first (the host library links it)

HERE = os.path.dirname(os.path.abspath(__file__))
# doubles per run in a packed planet list: DH_MAX_PLANETS x DH_PLANET_STRIDE
DH_PLANET_WORDS = 8 * 8


class HostRun:
    def __init__(self, config: str, par_path: str):
        _hip_lib()
        path = os.path.join(HERE, "host", f"libdisco_host_{config}.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: build with `python -m disco_amd.build`")
        L = C.CDLL(path)
        vp = C.c_void_p
        L.disco_new.restype = vp
        L.disco_delete.argtypes = [vp]
        L.disco_read_par.argtypes = [vp, C.c_char_p]
        L.disco_setup.argtypes = [vp]
        L.disco_real_cells.argtypes = [vp]
        L.disco_real_cells.restype = C.c_long
        L.disco_ntot.argtypes = [vp]
        L.disco_ntot.restype = C.c_long
        L.disco_device_context.argtypes = [vp]
        L.disco_device_context.restype = vp
        L.getmindt.argtypes = [vp]
        L.getmindt.restype = C.c_double
        L.timestep.argtypes = [vp, C.c_double]
        L.disco_write_dstate.argtypes = [vp, C.c_char_p]
        L.disco_last_error.restype = C.c_char_p
        L.disco_step_begin.argtypes = [vp, C.c_double]
        L.disco_step_end.argtypes = [vp, C.c_double]
        L.disco_pack_planets.argtypes = [vp, C.POINTER(C.c_double)]
        L.disco_diag_kind.argtypes = [vp]
        L.disco_add_diag_time.argtypes = [vp, C.c_double]
        L.disco_npl.argtypes = [vp]
        L.disco_get_diagnostics.argtypes = [vp, C.c_void_p]
        self.L = L
        self.d = L.disco_new()
        if L.disco_read_par(self.d, par_path.encode()) != 0:
            raise RuntimeError("disco_read_par failed")
        if L.disco_setup(self.d) != 0:
            raise RuntimeError(f"disco_setup failed: {L.disco_last_error().decode()}")

    @property
    def real_cells(self) -> int:
        return self.L.disco_real_cells(self.d)

    @property
    def ntot(self) -> int:
        return self.L.disco_ntot(self.d)

    @property
    def ctx(self):
        return self.L.disco_device_context(self.d)

    def getmindt(self) -> float:
        return self.L.getmindt(self.d)

    def timestep(self, dt: float) -> None:
        self.L.timestep(self.d, dt)

    # timestep's host parts, for stepping several runs together (Ensemble)
    def step_begin(self, dt: float) -> None:
        if self.L.disco_step_begin(self.d, dt) != 0:
            raise RuntimeError(f"disco_step_begin: {self.L.disco_last_error().decode()}")

    def step_end(self, dt: float) -> None:
        self.L.disco_step_end(self.d, dt)

    def pack_planets(self, out) -> None:
        self.L.disco_pack_planets(self.d, out)

    @property
    def npl(self) -> int:
        return self.L.disco_npl(self.d)

    @property
    def diag_kind(self) -> int:
        return self.L.disco_diag_kind(self.d)

    def add_diag_time(self, dt: float) -> None:
        self.L.disco_add_diag_time(self.d, dt)

    def diagnostics(self, n: int):
        """The Analysis module's accumulated Qrz (n = Nr * Nz * its quantity count)."""
        import numpy as np
        out = np.zeros(n)
        if self.L.disco_get_diagnostics(self.d, out.ctypes.data_as(C.c_void_p)) != 0:
            raise RuntimeError(f"disco_get_diagnostics: {self.L.disco_last_error().decode()}")
        return out

    def write(self, path: str) -> None:
        if self.L.disco_write_dstate(self.d, path.encode()) != 0:
            raise RuntimeError("disco_write_dstate failed")

    def close(self):
        if self.d:
            self.L.disco_delete(self.d)
            self.d = None


class Ensemble:
    """Host runs of one module selection stepped together on the device
    (dh_ens_*): every kernel of their timestep runs as one launch for all
    runs, each run with its own CFL dt, planets and diagnostics."""

    def __init__(self, runs, graphs: bool = True):
        self.lib = _hip_lib()
        self.runs = list(runs)
        n = len(self.runs)
        ctxs = (C.c_void_p * n)(*[r.ctx for r in self.runs])
        self.ens = C.c_void_p()
        if self.lib.dh_ens_create(ctxs, n, C.byref(self.ens)):
            raise RuntimeError(f"dh_ens_create: {self.lib.dh_last_error().decode()}")
        self._check(self.lib.dh_ens_enable_graphs(self.ens, 1 if graphs else 0))
        self.dts = (C.c_double * n)()
        self.npl = max(r.npl for r in self.runs)
        self.stride = DH_PLANET_WORDS
        self.pl = (C.c_double * (n * self.stride))()
        kinds = {r.diag_kind for r in self.runs}
        if len(kinds) != 1:
            raise RuntimeError("ensemble runs use different Analysis modules")
        self.kind = kinds.pop()

    def _check(self, rc):
        if rc:
            raise RuntimeError(self.lib.dh_last_error().decode())

    def step(self):
        """getmindt and timestep (with the diagnostics) of every run."""
        self._check(self.lib.dh_ens_getmindt(self.ens, self.dts))
        for r, dt in zip(self.runs, self.dts):
            r.step_begin(dt)
        self._check(self.lib.dh_ens_timestep(self.ens, self.dts))
        base = C.addressof(self.pl)
        for m, (r, dt) in enumerate(zip(self.runs, self.dts)):
            r.step_end(dt)
            r.pack_planets(C.cast(base + 8 * m * self.stride, C.POINTER(C.c_double)))
        self._check(self.lib.dh_ens_add_diagnostics(self.ens, self.kind, self.dts, self.pl,
                                                    self.npl))
        for r, dt in zip(self.runs, self.dts):
            r.add_diag_time(dt)

    def sync(self):
        self._check(self.lib.dh_ens_sync(self.ens))

    def close(self):
        if self.ens:
            self.lib.dh_ens_destroy(self.ens)
            self.ens = C.c_void_p()
