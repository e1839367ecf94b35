"""ctypes handle on a host library libdisco_host_<config>.so: the drop-in's
in.par reader, mesh set-up and the reference-signature timestep/getmindt."""
from __future__ import annotations

import ctypes as C
import os

from .capi import lib as _hip_lib  # loads libdisco_hip.so first (the host library links it)

HERE = os.path.dirname(os.path.abspath(__file__))


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

    def write(self, path: str) -> None:
        if self.L.disco_write_dstate(self.d, path.encode()) != 0:
            raise RuntimeError("disco_write_dstate failed")

    def close(self):
        if self.d:
            self.L.disco_delete(self.d)
            self.d = None
