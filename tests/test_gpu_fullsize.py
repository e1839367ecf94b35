"""The 3-D MHD disk at its benchmark size on the GPU (Nr=256, Nz=64,
22,004,544 real cells; too large for the oracle): 20 CFL steps through the
drop-in host layer, then properties that do not depend on a reference run.

* every primitive, conserved and flux field is finite;
* div B: the net magnetic flux out of every real cell, relative to the sum
  of its faces' absolute fluxes, stays at round-off (<= 1e-12).  The
  reference's face hand-over (bfields.c:693-814) keeps it there while the
  Keplerian shear moves a cell edge by less than one cell per step; at this
  resolution and CFL 0.5 it moves the innermost annuli's edges by about a
  fifth of a cell per step, and the test measures the largest shift;
* the Keplerian disk stays near equilibrium: density within 1% of its
  initial value in every real cell.
"""
import ctypes as C
import os
import tempfile

import numpy as np
import pytest

from disco_amd import capi, configs, state_io
from helpers import div_b

pytestmark = pytest.mark.gpu


def _faces(lib, ctx, dim):
    n = C.c_long()
    assert lib.dh_face_count(ctx, dim, C.byref(n)) == 0
    L = np.zeros(n.value, np.int32)
    R = np.zeros(n.value, np.int32)
    LR = np.zeros(n.value, np.int32)
    ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))  # noqa: E731
    assert lib.dh_get_faces(ctx, dim, ip(L), ip(R), ip(LR), None, None) == 0, \
        lib.dh_last_error().decode()
    return L, R, LR


def _state(run):
    d = "/dev/shm" if os.path.isdir("/dev/shm") else None
    fd, path = tempfile.mkstemp(prefix="fullsize_", suffix=".dstate", dir=d)
    os.close(fd)
    try:
        run.write(path)
        return state_io.read_state(path)
    finally:
        os.remove(path)


def test_mhd_disk_benchmark_size_keeps_div_b_and_equilibrium():
    from disco_amd.host_lib import HostRun
    cfg = configs.get("mhd_kep")
    par = cfg.par_dict()
    steps = 20
    d = tempfile.mkdtemp(prefix="fullsize_")
    run = HostRun("mhd_kep", state_io.write_par(d, par))
    print(f"case mhd_kep benchmark size: {run.real_cells} cells, {steps} CFL steps", flush=True)
    lib = capi.lib()
    ctx = C.c_void_p(run.ctx)
    try:
        a = _state(run)
        dts = []
        for _ in range(steps):
            dt = run.getmindt()
            dts.append(dt)
            run.timestep(dt)
        b = _state(run)
        faces = {dim: _faces(lib, ctx, dim) for dim in (1, 2)}
    finally:
        run.close()
    m = b.real_mask()
    for name in ("prim", "cons", "piph", "dphi", "phi_flux"):
        assert np.isfinite(getattr(b, name)).all(), name
    # the largest shift of a face position relative to the next annulus
    # over one step: (Omega(r_j) - Omega(r_j+1)) dt against the cell's dphi
    r = b.r_jph
    rc = 0.5 * (r[1:] + r[:-1])
    om = rc ** -1.5
    shear = np.abs(np.diff(om)).max() * max(dts)
    cell = (2 * np.pi / b.np_[:b.nr]).min()
    print(f"  dt {min(dts):.3e}..{max(dts):.3e}; largest shear per step {shear / cell:.3f} "
          f"of a cell", flush=True)
    assert shear < cell
    div = div_b(b, faces)[m]
    print(f"  div B (relative, max over real cells) {div.max():.2e}", flush=True)
    assert div.max() <= 1e-12
    drho = np.abs(b.prim[0, m] / a.prim[0, m] - 1.0).max()
    print(f"  max |rho / rho_0 - 1| {drho:.2e}", flush=True)
    assert drho < 1e-2
