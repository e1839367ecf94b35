"""Raw state files ("DSTATE01") and in.par writers.

A DSTATE01 file holds one full mesh state in structure-of-arrays order:

    char    magic[8] = "DSTATE01"
    int32   hdr[16]  = NQ, NF, Nr, Nz, NgRa, NgRb, NgZa, NgZb, Ntot, Npl, step, NUM_C, 0...
    float64 dhdr[8]  = t, phi_max, 0...
    int32   Np[Nr*Nz]
    float64 r_jph[Nr+1], z_kph[Nz+1]          (edge -1 first, as gridsetup.c:94-95)
    float64 prim[NQ][Ntot], cons[NQ][Ntot]
    float64 piph[Ntot], dphi[Ntot], wiph[Ntot]
    float64 Phi[NF][Ntot]
    float64 planets[Npl][8]  (r, phi, M, omega, vr, eps, type, 0)

Cells are ordered annulus by annulus (jk = j + Nr*k, the reference's
theCells[jk] index, paul.h:141) and then by i = 0..Np[jk]-1.  The in.par
writer produces the key/value file that readpar.c:44-188 of the reference
parses; every key is written, since a missing key is left at 0 (or, for
T_Times_2pi / P_Times_2pi, uninitialised) by the reference.
"""
from __future__ import annotations

import dataclasses
import math
import os
from typing import Dict, Optional

import numpy as np

MAGIC = b"DSTATE01"


@dataclasses.dataclass
class MeshState:
    nq: int
    nf: int
    nr: int
    nz: int
    ng: tuple  # (NgRa, NgRb, NgZa, NgZb)
    t: float
    phi_max: float
    np_: np.ndarray  # int32 [Nr*Nz]
    r_jph: np.ndarray  # [Nr+1], r_jph[0] is edge -1
    z_kph: np.ndarray  # [Nz+1]
    prim: np.ndarray  # [NQ, Ntot]
    cons: np.ndarray  # [NQ, Ntot]
    piph: np.ndarray
    dphi: np.ndarray
    wiph: np.ndarray
    phi_flux: np.ndarray  # [NF, Ntot]
    planets: np.ndarray  # [Npl, 8]
    step: int = 0
    num_c: int = 5

    @property
    def ntot(self) -> int:
        return int(self.piph.shape[0])

    @property
    def offsets(self) -> np.ndarray:
        off = np.zeros(self.nr * self.nz + 1, dtype=np.int64)
        off[1:] = np.cumsum(self.np_)
        return off

    def real_mask(self) -> np.ndarray:
        """True for real (non-ghost) cells."""
        nga, ngb, nza, nzb = self.ng
        m = np.zeros(self.ntot, dtype=bool)
        off = self.offsets
        for k in range(nza, self.nz - nzb):
            for j in range(nga, self.nr - ngb):
                jk = j + self.nr * k
                m[off[jk]:off[jk + 1]] = True
        return m

    def n_real(self) -> int:
        return int(self.real_mask().sum())

    def copy(self) -> "MeshState":
        return dataclasses.replace(
            self,
            np_=self.np_.copy(), r_jph=self.r_jph.copy(), z_kph=self.z_kph.copy(),
            prim=self.prim.copy(), cons=self.cons.copy(), piph=self.piph.copy(),
            dphi=self.dphi.copy(), wiph=self.wiph.copy(),
            phi_flux=self.phi_flux.copy(), planets=self.planets.copy())


def read_state(path: str) -> MeshState:
    with open(path, "rb") as f:
        buf = f.read()
    if buf[:8] != MAGIC:
        raise ValueError(f"{path}: not a DSTATE01 file")
    o = 8
    hdr = np.frombuffer(buf, dtype=np.int32, count=16, offset=o); o += 64
    dhdr = np.frombuffer(buf, dtype=np.float64, count=8, offset=o); o += 64
    nq, nf, nr, nz, ngra, ngrb, ngza, ngzb, ntot, npl, step, num_c = (int(v) for v in hdr[:12])

    def take(dtype, n):
        nonlocal o
        a = np.frombuffer(buf, dtype=dtype, count=n, offset=o).copy()
        o += a.nbytes
        return a

    np_ = take(np.int32, nr * nz)
    r_jph = take(np.float64, nr + 1)
    z_kph = take(np.float64, nz + 1)
    prim = take(np.float64, nq * ntot).reshape(nq, ntot)
    cons = take(np.float64, nq * ntot).reshape(nq, ntot)
    piph = take(np.float64, ntot)
    dphi = take(np.float64, ntot)
    wiph = take(np.float64, ntot)
    phi_flux = take(np.float64, nf * ntot).reshape(nf, ntot)
    planets = take(np.float64, npl * 8).reshape(npl, 8)
    if o != len(buf):
        raise ValueError(f"{path}: {len(buf) - o} trailing bytes")
    return MeshState(nq=nq, nf=nf, nr=nr, nz=nz, ng=(ngra, ngrb, ngza, ngzb),
                     t=float(dhdr[0]), phi_max=float(dhdr[1]), np_=np_,
                     r_jph=r_jph, z_kph=z_kph, prim=prim, cons=cons, piph=piph,
                     dphi=dphi, wiph=wiph, phi_flux=phi_flux, planets=planets,
                     step=step, num_c=num_c)


def read_diag(path: str):
    """(t_avg, Qrz[Nr*Nz, nq]) of a diag.bin: int32 nq, int32 Nr*Nz, double
    t_avg, double Qrz (the Analysis module's sums, analysis.c:52-106)."""
    with open(path, "rb") as f:
        nq, nann = np.frombuffer(f.read(8), dtype=np.int32)
        t_avg = float(np.frombuffer(f.read(8), dtype=np.float64)[0])
        q = np.frombuffer(f.read(8 * int(nq) * int(nann)), dtype=np.float64)
    return t_avg, q.reshape(int(nann), int(nq)).copy()


def write_state(path: str, s: MeshState) -> None:
    hdr = np.zeros(16, dtype=np.int32)
    hdr[:12] = [s.nq, s.nf, s.nr, s.nz, *s.ng, s.ntot, s.planets.shape[0], s.step, s.num_c]
    dhdr = np.zeros(8, dtype=np.float64)
    dhdr[0], dhdr[1] = s.t, s.phi_max
    with open(path, "wb") as f:
        f.write(MAGIC)
        for a in (hdr, dhdr, s.np_.astype(np.int32), s.r_jph, s.z_kph, s.prim,
                  s.cons, s.piph, s.dphi, s.wiph, s.phi_flux, s.planets):
            f.write(np.ascontiguousarray(a).tobytes())


# ---------------------------------------------------------------------------
# in.par presets
# ---------------------------------------------------------------------------

_DEFAULTS: Dict[str, object] = {
    "Restart": 0, "T_Start": 0.0, "T_End": 1.0e30, "T_Times_2pi": 0,
    "Num_Reports": 0, "Num_Snapshots": 0, "Num_Checkpoints": 0, "Use_Logtime": 0,
    "Num_R": 16, "Num_Z": 1, "aspect": 1.0, "Max_Aspect_Short": 1000.0,
    "Max_Aspect_Long": 1000.0,
    "R_Min": 0.0, "R_Max": 1.0, "Z_Min": -0.5, "Z_Max": 0.5, "R_Periodic": 0,
    "Z_Periodic": 0, "Phi_Max": 1.0, "P_Times_2pi": 1, "Z_Times_pi": 0,
    "Log_Zoning": 0, "Log_Radius": 0.1,
    "NoBC_Rmin": 1, "NoBC_Rmax": 0, "NoBC_Zmin": 0, "NoBC_Zmax": 0,
    "CFL": 0.5, "PLM": 1.5, "Max_DT": 0.0, "Timestep": 0, "Riemann_Solver": 1,
    "Mesh_Motion": 4, "Exact_Mesh_Omega": 1, "Exact_Mesh_Omega_Par": 0.0,
    "Absorbing_BC": 0, "Initial_Regrid": 0, "Density_Floor": 0.0,
    "Pressure_Floor": 0.0, "Constrained_Transport": 0,
    "Adiabatic_Index": 5.0 / 3.0, "Isothermal": 0, "Cs2_Profile": 0, "Cs2_Par": 0.0,
    "Energy_Omega": 0, "Energy_Omega_Par": 0.0, "Use_Viscosity": 0, "Viscosity": 0.0,
    "Use_As_Alpha": 0, "Mass_Ratio": 0.0, "Eccentricity": 0.0, "Drift_Rate": 0.0,
    "Drift_Exp": 0.66666667, "Grav2D": 0, "RotFrame": 0, "RotOmega": 0.0, "RotD": 0.0,
    "Mach_Number": 20.0, "Include_Atmos": 0,
    "Init_Par0": 0, "Init_Par1": 0.0, "Init_Par2": 0.0, "Init_Par3": 0.0,
    "Init_Par4": 0.0, "Init_Par5": 0.0, "Init_Par6": 0.0, "Init_Par7": 0.0,
    "Init_Par8": 0.0, "Noise_Type": 0, "Noise_Abs": 0.0, "Noise_Rel": 0.0,
    "Sink_Type": 0, "Sink_Par1": 0.0, "Sink_Par2": 0.0, "Sink_Par3": 0.0,
    "Sink_Par4": 0.0, "Nozzle_Type": 0, "Nozzle_Par1": 0.0, "Nozzle_Par2": 0.0,
    "Nozzle_Par3": 0.0, "Nozzle_Par4": 0.0,
    "Metric_Par0": 0, "Metric_Par1": 0.0, "Metric_Par2": 1.0, "Metric_Par3": 0.0,
    "Metric_Par4": 0.0,
}

# The 2-D Euler isentropic-vortex workload: r in [0,1] with the polar axis,
# Nr=128, gamma=5/3, eps=5, PLM 1.5, HLLC, RK2, CFL 0.5, rigid mesh rotation.
PRESETS: Dict[str, Dict[str, object]] = {
    "euler_yee": {
        "Num_R": 128, "Num_Z": 1, "R_Min": 0.0, "R_Max": 1.0, "Z_Min": -0.5,
        "Z_Max": 0.5, "NoBC_Rmin": 1, "CFL": 0.5, "PLM": 1.5, "Riemann_Solver": 1,
        "Mesh_Motion": 4, "Exact_Mesh_Omega": 1, "Adiabatic_Index": 5.0 / 3.0,
        "Init_Par1": 5.0,
    },
}


# Keys of this path with no reference counterpart (grmhd2 cons2prim's guess,
# stopping rule and model file, the recorded timestep, the launch tuner's
# steps); written only when given.
EXTRA_KEYS = ("C2P_Guess", "C2P_Stop", "C2P_Model", "Graphs", "Autotune")


def format_par(overrides: Optional[Dict[str, object]] = None) -> str:
    vals = dict(_DEFAULTS)
    if overrides:
        for k, v in overrides.items():
            if k not in vals and k not in EXTRA_KEYS:
                raise KeyError(f"unknown in.par key {k!r}")
            vals[k] = v
    lines = []
    for k, v in vals.items():
        if isinstance(v, float):
            lines.append(f"{k}\t{v!r}")
        else:
            lines.append(f"{k}\t{v}")
    return "\n".join(lines) + "\n"


def write_par(directory: str, overrides: Optional[Dict[str, object]] = None) -> str:
    os.makedirs(directory, exist_ok=True)
    path = os.path.join(directory, "in.par")
    with open(path, "w") as f:
        f.write(format_par(overrides))
    return path


def preset(name: str, **kw) -> Dict[str, object]:
    d = dict(PRESETS[name])
    d.update(kw)
    return d


def real_cells_formula(num_r: int, rmin: float, rmax: float, num_z: int,
                       aspect: float = 1.0, phimax: float = 2 * math.pi) -> int:
    """Real-cell count from the reference Np rule (gridsetup.c:120-140)."""
    n = 0
    for j in range(num_r):
        rp = rmin + (j + 1) / num_r * (rmax - rmin)
        rm = rmin + j / num_r * (rmax - rmin)
        dp = aspect * (rp - rm) / rp
        n += max(4, int(phimax / dp))
    return n * num_z
