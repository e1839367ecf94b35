"""Module selections and in.par presets of the supported workloads.

A Disco configuration is a link-time module selection (the reference's
Makefile_opt.in: HYDRO, BOUNDARY, INITIAL, PLANETS, NUM_C, NUM_N, CT_MODE)
plus an in.par file.  `Config` holds both; `dh_bc` says how each side of the
mesh is treated on the device (see DH_BC_* in include/disco_hip.h).
"""
from __future__ import annotations

import dataclasses
from typing import Dict, Optional, Tuple

# DH_BC_* (include/disco_hip.h)
BC_NONE, BC_HELD, BC_ZEROGRAD, BC_DIODE = 0, 1, 2, 3
HYDRO_EULER, HYDRO_MHD, HYDRO_GRMHD = 0, 1, 2
# DH_METRIC_* by Metric module name
METRICS = {"flat_rpz": 0, "schw_ks_eq_rpz": 1, "schw_ks_rpz": 2, "kerr_ks_rpz": 3}


@dataclasses.dataclass(frozen=True)
class Config:
    name: str
    hydro: str                 # "euler" | "mhd" | "grmhd2"
    initial: str               # plugin name (disco_amd/plugins/Initial or reference Initial/)
    boundary: str              # "polar" | "fixed" | "outflow3d" | "fixed_r_outflow_z"
    planets: str               # "none" | "sun" | ...
    num_c: int
    num_n: int
    ct_mode: int
    par: Tuple[Tuple[str, object], ...]
    analysis: str = "diag_euler"  # Diagnostics module (ANALYSIS)
    metric: str = ""           # Metric module (METRIC); "" is the hydro module's default

    @property
    def metric_module(self) -> str:
        if self.metric:
            return self.metric
        return "schw_ks_eq_rpz" if self.hydro == "grmhd2" else "flat_rpz"

    def par_dict(self, **overrides) -> Dict[str, object]:
        d = dict(self.par)
        d.update(overrides)
        return d

    @property
    def nq(self) -> int:
        return self.num_c + self.num_n

    @property
    def nf(self) -> int:
        return {0: 0, 1: 3, 2: 5}[self.ct_mode]

    def dh_bc(self, par: Dict[str, object]) -> Tuple[int, int, int, int]:
        """Per-side device boundary kind (rmin, rmax, zmin, zmax)."""
        nobc = [par.get("NoBC_Rmin", 0), par.get("NoBC_Rmax", 0),
                par.get("NoBC_Zmin", 0), par.get("NoBC_Zmax", 0)]
        if self.boundary in ("polar", "fixed"):
            sides = [BC_HELD] * 4
        elif self.boundary == "outflow3d":
            sides = [BC_DIODE] * 4
        elif self.boundary == "zerograd":
            sides = [BC_ZEROGRAD] * 4
        elif self.boundary == "fixed_r_outflow_z":
            sides = [BC_HELD, BC_HELD, BC_DIODE, BC_DIODE]
        else:
            raise ValueError(f"no device boundary for module {self.boundary!r}")
        if int(par.get("Num_Z", 1)) == 1:
            nobc[2] = nobc[3] = 1
        return tuple(BC_NONE if nobc[i] else sides[i] for i in range(4))


def _cfg(name, hydro, initial, boundary, planets, num_c, num_n, ct_mode, analysis="diag_euler",
         **par):
    return Config(name, hydro, initial, boundary, planets, num_c, num_n, ct_mode,
                  tuple(par.items()), analysis)


# DH_DIAG_* (include/disco_hip.h) and the number of quantities of each module
DIAG_KINDS = {"diag_euler": (1, 8), "diag_disk2d": (2, 5), "diag_mhdmri": (3, 18)}


CONFIGS: Dict[str, Config] = {}

# 2-D Newtonian Euler isentropic vortex: r in [0,1] with the polar axis,
# Nr=128 (51,809 cells), gamma=5/3, eps=5, PLM 1.5, HLLC, RK2, CFL 0.5,
# rigid mesh rotation (Mesh_Motion 4, Exact_Mesh_Omega 1).
CONFIGS["euler_yee"] = _cfg(
    "euler_yee", "euler", "yee_vortex", "polar", "none", 5, 0, 0,
    Num_R=128, Num_Z=1, R_Min=0.0, R_Max=1.0, Z_Min=-0.5, Z_Max=0.5,
    NoBC_Rmin=1, CFL=0.5, PLM=1.5, Riemann_Solver=1, Mesh_Motion=4,
    Exact_Mesh_Omega=1, Adiabatic_Index=5.0 / 3.0, Init_Par1=5.0)

CONFIGS["euler_yee_n1"] = dataclasses.replace(
    CONFIGS["euler_yee"], name="euler_yee_n1", num_n=1, analysis="diag_disk2d")

# The reference's own Gaussian vortex (Initial/vortex.c), HLLC, polar axis.
CONFIGS["euler_vortex"] = _cfg(
    "euler_vortex", "euler", "vortex", "polar", "none", 5, 0, 0,
    Num_R=32, Num_Z=1, R_Min=0.0, R_Max=3.0, Z_Min=-0.5, Z_Max=0.5,
    NoBC_Rmin=1, CFL=0.5, PLM=1.5, Riemann_Solver=1, Mesh_Motion=1,
    Adiabatic_Index=5.0 / 3.0)


# 3-D Newtonian MHD Keplerian disk (MRI seed): r in [0.5,2], z in [-0.25,0.25],
# Nr=256, Nz=64 (22,004,544 cells), rho=1, cs = v_K/20 (isothermal), vertical
# field at beta=1000, velocity noise 1e-3 cs (Noise_Type 3), HLLD, PLM 1.5,
# constrained transport, Keplerian mesh motion and energy frame.
CONFIGS["mhd_kep"] = _cfg(
    "mhd_kep", "mhd", "kep_mhd_disk", "fixed_r_outflow_z", "central", 8, 0, 2,
    analysis="diag_mhdmri",
    Num_R=256, Num_Z=64, R_Min=0.5, R_Max=2.0, Z_Min=-0.25, Z_Max=0.25,
    NoBC_Rmin=0, NoBC_Rmax=0, NoBC_Zmin=0, NoBC_Zmax=0, CFL=0.5, PLM=1.5,
    Riemann_Solver=2, Mesh_Motion=4, Exact_Mesh_Omega=2, Energy_Omega=2,
    Constrained_Transport=1, Isothermal=1, Cs2_Profile=4, Mach_Number=20.0,
    Metric_Par2=1.0, Grav2D=1, Adiabatic_Index=5.0 / 3.0, Init_Par1=1000.0,
    Noise_Type=3, Noise_Rel=5e-5, Noise_Abs=0.0)


# GRMHD: a magnetised gas disk on the equator of a Schwarzschild black hole
# (M = 1) in Kerr-Schild coordinates, r in [4, 12], z in [-0.5, 0.5], HLLC,
# PLM 1.5, gamma = 4/3, constrained transport, Keplerian mesh motion, the
# Noble et al. cons2prim (Hydro/grmhd2.c, Metric/schw_ks_eq_rpz.c).
CONFIGS["gr_disk"] = _cfg(
    "gr_disk", "grmhd2", "gr_disk", "fixed_r_outflow_z", "none", 8, 0, 2,
    Num_R=64, Num_Z=16, R_Min=4.0, R_Max=12.0, Z_Min=-0.5, Z_Max=0.5,
    NoBC_Rmin=0, NoBC_Rmax=0, NoBC_Zmin=0, NoBC_Zmax=0, CFL=0.5, PLM=1.5,
    Riemann_Solver=1, Mesh_Motion=4, Exact_Mesh_Omega=2, Energy_Omega=0,
    Constrained_Transport=1, Isothermal=0, Metric_Par2=1.0, Adiabatic_Index=4.0 / 3.0,
    Init_Par1=100.0, Density_Floor=1e-8, Pressure_Floor=1e-10)

# The GRMHD disk in the full (r, z) Schwarzschild metric (schw_ks_rpz): off
# the equator gravity has a vertical component, so the slab is not in
# equilibrium and falls towards z = 0.
CONFIGS["gr_disk3d"] = dataclasses.replace(CONFIGS["gr_disk"], name="gr_disk3d",
                                           metric="schw_ks_rpz")

# GRMHD: the Fishbone-Moncrief torus (inner edge r = 6, pressure maximum at
# r = 12, gamma = 4/3) around a Kerr black hole (M = 1, a = 0.9) in
# Kerr-Schild cylindrical coordinates (kerr_ks_rpz), poloidal field loops on
# the density contours at beta = 100 with divergence-free face fluxes plus a
# weak uniform vertical field (B^z = 1e-4: the reference's relativistic HLLD
# fails and leaves NaN where the field vanishes identically), 1% pressure
# noise, an atmosphere at rest in the normal frame (rho 1e-2, P 1e-5) with
# floors below it; r in [2, 40], z in [-20, 20], Nr = 192, Nz = 96, aspect
# 2.5 (at most 512 cells per annulus, 4.9M cells), relativistic HLLD, PLM
# 1.5, constrained transport, the mesh moving with the gas (Mesh_Motion 1),
# CFL 0.05.  With a thinner atmosphere (rho 1e-3 or less) the torus's edge
# cells empty to the floors within a few steps, where the reference's own
# run branches on round-off (and with rho 1e-4 fails).  The atmosphere
# between the hole and the torus falls into the hole and empties: cells
# there, at the density floor and falling at u_r of order -20, are left
# with a negative conserved density by the update (grmhd2's cons2prim has
# no positivity fix) and turn non-finite, at step 50 with CFL 0.2, step 60
# with CFL 0.1 and step 130 with CFL 0.05 (measured on the GPU); CFL 0.05
# keeps the 60 steps that the learned guess's online harvest (50 steps) and
# its timed comparison (10 more) take finite, with under 0.6% of the
# cons2prim solves unconverged.
CONFIGS["kerr_torus"] = _cfg(
    "kerr_torus", "grmhd2", "fm_torus", "fixed_r_outflow_z", "none", 8, 0, 2,
    Num_R=192, Num_Z=96, R_Min=2.0, R_Max=40.0, Z_Min=-20.0, Z_Max=20.0, aspect=2.5,
    NoBC_Rmin=0, NoBC_Rmax=0, NoBC_Zmin=0, NoBC_Zmax=0, CFL=0.05, PLM=1.5,
    Riemann_Solver=2, Mesh_Motion=1, Energy_Omega=0, Constrained_Transport=1, Isothermal=0,
    Metric_Par2=1.0, Metric_Par3=0.9, Adiabatic_Index=4.0 / 3.0, Init_Par1=6.0,
    Init_Par2=12.0, Init_Par3=100.0, Init_Par4=0.2, Init_Par5=1e-2, Init_Par6=1e-5,
    Init_Par7=1e-4, Density_Floor=1e-6, Pressure_Floor=1e-8, Noise_Type=2, Noise_Rel=1e-2,
    Noise_Abs=0.0)
CONFIGS["kerr_torus"] = dataclasses.replace(CONFIGS["kerr_torus"], metric="kerr_ks_rpz")

# One member of the disk-planet ensemble: a 2-D isothermal disk (Sigma =
# r^-0.5, h = 0.05) with alpha viscosity 1e-3 and a planet of mass ratio
# Mass_Ratio on a circular orbit at r = 1 softened by 0.6 h; HLLC, PLM 1.5,
# Keplerian mesh motion, held radial boundaries.
CONFIGS["disk_planet"] = _cfg(
    "disk_planet", "euler", "planet_disk", "fixed", "planet_circ", 5, 0, 0,
    Num_R=128, Num_Z=1, R_Min=0.4, R_Max=2.5, NoBC_Rmin=0, NoBC_Rmax=0, CFL=0.5, PLM=1.5,
    Riemann_Solver=1, Mesh_Motion=4, Exact_Mesh_Omega=2, Energy_Omega=0, Isothermal=1,
    Cs2_Profile=4, Mach_Number=20.0, Metric_Par2=1.0, Grav2D=1, Adiabatic_Index=5.0 / 3.0,
    Use_Viscosity=1, Use_As_Alpha=1, Viscosity=1e-3, Mass_Ratio=1e-3)

# The disk with a companion of mass ratio 1e-3 at r = 1.25 (Planets/central.c
# reads Mass_Ratio) and the Euler diagnostics, which read planet 1.
CONFIGS["mhd_kep_p2"] = dataclasses.replace(
    CONFIGS["mhd_kep"], name="mhd_kep_p2", analysis="diag_euler",
    par=CONFIGS["mhd_kep"].par + (("Mass_Ratio", 1e-3),))


# The memory-filling MHD disk's in.par on top of mhd_kep: r in [0.5, 3],
# z in [-0.3, 0.3], a zero-net-flux toroidal field at beta = 100
# (kep_mhd_disk with Init_Par0 = 1); the full size is Nr = 512, Nz = 160.
C5_PAR = {"R_Min": 0.5, "R_Max": 3.0, "Z_Min": -0.3, "Z_Max": 0.3, "Init_Par0": 1,
          "Init_Par1": 100.0}


def get(name: str) -> Config:
    return CONFIGS[name]


def ref_binary(cfg: Config, repo: Optional[str] = None) -> str:
    import os
    repo = repo or os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    return os.path.join(repo, "oracle", "_ref", f"disco_ref_{cfg.name}")
