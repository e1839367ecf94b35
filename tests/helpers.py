"""Shared helpers of the test suite."""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(HERE, "golden"))

from disco_amd import configs  # noqa: E402
from disco_amd.state_io import MeshState  # noqa: E402
import make_golden  # noqa: E402

GOLDEN = os.path.join(HERE, "golden")


def _state(z, prefix) -> MeshState:
    nq, nf, nr, nz = (int(v) for v in z["dims"])
    return MeshState(nq=nq, nf=nf, nr=nr, nz=nz, ng=tuple(int(v) for v in z["ng"]), t=0.0,
                     phi_max=float(z["phi_max"][0]), np_=z["np"].astype(np.int32),
                     r_jph=z["r_jph"], z_kph=z["z_kph"],
                     prim=z[f"{prefix}_prim"].copy(), cons=z[f"{prefix}_cons"].copy(),
                     piph=z[f"{prefix}_piph"].copy(), dphi=z[f"{prefix}_dphi"].copy(),
                     wiph=z[f"{prefix}_wiph"].copy(),
                     phi_flux=z[f"{prefix}_phi_flux"].copy(), planets=z["planets"].copy())


def load_golden(name):
    """(config, in.par dict, initial state, final state, dt list) of a fixture."""
    cname, over, steps = make_golden.CASES[name]
    cfg = configs.get(cname)
    par = cfg.par_dict(**over)
    z = np.load(os.path.join(GOLDEN, f"{name}.npz"))
    a, b = _state(z, "init"), _state(z, "final")
    dts = [float(v) for v in np.atleast_1d(z["dt"])]
    assert len(dts) == steps
    return cfg, par, a, b, dts


GOLDEN_CASES = list(make_golden.CASES)


def rel_l1(x: np.ndarray, y: np.ndarray) -> float:
    """Relative L1 difference sum|x-y| / sum|y| (0 when both are 0)."""
    den = np.abs(y).sum()
    num = np.abs(x - y).sum()
    return float(num / den) if den > 0 else float(num)


# Parity bar of the HIP path against the reference: relative L1 of every
# primitive (and conserved) field over real cells after the run.
PARITY_TOL = 1e-10


def cell_r(s: MeshState) -> np.ndarray:
    """Radial centroid of every cell (cylindrical.c:24-30)."""
    j = np.repeat(np.arange(len(s.np_)) % s.nr, s.np_)
    rp, rm = s.r_jph[j + 1], s.r_jph[j]
    return 2.0 * (rp * rp + rp * rm + rm * rm) / (3.0 * (rp + rm))


def field_errors(s: MeshState, ref: MeshState):
    """Relative L1 error over real cells of every primitive and conserved
    field, and of piph.  A component of a vector (velocity, momentum, B) is
    measured against the L1 norm of the whole vector, so that a component
    that vanishes in the solution (v_r of a steady vortex, B_r of a
    vertical field) is judged on the scale of the field it belongs to."""
    m = ref.real_mask()
    r = cell_r(ref)[m]
    out = {}

    def put(key, err, scale):
        den = float(np.abs(scale).sum())
        num = float(np.abs(err).sum())
        out[key] = num / den if den > 0 else num

    P, Ps = ref.prim[:, m], s.prim[:, m]
    C, Cs = ref.cons[:, m], s.cons[:, m]
    nc = ref.num_c
    for q in (0, 1):
        put(f"prim{q}", Ps[q] - P[q], P[q])
        put(f"cons{q}", Cs[q] - C[q], C[q])
    # velocity (v_r, r omega, v_z) and momentum (S_r, L / r, S_z)
    vmag = np.abs(P[2]) + np.abs(r * P[3]) + np.abs(P[4])
    put("prim2", Ps[2] - P[2], vmag)
    put("prim3", r * (Ps[3] - P[3]), vmag)
    put("prim4", Ps[4] - P[4], vmag)
    smag = np.abs(C[2]) + np.abs(C[3] / r) + np.abs(C[4])
    put("cons2", Cs[2] - C[2], smag)
    put("cons3", (Cs[3] - C[3]) / r, smag)
    put("cons4", Cs[4] - C[4], smag)
    if nc >= 8:
        bmag = np.abs(P[5]) + np.abs(P[6]) + np.abs(P[7])
        for q in (5, 6, 7):
            put(f"prim{q}", Ps[q] - P[q], bmag)
        # conserved B: (B_r dV / r, B_phi dV / r, B_z dV)
        w = np.array([r, r, np.ones_like(r)])
        cmag = sum(np.abs(C[q] * w[q - 5]) for q in (5, 6, 7))
        for q in (5, 6, 7):
            put(f"cons{q}", (Cs[q] - C[q]) * w[q - 5], cmag)
    for q in range(nc, ref.nq):
        put(f"prim{q}", Ps[q] - P[q], P[q])
        put(f"cons{q}", Cs[q] - C[q], C[q])
    put("piph", s.piph - ref.piph, ref.piph)
    return out


def compare_states(s: MeshState, ref: MeshState, tol: float = PARITY_TOL):
    """Asserts every field within tol; returns the worst relative L1."""
    errs = field_errors(s, ref)
    worst_key = max(errs, key=errs.get)
    worst = errs[worst_key]
    assert worst <= tol, f"relative L1 {worst:.3e} ({worst_key}) > {tol:.1e}: {errs}"
    return worst


def face_flux(phi_flux: np.ndarray, dim: int, L: np.ndarray, R: np.ndarray,
              LR: np.ndarray) -> np.ndarray:
    """Magnetic flux through every r (dim 1) or z (dim 2) face: the Phi slot
    of the face's owner (LRtype 0: L's outer slot, 1: R's inner slot)."""
    if dim == 1:
        return np.where(LR == 0, phi_flux[2][L], phi_flux[1][R])
    return np.where(LR == 0, phi_flux[4][L], phi_flux[3][R])


def div_b(s: MeshState, faces) -> np.ndarray:
    """Relative divergence of B per cell: the net outward flux through the
    cell's faces over the sum of their absolute fluxes.  faces = {dim: (L, R,
    LRtype)} of the state's current mesh."""
    nt = s.ntot
    net = np.zeros(nt)
    mag = np.zeros(nt)
    off = np.concatenate([[0], np.cumsum(s.np_)])
    cm = np.empty(nt, dtype=np.int64)
    for a in range(len(s.np_)):
        i = np.arange(s.np_[a])
        cm[off[a] + i] = off[a] + (i - 1) % s.np_[a]
    p0 = s.phi_flux[0]
    net += p0 - p0[cm]
    mag += np.abs(p0) + np.abs(p0[cm])
    for dim, (L, R, LR) in faces.items():
        f = face_flux(s.phi_flux, dim, L, R, LR)
        af = np.abs(f)
        net += np.bincount(L, weights=f, minlength=nt) - np.bincount(R, weights=f, minlength=nt)
        mag += np.bincount(L, weights=af, minlength=nt) + np.bincount(R, weights=af, minlength=nt)
    return np.abs(net) / np.where(mag > 0, mag, 1.0)


REPORT_COLUMNS = ("t", "Torque", "Power", "Fr", "rho_min", "rhoavg_min", "PsiR", "PsiI", "Mass",
                  "Mdot", "S_R", "L1_rho", "L1_isen", "L1_B", "Br2", "aM", "bM")


def read_report(path: str) -> np.ndarray:
    """report.dat rows (report.c:188-191): t and 16 columns each."""
    with open(path) as f:
        rows = [[float(x) for x in line.split()] for line in f if line.strip()]
    return np.array(rows).reshape(-1, len(REPORT_COLUMNS))


def assert_reports_match(got, want, rtol=2e-6):
    """Two report rows, equal to the 7 significant digits report.c prints
    (%le), with a floor of 1e-12 of the disk's mass for columns that are
    sums of cancelling terms (PsiR, PsiI of a nearly axisymmetric disk)."""
    floor = 1e-12 * max(abs(want[8]), 1.0)
    for name, a, b in zip(REPORT_COLUMNS, got, want):
        if not np.isfinite(b):
            assert a == b or (np.isnan(a) and np.isnan(b)), (name, a, b)
            continue
        assert abs(a - b) <= rtol * max(abs(a), abs(b)) + floor, (name, a, b)
