"""The Kerr metric in Kerr-Schild cylindrical coordinates (kerr_ks.h, the
kerr_ks_rpz Metric plugin and DH_METRIC_KERR_KS) pinned by what the
mathematics fixes, and the Fishbone-Moncrief torus built on it:

* at a = 0 every function equals the reference's schw_ks_rpz.c
  (Hydro/Metric/schw_ks_rpz.c:20-189) to round-off;
* gamma_ij gamma^jk = delta_i^k; det g_{mu nu} = -R^2 (sqrt(-g) = R, the
  plugin's metric_jacobian); det gamma_ij = R^2 (1 + H) in closed form; l is
  null;
* d g / dR, d g / dz, the lapse's and the shift's derivatives equal central
  differences of the functions themselves;
* the torus: h = 1 at its inner edge and the pressure's maximum on the
  equator at r_max; its initial face fluxes are divergence-free.
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from disco_amd import configs, state_io

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(REPO, "disco_amd", "csrc")
REF = "/root/reference"

DRIVER = r"""
#include <stdio.h>
#include <math.h>
#include "kerr_ks.h"
#ifdef WITH_SCHW
double schw_lapse(const double x[3]); void schw_shift(const double x[3], double b[3]);
void schw_gam(const double x[3], double g[9]); void schw_igam(const double x[3], double g[9]);
void schw_der_g(const double x[3], int i, double dg[16]);
void schw_der_lapse(const double x[3], double da[4]);
void schw_der_shift(const double x[3], double db[12]);
#endif
static double pts[][2] = {{3.0, 0.5}, {10.0, -4.0}, {2.2, 0.01}, {25.0, 15.0}, {6.0, 0.0},
                          {40.0, -20.0}};
#define NP (int)(sizeof pts / sizeof pts[0])
static void all(double M, double a, double R, double z, double *out)
{
   /* lapse, shift[3], gam[9], igam[9], der_lapse[4], der_shift[12], der_g R, z [32] */
   kks_point p = kks_eval(M, a, R, z, 1);
   int n = 0;
   out[n++] = kks_lapse(&p);
   kks_shift(&p, out + n); n += 3;
   kks_gam(&p, out + n); n += 9;
   kks_igam(&p, out + n); n += 9;
   kks_der_lapse(&p, out + n); n += 4;
   kks_der_shift(&p, out + n); n += 12;
   kks_der_g(&p, 1, out + n); n += 16;
   kks_der_g(&p, 3, out + n); n += 16;
}
int main(void)
{
   double M = 1.0;
#ifdef WITH_SCHW
   for (int n = 0; n < NP; ++n) {
      double x[3] = {pts[n][0], 0.3, pts[n][1]}, k[70], s[70];
      all(M, 0.0, x[0], x[2], k);
      int m = 0;
      s[m++] = schw_lapse(x);
      schw_shift(x, s + m); m += 3;
      schw_gam(x, s + m); m += 9;
      schw_igam(x, s + m); m += 9;
      schw_der_lapse(x, s + m); m += 4;
      schw_der_shift(x, s + m); m += 12;
      schw_der_g(x, 1, s + m); m += 16;
      schw_der_g(x, 3, s + m); m += 16;
      double e = 0.0, sc = 0.0;
      for (int i = 0; i < m; ++i) { e = fmax(e, fabs(k[i] - s[i])); sc = fmax(sc, fabs(s[i])); }
      printf("schw %d %.17g\n", n, e / sc);
   }
#endif
   double a = 0.9, h = 1e-5;
   for (int n = 0; n < NP; ++n) {
      double R = pts[n][0], z = pts[n][1];
      kks_point p = kks_eval(M, a, R, z, 1);
      double g[9], ig[9], e = 0.0;
      kks_gam(&p, g);
      kks_igam(&p, ig);
      for (int i = 0; i < 3; ++i)
         for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += g[3 * i + k] * ig[3 * k + j];
            e = fmax(e, fabs(s - (i == j)));
         }
      double det3 = g[0] * (g[4] * g[8] - g[5] * g[7]) - g[1] * (g[3] * g[8] - g[5] * g[6]) +
                    g[2] * (g[3] * g[7] - g[4] * g[6]);
      double al = kks_lapse(&p);
      double null = -1.0 + p.l[1] * p.l[1] + p.l[2] * p.l[2] / (R * R) + p.l[3] * p.l[3];
      printf("ident %d %.17g %.17g %.17g %.17g\n", n, e, (-al * al * det3) / (-R * R) - 1.0,
             det3 / (R * R * (1.0 + p.H)) - 1.0, null);
      /* central differences of every function, relative to its scale */
      double v[70], vp[70], vm[70], fd = 0.0, sc = 0.0;
      all(M, a, R, z, v);
      for (int d = 0; d < 2; ++d) {
         double dR = d == 0 ? h * R : 0.0, dz = d == 1 ? h * R : 0.0;
         all(M, a, R + dR, z + dz, vp);
         all(M, a, R - dR, z - dz, vm);
         double step = 2.0 * h * R;
         /* lapse -> der_lapse[1 or 3]; shift[j] -> der_shift[3 i + j]; gam -> der_g */
         int i4 = d == 0 ? 1 : 3;
         double c = (vp[0] - vm[0]) / step;
         fd = fmax(fd, fabs(c - v[22 + i4])); sc = fmax(sc, fabs(c));
         for (int j = 0; j < 3; ++j) {
            c = (vp[1 + j] - vm[1 + j]) / step;
            fd = fmax(fd, fabs(c - v[26 + 3 * i4 + j])); sc = fmax(sc, fabs(c));
         }
         int base = d == 0 ? 38 : 54;
         for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
               c = (vp[4 + 3 * i + j] - vm[4 + 3 * i + j]) / step;
               fd = fmax(fd, fabs(c - v[base + 4 * (i + 1) + (j + 1)]));
               sc = fmax(sc, fabs(c));
            }
         /* g_tt = -1 + H, g_ti = H l_i */
         kks_point pp = kks_eval(M, a, R + dR, z + dz, 0), pm = kks_eval(M, a, R - dR, z - dz, 0);
         c = (pp.H - pm.H) / step;
         fd = fmax(fd, fabs(c - v[base])); sc = fmax(sc, fabs(c));
         for (int i = 1; i < 4; ++i) {
            c = (pp.H * pp.l[i] - pm.H * pm.l[i]) / step;
            fd = fmax(fd, fabs(c - v[base + i])); sc = fmax(sc, fabs(c));
         }
      }
      printf("fd %d %.17g\n", n, fd / sc);
   }
   return 0;
}
"""

RENAME = ["metric_lapse", "metric_shift", "metric_gam", "metric_igam", "metric_der_g",
          "metric_der_lapse", "metric_der_shift", "metric_jacobian", "metric_killing",
          "setMetricParams"]


def _run_driver():
    d = tempfile.mkdtemp(prefix="kerr_")
    src = os.path.join(d, "drv.c")
    with open(src, "w") as f:
        f.write(DRIVER)
    objs, defs = [], []
    schw = os.path.join(REF, "Hydro", "Metric", "schw_ks_rpz.c")
    if os.path.exists(schw):
        ren = [f"-D{n}=schw_{n.replace('metric_', '')}" for n in RENAME]
        obj = os.path.join(d, "schw.o")
        subprocess.run(["gcc", "-O2", "-w", "-DNUM_C=8", "-DNUM_N=0", "-DCT_MODE=2", *ren, "-c",
                        schw, "-o", obj], check=True)
        objs.append(obj)
        defs.append("-DWITH_SCHW")
    exe = os.path.join(d, "drv")
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", *defs, f"-I{CSRC}", src, *objs, "-lm",
                    "-o", exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout
    rows = {}
    for ln in out.splitlines():
        k, n, *v = ln.split()
        rows.setdefault(k, []).append([float(x) for x in v])
    return rows


@pytest.fixture(scope="module")
def rows():
    return _run_driver()


def test_zero_spin_is_the_reference_schwarzschild_metric(rows):
    if "schw" not in rows:
        pytest.skip("reference sources not mounted")
    worst = max(r[0] for r in rows["schw"])
    print(f"a = 0 against schw_ks_rpz.c: worst relative difference {worst:.2e}")
    assert worst <= 1e-15


def test_kerr_identities(rows):
    for inv, detg, det3, null in rows["ident"]:
        assert abs(inv) <= 1e-14 and abs(detg) <= 1e-13 and abs(det3) <= 1e-13, (inv, detg, det3)
        assert abs(null) <= 1e-14


def test_kerr_derivatives_match_central_differences(rows):
    worst = max(r[0] for r in rows["fd"])
    print(f"derivatives against central differences: worst relative {worst:.2e}")
    assert worst <= 1e-7


def _initial_state(over):
    cfg = configs.get("kerr_torus")
    exe = configs.ref_binary(cfg)
    if not os.path.exists(exe):
        pytest.skip(f"{exe} not built")
    par = cfg.par_dict(**over)
    d = tempfile.mkdtemp(prefix="fmtorus_")
    state_io.write_par(d, par)
    subprocess.run([exe, d, "0"], cwd=d, check=True, capture_output=True)
    return cfg, par, state_io.read_state(os.path.join(d, "state_0000.bin"))


def test_fm_torus_edge_and_pressure_maximum():
    """On the equator the torus starts at r_in = 6 and its pressure peaks at
    r_max = 12 (Boyer-Lindquist r; R = sqrt(r^2 + a^2) in the cylindrical
    coordinates at z = 0); without noise the density there is 1."""
    cfg, par, s = _initial_state({"Num_R": 380, "Num_Z": 2, "Z_Min": -0.05, "Z_Max": 0.05,
                                  "Noise_Type": 0, "aspect": 40.0})
    from helpers import cell_r
    m = s.real_mask()
    r = cell_r(s)[m]
    rho, P = s.prim[0, m], s.prim[1, m]
    a = 0.9
    rbl = np.sqrt(r * r - a * a)
    torus = s.prim[3, m] != 0.0   # orbiting gas; the atmosphere has u_i = 0
    print(f"torus from r = {rbl[torus].min():.3f}; pressure maximum at r = "
          f"{rbl[np.argmax(P)]:.3f}, density there {rho.max():.4f}")
    dr = (par["R_Max"] - par["R_Min"]) / par["Num_R"]
    # h = 1 at r_in: the gas is there from r_in on, where its density (which
    # grows as (h - 1)^3 for gamma = 4/3) passes the atmosphere's
    assert 6.0 <= rbl[torus].min() <= 6.5
    assert abs(rbl[np.argmax(P)] - 12.0) <= 2 * dr
    assert 0.99 < rho.max() <= 1.0


def test_fm_torus_initial_field_is_divergence_free():
    """The face fluxes the reference's set_B_fields takes from the torus's
    initial() (line integrals of A_phi) cancel over every cell's faces."""
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpu_engine"))
    import build as cpu_build
    from disco_amd import capi
    from helpers import div_b
    cfg, par, s = _initial_state({"Num_R": 48, "Num_Z": 24})
    g = capi.DiscoHip(cfg, par, s, library=capi.load_library(cpu_build.build()))
    faces = {d: g.faces(d)[:3] for d in (1, 2)}
    g.close()
    div = div_b(s, faces)[s.real_mask()]
    bmax = np.abs(s.phi_flux).max()
    print(f"largest face flux {bmax:.3e}; div B (relative) {div.max():.2e}")
    assert bmax > 0.0 and div.max() <= 1e-12
