/*
 * disco_oracle.c -- CPU restatement of the reference Disco timestep.
 *
 * TEST INFRASTRUCTURE ONLY (see disco_oracle.h).  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 *
 * Every routine below restates one routine of the reference, cited as
 * file:line of HangChelseaSu/Disco, on the structure-of-arrays state of
 * include/disco_hip.h.  Loops keep the reference's order (including the
 * sequential scatter of face fluxes into both neighbours) and expressions keep
 * its association, so that the result equals the reference build bit for bit
 * when both are compiled with the same gcc flags (oracle/Makefile and
 * oracle/ref/Makefile both use -O3 -ffp-contract=off).
 *
 * Pinning: tests/test_oracle_ref.py compares whole runs with oracle/_ref
 * (bit-exact) and with the committed fixtures under tests/golden/; the
 * closed-form checks of tests/test_oracle_math.py cover prim<->cons
 * identity, the Riemann flux of equal states, face tiling and the slope
 * limiter's closed form.
 */
#include "disco_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define MAXQ 16
enum { RHO, PPP, URR, UPP, UZZ, BRR, BPP, BZZ };
enum { DDD, TAU, SRR, LLL, SZZ };
enum { R_HLL, R_HLLC, R_HLLD };

typedef struct {
   long L, R;
   double dxL, dxR;
   double cm[3];
   double dphi, dl, dA;
   double E, B;
   int LRtype;
   int flip_flag;
} or_face;

struct or_ctx {
   dh_params p;
   int nq, nf;
   int Nr, Nz, NgRa, NgRb, NgZa, NgZb;
   double phi_max;
   int *Np;
   long *off;
   double *rbuf, *zbuf; /* r_jph/z_kph with edge -1 at index 0 */
   double *r_jph, *z_kph; /* rbuf+1, zbuf+1: index -1 valid (gridsetup.c:94-95) */
   long ntot;
   double *prim, *cons, *rkcons, *gradr, *gradp, *gradz;
   double *piph, *dphi, *wiph;
   double *Phi, *rkPhi;
   double *Ecell, *Bcell, *Ephi, *temp; /* E[8], B[8], E_phi[4], tempDoub */
   or_face *faces[2];
   long *fI[2];
   int nftracks[2];
   double planets[2][DH_MAX_PLANETS * DH_PLANET_STRIDE];
   int npl[2];
   int cur_substep;
   int polar_sources;
};

#define PRIM(o, q, c) ((o)->prim[(long)(q) * (o)->ntot + (c)])
#define CONS(o, q, c) ((o)->cons[(long)(q) * (o)->ntot + (c)])

/* ------------------------------------------------------------------------ */
/* Geometry/cylindrical.c                                                    */
/* ------------------------------------------------------------------------ */

/* get_dp, cylindrical.c:10-15 */
static double g_dp(const or_ctx *o, double phip, double phim)
{
   double dp = phip - phim;
   while (dp < 0.0) dp += o->phi_max;
   while (dp > o->phi_max) dp -= o->phi_max;
   return dp;
}

/* get_signed_dp, cylindrical.c:17-22 */
static double g_sdp(const or_ctx *o, double phip, double phim)
{
   double dp = phip - phim;
   while (dp < -.5 * o->phi_max) dp += o->phi_max;
   while (dp > .5 * o->phi_max) dp -= o->phi_max;
   return dp;
}

/* get_centroid, cylindrical.c:24-30 */
static double g_centroid(double xp, double xm, int dim)
{
   if (dim == 1) return 2.0 * (xp * xp + xp * xm + xm * xm) / (3.0 * (xp + xm));
   return 0.5 * (xp + xm);
}

/* get_dL, cylindrical.c:32-41 */
static double g_dL(const or_ctx *o, const double *xp, const double *xm, int dim)
{
   double r = .5 * (xp[0] + xm[0]);
   double dphi = g_dp(o, xp[1], xm[1]);
   if (dim == 0) return r * dphi;
   if (dim == 1) return xp[0] - xm[0];
   return xp[2] - xm[2];
}

/* get_dA, cylindrical.c:43-54 */
static double g_dA(const or_ctx *o, const double *xp, const double *xm, int dim)
{
   double r = .5 * (xp[0] + xm[0]);
   double dr = xp[0] - xm[0];
   double dphi = g_dp(o, xp[1], xm[1]);
   double dz = xp[2] - xm[2];
   if (dim == 0) return dr * dz;
   if (dim == 1) return r * dphi * dz;
   return r * dr * dphi;
}

/* get_dV, cylindrical.c:56-64 */
static double g_dV(const or_ctx *o, const double *xp, const double *xm)
{
   double r = .5 * (xp[0] + xm[0]);
   double dr = xp[0] - xm[0];
   double dphi = g_dp(o, xp[1], xm[1]);
   double dz = xp[2] - xm[2];
   return r * dr * dphi * dz;
}

/* geom_polar_vec_adjust, cylindrical.c:175-183 (first component) */
static double g_polar_adjust(const or_ctx *o, const double *xp, const double *xm)
{
   double dphi = g_dp(o, xp[1], xm[1]);
   return sin(0.5 * dphi) / (0.5 * dphi);
}

/* ------------------------------------------------------------------------ */
/* omega.c                                                                   */
/* ------------------------------------------------------------------------ */

/* mesh_om, omega.c:43-60 */
static double mesh_om(const or_ctx *o, const double *x)
{
   double r = x[0];
   int ch = o->p.exact_mesh_omega;
   if (ch == 1) return 1.0;
   if (ch == 2) return pow(r, -1.5);
   if (ch == 3) {
      double n = 8.0;
      return 1. / pow(pow(r, 1.5 * n) + 1., 1. / n);
   }
   return 0.0;
}

/* get_om, omega.c:62-85 */
static double get_om(const or_ctx *o, const double *x)
{
   double r = x[0];
   double O0 = o->p.rot_omega;
   switch (o->p.energy_omega) {
   case 1: return 1.0 - O0;
   case 2: return 1.0 / pow(r, 1.5) - O0;
   case 3: {
      double n = 8.0;
      return 1. / pow(pow(r, 1.5 * n) + 1., 1. / n) - O0;
   }
   case 4: return 10. * exp(-.5 * r * r) - O0;
   default: return 0.0 - O0;
   }
}

/* get_om1, omega.c:87-110 */
static double get_om1(const or_ctx *o, const double *x)
{
   double r = x[0];
   switch (o->p.energy_omega) {
   case 1: return 0.0;
   case 2: return -1.5 / pow(r, 2.5);
   case 3: {
      double n = 8.0;
      return -1.5 * pow(r, -1 + 1.5 * n) / pow(pow(r, 1.5 * n) + 1., 1.0 + 1. / n);
   }
   case 4: return 10. * r * exp(-.5 * r * r);
   default: return 0.0;
   }
}

/* get_om2, omega.c:112-114 */
static double get_om2(const or_ctx *o, const double *x)
{
   (void)o; (void)x;
   return 0.0;
}

/* get_cs2, omega.c:116-140 (M = Metric_Par2, r1/H0 = Init_Par2/Init_Par4) */
static double get_cs2(const or_ctx *o, const double *x)
{
   double r = x[0];
   double Mach = o->p.mach;
   switch (o->p.cs2_profile) {
   case 1: return 1. / (Mach * Mach);
   case 2: {
      double nu = .5;
      return .5 / Mach / Mach / pow(r, 2. * nu);
   }
   case 3: {
      double r1 = o->p.init_par[2], H0 = o->p.init_par[4];
      return o->p.metric_par2 * H0 * H0 / (2 * r1 * r1 * r1);
   }
   case 4: {
      double v2 = o->p.metric_par2 / r;
      return v2 / (Mach * Mach);
   }
   default: return 1.0;
   }
}

/* ------------------------------------------------------------------------ */
/* Hydro/euler.c                                                             */
/* ------------------------------------------------------------------------ */

/* prim2cons, euler.c:36-62 */
static void e_prim2cons(const or_ctx *o, const double *prim, double *cons,
                        const double *x, double dV)
{
   double r = x[0];
   double rho = prim[RHO];
   double Pp = prim[PPP];
   double vr = prim[URR];
   double vp = prim[UPP] * r;
   double vz = prim[UZZ];
   double om = get_om(o, x);
   double vp_off = vp - om * r;
   double v2 = vr * vr + vp_off * vp_off + vz * vz;
   double rhoe = Pp / (o->p.gamma_law - 1.);
   cons[DDD] = rho * dV;
   cons[TAU] = (.5 * rho * v2 + rhoe) * dV;
   cons[SRR] = rho * vr * dV;
   cons[LLL] = r * rho * vp * dV;
   cons[SZZ] = rho * vz * dV;
   for (int q = o->p.num_c; q < o->nq; ++q) cons[q] = prim[q] * cons[DDD];
}

/* getUstar, euler.c:64-100 */
static void e_getUstar(const or_ctx *o, const double *prim, double *Ustar,
                       const double *x, double Sk, double Ss, const double *n)
{
   double r = x[0];
   double rho = prim[RHO];
   double vr = prim[URR];
   double vp = prim[UPP] * r;
   double vz = prim[UZZ];
   double Pp = prim[PPP];
   double om = get_om(o, x);
   double vp_off = vp - om * r;
   double v2 = vr * vr + vp_off * vp_off + vz * vz;
   double vn = vr * n[0] + vp * n[1] + vz * n[2];
   double vn_off = vn - om * r * n[1];
   double Ss_off = Ss + vn_off - vn;
   double rhoe = Pp / (o->p.gamma_law - 1.);
   double rhostar = rho * (Sk - vn) / (Sk - Ss);
   double Pstar = Pp * (Ss - vn) / (Sk - Ss);
   double Us = rhoe * (Sk - vn) / (Sk - Ss);
   Ustar[DDD] = rhostar;
   Ustar[SRR] = rhostar * (vr + (Ss - vn) * n[0]);
   Ustar[LLL] = r * rhostar * (vp + (Ss - vn) * n[1]);
   Ustar[SZZ] = rhostar * (vz + (Ss - vn) * n[2]);
   Ustar[TAU] = .5 * rhostar * v2 + Us + rhostar * Ss_off * (Ss - vn) + Pstar;
   for (int q = o->p.num_c; q < o->nq; ++q) Ustar[q] = prim[q] * Ustar[DDD];
}

/* cons2prim, euler.c:102-139 */
static void e_cons2prim(const or_ctx *o, const double *cons, double *prim,
                        const double *x, double dV)
{
   double r = x[0];
   double rho = cons[DDD] / dV;
   if (rho < o->p.rho_floor) rho = o->p.rho_floor;
   double Sr = cons[SRR] / dV;
   double Sp = cons[LLL] / dV / r;
   double Sz = cons[SZZ] / dV;
   double E = cons[TAU] / dV;
   double om = get_om(o, x);
   double vr = Sr / rho;
   double vp = Sp / rho;
   double vp_off = vp - om * r;
   double vz = Sz / rho;
   double KE = .5 * (Sr * vr + rho * vp_off * vp_off + Sz * vz);
   double rhoe = E - KE;
   double Pp = (o->p.gamma_law - 1.) * rhoe;
   if (Pp < o->p.p_floor * rho) Pp = o->p.p_floor * rho;
   if (o->p.isothermal) {
      double cs2 = get_cs2(o, x);
      Pp = cs2 * rho / o->p.gamma_law;
   }
   prim[RHO] = rho;
   prim[PPP] = Pp;
   prim[URR] = vr;
   prim[UPP] = vp / r;
   prim[UZZ] = vz;
   for (int q = o->p.num_c; q < o->nq; ++q) prim[q] = cons[q] / cons[DDD];
}

/* flux, euler.c:141-169 */
static void e_flux(const or_ctx *o, const double *prim, double *flux,
                   const double *x, const double *n)
{
   double r = x[0];
   double rho = prim[RHO];
   double Pp = prim[PPP];
   double vr = prim[URR];
   double vp = prim[UPP] * r;
   double vz = prim[UZZ];
   double om = get_om(o, x);
   double vn = vr * n[0] + vp * n[1] + vz * n[2];
   double wn = om * r * n[1];
   double vp_off = vp - om * r;
   double rhoe = Pp / (o->p.gamma_law - 1.);
   double v2 = vr * vr + vp_off * vp_off + vz * vz;
   flux[DDD] = rho * vn;
   flux[SRR] = rho * vr * vn + Pp * n[0];
   flux[LLL] = r * (rho * vp * vn + Pp * n[1]);
   flux[SZZ] = rho * vz * vn + Pp * n[2];
   flux[TAU] = (.5 * rho * v2 + rhoe + Pp) * vn - Pp * wn;
   for (int q = o->p.num_c; q < o->nq; ++q) flux[q] = prim[q] * flux[DDD];
}

/* source, euler.c:171-215 */
static void e_source(const or_ctx *o, const double *prim, double *cons,
                     const double *xp, const double *xm, double dVdt)
{
   double rp = xp[0];
   double rm = xm[0];
   double rho = prim[RHO];
   double Pp = prim[PPP];
   double r = g_centroid(rp, rm, 1);
   double z = g_centroid(xp[2], xm[2], 2);
   double r_1 = .5 * (rp + rm);
   double vr = prim[URR];
   double omega = prim[UPP];
   double vz = prim[UZZ];
   double x[3] = {r, 0.5 * (xm[1] + xp[1]), z};
   double centrifugal = rho * omega * omega * r;
   if (o->polar_sources) centrifugal *= g_polar_adjust(o, xp, xm);
   double press_bal = Pp / r_1;
   cons[SRR] += dVdt * (centrifugal + press_bal);
   double om = get_om(o, x);
   double om_r = get_om1(o, x);
   double om_z = get_om2(o, x);
   cons[TAU] += dVdt * rho * (r * vr * om * om - r * r * (omega - om) * (vr * om_r + vz * om_z));
}

/* alpha/constant kinematic viscosity, euler.c:219-227 */
static double e_nu(const or_ctx *o, const double *prim, double r)
{
   double nu = o->p.viscosity;
   if (o->p.alpha_flag) {
      double alpha = o->p.viscosity;
      double c = sqrt(o->p.gamma_law * prim[PPP] / prim[RHO]);
      double h = c * pow(r, 1.5);
      nu = alpha * c * h;
   }
   return nu;
}

/* visc_flux, euler.c:217-259 */
static void e_visc_flux(const or_ctx *o, const double *prim, const double *gradr,
                        const double *gradp, const double *gradz, double *flux,
                        const double *x, const double *n)
{
   double r = x[0];
   double nu = e_nu(o, prim, r);
   double rho = prim[RHO];
   double vr = prim[URR];
   double om = prim[UPP];
   double om_off = om - get_om(o, x);
   double vz = prim[UZZ];
   double divV_o_d = (gradr[URR] + gradp[UPP] + gradz[UZZ] + vr / r) / 3.0;
   double srr = gradr[URR] - divV_o_d;
   double spp = r * (r * gradp[UPP] + vr - r * divV_o_d);
   double szz = gradz[UZZ] - divV_o_d;
   double srp = 0.5 * (r * r * gradr[UPP] + gradp[URR]);
   double srz = 0.5 * (gradr[UZZ] + gradz[URR]);
   double spz = 0.5 * (gradp[UZZ] + r * r * gradz[UPP]);
   double nc[3] = {n[0], n[1] / r, n[2]};
   double srn = srr * nc[0] + srp * nc[1] + srz * nc[2];
   double spn = srp * nc[0] + spp * nc[1] + spz * nc[2];
   double szn = srz * nc[0] + spz * nc[1] + szz * nc[2];
   flux[SRR] = -2 * nu * rho * srn;
   flux[LLL] = -2 * nu * rho * spn;
   flux[SZZ] = -2 * nu * rho * szn;
   flux[TAU] = -2 * nu * rho * (vr * srn + om_off * spn + vz * szn);
}

/* visc_source, euler.c:261-295 */
static void e_visc_source(const or_ctx *o, const double *prim, const double *gradr,
                          const double *gradp, const double *gradz, double *cons,
                          const double *xp, const double *xm, double dVdt)
{
   double x[3] = {g_centroid(xp[0], xm[0], 1), g_centroid(xp[1], xm[1], 0),
                  g_centroid(xp[2], xm[2], 2)};
   double r = x[0];
   double nu = e_nu(o, prim, r);
   double rho = prim[RHO];
   double vr = prim[URR];
   double divV_o_d = (gradr[URR] + gradp[UPP] + gradz[UZZ] + vr / r) / 3.0;
   double spp = (r * gradp[UPP] + vr - r * divV_o_d) / (r * r * r);
   cons[SRR] += (-2 * r * rho * nu * spp) * dVdt;
   double srp = 0.5 * (r * r * gradr[UPP] + gradp[URR]);
   double spz = 0.5 * (gradp[UZZ] + r * r * gradz[UPP]);
   double om_r = get_om1(o, x);
   double om_z = get_om2(o, x);
   cons[TAU] += (2 * rho * nu * (srp * om_r + spz * om_z)) * dVdt;
}

/* vel, euler.c:297-330 */
static void e_vel(const or_ctx *o, const double *prim1, const double *prim2,
                  double *Sl, double *Sr, double *Ss, const double *n, const double *x)
{
   double r = x[0];
   double g = o->p.gamma_law;
   double P1 = prim1[PPP];
   double rho1 = prim1[RHO];
   double vn1 = prim1[URR] * n[0] + prim1[UPP] * n[1] * r + prim1[UZZ] * n[2];
   double cs1 = sqrt(g * P1 / rho1);
   double P2 = prim2[PPP];
   double rho2 = prim2[RHO];
   double vn2 = prim2[URR] * n[0] + prim2[UPP] * n[1] * r + prim2[UZZ] * n[2];
   double cs2 = sqrt(g * P2 / rho2);
   *Ss = (P2 - P1 + rho1 * vn1 * (-cs1) - rho2 * vn2 * cs2) / (rho1 * (-cs1) - rho2 * cs2);
   *Sr = cs1 + vn1;
   *Sl = -cs1 + vn1;
   if (*Sr < cs2 + vn2) *Sr = cs2 + vn2;
   if (*Sl > -cs2 + vn2) *Sl = -cs2 + vn2;
}

/* mindt, euler.c:332-380 */
static double e_mindt(const or_ctx *o, const double *prim, double w,
                      const double *xp, const double *xm)
{
   double r = g_centroid(xp[0], xm[0], 1);
   double Pp = prim[PPP];
   double rho = prim[RHO];
   double vp = (prim[UPP] - w) * r;
   double vr = prim[URR];
   double vz = prim[UZZ];
   double cs = sqrt(o->p.gamma_law * Pp / rho);
   double maxvr = cs + fabs(vr);
   double maxvp = cs + fabs(vp);
   double maxvz = cs + fabs(vz);
   double dtr = g_dL(o, xp, xm, 1) / maxvr;
   double dtp = g_dL(o, xp, xm, 0) / maxvp;
   double dtz = g_dL(o, xp, xm, 2) / maxvz;
   double dt = dtr;
   if (dt > dtp) dt = dtp;
   if (dt > dtz) dt = dtz;
   if (o->p.visc_flag) {
      double dL0 = g_dL(o, xp, xm, 0);
      double dL1 = g_dL(o, xp, xm, 1);
      double dL2 = g_dL(o, xp, xm, 2);
      double dx = dL0;
      if (dx > dL1) dx = dL1;
      if (dx > dL2) dx = dL2;
      double nu = e_nu(o, prim, r);
      double dt_visc = 0.5 * dx * dx / nu;
      if (dt > dt_visc) dt = dt_visc;
   }
   return dt;
}

/* ------------------------------------------------------------------------ */
/* Hydro/mhd.c (Newtonian MHD; B_r is conserved as B_r dV / r)               */
/* ------------------------------------------------------------------------ */

/* bfield_scale_factor, mhd.c:518-528 */
static double m_bscale(double x, int dim) { return dim == 0 ? 1.0 / x : 1.0; }

/* prim2cons, mhd.c:35-69 */
static void m_prim2cons(const or_ctx *o, const double *prim, double *cons,
                        const double *x, double dV)
{
   double r = x[0];
   double rho = prim[RHO];
   double Pp = prim[PPP];
   double vr = prim[URR];
   double vp = prim[UPP] * r;
   double vz = prim[UZZ];
   double om = get_om(o, x);
   double vp_off = vp - om * r;
   double Br = prim[BRR], Bp = prim[BPP], Bz = prim[BZZ];
   double v2 = vr * vr + vp_off * vp_off + vz * vz;
   double B2 = Br * Br + Bp * Bp + Bz * Bz;
   double rhoe = Pp / (o->p.gamma_law - 1.);
   cons[DDD] = rho * dV;
   cons[TAU] = (.5 * rho * v2 + rhoe + .5 * B2) * dV;
   cons[SRR] = rho * vr * dV;
   cons[LLL] = r * rho * vp * dV;
   cons[SZZ] = rho * vz * dV;
   cons[BRR] = Br * dV * m_bscale(r, 0);
   cons[BPP] = Bp * dV / r;
   cons[BZZ] = Bz * dV;
   for (int q = o->p.num_c; q < o->nq; ++q) cons[q] = prim[q] * cons[DDD];
}

/* cons2prim, mhd.c:138-185 */
static void m_cons2prim(const or_ctx *o, const double *cons, double *prim,
                        const double *x, double dV)
{
   double r = x[0];
   double rho = cons[DDD] / dV;
   if (rho < o->p.rho_floor) rho = o->p.rho_floor;
   double Sr = cons[SRR] / dV;
   double Sp = cons[LLL] / dV / r;
   double Sz = cons[SZZ] / dV;
   double E = cons[TAU] / dV;
   double om = get_om(o, x);
   double vr = Sr / rho;
   double vp = Sp / rho;
   double vp_off = vp - om * r;
   double vz = Sz / rho;
   double Br = cons[BRR] / (dV * m_bscale(r, 0));
   double Bp = cons[BPP] / dV * r;
   double Bz = cons[BZZ] / dV;
   double B2 = Br * Br + Bp * Bp + Bz * Bz;
   double KE = .5 * (Sr * vr + rho * vp_off * vp_off + Sz * vz);
   double rhoe = E - KE - .5 * B2;
   double Pp = (o->p.gamma_law - 1.) * rhoe;
   if (Pp < o->p.p_floor * rho) Pp = o->p.p_floor * rho;
   if (o->p.isothermal) {
      double cs2 = get_cs2(o, x);
      Pp = cs2 * rho / o->p.gamma_law;
   }
   prim[RHO] = rho;
   prim[PPP] = Pp;
   prim[URR] = vr;
   prim[UPP] = vp / r;
   prim[UZZ] = vz;
   prim[BRR] = Br;
   prim[BPP] = Bp;
   prim[BZZ] = Bz;
   for (int q = o->p.num_c; q < o->nq; ++q) prim[q] = cons[q] / cons[DDD];
}

/* source, mhd.c:225-271 (polar = NgRa == 0, mhd.c:21-22) */
static void m_source(const or_ctx *o, const double *prim, double *cons,
                     const double *xp, const double *xm, double dVdt)
{
   double rp = xp[0];
   double rm = xm[0];
   double dphi = g_dp(o, xp[1], xm[1]);
   double rho = prim[RHO];
   double Pp = prim[PPP];
   double r_1 = .5 * (rp + rm);
   double r2_3 = (rp * rp + rp * rm + rm * rm) / 3.;
   double vr = prim[URR];
   double omega = prim[UPP];
   double r = g_centroid(rp, rm, 1);
   double z = g_centroid(xp[2], xm[2], 2);
   double x[3] = {r, 0.5 * (xm[1] + xp[1]), z};
   double Br = prim[BRR], Bp = prim[BPP], Bz = prim[BZZ];
   double B2 = Br * Br + Bp * Bp + Bz * Bz;
   double centrifugal;
   if (o->NgRa == 0) {
      centrifugal = (rho * omega * omega * r2_3) / r_1 * sin(.5 * dphi) / (.5 * dphi);
      centrifugal -= Bp * Bp / r_1;
   } else {
      centrifugal = rho * omega * omega * r - Bp * Bp / r;
   }
   double press_bal = (Pp + .5 * B2) / r_1;
   cons[SRR] += dVdt * (centrifugal + press_bal);
   double om = get_om(o, x);
   double om1 = get_om1(o, x);
   cons[TAU] += dVdt * rho * vr * (om * om * r2_3 / r_1 - om1 * (omega - om) * r2_3);
}

/* mindt, mhd.c:452-483 */
static double m_mindt(const or_ctx *o, const double *prim, double w,
                      const double *xp, const double *xm)
{
   double r = g_centroid(xp[0], xm[0], 1);
   double Pp = prim[PPP];
   double rho = prim[RHO];
   double vp = (prim[UPP] - w) * r;
   double vr = prim[URR];
   double vz = prim[UZZ];
   double cs = sqrt(o->p.gamma_law * Pp / rho);
   double Br = prim[BRR], Bp = prim[BPP], Bz = prim[BZZ];
   double b2 = (Br * Br + Bp * Bp + Bz * Bz) / rho;
   double cf = sqrt(cs * cs + b2);
   double maxvr = cf + fabs(vr);
   double maxvp = cf + fabs(vp);
   double maxvz = cf + fabs(vz);
   double dtr = g_dL(o, xp, xm, 1) / maxvr;
   double dtp = g_dL(o, xp, xm, 0) / maxvp;
   double dtz = g_dL(o, xp, xm, 2) / maxvz;
   double dt = dtr;
   if (dt > dtp) dt = dtp;
   if (dt > dtz) dt = dtz;
   return dt;
}

/* flux_to_E, mhd.c:293-316: edge EMFs and normal fields from a face solve. */
static void m_flux_to_E(const double *Flux, const double *Ustr, const double *x,
                        double *E1, double *B1, double *E2, double *B2, int dim)
{
   double r = x[0];
   double irfac = 1.0 / m_bscale(x[0], 0);
   if (dim == 0) {
      *E1 = Flux[BRR] * irfac;
      *B1 = Ustr[BRR] * irfac;
      *E2 = -Flux[BZZ];
      *B2 = Ustr[BZZ];
   } else if (dim == 1) {
      *E1 = -Flux[BPP] * r;
      *B1 = Ustr[BRR] * irfac;
      *E2 = Flux[BZZ];
   } else {
      *E1 = Flux[BPP] * r;
      *B1 = Ustr[BZZ];
      *E2 = -Flux[BRR] * irfac;
   }
}

/* ------------------------------------------------------------------------ */
/* hlld.c: HLLD (Miyoshi & Kusano 2005) in the orthonormal frame             */
/* ------------------------------------------------------------------------ */

#define HLLD_TOL 1e-10

/* get_velocities, hlld.c:20-70 */
static void hd_velocities(double gam, const double *pL, const double *pR, const double *n,
                          double *vel, double *Pt_star)
{
   double rho_L = pL[RHO];
   double P_L = pL[PPP];
   double u_L = pL[URR] * n[0] + pL[UPP] * n[1] + pL[UZZ] * n[2];
   double Bn_L = pL[BRR] * n[0] + pL[BPP] * n[1] + pL[BZZ] * n[2];
   double B2_L = pL[BRR] * pL[BRR] + pL[BPP] * pL[BPP] + pL[BZZ] * pL[BZZ];
   double Pt_L = P_L + .5 * B2_L;
   double rho_R = pR[RHO];
   double P_R = pR[PPP];
   double u_R = pR[URR] * n[0] + pR[UPP] * n[1] + pR[UZZ] * n[2];
   double Bn_R = pR[BRR] * n[0] + pR[BPP] * n[1] + pR[BZZ] * n[2];
   double B2_R = pR[BRR] * pR[BRR] + pR[BPP] * pR[BPP] + pR[BZZ] * pR[BZZ];
   double Pt_R = P_R + .5 * B2_R;
   double Bn = .5 * (Bn_L + Bn_R);
   double cf2_L = .5 * (gam * P_L + B2_L + sqrt(fabs(pow(gam * P_L + B2_L, 2.) - 4. * gam * P_L * Bn_L * Bn_L))) / rho_L;
   double cf2_R = .5 * (gam * P_R + B2_R + sqrt(fabs(pow(gam * P_R + B2_R, 2.) - 4. * gam * P_R * Bn_R * Bn_R))) / rho_R;
   double cf2 = cf2_L > cf2_R ? cf2_L : cf2_R;
   double umin, umax;
   if (u_L < u_R) { umin = u_L; umax = u_R; } else { umin = u_R; umax = u_L; }
   double S_R = umax + sqrt(fabs(cf2));
   double S_L = umin - sqrt(fabs(cf2));
   double S_M = ((S_R - u_R) * rho_R * u_R - (S_L - u_L) * rho_L * u_L - Pt_R + Pt_L) /
                ((S_R - u_R) * rho_R - (S_L - u_L) * rho_L);
   double rhostar_L = rho_L * (S_L - u_L) / (S_L - S_M);
   double rhostar_R = rho_R * (S_R - u_R) / (S_R - S_M);
   double Ss_L = S_M - fabs(Bn) / sqrt(rhostar_L);
   double Ss_R = S_M + fabs(Bn) / sqrt(rhostar_R);
   vel[0] = S_L; vel[1] = Ss_L; vel[2] = S_M; vel[3] = Ss_R; vel[4] = S_R;
   double num = (S_R - u_R) * rho_R * Pt_L - (S_L - u_L) * rho_L * Pt_R +
                rho_L * rho_R * (S_R - u_R) * (S_L - u_L) * (u_R - u_L);
   double denom = (S_R - u_R) * rho_R - (S_L - u_L) * rho_L;
   *Pt_star = num / denom;
}

/* prim_to_cons, hlld.c:72-98 (lab frame, orthonormal components) */
static void hd_cons(double gam, const double *prim, double *cons)
{
   double rho = prim[RHO], Pp = prim[PPP];
   double vx = prim[URR], vy = prim[UPP], vz = prim[UZZ];
   double Bx = prim[BRR], By = prim[BPP], Bz = prim[BZZ];
   double v2 = vx * vx + vy * vy + vz * vz;
   double rhoe = Pp / (gam - 1.);
   double B2 = Bx * Bx + By * By + Bz * Bz;
   cons[DDD] = rho;
   cons[SRR] = rho * vx;
   cons[LLL] = rho * vy;
   cons[SZZ] = rho * vz;
   cons[TAU] = (.5 * rho * v2 + rhoe + .5 * B2);
   cons[BRR] = Bx;
   cons[BPP] = By;
   cons[BZZ] = Bz;
}

/* get_flux, hlld.c:100-133 */
static void hd_flux(double gam, const double *prim, double *flux, const double *n)
{
   double rho = prim[RHO], Pp = prim[PPP];
   double vx = prim[URR], vy = prim[UPP], vz = prim[UZZ];
   double Bx = prim[BRR], By = prim[BPP], Bz = prim[BZZ];
   double v2 = vx * vx + vy * vy + vz * vz;
   double vB = vx * Bx + vy * By + vz * Bz;
   double B2 = Bx * Bx + By * By + Bz * Bz;
   double vn = vx * n[0] + vy * n[1] + vz * n[2];
   double Bn = Bx * n[0] + By * n[1] + Bz * n[2];
   double rhoe = Pp / (gam - 1.);
   flux[DDD] = rho * vn;
   flux[SRR] = rho * vx * vn + (Pp + .5 * B2) * n[0] - Bx * Bn;
   flux[LLL] = rho * vy * vn + (Pp + .5 * B2) * n[1] - By * Bn;
   flux[SZZ] = rho * vz * vn + (Pp + .5 * B2) * n[2] - Bz * Bn;
   flux[TAU] = (.5 * rho * v2 + rhoe + Pp + B2) * vn - vB * Bn;
   flux[BRR] = Bx * vn - vx * Bn;
   flux[BPP] = By * vn - vy * Bn;
   flux[BZZ] = Bz * vn - vz * Bn;
}

/* get_single_star, hlld.c:135-193 */
static void hd_single_star(double gam, const double *prim, double *Ustar, double S_K,
                           double S_M, double Pt_star, const double *n)
{
   double rho = prim[RHO], vx = prim[URR], vy = prim[UPP], vz = prim[UZZ];
   double Pp = prim[PPP], Bx = prim[BRR], By = prim[BPP], Bz = prim[BZZ];
   double u_K = vx * n[0] + vy * n[1] + vz * n[2];
   double Bn = Bx * n[0] + By * n[1] + Bz * n[2];
   double vdotB = vx * Bx + vy * By + vz * Bz;
   double v2 = vx * vx + vy * vy + vz * vz;
   double B2 = Bx * Bx + By * By + Bz * Bz;
   double Pt = Pp + .5 * B2;
   double e = .5 * rho * v2 + Pp / (gam - 1.) + .5 * B2;
   double rhostar = rho * (S_K - u_K) / (S_K - S_M);
   double denom = rho * (S_K - u_K) * (S_K - S_M) - Bn * Bn;
   if (fabs(denom) < HLLD_TOL) denom = HLLD_TOL;
   double vx_star = vx - Bn * Bx * (S_M - u_K) / denom;
   double vy_star = vy - Bn * By * (S_M - u_K) / denom;
   double vz_star = vz - Bn * Bz * (S_M - u_K) / denom;
   double Bx_star = Bx * (rho * pow(S_K - u_K, 2.) - Bn * Bn) / denom;
   double By_star = By * (rho * pow(S_K - u_K, 2.) - Bn * Bn) / denom;
   double Bz_star = Bz * (rho * pow(S_K - u_K, 2.) - Bn * Bn) / denom;
   double vn_star = vx_star * n[0] + vy_star * n[1] + vz_star * n[2];
   double Bn_star = Bx_star * n[0] + By_star * n[1] + Bz_star * n[2];
   vx_star += (S_M - vn_star) * n[0];
   vy_star += (S_M - vn_star) * n[1];
   vz_star += (S_M - vn_star) * n[2];
   Bx_star += (Bn - Bn_star) * n[0];
   By_star += (Bn - Bn_star) * n[1];
   Bz_star += (Bn - Bn_star) * n[2];
   double vBstar = vx_star * Bx_star + vy_star * By_star + vz_star * Bz_star;
   double e_star = ((S_K - u_K) * e - Pt * u_K + Pt_star * S_M + Bn * (vdotB - vBstar)) / (S_K - S_M);
   Ustar[DDD] = rhostar;
   Ustar[SRR] = rhostar * vx_star;
   Ustar[LLL] = rhostar * vy_star;
   Ustar[SZZ] = rhostar * vz_star;
   Ustar[TAU] = e_star;
   Ustar[BRR] = Bx_star;
   Ustar[BPP] = By_star;
   Ustar[BZZ] = Bz_star;
}

/* get_double_star, hlld.c:195-280 */
static void hd_double_star(const double *UsL, const double *UsR, double *Uss,
                           const double *vel, const double *n, int LR)
{
   double rho_L = UsL[DDD], rho_R = UsR[DDD];
   double rhostar_K = LR == 0 ? rho_L : rho_R;
   double rrh_L = sqrt(rho_L), rrh_R = sqrt(rho_R), rrh_K = sqrt(rhostar_K);
   double Sx_L = UsL[SRR], Sy_L = UsL[LLL], Sz_L = UsL[SZZ];
   double Bx_L = UsL[BRR], By_L = UsL[BPP], Bz_L = UsL[BZZ];
   double vx_L = Sx_L / rho_L, vy_L = Sy_L / rho_L, vz_L = Sz_L / rho_L;
   double Sx_R = UsR[SRR], Sy_R = UsR[LLL], Sz_R = UsR[SZZ];
   double Bx_R = UsR[BRR], By_R = UsR[BPP], Bz_R = UsR[BZZ];
   double vx_R = Sx_R / rho_R, vy_R = Sy_R / rho_R, vz_R = Sz_R / rho_R;
   double Bn_L = Bx_L * n[0] + By_L * n[1] + Bz_L * n[2];
   double Bn_R = Bx_R * n[0] + By_R * n[1] + Bz_R * n[2];
   double Bn = .5 * (Bn_L + Bn_R);
   double signBn = Bn > 0.0 ? 1.0 : -1.0;
   double denom = rrh_L + rrh_R;
   double vx_ss = (rrh_L * vx_L + rrh_R * vx_R + (Bx_R - Bx_L) * signBn) / denom;
   double vy_ss = (rrh_L * vy_L + rrh_R * vy_R + (By_R - By_L) * signBn) / denom;
   double vz_ss = (rrh_L * vz_L + rrh_R * vz_R + (Bz_R - Bz_L) * signBn) / denom;
   double vn_ss = vx_ss * n[0] + vy_ss * n[1] + vz_ss * n[2];
   vx_ss += (vel[2] - vn_ss) * n[0];
   vy_ss += (vel[2] - vn_ss) * n[1];
   vz_ss += (vel[2] - vn_ss) * n[2];
   double Bx_ss = (rrh_L * Bx_R + rrh_R * Bx_L + rrh_L * rrh_R * (vx_R - vx_L) * signBn) / denom;
   double By_ss = (rrh_L * By_R + rrh_R * By_L + rrh_L * rrh_R * (vy_R - vy_L) * signBn) / denom;
   double Bz_ss = (rrh_L * Bz_R + rrh_R * Bz_L + rrh_L * rrh_R * (vz_R - vz_L) * signBn) / denom;
   double Bn_ss = Bx_ss * n[0] + By_ss * n[1] + Bz_ss * n[2];
   Bx_ss += (Bn - Bn_ss) * n[0];
   By_ss += (Bn - Bn_ss) * n[1];
   Bz_ss += (Bn - Bn_ss) * n[2];
   double vB_ss = vx_ss * Bx_ss + vy_ss * By_ss + vz_ss * Bz_ss;
   double vBstar, e_star;
   if (LR == 0) {
      vBstar = vx_L * Bx_L + vy_L * By_L + vz_L * Bz_L;
      e_star = UsL[TAU];
   } else {
      vBstar = vx_R * Bx_R + vy_R * By_R + vz_R * Bz_R;
      e_star = UsR[TAU];
   }
   double plusminus = LR == 1 ? 1.0 : -1.0;
   double e_ss = e_star + plusminus * rrh_K * (vBstar - vB_ss) * signBn;
   Uss[DDD] = rhostar_K;
   Uss[SRR] = rhostar_K * vx_ss;
   Uss[LLL] = rhostar_K * vy_ss;
   Uss[SZZ] = rhostar_K * vz_ss;
   Uss[TAU] = e_ss;
   Uss[BRR] = Bx_ss;
   Uss[BPP] = By_ss;
   Uss[BZZ] = Bz_ss;
}

/* get_Ustar_HLLD, hlld.c:282-401: flux F and state U of the fan at speed w */
static void hlld(const or_ctx *o, double w, const double *primL, const double *primR,
                 double *F, double *U, const double *x, const double *n)
{
   double gam = o->p.gamma_law;
   double vel[5], Pt_star;
   double hr = 1.0, hp = x[0], hz = 1.0;  /* get_scale_factor, cylindrical.c:66-71 */
   double pL[MAXQ], pR[MAXQ];
   int nq = o->nq, q;
   memcpy(pL, primL, nq * sizeof(double));
   memcpy(pR, primR, nq * sizeof(double));
   pL[URR] *= hr; pR[URR] *= hr;
   pL[UPP] *= hp; pR[UPP] *= hp;
   pL[UZZ] *= hz; pR[UZZ] *= hz;
   hd_velocities(gam, pL, pR, n, vel, &Pt_star);
   if (w < vel[2]) {
      double UL[8];
      hd_flux(gam, pL, F, n);
      hd_cons(gam, pL, UL);
      if (w < vel[0]) {
         hd_cons(gam, pL, U);
      } else if (w < vel[1]) {
         hd_single_star(gam, pL, U, vel[0], vel[2], Pt_star, n);
         for (q = 0; q < 8; ++q) F[q] += vel[0] * (U[q] - UL[q]);
      } else {
         double UsL[8], UsR[8];
         hd_single_star(gam, pL, UsL, vel[0], vel[2], Pt_star, n);
         hd_single_star(gam, pR, UsR, vel[4], vel[2], Pt_star, n);
         hd_double_star(UsL, UsR, U, vel, n, 0);
         for (q = 0; q < 8; ++q) F[q] += vel[1] * (U[q] - UsL[q]) + vel[0] * (UsL[q] - UL[q]);
      }
      for (q = 8; q < nq; q++) { U[q] = pL[q] * U[DDD]; F[q] = pL[q] * F[DDD]; }
   } else {
      double UR[8];
      hd_flux(gam, pR, F, n);
      hd_cons(gam, pR, UR);
      if (w > vel[4]) {
         hd_cons(gam, pR, U);
      } else if (w > vel[3]) {
         hd_single_star(gam, pR, U, vel[4], vel[2], Pt_star, n);
         for (q = 0; q < 8; ++q) F[q] += vel[4] * (U[q] - UR[q]);
      } else {
         double UsL[8], UsR[8];
         hd_single_star(gam, pL, UsL, vel[0], vel[2], Pt_star, n);
         hd_single_star(gam, pR, UsR, vel[4], vel[2], Pt_star, n);
         hd_double_star(UsL, UsR, U, vel, n, 1);
         for (q = 0; q < 8; ++q) F[q] += vel[3] * (U[q] - UsR[q]) + vel[4] * (UsR[q] - UR[q]);
      }
      for (q = 8; q < nq; q++) { U[q] = pR[q] * U[DDD]; F[q] = pR[q] * F[DDD]; }
   }
   double brfac = m_bscale(x[0], 0), bpfac = 1.0, bzfac = 1.0;
   F[SRR] *= hr; F[LLL] *= hp; F[SZZ] *= hz;
   F[BRR] *= brfac / hr; F[BPP] *= bpfac / hp; F[BZZ] *= bzfac / hz;
   U[SRR] *= hr; U[LLL] *= hp; U[SZZ] *= hz;
   U[BRR] *= brfac / hr; U[BPP] *= bpfac / hp; U[BZZ] *= bzfac / hz;
}

/* ------------------------------------------------------------------------ */
/* riemann.c                                                                 */
/* ------------------------------------------------------------------------ */

/* solve_riemann, riemann.c:216-318; returns F - w U* in Fw and, for MHD, the
 * EMFs of flux_to_E (mhd.c:293-316) for face direction dim (0 phi, 1 r, 2 z). */
static void solve_riemann(const or_ctx *o, const double *primL, const double *primR,
                          const double *gLr, const double *gLp, const double *gLz,
                          const double *gRr, const double *gRp, const double *gRz,
                          const double *x, const double *n, double w, double *Fw,
                          int dim, double *emf)
{
   int nq = o->nq, q;
   double Flux[MAXQ], Ustr[MAXQ];
   for (q = 0; q < nq; ++q) { Flux[q] = 0.0; Ustr[q] = 0.0; }
   int rs = o->p.riemann_solver;
   if (o->p.hydro == DH_HYDRO_MHD) {
      hlld(o, w, primL, primR, Flux, Ustr, x, n);
   } else if (rs == R_HLL || rs == R_HLLC) {
      double Sl, Sr, Ss;
      e_vel(o, primL, primR, &Sl, &Sr, &Ss, n, x);
      if (w < Sl) {
         e_flux(o, primL, Flux, x, n);
         e_prim2cons(o, primL, Ustr, x, 1.0);
      } else if (w > Sr) {
         e_flux(o, primR, Flux, x, n);
         e_prim2cons(o, primR, Ustr, x, 1.0);
      } else if (rs == R_HLL) {
         double Fl[MAXQ], Fr[MAXQ], Ul[MAXQ], Ur[MAXQ];
         double aL = Sr;
         double aR = -Sl;
         e_prim2cons(o, primL, Ul, x, 1.0);
         e_prim2cons(o, primR, Ur, x, 1.0);
         e_flux(o, primL, Fl, x, n);
         e_flux(o, primR, Fr, x, n);
         for (q = 0; q < nq; ++q) {
            Flux[q] = (aL * Fl[q] + aR * Fr[q] + aL * aR * (Ul[q] - Ur[q])) / (aL + aR);
            Ustr[q] = (aR * Ul[q] + aL * Ur[q] + Fl[q] - Fr[q]) / (aL + aR);
         }
      } else {
         double Uk[MAXQ], Fk[MAXQ];
         if (w < Ss) {
            e_prim2cons(o, primL, Uk, x, 1.0);
            e_getUstar(o, primL, Ustr, x, Sl, Ss, n);
            e_flux(o, primL, Fk, x, n);
            for (q = 0; q < nq; ++q) Flux[q] = Fk[q] + Sl * (Ustr[q] - Uk[q]);
         } else {
            e_prim2cons(o, primR, Uk, x, 1.0);
            e_getUstar(o, primR, Ustr, x, Sr, Ss, n);
            e_flux(o, primR, Fk, x, n);
            for (q = 0; q < nq; ++q) Flux[q] = Fk[q] + Sr * (Ustr[q] - Uk[q]);
         }
      }
   } else {
      fprintf(stderr, "oracle: HLLD needs the MHD module\n");
      abort();
   }
   if (o->p.visc_flag && o->p.hydro == DH_HYDRO_EULER) {
      double vFlux[MAXQ] = {0}, prim[MAXQ] = {0}, gr[MAXQ] = {0}, gp[MAXQ] = {0},
             gz[MAXQ] = {0};
      for (q = 0; q < nq; ++q) {
         prim[q] = .5 * (primL[q] + primR[q]);
         gr[q] = .5 * (gLr[q] + gRr[q]);
         gp[q] = .5 * (gLp[q] + gRp[q]);
         gz[q] = .5 * (gLz[q] + gRz[q]);
         vFlux[q] = 0.0;
      }
      e_visc_flux(o, prim, gr, gp, gz, vFlux, x, n);
      for (q = 0; q < nq; ++q) Flux[q] += vFlux[q];
   }
   for (q = 0; q < nq; ++q) Fw[q] = Flux[q] - w * Ustr[q];
   if (o->p.hydro == DH_HYDRO_MHD && emf) {
      double d0 = 0.0;
      m_flux_to_E(Flux, Ustr, x, &emf[0], &emf[1], &emf[2], dim == 0 ? &emf[3] : &d0, dim);
   }
}

/* ------------------------------------------------------------------------ */
/* Context                                                                   */
/* ------------------------------------------------------------------------ */

static void *xcalloc(size_t n, size_t s)
{
   void *p = calloc(n ? n : 1, s);
   if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
   return p;
}

or_ctx *or_create(const dh_grid *g, const dh_params *par)
{
   or_ctx *o = (or_ctx *)xcalloc(1, sizeof *o);
   o->p = *par;
   o->nq = par->num_c + par->num_n;
   o->nf = par->ct_mode == 1 ? 3 : (par->ct_mode == 2 ? 5 : 0);
   if (o->nq > MAXQ) { fprintf(stderr, "oracle: NQ too large\n"); abort(); }
   if (!(par->phi_order > 0.0)) { fprintf(stderr, "oracle: phi_order must be 2\n"); abort(); }
   if (par->hydro == DH_HYDRO_MHD && (par->ct_mode != 2 || par->riemann_solver != R_HLLD)) {
      fprintf(stderr, "oracle: MHD is restated with CT_MODE 2 and HLLD only\n");
      abort();
   }
   o->Nr = g->Nr; o->Nz = g->Nz;
   o->NgRa = g->NgRa; o->NgRb = g->NgRb; o->NgZa = g->NgZa; o->NgZb = g->NgZb;
   o->phi_max = g->phi_max;
   int na = o->Nr * o->Nz;
   o->Np = (int *)xcalloc(na, sizeof(int));
   o->off = (long *)xcalloc(na + 1, sizeof(long));
   for (int jk = 0; jk < na; ++jk) {
      o->Np[jk] = g->Np[jk];
      o->off[jk + 1] = o->off[jk] + g->Np[jk];
   }
   o->ntot = o->off[na];
   o->rbuf = (double *)xcalloc(o->Nr + 1, sizeof(double));
   o->zbuf = (double *)xcalloc(o->Nz + 1, sizeof(double));
   memcpy(o->rbuf, g->r_jph, (o->Nr + 1) * sizeof(double));
   memcpy(o->zbuf, g->z_kph, (o->Nz + 1) * sizeof(double));
   o->r_jph = o->rbuf + 1;
   o->z_kph = o->zbuf + 1;
   long nt = o->ntot;
   o->prim = (double *)xcalloc(nt * o->nq, sizeof(double));
   o->cons = (double *)xcalloc(nt * o->nq, sizeof(double));
   o->rkcons = (double *)xcalloc(nt * o->nq, sizeof(double));
   o->gradr = (double *)xcalloc(nt * o->nq, sizeof(double));
   o->gradp = (double *)xcalloc(nt * o->nq, sizeof(double));
   o->gradz = (double *)xcalloc(nt * o->nq, sizeof(double));
   o->piph = (double *)xcalloc(nt, sizeof(double));
   o->dphi = (double *)xcalloc(nt, sizeof(double));
   o->wiph = (double *)xcalloc(nt, sizeof(double));
   o->Phi = (double *)xcalloc(nt * (o->nf ? o->nf : 1), sizeof(double));
   o->rkPhi = (double *)xcalloc(nt * (o->nf ? o->nf : 1), sizeof(double));
   o->Ecell = (double *)xcalloc(nt * 8, sizeof(double));
   o->Bcell = (double *)xcalloc(nt * 8, sizeof(double));
   o->Ephi = (double *)xcalloc(nt * 4, sizeof(double));
   o->temp = (double *)xcalloc(nt, sizeof(double));
   /* get_num_rzFaces, faces_alt.c:5-8 */
   o->nftracks[0] = (o->Nr - 1) * o->Nz;
   o->nftracks[1] = (o->Nz - 1) * o->Nr;
   for (int d = 0; d < 2; ++d) {
      o->fI[d] = (long *)xcalloc(o->nftracks[d] + 1, sizeof(long));
      o->faces[d] = NULL;
   }
   /* euler.c:23-24 and planet.c:15-20: polar source corrections when the
    * cylindrical mesh reaches r = 0. */
   o->polar_sources = par->nobc_rmin == 1;
   return o;
}

void or_destroy(or_ctx *o)
{
   if (!o) return;
   free(o->Np); free(o->off); free(o->rbuf); free(o->zbuf);
   free(o->prim); free(o->cons); free(o->rkcons); free(o->gradr);
   free(o->gradp); free(o->gradz); free(o->piph); free(o->dphi); free(o->wiph);
   free(o->Phi); free(o->rkPhi);
   free(o->Ecell); free(o->Bcell); free(o->Ephi); free(o->temp);
   for (int d = 0; d < 2; ++d) { free(o->fI[d]); free(o->faces[d]); }
   free(o);
}

long or_ntot(const or_ctx *o) { return o->ntot; }

void or_set_state(or_ctx *o, const double *prim, const double *cons,
                  const double *piph, const double *dphi, const double *wiph,
                  const double *Phi)
{
   long nt = o->ntot;
   if (prim) memcpy(o->prim, prim, nt * o->nq * sizeof(double));
   if (cons) memcpy(o->cons, cons, nt * o->nq * sizeof(double));
   if (piph) memcpy(o->piph, piph, nt * sizeof(double));
   if (dphi) memcpy(o->dphi, dphi, nt * sizeof(double));
   if (wiph) memcpy(o->wiph, wiph, nt * sizeof(double));
   if (Phi && o->nf) memcpy(o->Phi, Phi, nt * o->nf * sizeof(double));
}

void or_get_state(const or_ctx *o, double *prim, double *cons, double *piph,
                  double *dphi, double *wiph, double *Phi)
{
   long nt = o->ntot;
   if (prim) memcpy(prim, o->prim, nt * o->nq * sizeof(double));
   if (cons) memcpy(cons, o->cons, nt * o->nq * sizeof(double));
   if (piph) memcpy(piph, o->piph, nt * sizeof(double));
   if (dphi) memcpy(dphi, o->dphi, nt * sizeof(double));
   if (wiph) memcpy(wiph, o->wiph, nt * sizeof(double));
   if (Phi && o->nf) memcpy(Phi, o->Phi, nt * o->nf * sizeof(double));
}

void or_set_planets(or_ctx *o, int substep, const double *pl, int npl)
{
   if (substep < 0 || substep > 1 || npl < 0 || npl > DH_MAX_PLANETS) abort();
   o->npl[substep] = npl;
   memcpy(o->planets[substep], pl, (size_t)npl * DH_PLANET_STRIDE * sizeof(double));
}

/* ------------------------------------------------------------------------ */
/* misc.c: per-cell drivers                                                  */
/* ------------------------------------------------------------------------ */

static void gather(const or_ctx *o, const double *soa, long c, double *v)
{
   for (int q = 0; q < o->nq; ++q) v[q] = soa[(long)q * o->ntot + c];
}

static void scatter(const or_ctx *o, double *soa, long c, const double *v)
{
   for (int q = 0; q < o->nq; ++q) soa[(long)q * o->ntot + c] = v[q];
}

/* getmindt, misc.c:39-84 */
double or_getmindt(or_ctx *o)
{
   double dt = o->p.max_dt / o->p.cfl;
   if (dt <= 0.0) dt = 1.0e100;
   double prim[MAXQ];
   for (int k = o->NgZa; k < o->Nz - o->NgZb; ++k)
      for (int j = o->NgRa; j < o->Nr - o->NgRb; ++j) {
         int jk = j + o->Nr * k;
         long c0 = o->off[jk];
         for (int i = 0; i < o->Np[jk]; ++i) {
            long c = c0 + i;
            double phip = o->piph[c];
            double phim = phip - o->dphi[c];
            double xp[3] = {o->r_jph[j], phip, o->z_kph[k]};
            double xm[3] = {o->r_jph[j - 1], phim, o->z_kph[k - 1]};
            int im = i - 1;
            if (i == 0) im = o->Np[jk] - 1;
            double w = .5 * (o->wiph[c0 + im] + o->wiph[c]);
            gather(o, o->prim, c, prim);
            double dtt = o->p.hydro == DH_HYDRO_MHD ? m_mindt(o, prim, w, xp, xm)
                                                    : e_mindt(o, prim, w, xp, xm);
            if (dt > dtt) dt = dtt;
         }
      }
   dt *= o->p.cfl;
   return dt;
}

/* set_wcell, misc.c:101-168 (get_omega = prim[UPP], euler.c:32-34) */
static void set_wcell(or_ctx *o)
{
   int mm = o->p.mesh_motion;
   for (int k = 0; k < o->Nz; ++k)
      for (int j = 0; j < o->Nr; ++j) {
         int jk = j + o->Nr * k;
         long c0 = o->off[jk];
         int np = o->Np[jk];
         for (int i = 0; i < np; ++i) {
            double w = 0.0;
            if (mm) {
               int ip = (i + 1) % np;
               double wL = PRIM(o, UPP, c0 + i);
               double wR = PRIM(o, UPP, c0 + ip);
               w = .5 * (wL + wR);
            }
            o->wiph[c0 + i] = w;
         }
      }
   if (mm == 3) {
      for (int jk = 0; jk < o->Nr * o->Nz; ++jk) {
         double w = 0.0;
         long c0 = o->off[jk];
         for (int i = 0; i < o->Np[jk]; ++i) w += o->wiph[c0 + i];
         w /= (double)o->Np[jk];
         for (int i = 0; i < o->Np[jk]; ++i) o->wiph[c0 + i] = w;
      }
   }
   if (mm == 4) {
      for (int k = 0; k < o->Nz; ++k) {
         double z = g_centroid(o->z_kph[k], o->z_kph[k - 1], 2);
         for (int j = 0; j < o->Nr; ++j) {
            int jk = j + o->Nr * k;
            double r = g_centroid(o->r_jph[j], o->r_jph[j - 1], 1);
            long c0 = o->off[jk];
            for (int i = 0; i < o->Np[jk]; ++i) {
               double phip = o->piph[c0 + i];
               double phim = phip - o->dphi[c0 + i];
               double x[3] = {r, 0.5 * (phim + phip), z};
               o->wiph[c0 + i] = mesh_om(o, x);
            }
         }
      }
   }
}

/* adjust_RK_cons, misc.c:226-244 */
static void adjust_RK_cons(or_ctx *o, double RK)
{
   long n = o->ntot * o->nq;
   for (long a = 0; a < n; ++a) o->cons[a] = (1. - RK) * o->cons[a] + RK * o->rkcons[a];
   long nfa = o->ntot * o->nf;
   for (long a = 0; a < nfa; ++a) o->Phi[a] = (1. - RK) * o->Phi[a] + RK * o->rkPhi[a];
}

/* minmod, plm.c:6-13 */
static double minmod(double a, double b, double c)
{
   double m = a;
   if (a * b < 0.0) m = 0.0;
   if (fabs(b) < fabs(m)) m = b;
   if (b * c < 0.0) m = 0.0;
   if (fabs(c) < fabs(m)) m = c;
   return m;
}

/* plm_phi, plm.c:15-51 */
static void plm_phi(or_ctx *o)
{
   double PLM = o->p.plm;
   for (int jk = 0; jk < o->Nr * o->Nz; ++jk) {
      long c0 = o->off[jk];
      int np = o->Np[jk];
      for (int i = 0; i < np; ++i) {
         int im = i - 1;
         if (im == -1) im = np - 1;
         int ip = (i + 1) % np;
         long c = c0 + i, cL = c0 + im, cR = c0 + ip;
         double dpL = o->dphi[cL];
         double dpC = o->dphi[c];
         double dpR = o->dphi[cR];
         for (int q = 0; q < o->nq; ++q) {
            double pL = PRIM(o, q, cL);
            double pC = PRIM(o, q, c);
            double pR = PRIM(o, q, cR);
            double sL = pC - pL;
            sL /= .5 * (dpC + dpL);
            double sR = pR - pC;
            sR /= .5 * (dpR + dpC);
            double sC = pR - pL;
            sC /= .5 * (dpL + dpR) + dpC;
            o->gradp[(long)q * o->ntot + c] = minmod(PLM * sL, sC, PLM * sR);
         }
      }
   }
}

/* Ring start per annulus: the first cell (scanning from i=0) whose piph lies in
 * the first quadrant right after one in the fourth (faces_alt.c:59-78,
 * bfields.c:899-918). */
static void ring_starts(const or_ctx *o, int *I0)
{
   double convert = 2. * M_PI / o->phi_max;
   for (int jk = 0; jk < o->Nr * o->Nz; ++jk) {
      int found = 0, quad_prev = 0;
      long c0 = o->off[jk];
      for (int i = 0; i < o->Np[jk] && !found; ++i) {
         double sn = sin(o->piph[c0 + i] * convert);
         double cs = cos(o->piph[c0 + i] * convert);
         if (sn > 0. && cs > 0. && quad_prev) {
            quad_prev = 0;
            found = 1;
            I0[jk] = i;
         } else if (sn < 0. && cs > 0.) {
            quad_prev = 1;
         }
      }
      if (!found) I0[jk] = 0;
   }
}

/* buildfaces (both passes) + addFace, faces_alt.c:10-179 */
static void build_faces(or_ctx *o, int dim)
{
   int d = dim - 1;
   int Nr = o->Nr, Nz = o->Nz;
   int *I0 = (int *)xcalloc(Nr * Nz, sizeof(int));
   ring_starts(o, I0);
   int Nrmax = dim == 2 ? Nr : Nr - 1;
   int Nzmax = dim == 1 ? Nz : Nz - 1;
   long *ntj = o->fI[d];
   long n = 0;
   for (int k = 0; k < Nzmax; ++k)
      for (int j = 0; j < Nrmax; ++j) {
         int JK = j + Nrmax * k;
         int jk = j + Nr * k;
         int jkp = dim == 2 ? j + Nr * (k + 1) : j + 1 + Nr * k;
         ntj[JK] = n;
         n += o->Np[jk] + o->Np[jkp];
      }
   ntj[Nrmax * Nzmax] = n;
   free(o->faces[d]);
   o->faces[d] = (or_face *)xcalloc(n, sizeof(or_face));
   or_face *F = o->faces[d];
   double Pmax = o->phi_max;
   n = 0;
   for (int k = 0; k < Nzmax; ++k)
      for (int j = 0; j < Nrmax; ++j) {
         int jp = j + 1, kp = k + 1;
         int jk = j + Nr * k;
         int jkp = dim == 2 ? j + Nr * kp : jp + Nr * k;
         double dxL, dxR;
         if (dim == 1) {
            double rm = g_centroid(o->r_jph[j], o->r_jph[j - 1], 1);
            double rp = g_centroid(o->r_jph[jp], o->r_jph[j], 1);
            dxL = o->r_jph[j] - rm;
            dxR = rp - o->r_jph[j];
         } else {
            double zm = g_centroid(o->z_kph[k], o->z_kph[k - 1], 2);
            double zp = g_centroid(o->z_kph[kp], o->z_kph[k], 2);
            dxL = o->z_kph[k] - zm;
            dxR = zp - o->z_kph[k];
         }
         double xp[3] = {o->r_jph[j], 0.0, o->z_kph[k]};
         double xm[3] = {o->r_jph[j], 0.0, o->z_kph[k - 1]};
         if (dim == 2) { xm[0] = o->r_jph[j - 1]; xm[2] = o->z_kph[k]; }
         int i = I0[jk], ip = I0[jkp];
         int nf = o->Np[jk] + o->Np[jkp];
         for (int f = 0; f < nf; ++f) {
            long cL = o->off[jk] + i, cR = o->off[jkp] + ip;
            double dphi = o->piph[cL] - o->piph[cR];
            while (dphi > .5 * Pmax) dphi -= Pmax;
            while (dphi < -.5 * Pmax) dphi += Pmax;
            int LR = dphi > 0. ? 1 : 0;
            dphi = o->piph[cL] - o->dphi[cL] - o->piph[cR] + o->dphi[cR];
            while (dphi > .5 * Pmax) dphi -= Pmax;
            while (dphi < -.5 * Pmax) dphi += Pmax;
            int LR_back = dphi < 0. ? 1 : 0;
            xp[1] = LR == 0 ? o->piph[cL] : o->piph[cR];
            xm[1] = LR_back == 0 ? o->piph[cL] - o->dphi[cL] : o->piph[cR] - o->dphi[cR];
            /* addFace, faces_alt.c:10-34 */
            or_face *fc = F + n;
            double dp = g_dp(o, xp[1], xm[1]);
            double phic = g_dp(o, xp[1], .5 * dp);
            double r = dim == 1 ? 0.5 * (xp[0] + xm[0]) : g_centroid(xp[0], xm[0], 1);
            double z = dim == 2 ? 0.5 * (xp[2] + xm[2]) : g_centroid(xp[2], xm[2], 2);
            fc->cm[0] = r; fc->cm[1] = phic; fc->cm[2] = z;
            fc->L = cL; fc->R = cR;
            fc->dxL = dxL; fc->dxR = dxR;
            fc->dphi = dp;
            fc->dA = g_dA(o, xp, xm, dim);
            fc->dl = g_dL(o, xp, xm, 3 - dim);
            fc->LRtype = LR;
            ++n;
            if (LR == 0) { ++i; if (i == o->Np[jk]) i = 0; }
            else { ++ip; if (ip == o->Np[jkp]) ip = 0; }
         }
      }
   free(I0);
}

/* geom_grad, cylindrical.c:141-173 (polar r=0 boundary, LR=0) */
static void geom_grad_polar(or_ctx *o, long c, const double *xp, const double *xm)
{
   if (xp[0] < 0.0 || fabs(xm[0]) > 1.0e-10 * xp[0]) return;
   double r = g_centroid(xp[0], xm[0], 1);
   double PLM = o->p.plm;
   for (int q = 0; q < o->nq; ++q) {
      double *g = &o->gradr[(long)q * o->ntot + c];
      if (q == URR || (o->p.num_c > BZZ && q == BRR)) {
         double SL = PRIM(o, q, c) / r;
         double S = *g;
         if (S * SL < 0.0) *g = 0.0;
         else if (fabs(PLM * SL) < fabs(S)) *g = PLM * SL;
      } else {
         *g = 0.0;
      }
   }
}

/* plm_trans + plm_geom_boundary, plm.c:56-232 */
static void plm_trans(or_ctx *o, int dim)
{
   int d = dim - 1;
   double *grad = dim == 1 ? o->gradr : o->gradz;
   long nt = o->ntot;
   memset(grad, 0, nt * o->nq * sizeof(double));
   double PLM = o->p.plm;
   if (PLM == 0.0) return;
   or_face *F = o->faces[d];
   long Nf = o->fI[d][o->nftracks[d]];
   for (long n = 0; n < Nf; ++n) {
      or_face *f = F + n;
      double phi = f->cm[1];
      long cL = f->L, cR = f->R;
      double phiL = o->piph[cL] - .5 * o->dphi[cL];
      double phiR = o->piph[cR] - .5 * o->dphi[cR];
      double dpL = g_sdp(o, phi, phiL);
      double dpR = g_sdp(o, phiR, phi);
      for (int q = 0; q < o->nq; ++q) {
         double WL = PRIM(o, q, cL) + dpL * o->gradp[q * nt + cL];
         double WR = PRIM(o, q, cR) - dpR * o->gradp[q * nt + cR];
         double S = (WR - WL) / (f->dxR + f->dxL);
         grad[q * nt + cL] += S * f->dA;
         grad[q * nt + cR] += S * f->dA;
      }
   }
   for (int j = 0; j < o->Nr; ++j)
      for (int k = 0; k < o->Nz; ++k) {
         int jk = j + o->Nr * k;
         for (int i = 0; i < o->Np[jk]; ++i) {
            long c = o->off[jk] + i;
            double phip = o->piph[c];
            double phim = phip - o->dphi[c];
            double xp[3] = {o->r_jph[j], phip, o->z_kph[k]};
            double xm[3] = {o->r_jph[j - 1], phim, o->z_kph[k - 1]};
            if (dim == 1) xm[0] = o->r_jph[j];
            else xm[2] = o->z_kph[k];
            double dAp = g_dA(o, xp, xm, dim);
            if (dim == 1) { xp[0] = o->r_jph[j - 1]; xm[0] = o->r_jph[j - 1]; }
            else { xp[2] = o->z_kph[k - 1]; xm[2] = o->z_kph[k - 1]; }
            double dAm = g_dA(o, xp, xm, dim);
            double dAtot = dAp + dAm;
            if ((dim == 1 && j == 0) || (dim == 2 && k == 0)) dAtot = dAp;
            if ((dim == 1 && j == o->Nr - 1) || (dim == 2 && k == o->Nz - 1)) dAtot = dAm;
            for (int q = 0; q < o->nq; ++q) grad[q * nt + c] /= dAtot;
         }
      }
   for (long n = 0; n < Nf; ++n) {
      or_face *f = F + n;
      double phi = f->cm[1];
      long cL = f->L, cR = f->R;
      double phiL = o->piph[cL] - .5 * o->dphi[cL];
      double phiR = o->piph[cR] - .5 * o->dphi[cR];
      double dpL = g_sdp(o, phi, phiL);
      double dpR = g_sdp(o, phiR, phi);
      for (int q = 0; q < o->nq; ++q) {
         double WL = PRIM(o, q, cL) + dpL * o->gradp[q * nt + cL];
         double WR = PRIM(o, q, cR) - dpR * o->gradp[q * nt + cR];
         double S = (WR - WL) / (f->dxR + f->dxL);
         double *gL = &grad[q * nt + cL], *gR = &grad[q * nt + cR];
         double SL = *gL, SR = *gR;
         if (S * SL < 0.0) *gL = 0.0;
         else if (fabs(PLM * S) < fabs(SL)) *gL = PLM * S;
         if (S * SR < 0.0) *gR = 0.0;
         else if (fabs(PLM * S) < fabs(SR)) *gR = PLM * S;
      }
   }
   /* plm_geom_boundary: only the r = 0 axis is a geometric boundary of the
    * cylindrical mesh (cylindrical.c:144-149 ignores the other sides). */
   if (dim == 1 && o->NgRa == 0) {
      int j = 0;
      for (int k = 0; k < o->Nz; ++k) {
         int jk = j + o->Nr * k;
         for (int i = 0; i < o->Np[jk]; ++i) {
            long c = o->off[jk] + i;
            double xp[3] = {o->r_jph[j], o->piph[c], o->z_kph[k]};
            double xm[3] = {o->r_jph[j - 1], o->piph[c] - o->dphi[c], o->z_kph[k - 1]};
            geom_grad_polar(o, c, xp, xm);
         }
      }
   }
}

static void add_delta(or_ctx *o, long c, const double *Fw, double dAdt, double sgn)
{
   long nt = o->ntot;
   for (int q = 0; q < o->nq; ++q) {
      if (sgn < 0) o->cons[q * nt + c] -= Fw[q] * dAdt;
      else o->cons[q * nt + c] += Fw[q] * dAdt;
   }
}

/* phi_flux + riemann_phi, misc.c:368-426, riemann.c:39-112 */
static void phi_flux(or_ctx *o, double dt)
{
   long nt = o->ntot;
   int jmin = o->NgRa, jmax = o->Nr - o->NgRb, kmin = o->NgZa, kmax = o->Nz - o->NgZb;
   int ct = o->p.hydro == DH_HYDRO_MHD && o->p.ct;
   if (ct) {
      jmin = o->NgRa == 0 ? 0 : o->NgRa - 1;
      jmax = o->NgRb == 0 ? o->Nr : o->Nr - o->NgRb + 1;
      kmin = o->NgZa == 0 ? 0 : o->NgZa - 1;
      kmax = o->NgZb == 0 ? o->Nz : o->Nz - o->NgZb + 1;
   }
   double n[3] = {0.0, 1.0, 0.0};
   for (int k = kmin; k < kmax; ++k) {
      double zp = o->z_kph[k], zm = o->z_kph[k - 1];
      double z = g_centroid(zp, zm, 2);
      for (int j = jmin; j < jmax; ++j) {
         double rp = o->r_jph[j], rm = o->r_jph[j - 1];
         double r = g_centroid(rp, rm, 1);
         int jk = j + o->Nr * k;
         long c0 = o->off[jk];
         int np = o->Np[jk];
         for (int i = 0; i < np; ++i) {
            int ip = i < np - 1 ? i + 1 : 0;
            long cL = c0 + i, cR = c0 + ip;
            double phi = o->piph[cL];
            double xp[3] = {rp, phi, zp};
            double xm[3] = {rm, phi, zm};
            double dA = g_dA(o, xp, xm, 0);
            double x[3] = {r, phi, z};
            double dAdt = dA * dt;
            double primL[MAXQ], primR[MAXQ], gp[MAXQ], gLr[MAXQ], gLz[MAXQ], gRr[MAXQ], gRz[MAXQ];
            for (int q = 0; q < o->nq; ++q) {
               primL[q] = PRIM(o, q, cL) + .5 * o->gradp[q * nt + cL] * o->dphi[cL];
               primR[q] = PRIM(o, q, cR) - .5 * o->gradp[q * nt + cR] * o->dphi[cR];
               gp[q] = (PRIM(o, q, cR) - PRIM(o, q, cL)) / (0.5 * (o->dphi[cR] + o->dphi[cL]));
               gLr[q] = o->gradr[q * nt + cL]; gLz[q] = o->gradz[q * nt + cL];
               gRr[q] = o->gradr[q * nt + cR]; gRz[q] = o->gradz[q * nt + cR];
            }
            double hn = x[0]; /* get_scale_factor(x, 0), cylindrical.c:66-71 */
            if (o->p.hydro == DH_HYDRO_MHD) {
               double Bp = .5 * (primL[BPP] + primR[BPP]);
               primL[BPP] = Bp;
               primR[BPP] = Bp;
            }
            double Fw[MAXQ], emf[4];
            solve_riemann(o, primL, primR, gLr, gp, gLz, gRr, gp, gRz, x, n,
                          hn * o->wiph[cL], Fw, 0, emf);
            add_delta(o, cL, Fw, dAdt, -1.0);
            add_delta(o, cR, Fw, dAdt, +1.0);
            if (o->p.hydro == DH_HYDRO_MHD) {
               /* riemann.c:86-109 (NUM_EDGES == 8) */
               double Ez = emf[0], Br = emf[1], Er = emf[2], Bz = emf[3];
               double *E = o->Ecell, *B = o->Bcell;
               E[0 * nt + cL] = .5 * Ez; E[1 * nt + cL] = .5 * Ez;
               E[2 * nt + cR] = .5 * Ez; E[3 * nt + cR] = .5 * Ez;
               B[0 * nt + cL] = .5 * Br; B[1 * nt + cL] = .5 * Br;
               B[2 * nt + cR] = .5 * Br; B[3 * nt + cR] = .5 * Br;
               E[4 * nt + cL] = .5 * Er; E[5 * nt + cL] = .5 * Er;
               E[6 * nt + cR] = .5 * Er; E[7 * nt + cR] = .5 * Er;
               B[4 * nt + cL] = .5 * Bz; B[5 * nt + cL] = .5 * Bz;
               B[6 * nt + cR] = .5 * Bz; B[7 * nt + cR] = .5 * Bz;
            }
         }
      }
   }
}

/* trans_flux + riemann_trans, misc.c:431-509, riemann.c:114-213 */
static void trans_flux(or_ctx *o, double dt, int dim)
{
   int d = dim - 1;
   long nt = o->ntot;
   int Nr = o->Nr, Nz = o->Nz;
   int jmin, jmax, kmin, kmax, Nfr;
   int ct = o->p.hydro == DH_HYDRO_MHD && o->p.ct;
   if (dim == 1) {
      Nfr = Nr - 1;
      jmin = o->NgRa == 0 ? 0 : o->NgRa - 1;
      jmax = o->NgRb == 0 ? Nr - 1 : Nr - o->NgRb;
      kmin = o->NgZa;
      kmax = Nz - o->NgZb;
      if (ct) {
         kmin = o->NgZa == 0 ? 0 : o->NgZa - 1;
         kmax = o->NgZb == 0 ? Nz : Nz - o->NgZb + 1;
      }
   } else {
      Nfr = Nr;
      jmin = o->NgRa;
      jmax = Nr - o->NgRb;
      kmin = o->NgZa == 0 ? 0 : o->NgZa - 1;
      kmax = o->NgZb == 0 ? Nz - 1 : Nz - o->NgZb;
      if (ct) {
         jmin = o->NgRa == 0 ? 0 : o->NgRa - 1;
         jmax = o->NgRb == 0 ? Nr : Nr - o->NgRb + 1;
      }
   }
   double *grad = dim == 1 ? o->gradr : o->gradz;
   double n[3] = {0.0, 0.0, 0.0};
   if (dim == 1) n[0] = 1.0; else n[2] = 1.0;
   for (int k = kmin; k < kmax; ++k)
      for (int j = jmin; j < jmax; ++j) {
         int JK = j + Nfr * k;
         for (long fidx = o->fI[d][JK]; fidx < o->fI[d][JK + 1]; ++fidx) {
            or_face *F = o->faces[d] + fidx;
            long cL = F->L, cR = F->R;
            double dAdt = F->dA * dt;
            double dxL = F->dxL, dxR = F->dxR;
            double phi = F->cm[1];
            double phiL = o->piph[cL] - .5 * o->dphi[cL];
            double phiR = o->piph[cR] - .5 * o->dphi[cR];
            double dpL = g_sdp(o, phi, phiL);
            double dpR = g_sdp(o, phiR, phi);
            double primL[MAXQ], primR[MAXQ], gd[MAXQ];
            double gLr[MAXQ], gLp[MAXQ], gLz[MAXQ], gRr[MAXQ], gRp[MAXQ], gRz[MAXQ];
            for (int q = 0; q < o->nq; ++q) {
               double gpL = o->gradp[q * nt + cL], gpR = o->gradp[q * nt + cR];
               primL[q] = PRIM(o, q, cL) + grad[q * nt + cL] * dxL + gpL * dpL;
               primR[q] = PRIM(o, q, cR) - grad[q * nt + cR] * dxR - gpR * dpR;
               gd[q] = ((PRIM(o, q, cR) - gpR * dpR) - (PRIM(o, q, cL) + gpL * dpL)) / (dxL + dxR);
               gLr[q] = o->gradr[q * nt + cL]; gLp[q] = gpL; gLz[q] = o->gradz[q * nt + cL];
               gRr[q] = o->gradr[q * nt + cR]; gRp[q] = gpR; gRz[q] = o->gradz[q * nt + cR];
            }
            if (o->p.hydro == DH_HYDRO_MHD) {
               int bt = dim == 1 ? BRR : BZZ;
               double Bavg = .5 * (primL[bt] + primR[bt]);
               primL[bt] = Bavg;
               primR[bt] = Bavg;
            }
            double Fw[MAXQ], emf[4];
            if (dim == 1)
               solve_riemann(o, primL, primR, gd, gLp, gLz, gd, gRp, gRz, F->cm, n, 0.0, Fw,
                             dim, emf);
            else
               solve_riemann(o, primL, primR, gLr, gLp, gd, gRr, gRp, gd, F->cm, n, 0.0, Fw,
                             dim, emf);
            add_delta(o, cL, Fw, dAdt, -1.0);
            add_delta(o, cR, Fw, dAdt, +1.0);
            if (o->p.hydro == DH_HYDRO_MHD) {
               /* riemann.c:166-212 (NUM_EDGES == 8, NUM_AZ_EDGES == 4) */
               double Erz = emf[0], Brz = emf[1], Ephi = emf[2];
               double fracL = F->dphi / o->dphi[cL];
               double fracR = F->dphi / o->dphi[cR];
               double *E = o->Ecell, *B = o->Bcell;
               int a0 = dim == 1 ? 1 : 5, a1 = dim == 1 ? 3 : 7;  /* L's r+/z+ edges */
               int b0 = dim == 1 ? 0 : 4, b1 = dim == 1 ? 2 : 6;  /* R's r-/z- edges */
               E[a0 * nt + cL] += .5 * Erz * fracL;
               E[a1 * nt + cL] += .5 * Erz * fracL;
               E[b0 * nt + cR] += .5 * Erz * fracR;
               E[b1 * nt + cR] += .5 * Erz * fracR;
               B[a0 * nt + cL] += .5 * Brz * fracL;
               B[a1 * nt + cL] += .5 * Brz * fracL;
               B[b0 * nt + cR] += .5 * Brz * fracR;
               B[b1 * nt + cR] += .5 * Brz * fracR;
               if (F->LRtype == 0) o->Ephi[(dim == 1 ? 1 : 3) * nt + cL] = Ephi;
               else o->Ephi[(dim == 1 ? 0 : 2) * nt + cR] = Ephi;
            }
         }
      }
}

/* fgrav, planet.c:56-76 */
static double fgrav(double M, double r, double eps, int type, double n)
{
   if (type == 0) {
      return M * pow(r, n - 1.) / pow(pow(r, n) + pow(eps, n), 1. + 1. / n);
   } else if (type == 1) {
      return M / ((r - 2 * M) * (r - 2 * M));
   } else if (type == 2) {
      return M;
   }
   return 0.0;
}

/* planetaryForce, planet.c:99-150 (mode 0) */
static void planetary_force(const or_ctx *o, const double *pl, double r, double phi,
                            double z, double *fr, double *fp, double *fz)
{
   double rp = pl[0], pp = pl[1];
   if (o->p.grav2d == 1) z = 0.0;
   else if (o->p.grav2d == 2) { rp = r; pp = phi; z = 1.0; }
   double cosp = cos(phi);
   double sinp = sin(phi);
   double dx = r * cosp - rp * cos(pp);
   double dy = r * sinp - rp * sin(pp);
   double script_r = sqrt(dx * dx + dy * dy + z * z);
   double script_r_perp = sqrt(dx * dx + dy * dy);
   double f1 = -fgrav(pl[2], script_r, pl[5], (int)pl[6], o->p.phi_order);
   double cosa = dx / script_r_perp;
   double sina = dy / script_r_perp;
   double cosap = cosa * cosp + sina * sinp;
   double sinap = sina * cosp - cosa * sinp;
   if (script_r_perp <= 0.0) { cosap = 0.0; sinap = 0.0; }
   double sint = script_r_perp / script_r;
   double cost = z / script_r;
   *fr = cosap * f1 * sint;
   *fp = sinap * f1 * sint;
   *fz = f1 * cost;
}

/* planet_src, planet.c:152-196 */
static void planet_src(const or_ctx *o, const double *pl, const double *prim,
                       double *cons, const double *xp, const double *xm, double dVdt)
{
   double rho = prim[RHO];
   double vr = prim[URR];
   double vz = prim[UZZ];
   double omega = prim[UPP];
   double dphi = g_dp(o, xp[1], xm[1]);
   double phi = xm[1] + 0.5 * dphi;
   double x[3] = {g_centroid(xp[0], xm[0], 1), phi, g_centroid(xp[2], xm[2], 2)};
   double F[3];
   planetary_force(o, pl, x[0], x[1], x[2], &F[0], &F[1], &F[2]);
   double hr = 1.0, hp = x[0], hz = 1.0;
   if (o->polar_sources) {
      double adj = g_polar_adjust(o, xp, xm);
      F[0] *= adj;
      F[1] *= adj;
   }
   cons[SRR] += rho * hr * F[0] * dVdt;
   cons[LLL] += rho * hp * F[1] * dVdt;
   cons[SZZ] += rho * hz * F[2] * dVdt;
   cons[TAU] += rho * (hr * F[0] * vr + hz * F[2] * vz + hp * F[1] * omega) * dVdt;
}

/* add_source, misc.c:538-580 (omega_src and sink_src are no-ops unless
 * RotFrame / Nozzle_Type are set; RotFrame is rejected at creation). */
static void add_source(or_ctx *o, double dt)
{
   long nt = o->ntot;
   int s = o->cur_substep;
   for (int k = o->NgZa; k < o->Nz - o->NgZb; ++k)
      for (int j = o->NgRa; j < o->Nr - o->NgRb; ++j) {
         int jk = j + o->Nr * k;
         for (int i = 0; i < o->Np[jk]; ++i) {
            long c = o->off[jk] + i;
            double phip = o->piph[c];
            double phim = o->piph[c] - o->dphi[c];
            double xp[3] = {o->r_jph[j], phip, o->z_kph[k]};
            double xm[3] = {o->r_jph[j - 1], phim, o->z_kph[k - 1]};
            double dV = g_dV(o, xp, xm);
            double prim[MAXQ], cons[MAXQ];
            gather(o, o->prim, c, prim);
            gather(o, o->cons, c, cons);
            if (o->p.hydro == DH_HYDRO_MHD) m_source(o, prim, cons, xp, xm, dV * dt);
            else e_source(o, prim, cons, xp, xm, dV * dt);
            for (int p = 0; p < o->npl[s]; ++p)
               planet_src(o, o->planets[s] + p * DH_PLANET_STRIDE, prim, cons, xp, xm, dV * dt);
            if (o->p.visc_flag) {
               double gr[MAXQ], gp[MAXQ], gz[MAXQ];
               gather(o, o->gradr, c, gr);
               gather(o, o->gradp, c, gp);
               gather(o, o->gradz, c, gz);
               e_visc_source(o, prim, gr, gp, gz, cons, xp, xm, dV * dt);
            }
            scatter(o, o->cons, c, cons);
         }
      }
   (void)nt;
}

/* move_cells, misc.c:256-268 */
static void move_cells(or_ctx *o, double dt)
{
   for (long c = 0; c < o->ntot; ++c) o->piph[c] += o->wiph[c] * dt;
}

/* clean_pi, misc.c:8-36 (cells only; planets live on the host) */
static void clean_pi(or_ctx *o)
{
   double pm = o->phi_max;
   for (long c = 0; c < o->ntot; ++c) {
      double phi = o->piph[c];
      while (phi > pm) phi -= pm;
      while (phi < 0.0) phi += pm;
      o->piph[c] = phi;
   }
}

/* calc_dp, misc.c:271-288 */
static void calc_dp(or_ctx *o)
{
   for (int jk = 0; jk < o->Nr * o->Nz; ++jk) {
      long c0 = o->off[jk];
      for (int i = 0; i < o->Np[jk]; ++i) {
         int im = i - 1;
         if (i == 0) im = o->Np[jk] - 1;
         o->dphi[c0 + i] = g_dp(o, o->piph[c0 + i], o->piph[c0 + im]);
      }
   }
}

/* calc_prim, misc.c:290-325 */
static void calc_prim(or_ctx *o)
{
   for (int k = o->NgZa; k < o->Nz - o->NgZb; ++k) {
      double zm = o->z_kph[k - 1], zp = o->z_kph[k];
      double z = g_centroid(zp, zm, 2);
      for (int j = o->NgRa; j < o->Nr - o->NgRb; ++j) {
         int jk = j + o->Nr * k;
         double rm = o->r_jph[j - 1], rp = o->r_jph[j];
         double r = g_centroid(rp, rm, 1);
         for (int i = 0; i < o->Np[jk]; ++i) {
            long c = o->off[jk] + i;
            double phip = o->piph[c];
            double phim = phip - o->dphi[c];
            double xp[3] = {rp, phip, zp};
            double xm[3] = {rm, phim, zm};
            double dV = g_dV(o, xp, xm);
            double x[3] = {r, 0.5 * (phim + phip), z};
            double cons[MAXQ], prim[MAXQ];
            gather(o, o->cons, c, cons);
            if (o->p.hydro == DH_HYDRO_MHD) m_cons2prim(o, cons, prim, x, dV);
            else e_cons2prim(o, cons, prim, x, dV);
            scatter(o, o->prim, c, prim);
         }
      }
   }
}

/* set_cells_copy_distant, boundary_functions.c:90-159: area-weighted copy of
 * annulus src onto annulus dst. */
static void copy_distant(or_ctx *o, int dst, int src)
{
   long nt = o->ntot;
   long c0 = o->off[dst], s0 = o->off[src];
   int Np = o->Np[dst], Np1 = o->Np[src];
   double pm = o->phi_max;
   for (int i = 0; i < Np; ++i)
      for (int q = 0; q < o->nq; ++q) o->prim[q * nt + c0 + i] = 0.0;
   int i1;
   double phi0 = o->piph[c0] - o->dphi[c0];
   for (i1 = 0; i1 < Np1; i1++) {
      double dphip = g_sdp(o, o->piph[s0 + i1], phi0);
      double dphim = dphip - o->dphi[s0 + i1];
      if (dphip > 0 && dphim <= 0.0) break;
   }
   if (i1 >= Np1) i1 = 0;
   for (int i = 0; i < Np; ++i) {
      double phip = o->piph[c0 + i];
      double phim = o->piph[c0 + i] - o->dphi[c0 + i];
      double phip1, phim1;
      do {
         phip1 = o->piph[s0 + i1];
         phim1 = o->piph[s0 + i1] - o->dphi[s0 + i1];
         while (phip1 < phim) { phip1 += pm; phim1 += pm; }
         while (phim1 > phip) { phip1 -= pm; phim1 -= pm; }
         double phi1 = phip < phip1 ? phip : phip1;
         double phi2 = phim > phim1 ? phim : phim1;
         double dphi = phi1 - phi2;
         for (int q = 0; q < o->nq; ++q)
            o->prim[q * nt + c0 + i] += o->prim[q * nt + s0 + i1] * dphi;
         i1 = i1 < Np1 - 1 ? i1 + 1 : 0;
      } while (phip1 < phip);
      i1 = i1 > 0 ? i1 - 1 : Np1 - 1;
   }
   for (int i = 0; i < Np; ++i)
      for (int q = 0; q < o->nq; ++q) o->prim[q * nt + c0 + i] /= o->dphi[c0 + i];
}

static void diode(or_ctx *o, int jk, int q, double sgn)
{
   long nt = o->ntot;
   for (int i = 0; i < o->Np[jk]; ++i) {
      double *v = &o->prim[q * nt + o->off[jk] + i];
      if (sgn * *v > 0) *v = 0.0;
   }
}

/* boundary_trans for the supported modules: held (fixed/polar with
 * phi-independent initial data: the reference rewrites the same values),
 * zero-gradient and diode (boundary_functions.c:269-441). */
static void boundaries(or_ctx *o, int dim)
{
   int Nr = o->Nr, Nz = o->Nz;
   if (dim == 1) {
      int b = o->p.bc_rmin;
      if (b == DH_BC_ZEROGRAD || b == DH_BC_DIODE)
         for (int k = 0; k < Nz; ++k)
            for (int j = o->NgRa - 1; j >= 0; --j) {
               copy_distant(o, j + Nr * k, j + 1 + Nr * k);
               if (b == DH_BC_DIODE) diode(o, j + Nr * k, URR, +1.0);
            }
      b = o->p.bc_rmax;
      if (b == DH_BC_ZEROGRAD || b == DH_BC_DIODE)
         for (int k = 0; k < Nz; ++k)
            for (int j = Nr - o->NgRb; j < Nr; ++j) {
               copy_distant(o, j + Nr * k, j - 1 + Nr * k);
               if (b == DH_BC_DIODE) diode(o, j + Nr * k, URR, -1.0);
            }
   } else {
      int b = o->p.bc_zmin;
      if (b == DH_BC_ZEROGRAD || b == DH_BC_DIODE)
         for (int j = 0; j < Nr; ++j)
            for (int k = o->NgZa - 1; k >= 0; --k) {
               copy_distant(o, j + Nr * k, j + Nr * (k + 1));
               if (b == DH_BC_DIODE) diode(o, j + Nr * k, UZZ, +1.0);
            }
      b = o->p.bc_zmax;
      if (b == DH_BC_ZEROGRAD || b == DH_BC_DIODE)
         for (int j = 0; j < Nr; ++j)
            for (int k = Nz - o->NgZb; k < Nz; ++k) {
               copy_distant(o, j + Nr * k, j + Nr * (k - 1));
               if (b == DH_BC_DIODE) diode(o, j + Nr * k, UZZ, -1.0);
            }
   }
}

/* ------------------------------------------------------------------------ */
/* bfields.c: constrained transport on the moving mesh (CT_MODE 2)          */
/* ------------------------------------------------------------------------ */

#define EC(e, c) o->Ecell[(long)(e) * o->ntot + (c)]
#define BC(e, c) o->Bcell[(long)(e) * o->ntot + (c)]
#define PHI(f, c) o->Phi[(long)(f) * o->ntot + (c)]
#define RKPHI(f, c) o->rkPhi[(long)(f) * o->ntot + (c)]

/* avg_Efields, bfields.c:392-644 */
static void avg_Efields(or_ctx *o)
{
   int Nr = o->Nr, Nz = o->Nz;
   for (int jk = 0; jk < Nr * Nz; ++jk) {
      long c0 = o->off[jk];
      int np = o->Np[jk];
      for (int i = 0; i < np; ++i) {
         long c = c0 + i, cp = c0 + (i + 1) % np;
         for (int s = 0; s < 2; ++s) {
            int e0 = 4 * s, e1 = 4 * s + 1, e2 = 4 * s + 2, e3 = 4 * s + 3;
            double El = .5 * (EC(e0, c) + EC(e2, cp));
            double Er = .5 * (EC(e1, c) + EC(e3, cp));
            EC(e0, c) = El; EC(e1, c) = Er; EC(e2, cp) = El; EC(e3, cp) = Er;
            double Bl = .5 * (BC(e0, c) + BC(e2, cp));
            double Br = .5 * (BC(e1, c) + BC(e3, cp));
            BC(e0, c) = Bl; BC(e1, c) = Br; BC(e2, cp) = Bl; BC(e3, cp) = Br;
         }
      }
   }
   for (int d = 0; d < 2; ++d) {
      int dim = d + 1;
      if ((dim == 1 && Nr <= 1) || (dim == 2 && Nz <= 1)) continue;
      or_face *F = o->faces[d];
      long Nf = o->fI[d][o->nftracks[d]];
      /* L's outer edge slot, R's inner edge slot, and R's/L's corner pairs */
      int eL = dim == 1 ? 1 : 5, eR = dim == 1 ? 0 : 4;
      for (long n = 0; n < Nf; ++n) {
         or_face *f = F + n;
         long c1 = f->LRtype == 0 ? f->L : f->R;
         long c2 = f->LRtype == 0 ? f->R : f->L;
         double dp1 = g_dp(o, o->piph[c2], o->piph[c1]);
         double dp2 = o->dphi[c2] - dp1;
         if (f->LRtype == 0) {
            double Eavg = (dp2 * EC(eR, c2) + dp1 * EC(eR + 2, c2)) / (dp1 + dp2);
            double Bavg = (dp2 * BC(eR, c2) + dp1 * BC(eR + 2, c2)) / (dp1 + dp2);
            f->E = .5 * (EC(eL, f->L) + Eavg);
            f->B = .5 * (BC(eL, f->L) + Bavg);
         } else {
            double Eavg = (dp2 * EC(eL, c2) + dp1 * EC(eL + 2, c2)) / (dp1 + dp2);
            double Bavg = (dp2 * BC(eL, c2) + dp1 * BC(eL + 2, c2)) / (dp1 + dp2);
            f->E = .5 * (EC(eR, f->R) + Eavg);
            f->B = .5 * (BC(eR, f->R) + Bavg);
         }
      }
      for (long n = 0; n < Nf; ++n) {
         or_face *f = F + n;
         if (f->LRtype == 0) { EC(eL, f->L) = f->E; BC(eL, f->L) = f->B; }
         else { EC(eR, f->R) = f->E; BC(eR, f->R) = f->B; }
         f->E = 0.0;
         f->B = 0.0;
      }
   }
   /* E along the r = 0 pole and the r/z coordinate boundaries */
   if (o->NgRa == 0) {
      for (int k = 0; k < Nz; ++k) {
         int jk = Nr * k;
         double E = 0.0;
         for (int i = 0; i < o->Np[jk]; ++i) E += EC(1, o->off[jk] + i);
         E /= o->Np[jk];
         for (int i = 0; i < o->Np[jk]; ++i) { EC(0, o->off[jk] + i) = E; BC(0, o->off[jk] + i) = 0.0; }
      }
   }
   if (o->NgRb == 0) {
      for (int k = 0; k < Nz; ++k) {
         int jk = Nr - 1 + Nr * k;
         double E = 0.0;
         for (int i = 0; i < o->Np[jk]; ++i) E += EC(0, o->off[jk] + i);
         E /= o->Np[jk];
         for (int i = 0; i < o->Np[jk]; ++i) { EC(1, o->off[jk] + i) = E; BC(1, o->off[jk] + i) = 0.0; }
      }
   }
   if (o->NgZa == 0) {
      for (int j = 0; j < Nr; ++j) {
         int jk = j;
         double E = 0.0;
         for (int i = 0; i < o->Np[jk]; ++i) E += EC(5, o->off[jk] + i);
         E /= o->Np[jk];
         for (int i = 0; i < o->Np[jk]; ++i) { EC(5, o->off[jk] + i) = E; BC(5, o->off[jk] + i) = 0.0; }
      }
   }
   if (o->NgZb == 0) {
      for (int j = 0; j < Nr; ++j) {
         int jk = j + Nr * (Nz - 1);
         double E = 0.0;
         for (int i = 0; i < o->Np[jk]; ++i) E += EC(4, o->off[jk] + i);
         E /= o->Np[jk];
         for (int i = 0; i < o->Np[jk]; ++i) { EC(4, o->off[jk] + i) = E; BC(4, o->off[jk] + i) = 0.0; }
      }
   }
   for (int jk = 0; jk < Nr * Nz; ++jk) {
      long c0 = o->off[jk];
      int np = o->Np[jk];
      for (int i = 0; i < np; ++i) {
         long c = c0 + i, cp = c0 + (i + 1) % np;
         EC(2, cp) = EC(0, c); EC(3, cp) = EC(1, c);
         BC(2, cp) = BC(0, c); BC(3, cp) = BC(1, c);
         EC(6, cp) = EC(4, c); EC(7, cp) = EC(5, c);
         BC(6, cp) = BC(4, c); BC(7, cp) = BC(5, c);
      }
   }
}

/* subtract_advective_B_fluxes, bfields.c:646-690: E -= (w x B) of the
 * sliding mesh on the phi-moving edges. */
static void subtract_advective(or_ctx *o)
{
   for (int k = 0; k < o->Nz; ++k)
      for (int j = 0; j < o->Nr; ++j) {
         int jk = j + o->Nr * k;
         double xc0 = g_centroid(o->r_jph[j], o->r_jph[j - 1], 1);
         double hL = o->r_jph[j - 1], hR = o->r_jph[j], hD = xc0, hU = xc0;
         for (int i = 0; i < o->Np[jk]; ++i) {
            long c = o->off[jk] + i;
            double wm = o->wiph[c], wp = o->wiph[c];
            EC(0, c) -= hL * wm * BC(0, c);
            EC(1, c) -= hR * wp * BC(1, c);
            EC(4, c) += hD * wm * BC(4, c);
            EC(5, c) += hU * wp * BC(5, c);
         }
      }
}

/* Phi slot of a face's owner (get_phi_pointer, bfields.c:716-740) */
static long phi_slot(const or_ctx *o, const or_face *f, int d)
{
   long nt = o->ntot;
   if (d == 0) return f->LRtype == 0 ? 2 * nt + f->L : 1 * nt + f->R;
   return f->LRtype == 0 ? 4 * nt + f->L : 3 * nt + f->R;
}

/* get_which4 / phi_switch, bfields.c:817-886 */
static int phi_switch(double dphi, double Pmax, int mode)
{
   while (dphi > .5 * Pmax) dphi -= Pmax;
   while (dphi < -.5 * Pmax) dphi += Pmax;
   if (mode == 1) dphi = -dphi;
   return dphi > 0. ? 1 : 0;
}

static int get_which4(double phi, double phiR, double phiU, double phiUR, int *LR_alt,
                      int *UD_alt, int mode, double Pmax)
{
   int LR_D = phi_switch(phi - phiR, Pmax, mode);
   int LR_U = phi_switch(phiU - phiUR, Pmax, mode);
   double phi1 = LR_D ? phiR : phi;
   double phi2 = LR_U ? phiUR : phiU;
   int UD = phi_switch(phi1 - phi2, Pmax, mode);
   int which4;
   if (UD == 0) which4 = LR_D == 0 ? 0 : 1;
   else which4 = LR_U == 0 ? 2 : 3;
   if (mode == 0) {
      *LR_alt = (which4 == 0 || which4 == 1) ? LR_U : LR_D;
      if (which4 == 0 || which4 == 2) *UD_alt = phi_switch(phiR - phiUR, Pmax, mode);
      else *UD_alt = phi_switch(phi - phiU, Pmax, mode);
   }
   return which4;
}

/* make_edge_adjust, bfields.c:888-1127: E_phi on the phi-directed edges where
 * four annuli meet, walked as a 4-way merge of the rings. */
static void make_edge_adjust(or_ctx *o, double dt)
{
   int Nr = o->Nr, Nz = o->Nz;
   double Pmax = o->phi_max;
   long nt = o->ntot;
   int *I0 = (int *)xcalloc(Nr * Nz, sizeof(int));
   ring_starts(o, I0);
   for (int k = 0; k < Nz - 1; ++k)
      for (int j = 0; j < Nr - 1; ++j) {
         int jk = j + Nr * k, jkR = j + 1 + Nr * k, jkU = j + Nr * (k + 1), jkUR = j + 1 + Nr * (k + 1);
         double xp[3] = {o->r_jph[j], 0.0, o->z_kph[k]};
         double xm[3] = {o->r_jph[j], 0.0, o->z_kph[k]};
         int i = I0[jk], iR = I0[jkR], iU = I0[jkU], iUR = I0[jkUR];
         int Ne = o->Np[jk] + o->Np[jkR] + o->Np[jkU] + o->Np[jkUR];
         for (int e = 0; e < Ne; ++e) {
            long c = o->off[jk] + i, cR = o->off[jkR] + iR;
            long cU = o->off[jkU] + iU, cUR = o->off[jkUR] + iUR;
            int LR_alt, UD_alt;
            int w4 = get_which4(o->piph[c], o->piph[cR], o->piph[cU], o->piph[cUR], &LR_alt,
                                &UD_alt, 0, Pmax);
            int w4b = get_which4(o->piph[c] - o->dphi[c], o->piph[cR] - o->dphi[cR],
                                 o->piph[cU] - o->dphi[cU], o->piph[cUR] - o->dphi[cUR],
                                 NULL, NULL, 1, Pmax);
            long PL, PR, PU, PD;
            double E;
            double *EP = o->Ephi;
            if (w4 == 0) {
               PL = 4 * nt + c; PD = 2 * nt + c;
               E = .25 * (EP[3 * nt + c] + EP[1 * nt + c]);
               if (LR_alt == 0) { PU = 2 * nt + cU; E += .25 * EP[1 * nt + cU]; }
               else { PU = 1 * nt + cUR; E += .25 * EP[0 * nt + cUR]; }
               if (UD_alt == 0) { PR = 4 * nt + cR; E += .25 * EP[3 * nt + cR]; }
               else { PR = 3 * nt + cUR; E += .25 * EP[2 * nt + cUR]; }
            } else if (w4 == 1) {
               PR = 4 * nt + cR; PD = 1 * nt + cR;
               E = .25 * (EP[3 * nt + cR] + EP[0 * nt + cR]);
               if (LR_alt == 0) { PU = 2 * nt + cU; E += .25 * EP[1 * nt + cU]; }
               else { PU = 1 * nt + cUR; E += .25 * EP[0 * nt + cUR]; }
               if (UD_alt == 0) { PL = 4 * nt + c; E += .25 * EP[3 * nt + c]; }
               else { PL = 3 * nt + cU; E += .25 * EP[2 * nt + cU]; }
            } else if (w4 == 2) {
               PL = 3 * nt + cU; PU = 2 * nt + cU;
               E = .25 * (EP[2 * nt + cU] + EP[1 * nt + cU]);
               if (LR_alt == 0) { PD = 2 * nt + c; E += .25 * EP[1 * nt + c]; }
               else { PD = 1 * nt + cR; E += .25 * EP[0 * nt + cR]; }
               if (UD_alt == 0) { PR = 4 * nt + cR; E += .25 * EP[3 * nt + cR]; }
               else { PR = 3 * nt + cUR; E += .25 * EP[2 * nt + cUR]; }
            } else {
               PR = 3 * nt + cUR; PU = 1 * nt + cUR;
               E = .25 * (EP[2 * nt + cUR] + EP[0 * nt + cUR]);
               if (LR_alt == 0) { PD = 2 * nt + c; E += .25 * EP[1 * nt + c]; }
               else { PD = 1 * nt + cR; E += .25 * EP[0 * nt + cR]; }
               if (UD_alt == 0) { PL = 4 * nt + c; E += .25 * EP[3 * nt + c]; }
               else { PL = 3 * nt + cU; E += .25 * EP[2 * nt + cU]; }
            }
            long cs[4] = {c, cR, cU, cUR};
            xp[1] = o->piph[cs[w4]];
            xm[1] = o->piph[cs[w4b]] - o->dphi[cs[w4b]];
            double dl = g_dL(o, xp, xm, 0);
            double Edldt = E * dl * dt;
            /* add_E_phi, bfields.c:384-389 */
            o->Phi[PL] -= Edldt;
            o->Phi[PR] += Edldt;
            o->Phi[PU] -= Edldt;
            o->Phi[PD] += Edldt;
            if (w4 == 0) { if (++i == o->Np[jk]) i = 0; }
            else if (w4 == 1) { if (++iR == o->Np[jkR]) iR = 0; }
            else if (w4 == 2) { if (++iU == o->Np[jkU]) iU = 0; }
            else { if (++iUR == o->Np[jkUR]) iUR = 0; }
         }
      }
   free(I0);
}

/* update_B_fluxes, bfields.c:313-382.  The track loop visits track jk-1 at
 * iteration jk (fIndex holds start offsets), so the last track of each
 * direction is not visited; that track lies in a ghost corner. */
static void update_B_fluxes(or_ctx *o, double dt)
{
   for (int d = 0; d < 2; ++d) {
      int dim = d + 1;
      if ((dim == 1 && o->Nr <= 1) || (dim == 2 && o->Nz <= 1)) continue;
      or_face *F = o->faces[d];
      long *Nf = o->fI[d];
      int Njk = o->nftracks[d];
      long n0 = 0;
      double sgn = dim == 1 ? 1.0 : -1.0;
      int eL = dim == 1 ? 1 : 5, eR = dim == 1 ? 0 : 4;
      for (int jk = 0; jk < Njk; ++jk) {
         for (long n = 0; n < Nf[jk] - n0; ++n) {
            or_face *f = F + n0 + n;
            long np = n + 1;
            if (np == Nf[jk] - n0) np = 0;
            or_face *fp = F + n0 + np;
            double dl = f->dl;
            double E;
            if (f->LRtype == 0) {
               E = EC(eL, f->L);
               o->Phi[phi_slot(o, f, d)] -= sgn * E * dl * dt;
               PHI(0, f->L) += sgn * E * dl * dt;
            } else {
               E = EC(eR, f->R);
               o->Phi[phi_slot(o, f, d)] -= sgn * E * dl * dt;
               PHI(0, f->R) -= sgn * E * dl * dt;
            }
            o->Phi[phi_slot(o, fp, d)] += sgn * E * dl * dt;
         }
         n0 = Nf[jk];
      }
   }
   if (o->Nz > 1) make_edge_adjust(o, dt);
}

/* check_flipped + flip_fluxes, bfields.c:693-814 (same track loop) */
static void flips(or_ctx *o, int d)
{
   or_face *F = o->faces[d];
   long *fI = o->fI[d];
   int Nt = o->nftracks[d];
   long n0 = 0;
   for (int jk = 0; jk < Nt; ++jk) {
      for (long n = 0; n < fI[jk] - n0; ++n) {
         or_face *f = F + n0 + n;
         long np = n + 1;
         if (np == fI[jk] - n0) np = 0;
         or_face *fp = F + n0 + np;
         double pm = f->LRtype == 0 ? o->piph[f->L] : o->piph[f->R];
         double pp = fp->LRtype == 0 ? o->piph[fp->L] : o->piph[fp->R];
         if (g_sdp(o, pp, pm) < 0.) fp->flip_flag = 1;
      }
      n0 = fI[jk];
   }
   n0 = 0;
   for (int jk = 0; jk < Nt; ++jk) {
      for (long n = 0; n < fI[jk] - n0; ++n) {
         or_face *f = F + n0 + n;
         if (!f->flip_flag) continue;
         long np = n + 1;
         if (np == fI[jk] - n0) np = 0;
         long nm = n - 1;
         if (nm == -1) nm = fI[jk] - n0 - 1;
         or_face *fp = F + n0 + np, *fm = F + n0 + nm;
         long P = phi_slot(o, f, d), Pm = phi_slot(o, fm, d), Pp = phi_slot(o, fp, d);
         double Phi = o->Phi[P], RK_Phi = o->rkPhi[P];
         o->Phi[P] = o->Phi[Pm] + Phi;
         o->rkPhi[P] = o->rkPhi[Pm] + RK_Phi;
         o->Phi[Pm] = -Phi;
         o->rkPhi[Pm] = -RK_Phi;
         o->Phi[Pp] = o->Phi[Pp] + Phi;
         o->rkPhi[Pp] = o->rkPhi[Pp] + RK_Phi;
         f->LRtype = !f->LRtype;
         fm->LRtype = !fm->LRtype;
      }
      n0 = fI[jk];
   }
}

/* B_faces_to_cells(type = 1), bfields.c:88-309: conserved B from the face
 * fluxes (area-weighted over each cell's r and z faces; B_phi from the two
 * phi faces). */
static void B_faces_to_cells(or_ctx *o)
{
   long nt = o->ntot;
   int Nr = o->Nr, Nz = o->Nz;
   for (long c = 0; c < nt; ++c) {
      o->temp[c] = 0.0;
      CONS(o, BRR, c) = 0.0;
      CONS(o, BPP, c) = 0.0;
      CONS(o, BZZ, c) = 0.0;
   }
   if (Nr > 1) {
      long Nf = o->fI[0][o->nftracks[0]];
      for (long n = 0; n < Nf; ++n) {
         or_face *f = o->faces[0] + n;
         double Phi = f->LRtype == 0 ? PHI(2, f->L) : PHI(1, f->R);
         o->temp[f->L] += f->dA;
         o->temp[f->R] += f->dA;
         CONS(o, BRR, f->L) += Phi;
         CONS(o, BRR, f->R) += Phi;
      }
   }
   for (int j = 0; j < Nr; ++j)
      for (int k = 0; k < Nz; ++k) {
         int jk = j + Nr * k;
         for (int i = 0; i < o->Np[jk]; ++i) {
            int im = i == 0 ? o->Np[jk] - 1 : i - 1;
            long c = o->off[jk] + i, cm = o->off[jk] + im;
            double xp[3] = {o->r_jph[j], o->piph[c], o->z_kph[k]};
            double xm[3] = {o->r_jph[j - 1], o->piph[cm], o->z_kph[k - 1]};
            double x0 = g_centroid(xp[0], xm[0], 1);
            double dA = g_dA(o, xp, xm, 0);
            double dV = g_dV(o, xp, xm);
            double rfac = m_bscale(x0, 0), hr = 1.0, hp = x0;
            CONS(o, BRR, c) *= dV * rfac / (o->temp[c] * hr);
            CONS(o, BPP, c) = .5 * (PHI(0, c) + PHI(0, cm)) * dV / (dA * hp);
         }
      }
   if (Nz > 1) {
      for (long c = 0; c < nt; ++c) o->temp[c] = 0.0;
      long Nf = o->fI[1][o->nftracks[1]];
      for (long n = 0; n < Nf; ++n) {
         or_face *f = o->faces[1] + n;
         double Phi = f->LRtype == 0 ? PHI(4, f->L) : PHI(3, f->R);
         o->temp[f->L] += f->dA;
         o->temp[f->R] += f->dA;
         CONS(o, BZZ, f->L) += Phi;
         CONS(o, BZZ, f->R) += Phi;
      }
      for (int j = 0; j < Nr; ++j)
         for (int k = 0; k < Nz; ++k) {
            int jk = j + Nr * k;
            for (int i = 0; i < o->Np[jk]; ++i) {
               int im = i == 0 ? o->Np[jk] - 1 : i - 1;
               long c = o->off[jk] + i, cm = o->off[jk] + im;
               double xp[3] = {o->r_jph[j], o->piph[c], o->z_kph[k]};
               double xm[3] = {o->r_jph[j - 1], o->piph[cm], o->z_kph[k - 1]};
               double dV = g_dV(o, xp, xm);
               double zfac = 1.0, hz = 1.0;
               CONS(o, BZZ, c) *= dV * zfac / (o->temp[c] * hz);
            }
         }
   }
}

/* onestep, onestep.c:84-228 */
static void onestep(or_ctx *o, double RK, double dt, int first_step)
{
   if (first_step) set_wcell(o);
   adjust_RK_cons(o, RK);
   plm_phi(o);
   if (o->Nr > 1) { build_faces(o, 1); plm_trans(o, 1); }
   if (o->Nz > 1) { build_faces(o, 2); plm_trans(o, 2); }
   phi_flux(o, dt);
   if (o->Nr > 1) trans_flux(o, dt, 1);
   if (o->Nz > 1) trans_flux(o, dt, 2);
   int mhd = o->p.hydro == DH_HYDRO_MHD;
   if (mhd) {
      avg_Efields(o);
      subtract_advective(o);
      update_B_fluxes(o, dt);
   }
   add_source(o, dt);
   if (first_step) {
      move_cells(o, dt);
      if (mhd) {
         if (o->Nr > 1) flips(o, 0);
         if (o->Nz > 1) flips(o, 1);
      }
   }
   clean_pi(o);
   calc_dp(o);
   if (mhd && o->p.ct) B_faces_to_cells(o);
   calc_prim(o);
   if (!o->p.r_periodic) boundaries(o, 1);
   if (o->Nz > 1 && !o->p.z_periodic) boundaries(o, 2);
}

/* timestep, timestep.c:7-45 */
void or_timestep(or_ctx *o, double dt)
{
   memcpy(o->rkcons, o->cons, o->ntot * o->nq * sizeof(double));
   if (o->nf) memcpy(o->rkPhi, o->Phi, o->ntot * o->nf * sizeof(double));
   o->cur_substep = 0;
   if (o->p.timestep_mode == 1) {
      onestep(o, 0.0, dt, 1);
   } else {
      onestep(o, 0.0, dt, 1);
      o->cur_substep = 1;
      onestep(o, 0.5, 0.5 * dt, 0);
   }
}

/* ------------------------------------------------------------------------ */
/* Stage-level entry points                                                  */
/* ------------------------------------------------------------------------ */

void or_plm_phi(or_ctx *o, double *gradp_out)
{
   plm_phi(o);
   if (gradp_out) memcpy(gradp_out, o->gradp, o->ntot * o->nq * sizeof(double));
}

long or_faces(or_ctx *o, int dim, int *L, int *R, int *LRtype, double *dA,
              double *phic, double *dphi)
{
   build_faces(o, dim);
   int d = dim - 1;
   long nf = o->fI[d][o->nftracks[d]];
   for (long n = 0; n < nf; ++n) {
      or_face *f = o->faces[d] + n;
      if (L) L[n] = (int)f->L;
      if (R) R[n] = (int)f->R;
      if (LRtype) LRtype[n] = f->LRtype;
      if (dA) dA[n] = f->dA;
      if (phic) phic[n] = f->cm[1];
      if (dphi) dphi[n] = f->dphi;
   }
   return nf;
}

void or_plm_trans(or_ctx *o, int dim, double *grad_out)
{
   plm_phi(o);
   build_faces(o, dim);
   plm_trans(o, dim);
   if (grad_out)
      memcpy(grad_out, dim == 1 ? o->gradr : o->gradz, o->ntot * o->nq * sizeof(double));
}

void or_riemann_flux(or_ctx *o, const double *primL, const double *primR,
                     const double *x, const double *n, double w, double *F)
{
   double z[MAXQ] = {0};
   solve_riemann(o, primL, primR, z, z, z, z, z, z, x, n, w, F, 0, NULL);
}

void or_prim2cons(or_ctx *o, const double *prim, double *cons, const double *x, double dV)
{
   if (o->p.hydro == DH_HYDRO_MHD) m_prim2cons(o, prim, cons, x, dV);
   else e_prim2cons(o, prim, cons, x, dV);
}

void or_cons2prim(or_ctx *o, const double *cons, double *prim, const double *x, double dV)
{
   if (o->p.hydro == DH_HYDRO_MHD) m_cons2prim(o, cons, prim, x, dV);
   else e_cons2prim(o, cons, prim, x, dV);
}
