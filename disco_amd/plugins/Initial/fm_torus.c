/*
 * Initial plugin (reference ABI, setICparams / initial): the
 * Fishbone-Moncrief torus (ApJ 207, 962, 1976) around a Kerr black hole of
 * mass M = Metric_Par2 and spin a = Metric_Par3, in the primitives of the
 * GRMHD module (covariant u_i, densitised field) on the Kerr-Schild
 * cylindrical coordinates of the kerr_ks_rpz Metric plugin.
 *
 * The torus: gas on circular orbits (u^r = u^theta = 0) with constant
 * ell = u^t u_phi, isentropic with P = K rho^gamma (gamma = Adiabatic_Index).
 * For such flow the Euler equation integrates to
 *     ln h = ln u^t - ell Omega + const,      Omega = u^phi / u^t,
 * h = 1 + gamma P / ((gamma - 1) rho).  In the frame of the zero angular
 * momentum observer (lapse alpha = sqrt(Sigma Delta / A), omega = 2 M a r / A,
 * g_phiphi = A sin^2(theta) / Sigma in Boyer-Lindquist coordinates) the
 * orbital four-velocity component U = u^(phi) solves
 *     U sqrt(1 + U^2) = ell alpha / sqrt(g_phiphi),
 * so U^2 = (sqrt(1 + 4 k^2) - 1) / 2 with k = ell alpha / sqrt(g_phiphi),
 * and u^t = sqrt(1 + U^2) / alpha, u^phi = omega u^t + U / sqrt(g_phiphi).
 * ell is that of the prograde circular geodesic at the pressure maximum
 * r_max = Init_Par2 (where the pressure gradient vanishes), the constant
 * puts h = 1 at the inner edge r_in = Init_Par1 on the equator, and K
 * makes the density 1 at the maximum.  The torus is where h > 1, r >= r_in
 * and its density exceeds the atmosphere's; elsewhere an atmosphere at rest
 * in the normal frame (u_i = 0) with density Init_Par5 and pressure
 * Init_Par6.  Since u^r = 0 the
 * Boyer-Lindquist u^t and u^phi are also the Kerr-Schild ones (the
 * coordinates differ by functions of r only), and the plugin lowers them
 * with the Kerr-Schild metric.
 *
 * Field: poloidal loops from the potential A_phi = A0 max(rho - Init_Par4, 0)
 * (Init_Par4 = 0.2 when 0), normalised so that the torus's largest
 * magnetic pressure b^2 / 2 is its largest gas pressure over beta =
 * Init_Par3 (100 when 0), plus Init_Par7 R^2 / 2, a uniform vertical field
 * of primitive B^z = Init_Par7 everywhere.  The field a face carries is set from A_phi's
 * line integrals (Stokes): an r face at R spanning [z-, z+] of its cell gets
 * B^r = -(A(R, z+) - A(R, z-)) / (R dz), a z face spanning [R-, R+] gets
 * B^z = (A(R+, z) - A(R-, z)) / (R_mid dR), with the cell edges of the
 * uniform grid (LogZoning 0) found from the face's centroid.  The fluxes of
 * every cell's faces then cancel to round-off: the initial field is
 * divergence-free, which constrained transport keeps.
 */
#include <math.h>

#include "paul.h"
#include "kerr_ks.h"

static double fm_M = 1.0, fm_a = 0.0, fm_gam = 4.0 / 3.0;
static double fm_rin = 6.0, fm_rmax = 12.0, fm_beta = 100.0, fm_cut = 0.2;
static double fm_rho_atm = 1e-4, fm_P_atm = 1e-6;
static double fm_ell, fm_lnq_in, fm_K, fm_A0 = 0.0, fm_Bv = 0.0;
static double fm_Rmin, fm_dR, fm_Zmin, fm_dz;

/* Boyer-Lindquist quantities of the torus flow at (r, sin^2 theta):
 * u^t, u^phi and ln(u^t) - ell Omega. */
static void fm_flow(double r, double s2, double *ut, double *uph, double *lnq)
{
   double a = fm_a, M = fm_M;
   double c2 = 1.0 - s2;
   double Sig = r * r + a * a * c2;
   double Del = r * r - 2.0 * M * r + a * a;
   double A = (r * r + a * a) * (r * r + a * a) - a * a * Del * s2;
   double al = sqrt(Sig * Del / A);
   double om = 2.0 * M * a * r / A;
   double gpp = A * s2 / Sig;
   double k = fm_ell * al / sqrt(gpp);
   double U2 = 0.5 * (sqrt(1.0 + 4.0 * k * k) - 1.0);
   *ut = sqrt(1.0 + U2) / al;
   *uph = om * (*ut) + sqrt(U2) / sqrt(gpp);
   *lnq = log(*ut) - fm_ell * (*uph) / (*ut);
}

/* Density and pressure of the torus at (R, z); 0 where it is outside. */
static int fm_gas(double R, double z, double *rho, double *P, double *ut, double *uph)
{
   double r = kks_radius(fm_a, R, z);
   if (r < fm_rin) return 0;
   double s2 = R * R / (r * r + fm_a * fm_a);
   double lnq;
   fm_flow(r, s2, ut, uph, &lnq);
   double h = exp(lnq - fm_lnq_in);
   if (!(h > 1.0)) return 0;
   *rho = pow((h - 1.0) * (fm_gam - 1.0) / (fm_K * fm_gam), 1.0 / (fm_gam - 1.0));
   *P = fm_K * pow(*rho, fm_gam);
   return *rho > fm_rho_atm;
}

static double fm_Aphi(double R, double z)
{
   double rho, P, ut, uph;
   double A = 0.5 * fm_Bv * R * R;
   if (!fm_gas(R, z, &rho, &P, &ut, &uph)) return A;
   return rho > fm_cut ? A + fm_A0 * (rho - fm_cut) : A;
}

/* Covariant spatial u_i (R, phi, z) of u^mu = (u^t, 0, u^phi, 0) in the
 * Kerr-Schild cylindrical metric. */
static void fm_lower(double R, double z, double ut, double uph, double *u)
{
   kks_point p = kks_eval(fm_M, fm_a, R, z, 0);
   for (int i = 1; i < 4; ++i) {
      double git = p.H * p.l[i];
      double gip = p.H * p.l[i] * p.l[2] + (i == 2 ? R * R : 0.0);
      u[i - 1] = git * ut + gip * uph;
   }
}

/* The field of a face (prim[BRR] for an r face through x, prim[BZZ] for a
 * z face through x). */
static double fm_Br_face(double R, double z)
{
   double zm = fm_Zmin + floor((z - fm_Zmin) / fm_dz) * fm_dz, zp = zm + fm_dz;
   return -(fm_Aphi(R, zp) - fm_Aphi(R, zm)) / (R * fm_dz);
}
static double fm_Bz_face(double R, double z)
{
   double Rm = fm_Rmin + floor((R - fm_Rmin) / fm_dR) * fm_dR, Rp = Rm + fm_dR;
   return (fm_Aphi(Rp, z) - fm_Aphi(Rm, z)) / (0.5 * (Rp + Rm) * (Rp - Rm));
}

void setICparams(struct domain *theDomain)
{
   struct param_list *L = &theDomain->theParList;
   fm_gam = L->Adiabatic_Index;
   fm_M = L->metricPar2;
   fm_a = L->metricPar3;
   if (L->initPar1 > 0.0) fm_rin = L->initPar1;
   if (L->initPar2 > 0.0) fm_rmax = L->initPar2;
   fm_beta = L->initPar3 > 0.0 ? L->initPar3 : 100.0;
   fm_cut = L->initPar4 > 0.0 ? L->initPar4 : 0.2;
   if (L->initPar5 > 0.0) fm_rho_atm = L->initPar5;
   if (L->initPar6 > 0.0) fm_P_atm = L->initPar6;
   fm_Bv = L->initPar7;
   fm_Rmin = L->rmin;
   fm_dR = (L->rmax - L->rmin) / L->Num_R;
   fm_Zmin = L->zmin;
   fm_dz = (L->zmax - L->zmin) / L->Num_Z;

   /* ell of the prograde circular geodesic at r_max (Boyer-Lindquist,
    * equator): Omega_K = sqrt(M) / (r^1.5 + a sqrt(M)) */
   double M = fm_M, a = fm_a, r = fm_rmax;
   double OmK = sqrt(M) / (pow(r, 1.5) + a * sqrt(M));
   double gtt = -(1.0 - 2.0 * M / r), gtp = -2.0 * M * a / r;
   double gpp = r * r + a * a + 2.0 * M * a * a / r;
   double ut2 = -1.0 / (gtt + 2.0 * gtp * OmK + gpp * OmK * OmK);
   fm_ell = ut2 * (gtp + gpp * OmK);
   double ut, uph, lnq_max;
   fm_flow(fm_rin, 1.0, &ut, &uph, &fm_lnq_in);
   fm_flow(fm_rmax, 1.0, &ut, &uph, &lnq_max);
   double hmax = exp(lnq_max - fm_lnq_in);
   fm_K = (hmax - 1.0) * (fm_gam - 1.0) / fm_gam;

   /* field amplitude: the largest b^2 / 2 over the torus with A0 = 1,
    * sampled on a lattice across the torus's equatorial extent */
   double Bv = fm_Bv;
   fm_Bv = 0.0;
   fm_A0 = 1.0;
   double Pmax = fm_K * 1.0, b2max = 0.0;
   double Rhi = L->rmax, zhi = fabs(L->zmin) > fabs(L->zmax) ? fabs(L->zmin) : fabs(L->zmax);
   int ns = 400;
   double hR = 1e-4 * fm_rmax;
   for (int i = 0; i < ns; ++i)
      for (int k = 0; k < ns; ++k) {
         double R = fm_rin + (Rhi - fm_rin) * (i + 0.5) / ns;
         double z = -zhi + 2.0 * zhi * (k + 0.5) / ns;
         double rho, P;
         if (!fm_gas(R, z, &rho, &P, &ut, &uph)) continue;
         double BR = -(fm_Aphi(R, z + hR) - fm_Aphi(R, z - hR)) / (2.0 * hR * R);
         double BZ = (fm_Aphi(R + hR, z) - fm_Aphi(R - hR, z)) / (2.0 * hR * R);
         kks_point p = kks_eval(fm_M, fm_a, R, z, 0);
         double al = kks_lapse(&p), g[9], u[3];
         kks_gam(&p, g);
         double B[3] = {al * BR, 0.0, al * BZ};  /* the primitive over sqrt(gamma) / r */
         double B2 = 0.0;
         for (int m = 0; m < 3; ++m)
            for (int n = 0; n < 3; ++n) B2 += g[3 * m + n] * B[m] * B[n];
         fm_lower(R, z, ut, uph, u);
         double uB = u[0] * B[0] + u[2] * B[2];
         double w = al * ut;
         double b2 = (B2 + uB * uB) / (w * w);
         if (b2 > b2max) b2max = b2;
      }
   fm_A0 = b2max > 0.0 ? sqrt(2.0 * Pmax / fm_beta / b2max) : 0.0;
   fm_Bv = Bv;
}

void initial(double *prim, double *x)
{
   double R = x[0], z = x[2];
   double rho, P, ut, uph, u[3] = {0.0, 0.0, 0.0};
   if (fm_gas(R, z, &rho, &P, &ut, &uph)) {
      fm_lower(R, z, ut, uph, u);
   } else {
      rho = fm_rho_atm;
      P = fm_P_atm;
   }
   prim[RHO] = rho;
   prim[PPP] = P;
   prim[URR] = u[0];
   prim[UPP] = u[1];
   prim[UZZ] = u[2];
   if (NUM_C > BZZ) {
      prim[BRR] = fm_Br_face(R, z);
      prim[BPP] = 0.0;
      prim[BZZ] = fm_Bz_face(R, z);
   }
   for (int q = NUM_C; q < NUM_Q; ++q) prim[q] = 0.0;
}
