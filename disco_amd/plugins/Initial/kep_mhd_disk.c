/*
 * Initial plugin: unstratified Keplerian MHD disk (reference plugin ABI,
 * setICparams / initial).
 *
 *   rho = 1, isothermal sound speed cs = v_K / Mach (Cs2_Profile 4, M =
 *   Metric_Par2), P = cs^2 rho / gamma (what the isothermal cons2prim sets),
 *   Omega from radial balance of gravity and the pressure gradient:
 *       Omega^2 = (M / r^3) (1 - 1 / (gamma Mach^2)),
 *   vertical field B_z at plasma beta = Init_Par1 (1000 when 0) measured at
 *   r = 1:  B_z = sqrt(2 P(1) / beta).
 *   Init_Par0 = 1 instead gives a zero-net-flux toroidal field
 *       B_phi = sqrt(2 P(r) / beta) sin(2 pi (r - R_Min) / (R_Max - R_Min)).
 *
 * No vertical gravity is assumed (use Grav2D 1), so the state is an exact
 * equilibrium up to the seeded velocity noise (Noise_Type 3 in in.par).
 */
#include "paul.h"

static double kd_gamma = 5.0 / 3.0;
static double kd_mach = 20.0;
static double kd_M = 1.0;
static double kd_beta = 1000.0;
static int kd_toroidal = 0;
static double kd_rmin = 0.5, kd_rmax = 2.0;

void setICparams(struct domain *theDomain)
{
   kd_gamma = theDomain->theParList.Adiabatic_Index;
   kd_mach = theDomain->theParList.Disk_Mach;
   kd_M = theDomain->theParList.metricPar2;
   kd_beta = theDomain->theParList.initPar1;
   if (kd_beta == 0.0) kd_beta = 1000.0;
   kd_toroidal = theDomain->theParList.initPar0;
   kd_rmin = theDomain->theParList.rmin;
   kd_rmax = theDomain->theParList.rmax;
}

void initial(double *prim, double *x)
{
   double r = x[0];
   double cs2 = kd_M / r / (kd_mach * kd_mach);
   double rho = 1.0;
   double P = cs2 * rho / kd_gamma;
   double om2 = kd_M / (r * r * r) * (1.0 - 1.0 / (kd_gamma * kd_mach * kd_mach));
   prim[RHO] = rho;
   prim[PPP] = P;
   prim[URR] = 0.0;
   prim[UPP] = sqrt(om2);
   prim[UZZ] = 0.0;
   if (NUM_C > BZZ) {
      prim[BRR] = 0.0;
      prim[BPP] = 0.0;
      prim[BZZ] = 0.0;
      if (kd_toroidal) {
         double B0 = sqrt(2.0 * P / kd_beta);
         prim[BPP] = B0 * sin(2.0 * M_PI * (r - kd_rmin) / (kd_rmax - kd_rmin));
      } else {
         double P1 = kd_M / (kd_mach * kd_mach) / kd_gamma;
         prim[BZZ] = sqrt(2.0 * P1 / kd_beta);
      }
   }
   int q;
   for (q = NUM_C; q < NUM_Q; ++q) prim[q] = r < 1.0 ? 1.0 : 0.0;
}
