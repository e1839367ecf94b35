/*
 * Initial plugin: a magnetised gas disk on the equator of a Schwarzschild
 * black hole of mass M = Metric_Par2, in the primitive variables of the
 * GRMHD module (reference plugin ABI, setICparams / initial).
 *
 *   rho = (r / 6M)^-1.5, P = 0.01 rho (adiabatic, Adiabatic_Index),
 *   the covariant four-velocity of a circular geodesic in Kerr-Schild
 *   coordinates (u^r = 0): u_t-coupled radial component
 *   u_r = g_rt u^t = 2M / (r sqrt(1 - 3M/r)), u_phi = sqrt(M r) / sqrt(1 - 3M/r),
 *   u_z = 0, and a vertical field B^z = sqrt(2 P(6M) / beta), beta = Init_Par1
 *   (100 when 0).  Init_Par0 = 1 starts the gas at rest in u_r instead
 *   (u_r = 0, not an orbit: it falls in).
 *
 * With u^r = 0 the disk is close to equilibrium (pressure and field are
 * small), so a step's changes are small and smooth.
 */
#include <math.h>

#include "paul.h"

static double gd_M = 1.0;
static double gd_beta = 100.0;
static int gd_infall = 0;

void setICparams(struct domain *theDomain)
{
   gd_M = theDomain->theParList.metricPar2;
   gd_beta = theDomain->theParList.initPar1;
   if (gd_beta == 0.0) gd_beta = 100.0;
   gd_infall = theDomain->theParList.initPar0;
}

void initial(double *prim, double *x)
{
   double r = x[0];
   double r0 = 6.0 * gd_M;
   double rho = pow(r / r0, -1.5);
   double P = 0.01 * rho;
   prim[RHO] = rho;
   prim[PPP] = P;
   double ut = 1.0 / sqrt(1.0 - 3.0 * gd_M / r);
   prim[URR] = gd_infall ? 0.0 : 2.0 * gd_M / r * ut;
   prim[UPP] = sqrt(gd_M * r) * ut;
   prim[UZZ] = 0.0;
   if (NUM_C > BZZ) {
      prim[BRR] = 0.0;
      prim[BPP] = 0.0;
      prim[BZZ] = sqrt(2.0 * 0.01 / gd_beta);
   }
   for (int q = NUM_C; q < NUM_Q; ++q) prim[q] = r < r0 ? 1.0 : 0.0;
}
