/*
 * Initial plugin: a magnetised gas disk on the equator of a Schwarzschild
 * black hole of mass M = Metric_Par2, in the primitive variables of the
 * GRMHD module (reference plugin ABI, setICparams / initial).
 *
 *   rho = (r / 6M)^-1.5, P = 0.01 rho (adiabatic, Adiabatic_Index),
 *   covariant four-velocity u_r = 0, u_z = 0 and the angular momentum of a
 *   circular geodesic, u_phi = sqrt(M r) / sqrt(1 - 3 M / r),
 *   vertical field B^z = sqrt(2 P(6M) / beta), beta = Init_Par1 (100 when 0).
 *
 * The state is not an equilibrium in Kerr-Schild coordinates (the gas starts
 * falling in), which makes every flux and source term of a step non-trivial.
 */
#include <math.h>

#include "paul.h"

static double gd_M = 1.0;
static double gd_beta = 100.0;

void setICparams(struct domain *theDomain)
{
   gd_M = theDomain->theParList.metricPar2;
   gd_beta = theDomain->theParList.initPar1;
   if (gd_beta == 0.0) gd_beta = 100.0;
}

void initial(double *prim, double *x)
{
   double r = x[0];
   double r0 = 6.0 * gd_M;
   double rho = pow(r / r0, -1.5);
   double P = 0.01 * rho;
   prim[RHO] = rho;
   prim[PPP] = P;
   prim[URR] = 0.0;
   prim[UPP] = sqrt(gd_M * r) / sqrt(1.0 - 3.0 * gd_M / r);
   prim[UZZ] = 0.0;
   if (NUM_C > BZZ) {
      prim[BRR] = 0.0;
      prim[BPP] = 0.0;
      prim[BZZ] = sqrt(2.0 * 0.01 / gd_beta);
   }
   for (int q = NUM_C; q < NUM_Q; ++q) prim[q] = r < r0 ? 1.0 : 0.0;
}
