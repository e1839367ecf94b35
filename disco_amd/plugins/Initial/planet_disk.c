/*
 * Initial plugin: a 2-D isothermal gas disk for disk-planet runs (reference
 * plugin ABI, setICparams / initial).
 *
 *   Sigma = r^-0.5, sound speed cs = v_K / Mach (aspect ratio h = 1/Mach,
 *   Mach = Mach_Number, the Cs2_Profile 4 profile with M = Metric_Par2),
 *   P = cs^2 Sigma / gamma (what the isothermal cons2prim sets), and the
 *   rotation of radial balance against the central mass and the pressure
 *   gradient (P ~ r^-1.5):  Omega^2 = (M / r^3) (1 - 1.5 / (gamma Mach^2)).
 *   A passive scalar, when there is one, marks r < 1.
 */
#include <math.h>

#include "paul.h"

static double pd_gamma = 5.0 / 3.0;
static double pd_mach = 20.0;
static double pd_M = 1.0;

void setICparams(struct domain *theDomain)
{
   pd_gamma = theDomain->theParList.Adiabatic_Index;
   pd_mach = theDomain->theParList.Disk_Mach;
   pd_M = theDomain->theParList.metricPar2;
}

void initial(double *prim, double *x)
{
   double r = x[0];
   double sigma = 1.0 / sqrt(r);
   double cs2 = pd_M / r / (pd_mach * pd_mach);
   double om2 = pd_M / (r * r * r) * (1.0 - 1.5 / (pd_gamma * pd_mach * pd_mach));
   prim[RHO] = sigma;
   prim[PPP] = cs2 * sigma / pd_gamma;
   prim[URR] = 0.0;
   prim[UPP] = sqrt(om2);
   prim[UZZ] = 0.0;
   int q;
   for (q = NUM_C; q < NUM_Q; ++q) prim[q] = r < 1.0 ? 1.0 : 0.0;
}
