/*
 * Initial plugin: isentropic (Yee et al. 1999) vortex centred on the polar
 * axis, in the reference plugin ABI (setICparams / initial, as in
 * Initial/vortex.c:4-7 of the reference).
 *
 *   rho_inf = P_inf = 1,   T = P/rho = 1 - (gamma-1) eps^2 / (8 gamma pi^2) e^{1-r^2}
 *   rho = T^{1/(gamma-1)},  P = rho T
 *   v_phi = eps/(2 pi) r e^{(1-r^2)/2}  ->  prim[UPP] = omega = v_phi / r
 *
 * Parameters: Adiabatic_Index (gamma), Init_Par1 (eps; 5 when left at 0).
 * Passive scalars (NUM_N > 0) are set to e^{-r^2}, so every field depends on r
 * only and the state is an exact steady state of the Euler equations.
 */
#include "paul.h"

static double yv_gamma = 5.0 / 3.0;
static double yv_eps = 5.0;

void setICparams(struct domain *theDomain)
{
   yv_gamma = theDomain->theParList.Adiabatic_Index;
   yv_eps = theDomain->theParList.initPar1;
   if (yv_eps == 0.0) yv_eps = 5.0;
}

void initial(double *prim, double *x)
{
   double r = x[0];
   double g = yv_gamma;
   double T = 1.0 - (g - 1.0) * yv_eps * yv_eps / (8.0 * g * M_PI * M_PI)
                        * exp(1.0 - r * r);
   double rho = pow(T, 1.0 / (g - 1.0));

   prim[RHO] = rho;
   prim[PPP] = rho * T;
   prim[URR] = 0.0;
   prim[UPP] = yv_eps / (2.0 * M_PI) * exp(0.5 * (1.0 - r * r));
   prim[UZZ] = 0.0;
   int q;
   for (q = NUM_C; q < NUM_Q; ++q) prim[q] = exp(-r * r);
}
