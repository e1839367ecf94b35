/* Planets plugin (plugin ABI of the reference Planets/ modules): a fixed
 * central mass M = 1 at r = 0 and a planet of mass ratio q = Mass_Ratio on a
 * circular Keplerian orbit at r = 1, softened by 0.6 h with the disk's aspect
 * ratio h = 1 / Mach_Number; moved analytically (phi += omega dt). */
#include <math.h>

#include "paul.h"

static double pc_q = 1e-3;
static double pc_eps = 0.03;

void setPlanetParams(struct domain *theDomain)
{
   pc_q = theDomain->theParList.Mass_Ratio;
   pc_eps = 0.6 / theDomain->theParList.Disk_Mach;
   theDomain->Npl = 2;
}

int planet_motion_analytic(void) { return 1; }

void initializePlanets(struct planet *thePlanets)
{
   thePlanets[0].M = 1.0;
   thePlanets[0].vr = 0.0;
   thePlanets[0].omega = 0.0;
   thePlanets[0].r = 0.0;
   thePlanets[0].phi = 0.0;
   thePlanets[0].eps = 0.0;
   thePlanets[0].type = PLPOINTMASS;
   thePlanets[1].M = pc_q;
   thePlanets[1].vr = 0.0;
   thePlanets[1].omega = sqrt(1.0 + pc_q);
   thePlanets[1].r = 1.0;
   thePlanets[1].phi = 0.0;
   thePlanets[1].eps = pc_eps;
   thePlanets[1].type = PLPOINTMASS;
}

/* The orbit is the planet's own (its omega), so runs with different mass
 * ratios can share the module. */
void movePlanets(struct planet *thePlanets, double t, double dt)
{
   thePlanets[1].phi += thePlanets[1].omega * dt;
}

void forcePlanets(struct planet *thePlanets, double dt) {}
