/* Planets plugin (plugin ABI of the reference Planets/ modules): a fixed,
 * unsoftened central point mass M = 1 at r = 0 and, when Mass_Ratio > 0, a
 * companion of mass Mass_Ratio on a circular Keplerian orbit at r = 1.25
 * (softening 0.05), moved analytically. */
#include <math.h>

#include "paul.h"

static double q_companion = 0.0;

void setPlanetParams(struct domain *theDomain)
{
   q_companion = theDomain->theParList.Mass_Ratio;
   theDomain->Npl = q_companion > 0.0 ? 2 : 1;
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
   if (q_companion > 0.0) {
      double r = 1.25;
      thePlanets[1].M = q_companion;
      thePlanets[1].vr = 0.0;
      thePlanets[1].omega = 1.0 / (r * sqrt(r));
      thePlanets[1].r = r;
      thePlanets[1].phi = 0.0;
      thePlanets[1].eps = 0.05;
      thePlanets[1].type = PLPOINTMASS;
   }
}

void movePlanets(struct planet *thePlanets, double t, double dt)
{
   if (q_companion > 0.0) thePlanets[1].phi += thePlanets[1].omega * dt;
}

void forcePlanets(struct planet *thePlanets, double dt) {}
