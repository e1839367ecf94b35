/* Planets plugin: one fixed, unsoftened central point mass M = 1 at r = 0
 * (plugin ABI of the reference Planets/ modules). */
#include "paul.h"

void setPlanetParams(struct domain *theDomain) { theDomain->Npl = 1; }

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
}

void movePlanets(struct planet *thePlanets, double t, double dt) {}

void forcePlanets(struct planet *thePlanets, double dt) {}
