/* Boundary plugin (reference Boundary/ ABI): ghost annuli in r are reset to
 * initial() every substep; in z they copy the adjacent annulus with the
 * outflow diode (v_z may only point out of the domain). */
#include "paul.h"
#include "boundary.h"

void boundary_trans(struct domain *theDomain, int dim)
{
   if (dim == 1) {
      boundary_fixed_rinn(theDomain);
      boundary_fixed_rout(theDomain);
   } else if (dim == 2) {
      boundary_zerograd_zbot(theDomain, 1);
      boundary_zerograd_ztop(theDomain, 1);
   }
}
