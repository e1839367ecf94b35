/* Link-time fix for building the reference with HYDRO=mhd: misc.c:571-572
 * calls visc_source(), which Hydro/mhd.c does not define at this revision.
 * MHD configurations run with Use_Viscosity 0, so it is never called. */
void visc_source(const double *prim, const double *gradr, const double *gradp,
                 const double *gradz, double *cons, const double *xp,
                 const double *xm, double dVdt)
{
}
