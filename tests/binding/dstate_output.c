/*
 * Output plugin (reference ABI: void output(struct domain *, char *)) for
 * the binding test: writes the cells of theDomain in the DSTATE01 layout
 * (disco_amd/state_io.py) to <filestart>.dstate.  Test infrastructure only.
 */
#include "paul.h"

void output(struct domain *d, char *filestart)
{
   char path[512];
   snprintf(path, sizeof path, "%s.dstate", filestart);
   FILE *f = fopen(path, "wb");
   if (!f) return;
   int Nr = d->Nr, Nz = d->Nz, Ntot = 0, a, i, q;
   for (a = 0; a < Nr * Nz; ++a) Ntot += d->Np[a];
   char magic[8] = {'D', 'S', 'T', 'A', 'T', 'E', '0', '1'};
   int hdr[16] = {0};
   hdr[0] = NUM_Q; hdr[1] = NUM_FACES; hdr[2] = Nr; hdr[3] = Nz;
   hdr[4] = d->NgRa; hdr[5] = d->NgRb; hdr[6] = d->NgZa; hdr[7] = d->NgZb;
   hdr[8] = Ntot; hdr[9] = d->Npl; hdr[10] = d->count_steps; hdr[11] = NUM_C;
   double dhdr[8] = {0};
   dhdr[0] = d->t;
   dhdr[1] = d->phi_max;
   fwrite(magic, 1, 8, f);
   fwrite(hdr, sizeof(int), 16, f);
   fwrite(dhdr, sizeof(double), 8, f);
   fwrite(d->Np, sizeof(int), Nr * Nz, f);
   fwrite(d->r_jph - 1, sizeof(double), Nr + 1, f);
   fwrite(d->z_kph - 1, sizeof(double), Nz + 1, f);
#define FIELD(expr)                                                  \
   for (a = 0; a < Nr * Nz; ++a)                                     \
      for (i = 0; i < d->Np[a]; ++i) {                               \
         struct cell *c = &d->theCells[a][i];                        \
         double v = (expr);                                          \
         fwrite(&v, sizeof(double), 1, f);                           \
      }
   for (q = 0; q < NUM_Q; ++q) FIELD(c->prim[q])
   for (q = 0; q < NUM_Q; ++q) FIELD(c->cons[q])
   FIELD(c->piph)
   FIELD(c->dphi)
   FIELD(c->wiph)
   for (q = 0; q < NUM_FACES; ++q) FIELD(c->Phi[q])
#undef FIELD
   for (a = 0; a < d->Npl; ++a) {
      struct planet *p = d->thePlanets + a;
      double v[8] = {p->r, p->phi, p->M, p->omega, p->vr, p->eps, (double)p->type, 0.0};
      fwrite(v, sizeof(double), 8, f);
   }
   fclose(f);
}
