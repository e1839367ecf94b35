/*
 * disco_oracle.h -- CPU restatement of the reference Disco timestep.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the HIP path is
 * compared against; only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg may load it.  It is never part of the product path.
 *
 * It keeps the reference's algorithm and floating-point evaluation order
 * (sequential scatter loops, same expression association) on the
 * structure-of-arrays state of include/disco_hip.h, so that on the same input
 * it reproduces the reference build (oracle/_ref) bit for bit.
 */
#ifndef DISCO_ORACLE_H
#define DISCO_ORACLE_H

#include "disco_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_ctx or_ctx;

or_ctx *or_create(const dh_grid *grid, const dh_params *par);
void or_destroy(or_ctx *o);
long or_ntot(const or_ctx *o);
void or_set_state(or_ctx *o, const double *prim, const double *cons,
                  const double *piph, const double *dphi, const double *wiph,
                  const double *Phi);
void or_get_state(const or_ctx *o, double *prim, double *cons, double *piph,
                  double *dphi, double *wiph, double *Phi);
void or_set_planets(or_ctx *o, int substep, const double *pl, int npl);
double or_getmindt(or_ctx *o);
void or_timestep(or_ctx *o, double dt);

/* Stage-level entry points (state in, stage result out) for unit checks. */
void or_plm_phi(or_ctx *o, double *gradp_out);
/* Builds the r (dim=1) or z (dim=2) face list; returns the face count and,
 * when the pointers are non-NULL, fills L, R (global cell indices), LRtype,
 * and the face geometry dA, phi-centre and phi-extent. */
long or_faces(or_ctx *o, int dim, int *L, int *R, int *LRtype, double *dA,
              double *phic, double *dphi);
/* plm_phi followed by plm_trans for dim (1 or 2): the limited slopes. */
void or_plm_trans(or_ctx *o, int dim, double *grad_out);
/* One Riemann solve in the orthonormal frame, as solve_riemann does it
 * (riemann.c:216-318) without the cons update: F - w*U* for unit area. */
void or_riemann_flux(or_ctx *o, const double *primL, const double *primR,
                     const double *x, const double *n, double w, double *F);
/* Hydro prim2cons/cons2prim at position x with volume dV. */
void or_prim2cons(or_ctx *o, const double *prim, double *cons, const double *x,
                  double dV);
void or_cons2prim(or_ctx *o, const double *cons, double *prim, const double *x,
                  double dV);

#ifdef __cplusplus
}
#endif
#endif
