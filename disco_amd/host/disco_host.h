/*
 * disco_host.h -- host layer of the drop-in (libdisco_host_<config>.so and the
 * `disco` driver), built per module selection like the reference's Makefile.
 *
 * Reference-signature entry points (what a Disco driver links against):
 *   double getmindt(struct domain *)            misc.c:39-84
 *   void   timestep(struct domain *, double)    timestep.c:7-45
 *   void   add_diagnostics / avg_diagnostics / zero_diagnostics
 *                                               analysis.c:8-106 (the sums
 *          live on the device: a binding leaves analysis.o out)
 * Helpers under their own names, so that a binding that keeps the
 * reference's domain.o and report.o (which define check_dt and report)
 * links without two definitions of one symbol:
 *   disco_check_dt        check_dt, domain.c:284-311
 *   disco_report          report, report.c:7-202, from the device's sums
 *   disco_possibly_report possiblyOutput's report schedule, domain.c:316-335
 * Set-up (readpar.c:44-188, gridsetup.c:15-144, domain.c:26-242):
 *   disco_read_par, disco_setup_grid, disco_setup
 * Device state <-> host mirror: disco_sync_to_host, disco_export_state.
 * Checkpoint and restart (raw DSTATE01 files): disco_write_dstate,
 * disco_restart_dstate.
 */
#ifndef DISCO_HOST_H
#define DISCO_HOST_H

#include "paul.h"

#ifdef __cplusplus
extern "C" {
#endif

int disco_read_par(struct domain *d, const char *file);
void disco_setup_grid(struct domain *d);
int disco_setup(struct domain *d);
int disco_sync_to_host(struct domain *d);
int disco_export_state(struct domain *d, double *prim, double *cons, double *piph,
                       double *dphi, double *wiph);
long disco_ntot(const struct domain *d);
void *disco_device_context(struct domain *d);
void disco_free(struct domain *d);
const char *disco_last_error(void);
struct domain *disco_new(void);
void disco_delete(struct domain *d);
long disco_real_cells(const struct domain *d);
int disco_write_dstate(struct domain *d, const char *path);
int disco_restart_dstate(struct domain *d, const char *path);

/* timestep's host parts around the device step, for a host that steps
 * several meshes together (dh_ens_*): disco_step_begin (planets for the
 * substeps), disco_step_end (t, planets, step count), and for the
 * diagnostics disco_pack_planets, disco_diag_kind, disco_add_diag_time. */
int disco_step_begin(struct domain *d, double dt);
void disco_step_end(struct domain *d, double dt);
void disco_pack_planets(struct domain *d, double *out);
int disco_diag_kind(struct domain *d);
int disco_npl(struct domain *d);
void disco_add_diag_time(struct domain *d, double dt);

double getmindt(struct domain *d);
void timestep(struct domain *d, double dt);
void disco_check_dt(struct domain *d, double *dt);
void add_diagnostics(struct domain *d, double dt);
/* report (report.c:7-202) appended to report.dat, and possiblyOutput's
 * schedule of it (domain.c:316-335). */
void disco_report(struct domain *d);
void disco_possibly_report(struct domain *d, int override);
void avg_diagnostics(struct domain *d);
void zero_diagnostics(struct domain *d);
/* theTools.Qrz after the last step (disco_sync_to_host refreshes it). */
int disco_get_diagnostics(struct domain *d, double *Qrz);

#ifdef __cplusplus
}
#endif
#endif
