/*
 * disco_hip.h -- C ABI of the MI355X timestep library (libdisco_hip.so).
 *
 * The library advances a Disco mesh (ragged annuli of cells, cylindrical
 * r/phi/z) by full RK2 timesteps on one GPU.  Every entry point takes plain
 * pointers and sizes; no C++ or framework types cross this boundary.
 *
 * Correspondence with the reference (HangChelseaSu/Disco):
 *   dh_timestep   replaces  void timestep(struct domain*, double dt)
 *                           (timestep.c:7-45: RK snapshot + onestep x2,
 *                            onestep.c:84-228)
 *   dh_getmindt   replaces  double getmindt(struct domain*)   (misc.c:39-84)
 *   dh_create / dh_upload_state / dh_download_state
 *                 replace the state ownership of struct domain
 *                 (theCells[jk][i], paul.h:134-203; domain.c:26-127): the
 *                 device holds the state, the host struct is a mirror that is
 *                 filled on demand.
 *   dh_set_planets  passes the planet list that planet_src reads
 *                 (planet.c:152-196) for each RK substep.
 * The reference-signature wrappers `timestep` and `getmindt` that a Disco
 * driver links against are in disco_amd/host/disco_host.c (declared in
 * disco_amd/host/disco_host.h).
 *
 * Layout of state arrays (all double precision, structure of arrays):
 *   cell c = off[jk] + i, with jk = j + Nr*k the annulus index (ghost annuli
 *   included, as theCells[jk] of the reference), i = 0..Np[jk]-1 and
 *   off[jk] = sum_{m<jk} Np[m].  Ntot = off[Nr*Nz].
 *   prim[q*Ntot + c], cons[q*Ntot + c]        q < NQ = num_c + num_n
 *   piph[c], dphi[c], wiph[c]
 *   Phi[f*Ntot + c]                           f < NF (0, 3 or 5 by ct_mode)
 *
 * Every function returns 0 on success and a nonzero DH_E* code otherwise;
 * dh_last_error() returns a message for the last failure on this thread.
 */
#ifndef DISCO_HIP_H
#define DISCO_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
   DH_OK = 0,
   DH_EINVAL = 1,      /* bad argument / unsupported configuration */
   DH_EHIP = 2,        /* HIP runtime error */
   DH_ENOMEM = 3,      /* device allocation failed */
   DH_ENODEV = 4       /* no GPU visible */
};

/* Hydro modules (reference Hydro/<name>.c). */
enum {
   DH_HYDRO_EULER = 0,  /* Hydro/euler.c */
   DH_HYDRO_MHD = 1,    /* Hydro/mhd.c */
   DH_HYDRO_GRMHD = 2   /* Hydro/grmhd2.c with Frame/normal.c and a DH_METRIC_SCHW_* or
                           DH_METRIC_KERR_KS metric (M = metric_par2) */
};

/* Metric modules (reference Hydro/Metric/<name>.c). */
enum {
   DH_METRIC_FLAT = 0,     /* flat_rpz.c (euler, mhd) */
   DH_METRIC_SCHW_KS_EQ = 1, /* schw_ks_eq_rpz.c: Schwarzschild in Kerr-Schild coordinates
                              evaluated on the equator (z = 0) */
   DH_METRIC_SCHW_KS = 2,  /* schw_ks_rpz.c: the same metric at (r, z) */
   DH_METRIC_KERR_KS = 3   /* the Kerr metric in Kerr-Schild form at (r, z), mass
                              metric_par2 and spin metric_par3 (no reference
                              module; disco_amd/plugins/Metric/kerr_ks_rpz.c is
                              the same metric written against Hydro/metric.h) */
};

/* The initial guess and stopping rule of grmhd2's cons2prim (noble2d,
 * grmhd2.c:1096-1314).  DH_C2P_GUESS_PREVIOUS and DH_C2P_STOP_REFERENCE are
 * the reference's: the previous primitives, and |d eta/eta| < 1e-12 followed
 * by 10 more iterations.  DH_C2P_GUESS_LEARNED takes (v^2, eta) from the
 * model given to dh_set_c2p_model (no counterpart in the reference), falling
 * back to the previous primitives where its output is unphysical or the solve
 * from it does not converge; DH_C2P_GUESS_BETTER takes the model's guess only
 * where, besides, its residual is below the previous primitives' (the better
 * of the two starts, each cell choosing); DH_C2P_STOP_RESIDUAL stops when
 * both residuals are below 1e-10 relative, without the extra iterations. */
enum { DH_C2P_GUESS_PREVIOUS = 0, DH_C2P_GUESS_LEARNED = 1, DH_C2P_GUESS_BETTER = 2 };
enum { DH_C2P_STOP_REFERENCE = 0, DH_C2P_STOP_RESIDUAL = 1 };

/* Boundary treatment of one side of the mesh, applied to ghost annuli after
 * every substep (reference onestep.c:194-210 -> Boundary/<name>.c).
 *   DH_BC_HELD:    ghost prims are left as uploaded.  This is exactly the
 *                  reference's fixed/polar boundaries (Boundary/fixed.c:24-60,
 *                  Boundary/polar.c:26-75) when initial() does not depend on
 *                  phi, which the host checks before choosing it.
 *   DH_BC_ZEROGRAD / DH_BC_DIODE: the reference's zero-gradient copy from the
 *                  adjacent annulus, optionally with the outflow diode
 *                  (boundary_functions.c:269-441, 90-159).
 *   DH_BC_NONE:    the side has no ghost annuli (NoBC_* = 1). */
enum { DH_BC_NONE = 0, DH_BC_HELD = 1, DH_BC_ZEROGRAD = 2, DH_BC_DIODE = 3 };

typedef struct dh_grid {
   int Nr, Nz;                  /* annuli including ghosts (paul.h:142) */
   int NgRa, NgRb, NgZa, NgZb;  /* ghost annuli on each side */
   const int *Np;               /* [Nr*Nz] cells per annulus */
   const double *r_jph;         /* [Nr+1]: radial edges, r_jph[0] is edge -1 */
   const double *z_kph;         /* [Nz+1]: vertical edges, z_kph[0] is edge -1 */
   double phi_max;              /* azimuthal period (2 pi for full disks) */
} dh_grid;

/* Run-time parameters: the in.par keys (readpar.c:70-155) and the compile-time
 * sizes (NUM_C, NUM_N, CT_MODE; paul.h:29-49) that the timestep reads. */
typedef struct dh_params {
   int num_c, num_n, ct_mode;   /* NUM_C, NUM_N, CT_MODE */
   int hydro;                   /* DH_HYDRO_* */
   int riemann_solver;          /* Riemann_Solver: 0 HLL, 1 HLLC, 2 HLLD */
   int mesh_motion;             /* Mesh_Motion: 0 fixed, 1 by cell, 3 avg, 4 exact */
   int exact_mesh_omega;        /* Exact_Mesh_Omega (omega.c:43-60) */
   int energy_omega;            /* Energy_Omega (omega.c:62-114) */
   int timestep_mode;           /* Timestep: 0 RK2, 1 forward Euler */
   int ct;                      /* Constrained_Transport */
   int isothermal;              /* Isothermal */
   int cs2_profile;             /* Cs2_Profile (omega.c:116-140) */
   int visc_flag, alpha_flag;   /* Use_Viscosity, Use_As_Alpha */
   int rot_frame;               /* RotFrame */
   int grav2d;                  /* Grav2D */
   int nobc_rmin, nobc_rmax, nobc_zmin, nobc_zmax;
   int r_periodic, z_periodic;
   int bc_rmin, bc_rmax, bc_zmin, bc_zmax;  /* DH_BC_* per side */
   int nozzle_type;             /* Nozzle_Type: 0 none, 1 the spherical nozzle
                                   of sink.c:35-70 */
   int metric;                  /* DH_METRIC_* (the METRIC module) */
   int c2p_guess;               /* DH_C2P_GUESS_*: grmhd2 cons2prim's initial guess */
   int c2p_stop;                /* DH_C2P_STOP_*: grmhd2 cons2prim's stopping rule */
   double gamma_law;            /* Adiabatic_Index */
   double plm;                  /* PLM */
   double cfl;                  /* CFL */
   double max_dt;               /* Max_DT */
   double rho_floor, p_floor;   /* Density_Floor, Pressure_Floor */
   double viscosity;            /* Viscosity */
   double mach;                 /* Mach_Number */
   double metric_par2;          /* Metric_Par2 (M in the Cs2 profiles) */
   double rot_omega, rot_d;     /* RotOmega, RotD */
   double init_par[9];          /* Init_Par0..8 (Init_Par0 as a double) */
   double phi_order;            /* PHI_ORDER of the softened point-mass force,
                                   2.0 (planet.c:4); a run-time value as there,
                                   so pow() is evaluated as the reference does */
   double metric_par3;          /* Metric_Par3 (the spin a of DH_METRIC_KERR_KS) */
   double nozzle_par[4];        /* Nozzle_Par1..4: centre radius, radius, speed,
                                   Mach number of the nozzle (sink.c:35-70) */
} dh_params;

/* Planet record passed to dh_set_planets: 8 doubles per planet,
 * {r, phi, M, omega, vr, eps, type, 0} (struct planet, paul.h:233-251). */
#define DH_PLANET_STRIDE 8
#define DH_MAX_PLANETS 8

typedef struct dh_ctx dh_ctx;

/* Allocates the device state for one mesh and builds the static tables
 * (annulus offsets, face-list offsets).  The grid arrays are copied. */
int dh_create(const dh_grid *grid, const dh_params *par, dh_ctx **out);
int dh_destroy(dh_ctx *ctx);

/* Number of cells (Ntot) and of conserved/face fields of a context. */
int dh_sizes(const dh_ctx *ctx, long *ntot, int *nq, int *nf);

/* Host <-> device copies of the whole state (layout above).  Any pointer may
 * be NULL to skip that field.  dphi is recomputed from piph on upload when
 * the pointer is NULL. */
int dh_upload_state(dh_ctx *ctx, const double *prim, const double *cons,
                    const double *piph, const double *dphi, const double *wiph,
                    const double *Phi);
int dh_download_state(dh_ctx *ctx, double *prim, double *cons, double *piph,
                      double *dphi, double *wiph, double *Phi);

/* Finishes the set-up of uploaded primitives (domain.c:175-241 and
 * main.c:53-62): cons = prim2cons(prim) and prim = cons2prim(cons) on real
 * cells, set_wcell on every cell, then the boundaries of every side.
 * Ghost prims of DH_BC_HELD sides must already hold their boundary values. */
int dh_init_cells(dh_ctx *ctx);

/* Number of r (dim 1) or z (dim 2) faces of the mesh: the sum over
 * neighbouring annulus pairs of Np[a] + Np[b] (faces_alt.c:80-96). */
int dh_face_count(dh_ctx *ctx, int dim, long *nfaces);

/* Builds the r (dim 1) or z (dim 2) faces of the current mesh and copies
 * them out: cell indices L and R, LRtype (0: the face's upper phi end is
 * L's), the face centroid cm[3*f..3*f+2] = {r, phi, z} and area dA.  Any
 * pointer may be NULL.  Replaces the face walk of setup_faces
 * (faces_alt.c:36-179) for a host that needs the faces, as set_B_fields
 * does to set the initial magnetic fluxes (bfields.c:45-78). */
int dh_get_faces(dh_ctx *ctx, int dim, int *L, int *R, int *LRtype, double *cm,
                 double *dA);

/* Planets seen by the sources of RK substep `substep` (0 or 1) of the next
 * dh_timestep.  npl <= DH_MAX_PLANETS. */
int dh_set_planets(dh_ctx *ctx, int substep, const double *planets, int npl);

/* CFL timestep: min over real cells of Hydro mindt, times CFL (misc.c:39-84). */
int dh_getmindt(dh_ctx *ctx, double *dt);

/* One full timestep (RK2 or forward Euler per timestep_mode). Asynchronous
 * with respect to the host; dh_getmindt / dh_download_state synchronise. */
int dh_timestep(dh_ctx *ctx, double dt);

/* Waits for all work queued on the context's stream. */
int dh_sync(dh_ctx *ctx);

/* Per-annulus diagnostics (the Analysis module): every step the reference
 * adds, for each annulus jk and quantity q, dt times the volume-weighted
 * mean of get_diagnostics(x, prim) over the annulus's cells to
 * Qrz[jk*nq + q] (analysis.c:52-106, called at timestep.c:42).  The device
 * accumulates into its own Qrz for the shipped modules below; planets are
 * the state after the step (the modules read planet 1, when Npl > 1). */
enum {
   DH_DIAG_EULER = 1,   /* Diagnostics/diag_euler.c,   8 quantities */
   DH_DIAG_DISK2D = 2,  /* Diagnostics/diag_disk2d.c,  5 quantities */
   DH_DIAG_MHDMRI = 3   /* Diagnostics/diag_mhdmri.c, 18 quantities */
};
#define DH_DIAG_MAXQ 18
int dh_add_diagnostics(dh_ctx *ctx, int kind, double dt, const double *planets, int npl);
/* Copies the accumulated Qrz[Nr*Nz*nq] out (nq of the kind last added). */
int dh_get_diagnostics(dh_ctx *ctx, double *Qrz);
/* Zeroes the accumulated Qrz (zero_diagnostics, analysis.c:8-26). */
int dh_zero_diagnostics(dh_ctx *ctx);
/* Divides the accumulated Qrz by t_avg (avg_diagnostics, analysis.c:28-49). */
int dh_avg_diagnostics(dh_ctx *ctx, double t_avg);

/* report (report.c:7-202) of the current state: out[16] receives the
 * report.dat columns after t, in its order: Torque, Power, Fr, rho_min,
 * rhoavg_min, PsiR, PsiI, Mass, Mdot, S_R, L1_rho, L1_isen, L1_B, Br2, aM,
 * bM.  The planet terms use planets 0 and 1 (DH_PLANET_STRIDE each) when
 * npl > 1, as the reference does when Npl > 1.  The sums run per annulus on
 * the device and over the annuli in the reference's order on the host, so
 * they equal the reference's to rounding. */
int dh_report(dh_ctx *ctx, const double *planets, int npl, double *out);

/* The learned initial guess of grmhd2's cons2prim (no counterpart in the
 * reference: an MLP 8 -> 64 -> 64 -> 2 run on MFMA).
 * dh_set_c2p_model loads its n = 4886 float parameters (layout in
 * disco_amd/csrc/dh_c2pnet.h) into a context created with
 * c2p_guess = DH_C2P_GUESS_LEARNED or _BETTER; dh_timestep refuses to run such a
 * context before.  dh_c2p_harvest writes the model's features
 * feat[8 * ntot] and targets targ[2 * ntot] of the current state (NaN
 * outside the real cells), the training data.  dh_c2p_stats sums the last
 * cons2prim's records over the real cells: out[4] = {cells, Newton
 * iterations, cells solved from the learned guess, cells that fell back to
 * the previous primitives}.  dh_c2p_predict writes the model's guesses
 * guess[2 * ntot] = (v^2, eta) per cell for the current state (NaN outside
 * the real cells).  dh_c2p_info writes the last cons2prim's record of every
 * cell, info[ntot] (0 outside the real cells): bits 0-5 the Newton
 * iterations (63: the last solve did not meet its stopping rule), 64 the
 * learned guess was used, 128 the solve fell back to the previous
 * primitives; dh_c2p_stats counts a solve that did not meet its rule as 63
 * iterations. */
int dh_set_c2p_model(dh_ctx *ctx, const float *params, long n);
int dh_c2p_harvest(dh_ctx *ctx, float *feat, float *targ);
int dh_c2p_predict(dh_ctx *ctx, float *guess);
int dh_c2p_stats(dh_ctx *ctx, long long *out);
int dh_c2p_info(dh_ctx *ctx, unsigned char *info);

/* Training of the learned guess on the device (no context): float32
 * forward and backward passes, deterministic per-parameter gradient sums
 * and Adam (disco_amd/csrc/dh_c2ptrain.h).  dh_c2p_train fits the model to
 * n rows of features X[n][8] and targets Y[n][2] (as dh_c2p_harvest writes
 * them) for `epochs` passes in batches of `batch` rows at learning rate lr
 * (cosine-decayed), from `seed`; it writes the 4886-float parameter block
 * and the last epoch's mean squared error.  dh_c2p_grad writes the gradient
 * of the mean squared error over all n rows with respect to the parameter
 * block (zero at the normalisation constants) and the error itself. */
int dh_c2p_train(const float *X, const float *Y, long n, int epochs, int batch, float lr,
                 unsigned long long seed, float *params, float *loss);
int dh_c2p_grad(const float *params, const float *X, const float *Y, long n, float *grad,
                float *loss);

/* dh_c2p_grad without any allocation, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc); no counterpart in the reference.
 * params[4886], X[n][8], Y[n][2], grad[4886] and loss[1] are DEVICE
 * pointers; the kernels are queued on `stream` (a hipStream_t; NULL is the
 * default stream) and the call returns without waiting for them.  n is at
 * most DH_C2P_NOALLOC_MAX_ROWS: the per-row activations live in one fixed
 * device array per GPU (4.2 MB, held until the process ends), shared by
 * every call on that GPU, so calls that overlap on one GPU race.  The
 * results equal dh_c2p_grad's bit for bit. */
#define DH_C2P_NOALLOC_MAX_ROWS 4096
int dh_c2p_grad_noalloc(const float *params, const float *X, const float *Y, long n,
                        float *grad, float *loss, void *stream);

/* Average device time (ms) of the kernels of the last `dh_timestep` calls,
 * accumulated per kernel family since dh_reset_timers; used by the benchmark
 * to report the dominant kernel.  names/ms/calls receive up to `cap` rows. */
int dh_enable_timers(dh_ctx *ctx, int on);
int dh_read_timers(dh_ctx *ctx, int cap, char (*names)[48], double *ms,
                   long *calls, int *nrows);

/* With on != 0, an RK2 `dh_timestep` is recorded once as a HIP graph and
 * replayed from then on (the step's dt and planets travel in device memory,
 * so replays take new values); results are bit-identical to the eager
 * launches.  Off by default; ignored while the timers are on and for
 * forward-Euler steps.  No counterpart in the reference (timestep.c:7-45
 * calls onestep twice); it removes the per-kernel launch cost of small
 * meshes.  The host build accepts and ignores it. */
int dh_enable_graphs(dh_ctx *ctx, int on);

/* Ensembles: independent meshes stepped together (no counterpart in the
 * reference, which runs one mesh per process).  The members are contexts of
 * one hydro module and field count, created and set up as usual (each with
 * its own grid, parameters, time step and planets); while they belong to an
 * ensemble their work runs on one stream.  Every kernel of a timestep is
 * launched once for all members: each member's launches are recorded in
 * the order dh_timestep makes them, and the launches of equal position
 * (the same kernel) become one launch whose blocks are dealt to the members
 * by a concatenated block table.
 *   dh_ens_getmindt   every member's CFL time step (getmindt, misc.c:39-84),
 *                     one reduction over all members' cells, one copy out;
 *   dh_ens_timestep   one timestep of every member with its own dt[m]
 *                     (timestep.c:7-45), with the planets each member was
 *                     given by dh_set_planets; recorded as one HIP graph
 *                     when dh_ens_enable_graphs is on and the launches are
 *                     the same as the step before;
 *   dh_ens_add_diagnostics  add_diagnostics (analysis.c:52-106) of every
 *                     member: planets[m * DH_MAX_PLANETS * DH_PLANET_STRIDE
 *                     ...] are member m's npl planets after the step;
 *   dh_ens_destroy    gives every member its own stream back.
 * The members' results equal what dh_timestep gives each of them alone. */
typedef struct dh_ens dh_ens;
int dh_ens_create(dh_ctx *const *members, int n, dh_ens **out);
int dh_ens_destroy(dh_ens *ens);
int dh_ens_getmindt(dh_ens *ens, double *dts);
int dh_ens_timestep(dh_ens *ens, const double *dts);
int dh_ens_add_diagnostics(dh_ens *ens, int kind, const double *dts, const double *planets,
                           int npl);
int dh_ens_enable_graphs(dh_ens *ens, int on);
int dh_ens_sync(dh_ens *ens);

/* Launch tuner (no counterpart in the reference, whose loops run in one
 * order on one core).  Every kernel of the step that runs over cells or
 * faces can be launched in several traversal orders: storage order with the
 * workgroups dealt to the XCDs round-robin (variant 0) or in contiguous
 * eighths (1), or whole annuli (annulus pairs) taken in strips of S layers
 * of z, the strips one after another, r after r within a strip and z
 * fastest (2, 3, ...; S = 4, 8, 16 or all layers).  Results are the same
 * in every order.  dh_autotune makes the next `steps` timesteps (eager,
 * never recorded) tuning steps: in tuning step s every such kernel runs
 * all its launches of the step in variant s mod (its variant count), timed
 * with events, so that each variant is timed over whole steps; then each
 * kernel keeps its fastest (steps at least the largest variant count time
 * them all; steps = 0 keeps the choices made so far).  Until a kernel is
 * tuned it runs in variant 0, or 1 for the kernels that gather from the
 * next annulus in r.
 * dh_read_tuning reports up to `cap` kernels: the chosen variant, its
 * strip height (0 for variants 0 and 1), and the mean launch time of the
 * chosen and of the untuned variant over the tuning steps (-1 if none).
 * dh_read_tuning_sweep reports every variant of the kernel `name`: its
 * strip height and mean launch time over the tuning steps (-1 if it was
 * not timed); *nvar is the kernel's variant count (0 for a name that has
 * not run), of which the first `cap` are written. */
int dh_autotune(dh_ctx *ctx, int steps);
int dh_read_tuning(dh_ctx *ctx, int cap, char (*names)[48], int *variant, int *strip,
                   double *ms_chosen, double *ms_default, int *nrows);
int dh_read_tuning_sweep(dh_ctx *ctx, const char *name, int cap, int *strip, double *ms,
                         int *nvar);

const char *dh_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* DISCO_HIP_H */
