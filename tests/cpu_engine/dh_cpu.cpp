// dh_cpu.cpp -- the timestep kernels of disco_amd/csrc built as host code,
// for the test suite only.
//
// Every kernel body (dh_kernels.h) and the launch sequence (dh_pipeline.h)
// are the ones the GPU library runs; this host executor runs each "launch"
// as a loop over its threads.  Built with -ffp-contract=off and the host
// libm, the results are expected to match the oracle bit for bit, which
// checks the owner-computes restatement (merge paths, gather order, edge
// walks) without a GPU.  The library exports the same dh_* entry points as
// libdisco_hip.so.  Nothing in the product loads it.
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "disco_hip.h"
#include "dh_pipeline.h"

using namespace dh;

struct HostExec {
   std::vector<void *> allocs;
   long flipped = 0;  // faces handed over so far (a check that flips ran)
   long bc_rings = 0;   // ghost rings walked so far
   long bc_misses = 0;  // of them, rings where the GPU kernel's guesses miss

   template <class K>
   int run(const char *, const K &k, long n)
   {
      for (long t = 0; t < n; ++t) k(t);
      return DH_OK;
   }
   // A kernel with launch variants (the GPU tuner's numbering: 0 and 1 the
   // storage order, 2 + s group order s).  DH_CPU_ORDER = v runs every such
   // kernel in variant v where the kernel has it (strip orders visit every
   // element once, in another sequence: the results must not change).
   template <class K>
   int run_tuned(const char *, const K &k, long n, const GroupOrder *ords, int nord, const int *)
   {
      const char *e = getenv("DH_CPU_ORDER");
      int v = e ? atoi(e) : 0;
      if (v >= 2 && v - 2 < nord) {
         Grouped<K> g{k, ords[v - 2]};
         for (long t = 0; t < n; ++t) g(t);
         return DH_OK;
      }
      for (long t = 0; t < n; ++t) k(t);
      return DH_OK;
   }
   int ring_start(const Mesh &m, const double *piph, int *I0)
   {
      for (int a = 0; a < m.nann; ++a) I0[a] = ring_start_of(m, piph, a);
      return DH_OK;
   }
   // The sequential ring walk, and a count of the rings where the guesses
   // of the GPU's parallel walk (k_bc_scan) would not all hold, so that
   // the GPU would walk the ring again.
   int bc_scan(const Mesh &m, int side, const double *piph, const double *dphi, int *start)
   {
      long rings = (long)bc_lines(m, side) * bc_layers(m, side);
      KBcScan k{m, side, piph, dphi, start};
      for (long t = 0; t < rings; ++t) {
         k(t);
         int lines = bc_lines(m, side);
         int dst, src;
         bc_pair(m, side, (int)(t % lines), (int)(t / lines), &dst, &src);
         int c0 = m.off[dst], s0 = m.off[src], Np = m.Np[dst], Np1 = m.Np[src];
         int rot = 0;
         for (int i = 1; i < Np1; ++i)
            if (piph[s0 + i] < piph[s0 + rot]) rot = i;
         bool miss = false;
         for (int i = 1; i < Np; ++i)
            miss |= bc_guess(piph, s0, Np1, rot, piph[c0 + i - 1], m.Pmax) != start[c0 + i];
         ++bc_rings;
         bc_misses += miss;
      }
      return DH_OK;
   }
   int flips(const Mesh &m, int dim, const Faces &F, const signed char *flag, double *Phi,
             double *rkPhi)
   {
      KFlipsTrack k{m, dim, F, flag, Phi, rkPhi};
      for (long JK = 0; JK < F.npairs; ++JK) {
         if (track_skipped(F.npairs, (int)JK)) continue;
         for (int f = F.fI[JK]; f < F.fI[JK + 1]; ++f) flipped += flag[f] != 0;
         k(JK);
      }
      return DH_OK;
   }
   int diag(const Mesh &m, const dh_params &P, int kind, const DiagPlanet &p1,
            const double *prim, const double *piph, const double *dphi, double dt, double *Qrz)
   {
      for (int a = 0; a < m.nann; ++a)
         diag_annulus(m, P, kind, p1, prim, piph, dphi, dt, a, Qrz);
      return DH_OK;
   }
   int report(const Mesh &m, const dh_params &P, const DiagPlanet &p0, const DiagPlanet &p1,
              int mhd, const double *prim, const double *piph, const double *dphi, double *R)
   {
      for (int a = 0; a < m.nann; ++a)
         if (is_real(m, a % m.Nr, a / m.Nr))
            report_annulus(m, P, p0, p1, mhd, prim, piph, dphi, a, R);
      return DH_OK;
   }
   template <int NQ, class H>
   int min_dt(const Mesh &m, const dh_params &P, const double *prim, const double *piph,
              const double *dphi, const double *wiph, double *out)
   {
      double v = HUGE_VAL;
      for (long c = 0; c < m.ntot; ++c) {
         double d = cell_mindt<NQ, H>(m, P, prim, piph, dphi, wiph, c);
         if (d < HUGE_VAL && d > 0.0 && d < v) v = d;
      }
      *out = v;
      return DH_OK;
   }
   // The learned cons2prim guesses, cell by cell, with the device's bfloat16
   // rounding (c2pnet_eval_host) in place of its MFMA products.
   template <int NQ>
   int c2p_guess(const Mesh &m, const dh_params &P, const double *cons, const double *piph,
                 const double *dphi, const float *prm, const unsigned short *, float *guess)
   {
      for (long c = 0; c < m.ntot; ++c) {
         int a = m.cjk[c];
         int j = a % m.Nr, k = a / m.Nr;
         if (!is_real(m, j, k)) {
            guess[2 * c] = guess[2 * c + 1] = NAN;
            continue;
         }
         C2pInv v = cell_c2p_invariants<NQ>(m, P, cons, piph, dphi, c, j, k);
         float f[C2PNET_IN], t[C2PNET_OUT];
         c2p_features(v, P.gamma_law, f);
         c2pnet_eval_host(prm, f, t);
         c2p_targets_to_guess(t, v.e, guess + 2 * c);
      }
      return DH_OK;
   }
   int release(void *p)
   {
      for (size_t i = 0; i < allocs.size(); ++i)
         if (allocs[i] == p) {
            allocs.erase(allocs.begin() + (long)i);
            free(p);
            return DH_OK;
         }
      return p ? fail(DH_EINVAL, "release of an unknown pointer") : DH_OK;
   }
   // With DH_CPU_POISON set, new memory is filled with 0xff bytes (NaN
   // doubles, -1 ints) instead of zeros, as device memory holds whatever was
   // there: a kernel that reads an element nothing wrote then shows it.
   int alloc(void **p, size_t bytes)
   {
      *p = calloc(bytes ? bytes : 8, 1);
      if (*p && getenv("DH_CPU_POISON")) memset(*p, 0xff, bytes ? bytes : 8);
      if (!*p) return fail(DH_ENOMEM, "out of host memory (%zu bytes)", bytes);
      allocs.push_back(*p);
      return DH_OK;
   }
   int put(void *dst, const void *src, size_t bytes)
   {
      memcpy(dst, src, bytes);
      return DH_OK;
   }
   int get(void *dst, const void *src, size_t bytes)
   {
      memcpy(dst, src, bytes);
      return DH_OK;
   }
   int copy(void *dst, const void *src, size_t bytes)
   {
      memcpy(dst, src, bytes);
      return DH_OK;
   }
   int zero(void *p, size_t bytes)
   {
      memset(p, 0, bytes);
      return DH_OK;
   }
   int sync() { return DH_OK; }
   int set_step(StepArgs *dst, const StepArgs &sa)
   {
      *dst = sa;
      return DH_OK;
   }
   template <class F>
   int replayable(const void *, F &&body)
   {
      return body();
   }
   // Whether small meshes run launches that share no data as one (the GPU
   // executor's rule); DH_CPU_MERGE=0 keeps them apart, to check that both
   // sequences give the same state.
   bool merges_launches() const
   {
      const char *e = getenv("DH_CPU_MERGE");
      return !(e && atoi(e) == 0);
   }
};

struct dh_ctx {
   HostExec x;
   Buffers b;
};

static void release(dh_ctx *c)
{
   if (!c) return;
   for (void *p : c->x.allocs) free(p);
   delete c;
}

template <int NQ, class H>
static int do_init(dh_ctx *c)
{
   Engine<HostExec> e{c->x, c->b};
   return e.template init<NQ, H>();
}
template <int NQ, class H>
static int do_timestep(dh_ctx *c, double dt)
{
   Engine<HostExec> e{c->x, c->b};
   return e.template timestep<NQ, H>(dt);
}
template <int NQ, class H>
static int do_mindt(dh_ctx *c, double *dt)
{
   Engine<HostExec> e{c->x, c->b};
   return e.template mindt<NQ, H>(dt);
}
template <int NQ, class H>
static int do_faces(dh_ctx *c)
{
   Engine<HostExec> e{c->x, c->b};
   return e.template faces<NQ, H>();
}

template <int NQ, class H>
static int do_grads(dh_ctx *c, int dim, double *out)
{
   Engine<HostExec> e{c->x, c->b};
   Buffers &b = c->b;
   KPlmPhi<NQ> kp{b.m, b.P.plm, b.prim, b.dphi, b.gradp};
   DH_TRY(c->x.run("plm_phi", kp, b.m.ntot));
   DH_TRY(e.template topology<NQ>(true));
   const double *g = dim == 0 ? b.gradp : dim == 1 ? b.gradr : b.gradz;
   memcpy(out, g, sizeof(double) * NQ * b.m.ntot);
   return DH_OK;
}

extern "C" {
// Stage check for the tests: phi (dim 0), r (1) or z (2) slopes of the
// current state.
int dhx_slopes(dh_ctx *c, int dim, double *out)
{
   if (!c || !out || dim < 0 || dim > 2) return fail(DH_EINVAL, "bad argument");
   return DH_DISPATCH(c->b, do_grads, c, dim, out);
}

// Stage check for the tests: one Riemann solve of the module the context
// was created for; F - w U* into Fw (riemann.c:216-318).
int dhx_riemann(dh_ctx *c, const double *pL, const double *pR, double r, const double *n,
                double w, double *Fw)
{
   if (!c) return fail(DH_EINVAL, "null context");
   double F[9], U[9];
   if (c->b.mhd && c->b.nq == 8) MHD<8>::riemann(c->b.P, pL, pR, r, 0.0, n, w, F, U);
   else if (!c->b.mhd && c->b.nq == 5) Euler<5>::riemann(c->b.P, pL, pR, r, 0.0, n, w, F, U);
   else return fail(DH_EINVAL, "dhx_riemann: NQ %d not wired", c->b.nq);
   for (int q = 0; q < c->b.nq; ++q) Fw[q] = F[q] - w * U[q];
   return DH_OK;
}
}


// Every group order of every kind visits each element of its range exactly
// once: returns the number of orders checked, or -1 at the first that does
// not.
extern "C" long dhx_check_orders(dh_ctx *c)
{
   if (!c) return -1;
   Buffers &b = c->b;
   long nkind[ORD_KINDS] = {b.m.ntot, b.F[0].n, b.F[1].n, (long)b.F[0].n + b.F[1].n};
   long checked = 0;
   for (int kind = 0; kind < ORD_KINDS; ++kind)
      for (int s = 0; s < b.ord[kind].n; ++s) {
         const GroupOrder &o = b.ord[kind].o[s];
         long n = nkind[kind];
         if (o.pos[o.ng] != n) return -1;
         std::vector<char> seen((size_t)n, 0);
         for (long t = 0; t < n; ++t) {
            long e = o.elem(t);
            if (e < 0 || e >= n || seen[(size_t)e]) return -1;
            seen[(size_t)e] = 1;
         }
         ++checked;
      }
   return checked;
}

extern "C" long dhx_flip_count(dh_ctx *c) { return c ? c->x.flipped : -1; }
extern "C" long dhx_bc_rings(dh_ctx *c) { return c ? c->x.bc_rings : -1; }
extern "C" long dhx_bc_misses(dh_ctx *c) { return c ? c->x.bc_misses : -1; }
// The galloping search of the merge steps on a table of predicate values.
extern "C" int dhx_last_true(int N, int g, const unsigned char *truth)
{
   return last_true(N, g, [&](int m) { return truth[m] != 0; });
}

// The training entry points own a temporary executor for the call.
static int with_exec(int (*fn)(dh_ctx *, void *), void *arg)
{
   dh_ctx *c = new dh_ctx();
   int rc = fn(c, arg);
   release(c);
   return rc;
}

extern "C" {

const char *dh_last_error(void) { return g_err; }

int dh_destroy(dh_ctx *c)
{
   release(c);
   return DH_OK;
}

int dh_create(const dh_grid *g, const dh_params *par, dh_ctx **out)
{
   if (!g || !par || !out) return fail(DH_EINVAL, "null argument");
   *out = nullptr;
   dh_ctx *c = new dh_ctx();
   int rc = create_buffers(c->x, c->b, g, *par);
   if (rc) {
      release(c);
      return rc;
   }
   *out = c;
   return DH_OK;
}

int dh_sizes(const dh_ctx *c, long *ntot, int *nq, int *nf)
{
   if (!c) return fail(DH_EINVAL, "null context");
   if (ntot) *ntot = c->b.m.ntot;
   if (nq) *nq = c->b.nq;
   if (nf) *nf = c->b.nf;
   return DH_OK;
}

int dh_upload_state(dh_ctx *c, const double *prim, const double *cons, const double *piph,
                    const double *dphi, const double *wiph, const double *Phi)
{
   if (!c) return fail(DH_EINVAL, "null context");
   return upload_state(c->x, c->b, prim, cons, piph, dphi, wiph, Phi);
}

int dh_download_state(dh_ctx *c, double *prim, double *cons, double *piph, double *dphi,
                      double *wiph, double *Phi)
{
   if (!c) return fail(DH_EINVAL, "null context");
   return download_state(c->x, c->b, prim, cons, piph, dphi, wiph, Phi);
}

int dh_init_cells(dh_ctx *c)
{
   if (!c) return fail(DH_EINVAL, "null context");
   return DH_DISPATCH(c->b, do_init, c);
}

int dh_face_count(dh_ctx *c, int dim, long *nfaces)
{
   if (!c || !nfaces || (dim != 1 && dim != 2)) return fail(DH_EINVAL, "bad argument");
   *nfaces = c->b.F[dim - 1].n;
   return DH_OK;
}

int dh_get_faces(dh_ctx *c, int dim, int *L, int *R, int *LRtype, double *cm, double *dA)
{
   if (!c || (dim != 1 && dim != 2)) return fail(DH_EINVAL, "bad argument");
   if (c->b.F[dim - 1].n == 0) return DH_OK;
   DH_TRY(DH_DISPATCH(c->b, do_faces, c));
   return export_faces(c->x, c->b, dim, L, R, LRtype, cm, dA);
}

int dh_set_planets(dh_ctx *c, int substep, const double *pl, int npl)
{
   if (!c) return fail(DH_EINVAL, "null context");
   return set_planets(c->b, substep, pl, npl);
}

int dh_getmindt(dh_ctx *c, double *dt)
{
   if (!c || !dt) return fail(DH_EINVAL, "null argument");
   return DH_DISPATCH(c->b, do_mindt, c, dt);
}

int dh_timestep(dh_ctx *c, double dt)
{
   if (!c) return fail(DH_EINVAL, "null context");
   if (!(dt > 0.0)) return fail(DH_EINVAL, "dt must be positive");
   return DH_DISPATCH(c->b, do_timestep, c, dt);
}

int dh_sync(dh_ctx *c) { return c ? DH_OK : fail(DH_EINVAL, "null context"); }

int dh_add_diagnostics(dh_ctx *c, int kind, double dt, const double *planets, int npl)
{
   if (!c) return fail(DH_EINVAL, "null context");
   return add_diagnostics(c->x, c->b, kind, dt, planets, npl);
}

int dh_get_diagnostics(dh_ctx *c, double *Qrz)
{
   if (!c || !Qrz) return fail(DH_EINVAL, "null argument");
   return get_diagnostics(c->x, c->b, Qrz);
}

int dh_zero_diagnostics(dh_ctx *c)
{
   if (!c) return fail(DH_EINVAL, "null context");
   return zero_diagnostics(c->x, c->b);
}

int dh_avg_diagnostics(dh_ctx *c, double t_avg)
{
   if (!c) return fail(DH_EINVAL, "null context");
   return avg_diagnostics(c->x, c->b, t_avg);
}

int dh_set_c2p_model(dh_ctx *c, const float *params, long n)
{
   if (!c) return fail(DH_EINVAL, "null context");
   return set_c2p_model(c->x, c->b, params, n);
}

int dh_c2p_harvest(dh_ctx *c, float *feat, float *targ)
{
   if (!c) return fail(DH_EINVAL, "null context");
   return c2p_harvest(c->x, c->b, feat, targ);
}

int dh_c2p_predict(dh_ctx *c, float *guess)
{
   if (!c) return fail(DH_EINVAL, "null context");
   return c2p_predict(c->x, c->b, guess);
}


struct TrainArgs {
   const float *X, *Y;
   long n;
   int epochs, batch;
   float lr;
   unsigned long long seed;
   float *params, *loss;
};

struct GradArgs {
   const float *params, *X, *Y;
   long n;
   float *grad, *loss;
};

int dh_c2p_train(const float *X, const float *Y, long n, int epochs, int batch, float lr,
                 unsigned long long seed, float *params, float *loss)
{
   TrainArgs a{X, Y, n, epochs, batch, lr, seed, params, loss};
   return with_exec([](dh_ctx *c, void *p) {
      TrainArgs *t = (TrainArgs *)p;
      return c2p_train(c->x, t->X, t->Y, t->n, t->epochs, t->batch, t->lr, t->seed, t->params,
                       t->loss);
   }, &a);
}

int dh_c2p_grad(const float *params, const float *X, const float *Y, long n, float *grad,
                float *loss)
{
   GradArgs a{params, X, Y, n, grad, loss};
   return with_exec([](dh_ctx *c, void *p) {
      GradArgs *t = (GradArgs *)p;
      return c2p_grad(c->x, t->params, t->X, t->Y, t->n, t->grad, t->loss);
   }, &a);
}

int dh_c2p_stats(dh_ctx *c, long long *out)
{
   if (!c) return fail(DH_EINVAL, "null context");
   return c2p_stats(c->x, c->b, out);
}

int dh_c2p_info(dh_ctx *c, unsigned char *info)
{
   if (!c) return fail(DH_EINVAL, "null context");
   return c2p_info(c->x, c->b, info);
}

int dh_enable_timers(dh_ctx *c, int) { return c ? DH_OK : fail(DH_EINVAL, "null context"); }
int dh_enable_graphs(dh_ctx *c, int) { return c ? DH_OK : fail(DH_EINVAL, "null context"); }
int dh_autotune(dh_ctx *c, int) { return c ? DH_OK : fail(DH_EINVAL, "null context"); }
// Ensembles group device launches; the host build has none.
int dh_ens_create(dh_ctx *const *, int, dh_ens **) { return fail(DH_ENODEV, "ensembles need the GPU library"); }
int dh_ens_destroy(dh_ens *) { return DH_OK; }
int dh_ens_getmindt(dh_ens *, double *) { return fail(DH_ENODEV, "ensembles need the GPU library"); }
int dh_ens_timestep(dh_ens *, const double *) { return fail(DH_ENODEV, "ensembles need the GPU library"); }
int dh_ens_add_diagnostics(dh_ens *, int, const double *, const double *, int)
{
   return fail(DH_ENODEV, "ensembles need the GPU library");
}
int dh_ens_enable_graphs(dh_ens *, int) { return fail(DH_ENODEV, "ensembles need the GPU library"); }
int dh_ens_sync(dh_ens *) { return fail(DH_ENODEV, "ensembles need the GPU library"); }
// The host build has no device array to work in.
int dh_c2p_grad_noalloc(const float *, const float *, const float *, long, float *, float *, void *)
{
   return fail(DH_ENODEV, "dh_c2p_grad_noalloc needs the GPU library");
}
int dh_read_tuning(dh_ctx *c, int, char (*)[48], int *, int *, double *, double *, int *nrows)
{
   if (!c || !nrows) return fail(DH_EINVAL, "bad arguments");
   *nrows = 0;
   return DH_OK;
}
int dh_read_tuning_sweep(dh_ctx *, const char *, int, int *, double *, int *nvar)
{
   if (nvar) *nvar = 0;
   return DH_OK;
}

int dh_report(dh_ctx *c, const double *planets, int npl, double *out)
{
   if (!c || !out) return fail(DH_EINVAL, "null argument");
   return report(c->x, c->b, planets, npl, out);
}

int dh_read_timers(dh_ctx *c, int, char (*)[48], double *, long *, int *nrows)
{
   if (!c) return fail(DH_EINVAL, "null context");
   if (nrows) *nrows = 0;
   return DH_OK;
}

}  // extern "C"
