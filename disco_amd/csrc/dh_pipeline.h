// dh_pipeline.h -- buffers of one mesh and the sequence of kernels that
// makes a timestep (timestep.c:7-45, onestep.c:84-228).
//
// The sequence is written once against an executor X that owns memory and
// launches: the HIP executor (dh_engine.hip) runs it on the GPU, and the
// test suite instantiates a host executor to run the same bodies on the CPU
// against the oracle.  X provides
//   run(name, kernel, n)          kernel(t) for t in [0, n)
//   ring_start(m, piph, I0)       start cell of every annulus
//   flips(m, dim, F, flag, Phi, rkPhi)   hand-overs of flagged faces per track
//   min_dt<NQ,H>(m, P, prim, piph, dphi, wiph, &dt)
//   diag(m, P, kind, pl1, prim, piph, dphi, dt, Qrz)   per-annulus diagnostics
//   c2p_guess<NQ>(m, P, cons, piph, dphi, prm, frag, guess)   learned cons2prim guesses
//   alloc/put/get/copy/zero/sync  memory
#pragma once
#include <cstdarg>
#include <cstdio>
#include <cmath>
#include <cstring>
#include <vector>

#include "dh_c2ptrain.h"
#include "dh_kernels.h"

namespace dh {

inline thread_local char g_err[512];

inline int fail(int code, const char *fmt, ...)
{
   va_list ap;
   va_start(ap, fmt);
   vsnprintf(g_err, sizeof g_err, fmt, ap);
   va_end(ap);
   return code;
}

#define DH_TRY(expr)                                                          \
   do {                                                                       \
      int rc_ = (expr);                                                       \
      if (rc_) return rc_;                                                    \
   } while (0)

// Traversal orders of the four element kinds the kernels run over: the
// cells (groups: annuli), the r faces and the z faces (groups: annulus
// pairs), and both face lists as one range (the edge kernels; an r pair's
// faces, then the z pair's of the same (j, k)).  Order s of a kind visits
// its groups (j, k) in strips of strip[s] layers of k: strip after strip,
// j after j within a strip, k fastest within j.  Layer k + 1 of an
// annulus is Nr annuli away in storage and j + 1 only one, so in storage
// order the z neighbours of a gather are far apart in time and in the
// strip orders both neighbours are close; the strip height trades the
// annuli between the two.  With one layer every strip order is the storage
// order, and a set has none.
enum { ORD_CELLS = 0, ORD_FACES_R = 1, ORD_FACES_Z = 2, ORD_EDGES = 3, ORD_KINDS = 4 };
constexpr int MAX_ORDERS = 4;
struct OrderSet {
   GroupOrder o[MAX_ORDERS];
   int strip[MAX_ORDERS];
   int n = 0;
};

// Meshes below this many cells (ghosts included) run some independent
// launches of a substep as one.
constexpr long SMALL_MESH_CELLS = 1L << 20;

struct Buffers {
   dh_params P;
   int nq = 0, nf = 0;
   bool mhd = false;
   Mesh m;
   int *Np = nullptr, *off = nullptr, *cjk = nullptr;
   double *r = nullptr, *z = nullptr;
   std::vector<int> h_Np, h_off;
   Faces F[2];
   int *rng[2] = {nullptr, nullptr};
   double *prim = nullptr, *cons = nullptr, *rkcons = nullptr;
   double *gradp = nullptr, *gradr = nullptr, *gradz = nullptr;
   double *piph = nullptr, *piph2 = nullptr, *dphi = nullptr, *dphi2 = nullptr;
   double *wiph = nullptr;
   double *dF[2] = {nullptr, nullptr}, *dFp = nullptr;
   int *I0 = nullptr;
   // constrained transport
   double *Phi = nullptr, *rkPhi = nullptr, *emf_phi = nullptr;
   double *emf_f[2] = {nullptr, nullptr};
   double *Epre = nullptr, *Eo = nullptr, *Eown = nullptr, *Ecarry = nullptr;
   signed char *flag[2] = {nullptr, nullptr};
   int *estep[2] = {nullptr, nullptr};  // merge steps of each face on its two edge lines
   int *lbase[2] = {nullptr, nullptr};  // first segment of the edge line whose first list is pair JK
   double *seg[2] = {nullptr, nullptr}; // E dl dt of every segment of every edge line
   int *bcstart = nullptr;          // where each ghost cell's overlap sum starts
   // Traversal orders the executor may launch a kernel in (OrderSet)
   OrderSet ord[ORD_KINDS];
   OrderSet none;  // storage order only (launches whose elements are not one kind)
   int bc_maxnp[4][8] = {{0}};      // largest Np of each side's ghost layers
   PlanetSet pl[2];
   StepArgs *step = nullptr;  // device copy of {dt, planets} of the current timestep
   // The ring starts, cell ranges, faces and edge-line steps in the buffers
   // are those of the current face positions (set by an RK2 step, whose
   // second substep builds them for positions it does not move).  Every
   // path that writes piph or dphi other than the timestep itself (state
   // upload, cell set-up, a restart, any new entry point) must clear it, or
   // the next step would take faces of the old positions.
   bool topo_valid = false;
   // diagnostics (analysis.c:52-106): Qrz[nann * DH_DIAG_MAXQ], allocated on
   // first use; the kind that accumulated it
   double *diagQ = nullptr;
   int diag_kind = 0;
   // report (report.c:7-202): per-annulus sums R[nann * RP_N], allocated on
   // first use
   double *repQ = nullptr;
   // grmhd2 cons2prim: every real cell's solve record (C2P_INFO_*); with the
   // learned guess its model (float parameters and bfloat16 MFMA fragments)
   // and the per-cell guesses (v^2, eta)
   unsigned char *c2p_info = nullptr;
   float *c2p_prm = nullptr, *c2p_guess = nullptr;
   unsigned short *c2p_frag = nullptr;
   bool c2p_model = false;
};

// Checks a grid and parameter set against what the engine implements.
inline int validate(const dh_grid *g, const dh_params &P, long *ntot_out)
{
   int nq = P.num_c + P.num_n;
   if (P.hydro != DH_HYDRO_GRMHD && P.metric != DH_METRIC_FLAT)
      return fail(DH_EINVAL, "euler and mhd run with the flat_rpz metric");
   if (P.hydro == DH_HYDRO_EULER) {
      if (P.num_c != 5) return fail(DH_EINVAL, "euler needs NUM_C = 5 (got %d)", P.num_c);
      if (P.ct_mode != 0) return fail(DH_EINVAL, "CT_MODE %d needs the MHD module", P.ct_mode);
      if (P.riemann_solver != 0 && P.riemann_solver != 1)
         return fail(DH_EINVAL, "Riemann_Solver %d not supported by euler", P.riemann_solver);
      if (nq < 5 || nq > 7) return fail(DH_EINVAL, "NUM_C+NUM_N = %d not supported", nq);
   } else if (P.hydro == DH_HYDRO_MHD) {
      if (P.num_c != 8) return fail(DH_EINVAL, "mhd needs NUM_C = 8 (got %d)", P.num_c);
      if (P.ct_mode != 2 || !P.ct)
         return fail(DH_EINVAL, "mhd is supported with CT_MODE 2 and Constrained_Transport 1");
      if (P.riemann_solver != 2) return fail(DH_EINVAL, "mhd is supported with HLLD only");
      if (P.visc_flag) return fail(DH_EINVAL, "mhd has no viscous terms (Use_Viscosity 0)");
      if (nq < 8 || nq > 9) return fail(DH_EINVAL, "NUM_C+NUM_N = %d not supported", nq);
      if (g->NgRa == 0 || g->NgRb == 0 || (g->Nz > 1 && (g->NgZa == 0 || g->NgZb == 0)))
         return fail(DH_EINVAL, "mhd needs ghost annuli on every side");
   } else if (P.hydro == DH_HYDRO_GRMHD) {
      if (P.num_c != 8) return fail(DH_EINVAL, "grmhd2 needs NUM_C = 8 (got %d)", P.num_c);
      if (P.ct_mode != 2 || !P.ct)
         return fail(DH_EINVAL, "grmhd2 is supported with CT_MODE 2 and Constrained_Transport 1");
      if (P.riemann_solver < 0 || P.riemann_solver > 2)
         return fail(DH_EINVAL, "Riemann_Solver %d not supported by grmhd2", P.riemann_solver);
      if (P.isothermal) return fail(DH_EINVAL, "grmhd2 is supported with the adiabatic cons2prim");
      if (P.visc_flag) return fail(DH_EINVAL, "grmhd2 has no viscous terms (Use_Viscosity 0)");
      if (nq < 8 || nq > 9) return fail(DH_EINVAL, "NUM_C+NUM_N = %d not supported", nq);
      if (g->NgRa == 0 || g->NgRb == 0 || (g->Nz > 1 && (g->NgZa == 0 || g->NgZb == 0)))
         return fail(DH_EINVAL, "grmhd2 needs ghost annuli on every side");
      if (P.mesh_motion == 3 || P.energy_omega)
         return fail(DH_EINVAL, "grmhd2 takes Mesh_Motion 0, 1 or 4 and no energy frame");
      if (P.metric != DH_METRIC_SCHW_KS_EQ && P.metric != DH_METRIC_SCHW_KS &&
          P.metric != DH_METRIC_KERR_KS)
         return fail(DH_EINVAL, "grmhd2 runs with the schw_ks_eq_rpz, schw_ks_rpz or kerr_ks_rpz metric");
      if (P.metric == DH_METRIC_KERR_KS && !(fabs(P.metric_par3) < P.metric_par2))
         return fail(DH_EINVAL, "kerr_ks_rpz needs a spin below the mass (|a| = %g, M = %g)",
                     P.metric_par3, P.metric_par2);
   } else {
      return fail(DH_EINVAL, "hydro module %d not supported", P.hydro);
   }
   if (P.c2p_guess != DH_C2P_GUESS_PREVIOUS && P.c2p_guess != DH_C2P_GUESS_LEARNED &&
       P.c2p_guess != DH_C2P_GUESS_BETTER)
      return fail(DH_EINVAL, "bad c2p_guess %d", P.c2p_guess);
   if (P.c2p_stop != DH_C2P_STOP_REFERENCE && P.c2p_stop != DH_C2P_STOP_RESIDUAL)
      return fail(DH_EINVAL, "bad c2p_stop %d", P.c2p_stop);
   if (P.hydro != DH_HYDRO_GRMHD &&
       (P.c2p_guess != DH_C2P_GUESS_PREVIOUS || P.c2p_stop != DH_C2P_STOP_REFERENCE))
      return fail(DH_EINVAL, "c2p_guess and c2p_stop apply to grmhd2 only");
   // the rotating frame's and the nozzle's sources read the Newtonian
   // primitives (omega, v); grmhd2's are covariant four-velocities
   if ((P.rot_frame || P.nozzle_type) && P.hydro == DH_HYDRO_GRMHD)
      return fail(DH_EINVAL, "RotFrame and Nozzle_Type apply to the Newtonian modules");
   if (P.nozzle_type != 0 && P.nozzle_type != 1)
      return fail(DH_EINVAL, "Nozzle_Type %d is not a nozzle of sink.c (0 or 1)", P.nozzle_type);
   if (P.nozzle_type == 1 && !(P.nozzle_par[1] > 0.0))
      return fail(DH_EINVAL, "Nozzle_Par2 (the nozzle's radius) must be positive");
   if (P.r_periodic || P.z_periodic) return fail(DH_EINVAL, "periodic boundaries are not supported");
   if (!(P.phi_order > 0.0)) return fail(DH_EINVAL, "phi_order must be positive (PHI_ORDER is 2)");
   if (g->Nr < 1 || g->Nz < 1) return fail(DH_EINVAL, "empty grid");
   if (!(g->phi_max > 0.0)) return fail(DH_EINVAL, "phi_max must be positive");
   int bcs[4] = {P.bc_rmin, P.bc_rmax, P.bc_zmin, P.bc_zmax};
   int ngs[4] = {g->NgRa, g->NgRb, g->NgZa, g->NgZb};
   for (int s = 0; s < 4; ++s) {
      if (bcs[s] < DH_BC_NONE || bcs[s] > DH_BC_DIODE)
         return fail(DH_EINVAL, "bad boundary kind %d", bcs[s]);
      if (ngs[s] > 0 && bcs[s] == DH_BC_NONE)
         return fail(DH_EINVAL, "side %d has ghost annuli but no boundary kind", s);
   }
   long ntot = 0;
   for (int a = 0; a < g->Nr * g->Nz; ++a) {
      if (g->Np[a] < 1) return fail(DH_EINVAL, "annulus %d has %d cells", a, g->Np[a]);
      ntot += g->Np[a];
   }
   // cells, faces (about two per cell and direction) and edge-line
   // segments (about two per face) are indexed in int32
   if (ntot >= (1L << 29)) return fail(DH_EINVAL, "mesh too large (%ld cells)", ntot);
   *ntot_out = ntot;
   return DH_OK;
}

template <class X>
int alloc_n(X &x, void *pp, size_t bytes)
{
   return x.alloc((void **)pp, bytes);
}

// The strip heights of the orders of a kind whose groups have K layers:
// 4, 8 and 16 layers where they are fewer than K, and all K (k fastest
// over the whole column).
inline std::vector<int> strip_heights(int K)
{
   std::vector<int> out;
   if (K < 2) return out;
   for (int S : {4, 8, 16})
      if (S < K) out.push_back(S);
   out.push_back(K);
   return out;
}

// One group order on the device from the host lists of the visited groups'
// storage starts and sizes.
template <class X>
int put_order(X &x, GroupOrder &o, const std::vector<int> &first, const std::vector<int> &size)
{
   int ng = (int)first.size();
   std::vector<int> pos(ng + 1, 0);
   for (int p = 0; p < ng; ++p) pos[p + 1] = pos[p] + size[p];
   long n = pos[ng];
   std::vector<int> chunk((size_t)(n / ORDER_CHUNK + 1), 0);
   for (int p = 0, c = 0; c < (int)chunk.size(); ++c) {
      long t = (long)c * ORDER_CHUNK;
      while (p < ng - 1 && pos[p + 1] <= t) ++p;
      chunk[c] = p;
   }
   int *d[3] = {nullptr, nullptr, nullptr};
   DH_TRY(alloc_n(x, &d[0], sizeof(int) * (size_t)(ng > 0 ? ng : 1)));
   DH_TRY(alloc_n(x, &d[1], sizeof(int) * (size_t)(ng + 1)));
   DH_TRY(alloc_n(x, &d[2], sizeof(int) * chunk.size()));
   if (ng > 0) DH_TRY(x.put(d[0], first.data(), sizeof(int) * (size_t)ng));
   DH_TRY(x.put(d[1], pos.data(), sizeof(int) * (size_t)(ng + 1)));
   DH_TRY(x.put(d[2], chunk.data(), sizeof(int) * chunk.size()));
   o.first = d[0];
   o.pos = d[1];
   o.chunk = d[2];
   o.ng = ng;
   return DH_OK;
}

// The strip orders of every kind (OrderSet).  fI holds the host copies of
// both face lists' pair offsets.
template <class X>
int build_orders(X &x, Buffers &b, const std::vector<int> *fI)
{
   const Mesh &m = b.m;
   int Nr = m.Nr, Nz = m.Nz;
   for (int S : strip_heights(Nz)) {
      // (j, k) of the cells' groups in this strip order
      std::vector<std::pair<int, int>> seq;
      for (int k0 = 0; k0 < Nz; k0 += S)
         for (int j = 0; j < Nr; ++j)
            for (int k = k0; k < k0 + S && k < Nz; ++k) seq.push_back({j, k});
      std::vector<int> first[ORD_KINDS], size[ORD_KINDS];
      int nfr = fI[0].empty() ? 0 : fI[0].back();
      for (auto jk : seq) {
         int j = jk.first, k = jk.second;
         int a = j + Nr * k;
         first[ORD_CELLS].push_back(b.h_off[a]);
         size[ORD_CELLS].push_back(b.h_Np[a]);
         for (int d = 0; d < 2; ++d) {
            int J = d == 0 ? Nr - 1 : Nr, K = d == 0 ? Nz : Nz - 1;
            if (j >= J || k >= K || fI[d].empty()) continue;
            int P = j + J * k;
            int f0 = fI[d][P], nf = fI[d][P + 1] - fI[d][P];
            first[ORD_FACES_R + d].push_back(f0);
            size[ORD_FACES_R + d].push_back(nf);
            first[ORD_EDGES].push_back(d == 0 ? f0 : nfr + f0);
            size[ORD_EDGES].push_back(nf);
         }
      }
      for (int kind = 0; kind < ORD_KINDS; ++kind) {
         OrderSet &os = b.ord[kind];
         if (os.n >= MAX_ORDERS) return fail(DH_EINVAL, "too many traversal orders");
         DH_TRY(put_order(x, os.o[os.n], first[kind], size[kind]));
         os.strip[os.n] = S;
         ++os.n;
      }
   }
   return DH_OK;
}

// Runs k over its n elements; the executor may take any order of the set
// (results do not depend on it).
template <class X, class K>
int run_ord(X &x, const char *name, const K &k, long n, const OrderSet &os)
{
   return x.run_tuned(name, k, n, os.o, os.n, os.strip);
}

// Allocates and fills every buffer of a validated mesh.
template <class X>
int create_buffers(X &x, Buffers &b, const dh_grid *g, const dh_params &P)
{
   long ntot = 0;
   DH_TRY(validate(g, P, &ntot));
   b.P = P;
   b.nq = P.num_c + P.num_n;
   b.mhd = P.hydro == DH_HYDRO_MHD || P.hydro == DH_HYDRO_GRMHD;
   b.nf = b.mhd ? 5 : 0;
   int nq = b.nq;
   int nann = g->Nr * g->Nz;
   b.h_Np.assign(g->Np, g->Np + nann);
   b.h_off.assign(nann + 1, 0);
   for (int a = 0; a < nann; ++a) b.h_off[a + 1] = b.h_off[a] + g->Np[a];
   std::vector<int> cjk(ntot);
   for (int a = 0; a < nann; ++a)
      for (int i = 0; i < g->Np[a]; ++i) cjk[b.h_off[a] + i] = a;
   DH_TRY(alloc_n(x, &b.Np, sizeof(int) * nann));
   DH_TRY(alloc_n(x, &b.off, sizeof(int) * (nann + 1)));
   DH_TRY(alloc_n(x, &b.cjk, sizeof(int) * ntot));
   DH_TRY(alloc_n(x, &b.r, sizeof(double) * (g->Nr + 1)));
   DH_TRY(alloc_n(x, &b.z, sizeof(double) * (g->Nz + 1)));
   DH_TRY(x.put(b.Np, g->Np, sizeof(int) * nann));
   DH_TRY(x.put(b.off, b.h_off.data(), sizeof(int) * (nann + 1)));
   DH_TRY(x.put(b.cjk, cjk.data(), sizeof(int) * ntot));
   DH_TRY(x.put(b.r, g->r_jph, sizeof(double) * (g->Nr + 1)));
   DH_TRY(x.put(b.z, g->z_kph, sizeof(double) * (g->Nz + 1)));

   Mesh &m = b.m;
   m.Nr = g->Nr;
   m.Nz = g->Nz;
   m.NgRa = g->NgRa;
   m.NgRb = g->NgRb;
   m.NgZa = g->NgZa;
   m.NgZb = g->NgZb;
   m.nann = nann;
   m.ntot = (int)ntot;
   m.Np = b.Np;
   m.off = b.off;
   m.cjk = b.cjk;
   m.r_jph = b.r + 1;
   m.z_kph = b.z + 1;
   m.Pmax = g->phi_max;

   // Face lists: the pair offsets are fixed because Np never changes
   // (buildfaces mode 0, faces_alt.c:80-96).
   std::vector<int> h_fI[2];
   for (int dim = 1; dim <= 2; ++dim) {
      int d = dim - 1;
      Faces &F = b.F[d];
      memset(&F, 0, sizeof F);
      int Nrmax = dim == 2 ? g->Nr : g->Nr - 1;
      int Nzmax = dim == 1 ? g->Nz : g->Nz - 1;
      if (Nrmax < 1 || Nzmax < 1) continue;
      std::vector<int> fI(Nrmax * Nzmax + 1, 0);
      long n = 0;
      for (int k = 0; k < Nzmax; ++k)
         for (int j = 0; j < Nrmax; ++j) {
            int jk = j + g->Nr * k;
            int jkp = dim == 2 ? j + g->Nr * (k + 1) : jk + 1;
            fI[j + Nrmax * k] = (int)n;
            n += g->Np[jk] + g->Np[jkp];
         }
      if (n >= (1L << 31)) return fail(DH_EINVAL, "too many faces (%ld)", n);
      fI[Nrmax * Nzmax] = (int)n;
      h_fI[d] = fI;
      F.npairs = Nrmax * Nzmax;
      F.n = (int)n;
      int *dfI = nullptr;
      DH_TRY(alloc_n(x, &dfI, sizeof(int) * fI.size()));
      DH_TRY(x.put(dfI, fI.data(), sizeof(int) * fI.size()));
      F.fI = dfI;
      std::vector<int> pf(n);
      for (int P = 0; P < F.npairs; ++P)
         for (int f = fI[P]; f < fI[P + 1]; ++f) pf[f] = P;
      int *dpf = nullptr;
      DH_TRY(alloc_n(x, &dpf, sizeof(int) * (size_t)n));
      DH_TRY(x.put(dpf, pf.data(), sizeof(int) * (size_t)n));
      F.pf = dpf;
      DH_TRY(alloc_n(x, &F.L, sizeof(int) * n));
      DH_TRY(alloc_n(x, &F.R, sizeof(int) * n));
      DH_TRY(alloc_n(x, &F.LR, sizeof(signed char) * n));
      DH_TRY(alloc_n(x, &F.dA, sizeof(double) * n));
      DH_TRY(alloc_n(x, &F.phic, sizeof(double) * n));
      DH_TRY(alloc_n(x, &F.dphi, sizeof(double) * n));
      DH_TRY(alloc_n(x, &F.up, sizeof(double) * n));
      DH_TRY(alloc_n(x, &b.dF[d], sizeof(double) * n * nq));
      DH_TRY(alloc_n(x, &b.rng[d], sizeof(int) * 2 * (size_t)ntot));
      if (b.mhd) {
         DH_TRY(alloc_n(x, &b.emf_f[d], sizeof(double) * 3 * n));
         DH_TRY(alloc_n(x, &b.flag[d], sizeof(signed char) * n));
         DH_TRY(alloc_n(x, &b.estep[d], sizeof(int) * 2 * n));
         // edge lines: the line whose first list is pair P holds the merge
         // of P's faces with the next pair's (r faces: k + 1, z faces: j + 1)
         std::vector<int> lb(F.npairs + 1, 0);
         long nseg = 0;
         for (int P = 0; P < F.npairs; ++P) {
            lb[P] = (int)nseg;
            Pair pr = pair_of(m, dim, P);
            EdgeRole er = edge_role(m, dim, pr.j, pr.k, 1);
            if (er.ok) nseg += (fI[P + 1] - fI[P]) + (fI[er.oJK + 1] - fI[er.oJK]);
         }
         if (nseg >= (1L << 31)) return fail(DH_EINVAL, "too many edge segments (%ld)", nseg);
         lb[F.npairs] = (int)nseg;
         DH_TRY(alloc_n(x, &b.lbase[d], sizeof(int) * lb.size()));
         DH_TRY(x.put(b.lbase[d], lb.data(), sizeof(int) * lb.size()));
         DH_TRY(alloc_n(x, &b.seg[d], sizeof(double) * (size_t)(nseg > 0 ? nseg : 1)));
         DH_TRY(x.zero(b.emf_f[d], sizeof(double) * 3 * n));
      }
   }
   DH_TRY(build_orders(x, b, h_fI));
   size_t qb = sizeof(double) * (size_t)ntot * nq;
   size_t cb = sizeof(double) * (size_t)ntot;
   DH_TRY(alloc_n(x, &b.prim, qb));
   DH_TRY(alloc_n(x, &b.cons, qb));
   DH_TRY(alloc_n(x, &b.rkcons, qb));
   DH_TRY(x.zero(b.rkcons, qb));
   DH_TRY(alloc_n(x, &b.gradp, qb));
   DH_TRY(alloc_n(x, &b.gradr, 2 * qb));  // r and z slopes in one block
   b.gradz = b.gradr + (size_t)nq * ntot;
   DH_TRY(alloc_n(x, &b.dFp, qb));
   DH_TRY(alloc_n(x, &b.piph, cb));
   DH_TRY(alloc_n(x, &b.piph2, cb));
   DH_TRY(alloc_n(x, &b.dphi, cb));
   DH_TRY(alloc_n(x, &b.dphi2, cb));
   DH_TRY(alloc_n(x, &b.wiph, cb));
   DH_TRY(alloc_n(x, &b.I0, sizeof(int) * nann));
   m.I0 = b.I0;
   DH_TRY(alloc_n(x, &b.bcstart, sizeof(int) * (size_t)ntot));
   DH_TRY(x.zero(b.bcstart, sizeof(int) * (size_t)ntot));
   {
      int ngs[4] = {g->NgRa, g->NgRb, g->NgZa, g->NgZb};
      for (int side = 0; side < 4; ++side) {
         if (ngs[side] > 8) return fail(DH_EINVAL, "more than 8 ghost annuli on side %d", side);
         int lines = side < 2 ? g->Nz : g->Nr;
         for (int gl = 0; gl < ngs[side]; ++gl) {
            int mx = 0;
            for (int line = 0; line < lines; ++line) {
               int dst, src;
               bc_pair(m, side, line, gl, &dst, &src);
               if (g->Np[dst] > mx) mx = g->Np[dst];
            }
            b.bc_maxnp[side][gl] = mx;
         }
      }
   }
   DH_TRY(x.zero(b.gradr, qb));
   DH_TRY(x.zero(b.gradz, qb));
   DH_TRY(x.zero(b.gradp, qb));
   DH_TRY(x.zero(b.dFp, qb));
   DH_TRY(x.zero(b.wiph, cb));
   DH_TRY(x.zero(b.I0, sizeof(int) * nann));
   if (b.mhd) {
      DH_TRY(alloc_n(x, &b.Phi, 5 * cb));
      DH_TRY(alloc_n(x, &b.rkPhi, 5 * cb));
      DH_TRY(alloc_n(x, &b.emf_phi, 4 * cb));
      // The edge EMFs of a substep live from edges_collect to edge_segments,
      // after the last reader of the slopes (the fluxes; the update reads
      // them only for viscosity, which the MHD modules do not have) and
      // before plm_phi and the gradients of the next substep rewrite them:
      // Epre (16 per cell) lives in the r and z slopes, Eo and Eown (4 + 4)
      // in the phi slopes.
      b.Epre = b.gradr;
      b.Eo = b.gradp;
      b.Eown = b.gradp + 4 * (size_t)ntot;
      DH_TRY(alloc_n(x, &b.Ecarry, 16 * cb));
      DH_TRY(x.zero(b.Phi, 5 * cb));
      DH_TRY(x.zero(b.rkPhi, 5 * cb));
      DH_TRY(x.zero(b.emf_phi, 4 * cb));
      DH_TRY(x.zero(b.Ecarry, 16 * cb));
   }
   if (P.hydro == DH_HYDRO_GRMHD) {
      DH_TRY(alloc_n(x, &b.c2p_info, (size_t)ntot));
      DH_TRY(x.zero(b.c2p_info, (size_t)ntot));
      if (P.c2p_guess != DH_C2P_GUESS_PREVIOUS) {
         DH_TRY(alloc_n(x, &b.c2p_prm, sizeof(float) * C2PNET_NPARAM));
         DH_TRY(alloc_n(x, &b.c2p_frag, sizeof(unsigned short) * C2PNET_NFRAG));
         DH_TRY(alloc_n(x, &b.c2p_guess, sizeof(float) * 2 * (size_t)ntot));
      }
   }
   b.pl[0].n = b.pl[1].n = 0;
   DH_TRY(alloc_n(x, &b.step, sizeof(StepArgs)));
   return DH_OK;
}

template <class X>
struct Engine {
   X &x;
   Buffers &b;

   // Ring starts, cell ranges and faces of both directions, and the r/z
   // slopes when grads is set (plm.c:56-232).  With reuse the buffers
   // already hold the faces of the current positions and only the slopes
   // are formed.
   template <int NQ>
   int topology(bool grads, bool reuse = false)
   {
      const Mesh &m = b.m;
      if (!dim_active(m, 1) && !dim_active(m, 2)) return DH_OK;
      if (!reuse) DH_TRY(x.ring_start(m, b.piph, b.I0));
      for (int dim = 1; dim <= 2; ++dim) {
         int d = dim - 1;
         if (!dim_active(m, dim)) continue;
         double *grad = dim == 1 ? b.gradr : b.gradz;
         bool slopes = grads && b.P.plm != 0.0;
         if (!reuse && dim == 1) {
            DH_TRY(run_ord(x, "ranges_r", KCellRanges<1>{m, b.piph, b.rng[d]}, m.ntot, b.ord[ORD_CELLS]));
            DH_TRY(run_ord(x, "faces_r", KFaces<1>{m, b.F[d], b.rng[d], b.piph, b.dphi}, m.ntot,
                           b.ord[ORD_CELLS]));
         } else if (!reuse) {
            DH_TRY(run_ord(x, "ranges_z", KCellRanges<2>{m, b.piph, b.rng[d]}, m.ntot, b.ord[ORD_CELLS]));
            DH_TRY(run_ord(x, "faces_z", KFaces<2>{m, b.F[d], b.rng[d], b.piph, b.dphi}, m.ntot,
                           b.ord[ORD_CELLS]));
         }
         if (!grads) continue;
         if (!slopes) {
            DH_TRY(x.zero(grad, sizeof(double) * NQ * m.ntot));
            continue;
         }
         if (dim == 1) {
            KGrad<NQ, 1> kg{m, b.P.plm, b.P.num_c, b.F[d], b.rng[d], b.piph, b.dphi, b.prim,
                            b.gradp, grad};
            DH_TRY(run_ord(x, "grad_r", kg, m.ntot, b.ord[ORD_CELLS]));
         } else {
            KGrad<NQ, 2> kg{m, b.P.plm, b.P.num_c, b.F[d], b.rng[d], b.piph, b.dphi, b.prim,
                            b.gradp, grad};
            DH_TRY(run_ord(x, "grad_z", kg, m.ntot, b.ord[ORD_CELLS]));
         }
      }
      return DH_OK;
   }

   // Zero-gradient / diode ghost annuli (onestep.c:194-210): r sides, then z.
   template <int NQ>
   int boundaries()
   {
      const Mesh &m = b.m;
      const dh_params &P = b.P;
      int bcs[4] = {P.bc_rmin, P.bc_rmax, P.bc_zmin, P.bc_zmax};
      for (int side = 0; side < 4; ++side) {
         if (side >= 2 && m.Nz <= 1) continue;
         if (bcs[side] != DH_BC_ZEROGRAD && bcs[side] != DH_BC_DIODE) continue;
         int lines = bc_lines(m, side), layers = bc_layers(m, side);
         if (layers == 0) continue;
         DH_TRY(x.bc_scan(m, side, b.piph, b.dphi, b.bcstart));
         for (int g = 0; g < layers; ++g) {
            int maxnp = b.bc_maxnp[side][g];
            KBcFill<NQ> k{m, side, g, bcs[side] == DH_BC_DIODE ? 1 : 0, b.bcstart, b.prim,
                          b.piph, b.dphi, maxnp};
            DH_TRY(x.run("bc_fill", k, (long)lines * maxnp));
         }
      }
      return DH_OK;
   }

   // The conserved state and the magnetic fluxes of a substep: in, the RK
   // blend's other state, and out (the buffer that holds the substep's
   // result, which the kernels after the update read); pflip is the copy of
   // the fluxes that the face flips hand over alongside out (the reference's
   // RK_Phi), a buffer other than pout.
   struct SubState {
      double *cin, *crk, *cout, *pin, *prk, *pout, *pflip;
   };

   // One substep (onestep.c:84-228); its dt and planets are b.step's
   // entries sub.  reuse: the faces of the current positions are already
   // built (buildfaces would rebuild the same ones).
   template <int NQ, class H>
   int substep(double RK, int first_step, int sub, bool reuse, const SubState &st)
   {
      const double *dt = &b.step->dt[sub];
      constexpr bool mhd = IsMHD<H>::value;
      const Mesh &m = b.m;
      const dh_params &P = b.P;
      // On small meshes, where every launch waits mostly on its own
      // latency, launches that share no data go as one: the mesh velocity
      // (wiph) with the phi slopes (gradp), the phi with the r fluxes.
      const bool merge = x.merges_launches() && m.ntot < SMALL_MESH_CELLS;
      KWcell<NQ, H> kw{m, P, b.prim, b.wiph};
      KPlmPhi<NQ> kpl{m, P.plm, b.prim, b.dphi, b.gradp};
      if (first_step && merge && P.mesh_motion != 3) {
         DH_TRY(run_ord(x, "wcell_plm_phi", KBoth<KWcell<NQ, H>, KPlmPhi<NQ>>{kw, kpl, m.ntot},
                        2 * m.ntot, b.none));
      } else {
         if (first_step) {
            DH_TRY(run_ord(x, "wcell", kw, m.ntot, b.ord[ORD_CELLS]));
            if (P.mesh_motion == 3) DH_TRY(x.run("wcell", KWcellAvg{m, b.wiph}, m.nann));
         }
         DH_TRY(run_ord(x, "plm_phi", kpl, m.ntot, b.ord[ORD_CELLS]));
      }
      DH_TRY(topology<NQ>(true, reuse));
      KFluxPhi<NQ, H> kp{m, P, dt, b.prim, b.gradp, b.gradr, b.gradz, b.piph, b.dphi, b.wiph,
                         b.dFp, b.emf_phi};
      const bool merge_flux = merge && dim_active(m, 1);
      if (!merge_flux) DH_TRY(run_ord(x, "flux_phi", kp, m.ntot, b.ord[ORD_CELLS]));
      for (int dim = 1; dim <= 2; ++dim) {
         int d = dim - 1;
         if (!dim_active(m, dim)) continue;
         if (dim == 1) {
            KFluxTrans<NQ, H, 1> kt{m, P, dt, b.F[d], b.prim, b.gradp, b.gradr, b.gradz,
                                    b.piph, b.dphi, b.dF[d], b.emf_f[d]};
            if (merge_flux)
               DH_TRY(run_ord(x, "flux_phi_r",
                              KBoth<KFluxPhi<NQ, H>, KFluxTrans<NQ, H, 1>>{kp, kt, m.ntot},
                              m.ntot + b.F[d].n, b.none));
            else
               DH_TRY(run_ord(x, "flux_r", kt, b.F[d].n, b.ord[ORD_FACES_R]));
         } else {
            KFluxTrans<NQ, H, 2> kt{m, P, dt, b.F[d], b.prim, b.gradp, b.gradr, b.gradz,
                                    b.piph, b.dphi, b.dF[d], b.emf_f[d]};
            DH_TRY(run_ord(x, "flux_z", kt, b.F[d].n, b.ord[ORD_FACES_Z]));
         }
      }
      if (mhd) {
         int ct = P.ct;
         KEdgesCollect kc{m, ct, b.dphi, b.emf_phi, b.F[0], b.F[1], b.rng[0], b.rng[1],
                          b.emf_f[0], b.emf_f[1], b.Ecarry, b.Epre};
         DH_TRY(run_ord(x, "edges_collect", kc, m.ntot, b.ord[ORD_CELLS]));
         EdgeArgs ea{m, b.F[0], b.F[1], b.rng[0], b.rng[1], b.Epre, b.piph, b.dphi};
         KEdgesFinal kef{ea, ct, b.wiph, b.emf_f[0], b.emf_f[1], b.Eo, b.Eown, b.Ecarry};
         DH_TRY(run_ord(x, "edges_final", kef, m.ntot, b.ord[ORD_CELLS]));
         KUpdateB kb{m, RK, dt, b.F[0], b.F[1], b.rng[0], b.rng[1], b.Eo, st.pin, st.prk, st.pout};
         DH_TRY(run_ord(x, "update_B", kb, m.ntot, b.ord[ORD_CELLS]));
         if (dim_active(m, 2)) {
            if (dim_active(m, 1) && !reuse)
               DH_TRY(run_ord(x, "line_steps_r", KLineSteps<1>{m, b.F[0], b.estep[0]}, b.F[0].n,
                              b.ord[ORD_FACES_R]));
            if (!reuse)
               DH_TRY(run_ord(x, "line_steps_z", KLineSteps<2>{m, b.F[1], b.estep[1]}, b.F[1].n,
                              b.ord[ORD_FACES_Z]));
            KEdgeSegments ks{m, dt, b.F[0], b.F[1], b.estep[0], b.estep[1], b.lbase[0], b.lbase[1],
                             b.piph, b.dphi, b.Eown, b.seg[0], b.seg[1]};
            DH_TRY(run_ord(x, "edge_segments", ks, (long)b.F[0].n + b.F[1].n, b.ord[ORD_EDGES]));
            KEdgeAdjust ke{m, b.F[0], b.F[1], b.estep[0], b.estep[1], b.lbase[0], b.lbase[1],
                           b.seg[0], b.seg[1], st.pout};
            DH_TRY(run_ord(x, "edge_adjust", ke, (long)b.F[0].n + b.F[1].n, b.ord[ORD_EDGES]));
         }
      }
      KUpdate<NQ, H> ku{m, P, RK, dt, first_step, &b.step->pl[sub], b.prim, b.prim, st.cin,
                        st.cout, st.crk, b.gradr, b.gradp, b.gradz, b.piph, b.dphi, b.wiph,
                        b.piph2, b.dphi2, b.dFp, b.dF[0], b.dF[1], b.F[0].fI, b.F[1].fI,
                        b.rng[0], b.rng[1], b.F[0].n, b.F[1].n};
      DH_TRY(run_ord(x, "update", ku, m.ntot, b.ord[ORD_CELLS]));
      if (mhd) {
         if (first_step) {
            for (int dim = 1; dim <= 2; ++dim) {
               int d = dim - 1;
               if (!dim_active(m, dim)) continue;
               if (dim == 1)
                  DH_TRY(run_ord(x, "flip_flags_r",
                                 KFlipFlags<1>{m, b.F[d], b.piph, b.wiph, dt, b.flag[d]},
                                 b.F[d].n, b.ord[ORD_FACES_R]));
               else
                  DH_TRY(run_ord(x, "flip_flags_z",
                                 KFlipFlags<2>{m, b.F[d], b.piph, b.wiph, dt, b.flag[d]},
                                 b.F[d].n, b.ord[ORD_FACES_Z]));
               DH_TRY(x.flips(m, dim, b.F[d], b.flag[d], st.pout, st.pflip));
            }
         }
         // with the learned guess cons2prim waits for the model's guesses,
         // which need the cells' conserved field first
         bool learned = IsGR<H>::value && P.c2p_guess != DH_C2P_GUESS_PREVIOUS;
         KBfacesC2p<NQ, H> kbc{m, P, b.F[0], b.F[1], b.rng[0], b.rng[1], b.piph2, b.dphi2,
                               st.pout, st.cout, b.prim, 1, learned ? 0 : 1, b.c2p_info};
         DH_TRY(run_ord(x, "bfaces_c2p", kbc, m.ntot, b.ord[ORD_CELLS]));
         if constexpr (IsGR<H>::value) {
            if (learned) {
               DH_TRY(x.template c2p_guess<NQ>(m, P, st.cout, b.piph2, b.dphi2, b.c2p_prm,
                                               b.c2p_frag, b.c2p_guess));
               KC2p<NQ> kc{m, P, b.piph2, b.dphi2, st.cout, b.prim, b.c2p_guess, b.c2p_info};
               DH_TRY(run_ord(x, "c2p", kc, m.ntot, b.ord[ORD_CELLS]));
            }
         }
      }
      std::swap(b.piph, b.piph2);
      std::swap(b.dphi, b.dphi2);
      return boundaries<NQ>();
   }

   // timestep (timestep.c:7-45): RK2 unless Timestep = 1 (forward Euler).
   // The step's dt and planets go to the device first; the rest launches
   // the same kernels with the same arguments every step, so the executor
   // may replay it as a recorded graph.  An RK2 step swaps the face
   // positions twice and so leaves every buffer pointer where it was; a
   // forward-Euler step does not (then the step is never recorded).
   // The first substep starts from the positions the last RK2 step's second
   // substep left unmoved, so it takes that substep's faces instead of
   // rebuilding them; a step with nothing to take is not recorded either.
   template <int NQ, class H>
   int timestep(double dt)
   {
      if (b.P.c2p_guess != DH_C2P_GUESS_PREVIOUS && !b.c2p_model)
         return fail(DH_EINVAL, "the learned cons2prim guess has no model (dh_set_c2p_model)");
      StepArgs sa;
      memset(&sa, 0, sizeof sa);
      sa.dt[0] = dt;
      sa.dt[1] = 0.5 * dt;
      sa.pl[0] = b.pl[0];
      sa.pl[1] = b.pl[1];
      DH_TRY(x.set_step(b.step, sa));
      bool rk2 = b.P.timestep_mode != 1;
      bool reuse = b.topo_valid;
      b.topo_valid = false;
      DH_TRY(x.replayable(rk2 && reuse ? (const void *)b.piph : nullptr, [&]() -> int {
         // RK2 without copies: the first substep reads the step's start
         // state from cons (Phi) and writes its result to rkcons (rkPhi);
         // its face flips hand over both, the start state standing for the
         // reference's RK copy; the second substep blends its input with the
         // start state and writes back to cons (Phi).  Forward Euler updates
         // in place (its RK copy is never read: the flips' second array is
         // the spare rkPhi).
         if (rk2) {
            DH_TRY((substep<NQ, H>(0.0, 1, 0, reuse,
                                   {b.cons, b.cons, b.rkcons, b.Phi, b.Phi, b.rkPhi, b.Phi})));
            DH_TRY((substep<NQ, H>(0.5, 0, 1, false,
                                   {b.rkcons, b.cons, b.cons, b.rkPhi, b.Phi, b.Phi, nullptr})));
         } else {
            DH_TRY((substep<NQ, H>(0.0, 1, 0, reuse,
                                   {b.cons, b.cons, b.cons, b.Phi, b.Phi, b.Phi, b.rkPhi})));
         }
         return DH_OK;
      }));
      b.topo_valid = rk2;
      return DH_OK;
   }

   // setupCells after the initial prims are in place (domain.c:139-242):
   // cons on real cells; for MHD, cell fields from the face fluxes the host
   // set (set_B_fields' B_faces_to_cells, bfields.c:80-81) and cons2prim;
   // then face velocities and ghost annuli.
   template <int NQ, class H>
   int init()
   {
      const Mesh &m = b.m;
      b.topo_valid = false;
      DH_TRY(x.run("init", KInitCells<NQ, H>{m, b.P, b.prim, b.cons, b.piph, b.dphi}, m.ntot));
      if constexpr (IsMHD<H>::value) {
         // The reference also sets the primitive field of every cell here
         // (B_faces_to_cells type 0), but real cells then take cons2prim's
         // and ghost cells their boundary values, so no trace of it remains.
         DH_TRY(topology<NQ>(false));
         KBfacesC2p<NQ, H> k1{m, b.P, b.F[0], b.F[1], b.rng[0], b.rng[1], b.piph, b.dphi,
                              b.Phi, b.cons, b.prim, 1, 1, nullptr};
         DH_TRY(x.run("init", k1, m.ntot));
      }
      DH_TRY(x.run("init", KWcell<NQ, H>{m, b.P, b.prim, b.wiph}, m.ntot));
      if (b.P.mesh_motion == 3) DH_TRY(x.run("init", KWcellAvg{m, b.wiph}, m.nann));
      DH_TRY(boundaries<NQ>());
      return x.sync();
   }

   // getmindt (misc.c:39-84)
   template <int NQ, class H>
   int mindt(double *dt)
   {
      double v = HUGE_VAL;
      DH_TRY((x.template min_dt<NQ, H>(b.m, b.P, b.prim, b.piph, b.dphi, b.wiph, &v)));
      double d0 = b.P.max_dt / b.P.cfl;
      if (d0 <= 0.0) d0 = 1.0e100;
      if (v < d0) d0 = v;
      *dt = d0 * b.P.cfl;
      return DH_OK;
   }

   // Faces of the current mesh, with no slopes (for a host that sets the
   // initial face fluxes).
   template <int NQ, class H>
   int faces()
   {
      DH_TRY(topology<NQ>(false));
      return x.sync();
   }
};

// Dispatch on the hydro module and the number of fields (NUM_Q).
#define DH_DISPATCH(b, FN, ...)                                               \
   ([&]() -> int {                                                            \
      if ((b).P.hydro == DH_HYDRO_GRMHD) {                                    \
         switch ((b).nq) {                                                    \
         case 8: return FN<8, GRMHD<8>>(__VA_ARGS__);                         \
         case 9: return FN<9, GRMHD<9>>(__VA_ARGS__);                         \
         default: return fail(DH_EINVAL, "unsupported NQ %d", (b).nq);       \
         }                                                                    \
      }                                                                       \
      if ((b).mhd) {                                                          \
         switch ((b).nq) {                                                    \
         case 8: return FN<8, MHD<8>>(__VA_ARGS__);                           \
         case 9: return FN<9, MHD<9>>(__VA_ARGS__);                           \
         default: return fail(DH_EINVAL, "unsupported NQ %d", (b).nq);       \
         }                                                                    \
      }                                                                       \
      switch ((b).nq) {                                                       \
      case 5: return FN<5, Euler<5>>(__VA_ARGS__);                            \
      case 6: return FN<6, Euler<6>>(__VA_ARGS__);                            \
      case 7: return FN<7, Euler<7>>(__VA_ARGS__);                            \
      default: return fail(DH_EINVAL, "unsupported NQ %d", (b).nq);          \
      }                                                                       \
   }())

// Host <-> executor copies of the whole state; any pointer may be null.
// dphi is recomputed from piph (calc_dp, misc.c:271-288) when it is null.
template <class X>
int upload_state(X &x, Buffers &b, const double *prim, const double *cons, const double *piph,
                 const double *dphi, const double *wiph, const double *Phi)
{
   size_t nt = (size_t)b.m.ntot, qb = sizeof(double) * nt * b.nq;
   if (prim) DH_TRY(x.put(b.prim, prim, qb));
   if (cons) DH_TRY(x.put(b.cons, cons, qb));
   if (piph) DH_TRY(x.put(b.piph, piph, sizeof(double) * nt));
   if (piph || dphi) b.topo_valid = false;
   if (wiph) DH_TRY(x.put(b.wiph, wiph, sizeof(double) * nt));
   if (Phi && b.nf) DH_TRY(x.put(b.Phi, Phi, sizeof(double) * nt * b.nf));
   if (dphi) {
      DH_TRY(x.put(b.dphi, dphi, sizeof(double) * nt));
   } else if (piph) {
      std::vector<double> dp(nt);
      for (int a = 0; a < b.m.nann; ++a) {
         int c0 = b.h_off[a], np = b.h_Np[a];
         for (int i = 0; i < np; ++i) {
            double v = piph[c0 + i] - piph[c0 + (i == 0 ? np - 1 : i - 1)];
            while (v < 0.0) v += b.m.Pmax;
            while (v > b.m.Pmax) v -= b.m.Pmax;
            dp[c0 + i] = v;
         }
      }
      DH_TRY(x.put(b.dphi, dp.data(), sizeof(double) * nt));
   }
   return DH_OK;
}

template <class X>
int download_state(X &x, Buffers &b, double *prim, double *cons, double *piph, double *dphi,
                   double *wiph, double *Phi)
{
   size_t nt = (size_t)b.m.ntot, qb = sizeof(double) * nt * b.nq;
   if (prim) DH_TRY(x.get(prim, b.prim, qb));
   if (cons) DH_TRY(x.get(cons, b.cons, qb));
   if (piph) DH_TRY(x.get(piph, b.piph, sizeof(double) * nt));
   if (dphi) DH_TRY(x.get(dphi, b.dphi, sizeof(double) * nt));
   if (wiph) DH_TRY(x.get(wiph, b.wiph, sizeof(double) * nt));
   if (Phi && b.nf) DH_TRY(x.get(Phi, b.Phi, sizeof(double) * nt * b.nf));
   return DH_OK;
}

// add_diagnostics (analysis.c:52-106) of a shipped module into the device
// Qrz; a different kind than the one accumulated so far starts from zero.
template <class X>
int add_diagnostics(X &x, Buffers &b, int kind, double dt, const double *planets, int npl)
{
   if (!diag_count(kind)) return fail(DH_EINVAL, "unknown diagnostics kind %d", kind);
   if (kind == DH_DIAG_MHDMRI && !b.mhd)
      return fail(DH_EINVAL, "diag_mhdmri needs the MHD fields (NUM_C 8)");
   if (npl < 0 || npl > DH_MAX_PLANETS) return fail(DH_EINVAL, "npl %d out of range", npl);
   if (npl > 1 && !planets) return fail(DH_EINVAL, "null planet list");
   size_t bytes = sizeof(double) * (size_t)b.m.nann * DH_DIAG_MAXQ;
   if (!b.diagQ) {
      DH_TRY(x.alloc((void **)&b.diagQ, bytes));
      DH_TRY(x.zero(b.diagQ, bytes));
   } else if (b.diag_kind != kind) {
      DH_TRY(x.zero(b.diagQ, bytes));
   }
   b.diag_kind = kind;
   DiagPlanet p1{};
   p1.on = npl > 1;
   if (p1.on)
      for (int q = 0; q < DH_PLANET_STRIDE; ++q) p1.v[q] = planets[DH_PLANET_STRIDE + q];
   return x.diag(b.m, b.P, kind, p1, b.prim, b.piph, b.dphi, dt, b.diagQ);
}

// report (report.c:7-202) of the current state: out[16] receives report.dat's
// columns after t: Torque, Power, Fr, rho_min, rhoavg_min, PsiR, PsiI, Mass,
// Mdot, S_R, L1_rho, L1_isen, L1_B, Br2, aM, bM.  The device sums each real
// annulus; the annuli are added here in the reference's order (j outer, k
// inner), so the totals differ from its cell-by-cell sums by rounding only.
template <class X>
int report(X &x, Buffers &b, const double *planets, int npl, double *out)
{
   if (npl < 0 || npl > DH_MAX_PLANETS) return fail(DH_EINVAL, "npl %d out of range", npl);
   if (npl > 1 && !planets) return fail(DH_EINVAL, "null planet list");
   const Mesh &m = b.m;
   size_t n = (size_t)m.nann * RP_N;
   if (!b.repQ) DH_TRY(x.alloc((void **)&b.repQ, sizeof(double) * n));
   DiagPlanet p0{}, p1{};
   p1.on = npl > 1;
   if (p1.on)
      for (int q = 0; q < DH_PLANET_STRIDE; ++q) {
         p0.v[q] = planets[q];
         p1.v[q] = planets[DH_PLANET_STRIDE + q];
      }
   DH_TRY(x.report(m, b.P, p0, p1, b.mhd ? 1 : 0, b.prim, b.piph, b.dphi, b.repQ));
   std::vector<double> R(n);
   DH_TRY(x.get(R.data(), b.repQ, sizeof(double) * n));
   double S[RP_N] = {0.0}, rho_min = HUGE_VAL, rhoavg_min = HUGE_VAL;
   for (int j = m.NgRa; j < m.Nr - m.NgRb; ++j) {
      double mass = 0.0, vol = 0.0;
      for (int k = m.NgZa; k < m.Nz - m.NgZb; ++k) {
         const double *r = &R[(size_t)(j + m.Nr * k) * RP_N];
         for (int q = 0; q < RP_RHOMIN; ++q) S[q] += r[q];
         mass += r[RP_MASS];
         vol += r[RP_VOL];
         if (rho_min > r[RP_RHOMIN]) rho_min = r[RP_RHOMIN];
      }
      double rho_avg = mass / vol;
      if (rhoavg_min > rho_avg) rhoavg_min = rho_avg;
   }
   double vol = S[RP_VOL];
   double v[16] = {S[RP_TORQUE], S[RP_POWER], S[RP_FR], rho_min, rhoavg_min, S[RP_PSIR],
                   S[RP_PSII], S[RP_MASS], S[RP_MDOT] / vol, S[RP_SR] / S[RP_S0],
                   S[RP_L1RHO] / vol, S[RP_L1ISEN] / vol, S[RP_L1B] / vol, S[RP_BR2],
                   S[RP_BRBP] / S[RP_PDV], S[RP_PDV] / S[RP_B2]};
   for (int q = 0; q < 16; ++q) out[q] = v[q];
   return DH_OK;
}

template <class X>
int get_diagnostics(X &x, Buffers &b, double *Qrz)
{
   int nq = diag_count(b.diag_kind);
   size_t n = (size_t)b.m.nann * nq;
   if (!b.diagQ || !nq) {
      for (size_t i = 0; i < n; ++i) Qrz[i] = 0.0;
      return DH_OK;
   }
   return x.get(Qrz, b.diagQ, sizeof(double) * n);
}

template <class X>
int avg_diagnostics(X &x, Buffers &b, double t_avg)
{
   if (!b.diagQ) return DH_OK;
   long n = (long)b.m.nann * diag_count(b.diag_kind);
   return x.run("diag_avg", KDiagAvg{b.diagQ, t_avg}, n);
}

template <class X>
int zero_diagnostics(X &x, Buffers &b)
{
   if (!b.diagQ) return DH_OK;
   return x.zero(b.diagQ, sizeof(double) * (size_t)b.m.nann * DH_DIAG_MAXQ);
}

// Planets seen by the sources of one RK substep.
inline int set_planets(Buffers &b, int substep, const double *pl, int npl)
{
   if (substep < 0 || substep > 1) return fail(DH_EINVAL, "substep must be 0 or 1");
   if (npl < 0 || npl > DH_MAX_PLANETS) return fail(DH_EINVAL, "npl %d out of range", npl);
   if (npl > 0 && !pl) return fail(DH_EINVAL, "null planet list");
   PlanetSet &s = b.pl[substep];
   s.n = 0;
   for (int p = 0; p < npl; ++p) {
      const double *v = pl + p * DH_PLANET_STRIDE;
      // A planet of zero mass adds exactly zero force (planet.c:56-76).
      if (v[2] == 0.0 && ((int)v[6] == 0 || (int)v[6] == 1)) continue;
      for (int q = 0; q < DH_PLANET_STRIDE; ++q) s.v[s.n * DH_PLANET_STRIDE + q] = v[q];
      s.n++;
   }
   return DH_OK;
}

// Host-side face description for the host layer (faces_alt.c:10-34 addFace):
// cells, orientation, centroid {r, phi, z} and area of every face.
template <class X>
int export_faces(X &x, Buffers &b, int dim, int *L, int *R, int *LRtype, double *cm, double *dA)
{
   int d = dim - 1;
   const Faces &F = b.F[d];
   size_t n = (size_t)F.n;
   if (n == 0) return DH_OK;
   std::vector<int> hL(n), hR(n), fI(F.npairs + 1);
   std::vector<signed char> hLR(n);
   std::vector<double> hphic(n), hdA(n), rj(b.m.Nr + 1), zk(b.m.Nz + 1);
   DH_TRY(x.get(hL.data(), F.L, sizeof(int) * n));
   DH_TRY(x.get(hR.data(), F.R, sizeof(int) * n));
   DH_TRY(x.get(hLR.data(), F.LR, n));
   DH_TRY(x.get(hphic.data(), F.phic, sizeof(double) * n));
   DH_TRY(x.get(hdA.data(), F.dA, sizeof(double) * n));
   DH_TRY(x.get(fI.data(), F.fI, sizeof(int) * fI.size()));
   DH_TRY(x.get(rj.data(), b.r, sizeof(double) * rj.size()));
   DH_TRY(x.get(zk.data(), b.z, sizeof(double) * zk.size()));
   const double *r_jph = rj.data() + 1, *z_kph = zk.data() + 1;
   int Nrmax = dim == 2 ? b.m.Nr : b.m.Nr - 1;
   for (int P = 0; P < F.npairs; ++P) {
      int j = P % Nrmax, k = P / Nrmax;
      double r, z;
      if (dim == 1) {
         r = .5 * (r_jph[j] + r_jph[j]);
         z = .5 * (z_kph[k] + z_kph[k - 1]);
      } else {
         double xp = r_jph[j], xm = r_jph[j - 1];
         r = 2.0 * (xp * xp + xp * xm + xm * xm) / (3.0 * (xp + xm));
         z = .5 * (z_kph[k] + z_kph[k]);
      }
      for (int f = fI[P]; f < fI[P + 1]; ++f) {
         if (L) L[f] = hL[f];
         if (R) R[f] = hR[f];
         if (LRtype) LRtype[f] = hLR[f];
         if (cm) {
            cm[3 * (size_t)f] = r;
            cm[3 * (size_t)f + 1] = hphic[f];
            cm[3 * (size_t)f + 2] = z;
         }
         if (dA) dA[f] = hdA[f];
      }
   }
   return DH_OK;
}

// The learned cons2prim guess's parameters (dh_set_c2p_model): validated,
// kept in float32 for the biases and normalisation and as bfloat16 MFMA
// fragments for the weights.
template <class X>
int set_c2p_model(X &x, Buffers &b, const float *prm, long n)
{
   if (b.P.hydro != DH_HYDRO_GRMHD || b.P.c2p_guess == DH_C2P_GUESS_PREVIOUS)
      return fail(DH_EINVAL, "the context was not created with a learned c2p_guess");
   if (!prm || n != C2PNET_NPARAM)
      return fail(DH_EINVAL, "a c2p model has %d parameters (got %ld)", (int)C2PNET_NPARAM, n);
   for (long i = 0; i < n; ++i)
      if (!std::isfinite(prm[i])) return fail(DH_EINVAL, "c2p model parameter %ld is not finite", i);
   for (int i = 0; i < C2PNET_IN; ++i)
      if (!(prm[C2PNET_IN_SD + i] != 0.0f)) return fail(DH_EINVAL, "c2p model in_sd[%d] is zero", i);
   std::vector<unsigned short> frag(C2PNET_NFRAG);
   c2pnet_fragments(prm, frag.data());
   DH_TRY(x.put(b.c2p_prm, prm, sizeof(float) * C2PNET_NPARAM));
   DH_TRY(x.put(b.c2p_frag, frag.data(), sizeof(unsigned short) * C2PNET_NFRAG));
   b.c2p_model = true;
   return DH_OK;
}

// Training data of the learned guess from the current state (KC2pHarvest):
// feat[8 * ntot], targ[2 * ntot] host arrays, NaN outside the real cells.
template <class X>
int c2p_harvest(X &x, Buffers &b, float *feat, float *targ)
{
   if (b.P.hydro != DH_HYDRO_GRMHD) return fail(DH_EINVAL, "c2p harvest needs grmhd2");
   if (!feat || !targ) return fail(DH_EINVAL, "null argument");
   size_t nt = (size_t)b.m.ntot;
   float *df = nullptr, *dt = nullptr;
   DH_TRY(alloc_n(x, &df, sizeof(float) * C2PNET_IN * nt));
   DH_TRY(alloc_n(x, &dt, sizeof(float) * 2 * nt));
   int rc = DH_OK;
   if (b.nq == 8)
      rc = x.run("c2p_harvest", KC2pHarvest<8>{b.m, b.P, b.cons, b.prim, b.piph, b.dphi, df, dt}, b.m.ntot);
   else
      rc = x.run("c2p_harvest", KC2pHarvest<9>{b.m, b.P, b.cons, b.prim, b.piph, b.dphi, df, dt}, b.m.ntot);
   if (!rc) rc = x.get(feat, df, sizeof(float) * C2PNET_IN * nt);
   if (!rc) rc = x.get(targ, dt, sizeof(float) * 2 * nt);
   int rf = x.release(df), rt = x.release(dt);
   return rc ? rc : (rf ? rf : rt);
}

// The model's guesses (v^2, eta) for the current state, guess[2 * ntot]
// (NaN outside the real cells): the values cons2prim would start from.
template <class X>
int c2p_predict(X &x, Buffers &b, float *guess)
{
   if (!b.c2p_model) return fail(DH_EINVAL, "no learned cons2prim model (dh_set_c2p_model)");
   if (!guess) return fail(DH_EINVAL, "null argument");
   if (b.nq == 8)
      DH_TRY(x.template c2p_guess<8>(b.m, b.P, b.cons, b.piph, b.dphi, b.c2p_prm, b.c2p_frag,
                                     b.c2p_guess));
   else
      DH_TRY(x.template c2p_guess<9>(b.m, b.P, b.cons, b.piph, b.dphi, b.c2p_prm, b.c2p_frag,
                                     b.c2p_guess));
   return x.get(guess, b.c2p_guess, sizeof(float) * 2 * (size_t)b.m.ntot);
}

// Totals over the real cells of the last cons2prim's records: {cells,
// Newton iterations, cells solved from the learned guess, cells that fell
// back to the previous primitives}.
template <class X>
int c2p_stats(X &x, Buffers &b, long long *out)
{
   if (b.P.hydro != DH_HYDRO_GRMHD) return fail(DH_EINVAL, "c2p statistics need grmhd2");
   if (!out) return fail(DH_EINVAL, "null argument");
   std::vector<unsigned char> info((size_t)b.m.ntot);
   DH_TRY(x.get(info.data(), b.c2p_info, info.size()));
   long long s[4] = {0, 0, 0, 0};
   for (int a = 0; a < b.m.nann; ++a) {
      if (!is_real(b.m, a % b.m.Nr, a / b.m.Nr)) continue;
      for (int c = b.h_off[a]; c < b.h_off[a + 1]; ++c) {
         unsigned char v = info[c];
         s[0]++;
         s[1] += v & C2P_INFO_ITERS;
         s[2] += (v & C2P_INFO_LEARNED) ? 1 : 0;
         s[3] += (v & C2P_INFO_FALLBACK) ? 1 : 0;
      }
   }
   for (int i = 0; i < 4; ++i) out[i] = s[i];
   return DH_OK;
}

// The last cons2prim's record of every cell (C2P_INFO_*; 0 outside the
// real cells).
template <class X>
int c2p_info(X &x, Buffers &b, unsigned char *out)
{
   if (b.P.hydro != DH_HYDRO_GRMHD) return fail(DH_EINVAL, "c2p records need grmhd2");
   if (!out) return fail(DH_EINVAL, "null argument");
   std::vector<unsigned char> info((size_t)b.m.ntot);
   DH_TRY(x.get(info.data(), b.c2p_info, info.size()));
   for (int a = 0; a < b.m.nann; ++a) {
      bool real = is_real(b.m, a % b.m.Nr, a / b.m.Nr);
      for (int c = b.h_off[a]; c < b.h_off[a + 1]; ++c) out[c] = real ? info[c] : 0;
   }
   return DH_OK;
}

// splitmix64 (Steele, Lea & Flood 2014): the seeded stream of the training
// set-up (initial weights, batch order).
struct SplitMix {
   unsigned long long s;
   unsigned long long next()
   {
      unsigned long long z = (s += 0x9E3779B97F4A7C15ULL);
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
      return z ^ (z >> 31);
   }
   double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

// Normalisation of a training set: column means and standard deviations of
// the features and targets (a constant column gets 1).
inline void c2p_normalisation(const float *Xh, const float *Yh, long n, float *prm)
{
   for (int c = 0; c < C2PNET_IN + C2PNET_OUT; ++c) {
      bool in = c < C2PNET_IN;
      int w = in ? C2PNET_IN : C2PNET_OUT, k = in ? c : c - C2PNET_IN;
      const float *a = in ? Xh : Yh;
      double s = 0.0, s2 = 0.0;
      for (long i = 0; i < n; ++i) s += a[(size_t)i * w + k];
      double mu = n ? s / n : 0.0;
      for (long i = 0; i < n; ++i) {
         double d = a[(size_t)i * w + k] - mu;
         s2 += d * d;
      }
      double sd = n ? sqrt(s2 / n) : 1.0;
      float fsd = (float)sd;
      if (!(fsd > (in ? 1e-6f : 1e-12f))) fsd = 1.0f;
      prm[(in ? C2PNET_IN_MU : C2PNET_OUT_MU) + k] = (float)mu;
      prm[(in ? C2PNET_IN_SD : C2PNET_OUT_SD) + k] = fsd;
   }
}

// The gradient of the mean loss over all n rows (host arrays) with respect
// to the parameter block (zero at the normalisation constants), and the
// loss.
template <class X>
int c2p_grad(X &x, const float *prm, const float *Xh, const float *Yh, long n, float *grad_out,
             float *loss_out)
{
   if (!prm || !Xh || !Yh || !grad_out || n < 1 || n > (1L << 30))
      return fail(DH_EINVAL, "bad argument");
   float *dX, *dY, *dp, *act, *loss, *g;
   int *idx;
   DH_TRY(alloc_n(x, &dX, sizeof(float) * C2PNET_IN * n));
   DH_TRY(alloc_n(x, &dY, sizeof(float) * C2PNET_OUT * n));
   DH_TRY(alloc_n(x, &dp, sizeof(float) * C2PNET_NPARAM));
   DH_TRY(alloc_n(x, &idx, sizeof(int) * n));
   DH_TRY(alloc_n(x, &act, sizeof(float) * C2PNET_ROW * n));
   DH_TRY(alloc_n(x, &loss, sizeof(float) * n));
   DH_TRY(alloc_n(x, &g, sizeof(float) * C2PNET_NTRAIN));
   std::vector<int> hidx(n);
   for (long i = 0; i < n; ++i) hidx[i] = (int)i;
   DH_TRY(x.put(dX, Xh, sizeof(float) * C2PNET_IN * n));
   DH_TRY(x.put(dY, Yh, sizeof(float) * C2PNET_OUT * n));
   DH_TRY(x.put(dp, prm, sizeof(float) * C2PNET_NPARAM));
   DH_TRY(x.put(idx, hidx.data(), sizeof(int) * n));
   MlpBatch B{dp, dX, dY, idx, (int)n, act, loss};
   DH_TRY(x.run("c2p_row", KMlpRow{B}, n));
   DH_TRY(x.run("c2p_grad", KMlpGrad{B, g}, C2PNET_NTRAIN));
   std::vector<float> hg(C2PNET_NTRAIN), hl(n);
   DH_TRY(x.get(hg.data(), g, sizeof(float) * C2PNET_NTRAIN));
   DH_TRY(x.get(hl.data(), loss, sizeof(float) * n));
   for (int q = 0; q < C2PNET_NPARAM; ++q)
      grad_out[q] = (q >= C2PNET_TRAIN0 && q < C2PNET_TRAIN0 + C2PNET_NTRAIN)
                        ? hg[q - C2PNET_TRAIN0] : 0.0f;
   if (loss_out) {
      double s = 0.0;
      for (long i = 0; i < n; ++i) s += hl[i];
      *loss_out = (float)(s / (2.0 * n));
   }
   return DH_OK;
}

// Trains the learned guess on (X, Y) (host arrays, n rows): normalisation
// from the data; weights and biases drawn uniformly in +-1/sqrt(fan_in)
// from the seed; Adam (0.9, 0.999, 1e-8) at a rate decayed over the epochs
// as lr (1 + cos(pi e / epochs)) / 2; batches of a seeded permutation of
// the rows each epoch.  Writes the parameter block and the mean loss of the
// last epoch.
template <class X>
int c2p_train(X &x, const float *Xh, const float *Yh, long n, int epochs, int batch, float lr,
              unsigned long long seed, float *prm_out, float *loss_out)
{
   if (!Xh || !Yh || !prm_out || n < 1 || n > (1L << 30) || epochs < 1 || batch < 1)
      return fail(DH_EINVAL, "bad argument");
   if (batch > n) batch = (int)n;
   std::vector<float> prm(C2PNET_NPARAM, 0.0f);
   c2p_normalisation(Xh, Yh, n, prm.data());
   SplitMix rng{seed};
   auto init = [&](int w, int b, int fan_in, int count) {
      double bound = 1.0 / sqrt((double)fan_in);
      for (int i = 0; i < count; ++i) prm[w + i] = (float)((2.0 * rng.uniform() - 1.0) * bound);
      for (int i = 0; i < count / fan_in; ++i) prm[b + i] = (float)((2.0 * rng.uniform() - 1.0) * bound);
   };
   init(C2PNET_W1, C2PNET_B1, C2PNET_IN, C2PNET_IN * C2PNET_H);
   init(C2PNET_W2, C2PNET_B2, C2PNET_H, C2PNET_H * C2PNET_H);
   init(C2PNET_W3, C2PNET_B3, C2PNET_H, C2PNET_H * C2PNET_OUT);
   float *dX, *dY, *dp, *m, *v, *g, *act, *loss, *eloss;
   int *idx;
   DH_TRY(alloc_n(x, &dX, sizeof(float) * C2PNET_IN * n));
   DH_TRY(alloc_n(x, &dY, sizeof(float) * C2PNET_OUT * n));
   DH_TRY(alloc_n(x, &dp, sizeof(float) * C2PNET_NPARAM));
   DH_TRY(alloc_n(x, &m, sizeof(float) * C2PNET_NTRAIN));
   DH_TRY(alloc_n(x, &v, sizeof(float) * C2PNET_NTRAIN));
   DH_TRY(alloc_n(x, &g, sizeof(float) * C2PNET_NTRAIN));
   DH_TRY(alloc_n(x, &act, sizeof(float) * C2PNET_ROW * batch));
   DH_TRY(alloc_n(x, &loss, sizeof(float) * batch));
   DH_TRY(alloc_n(x, &eloss, sizeof(float)));
   DH_TRY(alloc_n(x, &idx, sizeof(int) * n));
   DH_TRY(x.put(dX, Xh, sizeof(float) * C2PNET_IN * n));
   DH_TRY(x.put(dY, Yh, sizeof(float) * C2PNET_OUT * n));
   DH_TRY(x.put(dp, prm.data(), sizeof(float) * C2PNET_NPARAM));
   DH_TRY(x.zero(m, sizeof(float) * C2PNET_NTRAIN));
   DH_TRY(x.zero(v, sizeof(float) * C2PNET_NTRAIN));
   std::vector<int> perm(n);
   for (long i = 0; i < n; ++i) perm[i] = (int)i;
   const float b1 = 0.9f, b2 = 0.999f;
   double p1 = 1.0, p2 = 1.0;
   float last = 0.0f;
   for (int e = 0; e < epochs; ++e) {
      for (long i = n - 1; i > 0; --i) {
         long j = (long)(rng.next() % (unsigned long long)(i + 1));
         int t = perm[i];
         perm[i] = perm[j];
         perm[j] = t;
      }
      DH_TRY(x.put(idx, perm.data(), sizeof(int) * n));
      DH_TRY(x.zero(eloss, sizeof(float)));
      float lre = (float)(lr * 0.5 * (1.0 + cos(M_PI * e / epochs)));
      for (long s0 = 0; s0 < n; s0 += batch) {
         int nb = (int)(n - s0 < batch ? n - s0 : batch);
         MlpBatch B{dp, dX, dY, idx + s0, nb, act, loss};
         p1 *= b1;
         p2 *= b2;
         DH_TRY(x.run("c2p_row", KMlpRow{B}, nb));
         DH_TRY(x.run("c2p_grad", KMlpGrad{B, g}, C2PNET_NTRAIN));
         DH_TRY(x.run("c2p_adam", KAdam{dp, m, v, g, lre, b1, b2, 1e-8f, (float)(1.0 - p1),
                                        (float)(1.0 - p2)}, C2PNET_NTRAIN));
         DH_TRY(x.run("c2p_loss", KLossSum{loss, nb, eloss}, 1));
      }
      DH_TRY(x.get(&last, eloss, sizeof(float)));
      last = (float)(last / (2.0 * n));
   }
   DH_TRY(x.get(prm_out, dp, sizeof(float) * C2PNET_NPARAM));
   if (loss_out) *loss_out = last;
   return DH_OK;
}

}  // namespace dh
