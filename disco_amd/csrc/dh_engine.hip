// dh_engine.hip -- MI355X (gfx950) timestep engine and the dh_* C ABI.
//
// One RK2 timestep of the reference (timestep.c:7-45 -> onestep.c:84-228) is
// a short sequence of kernels over the structure-of-arrays state.  Scatter
// loops of the reference (face fluxes added to both neighbours, slope sums)
// become owner-computes gathers: faces are computed once into face-indexed
// buffers, and each cell sums its own faces in the reference's order.  No
// atomics touch the state, so a step is deterministic.
//
// Substep pipeline (one launch each):
//   k_wcell        set_wcell, first substep only       (misc.c:101-168)
//   k_plm_phi      phi slopes, all cells               (plm.c:15-51)
//   k_ring_start   start cell of every ring            (faces_alt.c:59-78)
//   k_face_slopes  r|z faces + face slopes             (faces_alt.c:36-179, plm.c:84-110)
//   k_grad         r|z slopes: gather, normalise, limit, axis fix (plm.c:112-232)
//   k_flux_phi     phi-face Riemann fluxes              (misc.c:368-426, riemann.c:39-112)
//   k_flux_trans   r|z-face Riemann fluxes              (misc.c:431-509, riemann.c:114-213)
//   k_update       RK blend + flux gather + sources + mesh motion + cons2prim
//                  (misc.c:226-244, 538-580, 256-325)
//   k_bc_copy      zero-gradient ghost annuli           (boundary_functions.c:90-159, 269-441)
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "disco_hip.h"
#include "dh_euler.h"
#include "dh_geom.h"
#include "dh_mesh.h"

using namespace dh;

// ---------------------------------------------------------------------------
// Errors
// ---------------------------------------------------------------------------

static thread_local char g_err[512];

static int fail(int code, const char *fmt, ...)
{
   va_list ap;
   va_start(ap, fmt);
   vsnprintf(g_err, sizeof g_err, fmt, ap);
   va_end(ap);
   return code;
}

#define HIPCHK(x)                                                              \
   do {                                                                        \
      hipError_t e_ = (x);                                                     \
      if (e_ != hipSuccess)                                                    \
         return fail(DH_EHIP, "%s failed: %s", #x, hipGetErrorString(e_));     \
   } while (0)

static constexpr int TPB = 256;
static inline int nblocks(long n) { return (int)((n + TPB - 1) / TPB); }

struct PlanetSet {
   double v[DH_MAX_PLANETS * DH_PLANET_STRIDE];
   int n;
};

// ---------------------------------------------------------------------------
// Small device helpers
// ---------------------------------------------------------------------------

DHD bool is_real(const Mesh &m, int j, int k)
{
   return j >= m.NgRa && j < m.Nr - m.NgRb && k >= m.NgZa && k < m.Nz - m.NgZb;
}

template <int NQ>
DHD void load_q(const double *soa, size_t nt, int c, double *v)
{
#pragma unroll
   for (int q = 0; q < NQ; ++q) v[q] = soa[q * nt + c];
}

// Element index of cell c (i-th of its annulus) in its ring.
DHD int ring_elem(const Ring &R, int i)
{
   int x = i - R.I0;
   return x < 0 ? x + R.N : x;
}

// ---------------------------------------------------------------------------
// k_wcell: face angular velocity wiph (misc.c:101-168)
// ---------------------------------------------------------------------------

template <int NQ, class H>
__global__ void k_wcell(Mesh m, dh_params P, const double *prim, double *wiph)
{
   int c = blockIdx.x * blockDim.x + threadIdx.x;
   if (c >= m.ntot) return;
   int a = m.cjk[c];
   double w = 0.0;
   if (P.mesh_motion == 4) {
      int j = a % m.Nr;
      w = mesh_om(P, centroid_r(m.r_jph[j], m.r_jph[j - 1]));
   } else if (P.mesh_motion) {
      size_t nt = (size_t)m.ntot;
      int c0 = m.off[a], np = m.Np[a];
      int cp = (c - c0 + 1 == np) ? c0 : c + 1;
      double pc[NQ], pp[NQ];
      load_q<NQ>(prim, nt, c, pc);
      load_q<NQ>(prim, nt, cp, pp);
      w = .5 * (H::omega(pc) + H::omega(pp));
   }
   wiph[c] = w;
}

// Mesh_Motion 3: every face of an annulus moves with the annulus mean.
__global__ void k_wcell_avg(Mesh m, double *wiph)
{
   int a = blockIdx.x * blockDim.x + threadIdx.x;
   if (a >= m.nann) return;
   int c0 = m.off[a], np = m.Np[a];
   double w = 0.0;
   for (int i = 0; i < np; ++i) w += wiph[c0 + i];
   w /= (double)np;
   for (int i = 0; i < np; ++i) wiph[c0 + i] = w;
}

// ---------------------------------------------------------------------------
// k_plm_phi: generalised-minmod slope along phi (plm.c:6-51)
// ---------------------------------------------------------------------------

DHD double minmod(double a, double b, double c)
{
   double m = a;
   if (a * b < 0.0) m = 0.0;
   if (fabs(b) < fabs(m)) m = b;
   if (b * c < 0.0) m = 0.0;
   if (fabs(c) < fabs(m)) m = c;
   return m;
}

template <int NQ>
__global__ void k_plm_phi(Mesh m, double PLM, const double *prim, const double *dphi,
                          double *gradp)
{
   int c = blockIdx.x * blockDim.x + threadIdx.x;
   if (c >= m.ntot) return;
   size_t nt = (size_t)m.ntot;
   int a = m.cjk[c];
   int c0 = m.off[a], np = m.Np[a];
   int i = c - c0;
   int cL = i == 0 ? c0 + np - 1 : c - 1;
   int cR = i == np - 1 ? c0 : c + 1;
   double dpL = dphi[cL], dpC = dphi[c], dpR = dphi[cR];
#pragma unroll
   for (int q = 0; q < NQ; ++q) {
      double pL = prim[q * nt + cL];
      double pC = prim[q * nt + c];
      double pR = prim[q * nt + cR];
      double sL = pC - pL;
      sL /= .5 * (dpC + dpL);
      double sR = pR - pC;
      sR /= .5 * (dpR + dpC);
      double sC = pR - pL;
      sC /= .5 * (dpL + dpR) + dpC;
      gradp[q * nt + c] = minmod(PLM * sL, sC, PLM * sR);
   }
}

// ---------------------------------------------------------------------------
// k_ring_start: per annulus, the first cell (from i = 0) in the first
// quadrant that follows a cell in the fourth (faces_alt.c:59-78).  One
// wavefront per annulus scans 64 cells at a time with ballots.
// ---------------------------------------------------------------------------

__global__ void k_ring_start(Mesh m, const double *piph, int *I0)
{
   int a = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
   int lane = threadIdx.x & 63;
   if (a >= m.nann) return;
   int c0 = m.off[a], np = m.Np[a];
   double convert = 2. * M_PI / m.Pmax;
   bool seen4 = false;
   int found = -1;
   for (int base = 0; base < np; base += 64) {
      int i = base + lane;
      int quad = 0;
      if (i < np) {
         double sn, cs;
         sincos(piph[c0 + i] * convert, &sn, &cs);
         if (sn > 0. && cs > 0.) quad = 1;
         else if (sn < 0. && cs > 0.) quad = 4;
      }
      unsigned long long b4 = __ballot(quad == 4);
      unsigned long long b1 = __ballot(quad == 1);
      unsigned long long cand = 0;
      if (seen4) {
         cand = b1;
      } else if (b4) {
         int f4 = __ffsll((long long)b4) - 1;
         cand = f4 >= 63 ? 0ULL : (b1 & (~0ULL << (f4 + 1)));
      }
      if (cand) {
         found = base + __ffsll((long long)cand) - 1;
         break;
      }
      if (b4) seen4 = true;
   }
   if (lane == 0) I0[a] = found < 0 ? 0 : found;
}

// ---------------------------------------------------------------------------
// Face geometry of the pair's d-th face (faces_alt.c:10-34, 135-175)
// ---------------------------------------------------------------------------

struct FaceGeo {
   int cL, cR, LR;
   double phic, dp, dA, r, z;
};

DHD FaceGeo face_geometry(const Mesh &m, int dim, const Pair &pr, const Ring &A,
                          const Ring &B, int d, const double *piph, const double *dphi)
{
   double Pmax = m.Pmax;
   int x = merge_x(A, B, d, Pmax);
   FaceGeo g;
   g.cL = ring_cell(A, x);
   g.cR = ring_cell(B, d - x);
   double pL = piph[g.cL], pR = piph[g.cR];
   double dL = dphi[g.cL], dR = dphi[g.cR];
   double w = wrap_half(pL - pR, Pmax);
   g.LR = w > 0. ? 1 : 0;
   w = wrap_half(pL - dL - pR + dR, Pmax);
   int LRb = w < 0. ? 1 : 0;
   double xp1 = g.LR == 0 ? pL : pR;
   double xm1 = LRb == 0 ? pL - dL : pR - dR;
   g.dp = get_dp(xp1, xm1, Pmax);
   g.phic = get_dp(xp1, .5 * g.dp, Pmax);
   int j = pr.j, k = pr.k;
   if (dim == 1) {
      double r = .5 * (m.r_jph[j] + m.r_jph[j]);
      double dz = m.z_kph[k] - m.z_kph[k - 1];
      g.dA = r * g.dp * dz;
      g.r = r;
      g.z = centroid_z(m.z_kph[k], m.z_kph[k - 1]);
   } else {
      double r = .5 * (m.r_jph[j] + m.r_jph[j - 1]);
      double dr = m.r_jph[j] - m.r_jph[j - 1];
      g.dA = r * dr * g.dp;
      g.r = centroid_r(m.r_jph[j], m.r_jph[j - 1]);
      g.z = .5 * (m.z_kph[k] + m.z_kph[k]);
   }
   return g;
}

// dxL, dxR: distances from the cell centroids to the face (faces_alt.c:107-123)
DHD void pair_dx(const Mesh &m, int dim, const Pair &pr, double *dxL, double *dxR)
{
   int j = pr.j, k = pr.k;
   if (dim == 1) {
      double rm = centroid_r(m.r_jph[j], m.r_jph[j - 1]);
      double rp = centroid_r(m.r_jph[j + 1], m.r_jph[j]);
      *dxL = m.r_jph[j] - rm;
      *dxR = rp - m.r_jph[j];
   } else {
      double zm = centroid_z(m.z_kph[k], m.z_kph[k - 1]);
      double zp = centroid_z(m.z_kph[k + 1], m.z_kph[k]);
      *dxL = m.z_kph[k] - zm;
      *dxR = zp - m.z_kph[k];
   }
}

// ---------------------------------------------------------------------------
// k_face_slopes: every r (dim 1) or z (dim 2) face, with its one-sided slope
// S = (W_R - W_L)/(dx_L + dx_R) from phi-reconstructed states (plm.c:84-110).
// ---------------------------------------------------------------------------

template <int NQ>
__global__ void k_face_slopes(Mesh m, int dim, const int *fI, int npairs, int nfaces,
                              const int *I0, const double *piph, const double *dphi,
                              const double *prim, const double *gradp, int *fL, int *fR,
                              double *fdA, double *fphic, double *S)
{
   int f = blockIdx.x * blockDim.x + threadIdx.x;
   if (f >= nfaces) return;
   size_t nt = (size_t)m.ntot, nfs = (size_t)nfaces;
   int P = pair_of_face(fI, npairs, f);
   Pair pr = pair_of(m, dim, P);
   Ring A = make_ring(piph, m, I0, pr.a);
   Ring B = make_ring(piph, m, I0, pr.b);
   FaceGeo g = face_geometry(m, dim, pr, A, B, f - fI[P], piph, dphi);
   fL[f] = g.cL;
   fR[f] = g.cR;
   fdA[f] = g.dA;
   fphic[f] = g.phic;
   double dxL, dxR;
   pair_dx(m, dim, pr, &dxL, &dxR);
   double phiL = piph[g.cL] - .5 * dphi[g.cL];
   double phiR = piph[g.cR] - .5 * dphi[g.cR];
   double dpL = get_sdp(g.phic, phiL, m.Pmax);
   double dpR = get_sdp(phiR, g.phic, m.Pmax);
#pragma unroll
   for (int q = 0; q < NQ; ++q) {
      double WL = prim[q * nt + g.cL] + dpL * gradp[q * nt + g.cL];
      double WR = prim[q * nt + g.cR] - dpR * gradp[q * nt + g.cR];
      S[q * nfs + f] = (WR - WL) / (dxR + dxL);
   }
}

// Inner/outer neighbour pairs of a cell for dim (1: r, 2: z).
struct CellPairs {
   bool in, out;
   int fbase_in, fbase_out;
   FaceRange rin, rout;
};

DHD CellPairs cell_pairs(const Mesh &m, int dim, int a, int i, const int *fI,
                         const int *I0, const double *piph)
{
   CellPairs cp;
   int j = a % m.Nr, k = a / m.Nr;
   Ring self = make_ring(piph, m, I0, a);
   int xs = ring_elem(self, i);
   if (dim == 1) {
      cp.in = j > 0;
      cp.out = j < m.Nr - 1;
   } else {
      cp.in = k > 0;
      cp.out = k < m.Nz - 1;
   }
   if (cp.in) {
      int ain = dim == 1 ? a - 1 : a - m.Nr;
      int JK = dim == 1 ? (j - 1) + (m.Nr - 1) * k : j + m.Nr * (k - 1);
      Ring A = make_ring(piph, m, I0, ain);
      cp.rin = ranges_as_b(A, self, xs, m.Pmax);
      cp.fbase_in = fI[JK];
   }
   if (cp.out) {
      int aout = dim == 1 ? a + 1 : a + m.Nr;
      int JK = dim == 1 ? j + (m.Nr - 1) * k : j + m.Nr * k;
      Ring B = make_ring(piph, m, I0, aout);
      cp.rout = ranges_as_a(self, xs, B, m.Pmax);
      cp.fbase_out = fI[JK];
   }
   return cp;
}

// ---------------------------------------------------------------------------
// k_grad: r/z slope of every cell: area-weighted mean of its face slopes,
// normalised by the analytic face area, then limited by every face slope
// (plm.c:56-195); cells on the r = 0 axis get the polar limiter
// (plm.c:197-232, cylindrical.c:141-173).
// ---------------------------------------------------------------------------

template <int NQ>
__global__ void k_grad(Mesh m, int dim, double PLM, const int *fI, int nfaces,
                       const int *I0, const double *piph, const double *dphi,
                       const double *prim, const double *S, const double *fdA,
                       double *grad)
{
   int c = blockIdx.x * blockDim.x + threadIdx.x;
   if (c >= m.ntot) return;
   size_t nt = (size_t)m.ntot, nfs = (size_t)nfaces;
   int a = m.cjk[c];
   int i = c - m.off[a];
   int j = a % m.Nr, k = a / m.Nr;
   CellPairs cp = cell_pairs(m, dim, a, i, fI, I0, piph);
   double acc[NQ];
#pragma unroll
   for (int q = 0; q < NQ; ++q) acc[q] = 0.0;

   auto add_range = [&](int base, int lo, int hi) {
      for (int f = base + lo; f <= base + hi; ++f) {
         double dA = fdA[f];
#pragma unroll
         for (int q = 0; q < NQ; ++q) acc[q] += S[q * nfs + f] * dA;
      }
   };
   if (cp.in) {
      add_range(cp.fbase_in, cp.rin.h0, cp.rin.h1);
      add_range(cp.fbase_in, cp.rin.t0, cp.rin.t1);
   }
   if (cp.out) {
      add_range(cp.fbase_out, cp.rout.h0, cp.rout.h1);
      add_range(cp.fbase_out, cp.rout.t0, cp.rout.t1);
   }
   // Divide by the analytic area of the cell's two faces in this direction.
   double phip = piph[c];
   double phim = phip - dphi[c];
   double dp = get_dp(phip, phim, m.Pmax);
   double dAp, dAm;
   if (dim == 1) {
      double dz = m.z_kph[k] - m.z_kph[k - 1];
      dAp = .5 * (m.r_jph[j] + m.r_jph[j]) * dp * dz;
      dAm = .5 * (m.r_jph[j - 1] + m.r_jph[j - 1]) * dp * dz;
   } else {
      double r = .5 * (m.r_jph[j] + m.r_jph[j - 1]);
      double dr = m.r_jph[j] - m.r_jph[j - 1];
      dAp = r * dr * dp;
      dAm = dAp;
   }
   double dAtot = dAp + dAm;
   if (!cp.in) dAtot = dAp;
   if (!cp.out) dAtot = dAm;
   double g[NQ];
#pragma unroll
   for (int q = 0; q < NQ; ++q) g[q] = acc[q] / dAtot;

   // Limit against every face slope; the outcome does not depend on order.
   auto limit_range = [&](int base, int lo, int hi) {
      for (int f = base + lo; f <= base + hi; ++f) {
#pragma unroll
         for (int q = 0; q < NQ; ++q) {
            double s = S[q * nfs + f];
            if (s * g[q] < 0.0) g[q] = 0.0;
            else if (fabs(PLM * s) < fabs(g[q])) g[q] = PLM * s;
         }
      }
   };
   if (cp.in) {
      limit_range(cp.fbase_in, cp.rin.h0, cp.rin.h1);
      limit_range(cp.fbase_in, cp.rin.t0, cp.rin.t1);
   }
   if (cp.out) {
      limit_range(cp.fbase_out, cp.rout.h0, cp.rout.h1);
      limit_range(cp.fbase_out, cp.rout.t0, cp.rout.t1);
   }
   // Polar axis (r = 0 is a coordinate boundary without ghosts).
   if (dim == 1 && j == 0 && m.NgRa == 0) {
      double xp0 = m.r_jph[0], xm0 = m.r_jph[-1];
      if (!(xp0 < 0.0 || fabs(xm0) > 1.0e-10 * xp0)) {
         double r = centroid_r(xp0, xm0);
#pragma unroll
         for (int q = 0; q < NQ; ++q) {
            if (q == URR) {
               double SL = prim[q * nt + c] / r;
               double Sg = g[q];
               if (Sg * SL < 0.0) g[q] = 0.0;
               else if (fabs(PLM * SL) < fabs(Sg)) g[q] = PLM * SL;
            } else {
               g[q] = 0.0;
            }
         }
      }
   }
#pragma unroll
   for (int q = 0; q < NQ; ++q) grad[q * nt + c] = g[q];
}

// ---------------------------------------------------------------------------
// Viscous flux of an interface (riemann.c:291-309)
// ---------------------------------------------------------------------------

template <int NQ, class H>
DHD void add_visc(const dh_params &P, const double *pL, const double *pR, const double *gr,
                  const double *gp, const double *gz, double r, const double *n,
                  double *Flux)
{
   double pa[NQ], vF[NQ];
#pragma unroll
   for (int q = 0; q < NQ; ++q) {
      pa[q] = .5 * (pL[q] + pR[q]);
      vF[q] = 0.0;
   }
   H::visc_flux(P, pa, gr, gp, gz, vF, r, n);
#pragma unroll
   for (int q = 0; q < NQ; ++q) Flux[q] += vF[q];
}

// ---------------------------------------------------------------------------
// k_flux_phi: flux through the face between cell i and i+1 of every real
// annulus (misc.c:368-426, riemann.c:39-112), stored as (F - w U*) dA dt.
// ---------------------------------------------------------------------------

template <int NQ, class H>
__global__ void k_flux_phi(Mesh m, dh_params P, double dt, const double *prim,
                           const double *gradp, const double *gradr, const double *gradz,
                           const double *piph, const double *dphi, const double *wiph,
                           double *dFp)
{
   int c = blockIdx.x * blockDim.x + threadIdx.x;
   if (c >= m.ntot) return;
   int a = m.cjk[c];
   int j = a % m.Nr, k = a / m.Nr;
   if (!is_real(m, j, k)) return;
   size_t nt = (size_t)m.ntot;
   int c0 = m.off[a], np = m.Np[a];
   int cR = (c - c0 == np - 1) ? c0 : c + 1;
   double rp = m.r_jph[j], rm = m.r_jph[j - 1];
   double zp = m.z_kph[k], zm = m.z_kph[k - 1];
   double r = centroid_r(rp, rm);
   double dAdt = (rp - rm) * (zp - zm) * dt;
   double dL = dphi[c], dR = dphi[cR];
   double pL[NQ], pR[NQ];
#pragma unroll
   for (int q = 0; q < NQ; ++q) {
      pL[q] = prim[q * nt + c] + .5 * gradp[q * nt + c] * dL;
      pR[q] = prim[q * nt + cR] - .5 * gradp[q * nt + cR] * dR;
   }
   const double n[3] = {0.0, 1.0, 0.0};
   double w = r * wiph[c];
   double Flux[NQ], Ustr[NQ];
   H::riemann(P, pL, pR, r, n, w, Flux, Ustr);
   if (P.visc_flag) {
      double gr[NQ], gp[NQ], gz[NQ];
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
         double g = (prim[q * nt + cR] - prim[q * nt + c]) / (0.5 * (dR + dL));
         gr[q] = .5 * (gradr[q * nt + c] + gradr[q * nt + cR]);
         gp[q] = .5 * (g + g);
         gz[q] = .5 * (gradz[q * nt + c] + gradz[q * nt + cR]);
      }
      add_visc<NQ, H>(P, pL, pR, gr, gp, gz, r, n, Flux);
   }
#pragma unroll
   for (int q = 0; q < NQ; ++q) dFp[q * nt + c] = (Flux[q] - w * Ustr[q]) * dAdt;
}

// Pair ranges whose faces carry fluxes (trans_flux, misc.c:443-490, no CT).
DHD bool pair_has_flux(const Mesh &m, int dim, const Pair &pr)
{
   int jmin, jmax, kmin, kmax;
   if (dim == 1) {
      jmin = m.NgRa == 0 ? 0 : m.NgRa - 1;
      jmax = m.NgRb == 0 ? m.Nr - 1 : m.Nr - m.NgRb;
      kmin = m.NgZa;
      kmax = m.Nz - m.NgZb;
   } else {
      jmin = m.NgRa;
      jmax = m.Nr - m.NgRb;
      kmin = m.NgZa == 0 ? 0 : m.NgZa - 1;
      kmax = m.NgZb == 0 ? m.Nz - 1 : m.Nz - m.NgZb;
   }
   return pr.j >= jmin && pr.j < jmax && pr.k >= kmin && pr.k < kmax;
}

// ---------------------------------------------------------------------------
// k_flux_trans: flux through every r (dim 1) or z (dim 2) face of the flux
// range (riemann.c:114-213), stored as F dA dt (faces do not move: w = 0).
// ---------------------------------------------------------------------------

template <int NQ, class H>
__global__ void k_flux_trans(Mesh m, dh_params P, int dim, double dt, const int *fI,
                             int npairs, int nfaces, const int *fL, const int *fR,
                             const double *fdA, const double *fphic, const double *prim,
                             const double *gradp, const double *gradr, const double *gradz,
                             const double *piph, const double *dphi, double *dF)
{
   int f = blockIdx.x * blockDim.x + threadIdx.x;
   if (f >= nfaces) return;
   int Pi = pair_of_face(fI, npairs, f);
   Pair pr = pair_of(m, dim, Pi);
   if (!pair_has_flux(m, dim, pr)) return;
   size_t nt = (size_t)m.ntot, nfs = (size_t)nfaces;
   int cL = fL[f], cR = fR[f];
   double dxL, dxR;
   pair_dx(m, dim, pr, &dxL, &dxR);
   double phi = fphic[f];
   double dpL = get_sdp(phi, piph[cL] - .5 * dphi[cL], m.Pmax);
   double dpR = get_sdp(piph[cR] - .5 * dphi[cR], phi, m.Pmax);
   const double *grad = dim == 1 ? gradr : gradz;
   double pL[NQ], pR[NQ];
#pragma unroll
   for (int q = 0; q < NQ; ++q) {
      pL[q] = prim[q * nt + cL] + grad[q * nt + cL] * dxL + gradp[q * nt + cL] * dpL;
      pR[q] = prim[q * nt + cR] - grad[q * nt + cR] * dxR - gradp[q * nt + cR] * dpR;
   }
   double n[3] = {0.0, 0.0, 0.0};
   n[dim == 1 ? 0 : 2] = 1.0;
   double r;
   int j = pr.j;
   if (dim == 1) r = .5 * (m.r_jph[j] + m.r_jph[j]);
   else r = centroid_r(m.r_jph[j], m.r_jph[j - 1]);
   double Flux[NQ], Ustr[NQ];
   H::riemann(P, pL, pR, r, n, 0.0, Flux, Ustr);
   if (P.visc_flag) {
      double gd[NQ], gr[NQ], gp[NQ], gz[NQ];
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
         double gpL = gradp[q * nt + cL], gpR = gradp[q * nt + cR];
         gd[q] = ((prim[q * nt + cR] - gpR * dpR) - (prim[q * nt + cL] + gpL * dpL)) / (dxL + dxR);
         gp[q] = .5 * (gpL + gpR);
         if (dim == 1) {
            gr[q] = .5 * (gd[q] + gd[q]);
            gz[q] = .5 * (gradz[q * nt + cL] + gradz[q * nt + cR]);
         } else {
            gr[q] = .5 * (gradr[q * nt + cL] + gradr[q * nt + cR]);
            gz[q] = .5 * (gd[q] + gd[q]);
         }
      }
      add_visc<NQ, H>(P, pL, pR, gr, gp, gz, r, n, Flux);
   }
   double dAdt = fdA[f] * dt;
#pragma unroll
   for (int q = 0; q < NQ; ++q) dF[q * nfs + f] = (Flux[q] - 0.0 * Ustr[q]) * dAdt;
}

// ---------------------------------------------------------------------------
// Point-mass gravity (planet.c:56-196)
// ---------------------------------------------------------------------------

DHD double fgrav(double M, double r, double eps, int type)
{
   if (type == 0) {
      double n = 2.0;
      return M * pow(r, n - 1.) / pow(pow(r, n) + pow(eps, n), 1. + 1. / n);
   } else if (type == 1) {
      return M / ((r - 2 * M) * (r - 2 * M));
   } else if (type == 2) {
      return M;
   }
   return 0.0;
}

DHD void planet_src(const dh_params &P, const double *pl, const double *prim, double *cons,
                    double rp, double rm, double phip, double phim, double zp, double zm,
                    double dVdt, int polar, double Pmax)
{
   double rho = prim[RHO], vr = prim[URR], vz = prim[UZZ], omega = prim[UPP];
   double dphi = get_dp(phip, phim, Pmax);
   double phi = phim + 0.5 * dphi;
   double r = centroid_r(rp, rm);
   double z = centroid_z(zp, zm);
   double rpl = pl[0], ppl = pl[1];
   if (P.grav2d == 1) z = 0.0;
   else if (P.grav2d == 2) { rpl = r; ppl = phi; z = 1.0; }
   double cosp = cos(phi), sinp = sin(phi);
   double dx = r * cosp - rpl * cos(ppl);
   double dy = r * sinp - rpl * sin(ppl);
   double script_r = sqrt(dx * dx + dy * dy + z * z);
   double script_r_perp = sqrt(dx * dx + dy * dy);
   double f1 = -fgrav(pl[2], script_r, pl[5], (int)pl[6]);
   double cosa = dx / script_r_perp;
   double sina = dy / script_r_perp;
   double cosap = cosa * cosp + sina * sinp;
   double sinap = sina * cosp - cosa * sinp;
   if (script_r_perp <= 0.0) { cosap = 0.0; sinap = 0.0; }
   double sint = script_r_perp / script_r;
   double cost = z / script_r;
   double F0 = cosap * f1 * sint, F1 = sinap * f1 * sint, F2 = f1 * cost;
   if (polar) {
      double adj = sin(0.5 * dphi) / (0.5 * dphi);
      F0 *= adj;
      F1 *= adj;
   }
   double hr = 1.0, hp = r, hz = 1.0;
   cons[SRR] += rho * hr * F0 * dVdt;
   cons[LLL] += rho * hp * F1 * dVdt;
   cons[SZZ] += rho * hz * F2 * dVdt;
   cons[TAU] += rho * (hr * F0 * vr + hz * F2 * vz + hp * F1 * omega) * dVdt;
}

// ---------------------------------------------------------------------------
// k_update: RK blend (misc.c:226-244), flux gather in the reference's order
// (phi, then r, then z faces; each cell's inner pair before its outer pair),
// sources (misc.c:538-580), mesh motion (misc.c:256-288) and cons2prim
// (misc.c:290-325).  Ghost cells only move.
// ---------------------------------------------------------------------------

struct UpdateArgs {
   const double *prim_in;
   double *prim;
   double *cons;
   const double *rkcons;
   const double *gradr, *gradp, *gradz;
   const double *piph, *dphi, *wiph;
   double *piph_out, *dphi_out;
   const double *dFp, *dFr, *dFz;
   const int *fIr, *fIz;
   int nfr, nfz;
   const int *I0;
};

template <int NQ>
DHD void gather_range(const double *dF, size_t nfs, int base, int lo, int hi, double sgn,
                      double *cons)
{
   for (int f = base + lo; f <= base + hi; ++f) {
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
         if (sgn > 0) cons[q] += dF[q * nfs + f];
         else cons[q] -= dF[q * nfs + f];
      }
   }
}

template <int NQ>
DHD void gather_pairs(const Mesh &m, int dim, int a, int i, const int *fI, int nf,
                      const int *I0, const double *piph, const double *dF, double *cons)
{
   CellPairs cp = cell_pairs(m, dim, a, i, fI, I0, piph);
   size_t nfs = (size_t)nf;
   if (cp.in) {
      gather_range<NQ>(dF, nfs, cp.fbase_in, cp.rin.h0, cp.rin.h1, +1.0, cons);
      gather_range<NQ>(dF, nfs, cp.fbase_in, cp.rin.t0, cp.rin.t1, +1.0, cons);
   }
   if (cp.out) {
      gather_range<NQ>(dF, nfs, cp.fbase_out, cp.rout.h0, cp.rout.h1, -1.0, cons);
      gather_range<NQ>(dF, nfs, cp.fbase_out, cp.rout.t0, cp.rout.t1, -1.0, cons);
   }
}

DHD double clean_phi(double phi, double Pmax)
{
   if (!isfinite(phi)) return phi;
   if (fabs(phi) > 64.0 * Pmax) phi = fmod(phi, Pmax);
   while (phi > Pmax) phi -= Pmax;
   while (phi < 0.0) phi += Pmax;
   return phi;
}

template <int NQ, class H>
__global__ void k_update(Mesh m, dh_params P, UpdateArgs u, double RK, double dt,
                         int first_step, PlanetSet pls)
{
   int c = blockIdx.x * blockDim.x + threadIdx.x;
   if (c >= m.ntot) return;
   size_t nt = (size_t)m.ntot;
   int a = m.cjk[c];
   int c0 = m.off[a], np = m.Np[a];
   int i = c - c0;
   int j = a % m.Nr, k = a / m.Nr;
   int cm = i == 0 ? c0 + np - 1 : c - 1;
   double Pmax = m.Pmax;

   // Mesh motion of this cell and of its phi neighbour (move_cells only on
   // the first substep; clean_pi and calc_dp every substep).
   double pp = u.piph[c], pm = u.piph[cm];
   if (first_step) {
      pp += u.wiph[c] * dt;
      pm += u.wiph[cm] * dt;
   }
   pp = clean_phi(pp, Pmax);
   pm = clean_phi(pm, Pmax);
   double dp_new = get_dp(pp, pm, Pmax);
   u.piph_out[c] = pp;
   u.dphi_out[c] = dp_new;
   if (!is_real(m, j, k)) return;

   double cons[NQ];
#pragma unroll
   for (int q = 0; q < NQ; ++q) {
      double v = u.cons[q * nt + c];
      if (RK != 0.0) v = (1. - RK) * v + RK * u.rkcons[q * nt + c];
      cons[q] = v;
   }
   // phi faces: cell i is R of face i-1 and L of face i; the reference loop
   // reaches face i-1 first except for i = 0, whose R face is the last one.
   if (i == 0) {
#pragma unroll
      for (int q = 0; q < NQ; ++q) cons[q] -= u.dFp[q * nt + c];
#pragma unroll
      for (int q = 0; q < NQ; ++q) cons[q] += u.dFp[q * nt + cm];
   } else {
#pragma unroll
      for (int q = 0; q < NQ; ++q) cons[q] += u.dFp[q * nt + cm];
#pragma unroll
      for (int q = 0; q < NQ; ++q) cons[q] -= u.dFp[q * nt + c];
   }
   if (m.Nr > 1) gather_pairs<NQ>(m, 1, a, i, u.fIr, u.nfr, u.I0, u.piph, u.dFr, cons);
   if (m.Nz > 1) gather_pairs<NQ>(m, 2, a, i, u.fIz, u.nfz, u.I0, u.piph, u.dFz, cons);

   // Sources at the cell's pre-move position.
   double rp = m.r_jph[j], rm = m.r_jph[j - 1];
   double zp = m.z_kph[k], zm = m.z_kph[k - 1];
   double phip = u.piph[c];
   double phim = phip - u.dphi[c];
   double dV = cell_dV(rp, rm, phip, phim, zp, zm, Pmax);
   double dVdt = dV * dt;
   int polar = P.nobc_rmin == 1;
   double prim[NQ];
   load_q<NQ>(u.prim_in, nt, c, prim);
   H::source(P, prim, cons, rp, rm, phip, phim, zp, zm, dVdt, polar, Pmax);
   for (int p = 0; p < pls.n; ++p)
      planet_src(P, pls.v + p * DH_PLANET_STRIDE, prim, cons, rp, rm, phip, phim, zp, zm,
                 dVdt, polar, Pmax);
   if (P.visc_flag) {
      double gr[NQ], gp[NQ], gz[NQ];
      load_q<NQ>(u.gradr, nt, c, gr);
      load_q<NQ>(u.gradp, nt, c, gp);
      load_q<NQ>(u.gradz, nt, c, gz);
      H::visc_source(P, prim, gr, gp, gz, cons, centroid_r(rp, rm), dVdt);
   }
#pragma unroll
   for (int q = 0; q < NQ; ++q) u.cons[q * nt + c] = cons[q];

   // cons2prim at the moved position.
   double r = centroid_r(rp, rm);
   double phim_n = pp - dp_new;
   double dVn = cell_dV(rp, rm, pp, phim_n, zp, zm, Pmax);
   double out[NQ];
   H::cons2prim(P, cons, out, r, dVn);
#pragma unroll
   for (int q = 0; q < NQ; ++q) u.prim[q * nt + c] = out[q];
}

// ---------------------------------------------------------------------------
// k_bc_copy: zero-gradient ghost annuli, sequentially per column of ghost
// annuli as the reference does (boundary_functions.c:90-159, 269-441).
//   side 0: r inner, 1: r outer, 2: z bottom, 3: z top.
// ---------------------------------------------------------------------------

template <int NQ>
__device__ void copy_distant(const Mesh &m, double *prim, const double *piph,
                             const double *dphi, int dst, int src)
{
   size_t nt = (size_t)m.ntot;
   double pm = m.Pmax;
   int c0 = m.off[dst], s0 = m.off[src];
   int Np = m.Np[dst], Np1 = m.Np[src];
   for (int i = 0; i < Np; ++i)
      for (int q = 0; q < NQ; ++q) prim[q * nt + c0 + i] = 0.0;
   int i1;
   double phi0 = piph[c0] - dphi[c0];
   for (i1 = 0; i1 < Np1; i1++) {
      double dphip = get_sdp(piph[s0 + i1], phi0, pm);
      double dphim = dphip - dphi[s0 + i1];
      if (dphip > 0 && dphim <= 0.0) break;
   }
   if (i1 >= Np1) i1 = 0;
   for (int i = 0; i < Np; ++i) {
      double phip = piph[c0 + i];
      double phim = piph[c0 + i] - dphi[c0 + i];
      double phip1, phim1;
      double acc[NQ];
      for (int q = 0; q < NQ; ++q) acc[q] = 0.0;
      int guard = 0;
      do {
         phip1 = piph[s0 + i1];
         phim1 = piph[s0 + i1] - dphi[s0 + i1];
         if (isfinite(phip1) && isfinite(phim1) && isfinite(phip) && isfinite(phim)) {
            while (phip1 < phim) { phip1 += pm; phim1 += pm; }
            while (phim1 > phip) { phip1 -= pm; phim1 -= pm; }
         }
         double phi1 = phip < phip1 ? phip : phip1;
         double phi2 = phim > phim1 ? phim : phim1;
         double d = phi1 - phi2;
         for (int q = 0; q < NQ; ++q) acc[q] += prim[q * nt + s0 + i1] * d;
         i1 = i1 < Np1 - 1 ? i1 + 1 : 0;
      } while (phip1 < phip && ++guard <= Np1);
      i1 = i1 > 0 ? i1 - 1 : Np1 - 1;
      for (int q = 0; q < NQ; ++q) prim[q * nt + c0 + i] = acc[q] / dphi[c0 + i];
   }
}

template <int NQ>
__global__ void k_bc_copy(Mesh m, int side, int diode, double *prim, const double *piph,
                          const double *dphi)
{
   int t = blockIdx.x * blockDim.x + threadIdx.x;
   size_t nt = (size_t)m.ntot;
   if (side < 2) {
      if (t >= m.Nz) return;
      int k = t;
      if (side == 0) {
         for (int j = m.NgRa - 1; j >= 0; --j) {
            int dst = j + m.Nr * k;
            copy_distant<NQ>(m, prim, piph, dphi, dst, dst + 1);
            if (diode)
               for (int i = 0; i < m.Np[dst]; ++i) {
                  double *v = &prim[URR * nt + m.off[dst] + i];
                  if (*v > 0) *v = 0.0;
               }
         }
      } else {
         for (int j = m.Nr - m.NgRb; j < m.Nr; ++j) {
            int dst = j + m.Nr * k;
            copy_distant<NQ>(m, prim, piph, dphi, dst, dst - 1);
            if (diode)
               for (int i = 0; i < m.Np[dst]; ++i) {
                  double *v = &prim[URR * nt + m.off[dst] + i];
                  if (*v < 0) *v = 0.0;
               }
         }
      }
   } else {
      if (t >= m.Nr) return;
      int j = t;
      if (side == 2) {
         for (int k = m.NgZa - 1; k >= 0; --k) {
            int dst = j + m.Nr * k;
            copy_distant<NQ>(m, prim, piph, dphi, dst, dst + m.Nr);
            if (diode)
               for (int i = 0; i < m.Np[dst]; ++i) {
                  double *v = &prim[UZZ * nt + m.off[dst] + i];
                  if (*v > 0) *v = 0.0;
               }
         }
      } else {
         for (int k = m.Nz - m.NgZb; k < m.Nz; ++k) {
            int dst = j + m.Nr * k;
            copy_distant<NQ>(m, prim, piph, dphi, dst, dst - m.Nr);
            if (diode)
               for (int i = 0; i < m.Np[dst]; ++i) {
                  double *v = &prim[UZZ * nt + m.off[dst] + i];
                  if (*v < 0) *v = 0.0;
               }
         }
      }
   }
}

// ---------------------------------------------------------------------------
// k_init_cells: cons from prim and back on real cells (domain.c:175-237).
// ---------------------------------------------------------------------------

template <int NQ, class H>
__global__ void k_init_cells(Mesh m, dh_params P, double *prim, double *cons,
                             const double *piph, const double *dphi)
{
   int c = blockIdx.x * blockDim.x + threadIdx.x;
   if (c >= m.ntot) return;
   int a = m.cjk[c];
   int j = a % m.Nr, k = a / m.Nr;
   if (!is_real(m, j, k)) return;
   size_t nt = (size_t)m.ntot;
   double rp = m.r_jph[j], rm = m.r_jph[j - 1];
   double zp = m.z_kph[k], zm = m.z_kph[k - 1];
   double phip = piph[c];
   double phim = phip - dphi[c];
   double dV = cell_dV(rp, rm, phip, phim, zp, zm, m.Pmax);
   double r = centroid_r(rp, rm);
   double pr[NQ], cs[NQ];
   load_q<NQ>(prim, nt, c, pr);
   H::prim2cons(P, pr, cs, r, dV);
   H::cons2prim(P, cs, pr, r, dV);
#pragma unroll
   for (int q = 0; q < NQ; ++q) {
      cons[q * nt + c] = cs[q];
      prim[q * nt + c] = pr[q];
   }
}

// ---------------------------------------------------------------------------
// k_mindt: CFL reduction over real cells (misc.c:39-84).  Positive doubles
// order like their bit patterns, so the minimum is an unsigned 64-bit min.
// ---------------------------------------------------------------------------

template <int NQ, class H>
__global__ void k_mindt(Mesh m, dh_params P, const double *prim, const double *piph,
                        const double *dphi, const double *wiph, unsigned long long *out)
{
   __shared__ double red[TPB / 64];
   int c = blockIdx.x * blockDim.x + threadIdx.x;
   double dtc = HUGE_VAL;
   if (c < m.ntot) {
      int a = m.cjk[c];
      int j = a % m.Nr, k = a / m.Nr;
      if (is_real(m, j, k)) {
         size_t nt = (size_t)m.ntot;
         int c0 = m.off[a], np = m.Np[a];
         int cm = (c == c0) ? c0 + np - 1 : c - 1;
         double w = .5 * (wiph[cm] + wiph[c]);
         double phip = piph[c];
         double phim = phip - dphi[c];
         double pr[NQ];
         load_q<NQ>(prim, nt, c, pr);
         dtc = H::mindt(P, pr, w, m.r_jph[j], m.r_jph[j - 1], phip, phim, m.z_kph[k],
                        m.z_kph[k - 1], m.Pmax);
      }
   }
   for (int o = 32; o > 0; o >>= 1) {
      double v = __shfl_down(dtc, o, 64);
      dtc = v < dtc ? v : dtc;
   }
   int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
   if (lane == 0) red[wv] = dtc;
   __syncthreads();
   if (threadIdx.x == 0) {
      double v = red[0];
      for (int w = 1; w < TPB / 64; ++w) v = red[w] < v ? red[w] : v;
      if (v < HUGE_VAL && v > 0.0) atomicMin(out, (unsigned long long)__double_as_longlong(v));
   }
}

// ---------------------------------------------------------------------------
// Host context
// ---------------------------------------------------------------------------

struct TimerRow {
   std::string name;
   std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
   double ms = 0.0;
   long calls = 0;
};

struct dh_ctx {
   dh_params P;
   int nq, nf;
   Mesh m;
   int *d_Np = nullptr, *d_off = nullptr, *d_cjk = nullptr;
   double *d_r = nullptr, *d_z = nullptr;
   std::vector<int> h_Np, h_off;
   int npairs[2] = {0, 0}, nfaces[2] = {0, 0};
   int *d_fI[2] = {nullptr, nullptr};
   double *prim = nullptr, *cons = nullptr, *rkcons = nullptr;
   double *gradp = nullptr, *gradr = nullptr, *gradz = nullptr;
   double *piph = nullptr, *piph2 = nullptr, *dphi = nullptr, *dphi2 = nullptr;
   double *wiph = nullptr;
   int *fL[2] = {nullptr, nullptr}, *fR[2] = {nullptr, nullptr};
   double *fdA[2] = {nullptr, nullptr}, *fphic[2] = {nullptr, nullptr};
   double *S = nullptr, *dF[2] = {nullptr, nullptr}, *dFp = nullptr;
   int *I0 = nullptr;
   unsigned long long *d_mindt = nullptr;
   PlanetSet pl[2];
   hipStream_t stream = nullptr;
   bool timers = false;
   std::vector<TimerRow> trows;
   std::vector<void *> allocs;
};

static int dalloc(dh_ctx *x, void **p, size_t bytes)
{
   if (bytes == 0) bytes = 8;
   hipError_t e = hipMalloc(p, bytes);
   if (e != hipSuccess)
      return fail(DH_ENOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
   x->allocs.push_back(*p);
   return DH_OK;
}

#define DALLOC(ptr, bytes)                                                   \
   do {                                                                      \
      int rc_ = dalloc(x, (void **)&(ptr), (bytes));                         \
      if (rc_) { dh_destroy(x); return rc_; }                                \
   } while (0)

// Launch bracket: when timers are on, every launch is framed by events and
// accounted to its kernel family.
struct Launch {
   dh_ctx *x;
   int row = -1;
   hipEvent_t e0{}, e1{};
   Launch(dh_ctx *ctx, const char *name) : x(ctx)
   {
      if (!x->timers) return;
      for (size_t i = 0; i < x->trows.size(); ++i)
         if (x->trows[i].name == name) row = (int)i;
      if (row < 0) {
         x->trows.push_back(TimerRow());
         x->trows.back().name = name;
         row = (int)x->trows.size() - 1;
      }
      (void)hipEventCreate(&e0);
      (void)hipEventCreate(&e1);
      (void)hipEventRecord(e0, x->stream);
   }
   ~Launch()
   {
      if (row < 0) return;
      (void)hipEventRecord(e1, x->stream);
      x->trows[row].pending.push_back({e0, e1});
   }
};

#define LAUNCH_CHECK()                                                        \
   do {                                                                       \
      hipError_t e_ = hipGetLastError();                                      \
      if (e_ != hipSuccess)                                                   \
         return fail(DH_EHIP, "kernel launch failed: %s", hipGetErrorString(e_)); \
   } while (0)

// ---------------------------------------------------------------------------
// Substep (onestep.c:84-228)
// ---------------------------------------------------------------------------

template <int NQ>
static int boundaries(dh_ctx *x);

template <int NQ, class H>
static int substep(dh_ctx *x, double RK, double dt, int first_step, int sub)
{
   Mesh &m = x->m;
   dh_params &P = x->P;
   hipStream_t s = x->stream;
   int nb = nblocks(m.ntot);
   if (first_step) {
      {
         Launch L(x, "wcell");
         k_wcell<NQ, H><<<nb, TPB, 0, s>>>(m, P, x->prim, x->wiph);
      }
      LAUNCH_CHECK();
      if (P.mesh_motion == 3) {
         k_wcell_avg<<<nblocks(m.nann), TPB, 0, s>>>(m, x->wiph);
         LAUNCH_CHECK();
      }
   }
   {
      Launch L(x, "plm_phi");
      k_plm_phi<NQ><<<nb, TPB, 0, s>>>(m, P.plm, x->prim, x->dphi, x->gradp);
   }
   LAUNCH_CHECK();
   if (m.Nr > 1 || m.Nz > 1) {
      Launch L(x, "ring_start");
      k_ring_start<<<(m.nann + 3) / 4, 256, 0, s>>>(m, x->piph, x->I0);
   }
   LAUNCH_CHECK();
   for (int dim = 1; dim <= 2; ++dim) {
      int d = dim - 1;
      if ((dim == 1 && m.Nr <= 1) || (dim == 2 && m.Nz <= 1)) continue;
      double *grad = dim == 1 ? x->gradr : x->gradz;
      if (P.plm == 0.0) {
         HIPCHK(hipMemsetAsync(grad, 0, sizeof(double) * NQ * m.ntot, s));
      }
      {
         Launch L(x, dim == 1 ? "face_slopes_r" : "face_slopes_z");
         k_face_slopes<NQ><<<nblocks(x->nfaces[d]), TPB, 0, s>>>(
             m, dim, x->d_fI[d], x->npairs[d], x->nfaces[d], x->I0, x->piph, x->dphi,
             x->prim, x->gradp, x->fL[d], x->fR[d], x->fdA[d], x->fphic[d], x->S);
      }
      LAUNCH_CHECK();
      if (P.plm != 0.0) {
         Launch L(x, dim == 1 ? "grad_r" : "grad_z");
         k_grad<NQ><<<nb, TPB, 0, s>>>(m, dim, P.plm, x->d_fI[d], x->nfaces[d], x->I0,
                                       x->piph, x->dphi, x->prim, x->S, x->fdA[d], grad);
      }
      LAUNCH_CHECK();
   }
   {
      Launch L(x, "flux_phi");
      k_flux_phi<NQ, H><<<nb, TPB, 0, s>>>(m, P, dt, x->prim, x->gradp, x->gradr, x->gradz,
                                           x->piph, x->dphi, x->wiph, x->dFp);
   }
   LAUNCH_CHECK();
   for (int dim = 1; dim <= 2; ++dim) {
      int d = dim - 1;
      if ((dim == 1 && m.Nr <= 1) || (dim == 2 && m.Nz <= 1)) continue;
      Launch L(x, dim == 1 ? "flux_r" : "flux_z");
      k_flux_trans<NQ, H><<<nblocks(x->nfaces[d]), TPB, 0, s>>>(
          m, P, dim, dt, x->d_fI[d], x->npairs[d], x->nfaces[d], x->fL[d], x->fR[d],
          x->fdA[d], x->fphic[d], x->prim, x->gradp, x->gradr, x->gradz, x->piph, x->dphi,
          x->dF[d]);
   }
   LAUNCH_CHECK();
   UpdateArgs u;
   u.prim_in = x->prim;
   u.prim = x->prim;
   u.cons = x->cons;
   u.rkcons = x->rkcons;
   u.gradr = x->gradr;
   u.gradp = x->gradp;
   u.gradz = x->gradz;
   u.piph = x->piph;
   u.dphi = x->dphi;
   u.wiph = x->wiph;
   u.piph_out = x->piph2;
   u.dphi_out = x->dphi2;
   u.dFp = x->dFp;
   u.dFr = x->dF[0];
   u.dFz = x->dF[1];
   u.fIr = x->d_fI[0];
   u.fIz = x->d_fI[1];
   u.nfr = x->nfaces[0];
   u.nfz = x->nfaces[1];
   u.I0 = x->I0;
   {
      Launch L(x, "update");
      k_update<NQ, H><<<nb, TPB, 0, s>>>(m, P, u, RK, dt, first_step, x->pl[sub]);
   }
   LAUNCH_CHECK();
   std::swap(x->piph, x->piph2);
   std::swap(x->dphi, x->dphi2);
   return boundaries<NQ>(x);
}

// Boundaries (onestep.c:194-210): r sides, then z sides.
template <int NQ>
static int boundaries(dh_ctx *x)
{
   Mesh &m = x->m;
   dh_params &P = x->P;
   hipStream_t s = x->stream;
   int bcs[4] = {P.bc_rmin, P.bc_rmax, P.bc_zmin, P.bc_zmax};
   for (int side = 0; side < 4; ++side) {
      if (side >= 2 && (m.Nz <= 1 || P.z_periodic)) continue;
      if (side < 2 && P.r_periodic) continue;
      if (bcs[side] != DH_BC_ZEROGRAD && bcs[side] != DH_BC_DIODE) continue;
      int n = side < 2 ? m.Nz : m.Nr;
      Launch L(x, "bc_copy");
      k_bc_copy<NQ><<<(n + 63) / 64, 64, 0, s>>>(m, side, bcs[side] == DH_BC_DIODE, x->prim,
                                                 x->piph, x->dphi);
   }
   LAUNCH_CHECK();
   return DH_OK;
}

template <int NQ, class H>
static int init_impl(dh_ctx *x)
{
   Mesh &m = x->m;
   int nb = nblocks(m.ntot);
   k_init_cells<NQ, H><<<nb, TPB, 0, x->stream>>>(m, x->P, x->prim, x->cons, x->piph, x->dphi);
   LAUNCH_CHECK();
   k_wcell<NQ, H><<<nb, TPB, 0, x->stream>>>(m, x->P, x->prim, x->wiph);
   LAUNCH_CHECK();
   if (x->P.mesh_motion == 3) {
      k_wcell_avg<<<nblocks(m.nann), TPB, 0, x->stream>>>(m, x->wiph);
      LAUNCH_CHECK();
   }
   int rc = boundaries<NQ>(x);
   if (rc) return rc;
   HIPCHK(hipStreamSynchronize(x->stream));
   return DH_OK;
}

template <int NQ, class H>
static int timestep_impl(dh_ctx *x, double dt)
{
   HIPCHK(hipMemcpyAsync(x->rkcons, x->cons, sizeof(double) * NQ * x->m.ntot,
                         hipMemcpyDeviceToDevice, x->stream));
   int rc = substep<NQ, H>(x, 0.0, dt, 1, 0);
   if (rc) return rc;
   if (x->P.timestep_mode != 1) rc = substep<NQ, H>(x, 0.5, 0.5 * dt, 0, 1);
   return rc;
}

template <int NQ, class H>
static int mindt_impl(dh_ctx *x, double *dt)
{
   unsigned long long init = 0x7FF0000000000000ULL;  // +inf
   HIPCHK(hipMemcpyAsync(x->d_mindt, &init, sizeof init, hipMemcpyHostToDevice, x->stream));
   {
      Launch L(x, "mindt");
      k_mindt<NQ, H><<<nblocks(x->m.ntot), TPB, 0, x->stream>>>(
          x->m, x->P, x->prim, x->piph, x->dphi, x->wiph, x->d_mindt);
   }
   LAUNCH_CHECK();
   unsigned long long bits = 0;
   HIPCHK(hipMemcpyAsync(&bits, x->d_mindt, sizeof bits, hipMemcpyDeviceToHost, x->stream));
   HIPCHK(hipStreamSynchronize(x->stream));
   double v;
   memcpy(&v, &bits, sizeof v);
   double d0 = x->P.max_dt / x->P.cfl;
   if (d0 <= 0.0) d0 = 1.0e100;
   if (v < d0) d0 = v;
   *dt = d0 * x->P.cfl;
   return DH_OK;
}

// Dispatch on the number of fields (the reference's NUM_Q).
#define DISPATCH(x, FN, ...)                                                  \
   ([&]() -> int {                                                            \
      switch ((x)->nq) {                                                      \
      case 5: return FN<5, Euler<5>>(x, ##__VA_ARGS__);                       \
      case 6: return FN<6, Euler<6>>(x, ##__VA_ARGS__);                       \
      case 7: return FN<7, Euler<7>>(x, ##__VA_ARGS__);                       \
      default: return fail(DH_EINVAL, "unsupported NQ %d", (x)->nq);         \
      }                                                                       \
   }())

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

extern "C" {

const char *dh_last_error(void) { return g_err; }

int dh_destroy(dh_ctx *x)
{
   if (!x) return DH_OK;
   if (x->stream) (void)hipStreamSynchronize(x->stream);
   for (auto &row : x->trows)
      for (auto &pe : row.pending) {
         (void)hipEventDestroy(pe.first);
         (void)hipEventDestroy(pe.second);
      }
   for (void *p : x->allocs) (void)hipFree(p);
   if (x->stream) (void)hipStreamDestroy(x->stream);
   delete x;
   return DH_OK;
}

int dh_create(const dh_grid *g, const dh_params *par, dh_ctx **out)
{
   if (!g || !par || !out) return fail(DH_EINVAL, "null argument");
   *out = nullptr;
   int ndev = 0;
   if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
      return fail(DH_ENODEV, "no GPU visible");
   const dh_params &P = *par;
   if (P.hydro != DH_HYDRO_EULER) return fail(DH_EINVAL, "hydro module %d not supported", P.hydro);
   if (P.num_c != 5) return fail(DH_EINVAL, "euler needs NUM_C = 5 (got %d)", P.num_c);
   if (P.ct_mode != 0) return fail(DH_EINVAL, "CT_MODE %d needs the MHD module", P.ct_mode);
   if (P.riemann_solver != 0 && P.riemann_solver != 1)
      return fail(DH_EINVAL, "Riemann_Solver %d not supported by euler", P.riemann_solver);
   if (P.rot_frame) return fail(DH_EINVAL, "RotFrame is not supported");
   if (P.nozzle_type) return fail(DH_EINVAL, "Nozzle_Type is not supported");
   if (P.r_periodic || P.z_periodic) return fail(DH_EINVAL, "periodic boundaries are not supported");
   int nq = P.num_c + P.num_n;
   if (nq < 5 || nq > 7) return fail(DH_EINVAL, "NUM_C+NUM_N = %d not supported", nq);
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
   if (ntot >= (1L << 31) / 9) return fail(DH_EINVAL, "mesh too large (%ld cells)", ntot);

   dh_ctx *x = new dh_ctx();
   x->P = P;
   x->nq = nq;
   x->nf = 0;
   HIPCHK(hipStreamCreate(&x->stream));
   int nann = g->Nr * g->Nz;
   x->h_Np.assign(g->Np, g->Np + nann);
   x->h_off.assign(nann + 1, 0);
   for (int a = 0; a < nann; ++a) x->h_off[a + 1] = x->h_off[a] + g->Np[a];
   std::vector<int> cjk(ntot);
   for (int a = 0; a < nann; ++a)
      for (int i = 0; i < g->Np[a]; ++i) cjk[x->h_off[a] + i] = a;

   DALLOC(x->d_Np, sizeof(int) * nann);
   DALLOC(x->d_off, sizeof(int) * (nann + 1));
   DALLOC(x->d_cjk, sizeof(int) * ntot);
   DALLOC(x->d_r, sizeof(double) * (g->Nr + 1));
   DALLOC(x->d_z, sizeof(double) * (g->Nz + 1));
   HIPCHK(hipMemcpy(x->d_Np, g->Np, sizeof(int) * nann, hipMemcpyHostToDevice));
   HIPCHK(hipMemcpy(x->d_off, x->h_off.data(), sizeof(int) * (nann + 1), hipMemcpyHostToDevice));
   HIPCHK(hipMemcpy(x->d_cjk, cjk.data(), sizeof(int) * ntot, hipMemcpyHostToDevice));
   HIPCHK(hipMemcpy(x->d_r, g->r_jph, sizeof(double) * (g->Nr + 1), hipMemcpyHostToDevice));
   HIPCHK(hipMemcpy(x->d_z, g->z_kph, sizeof(double) * (g->Nz + 1), hipMemcpyHostToDevice));

   Mesh &m = x->m;
   m.Nr = g->Nr; m.Nz = g->Nz;
   m.NgRa = g->NgRa; m.NgRb = g->NgRb; m.NgZa = g->NgZa; m.NgZb = g->NgZb;
   m.nann = nann;
   m.ntot = (int)ntot;
   m.Np = x->d_Np;
   m.off = x->d_off;
   m.cjk = x->d_cjk;
   m.r_jph = x->d_r + 1;
   m.z_kph = x->d_z + 1;
   m.Pmax = g->phi_max;

   // Face lists: pair offsets are static because Np never changes
   // (buildfaces mode 0, faces_alt.c:80-96).
   for (int dim = 1; dim <= 2; ++dim) {
      int d = dim - 1;
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
      fI[Nrmax * Nzmax] = (int)n;
      x->npairs[d] = Nrmax * Nzmax;
      x->nfaces[d] = (int)n;
      DALLOC(x->d_fI[d], sizeof(int) * fI.size());
      HIPCHK(hipMemcpy(x->d_fI[d], fI.data(), sizeof(int) * fI.size(), hipMemcpyHostToDevice));
      DALLOC(x->fL[d], sizeof(int) * n);
      DALLOC(x->fR[d], sizeof(int) * n);
      DALLOC(x->fdA[d], sizeof(double) * n);
      DALLOC(x->fphic[d], sizeof(double) * n);
      DALLOC(x->dF[d], sizeof(double) * n * nq);
   }
   int nfmax = x->nfaces[0] > x->nfaces[1] ? x->nfaces[0] : x->nfaces[1];
   DALLOC(x->S, sizeof(double) * (size_t)nfmax * nq);
   size_t qb = sizeof(double) * (size_t)ntot * nq;
   DALLOC(x->prim, qb);
   DALLOC(x->cons, qb);
   DALLOC(x->rkcons, qb);
   DALLOC(x->gradp, qb);
   DALLOC(x->gradr, qb);
   DALLOC(x->gradz, qb);
   DALLOC(x->dFp, qb);
   DALLOC(x->piph, sizeof(double) * ntot);
   DALLOC(x->piph2, sizeof(double) * ntot);
   DALLOC(x->dphi, sizeof(double) * ntot);
   DALLOC(x->dphi2, sizeof(double) * ntot);
   DALLOC(x->wiph, sizeof(double) * ntot);
   DALLOC(x->I0, sizeof(int) * nann);
   DALLOC(x->d_mindt, sizeof(unsigned long long));
   HIPCHK(hipMemset(x->gradr, 0, qb));
   HIPCHK(hipMemset(x->gradz, 0, qb));
   HIPCHK(hipMemset(x->gradp, 0, qb));
   HIPCHK(hipMemset(x->I0, 0, sizeof(int) * nann));
   x->pl[0].n = x->pl[1].n = 0;
   *out = x;
   return DH_OK;
}

int dh_sizes(const dh_ctx *x, long *ntot, int *nq, int *nf)
{
   if (!x) return fail(DH_EINVAL, "null context");
   if (ntot) *ntot = x->m.ntot;
   if (nq) *nq = x->nq;
   if (nf) *nf = x->nf;
   return DH_OK;
}

int dh_upload_state(dh_ctx *x, const double *prim, const double *cons, const double *piph,
                    const double *dphi, const double *wiph, const double *Phi)
{
   if (!x) return fail(DH_EINVAL, "null context");
   size_t nt = (size_t)x->m.ntot, qb = sizeof(double) * nt * x->nq;
   if (prim) HIPCHK(hipMemcpyAsync(x->prim, prim, qb, hipMemcpyHostToDevice, x->stream));
   if (cons) HIPCHK(hipMemcpyAsync(x->cons, cons, qb, hipMemcpyHostToDevice, x->stream));
   if (piph) HIPCHK(hipMemcpyAsync(x->piph, piph, sizeof(double) * nt, hipMemcpyHostToDevice, x->stream));
   if (wiph) HIPCHK(hipMemcpyAsync(x->wiph, wiph, sizeof(double) * nt, hipMemcpyHostToDevice, x->stream));
   if (dphi) {
      HIPCHK(hipMemcpyAsync(x->dphi, dphi, sizeof(double) * nt, hipMemcpyHostToDevice, x->stream));
   } else if (piph) {
      std::vector<double> dp(nt);
      for (int a = 0; a < x->m.nann; ++a) {
         int c0 = x->h_off[a], np = x->h_Np[a];
         for (int i = 0; i < np; ++i) {
            double pp = piph[c0 + i], pm = piph[c0 + (i == 0 ? np - 1 : i - 1)];
            double v = pp - pm;
            while (v < 0.0) v += x->m.Pmax;
            while (v > x->m.Pmax) v -= x->m.Pmax;
            dp[c0 + i] = v;
         }
      }
      HIPCHK(hipMemcpy(x->dphi, dp.data(), sizeof(double) * nt, hipMemcpyHostToDevice));
   }
   (void)Phi;
   HIPCHK(hipStreamSynchronize(x->stream));
   return DH_OK;
}

int dh_download_state(dh_ctx *x, double *prim, double *cons, double *piph, double *dphi,
                      double *wiph, double *Phi)
{
   if (!x) return fail(DH_EINVAL, "null context");
   size_t nt = (size_t)x->m.ntot, qb = sizeof(double) * nt * x->nq;
   if (prim) HIPCHK(hipMemcpyAsync(prim, x->prim, qb, hipMemcpyDeviceToHost, x->stream));
   if (cons) HIPCHK(hipMemcpyAsync(cons, x->cons, qb, hipMemcpyDeviceToHost, x->stream));
   if (piph) HIPCHK(hipMemcpyAsync(piph, x->piph, sizeof(double) * nt, hipMemcpyDeviceToHost, x->stream));
   if (dphi) HIPCHK(hipMemcpyAsync(dphi, x->dphi, sizeof(double) * nt, hipMemcpyDeviceToHost, x->stream));
   if (wiph) HIPCHK(hipMemcpyAsync(wiph, x->wiph, sizeof(double) * nt, hipMemcpyDeviceToHost, x->stream));
   (void)Phi;
   HIPCHK(hipStreamSynchronize(x->stream));
   return DH_OK;
}

int dh_set_planets(dh_ctx *x, int substep_, const double *pl, int npl)
{
   if (!x) return fail(DH_EINVAL, "null context");
   if (substep_ < 0 || substep_ > 1) return fail(DH_EINVAL, "substep must be 0 or 1");
   if (npl < 0 || npl > DH_MAX_PLANETS) return fail(DH_EINVAL, "npl %d out of range", npl);
   if (npl > 0 && !pl) return fail(DH_EINVAL, "null planet list");
   PlanetSet &s = x->pl[substep_];
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

int dh_init_cells(dh_ctx *x)
{
   if (!x) return fail(DH_EINVAL, "null context");
   return DISPATCH(x, init_impl);
}

int dh_getmindt(dh_ctx *x, double *dt)
{
   if (!x || !dt) return fail(DH_EINVAL, "null argument");
   return DISPATCH(x, mindt_impl, dt);
}

int dh_timestep(dh_ctx *x, double dt)
{
   if (!x) return fail(DH_EINVAL, "null context");
   if (!(dt > 0.0)) return fail(DH_EINVAL, "dt must be positive");
   return DISPATCH(x, timestep_impl, dt);
}

int dh_sync(dh_ctx *x)
{
   if (!x) return fail(DH_EINVAL, "null context");
   HIPCHK(hipStreamSynchronize(x->stream));
   return DH_OK;
}

int dh_enable_timers(dh_ctx *x, int on)
{
   if (!x) return fail(DH_EINVAL, "null context");
   x->timers = on != 0;
   return DH_OK;
}

int dh_read_timers(dh_ctx *x, int cap, char (*names)[48], double *ms, long *calls, int *nrows)
{
   if (!x) return fail(DH_EINVAL, "null context");
   HIPCHK(hipStreamSynchronize(x->stream));
   int n = 0;
   for (auto &row : x->trows) {
      for (auto &pe : row.pending) {
         float t = 0.f;
         (void)hipEventElapsedTime(&t, pe.first, pe.second);
         row.ms += t;
         row.calls++;
         (void)hipEventDestroy(pe.first);
         (void)hipEventDestroy(pe.second);
      }
      row.pending.clear();
      if (n < cap) {
         snprintf(names[n], 48, "%s", row.name.c_str());
         ms[n] = row.ms;
         calls[n] = row.calls;
      }
      n++;
   }
   if (nrows) *nrows = n < cap ? n : cap;
   return DH_OK;
}

}  // extern "C"
