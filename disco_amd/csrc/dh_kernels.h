// dh_kernels.h -- the per-thread bodies of every timestep kernel.
//
// Each kernel is a functor: its members are the launch arguments and
// operator()(t) is the work of thread t (a cell, a face, a face track or a
// boundary line).  The engine launches them through one generic __global__
// template; the same bodies also build as host code, which the test suite
// runs on the CPU against the oracle to check indexing and summation order
// before anything reaches the GPU.
//
// Scatter loops of the reference (face fluxes added to both neighbours,
// slope sums, edge EMFs, magnetic-flux updates) are restated as
// owner-computes gathers: every output element is written by one thread,
// which sums its contributions in the reference's order, so a step is
// deterministic and needs no atomics.
#pragma once
#include "dh_c2pnet.h"
#include "dh_ct.h"
#include "dh_euler.h"
#include "dh_geom.h"
#include "dh_grmhd.h"
#include "dh_mesh.h"
#include "dh_mhd.h"

namespace dh {

struct PlanetSet {
   double v[DH_MAX_PLANETS * DH_PLANET_STRIDE];
   int n;
};

// What changes from one timestep to the next, kept in device memory so that
// a captured timestep replays with the new values: each substep's dt and
// planets.  The kernels take pointers into it.
struct StepArgs {
   double dt[2];
   PlanetSet pl[2];
};

// Planet 1 as the Diagnostics modules read it (theDomain->thePlanets + 1,
// used when Npl > 1).
struct DiagPlanet {
   double v[DH_PLANET_STRIDE];
   int on;
};

template <class H>
struct IsMHD {
   static constexpr bool value = H::NC == 8;
};

// The conserved components a cell's fluxes update.  With constrained
// transport (the MHD modules) the cell's field components are set from its
// faces' magnetic fluxes after the update (KBfacesC2p, bfields.c:88-309),
// so their face fluxes are neither stored nor gathered.
template <class H>
DHD constexpr bool flux_updates(int q)
{
   return !(H::NC == 8 && q >= BRR && q <= BZZ);
}

template <class H>
struct IsGR {
   static constexpr bool value = false;
};
template <int NQ>
struct IsGR<GRMHD<NQ>> {
   static constexpr bool value = true;
};

// Traversal orders of a kernel's elements.  The elements come in groups
// that are contiguous in storage (the cells of an annulus, the faces of an
// annulus pair), and an order visits whole groups in some sequence: thread
// t does element elem(t).  Every element's work is independent of when it
// runs (owner-computes, no atomics), so an order changes only which
// elements share an L2 at a time, never a result.
//   first[p]  storage index of the first element of the p-th visited group
//   pos[p]    thread index at which that group starts (pos[ng] = n)
//   chunk[c]  the p whose range holds thread ORDER_CHUNK * c, so that a
//             thread finds its group after a step or two from there
constexpr int ORDER_CHUNK = 256;
struct GroupOrder {
   const int *first, *pos, *chunk;
   int ng;
   DHD long elem(long t) const
   {
      int p = chunk[t / ORDER_CHUNK];
      while (pos[p + 1] <= t) ++p;
      return first[p] + (t - pos[p]);
   }
};

// A kernel run in a group order (its workgroups go to the XCDs in
// contiguous runs, so that neighbouring groups of the order share an L2).
template <class K>
struct Grouped {
   K k;
   GroupOrder o;
   DHD void operator()(long t) const { k(o.elem(t)); }
};

// Two kernels that share no data in one launch: threads [0, n1) run the
// first over its elements, the rest the second over its own.
template <class K1, class K2>
struct KBoth {
   K1 k1;
   K2 k2;
   long n1;
   DHD void operator()(long t) const
   {
      if (t < n1) k1(t);
      else k2(t - n1);
   }
};

// One direction's face list: faces of annulus pair P are fI[P] .. fI[P+1]-1.
struct Faces {
   int n, npairs;
   const int *fI;   // [npairs+1] first face of every annulus pair
   const int *pf;   // [n] annulus pair of every face (static, like fI)
   int *L, *R;
   signed char *LR;
   double *dA, *phic, *dphi;
   double *up;  // upper phi edge of the face (its owner's piph)
};

// ---------------------------------------------------------------------------
// Small helpers
// ---------------------------------------------------------------------------

DHD bool is_real(const Mesh &m, int j, int k)
{
   return j >= m.NgRa && j < m.Nr - m.NgRb && k >= m.NgZa && k < m.Nz - m.NgZb;
}

// Annuli whose phi faces carry fluxes: the real ones, plus one ghost layer
// when constrained transport needs the edge EMFs there (misc.c:368-399).
DHD bool phi_in_range(const Mesh &m, int ct, int j, int k)
{
   if (!ct) return is_real(m, j, k);
   int jmin = m.NgRa == 0 ? 0 : m.NgRa - 1;
   int jmax = m.NgRb == 0 ? m.Nr : m.Nr - m.NgRb + 1;
   int kmin = m.NgZa == 0 ? 0 : m.NgZa - 1;
   int kmax = m.NgZb == 0 ? m.Nz : m.Nz - m.NgZb + 1;
   return j >= jmin && j < jmax && k >= kmin && k < kmax;
}

// Annulus pairs whose faces carry fluxes (trans_flux, misc.c:431-490).
DHD bool pair_has_flux(const Mesh &m, int ct, int dim, int pj, int pk)
{
   int jmin, jmax, kmin, kmax;
   if (dim == 1) {
      jmin = m.NgRa == 0 ? 0 : m.NgRa - 1;
      jmax = m.NgRb == 0 ? m.Nr - 1 : m.Nr - m.NgRb;
      kmin = m.NgZa;
      kmax = m.Nz - m.NgZb;
      if (ct) {
         kmin = m.NgZa == 0 ? 0 : m.NgZa - 1;
         kmax = m.NgZb == 0 ? m.Nz : m.Nz - m.NgZb + 1;
      }
   } else {
      jmin = m.NgRa;
      jmax = m.Nr - m.NgRb;
      kmin = m.NgZa == 0 ? 0 : m.NgZa - 1;
      kmax = m.NgZb == 0 ? m.Nz - 1 : m.Nz - m.NgZb;
      if (ct) {
         jmin = m.NgRa == 0 ? 0 : m.NgRa - 1;
         jmax = m.NgRb == 0 ? m.Nr : m.Nr - m.NgRb + 1;
      }
   }
   return pj >= jmin && pj < jmax && pk >= kmin && pk < kmax;
}

DHD bool dim_active(const Mesh &m, int dim) { return dim == 1 ? m.Nr > 1 : m.Nz > 1; }

template <int NQ>
DHD void load_q(const double *soa, size_t nt, long c, double *v)
{
#pragma unroll
   for (int q = 0; q < NQ; ++q) v[q] = soa[q * nt + c];
}

// Element index of the i-th cell of an annulus in its ring.
DHD int ring_elem(const Ring &R, int i)
{
   int x = i - R.I0;
   return x < 0 ? x + R.N : x;
}

// clean_pi (misc.c:8-36) for one cell.
DHD double clean_phi(double phi, double Pmax)
{
   if (!std::isfinite(phi)) return phi;
   if (fabs(phi) > 64.0 * Pmax) phi = fmod(phi, Pmax);
   while (phi > Pmax) phi -= Pmax;
   while (phi < 0.0) phi += Pmax;
   return phi;
}

DHD double minmod(double a, double b, double c)
{
   double m = a;
   if (a * b < 0.0) m = 0.0;
   if (fabs(b) < fabs(m)) m = b;
   if (b * c < 0.0) m = 0.0;
   if (fabs(c) < fabs(m)) m = c;
   return m;
}

// Start cell of annulus a: the first cell (from i = 0) in the first
// quadrant that comes after a cell in the fourth (faces_alt.c:59-78).
DHD int ring_start_of(const Mesh &m, const double *piph, int a)
{
   int c0 = m.off[a], np = m.Np[a];
   double convert = 2. * M_PI / m.Pmax;
   bool seen4 = false;
   for (int i = 0; i < np; ++i) {
      double sn, cs;
      sincos(piph[c0 + i] * convert, &sn, &cs);
      if (sn > 0. && cs > 0. && seen4) return i;
      if (sn < 0. && cs > 0.) seen4 = true;
   }
   return 0;
}

// ---------------------------------------------------------------------------
// set_wcell (misc.c:101-168)
// ---------------------------------------------------------------------------

template <int NQ, class H>
struct KWcell {
   Mesh m;
   dh_params P;
   const double *prim;
   double *wiph;
   DHD void operator()(long c) const
   {
      int a = m.cjk[c];
      double w = 0.0;
      if (P.mesh_motion == 4) {
         int j = a % m.Nr;
         w = mesh_om(P, centroid_r(m.r_jph[j], m.r_jph[j - 1]));
      } else if (P.mesh_motion) {
         size_t nt = (size_t)m.ntot;
         int c0 = m.off[a], np = m.Np[a];
         long cp = (c - c0 + 1 == np) ? c0 : c + 1;
         double pc[NQ], pp[NQ];
         load_q<NQ>(prim, nt, c, pc);
         load_q<NQ>(prim, nt, cp, pp);
         int j = a % m.Nr, k = a / m.Nr;
         double r = centroid_r(m.r_jph[j], m.r_jph[j - 1]);
         double z = centroid_z(m.z_kph[k], m.z_kph[k - 1]);
         w = .5 * (H::omega(P, pc, r, z) + H::omega(P, pp, r, z));
      }
      wiph[c] = w;
   }
};

// Mesh_Motion 3: every face of an annulus moves with the annulus mean.
struct KWcellAvg {
   Mesh m;
   double *wiph;
   DHD void operator()(long a) const
   {
      int c0 = m.off[a], np = m.Np[a];
      double w = 0.0;
      for (int i = 0; i < np; ++i) w += wiph[c0 + i];
      w /= (double)np;
      for (int i = 0; i < np; ++i) wiph[c0 + i] = w;
   }
};

// ---------------------------------------------------------------------------
// Generalised-minmod slope along phi (plm.c:6-51)
// ---------------------------------------------------------------------------

// The limited phi slopes of cell c from its two phi neighbours.
template <int NQ>
DHD void plm_phi_cell(const Mesh &m, double PLM, const double *prim, const double *dphi, long c,
                      double *g)
{
   size_t nt = (size_t)m.ntot;
   int a = m.cjk[c];
   int c0 = m.off[a], np = m.Np[a];
   int i = (int)(c - c0);
   long cL = i == 0 ? c0 + np - 1 : c - 1;
   long cR = i == np - 1 ? c0 : c + 1;
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
      g[q] = minmod(PLM * sL, sC, PLM * sR);
   }
}

template <int NQ>
struct KPlmPhi {
   Mesh m;
   double PLM;
   const double *prim, *dphi;
   double *gradp;
   DHD void operator()(long c) const
   {
      size_t nt = (size_t)m.ntot;
      double g[NQ];
      plm_phi_cell<NQ>(m, PLM, prim, dphi, c, g);
#pragma unroll
      for (int q = 0; q < NQ; ++q) gradp[q * nt + c] = g[q];
   }
};

// ---------------------------------------------------------------------------
// Faces (faces_alt.c:10-34, 135-175)
// ---------------------------------------------------------------------------

struct FaceGeo {
   int cL, cR, LR;
   double phic, dp, dA;
};

// addFace (faces_alt.c:10-34, 135-175) for the face of pair (j, k) between
// cells of upper phi edges pL, pR and extents dL, dR (L in the inner or lower
// annulus): its orientation, phi extent, phi centre and area.  The slope
// and flux kernels read the stored centre and area (formed there again
// from the positions, the r-face fluxes ran 1.6 times slower).
DHD void face_shape(const Mesh &m, int dim, int j, int k, double pL, double dL, double pR,
                    double dR, int *LR, double *dp, double *phic, double *dA)
{
   double Pmax = m.Pmax;
   double w = wrap_half(pL - pR, Pmax);
   *LR = w > 0. ? 1 : 0;
   w = wrap_half(pL - dL - pR + dR, Pmax);
   int LRb = w < 0. ? 1 : 0;
   double xp1 = *LR == 0 ? pL : pR;
   double xm1 = LRb == 0 ? pL - dL : pR - dR;
   *dp = get_dp(xp1, xm1, Pmax);
   *phic = get_dp(xp1, .5 * *dp, Pmax);
   if (dim == 1) {
      double r = .5 * (m.r_jph[j] + m.r_jph[j]);
      double dz = m.z_kph[k] - m.z_kph[k - 1];
      *dA = r * *dp * dz;
   } else {
      double r = .5 * (m.r_jph[j] + m.r_jph[j - 1]);
      double dr = m.r_jph[j] - m.r_jph[j - 1];
      *dA = r * dr * *dp;
   }
}

// The same for the face between cells cL and cR of pair (j, k).
DHD FaceGeo face_geometry(const Mesh &m, int dim, int j, int k, int cL, int cR,
                          const double *piph, const double *dphi)
{
   FaceGeo g;
   g.cL = cL;
   g.cR = cR;
   face_shape(m, dim, j, k, piph[cL], dphi[cL], piph[cR], dphi[cR], &g.LR, &g.dp, &g.phic,
              &g.dA);
   return g;
}

// Distances from the two cell centroids to the faces of a pair (faces_alt.c:107-123).
DHD void pair_dx(const Mesh &m, int dim, int j, int k, double *dxL, double *dxR)
{
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

// Every r (dim 1) or z (dim 2) face (faces_alt.c:36-179).  One thread per
// cell of the inner (lower) annulus of a pair writes the faces the cell
// bounds there: face h of that range lies between the cell and element
// h - x of the outer ring (x = the cell's element, or the ring length for
// the start cell's tail).
template <int DIM>
struct KFaces {
   Mesh m;
   static constexpr int dim = DIM;
   static constexpr bool xcd_contiguous = true;
   Faces F;
   const int *rng;
   const double *piph, *dphi;
   DHD void face(int f, int j, int k, int cL, int cR) const
   {
      FaceGeo g = face_geometry(m, dim, j, k, cL, cR, piph, dphi);
      F.L[f] = g.cL;
      F.R[f] = g.cR;
      F.LR[f] = (signed char)g.LR;
      F.up[f] = piph[g.LR == 0 ? g.cL : g.cR];
      F.dA[f] = g.dA;
      F.phic[f] = g.phic;
      F.dphi[f] = g.dp;
   }
   DHD void operator()(long c) const
   {
      int a = m.cjk[c];
      int j = a % m.Nr, k = a / m.Nr;
      if (dim == 1 ? j >= m.Nr - 1 : k >= m.Nz - 1) return;
      CellRanges cr = load_ranges(m, dim, rng, F.fI, c, a);
      int b = dim == 1 ? a + 1 : a + m.Nr;
      int bN = m.Np[b], b0 = m.off[b], bI0 = m.I0[b];
      int aN = m.Np[a];
      int x = (int)(c - m.off[a]) - m.I0[a];
      if (x < 0) x += aN;
      for (int part = 0; part < 2; ++part) {
         int lo = part == 0 ? cr.out_h0 : cr.out_t0, hi = part == 0 ? cr.out_h1 : cr.out_t1;
         int consumed = part == 0 ? x : aN;
         for (int h = lo; h <= hi; ++h) {
            int y = h - consumed;
            if (y == bN) y = 0;
            int ib = bI0 + y;
            if (ib >= bN) ib -= bN;
            face(cr.base_out + h, j, k, (int)c, b0 + ib);
         }
      }
   }
};

// The merge steps at which each cell is consumed in its inner and outer pair
// of one direction (one merge-path search each, dh_mesh.h); every face range
// of the cell follows from them (load_ranges, dh_ct.h).
template <int DIM>
struct KCellRanges {
   Mesh m;
   static constexpr int dim = DIM;
   const double *piph;
   int *rng;
   DHD void operator()(long c) const
   {
      size_t nt = (size_t)m.ntot;
      int a = m.cjk[c];
      int i = (int)(c - m.off[a]);
      int j = a % m.Nr, k = a / m.Nr;
      Ring self = make_ring(piph, m, m.I0, a);
      int xs = ring_elem(self, i);
      bool in = dim == 1 ? j > 0 : k > 0;
      bool out = dim == 1 ? j < m.Nr - 1 : k < m.Nz - 1;
      int s_in = 0, s_out = 0;
      if (in) {
         Ring A = make_ring(piph, m, m.I0, dim == 1 ? a - 1 : a - m.Nr);
         s_in = b_step(A, self, xs, m.Pmax);
      }
      if (out) {
         Ring B = make_ring(piph, m, m.I0, dim == 1 ? a + 1 : a + m.Nr);
         s_out = a_step(self, xs, B, m.Pmax);
      }
      rng[c] = s_in;
      rng[nt + c] = s_out;
   }
};

// ---------------------------------------------------------------------------
// r/z slope of every cell: area-weighted mean of its face slopes over the
// analytic face area, limited by every face slope (plm.c:56-195); cells on
// the r = 0 axis get the polar limiter (plm.c:197-232, cylindrical.c:141-173).
// ---------------------------------------------------------------------------

template <int NQ, int DIM>
struct KGrad {
   Mesh m;
   static constexpr int dim = DIM;
   static constexpr bool xcd_contiguous = DIM == 1;
   double PLM;
   int num_c;
   Faces F;
   const int *rng;
   const double *piph, *dphi, *prim, *gradp;
   double *grad;
   // The limiter of plm.c:160-195 visits the face slopes s in order with
   //   if (s g < 0) g = 0; else if (|PLM s| < |g|) g = PLM s;
   // its outcome is 0 if any s has the sign opposite to the mean g0, else
   // the first PLM s of least magnitude when that is below |g0|, else g0
   // (NaN slopes never act).  So one pass over the faces keeps, per field,
   // the area-weighted sum, the first least PLM s, and which signs occurred.
   struct Acc {
      double sum[NQ], best[NQ];
      unsigned pos = 0, neg = 0;
   };
   // What one face brings from the other cell o: o's phi edge and extent
   // and its state (prim, gradp).
   struct FaceIn {
      long o;
      double phic, dA, po, dpo, p[NQ], g[NQ];
   };
   DHD void load_face(int f, bool own_left, FaceIn &d) const
   {
      size_t nt = (size_t)m.ntot;
      d.o = own_left ? F.R[f] : F.L[f];
      d.phic = F.phic[f];
      d.dA = F.dA[f];
      d.po = piph[d.o];
      d.dpo = dphi[d.o];
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
         d.p[q] = prim[q * nt + d.o];
         d.g[q] = gradp[q * nt + d.o];
      }
   }
   // A face of pair (pj, pk) whose distances from the centroids are dxL,
   // dxR; the cell is its L (own_left) or R, with state (pc, gc) and phi
   // edge and extent (pcc, dpc).  The one-sided slope is
   // S = (W_R - W_L)/(dx_L + dx_R) of the phi-reconstructed states at the
   // face's phi centre (plm.c:84-110), weighted by the face's area.
   DHD void add_face(const FaceIn &d, bool own_left, int pj, int pk, double pcc, double dpc,
                     const double *pc, const double *gc, double dxL, double dxR, Acc &A) const
   {
      double pL = own_left ? pcc : d.po, dL = own_left ? dpc : d.dpo;
      double pR = own_left ? d.po : pcc, dR = own_left ? d.dpo : dpc;
      double phic = d.phic, dA = d.dA;
      double dpL = get_sdp(phic, pL - .5 * dL, m.Pmax);
      double dpR = get_sdp(pR - .5 * dR, phic, m.Pmax);
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
         double WL = own_left ? pc[q] + dpL * gc[q] : d.p[q] + dpL * d.g[q];
         double WR = own_left ? d.p[q] - dpR * d.g[q] : pc[q] - dpR * gc[q];
         double S = (WR - WL) / (dxR + dxL);
         A.sum[q] += S * dA;
         double v = PLM * S;
         if (fabs(v) < fabs(A.best[q])) A.best[q] = v;
         if (S > 0.0) A.pos |= 1u << q;
         if (S < 0.0) A.neg |= 1u << q;
      }
   }
   DHD void operator()(long c) const
   {
      size_t nt = (size_t)m.ntot;
      int a = m.cjk[c];
      int j = a % m.Nr, k = a / m.Nr;
      CellRanges cr = load_ranges(m, dim, rng, F.fI, c, a);
      double dxLi = 0, dxRi = 0, dxLo = 0, dxRo = 0;
      if (cr.in) pair_dx(m, dim, dim == 1 ? j - 1 : j, dim == 1 ? k : k - 1, &dxLi, &dxRi);
      if (cr.out) pair_dx(m, dim, j, k, &dxLo, &dxRo);
      double pc[NQ], gc[NQ];
      load_q<NQ>(prim, nt, c, pc);
      load_q<NQ>(gradp, nt, c, gc);
      double pcc = piph[c], dpc = dphi[c];
      Acc A;
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
         A.sum[q] = 0.0;
         A.best[q] = HUGE_VAL;
      }
      // the reference's order: the inner pair (the cell is R), then the outer
      // pair (the cell is L), each head before tail; one face ahead (the
      // kernel waits on its gathers: two faces in flight per thread beat
      // the occupancy the extra registers cost)
      walk_ahead<FaceIn>(
          pair_seq(cr), [&](int f, int r, FaceIn &d) { load_face(f, r >= 2, d); },
          [&](const FaceIn &d, int r) {
             bool own_left = r >= 2;
             // the outer pair is (j, k); the inner one (j-1, k) or (j, k-1)
             int pj = own_left || dim == 2 ? j : j - 1, pk = own_left || dim == 1 ? k : k - 1;
             add_face(d, own_left, pj, pk, pcc, dpc, pc, gc, own_left ? dxLo : dxLi,
                      own_left ? dxRo : dxRi, A);
          });
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
      if (!cr.in) dAtot = dAp;
      if (!cr.out) dAtot = dAm;
      double g[NQ];
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
         double g0 = A.sum[q] / dAtot;
         bool opp = (g0 > 0.0 && (A.neg >> q & 1u)) || (g0 < 0.0 && (A.pos >> q & 1u));
         g[q] = opp ? 0.0 : fabs(A.best[q]) < fabs(g0) ? A.best[q] : g0;
      }
      if (dim == 1 && j == 0 && m.NgRa == 0) {
         double xp0 = m.r_jph[0], xm0 = m.r_jph[-1];
         if (!(xp0 < 0.0 || fabs(xm0) > 1.0e-10 * xp0)) {
            double r = centroid_r(xp0, xm0);
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
               if (q == URR || (num_c > BZZ && q == BRR)) {
                  double SL = pc[q] / r;
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
};

// ---------------------------------------------------------------------------
// Riemann fluxes
// ---------------------------------------------------------------------------

// Viscous flux of an interface (riemann.c:291-309)
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

// Flux through the face between cell i and i+1 of every annulus in range
// (misc.c:368-426, riemann.c:39-112), stored as (F - w U*) dA dt; for MHD
// also the face's {Ez, Br, Er, Bz} (riemann.c:66-109).
template <int NQ, class H>
struct KFluxPhi {
   Mesh m;
   dh_params P;
   const double *dtp;
   const double *prim, *gradp, *gradr, *gradz, *piph, *dphi, *wiph;
   double *dFp, *emf_phi;
   DHD bool in_range(int j, int k) const { return phi_in_range(m, IsMHD<H>::value && P.ct, j, k); }
   DHD void operator()(long c) const
   {
      int a = m.cjk[c];
      int j = a % m.Nr, k = a / m.Nr;
      if (!in_range(j, k)) return;
      size_t nt = (size_t)m.ntot;
      int c0 = m.off[a], np = m.Np[a];
      long cR = (c - c0 == np - 1) ? c0 : c + 1;
      double gC[NQ], gR[NQ];
      load_q<NQ>(gradp, nt, c, gC);
      load_q<NQ>(gradp, nt, cR, gR);
      flux(c, cR, j, k, gC, gR);
   }
   // The face between c and its upper phi neighbour cR, whose phi slopes are
   // gC and gR.
   DHD void flux(long c, long cR, int j, int k, const double *gC, const double *gR) const
   {
      constexpr bool mhd = IsMHD<H>::value;
      size_t nt = (size_t)m.ntot;
      double rp = m.r_jph[j], rm = m.r_jph[j - 1];
      double zp = m.z_kph[k], zm = m.z_kph[k - 1];
      double r = centroid_r(rp, rm);
      double dAdt = (rp - rm) * (zp - zm) * *dtp;
      double dL = dphi[c], dR = dphi[cR];
      double pL[NQ], pR[NQ];
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
         pL[q] = prim[q * nt + c] + .5 * gC[q] * dL;
         pR[q] = prim[q * nt + cR] - .5 * gR[q] * dR;
      }
      if (mhd) {
         double Bp = .5 * (pL[BPP] + pR[BPP]);
         pL[BPP] = Bp;
         pR[BPP] = Bp;
      }
      const double n[3] = {0.0, 1.0, 0.0};
      double w = r * wiph[c];
      double Flux[NQ], Ustr[NQ];
      H::riemann(P, pL, pR, r, centroid_z(zp, zm), n, w, Flux, Ustr);
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
      for (int q = 0; q < NQ; ++q)
         if (flux_updates<H>(q)) dFp[q * nt + c] = (Flux[q] - w * Ustr[q]) * dAdt;
      if constexpr (mhd) {
         double e[4];
         H::flux_to_E(Flux, Ustr, r, 0, e);
#pragma unroll
         for (int s = 0; s < 4; ++s) emf_phi[s * nt + c] = e[s];
      }
   }
};

// Flux through every r (dim 1) or z (dim 2) face of the flux range
// (riemann.c:114-213), stored as F dA dt (these faces do not move: w = 0);
// for MHD also the face's {E_rz, B_rz, E_phi}.
template <int NQ, class H, int DIM>
struct KFluxTrans {
   Mesh m;
   dh_params P;
   static constexpr int dim = DIM;
   const double *dtp;
   Faces F;
   const double *prim, *gradp, *gradr, *gradz, *piph, *dphi;
   double *dF, *emf_face;
   DHD void operator()(long f) const
   {
      constexpr bool mhd = IsMHD<H>::value;
      int Pi = F.pf[f];
      Pair pr = pair_of(m, dim, Pi);
      if (!pair_has_flux(m, mhd && P.ct, dim, pr.j, pr.k)) return;
      size_t nt = (size_t)m.ntot, nfs = (size_t)F.n;
      int cL = F.L[f], cR = F.R[f];
      double dxL, dxR;
      pair_dx(m, dim, pr.j, pr.k, &dxL, &dxR);
      double phi = F.phic[f];
      double dpL = get_sdp(phi, piph[cL] - .5 * dphi[cL], m.Pmax);
      double dpR = get_sdp(piph[cR] - .5 * dphi[cR], phi, m.Pmax);
      const double *grad = dim == 1 ? gradr : gradz;
      double pL[NQ], pR[NQ];
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
         pL[q] = prim[q * nt + cL] + grad[q * nt + cL] * dxL + gradp[q * nt + cL] * dpL;
         pR[q] = prim[q * nt + cR] - grad[q * nt + cR] * dxR - gradp[q * nt + cR] * dpR;
      }
      if (mhd) {
         int bt = dim == 1 ? BRR : BZZ;
         double Bavg = .5 * (pL[bt] + pR[bt]);
         pL[bt] = Bavg;
         pR[bt] = Bavg;
      }
      double n[3] = {0.0, 0.0, 0.0};
      n[dim == 1 ? 0 : 2] = 1.0;
      int j = pr.j;
      // the face centroid (faces_alt.c:11-18)
      int k = pr.k;
      double r = dim == 1 ? .5 * (m.r_jph[j] + m.r_jph[j]) : centroid_r(m.r_jph[j], m.r_jph[j - 1]);
      double z = dim == 2 ? .5 * (m.z_kph[k] + m.z_kph[k]) : centroid_z(m.z_kph[k], m.z_kph[k - 1]);
      double Flux[NQ], Ustr[NQ];
      H::riemann(P, pL, pR, r, z, n, 0.0, Flux, Ustr);
      if (P.visc_flag) {
         double gr[NQ], gp[NQ], gz[NQ];
#pragma unroll
         for (int q = 0; q < NQ; ++q) {
            double gpL = gradp[q * nt + cL], gpR = gradp[q * nt + cR];
            double gd = ((prim[q * nt + cR] - gpR * dpR) - (prim[q * nt + cL] + gpL * dpL)) / (dxL + dxR);
            gp[q] = .5 * (gpL + gpR);
            if (dim == 1) {
               gr[q] = .5 * (gd + gd);
               gz[q] = .5 * (gradz[q * nt + cL] + gradz[q * nt + cR]);
            } else {
               gr[q] = .5 * (gradr[q * nt + cL] + gradr[q * nt + cR]);
               gz[q] = .5 * (gd + gd);
            }
         }
         add_visc<NQ, H>(P, pL, pR, gr, gp, gz, r, n, Flux);
      }
      double dAdt = F.dA[f] * *dtp;
#pragma unroll
      for (int q = 0; q < NQ; ++q)
         if (flux_updates<H>(q)) dF[q * nfs + f] = (Flux[q] - 0.0 * Ustr[q]) * dAdt;
      if constexpr (mhd) {
         double e[4];
         H::flux_to_E(Flux, Ustr, r, dim, e);
         emf_face[0 * nfs + f] = e[0];
         emf_face[1 * nfs + f] = e[1];
         emf_face[2 * nfs + f] = e[2];
      }
   }
};

// ---------------------------------------------------------------------------
// Point-mass gravity (planet.c:56-196)
// ---------------------------------------------------------------------------

// n is PHI_ORDER (planet.c:4), read at run time as the reference does: with
// a constant the compiler would turn pow(r, 2.0) into r * r, which is not
// always glibc's pow.
DHD double fgrav(double M, double r, double eps, int type, double n)
{
   if (type == 0) {
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
   // sin and cos of one angle come from one sincos call, as the reference's
   // optimised build computes them.
   double sinp, cosp, sinpl, cospl;
   sincos(phi, &sinp, &cosp);
   sincos(ppl, &sinpl, &cospl);
   double dx = r * cosp - rpl * cospl;
   double dy = r * sinp - rpl * sinpl;
   double script_r = sqrt(dx * dx + dy * dy + z * z);
   double script_r_perp = sqrt(dx * dx + dy * dy);
   double f1 = -fgrav(pl[2], script_r, pl[5], (int)pl[6], P.phi_order);
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

// omega_src (rotframe.c:32-82) in cylindrical coordinates: the centrifugal
// and Coriolis forces of a frame rotating at RotOmega about an axis at
// distance RotD, at the cell's centroid (get_centroid_arr).  The vector
// conversions of the reference are kept step by step (to the orthonormal
// basis and back), so that the sums round as there.
DHD void omega_src(const dh_params &P, const double *prim, double *cons, double rp, double rm,
                   double phip, double phim, double zp, double zm, double dVdt)
{
   double r = centroid_r(rp, rm);
   double phi = 0.5 * (phip + phim);
   double rho = prim[RHO];
   // get_vec_covariant (orthonormal), then get_vec_contravariant
   double v[3] = {prim[URR], prim[UPP] * r, prim[UZZ]};
   double vrpz[3] = {v[0] / 1.0, v[1] / r, v[2] / 1.0};
   // omegaForce (rotframe.c:19-30); cos and sin of one angle from one sincos
   // call, as the reference's optimised build evaluates them
   double O0 = P.rot_omega, d = P.rot_d;
   double sn, cs;
   sincos(phi, &sn, &cs);
   double Rr = r - d * cs;
   double Rp = d * sn;
   double f[3] = {O0 * O0 * Rr + 2. * O0 * r * vrpz[1], O0 * O0 * Rp - 2. * O0 * vrpz[0], 0.0};
   // the energy-frame velocity, still orthonormal
   v[1] -= r * get_om(P, r);
   double vf = v[0] * f[0] + v[1] * f[1] + v[2] * f[2];
   // back to covariant components
   f[0] = f[0] * 1.0;
   f[1] = f[1] * r;
   f[2] = f[2] * 1.0;
   cons[SRR] += rho * f[0] * dVdt;
   cons[LLL] += rho * f[1] * dVdt;
   cons[SZZ] += rho * f[2] * dVdt;
   cons[TAU] += rho * vf * dVdt;
}

// sink_src (sink.c:35-70), Nozzle_Type 1: mass injected at a uniform rate
// (Mdot = 1) into a sphere of radius dR about (R0, 0, 0), or the disk of its
// cross-section in 2-D, with azimuthal speed v and sound speed v / Mach;
// cell midpoints (not centroids), phi taken into (-pi, pi].
template <int NQ, int NC>
DHD void sink_src(const dh_params &P, bool twoD, double *cons, double rp, double rm, double phip,
                  double phim, double zp, double zm, double dVdt)
{
   if (P.nozzle_type != 1) return;
   double R0 = P.nozzle_par[0], dR = P.nozzle_par[1], v = P.nozzle_par[2],
          mach = P.nozzle_par[3];
   double r = 0.5 * (rp + rm);
   double phi = 0.5 * (phip + phim);
   double z = 0.5 * (zp + zm);
   if (phi > M_PI) phi -= 2 * M_PI;
   if ((r - R0) * (r - R0) + r * r * phi * phi + z * z > dR * dR) return;
   double Mdot = 1.0;
   double rhodot = Mdot * 3.0 / (4.0 * M_PI * dR * dR * dR);
   if (twoD) rhodot *= 2 * sqrt(dR * dR - (r - R0) * (r - R0) - r * r * phi * phi);
   if (rhodot * dVdt > cons[DDD]) rhodot = cons[DDD] / dVdt;
   double om = get_om(P, r);
   double cs2 = v * v / (mach * mach);
   double gl = P.gamma_law;
   double eps = cs2 / (gl * (gl - 1));
   cons[DDD] += rhodot * dVdt;
   cons[LLL] += rhodot * r * v * dVdt;
   cons[TAU] += (0.5 * rhodot * (v - r * om) * (v - r * om) + rhodot * eps) * dVdt;
   if constexpr (NQ > NC) cons[NC] += rhodot * dVdt;
}

// ---------------------------------------------------------------------------
// Cell update: RK blend (misc.c:226-244), flux gather in the reference's
// order (phi, then r, then z faces; each cell's inner pair before its outer
// pair), sources (misc.c:538-580), mesh motion (misc.c:256-288) and, for
// hydro, cons2prim (misc.c:290-325).  Ghost cells only move.
// ---------------------------------------------------------------------------

template <int NQ, class H>
DHD void gather_range(const double *dF, size_t nfs, int base, int lo, int hi, bool add,
                      double *cons)
{
   for (int f = base + lo; f <= base + hi; ++f) {
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
         if (!flux_updates<H>(q)) continue;
         if (add) cons[q] += dF[q * nfs + f];
         else cons[q] -= dF[q * nfs + f];
      }
   }
}

template <int NQ, class H>
DHD void gather_pairs(const Mesh &m, int dim, long c, int a, const int *fI, const int *rng,
                      int nf, const double *dF, double *cons)
{
   CellRanges cr = load_ranges(m, dim, rng, fI, c, a);
   size_t nfs = (size_t)nf;
   if (cr.in) {
      gather_range<NQ, H>(dF, nfs, cr.base_in, cr.in_h0, cr.in_h1, true, cons);
      gather_range<NQ, H>(dF, nfs, cr.base_in, cr.in_t0, cr.in_t1, true, cons);
   }
   if (cr.out) {
      gather_range<NQ, H>(dF, nfs, cr.base_out, cr.out_h0, cr.out_h1, false, cons);
      gather_range<NQ, H>(dF, nfs, cr.base_out, cr.out_t0, cr.out_t1, false, cons);
   }
}

template <int NQ, class H>
struct KUpdate {
   Mesh m;
   dh_params P;
   double RK;
   const double *dtp;
   int first_step;
   const PlanetSet *pls;
   const double *prim_in;
   double *prim;
   // the substep's conserved state in, the RK blend's other state, and the
   // updated state out (cons_out may be cons)
   const double *cons;
   double *cons_out;
   const double *rkcons, *gradr, *gradp, *gradz, *piph, *dphi, *wiph;
   double *piph_out, *dphi_out;
   const double *dFp, *dFr, *dFz;
   const int *fIr, *fIz, *rngr, *rngz;
   int nfr, nfz;
   DHD void operator()(long c) const
   {
      constexpr bool mhd = IsMHD<H>::value;
      size_t nt = (size_t)m.ntot;
      int a = m.cjk[c];
      int c0 = m.off[a], np = m.Np[a];
      int i = (int)(c - c0);
      int j = a % m.Nr, k = a / m.Nr;
      long cm = i == 0 ? c0 + np - 1 : c - 1;
      double Pmax = m.Pmax;
      const double dt = *dtp;

      // Mesh motion of this cell and of its phi neighbour: move_cells on the
      // first substep only, clean_pi and calc_dp on every substep.
      double pp = piph[c], pm = piph[cm];
      if (first_step) {
         pp += wiph[c] * dt;
         pm += wiph[cm] * dt;
      }
      pp = clean_phi(pp, Pmax);
      pm = clean_phi(pm, Pmax);
      double dp_new = get_dp(pp, pm, Pmax);
      piph_out[c] = pp;
      dphi_out[c] = dp_new;
      if (!is_real(m, j, k)) return;

      // the field components (constrained transport) are left to
      // KBfacesC2p, which sets them from the faces' magnetic fluxes
      double cs[NQ];
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
         cs[q] = 0.0;
         if (!flux_updates<H>(q)) continue;
         double v = cons[q * nt + c];
         if (RK != 0.0) v = (1. - RK) * v + RK * rkcons[q * nt + c];
         cs[q] = v;
      }
      // phi faces: cell i is R of face i-1 and L of face i; the reference
      // reaches face i-1 first except for i = 0, whose R face is the last one.
      if (i == 0) {
#pragma unroll
         for (int q = 0; q < NQ; ++q)
            if (flux_updates<H>(q)) cs[q] -= dFp[q * nt + c];
#pragma unroll
         for (int q = 0; q < NQ; ++q)
            if (flux_updates<H>(q)) cs[q] += dFp[q * nt + cm];
      } else {
#pragma unroll
         for (int q = 0; q < NQ; ++q)
            if (flux_updates<H>(q)) cs[q] += dFp[q * nt + cm];
#pragma unroll
         for (int q = 0; q < NQ; ++q)
            if (flux_updates<H>(q)) cs[q] -= dFp[q * nt + c];
      }
      if (m.Nr > 1) gather_pairs<NQ, H>(m, 1, c, a, fIr, rngr, nfr, dFr, cs);
      if (m.Nz > 1) gather_pairs<NQ, H>(m, 2, c, a, fIz, rngz, nfz, dFz, cs);

      // Sources at the cell's position before the move.
      double rp = m.r_jph[j], rm = m.r_jph[j - 1];
      double zp = m.z_kph[k], zm = m.z_kph[k - 1];
      double phip = piph[c];
      double phim = phip - dphi[c];
      double dV = cell_dV(rp, rm, phip, phim, zp, zm, Pmax);
      double dVdt = dV * dt;
      int polar = mhd ? (m.NgRa == 0) : (P.nobc_rmin == 1);
      double pr[NQ];
      load_q<NQ>(prim_in, nt, c, pr);
      H::source(P, pr, cs, rp, rm, phip, phim, zp, zm, dVdt, polar, Pmax);
      for (int p = 0; p < pls->n; ++p)
         planet_src(P, pls->v + p * DH_PLANET_STRIDE, pr, cs, rp, rm, phip, phim, zp, zm, dVdt,
                    P.nobc_rmin == 1, Pmax);
      if (P.visc_flag) {
         double gr[NQ], gp[NQ], gz[NQ];
         load_q<NQ>(gradr, nt, c, gr);
         load_q<NQ>(gradp, nt, c, gp);
         load_q<NQ>(gradz, nt, c, gz);
         H::visc_source(P, pr, gr, gp, gz, cs, centroid_r(rp, rm), dVdt);
      }
      if (P.rot_frame) omega_src(P, pr, cs, rp, rm, phip, phim, zp, zm, dVdt);
      sink_src<NQ, H::NC>(P, m.Nz == 1, cs, rp, rm, phip, phim, zp, zm, dVdt);
#pragma unroll
      for (int q = 0; q < NQ; ++q)
         if (flux_updates<H>(q)) cons_out[q * nt + c] = cs[q];
      if constexpr (!mhd) {
         // cons2prim at the moved position.
         double r = centroid_r(rp, rm);
         double dVn = cell_dV(rp, rm, pp, pp - dp_new, zp, zm, Pmax);
         double out[NQ];
         H::cons2prim(P, cs, out, r, centroid_z(zp, zm), dVn);
#pragma unroll
         for (int q = 0; q < NQ; ++q) prim[q * nt + c] = out[q];
      }
   }
};

// ---------------------------------------------------------------------------
// Constrained transport (CT_MODE 2), see dh_ct.h for the state layout.
// Edge slots per cell: E[0..3] z-directed at (r-,phi+), (r+,phi+), (r-,phi-),
// (r+,phi-); E[4..7] r-directed at (z-,phi+), (z+,phi+), (z-,phi-), (z+,phi-).
// Ed[s*nt+c] holds E[s], Ed[(8+s)*nt+c] the matching normal field B[s].
// ---------------------------------------------------------------------------

// Edge EMFs of every cell after the Riemann solves (riemann.c:66-211): the
// phi faces set E[0..7] (= half their EMF), then every r and z face of the
// flux range adds half its EMF, weighted by its phi extent over the cell's.
// Cells outside the phi-flux range keep the values they ended the previous
// substep with (Ecarry), as the reference's cell arrays do.
struct KEdgesCollect {
   Mesh m;
   int ct;
   const double *dphi, *emf_phi;
   Faces Fr, Fz;
   const int *rngr, *rngz;
   const double *emf_r, *emf_z, *Ecarry;
   double *Epre;
   DHD void operator()(long c) const
   {
      size_t nt = (size_t)m.ntot;
      int a = m.cjk[c];
      int c0 = m.off[a], np = m.Np[a];
      long cm = c == c0 ? c0 + np - 1 : c - 1;
      int j = a % m.Nr, k = a / m.Nr;
      double E[8], B[8];
      if (phi_in_range(m, ct, j, k)) {
         double Ez = emf_phi[0 * nt + c], Br = emf_phi[1 * nt + c];
         double Er = emf_phi[2 * nt + c], Bz = emf_phi[3 * nt + c];
         double Ezm = emf_phi[0 * nt + cm], Brm = emf_phi[1 * nt + cm];
         double Erm = emf_phi[2 * nt + cm], Bzm = emf_phi[3 * nt + cm];
         E[0] = E[1] = .5 * Ez;
         E[2] = E[3] = .5 * Ezm;
         B[0] = B[1] = .5 * Br;
         B[2] = B[3] = .5 * Brm;
         E[4] = E[5] = .5 * Er;
         E[6] = E[7] = .5 * Erm;
         B[4] = B[5] = .5 * Bz;
         B[6] = B[7] = .5 * Bzm;
      } else {
#pragma unroll
         for (int s = 0; s < 8; ++s) {
            E[s] = Ecarry[s * nt + c];
            B[s] = Ecarry[(8 + s) * nt + c];
         }
      }
      double dpc = dphi[c];
      for (int dim = 1; dim <= 2; ++dim) {
         if (!dim_active(m, dim)) continue;
         const Faces &F = dim == 1 ? Fr : Fz;
         const double *em = dim == 1 ? emf_r : emf_z;
         size_t nfs = (size_t)F.n;
         CellRanges cr = load_ranges(m, dim, dim == 1 ? rngr : rngz, F.fI, c, a);
         // as R of the inner pair: edges 0,2 (r) or 4,6 (z);
         // as L of the outer pair: edges 1,3 (r) or 5,7 (z).
         int i0 = dim == 1 ? 0 : 4, o0 = dim == 1 ? 1 : 5;
         auto acc = [&](int base, int lo, int hi, int e0) {
            for (int f = base + lo; f <= base + hi; ++f) {
               double frac = F.dphi[f] / dpc;
               double de = .5 * em[0 * nfs + f] * frac;
               double db = .5 * em[1 * nfs + f] * frac;
               E[e0] += de;
               E[e0 + 2] += de;
               B[e0] += db;
               B[e0 + 2] += db;
            }
         };
         if (cr.in) {
            int pj = dim == 1 ? j - 1 : j, pk = dim == 1 ? k : k - 1;
            if (pair_has_flux(m, ct, dim, pj, pk)) {
               acc(cr.base_in, cr.in_h0, cr.in_h1, i0);
               acc(cr.base_in, cr.in_t0, cr.in_t1, i0);
            }
         }
         if (cr.out && pair_has_flux(m, ct, dim, j, k)) {
            acc(cr.base_out, cr.out_h0, cr.out_h1, o0);
            acc(cr.base_out, cr.out_t0, cr.out_t1, o0);
         }
      }
#pragma unroll
      for (int s = 0; s < 8; ++s) {
         Epre[s * nt + c] = E[s];
         Epre[(8 + s) * nt + c] = B[s];
      }
   }
};

// avg_Efields, first part (bfields.c:400-441), evaluated where it is read:
// each phi-edge takes the mean of the values its two cells hold.  Slot s of
// E (h = 0) or B (h = 1) of cell c.
DHD double edge_avg(const Mesh &m, const double *Epre, int h, int s, long c)
{
   size_t nt = (size_t)m.ntot;
   const double *X = Epre + (size_t)h * 8 * nt;
   int a = m.cjk[c];
   int c0 = m.off[a], np = m.Np[a];
   int e = s & 3, base = s & 4;
   if (e < 2) {
      long cp = c - c0 == np - 1 ? c0 : c + 1;
      return .5 * (X[(base + e) * nt + c] + X[(base + e + 2) * nt + cp]);
   }
   long cm = c == c0 ? c0 + np - 1 : c - 1;
   return .5 * (X[(base + e - 2) * nt + cm] + X[(base + e) * nt + c]);
}

// avg_Efields, second part (bfields.c:443-531): the edge at the upper end of
// face f (of direction dim) is the mean of its owner's value and the other
// cell's two corner values interpolated in phi; E in *fE, B in *fB.  Each
// face's edge is read by one cell, the face's owner (edges_post_face), which
// forms it where it reads it.
DHD void face_emf(const Mesh &m, int dim, const Faces &F, const double *piph,
                  const double *dphi, const double *Epre, int f, double *fE, double *fB)
{
   int L = F.L[f], R = F.R[f];
   int LR = F.LR[f];
   int c1 = LR == 0 ? L : R, c2 = LR == 0 ? R : L;
   double dp1 = get_dp(piph[c2], piph[c1], m.Pmax);
   double dp2 = dphi[c2] - dp1;
   int eL = dim == 1 ? 1 : 5, eR = dim == 1 ? 0 : 4;
   // LRtype 0: the face ends at L's upper edge; R's corners are interpolated
   int eo = LR == 0 ? eL : eR, ec = LR == 0 ? eR : eL;
   int co = LR == 0 ? L : R;
   double Ea = (dp2 * edge_avg(m, Epre, 0, ec, c2) + dp1 * edge_avg(m, Epre, 0, ec + 2, c2)) / (dp1 + dp2);
   double Ba = (dp2 * edge_avg(m, Epre, 1, ec, c2) + dp1 * edge_avg(m, Epre, 1, ec + 2, c2)) / (dp1 + dp2);
   *fE = .5 * (edge_avg(m, Epre, 0, eo, co) + Ea);
   *fB = .5 * (edge_avg(m, Epre, 1, eo, co) + Ba);
}

// The four phi+ edges of a cell after avg_Efields, before the advective
// subtraction: {E0, E1, E4, E5} then {B0, B1, B4, B5}.  An edge bounding
// the face the cell owns takes that face's value (bfields.c:533-560).
struct EdgeArgs {
   Mesh m;
   Faces Fr, Fz;
   const int *rngr, *rngz;
   const double *Epre, *piph, *dphi;
};

DHD void edges_post_face(const EdgeArgs &g, long c, double *out)
{
   const Mesh &m = g.m;
   size_t nt = (size_t)m.ntot;
   int a = m.cjk[c];
   const int slots[4] = {0, 1, 4, 5};
#pragma unroll
   for (int s = 0; s < 4; ++s) {
      out[s] = edge_avg(m, g.Epre, 0, slots[s], c);
      out[4 + s] = edge_avg(m, g.Epre, 1, slots[s], c);
   }
   if (m.Nr > 1) {
      CellRanges cr = load_ranges(m, 1, g.rngr, g.Fr.fI, c, a);
      if (cr.in) face_emf(m, 1, g.Fr, g.piph, g.dphi, g.Epre, cr.base_in + cr.in_h1, &out[0], &out[4]);
      if (cr.out) face_emf(m, 1, g.Fr, g.piph, g.dphi, g.Epre, cr.base_out + cr.out_h1, &out[1], &out[5]);
   }
   if (m.Nz > 1) {
      CellRanges cr = load_ranges(m, 2, g.rngz, g.Fz.fI, c, a);
      if (cr.in) face_emf(m, 2, g.Fz, g.piph, g.dphi, g.Epre, cr.base_in + cr.in_h1, &out[2], &out[6]);
      if (cr.out) face_emf(m, 2, g.Fz, g.piph, g.dphi, g.Epre, cr.base_out + cr.out_h1, &out[3], &out[7]);
   }
}

// Owner edges with the mesh-shear correction E -= h w B
// (subtract_advective_B_fluxes, bfields.c:646-690): Eo = {E0, E1, E4, E5}.
// Also E_phi of each owned face, Eown = {r-in, r-out, z-in, z-out}, and the
// full edge state of cells outside the phi-flux range (Ecarry).
struct KEdgesFinal {
   EdgeArgs g;
   int ct;
   const double *wiph, *emf_r, *emf_z;
   double *Eo, *Eown, *Ecarry;
   DHD void operator()(long c) const
   {
      const Mesh &m = g.m;
      size_t nt = (size_t)m.ntot;
      int a = m.cjk[c];
      int j = a % m.Nr, k = a / m.Nr;
      double v[8];
      edges_post_face(g, c, v);
      double own[4] = {0.0, 0.0, 0.0, 0.0};
      if (m.Nr > 1) {
         size_t nfs = (size_t)g.Fr.n;
         CellRanges cr = load_ranges(m, 1, g.rngr, g.Fr.fI, c, a);
         if (cr.in) own[0] = emf_r[2 * nfs + cr.base_in + cr.in_h1];
         if (cr.out) own[1] = emf_r[2 * nfs + cr.base_out + cr.out_h1];
      }
      if (m.Nz > 1) {
         size_t nfs = (size_t)g.Fz.n;
         CellRanges cr = load_ranges(m, 2, g.rngz, g.Fz.fI, c, a);
         if (cr.in) own[2] = emf_z[2 * nfs + cr.base_in + cr.in_h1];
         if (cr.out) own[3] = emf_z[2 * nfs + cr.base_out + cr.out_h1];
      }
      double xc0 = centroid_r(m.r_jph[j], m.r_jph[j - 1]);
      double hL = m.r_jph[j - 1], hR = m.r_jph[j], hD = xc0, hU = xc0;
      double w = wiph[c];
      double E0 = v[0], E1 = v[1], E4 = v[2], E5 = v[3];
      E0 -= hL * w * v[4];
      E1 -= hR * w * v[5];
      E4 += hD * w * v[6];
      E5 += hU * w * v[7];
      Eo[0 * nt + c] = E0;
      Eo[1 * nt + c] = E1;
      Eo[2 * nt + c] = E4;
      Eo[3 * nt + c] = E5;
#pragma unroll
      for (int s = 0; s < 4; ++s) Eown[s * nt + c] = own[s];
      if (!phi_in_range(m, ct, j, k)) {
         // The phi- corner slots were copied from the lower phi neighbour
         // before the subtraction (bfields.c:631-644).
         int c0 = m.off[a], np = m.Np[a];
         long cm = c == c0 ? c0 + np - 1 : c - 1;
         double vm[8];
         edges_post_face(g, cm, vm);
         double Ec[8] = {E0, E1, vm[0], vm[1], E4, E5, vm[2], vm[3]};
         double Bc[8] = {v[4], v[5], vm[4], vm[5], v[6], v[7], vm[6], vm[7]};
#pragma unroll
         for (int s = 0; s < 8; ++s) {
            Ecarry[s * nt + c] = Ec[s];
            Ecarry[(8 + s) * nt + c] = Bc[s];
         }
      }
   }
};

// The edge EMF of the face a cell owns, by the face's orientation.
DHD double owner_edge(const double *Eo, size_t nt, int dim, int LR, int L, int R)
{
   if (dim == 1) return LR == 0 ? Eo[1 * nt + L] : Eo[0 * nt + R];
   return LR == 0 ? Eo[3 * nt + L] : Eo[2 * nt + R];
}

// update_B_fluxes (bfields.c:313-382) per cell: the RK blend of Phi
// (misc.c:238-240), then E dl dt of the edges that bound the cell's own
// faces, in the reference's order.
struct KUpdateB {
   Mesh m;
   double RK;
   const double *dtp;
   Faces Fr, Fz;
   const int *rngr, *rngz;
   const double *Eo;
   // the substep's fluxes in, the RK blend's other fluxes, the updated
   // fluxes out (Phi_out may be Phi)
   const double *Phi;
   const double *rkPhi;
   double *Phi_out;
   DHD void operator()(long c) const
   {
      size_t nt = (size_t)m.ntot;
      const double dt = *dtp;
      int a = m.cjk[c];
      int j = a % m.Nr, k = a / m.Nr;
      double ph[5];
#pragma unroll
      for (int s = 0; s < 5; ++s) {
         double v = Phi[s * nt + c];
         if (RK != 0.0) v = (1. - RK) * v + RK * rkPhi[s * nt + c];
         ph[s] = v;
      }
      for (int dim = 1; dim <= 2; ++dim) {
         if (!dim_active(m, dim)) continue;
         const Faces &F = dim == 1 ? Fr : Fz;
         CellRanges cr = load_ranges(m, dim, dim == 1 ? rngr : rngz, F.fI, c, a);
         // r faces: Phi[own] -= E dl dt, the next face's Phi += E dl dt,
         // the owner's phi face +/-; z faces the same with the sign flipped.
         double sg = dim == 1 ? 1.0 : -1.0;
         for (int side = 0; side < 2; ++side) {
            if (side == 0 ? !cr.in : !cr.out) continue;
            int pj = j, pk = k;
            if (side == 0) {
               if (dim == 1) pj = j - 1;
               else pk = k - 1;
            }
            int JK = pair_index(m, dim, pj, pk);
            if (track_skipped(F.npairs, JK)) continue;
            int base = side == 0 ? cr.base_in : cr.base_out;
            int nfp = F.fI[JK + 1] - base;
            int own = side == 0 ? cr.in_h1 : cr.out_h1;
            int prev = own == 0 ? nfp - 1 : own - 1;
            int fo = base + own, fp = base + prev;
            double dl = dim == 1 ? m.z_kph[pk] - m.z_kph[pk - 1] : m.r_jph[pj] - m.r_jph[pj - 1];
            double Eown = owner_edge(Eo, nt, dim, F.LR[fo], F.L[fo], F.R[fo]);
            double Eprev = owner_edge(Eo, nt, dim, F.LR[fp], F.L[fp], F.R[fp]);
            double x_own = sg * Eown * dl * dt, x_prev = sg * Eprev * dl * dt;
            int s = dim == 1 ? (side == 0 ? 1 : 2) : (side == 0 ? 3 : 4);
            if (own == 0) {
               ph[s] -= x_own;
               ph[s] += x_prev;
            } else {
               ph[s] += x_prev;
               ph[s] -= x_own;
            }
            if (side == 0) ph[0] -= x_own;
            else ph[0] += x_own;
         }
      }
#pragma unroll
      for (int s = 0; s < 5; ++s) Phi_out[s * nt + c] = ph[s];
   }
};

// make_edge_adjust (bfields.c:888-1127) per face.  The reference walks the
// four rings around each phi-directed edge line; each step is the overlap
// of the current faces on either side of the line, so a face finds its own
// steps by a merge path between its list and the facing list, and adds
// +-E dl dt for each, in the reference's line and step order.
//   r face of pair (j,k): U role on line (j,k-1) [-], D role on line (j,k) [+]
//   z face of pair (j,k): R role on line (j-1,k) [+], L role on line (j,k) [-]
// The facing list of a role, and whether this face's list is the first of
// the line's merge.
struct EdgeRole {
   bool ok, self_first;
   int oJK, lj;
   double sgn;
};

DHD EdgeRole edge_role(const Mesh &m, int dim, int j, int k, int role)
{
   EdgeRole r;
   r.ok = true;
   if (dim == 1) {
      r.lj = j;
      if (role == 0) {
         r.ok = k > 0;
         r.oJK = pair_index(m, 1, j, k - 1); r.self_first = false; r.sgn = -1.0;
      } else {
         r.ok = k < m.Nz - 1;
         r.oJK = pair_index(m, 1, j, k + 1); r.self_first = true; r.sgn = +1.0;
      }
   } else {
      if (role == 0) {
         r.ok = j > 0;
         r.oJK = pair_index(m, 2, j - 1, k); r.self_first = false; r.sgn = +1.0; r.lj = j - 1;
      } else {
         r.ok = j < m.Nr - 1;
         r.oJK = pair_index(m, 2, j + 1, k); r.self_first = true; r.sgn = -1.0; r.lj = j;
      }
   }
   return r;
}

// The merge step at which each face is consumed on each of its two edge
// lines (one merge-path search per role); a face's steps on a line run from
// the step after its predecessor's to its own (cf. load_ranges).
template <int DIM>
struct KLineSteps {
   Mesh m;
   Faces F;
   int *estep;  // [2][F.n]
   DHD void operator()(long t) const
   {
      int f = (int)t;
      int JK = F.pf[f];
      Pair pr = pair_of(m, DIM, JK);
      int base = F.fI[JK];
      FaceRing self = make_face_ring(F.up, base, F.fI[JK + 1] - base);
      int x = f - base;
      for (int role = 0; role < 2; ++role) {
         EdgeRole er = edge_role(m, DIM, pr.j, pr.k, role);
         int step = 0;
         if (er.ok) {
            int obase = F.fI[er.oJK];
            FaceRing other = make_face_ring(F.up, obase, F.fI[er.oJK + 1] - obase);
            step = er.self_first ? a_step(self, x, other, m.Pmax) : b_step(other, self, x, m.Pmax);
         }
         estep[(size_t)role * F.n + f] = step;
      }
   }
};

// The segments of every edge line, each formed once: a line's steps are
// shared out among the faces of its first list (each step has exactly one
// current face in each list), and each of those faces stores E dl dt of its
// steps at seg[lbase[JK] + step], JK the first list's pair.
struct KEdgeSegments {
   Mesh m;
   const double *dtp;
   Faces Fr, Fz;
   const int *estep_r, *estep_z, *lbase_r, *lbase_z;
   const double *piph, *dphi, *Eown;
   double *seg_r, *seg_z;
   DHD void operator()(long t) const
   {
      size_t nt = (size_t)m.ntot;
      const double dt = *dtp;
      int dim = t < Fr.n ? 1 : 2;
      const Faces &F = dim == 1 ? Fr : Fz;
      const int *estep = dim == 1 ? estep_r : estep_z;
      int f = (int)(dim == 1 ? t : t - Fr.n);
      int JK = F.pf[f];
      Pair pr = pair_of(m, dim, JK);
      EdgeRole er = edge_role(m, dim, pr.j, pr.k, 1);  // the role where F's list is first
      if (!er.ok) return;
      double *seg = (dim == 1 ? seg_r : seg_z) + (dim == 1 ? lbase_r : lbase_z)[JK];
      int base = F.fI[JK];
      int N = F.fI[JK + 1] - base;
      int x = f - base;
      int fp = base + (x == 0 ? N - 1 : x - 1);
      int obase = F.fI[er.oJK];
      int oN = F.fI[er.oJK + 1] - obase;
      int h0, h1 = estep[(size_t)F.n + f], t0, t1;
      range_from_steps(h1, estep[(size_t)F.n + fp], x == 0, N + oN, &h0, &t0, &t1);
      double rline = m.r_jph[er.lj];
      // Corner positions: an r face (c, cR) is the line's first face and
      // the facing one (cU, cUR); a z face (c, cU), the facing one (cR, cUR).
      int oL = dim == 1 ? 2 : 1;
      EdgeCell c0 = edge_cell(piph, dphi, Eown, nt, F.L[f], 0);
      EdgeCell c1 = edge_cell(piph, dphi, Eown, nt, F.R[f], dim == 1 ? 1 : 2);
      for (int part = 0; part < 2; ++part) {
         int lo = part == 0 ? h0 : t0, hi = part == 0 ? h1 : t1;
         // while this face is current the merge has consumed x of its list
         // (all N in the tail): the facing face at step e is element e - x
         // (e - N) of the other list
         int consumed = part == 0 ? x : N;
         for (int e = lo; e <= hi; ++e) {
            int y = e - consumed;
            int g = obase + (y == oN ? 0 : y);
            EdgeCell c2 = edge_cell(piph, dphi, Eown, nt, F.L[g], oL);
            EdgeCell c3 = edge_cell(piph, dphi, Eown, nt, F.R[g], 3);
            double E, dl;
            if (dim == 1) edge_segment(c0, c1, c2, c3, rline, m.Pmax, &E, &dl);
            else edge_segment(c0, c2, c1, c3, rline, m.Pmax, &E, &dl);
            seg[e] = E * dl * dt;
         }
      }
   }
};

// make_edge_adjust (bfields.c:888-1127) per face: Phi of the face's owner
// slot takes +-E dl dt of each segment of its two edge lines while it is
// current, in the reference's line and step order, from the stored
// segments.
struct KEdgeAdjust {
   Mesh m;
   Faces Fr, Fz;
   const int *estep_r, *estep_z, *lbase_r, *lbase_z;
   const double *seg_r, *seg_z;
   double *Phi;
   DHD void operator()(long t) const
   {
      size_t nt = (size_t)m.ntot;
      int dim = t < Fr.n ? 1 : 2;
      const Faces &F = dim == 1 ? Fr : Fz;
      const int *estep = dim == 1 ? estep_r : estep_z;
      const int *lbase = dim == 1 ? lbase_r : lbase_z;
      const double *segd = dim == 1 ? seg_r : seg_z;
      int f = (int)(dim == 1 ? t : t - Fr.n);
      int JK = F.pf[f];
      Pair pr = pair_of(m, dim, JK);
      int base = F.fI[JK];
      int N = F.fI[JK + 1] - base;
      int x = f - base;
      int fp = base + (x == 0 ? N - 1 : x - 1);
      size_t slot = phi_slot(nt, dim, F.LR[f], F.L[f], F.R[f]);
      double v = Phi[slot];
      for (int role = 0; role < 2; ++role) {
         EdgeRole er = edge_role(m, dim, pr.j, pr.k, role);
         if (!er.ok) continue;
         int oN = F.fI[er.oJK + 1] - F.fI[er.oJK];
         const double *seg = segd + lbase[er.self_first ? JK : er.oJK];
         int h0, h1 = estep[(size_t)role * F.n + f], t0, t1;
         range_from_steps(h1, estep[(size_t)role * F.n + fp], x == 0, N + oN, &h0, &t0, &t1);
         for (int part = 0; part < 2; ++part) {
            int lo = part == 0 ? h0 : t0, hi = part == 0 ? h1 : t1;
            for (int e = lo; e <= hi; ++e) {
               if (er.sgn > 0) v += seg[e];
               else v -= seg[e];
            }
         }
      }
      Phi[slot] = v;
   }
};

// check_flipped (bfields.c:693-712): after the mesh moved, a face whose
// upper end now lies before its predecessor's has vanished.  Positions are
// the moved ones before clean_pi, as the reference has them.
template <int DIM>
struct KFlipFlags {
   Mesh m;
   static constexpr int dim = DIM;
   Faces F;
   const double *piph, *wiph;
   const double *dtp;
   signed char *flag;
   DHD double moved_up(int f) const
   {
      int o = F.LR[f] == 0 ? F.L[f] : F.R[f];
      return piph[o] + wiph[o] * *dtp;
   }
   DHD void operator()(long f) const
   {
      int JK = F.pf[f];
      int base = F.fI[JK], N = F.fI[JK + 1] - base;
      int n = (int)f - base;
      int fm = base + (n == 0 ? N - 1 : n - 1);
      flag[f] = get_sdp(moved_up((int)f), moved_up(fm), m.Pmax) < 0. ? 1 : 0;
   }
};

// flip_fluxes for flagged face n of a track (bfields.c:742-814): its flux
// joins its predecessor's, the predecessor's is handed on negated, and both
// faces change owner.
DHD void flip_one(int dim, size_t nt, Faces F, int base, int N, int n, double *Phi,
                  double *rkPhi)
{
   int f = base + n;
   int fm = base + (n == 0 ? N - 1 : n - 1);
   int fp = base + (n == N - 1 ? 0 : n + 1);
   size_t P = phi_slot(nt, dim, F.LR[f], F.L[f], F.R[f]);
   size_t Pm = phi_slot(nt, dim, F.LR[fm], F.L[fm], F.R[fm]);
   size_t Pp = phi_slot(nt, dim, F.LR[fp], F.L[fp], F.R[fp]);
   double ph = Phi[P], rk = rkPhi[P];
   Phi[P] = Phi[Pm] + ph;
   rkPhi[P] = rkPhi[Pm] + rk;
   Phi[Pm] = -ph;
   rkPhi[Pm] = -rk;
   Phi[Pp] = Phi[Pp] + ph;
   rkPhi[Pp] = rkPhi[Pp] + rk;
   F.LR[f] = (signed char)!F.LR[f];
   F.LR[fm] = (signed char)!F.LR[fm];
}

// One face track, sequentially (the host form of the wavefront kernel).
struct KFlipsTrack {
   Mesh m;
   int dim;
   Faces F;
   const signed char *flag;
   double *Phi, *rkPhi;
   DHD void operator()(long JK) const
   {
      if (track_skipped(F.npairs, (int)JK)) return;
      int base = F.fI[JK], N = F.fI[JK + 1] - base;
      for (int n = 0; n < N; ++n)
         if (flag[base + n]) flip_one(dim, (size_t)m.ntot, F, base, N, n, Phi, rkPhi);
   }
};

// B_faces_to_cells (bfields.c:88-309) and, for real cells, cons2prim
// (misc.c:290-325): cell B from the fluxes through the cell's faces
// (ownership after the flips) -- the primitive field (type 0) or the
// conserved one over the moved geometry (type 1).
template <int NQ, class H>
struct KBfacesC2p {
   Mesh m;
   dh_params P;
   Faces Fr, Fz;
   const int *rngr, *rngz;
   const double *piph_new, *dphi_new, *Phi;
   double *cons, *prim;
   int type;    // 0: primitive B (set_B_fields), 1: conserved B
   int do_c2p;  // cons2prim of real cells afterwards
   unsigned char *info;  // GRMHD: each cons2prim's record (C2P_INFO_*), or null
   DHD void operator()(long c) const
   {
      size_t nt = (size_t)m.ntot;
      int a = m.cjk[c];
      int c0 = m.off[a], np = m.Np[a];
      long cm = c == c0 ? c0 + np - 1 : c - 1;
      int j = a % m.Nr, k = a / m.Nr;
      double rp = m.r_jph[j], rm = m.r_jph[j - 1];
      double zp = m.z_kph[k], zm = m.z_kph[k - 1];
      double x0 = centroid_r(rp, rm);
      double dV = cell_dV(rp, rm, piph_new[c], piph_new[cm], zp, zm, m.Pmax);
      double dA0 = (rp - rm) * (zp - zm);
      double Br = 0.0, Bz = 0.0;
      for (int dim = 1; dim <= 2; ++dim) {
         if (!dim_active(m, dim)) continue;
         const Faces &F = dim == 1 ? Fr : Fz;
         CellRanges cr = load_ranges(m, dim, dim == 1 ? rngr : rngz, F.fI, c, a);
         double sum = 0.0, area = 0.0;
         auto add = [&](int base, int lo, int hi) {
            for (int f = base + lo; f <= base + hi; ++f) {
               area += F.dA[f];
               sum += Phi[phi_slot(nt, dim, F.LR[f], F.L[f], F.R[f])];
            }
         };
         if (cr.in) {
            add(cr.base_in, cr.in_h0, cr.in_h1);
            add(cr.base_in, cr.in_t0, cr.in_t1);
         }
         if (cr.out) {
            add(cr.base_out, cr.out_h0, cr.out_h1);
            add(cr.base_out, cr.out_t0, cr.out_t1);
         }
         if (type == 0) {
            if (dim == 1) Br = sum / area;
            else Bz = sum / area;
         } else if (dim == 1) {
            Br = sum * (dV * (1.0 / x0) / (area * 1.0));
         } else {
            Bz = sum * (dV * 1.0 / (area * 1.0));
         }
      }
      if (type == 0) {
         prim[BRR * nt + c] = Br;
         prim[BPP * nt + c] = .5 * (Phi[0 * nt + c] + Phi[0 * nt + cm]) / dA0;
         prim[BZZ * nt + c] = Bz;
         return;
      }
      double Bp = .5 * (Phi[0 * nt + c] + Phi[0 * nt + cm]) * dV / (dA0 * x0);
      cons[BRR * nt + c] = Br;
      cons[BPP * nt + c] = Bp;
      cons[BZZ * nt + c] = Bz;
      if (!do_c2p || !is_real(m, j, k)) return;
      double cs[NQ], out[NQ];
      load_q<NQ>(cons, nt, c, cs);
      double dVc = cell_dV(rp, rm, piph_new[c], piph_new[c] - dphi_new[c], zp, zm, m.Pmax);
      load_q<NQ>(prim, nt, c, out);  // the previous primitives: an iterative solve's guess
      if constexpr (IsGR<H>::value)
         H::cons2prim_guess(P, cs, out, x0, centroid_z(zp, zm), dVc, nullptr,
                            info ? info + c : nullptr);
      else
         H::cons2prim(P, cs, out, x0, centroid_z(zp, zm), dVc);
#pragma unroll
      for (int q = 0; q < NQ; ++q) prim[q * nt + c] = out[q];
   }
};

// cons2prim of the real cells alone (the last part of KBfacesC2p), for
// grmhd2 with a learned initial guess: guess[2c], guess[2c+1] = (v^2, eta)
// of cell c from k_c2p_guess, which runs between the two.
template <int NQ>
struct KC2p {
   Mesh m;
   dh_params P;
   const double *piph_new, *dphi_new, *cons;
   double *prim;
   const float *guess;
   unsigned char *info;
   DHD void operator()(long c) const
   {
      size_t nt = (size_t)m.ntot;
      int a = m.cjk[c];
      int j = a % m.Nr, k = a / m.Nr;
      if (!is_real(m, j, k)) return;
      double rp = m.r_jph[j], rm = m.r_jph[j - 1];
      double zp = m.z_kph[k], zm = m.z_kph[k - 1];
      double cs[NQ], out[NQ];
      load_q<NQ>(cons, nt, c, cs);
      load_q<NQ>(prim, nt, c, out);
      double dVc = cell_dV(rp, rm, piph_new[c], piph_new[c] - dphi_new[c], zp, zm, m.Pmax);
      GRMHD<NQ>::cons2prim_guess(P, cs, out, centroid_r(rp, rm), centroid_z(zp, zm), dVc,
                                 guess + 2 * c, info ? info + c : nullptr);
#pragma unroll
      for (int q = 0; q < NQ; ++q) prim[q * nt + c] = out[q];
   }
};

// The invariants of cell c's conserved state at its centroid, and its
// volume, as cons2prim forms them.
template <int NQ>
DHD C2pInv cell_c2p_invariants(const Mesh &m, const dh_params &P, const double *cons,
                               const double *piph, const double *dphi, long c, int j, int k)
{
   size_t nt = (size_t)m.ntot;
   double rp = m.r_jph[j], rm = m.r_jph[j - 1];
   double zp = m.z_kph[k], zm = m.z_kph[k - 1];
   double dV = cell_dV(rp, rm, piph[c], piph[c] - dphi[c], zp, zm, m.Pmax);
   double cs[NQ];
   load_q<NQ>(cons, nt, c, cs);
#pragma unroll
   for (int q = 0; q < NQ; ++q) cs[q] /= dV;
   double r = centroid_r(rp, rm);
   C2pInv v;
   c2p_invariants(gr_metric(P, r, centroid_z(zp, zm)), r, cs, v);
   return v;
}

// Training data of the learned guess: for every real cell the model's
// features (feat[8c..8c+7]) and targets (targ[2c], targ[2c+1]: v^2 and
// log(eta / e)) of its current state, whose primitives cons2prim produced
// from its conserved state; other cells get NaN.
template <int NQ>
struct KC2pHarvest {
   Mesh m;
   dh_params P;
   const double *cons, *prim, *piph, *dphi;
   float *feat, *targ;
   DHD void operator()(long c) const
   {
      size_t nt = (size_t)m.ntot;
      int a = m.cjk[c];
      int j = a % m.Nr, k = a / m.Nr;
      if (!is_real(m, j, k)) {
         for (int i = 0; i < C2PNET_IN; ++i) feat[C2PNET_IN * c + i] = NAN;
         targ[2 * c] = targ[2 * c + 1] = NAN;
         return;
      }
      C2pInv v = cell_c2p_invariants<NQ>(m, P, cons, piph, dphi, c, j, k);
      c2p_features(v, P.gamma_law, feat + C2PNET_IN * c);
      double gl = P.gamma_law;
      double l[3] = {prim[URR * nt + c], prim[UPP * nt + c], prim[UZZ * nt + c]};
      double u2 = 0.0;
      for (int i = 0; i < 3; ++i)
         for (int jj = 0; jj < 3; ++jj) u2 += v.g.igam[3 * i + jj] * l[i] * l[jj];
      double w = sqrt(1.0 + u2);
      double eta = w * (1.0 + gl / (gl - 1.0) * prim[PPP * nt + c] / prim[RHO * nt + c]);
      targ[2 * c] = (float)(u2 / (1.0 + u2));
      targ[2 * c + 1] = (float)log(eta / v.e);
   }
};

// ---------------------------------------------------------------------------
// Zero-gradient ghost annuli (set_cells_copy_distant, boundary_functions.c:
// 90-159, and the diode of outflow3d, boundary_functions.c:269-441): each
// ghost cell is the phi-overlap-weighted mean of the cells of the next
// annulus inward.  The reference walks both rings once per ghost annulus;
// here one thread per ghost annulus runs that walk over positions only and
// records where each ghost cell's overlap sum starts, then one thread per
// ghost cell forms the sum, layer by layer (a layer copies from the one
// inside it).
//   side 0: r inner, 1: r outer, 2: z bottom, 3: z top.
// ---------------------------------------------------------------------------

// Destination and source annulus of ghost layer g of one line of a side.
DHD void bc_pair(const Mesh &m, int side, int line, int g, int *dst, int *src)
{
   if (side == 0) { int j = m.NgRa - 1 - g; *dst = j + m.Nr * line; *src = *dst + 1; }
   else if (side == 1) { int j = m.Nr - m.NgRb + g; *dst = j + m.Nr * line; *src = *dst - 1; }
   else if (side == 2) { int k = m.NgZa - 1 - g; *dst = line + m.Nr * k; *src = *dst + m.Nr; }
   else { int k = m.Nz - m.NgZb + g; *dst = line + m.Nr * k; *src = *dst - m.Nr; }
}

DHD int bc_lines(const Mesh &m, int side) { return side < 2 ? m.Nz : m.Nr; }
DHD int bc_layers(const Mesh &m, int side)
{
   return side == 0 ? m.NgRa : side == 1 ? m.NgRb : side == 2 ? m.NgZa : m.NgZb;
}

// Overlap in phi of destination cell (phim, phip] with source cell i1,
// shifted by whole periods next to it as the reference does; returns the
// shifted upper edge of the source cell.
DHD double bc_overlap(double phim, double phip, const double *piph, const double *dphi, int s,
                      double pm, double *d)
{
   double phip1 = piph[s];
   double phim1 = piph[s] - dphi[s];
   if (std::isfinite(phip1) && std::isfinite(phim1) && std::isfinite(phip) && std::isfinite(phim)) {
      if (phip1 < phim - 64.0 * pm || phim1 > phip + 64.0 * pm) { *d = 0.0; return phip; }
      while (phip1 < phim) { phip1 += pm; phim1 += pm; }
      while (phim1 > phip) { phip1 -= pm; phim1 -= pm; }
   }
   double phi1 = phip < phip1 ? phip : phip1;
   double phi2 = phim > phim1 ? phim : phim1;
   *d = phi1 - phi2;
   return phip1;
}

// Whether source element i1 holds the lower edge phi0 of a ghost ring's
// first cell, the test that starts the ring walk.
DHD bool bc_holds_start(const double *piph, const double *dphi, int s0, int i1, double phi0,
                        double pm)
{
   double dphip = get_sdp(piph[s0 + i1], phi0, pm);
   double dphim = dphip - dphi[s0 + i1];
   return dphip > 0 && dphim <= 0.0;
}

// The walk over one ghost cell c from source element i1: the element of the
// last source cell it overlaps, which is where the next ghost cell's sum
// begins.
DHD int bc_walk_cell(const double *piph, const double *dphi, long c, int s0, int Np1, int i1,
                     double pm)
{
   double phip = piph[c];
   double phim = piph[c] - dphi[c];
   double phip1, d;
   int guard = 0;
   do {
      phip1 = bc_overlap(phim, phip, piph, dphi, s0 + i1, pm, &d);
      i1 = i1 < Np1 - 1 ? i1 + 1 : 0;
   } while (phip1 < phip && ++guard <= Np1);
   return i1 > 0 ? i1 - 1 : Np1 - 1;
}

// The ring walk over positions: start[c] = the source element (index within
// its annulus) where ghost cell c's overlap sum begins.  Ring t is line
// t % lines of layer t / lines.
struct KBcScan {
   Mesh m;
   int side;
   const double *piph, *dphi;
   int *start;
   DHD void operator()(long t) const
   {
      int lines = bc_lines(m, side);
      int line = (int)(t % lines), g = (int)(t / lines);
      int dst, src;
      bc_pair(m, side, line, g, &dst, &src);
      double pm = m.Pmax;
      int c0 = m.off[dst], s0 = m.off[src];
      int Np = m.Np[dst], Np1 = m.Np[src];
      int i1;
      double phi0 = piph[c0] - dphi[c0];
      for (i1 = 0; i1 < Np1; i1++)
         if (bc_holds_start(piph, dphi, s0, i1, phi0, pm)) break;
      if (i1 >= Np1) i1 = 0;
      for (int i = 0; i < Np; ++i) {
         start[c0 + i] = i1;
         i1 = bc_walk_cell(piph, dphi, c0 + i, s0, Np1, i1, pm);
      }
   }
};

// A guess of the walk's state on entering a ghost cell whose lower edge is
// x: the source cell that holds x, found by bisection over the source ring
// read from its lowest upper edge (element rot) on.  Used by the parallel
// form of the walk, which checks every guess against one step of the walk.
DHD int bc_guess(const double *piph, int s0, int Np1, int rot, double x, double pm)
{
   if (!(x >= 0.0 && x < pm)) x -= floor(x / pm) * pm;
   int lo = 0, hi = Np1;  // first t in [0, Np1) with U(t) >= x, else Np1
   while (lo < hi) {
      int mid = (lo + hi) >> 1;
      int k = rot + mid;
      if (k >= Np1) k -= Np1;
      if (piph[s0 + k] >= x) hi = mid;
      else lo = mid + 1;
   }
   int t = lo == Np1 ? 0 : lo;
   int k = rot + t;
   return k >= Np1 ? k - Np1 : k;
}

// One ghost cell of layer g: the overlap-weighted sum over the source cells
// from its start, in the walk's order, then the diode.
template <int NQ>
struct KBcFill {
   Mesh m;
   int side, g, diode;
   const int *start;
   double *prim;
   const double *piph, *dphi;
   int maxnp;  // cells per line slot: the largest Np of the layer's annuli
   DHD void operator()(long t) const
   {
      size_t nt = (size_t)m.ntot;
      int line = (int)(t / maxnp), i = (int)(t % maxnp);
      int dst, src;
      bc_pair(m, side, line, g, &dst, &src);
      int Np = m.Np[dst];
      if (i >= Np) return;
      int c = m.off[dst] + i;
      int s0 = m.off[src], Np1 = m.Np[src];
      double pm = m.Pmax;
      double phip = piph[c];
      double phim = piph[c] - dphi[c];
      double acc[NQ];
#pragma unroll
      for (int q = 0; q < NQ; ++q) acc[q] = 0.0;
      int i1 = start[c];
      double phip1, d;
      int guard = 0;
      do {
         phip1 = bc_overlap(phim, phip, piph, dphi, s0 + i1, pm, &d);
#pragma unroll
         for (int q = 0; q < NQ; ++q) acc[q] += prim[q * nt + s0 + i1] * d;
         i1 = i1 < Np1 - 1 ? i1 + 1 : 0;
      } while (phip1 < phip && ++guard <= Np1);
#pragma unroll
      for (int q = 0; q < NQ; ++q) prim[q * nt + c] = acc[q] / dphi[c];
      if (diode) {
         int qd = side < 2 ? URR : UZZ;
         double sgn = (side == 0 || side == 2) ? 1.0 : -1.0;
         double *v = &prim[qd * nt + c];
         if (sgn * *v > 0) *v = 0.0;
      }
   }
};

// ---------------------------------------------------------------------------
// Initial cons from prim on real cells (domain.c:175-237); hydro also maps
// back to prim there.
// ---------------------------------------------------------------------------

template <int NQ, class H>
struct KInitCells {
   Mesh m;
   dh_params P;
   double *prim, *cons;
   const double *piph, *dphi;
   DHD void operator()(long c) const
   {
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
      double z = centroid_z(zp, zm);
      H::prim2cons(P, pr, cs, r, z, dV);
      if constexpr (!IsMHD<H>::value) H::cons2prim(P, cs, pr, r, z, dV);
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
         cons[q * nt + c] = cs[q];
         prim[q * nt + c] = pr[q];
      }
   }
};

// ---------------------------------------------------------------------------
// Diagnostics (analysis.c:52-106 with the shipped Diagnostics modules)
// ---------------------------------------------------------------------------

// planetaryForce (planet.c:99-149): force per unit mass of planet pl
// (DH_PLANET_STRIDE layout: r, phi, M, omega, vr, eps, type) at (r, phi, z),
// its in-plane part along the cell's (r, phi) directions for mode 0 and along
// the planet's for mode 1.
DHD void planetary_force(const dh_params &P, DiagPlanet pl, double r, double phi, double z,
                         double *fr, double *fp, double *fz, int mode = 0)
{
   double rp = pl.v[0], pp = pl.v[1];
   if (P.grav2d == 1) z = 0.0;
   else if (P.grav2d == 2) { rp = r; pp = phi; z = 1.0; }
   double sinp, cosp, sinpp, cospp;
   sincos(phi, &sinp, &cosp);
   sincos(pp, &sinpp, &cospp);
   double dx = r * cosp - rp * cospp;
   double dy = r * sinp - rp * sinpp;
   double script_r = sqrt(dx * dx + dy * dy + z * z);
   double script_r_perp = sqrt(dx * dx + dy * dy);
   double f1 = -fgrav(pl.v[2], script_r, pl.v[5], (int)pl.v[6], P.phi_order);
   double cosa = dx / script_r_perp;
   double sina = dy / script_r_perp;
   double cosap = cosa * cosp + sina * sinp;
   double sinap = sina * cosp - cosa * sinp;
   if (mode == 1) {
      cosap = cosa * cospp + sina * sinpp;
      sinap = sina * cospp - cosa * sinpp;
   }
   if (script_r_perp <= 0.0) { cosap = 0.0; sinap = 0.0; }
   double sint = script_r_perp / script_r;
   double cost = z / script_r;
   *fr = cosap * f1 * sint;
   *fp = sinap * f1 * sint;
   *fz = f1 * cost;
}

// Primitive fields a module reads: rho..u_z, and B for the MHD module.
DHD int diag_fields(int kind) { return kind == DH_DIAG_MHDMRI ? BZZ + 1 : UZZ + 1; }

DHD int diag_count(int kind)
{
   return kind == DH_DIAG_EULER ? 8 : kind == DH_DIAG_DISK2D ? 5 : kind == DH_DIAG_MHDMRI ? 18 : 0;
}

// get_diagnostics of the module `kind` at x = {r, phi, z} (Diagnostics/
// diag_euler.c:17-45, diag_disk2d.c:17-39, diag_mhdmri.c:17-59); p1 is
// planet 1 (p1.on when Npl > 1).
DHD void diag_cell(int kind, const dh_params &P, DiagPlanet p1, double r, double phi, double z,
                   const double *prim, double *Q)
{
   double rho = prim[RHO], Pp = prim[PPP], vr = prim[URR], omega = prim[UPP], vz = prim[UZZ];
   if (kind == DH_DIAG_MHDMRI) {
      double gam = P.gamma_law;
      double Br = prim[BRR], Bp = prim[BPP], Bz = prim[BZZ];
      double cs2 = gam * Pp / rho;
      double B2 = Br * Br + Bp * Bp + Bz * Bz;
      double sr = rho * vr;
      double sp = r * r * rho * omega;
      double sz = rho * vz;
      double e = 0.5 * rho * (vr * vr + r * r * omega * omega + vz * vz) + Pp / (gam - 1.0) +
                 0.5 * B2;
      Q[0] = rho; Q[1] = Pp; Q[2] = vr; Q[3] = omega; Q[4] = vz;
      Q[5] = Br; Q[6] = Bp; Q[7] = Bz;
      Q[8] = rho; Q[9] = e; Q[10] = sr; Q[11] = sp; Q[12] = sz;
      Q[13] = rho * vr * omega;
      Q[14] = Br * Bp;
      Q[15] = B2;
      Q[16] = sqrt(cs2);
      Q[17] = sqrt(B2 / rho);
      return;
   }
   double Fr, Fp = 0.0, Fz;
   if (p1.on) planetary_force(P, p1, r, phi, kind == DH_DIAG_DISK2D ? 0.0 : z, &Fr, &Fp, &Fz);
   Q[0] = rho;
   Q[1] = 2. * M_PI * r * rho * vr;
   if (kind == DH_DIAG_DISK2D) {
      Q[2] = 2. * M_PI * r * rho * (r * Fp);
      Q[3] = 2. * M_PI * r * rho * (r * r * omega * vr);
      Q[4] = omega;
      return;
   }
   Q[2] = 2. * M_PI * r * rho * vz;
   Q[3] = 2. * M_PI * r * rho * (r * Fp);
   Q[4] = 2. * M_PI * r * rho * (r * r * omega * vr);
   Q[5] = omega;
   Q[6] = 2 * M_PI * r * rho * r * r * omega;
   Q[7] = Pp;
}

// Position and volume of cell c as add_diagnostics forms them: the midpoint
// of the cell's extents (not the centroid) and get_dV.
DHD void diag_geometry(const Mesh &m, const double *piph, const double *dphi, long c, int j,
                       int k, double *x, double *dV)
{
   double phip = piph[c];
   double phim = phip - dphi[c];
   double rp = m.r_jph[j], rm = m.r_jph[j - 1], zp = m.z_kph[k], zm = m.z_kph[k - 1];
   x[0] = .5 * (rp + rm);
   x[1] = .5 * (phip + phim);
   x[2] = .5 * (zp + zm);
   *dV = cell_dV(rp, rm, phip, phim, zp, zm, m.Pmax);
}

// One annulus, in the reference's order: Qrz[a*nq + q] += (sum_i Q_q dV) dt / sum_i dV.
DHD void diag_annulus(const Mesh &m, const dh_params &P, int kind, DiagPlanet p1,
                      const double *prim, const double *piph, const double *dphi, double dt,
                      int a, double *Qrz)
{
   int nq = diag_count(kind);
   size_t nt = (size_t)m.ntot;
   int j = a % m.Nr, k = a / m.Nr;
   double sum[DH_DIAG_MAXQ], vol = 0.0;
   for (int q = 0; q < nq; ++q) sum[q] = 0.0;
   for (int i = 0; i < m.Np[a]; ++i) {
      long c = m.off[a] + i;
      double x[3], dV, p[BZZ + 1], Q[DH_DIAG_MAXQ];
      diag_geometry(m, piph, dphi, c, j, k, x, &dV);
      for (int q = 0; q < diag_fields(kind); ++q) p[q] = prim[q * nt + c];
      diag_cell(kind, P, p1, x[0], x[1], x[2], p, Q);
      for (int q = 0; q < nq; ++q) sum[q] += Q[q] * dV;
      vol += dV;
   }
   for (int q = 0; q < nq; ++q) Qrz[(size_t)a * nq + q] += sum[q] * dt / vol;
}

// report (report.c:7-202): the global sums of one annulus's real cells.
// The host adds the annuli in the reference's (j outer, k inner) order and
// forms the report.dat columns from the totals.
enum {
   RP_PSIR, RP_PSII, RP_L1ISEN, RP_L1RHO, RP_L1B, RP_BR2, RP_BRBP, RP_B2, RP_PDV, RP_MDOT,
   RP_VOL, RP_MASS, RP_SR, RP_S0, RP_TORQUE, RP_POWER, RP_FR,
   RP_RHOMIN,  // a minimum, not a sum
   RP_N
};

// One cell's terms (report.c:76-145; rho0 = 1).  p0, p1: planets 0 and 1,
// used when p1.on (Npl > 1).
DHD void report_cell(const Mesh &m, const dh_params &P, const DiagPlanet &p0,
                     const DiagPlanet &p1, int mhd, const double *prim, const double *piph,
                     const double *dphi, long c, int j, int k, double *o)
{
   size_t nt = (size_t)m.ntot;
   double phi = piph[c] - .5 * dphi[c];
   double Pp = prim[PPP * nt + c], rho = prim[RHO * nt + c];
   double phip = piph[c], phim = phip - dphi[c];
   double rp = m.r_jph[j], rm = m.r_jph[j - 1];
   double dV = cell_dV(rp, rm, phip, phim, m.z_kph[k], m.z_kph[k - 1], m.Pmax);
   double r = centroid_r(rp, rm);
   o[RP_PSIR] = rho * dV * cos(phi);
   o[RP_PSII] = rho * dV * sin(phi);
   o[RP_L1ISEN] = fabs(Pp / pow(rho, P.gamma_law) - 1.) * dV;
   o[RP_L1RHO] = fabs(rho - 1.) * dV;
   o[RP_L1B] = o[RP_BR2] = o[RP_BRBP] = o[RP_B2] = 0.0;
   if (mhd) {
      double Br = prim[BRR * nt + c], Bp = prim[BPP * nt + c], Bz = prim[BZZ * nt + c];
      o[RP_L1B] = fabs(Br) * dV;
      o[RP_BR2] = .5 * Br * Br * dV;
      o[RP_BRBP] = Br * Bp * dV;
      o[RP_B2] = .5 * (Br * Br + Bp * Bp + Bz * Bz) * dV;
   }
   o[RP_PDV] = Pp * dV;
   o[RP_MDOT] = 2. * M_PI * r * rho * dV * prim[URR * nt + c];
   o[RP_VOL] = dV;
   o[RP_MASS] = rho * dV;
   double rho4 = pow(rho, 4.);
   o[RP_SR] = rho4 * r * dV;
   o[RP_S0] = rho4 * dV;
   o[RP_TORQUE] = o[RP_POWER] = o[RP_FR] = 0.0;
   if (p1.on) {
      double fr, fp, fz, fp2;
      planetary_force(P, p0, r, phi, 0.0, &fr, &fp2, &fz, 1);
      planetary_force(P, p1, r, phi, 0.0, &fr, &fp, &fz, 1);
      double rpl = p1.v[0], om = p1.v[3], vr = p1.v[4];
      o[RP_TORQUE] = -((rho - 1.0) * rpl * fp * dV);
      o[RP_POWER] = -((rho - 1.0) * (rpl * om * fp + vr * fr) * dV);
      o[RP_FR] = -((rho - 1.0) * fr * dV);
   }
   o[RP_RHOMIN] = rho;
}

// One annulus a, its cells in order: R[a * RP_N + q].
DHD void report_annulus(const Mesh &m, const dh_params &P, const DiagPlanet &p0,
                        const DiagPlanet &p1, int mhd, const double *prim, const double *piph,
                        const double *dphi, int a, double *R)
{
   int j = a % m.Nr, k = a / m.Nr;
   double sum[RP_N], o[RP_N];
   for (int q = 0; q < RP_RHOMIN; ++q) sum[q] = 0.0;
   sum[RP_RHOMIN] = HUGE_VAL;
   for (int i = 0; i < m.Np[a]; ++i) {
      report_cell(m, P, p0, p1, mhd, prim, piph, dphi, m.off[a] + i, j, k, o);
      for (int q = 0; q < RP_RHOMIN; ++q) sum[q] += o[q];
      if (sum[RP_RHOMIN] > o[RP_RHOMIN]) sum[RP_RHOMIN] = o[RP_RHOMIN];
   }
   for (int q = 0; q < RP_N; ++q) R[(size_t)a * RP_N + q] = sum[q];
}

// avg_diagnostics (analysis.c:28-49): every accumulated value over t_avg.
struct KDiagAvg {
   double *Qrz;
   double t_avg;
   DHD void operator()(long i) const { Qrz[i] /= t_avg; }
};

// CFL time of one real cell (misc.c:39-84); HUGE_VAL for ghosts.
template <int NQ, class H>
DHD double cell_mindt(const Mesh &m, const dh_params &P, const double *prim, const double *piph,
                      const double *dphi, const double *wiph, long c)
{
   int a = m.cjk[c];
   int j = a % m.Nr, k = a / m.Nr;
   if (!is_real(m, j, k)) return HUGE_VAL;
   size_t nt = (size_t)m.ntot;
   int c0 = m.off[a], np = m.Np[a];
   long cm = (c == c0) ? c0 + np - 1 : c - 1;
   double w = .5 * (wiph[cm] + wiph[c]);
   double phip = piph[c];
   double phim = phip - dphi[c];
   double pr[NQ];
   load_q<NQ>(prim, nt, c, pr);
   return H::mindt(P, pr, w, m.r_jph[j], m.r_jph[j - 1], phip, phim, m.z_kph[k], m.z_kph[k - 1],
                   m.Pmax);
}

}  // namespace dh
