// dh_mesh.h -- device view of the ragged annulus mesh and the face topology
// between two neighbouring annuli.
//
// The reference builds the r- and z-faces between annuli A and B by walking
// both rings of cells from a start cell (faces_alt.c:36-179): at every step it
// emits the overlap of the current A and B cells as a face and advances
// whichever cell ends first in phi (ties advance A).  That walk is a merge of
// the two sequences of cell upper edges, so on the GPU every face is found on
// its own by a merge-path binary search, and every cell finds the contiguous
// range of faces it owns by counting how many cells of the other ring end
// before it.  No face list has to be built serially, and the results are the
// reference's faces in the reference's order.
#pragma once
#include "dh_geom.h"

namespace dh {

struct Mesh {
   int Nr, Nz, NgRa, NgRb, NgZa, NgZb;
   int nann;          // Nr*Nz annuli
   int ntot;          // cells incl. ghosts
   const int *Np;     // [nann]
   const int *off;    // [nann+1]
   const int *cjk;    // [ntot] annulus of each cell
   const double *r_jph;  // index -1 valid
   const double *z_kph;  // index -1 valid
   double Pmax;
   const int *I0;        // [nann] ring start cells of the current substep
};

// One ring of cells read from its start cell I0: element x (0 <= x <= N) is
// cell (I0 + x) mod N; element N is the start cell again (a full turn later).
struct Ring {
   const double *piph;  // global piph array
   int off, N, I0;
   double p0;           // piph of the start cell
};

DHD Ring make_ring(const double *piph, const Mesh &m, const int *I0, int a)
{
   Ring R;
   R.piph = piph;
   R.off = m.off[a];
   R.N = m.Np[a];
   R.I0 = I0[a];
   R.p0 = piph[R.off + R.I0];
   return R;
}

DHD int ring_cell(const Ring &R, int x)
{
   int i = R.I0 + x;
   if (i >= R.N) i -= R.N;
   return R.off + i;
}

// A face list of one annulus pair read as a ring: element x (0 <= x <= N) is
// face f0 + (x mod N); its key is the face's upper phi edge, the upper edge
// of whichever of its two cells ends first (faces_alt.c:147-151), stored per
// face when the faces are built.
struct FaceRing {
   const double *up;
   int f0, N;
   double p0;
};

DHD FaceRing make_face_ring(const double *up, int f0, int N)
{
   FaceRing R;
   R.up = up;
   R.f0 = f0;
   R.N = N;
   R.p0 = up[f0];
   return R;
}

DHD double ring_key(const Ring &R, int x) { return R.piph[ring_cell(R, x)]; }
DHD double ring_key(const FaceRing &R, int x) { return R.up[R.f0 + (x == R.N ? 0 : x)]; }

// True when A's element x ends (in phi) no later than B's element y, i.e.
// the reference walk advances A first (faces_alt.c:135-145: LR = 0 when the
// wrapped difference of upper edges is not positive; bfields.c:817-866 uses
// the same rule).  Elements are compared through the reference's wrapped
// difference whenever they are within half a period of each other in the
// unwrapped order (always the case for pairs the walk visits); far-apart
// elements, which the binary searches also probe, are ordered by their
// unwrapped positions.
template <class RA, class RB>
DHD bool a_first(const RA &A, int x, const RB &B, int y, double Pmax)
{
   if (y == B.N) return x < A.N;
   if (x == A.N) return false;
   double pa = ring_key(A, x);
   double pb = ring_key(B, y);
   double ua = (x > 0 && pa < A.p0) ? pa + Pmax : pa;
   double ub = (y > 0 && pb < B.p0) ? pb + Pmax : pb;
   double D = ua - ub;
   if (fabs(D) < 0.5 * Pmax) {
      double w = wrap_half(pa - pb, Pmax);
      return !(w > 0.);
   }
   return D < 0.;
}

// Number of A elements consumed before step d of the walk (merge path).
template <class RA, class RB>
DHD int merge_x(const RA &A, const RB &B, int d, double Pmax)
{
   int lo = d - B.N > 0 ? d - B.N : 0;
   int hi = d < A.N ? d : A.N;
   while (lo < hi) {
      int mid = (lo + hi + 1) >> 1;
      if (a_first(A, mid - 1, B, d - mid, Pmax)) lo = mid;
      else hi = mid - 1;
   }
   return lo;
}

// Step at which A's element x (x < N_A) is consumed: x plus the number of B
// elements that end before it.
template <class RA, class RB>
DHD int a_step(const RA &A, int x, const RB &B, double Pmax)
{
   int lo = 0, hi = B.N;
   while (lo < hi) {
      int mid = (lo + hi + 1) >> 1;
      if (!a_first(A, x, B, mid - 1, Pmax)) lo = mid;
      else hi = mid - 1;
   }
   return x + lo;
}

// Step at which B's element y (y < N_B) is consumed.
template <class RA, class RB>
DHD int b_step(const RA &A, const RB &B, int y, double Pmax)
{
   int lo = 0, hi = A.N;
   while (lo < hi) {
      int mid = (lo + hi + 1) >> 1;
      if (a_first(A, mid - 1, B, y, Pmax)) lo = mid;
      else hi = mid - 1;
   }
   return y + lo;
}

// Faces (indices within the pair's list, inclusive ranges) owned by one cell:
// a head range and, for the ring's start cell only, a tail range at the end
// of the list (the walk ends on the start cell again).  Empty: lo > hi.
struct FaceRange {
   int h0, h1, t0, t1;
};

// Ranges of A's element x (the cell is the L side of these faces).
template <class RA, class RB>
DHD FaceRange ranges_as_a(const RA &A, int x, const RB &B, double Pmax)
{
   FaceRange r;
   int nf = A.N + B.N;
   r.h0 = x == 0 ? 0 : a_step(A, x - 1, B, Pmax) + 1;
   r.h1 = a_step(A, x, B, Pmax);
   r.t0 = 1;
   r.t1 = 0;
   if (x == 0) {
      r.t0 = (A.N > 1 ? a_step(A, A.N - 1, B, Pmax) : r.h1) + 1;
      r.t1 = nf - 1;
   }
   return r;
}

// Ranges of B's element y (the cell is the R side of these faces).
template <class RA, class RB>
DHD FaceRange ranges_as_b(const RA &A, const RB &B, int y, double Pmax)
{
   FaceRange r;
   int nf = A.N + B.N;
   r.h0 = y == 0 ? 0 : b_step(A, B, y - 1, Pmax) + 1;
   r.h1 = b_step(A, B, y, Pmax);
   r.t0 = 1;
   r.t1 = 0;
   if (y == 0) {
      r.t0 = (B.N > 1 ? b_step(A, B, B.N - 1, Pmax) : r.h1) + 1;
      r.t1 = nf - 1;
   }
   return r;
}

// Annulus pair of a face list: dim 1 pairs (j, j+1), dim 2 pairs (k, k+1)
// (faces_alt.c:80-102 ordering: k outer, j inner).
struct Pair {
   int a, b;      // annulus indices (L side, R side)
   int j, k;      // annulus a's (j, k)
};

DHD Pair pair_of(const Mesh &m, int dim, int JK)
{
   Pair p;
   if (dim == 1) {
      int Nrm = m.Nr - 1;
      p.j = JK % Nrm;
      p.k = JK / Nrm;
      p.a = p.j + m.Nr * p.k;
      p.b = p.a + 1;
   } else {
      p.j = JK % m.Nr;
      p.k = JK / m.Nr;
      p.a = p.j + m.Nr * p.k;
      p.b = p.a + m.Nr;
   }
   return p;
}

// Largest pair index P with fI[P] <= f.
DHD int pair_of_face(const int *fI, int npairs, int f)
{
   int lo = 0, hi = npairs - 1;
   while (lo < hi) {
      int mid = (lo + hi + 1) >> 1;
      if (fI[mid] <= f) lo = mid;
      else hi = mid - 1;
   }
   return lo;
}

}  // namespace dh
