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

// Position of element x (x < N) in its ring's unwrapped order, as a_first
// reads it.
template <class R>
DHD double ring_pos(const R &A, int x, double Pmax)
{
   double p = ring_key(A, x);
   return (x > 0 && p < A.p0) ? p + Pmax : p;
}

// Estimate of how many elements of ring B lie below unwrapped position u:
// the elements are spread nearly evenly over one period from B's start.
template <class R>
DHD int count_below(const R &B, double u, double Pmax)
{
   double g = (u - B.p0) * (double)B.N / Pmax;
   if (!(g > 0.0)) return 0;
   if (g >= (double)B.N) return B.N;
   return (int)ceil(g);
}

// Largest m in [0, N] with P(m), for P monotone (true up to the answer,
// false after) with P(0) taken as true: a galloping search out from the
// estimate g, then bisection inside the bracket it finds.  For monotone P
// the result is the plain bisection's, in a few probes when g is close.
template <class Pred>
DHD int last_true(int N, int g, const Pred &P)
{
   int lo, hi;
   if (g < 0) g = 0;
   if (g > N) g = N;
   if (g == 0 || P(g)) {
      lo = g;
      hi = N;
      for (int step = 1;; step <<= 1) {
         int probe = lo + step;
         if (probe > N) break;
         if (P(probe)) {
            lo = probe;
         } else {
            hi = probe - 1;
            break;
         }
      }
   } else {
      hi = g - 1;
      lo = 0;
      for (int step = 1;; step <<= 1) {
         int probe = g - step;
         if (probe <= 0) break;
         if (P(probe)) {
            lo = probe;
            break;
         }
         hi = probe - 1;
      }
   }
   while (lo < hi) {
      int mid = (lo + hi + 1) >> 1;
      if (P(mid)) lo = mid;
      else hi = mid - 1;
   }
   return lo;
}

// Step at which A's element x (x < N_A) is consumed: x plus the number of B
// elements that end before it.
template <class RA, class RB>
DHD int a_step(const RA &A, int x, const RB &B, double Pmax)
{
   int g = count_below(B, ring_pos(A, x, Pmax), Pmax);
   return x + last_true(B.N, g, [&](int mid) { return !a_first(A, x, B, mid - 1, Pmax); });
}

// Step at which B's element y (y < N_B) is consumed.
template <class RA, class RB>
DHD int b_step(const RA &A, const RB &B, int y, double Pmax)
{
   int g = count_below(A, ring_pos(B, y, Pmax), Pmax);
   return y + last_true(A.N, g, [&](int mid) { return a_first(A, mid - 1, B, y, Pmax); });
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

}  // namespace dh
