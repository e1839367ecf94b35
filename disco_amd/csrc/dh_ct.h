// dh_ct.h -- constrained transport (CT_MODE 2) on the moving mesh, as
// owner-computes kernels (device side of the reference bfields.c).
//
// State: Phi[5][c] per cell, the magnetic flux through the cell's upper phi
// face (0), its owned inner/outer r faces (1, 2) and inner/outer z faces
// (3, 4).  A face between annuli is owned by the cell whose upper phi edge
// bounds it (faces_alt.c:147-151: LRtype 0 -> L owns it in slot 2 or 4,
// LRtype 1 -> R owns it in slot 1 or 3); every cell owns exactly one face on
// each side.  EMFs live on cell edges: E[0..3] on the z-directed edges at
// (r-, phi+), (r+, phi+), (r-, phi-), (r+, phi-), E[4..7] on the r-directed
// edges at (z-, phi+), (z+, phi+), (z-, phi-), (z+, phi-) (riemann.c:66-211),
// and E_phi on the phi-directed edges where four annuli meet.
//
// The reference updates these with sequential scatter loops over faces and
// edges.  Here every output slot is computed by its owner (dh_kernels.h):
//   KEdgesCollect  edge EMFs after the Riemann solves     (riemann.c:66-211)
//   KEdgesFinal    owner edges: the phi-neighbour averages and each owned
//                  face's interpolated edge EMF, formed where they are read
//                  (bfields.c:400-531), then the advective subtraction
//                  (bfields.c:533-690)
//   KUpdateB       Phi += E dl dt of the face edges       (bfields.c:313-382)
//   KLineSteps, KEdgeAdjust  E_phi edges: each face's merge steps on its two
//                  edge lines, then its segments; a merge of two face lists
//                  instead of the reference's 4-way ring walk (bfields.c:888-1127)
//   KFlipFlags, k_flips  face re-association after the mesh moved (bfields.c:693-814)
//   KBfacesC2p     Phi -> cell B, then cons2prim           (bfields.c:88-309)
#pragma once
#include "dh_mesh.h"

namespace dh {

// The faces of a cell in its inner and outer pair of one direction, as
// inclusive head/tail ranges of the pair's face list: a head range and, for
// the ring's start cell only, a tail range at the end of the list (the walk
// ends on the start cell again).  Empty ranges have lo > hi.
struct CellRanges {
   int in_h0, in_h1, in_t0, in_t1, out_h0, out_h1, out_t0, out_t1;
   int base_in, base_out;   // first face of the inner / outer pair
   bool in, out;            // pair exists
};

// rng[0*nt + c] is the merge step at which cell c is consumed in its inner
// pair (where its ring is the second), rng[1*nt + c] the same in its outer
// pair (its ring first).  A cell's faces run from the step after its phi
// predecessor's to its own; the ring's start cell also owns the faces after
// the last cell's step (the walk ends on it again).
DHD void range_from_steps(int h1, int hp, bool start, int nf, int *h0, int *t0, int *t1)
{
   if (start) {
      *h0 = 0;
      *t0 = hp + 1;
      *t1 = nf - 1;
   } else {
      *h0 = hp + 1;
      *t0 = 1;
      *t1 = 0;
   }
   (void)h1;
}

DHD CellRanges load_ranges(const Mesh &m, int dim, const int *rng, const int *fI, long c, int a)
{
   size_t nt = (size_t)m.ntot;
   // The ring predecessor's steps: those of the storage predecessor c - 1,
   // read together with the cell's own (before the annulus offset is
   // known), except for the annulus's first cell, whose predecessor is its
   // last.
   long cs = c > 0 ? c - 1 : c;
   int hp_in = rng[cs], hp_out = rng[nt + cs];
   int h_in = rng[c], h_out = rng[nt + c];
   int c0 = m.off[a], np = m.Np[a];
   int i = (int)(c - c0);
   if (i == 0) {
      hp_in = rng[c0 + np - 1];
      hp_out = rng[nt + c0 + np - 1];
   }
   bool start = i == m.I0[a];
   int j = a % m.Nr, k = a / m.Nr;
   CellRanges r;
   int Jin, Jout;
   if (dim == 1) {
      r.in = j > 0;
      r.out = j < m.Nr - 1;
      Jin = (j - 1) + (m.Nr - 1) * k;
      Jout = j + (m.Nr - 1) * k;
   } else {
      r.in = k > 0;
      r.out = k < m.Nz - 1;
      Jin = j + m.Nr * (k - 1);
      Jout = j + m.Nr * k;
   }
   r.base_in = r.base_out = 0;
   r.in_h0 = r.in_t0 = r.out_h0 = r.out_t0 = 1;
   r.in_h1 = r.in_t1 = r.out_h1 = r.out_t1 = 0;
   if (r.in) {
      r.base_in = fI[Jin];
      r.in_h1 = h_in;
      range_from_steps(r.in_h1, hp_in, start, fI[Jin + 1] - r.base_in, &r.in_h0, &r.in_t0,
                       &r.in_t1);
   }
   if (r.out) {
      r.base_out = fI[Jout];
      r.out_h1 = h_out;
      range_from_steps(r.out_h1, hp_out, start, fI[Jout + 1] - r.base_out, &r.out_h0,
                       &r.out_t0, &r.out_t1);
   }
   return r;
}

// A cell's faces in both directions as one sequence in the reference's
// order: r inner head, r inner tail, r outer head, r outer tail, then the
// same four for z (ranges of absolute face indices, empty ones lo > hi).
// Position v of the sequence maps to its range and face through compares
// unrolled over the eight ranges, so that the ranges stay in registers.
// A kernel walks the sequence one face ahead of the face it adds: the next
// face's loads are in flight while the current one is summed, and the sums
// keep their order.
struct FaceSeq {
   int lo[8], end[8];  // range k holds positions end[k-1] .. end[k]-1
   int n;
   DHD int face(int v, int *k) const
   {
      int f = 0, b = 0;
      *k = 0;
#pragma unroll
      for (int r = 0; r < 8; ++r) {
         if (v >= b && v < end[r]) {
            f = lo[r] + (v - b);
            *k = r;
         }
         b = end[r];
      }
      return f;
   }
};

// Ranges r0 .. r0+3 of a sequence from one direction's CellRanges (the
// inner pair's when `in`, the outer pair's when `out`, else empty); b
// counts the positions so far.
DHD void seq_put(FaceSeq &s, int r0, const CellRanges &cr, bool in, bool out, int &b)
{
   int lo[4] = {1, 1, 1, 1}, hi[4] = {0, 0, 0, 0};
   if (in) {
      lo[0] = cr.base_in + cr.in_h0;
      hi[0] = cr.base_in + cr.in_h1;
      lo[1] = cr.base_in + cr.in_t0;
      hi[1] = cr.base_in + cr.in_t1;
   }
   if (out) {
      lo[2] = cr.base_out + cr.out_h0;
      hi[2] = cr.base_out + cr.out_h1;
      lo[3] = cr.base_out + cr.out_t0;
      hi[3] = cr.base_out + cr.out_t1;
   }
#pragma unroll
   for (int r = 0; r < 4; ++r) {
      int cnt = hi[r] - lo[r] + 1;
      b += cnt > 0 ? cnt : 0;
      s.lo[r0 + r] = lo[r];
      s.end[r0 + r] = b;
   }
}

// One direction's faces (ranges 0-3; 4-7 empty).
DHD FaceSeq pair_seq(const CellRanges &cr)
{
   FaceSeq s;
   CellRanges none;
   none.in = none.out = false;
   int b = 0;
   seq_put(s, 0, cr, cr.in, cr.out, b);
   seq_put(s, 4, none, false, false, b);
   s.n = b;
   return s;
}

// Walks a face sequence one face ahead: ld(f, k, d) loads face f (of range
// k) into d and use(d, k) consumes it, in the sequence's order; the next
// face's loads are issued before the current face is used, so that two
// faces' loads are in flight where a plain loop would wait for each.
template <class D, class L, class U>
DHD void walk_ahead(const FaceSeq &s, L ld, U use)
{
   if (s.n == 0) return;
   int k;
   D cur, nxt;
   ld(s.face(0, &k), k, cur);
   for (int v = 1;; ++v) {
      int kc = k;
      bool more = v < s.n;
      if (more) ld(s.face(v, &k), k, nxt);
      use(cur, kc);
      if (!more) break;
      cur = nxt;
   }
}

// The track loops of update_B_fluxes and check_flipped/flip_fluxes
// (bfields.c:313-382, 693-814) visit track jk-1 at iteration jk, so the last
// face track of each direction (a ghost corner) is never visited.
DHD bool track_skipped(int npairs, int JK) { return JK == npairs - 1; }

DHD int pair_index(const Mesh &m, int dim, int j, int k)
{
   return dim == 1 ? j + (m.Nr - 1) * k : j + m.Nr * k;
}

// Slot of the face's owner: the reference's get_phi_pointer (bfields.c:716-740).
DHD size_t phi_slot(size_t nt, int dim, int LR, int L, int R)
{
   if (dim == 1) return LR == 0 ? 2 * nt + L : 1 * nt + R;
   return LR == 0 ? 4 * nt + L : 3 * nt + R;
}

// phi_switch (bfields.c:817-829)
DHD int phi_switch(double dphi, double Pmax, int mode)
{
   dphi = wrap_half(dphi, Pmax);
   if (mode == 1) dphi = -dphi;
   return dphi > 0. ? 1 : 0;
}

// get_which4 (bfields.c:831-886)
DHD int get_which4(double phi, double phiR, double phiU, double phiUR, int *LR_alt,
                   int *UD_alt, int mode, double Pmax)
{
   int LR_D = phi_switch(phi - phiR, Pmax, mode);
   int LR_U = phi_switch(phiU - phiUR, Pmax, mode);
   double phi1 = LR_D ? phiR : phi;
   double phi2 = LR_U ? phiUR : phiU;
   int UD = phi_switch(phi1 - phi2, Pmax, mode);
   int w4;
   if (UD == 0) w4 = LR_D == 0 ? 0 : 1;
   else w4 = LR_U == 0 ? 2 : 3;
   if (mode == 0) {
      *LR_alt = (w4 == 0 || w4 == 1) ? LR_U : LR_D;
      if (w4 == 0 || w4 == 2) *UD_alt = phi_switch(phiR - phiUR, Pmax, mode);
      else *UD_alt = phi_switch(phi - phiU, Pmax, mode);
   }
   return w4;
}

// The four cells around a phi-directed edge line at (r_jph[j], z_kph[k]):
// position 0 = c (j,k), 1 = cR (j+1,k), 2 = cU (j,k+1), 3 = cUR (j+1,k+1).
// What the edge needs of each: its upper and lower phi edges and the E_phi
// of the two faces it owns that touch the line, taken from Eown[s][c]
// (s = 0 r-in, 1 r-out, 2 z-in, 3 z-out, the reference's E_phi[0..3]):
// ea is the r face (r-out for positions 0 and 2, r-in for 1 and 3), eb the
// z face (z-out for 0 and 1, z-in for 2 and 3).
struct EdgeCell {
   double up, lo, ea, eb;
};

DHD EdgeCell edge_cell(const double *piph, const double *dphi, const double *Eown, size_t nt,
                       long cell, int pos)
{
   EdgeCell q;
   q.up = piph[cell];
   q.lo = piph[cell] - dphi[cell];
   q.ea = Eown[(size_t)((pos & 1) ? 0 : 1) * nt + cell];
   q.eb = Eown[(size_t)((pos & 2) ? 2 : 3) * nt + cell];
   return q;
}


// One segment of the phi-directed edge line of radius r where the four
// cells meet: its EMF and length, exactly as make_edge_adjust forms them
// (bfields.c:948-1116).
DHD void edge_segment(EdgeCell c, EdgeCell cR, EdgeCell cU,
                      EdgeCell cUR, double r, double Pmax, double *E, double *dl)
{
   int LR_alt, UD_alt;
   int w4 = get_which4(c.up, cR.up, cU.up, cUR.up, &LR_alt, &UD_alt, 0, Pmax);
   int w4b = get_which4(c.lo, cR.lo, cU.lo, cUR.lo, nullptr, nullptr, 1, Pmax);
   double e;
   if (w4 == 0) {
      e = .25 * (c.eb + c.ea);
      e += LR_alt == 0 ? .25 * cU.ea : .25 * cUR.ea;
      e += UD_alt == 0 ? .25 * cR.eb : .25 * cUR.eb;
   } else if (w4 == 1) {
      e = .25 * (cR.eb + cR.ea);
      e += LR_alt == 0 ? .25 * cU.ea : .25 * cUR.ea;
      e += UD_alt == 0 ? .25 * c.eb : .25 * cU.eb;
   } else if (w4 == 2) {
      e = .25 * (cU.eb + cU.ea);
      e += LR_alt == 0 ? .25 * c.ea : .25 * cR.ea;
      e += UD_alt == 0 ? .25 * cR.eb : .25 * cUR.eb;
   } else {
      e = .25 * (cUR.eb + cUR.ea);
      e += LR_alt == 0 ? .25 * c.ea : .25 * cR.ea;
      e += UD_alt == 0 ? .25 * c.eb : .25 * cU.eb;
   }
   double xp1 = w4 == 0 ? c.up : w4 == 1 ? cR.up : w4 == 2 ? cU.up : cUR.up;
   double xm1 = w4b == 0 ? c.lo : w4b == 1 ? cR.lo : w4b == 2 ? cU.lo : cUR.lo;
   *E = e;
   *dl = .5 * (r + r) * get_dp(xp1, xm1, Pmax);
}

}  // namespace dh
