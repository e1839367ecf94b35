// dh_grmhd.h -- general-relativistic MHD on the moving cylindrical mesh
// (device side of the reference Hydro/grmhd2.c) in the Schwarzschild metric
// in Kerr-Schild cylindrical coordinates (t, r, phi, z), either at (r, z)
// (Hydro/Metric/schw_ks_rpz.c) or evaluated on the equator, z = 0
// (Hydro/Metric/schw_ks_eq_rpz.c), with the Eulerian observer frame
// (Hydro/Frame/normal.c).
//
// Primitive layout {rho, P, u_r, u_phi, u_z, B^r, B^phi, B^z, passives...}
// with covariant four-velocity components u_i and the densitised field;
// conserved {D, tau, S_r, S_phi, S_z, B^r dV / r, B^phi dV, B^z dV, ...}
// (grmhd2.c:88-180).  The metric is axisymmetric, so a module function needs
// the position's r and z (the reference's x[0] and x[2]).
//
// Riemann solvers: HLL and HLLC (riemann.c:216-292 with grmhd2.c's vel,
// flux, prim2cons and getUstar) and the relativistic HLLD of hlld_grmhd.c
// (dh_grhlld.h); cons2prim: the 2-D Newton scheme of Noble et al.
// (grmhd2.c:1096-1314).  The expressions keep the reference's
// association so that the host build reproduces the reference bit for bit.
#pragma once
#include "dh_geom.h"
#include "dh_grhlld.h"
#define KKS_FN DHD
#include "kerr_ks.h"

namespace dh {

// Metric (schw_ks_rpz.c:20-189, schw_ks_eq_rpz.c:19-189) and frame
// (normal.c:9-42) at (r, z), with M = Metric_Par2.  gam / igam are the spatial
// metric and its inverse, be the shift, al the lapse, U the frame
// four-velocity.  The equatorial module is the same metric at z = 0: its
// lapse takes R = r where this one takes sqrt(r^2 + z^2), which at z = 0 is r
// exactly (the square root of a rounded square is the number itself).
struct GRMetric {
   double al, be[3], gam[9], igam[9], jac, sqrtgam, U[4];
};

// The z at which the metric module evaluates a position (the equatorial
// module evaluates every position at z = 0).
DHD double gr_z(const dh_params &P, double z) { return P.metric == DH_METRIC_SCHW_KS_EQ ? 0.0 : z; }

DHD double gr_lapse(double M, double r, double z)
{
   double R = sqrt(r * r + z * z);
   return 1.0 / sqrt(1.0 + 2 * M / R);
}

DHD void gr_shift(double M, double r, double z, double *b)
{
   double R = sqrt(r * r + z * z);
   double a2 = 1.0 / (1.0 + 2 * M / R);
   b[0] = 2 * M * r / (R * R) * a2;
   b[1] = 0.0;
   b[2] = 2 * M * z / (R * R) * a2;
}

DHD void gr_gam(double M, double r, double z, double *g)
{
   double R = sqrt(r * r + z * z);
   double H = 2 * M / R;
   double st = r / R, ct = z / R;
   g[0] = 1.0 + H * st * st;
   g[1] = 0.0;
   g[2] = H * st * ct;
   g[3] = 0.0;
   g[4] = r * r;
   g[5] = 0.0;
   g[6] = H * st * ct;
   g[7] = 0.0;
   g[8] = 1.0 + H * ct * ct;
}

DHD void gr_igam(double M, double r, double z, double *g)
{
   double R = sqrt(r * r + z * z);
   double H = 2 * M / R;
   double st = r / R, ct = z / R;
   double a2 = 1.0 / (1.0 + H);
   g[0] = (1.0 + H * ct * ct) * a2;
   g[1] = 0.0;
   g[2] = -H * st * ct * a2;
   g[3] = 0.0;
   g[4] = 1.0 / (r * r);
   g[5] = 0.0;
   g[6] = -H * st * ct * a2;
   g[7] = 0.0;
   g[8] = (1.0 + H * st * st) * a2;
}

// d g_{mu nu} / d x^i for i = 1 (r) and 3 (z); zero otherwise.
DHD void gr_der_g(double M, double r, double z, int i, double *dg)
{
   for (int mu = 0; mu < 16; ++mu) dg[mu] = 0.0;
   if (i != 1 && i != 3) return;
   double R = sqrt(r * r + z * z);
   double st = r / R, ct = z / R;
   double MoR2 = M / (R * R), MoR3 = M / (R * R * R), MoR4 = M / (R * R * R * R);
   if (i == 1) {
      dg[0] = -2 * MoR2 * st;
      dg[1] = 2 * MoR2 * (1.0 - 2 * st * st);
      dg[3] = -4 * MoR3 * z * st;
      dg[5] = 2 * MoR3 * r * (2.0 - 3.0 * st * st);
      dg[7] = 2 * MoR3 * z * (1.0 - 3.0 * st * st);
      dg[10] = 2 * r;
      dg[15] = -6 * MoR4 * z * z * st;
   } else {
      dg[0] = -2 * MoR2 * ct;
      dg[1] = -4 * MoR3 * r * ct;
      dg[3] = 2 * MoR2 * (1.0 - 2 * ct * ct);
      dg[5] = -6 * MoR4 * r * r * ct;
      dg[7] = 2 * MoR3 * r * (1.0 - 3.0 * ct * ct);
      dg[15] = 2 * MoR3 * z * (2.0 - 3.0 * ct * ct);
   }
   dg[4] = dg[1];
   dg[12] = dg[3];
   dg[13] = dg[7];
}

// Killing directions of the metric: t and phi (x^0, x^2).
DHD bool gr_killing(int mu) { return !(mu == 1 || mu == 3); }

DHD void gr_der_lapse(double M, double r, double z, double *da)
{
   double R = sqrt(r * r + z * z);
   double st = r / R, ct = z / R;
   double MoR2 = M / (R * R);
   double a = 1.0 / sqrt(1.0 + 2 * M / R);
   da[0] = 0.0;
   da[1] = a * a * a * MoR2 * st;
   da[2] = 0.0;
   da[3] = a * a * a * MoR2 * ct;
}

DHD void gr_der_shift(double M, double r, double z, double *db)
{
   double R = sqrt(r * r + z * z);
   double st = r / R, ct = z / R;
   double denom = 1.0 / (R * R + 2 * M * R);
   double d_denom = -denom * denom * 2 * (R + M);
   for (int i = 0; i < 3; ++i) db[i] = 0;
   db[3] = 2 * M * denom + 2 * M * r * d_denom * st;
   db[4] = 0.0;
   db[5] = 2 * M * z * d_denom * st;
   for (int i = 0; i < 3; ++i) db[6 + i] = 0;
   db[9] = 2 * M * r * d_denom * ct;
   db[10] = 0.0;
   db[11] = 2 * M * denom + 2 * M * z * d_denom * ct;
}

// The Eulerian observer (normal.c:9-42).
DHD void gr_frame_U(double M, double r, double z, double *U)
{
   double b[3];
   double ia = 1.0 / gr_lapse(M, r, z);
   gr_shift(M, r, z, b);
   U[0] = ia;
   U[1] = -ia * b[0];
   U[2] = -ia * b[1];
   U[3] = -ia * b[2];
}

DHD void gr_frame_der_U(double M, double r, double z, double *dU)
{
   double da[4], b[3], db[12];
   double ia = 1.0 / gr_lapse(M, r, z);
   gr_der_lapse(M, r, z, da);
   gr_shift(M, r, z, b);
   gr_der_shift(M, r, z, db);
   for (int i = 0; i < 16; ++i) dU[i] = 0.0;
   for (int i = 0; i < 4; ++i) {
      if (gr_killing(i)) continue;
      dU[4 * i] = -ia * ia * da[i];
      for (int j = 0; j < 3; ++j) dU[4 * i + j + 1] = b[j] * ia * ia * da[i] - db[3 * i + j] * ia;
   }
}

// The metric at a position (r, z); z is the reference's x[2], which the
// equatorial module ignores.  The Kerr metric (DH_METRIC_KERR_KS,
// kerr_ks.h) has the same metric_jacobian (= r) and the same normal frame.
DHD GRMetric gr_metric(const dh_params &P, double r, double z_pos)
{
   double M = P.metric_par2, z = gr_z(P, z_pos);
   GRMetric g;
   if (P.metric == DH_METRIC_KERR_KS) {
      kks_point p = kks_eval(M, P.metric_par3, r, z, 0);
      g.al = kks_lapse(&p);
      kks_shift(&p, g.be);
      kks_gam(&p, g.gam);
      kks_igam(&p, g.igam);
      double ia = 1.0 / g.al;
      g.U[0] = ia;
      for (int i = 0; i < 3; ++i) g.U[i + 1] = -ia * g.be[i];
   } else {
      g.al = gr_lapse(M, r, z);
      gr_shift(M, r, z, g.be);
      gr_gam(M, r, z, g.gam);
      gr_igam(M, r, z, g.igam);
      gr_frame_U(M, r, z, g.U);
   }
   g.jac = r / r;  // metric_jacobian (= r) over r
   g.sqrtgam = g.jac / g.al;
   return g;
}

// d g_{mu nu} / d x^i of the metric module at (r, z) (z as gr_z gives it).
DHD void gr_metric_der_g(const dh_params &P, double r, double z, int i, double *dg)
{
   if (P.metric == DH_METRIC_KERR_KS) {
      kks_point p = kks_eval(P.metric_par2, P.metric_par3, r, z, 1);
      kks_der_g(&p, i, dg);
   } else {
      gr_der_g(P.metric_par2, r, z, i, dg);
   }
}

// frame_der_U (normal.c:23-42) of the metric module at (r, z).
DHD void gr_metric_frame_der_U(const dh_params &P, double r, double z, double *dU)
{
   if (P.metric != DH_METRIC_KERR_KS) {
      gr_frame_der_U(P.metric_par2, r, z, dU);
      return;
   }
   kks_point p = kks_eval(P.metric_par2, P.metric_par3, r, z, 1);
   double da[4], b[3], db[12];
   double ia = 1.0 / kks_lapse(&p);
   kks_der_lapse(&p, da);
   kks_shift(&p, b);
   kks_der_shift(&p, db);
   for (int i = 0; i < 16; ++i) dU[i] = 0.0;
   for (int i = 0; i < 4; ++i) {
      if (gr_killing(i)) continue;
      dU[4 * i] = -ia * ia * da[i];
      for (int j = 0; j < 3; ++j) dU[4 * i + j + 1] = b[j] * ia * ia * da[i] - db[3 * i + j] * ia;
   }
}

// The fluid of a primitive state in a metric: the quantities prim2cons,
// flux and source share (grmhd2.c:88-140, 272-320).
struct GRState {
   double rho, Pp, l[3], B[3];
   double igaml[3], u2, w, u0, u[3], l0, uU;
   double uB, b0, bd[3], B2, UB, b2, Ub;
};

DHD GRState gr_state(double gamma_law, const GRMetric &g, double r, const double *prim)
{
   GRState s;
   s.rho = prim[RHO];
   s.Pp = prim[PPP];
   s.l[0] = prim[URR];
   s.l[1] = prim[UPP];
   s.l[2] = prim[UZZ];
   s.B[0] = prim[BRR] / g.sqrtgam;
   s.B[1] = prim[BPP] / (r * g.sqrtgam);
   s.B[2] = prim[BZZ] / g.sqrtgam;
   for (int i = 0; i < 3; ++i) {
      s.igaml[i] = 0.0;
      for (int j = 0; j < 3; ++j) s.igaml[i] += g.igam[3 * i + j] * s.l[j];
   }
   s.u2 = s.l[0] * s.igaml[0] + s.l[1] * s.igaml[1] + s.l[2] * s.igaml[2];
   s.w = sqrt(1.0 + s.u2);
   s.u0 = s.w / g.al;
   for (int i = 0; i < 3; ++i) s.u[i] = s.igaml[i] - g.be[i] * s.u0;
   s.l0 = -g.al * s.w + g.be[0] * s.l[0] + g.be[1] * s.l[1] + g.be[2] * s.l[2];
   s.uU = g.U[0] * s.l0 + g.U[1] * s.l[0] + g.U[2] * s.l[1] + g.U[3] * s.l[2];
   s.uB = s.l[0] * s.B[0] + s.l[1] * s.B[1] + s.l[2] * s.B[2];
   s.b0 = s.uB / g.al;
   for (int i = 0; i < 3; ++i)
      s.bd[i] = (g.gam[3 * i] * s.B[0] + g.gam[3 * i + 1] * s.B[1] + g.gam[3 * i + 2] * s.B[2] +
                 s.l[i] * s.uB) / s.w;
   s.B2 = 0.0;
   s.UB = 0.0;
   for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
         s.B2 += g.gam[3 * i + j] * s.B[i] * s.B[j];
         s.UB += g.gam[3 * i + j] * s.B[i] * (g.U[0] * g.be[j] + g.U[j + 1]);
      }
   s.b2 = (s.B2 + s.uB * s.uB) / (s.w * s.w);
   s.Ub = (s.UB + s.uU * s.uB) / s.w;
   (void)gamma_law;
   return s;
}

// The quantities of a conserved state (per unit volume) that fix the
// noble2d solution (grmhd2.c:1140-1168): s2 = S.S / D^2, e, Q = sqrt(gam)
// B^2 / D and psi = sqrt(gam) (B.S)^2 / D^3, with the metric they were
// formed in.  The adiabatic solution (v^2, eta = W h) depends on these four
// and the adiabatic index only.
struct C2pInv {
   GRMetric g;
   double D, S[3], tau, B[3], s2, Us, BS, B2, Q, psi, e;
};

DHD void c2p_invariants(const GRMetric &g, double r, const double *cons, C2pInv &v)
{
   v.g = g;
   const double *be = g.be, *gam = g.gam, *igam = g.igam, *U = g.U;
   double sqrtgam = g.sqrtgam;
   v.D = cons[DDD];
   v.S[0] = cons[SRR];
   v.S[1] = cons[LLL];
   v.S[2] = cons[SZZ];
   v.tau = cons[TAU];
   v.B[0] = r * cons[BRR] / sqrtgam;
   v.B[1] = cons[BPP] / sqrtgam;
   v.B[2] = cons[BZZ] / sqrtgam;
   const double *S = v.S, *B = v.B;
   double D = v.D;
   double s2 = 0.0, Us = 0.0, BS = 0.0, B2 = 0.0;
   for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) s2 += igam[3 * i + j] * S[i] * S[j];
   s2 /= D * D;
   for (int i = 0; i < 3; ++i) Us += S[i] * (be[i] * U[0] + U[i + 1]);
   Us /= D;
   for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) B2 += gam[3 * i + j] * B[i] * B[j];
   double Q = sqrtgam * B2 / D;
   for (int i = 0; i < 3; ++i) BS += B[i] * S[i];
   double psi = sqrtgam * BS * BS / (D * D * D);
   v.s2 = s2;
   v.Us = Us;
   v.BS = BS;
   v.B2 = B2;
   v.Q = Q;
   v.psi = psi;
   v.e = (v.tau / D + Us + 1.0) / (g.al * U[0]);
}

// The record cons2prim leaves per cell: Newton iterations (at most 60: 30
// from the learned guess and 30 after a fall-back), C2P_INFO_ITERS itself
// when the last solve did not meet its stopping rule, and whether the
// learned guess was used or fell back.
enum { C2P_INFO_ITERS = 63, C2P_INFO_LEARNED = 64, C2P_INFO_FALLBACK = 128 };
// Relative residual of the residual stop rule (DH_C2P_STOP_RESIDUAL).
constexpr double C2P_RES_TOL = 1.0e-10;

DHD bool c2p_guess_ok(const float *g)
{
   return g[0] >= 0.0f && g[0] < 1.0f && g[1] >= 1.0f && g[1] < 3.0e38f;
}

// The Newton iteration of noble2d (grmhd2.c:1170-1240) from (v2, eta), at
// most 30 iterations.  stop = DH_C2P_STOP_REFERENCE: the reference's rule,
// |d eta / eta| < 1e-12 and then Nextra = 10 more iterations;
// DH_C2P_STOP_RESIDUAL: both residuals below C2P_RES_TOL relative to
// (eta + Q)^2 and eta, no extra iterations.  Leaves the last iterate in
// (v2, eta) and the iteration count in it; returns whether the rule was met.
DHD bool noble2d(const C2pInv &v, double n, int stop, double &v2_io, double &eta_io, int &it)
{
   const double prec = 1.0e-12;
   const double max_iter = 30;
   const double Nextra = 10;
   double Q = v.Q, psi = v.psi, s2 = v.s2, e = v.e;
   double v21 = v2_io, eta1 = eta_io;
   double v2, eta, w;
   int clean = -1;
   bool met = false;
   while (1) {
      v2 = v21;
      eta = eta1;
      w = 1.0 / sqrt(1.0 - v2);
      double fa = (eta + Q) * (eta + Q) * v2 - psi * (2 * eta + Q) / (eta * eta) - s2;
      double fb = eta - n * (1 - v2) * (eta - w) + 0.5 * Q * (1 + v2) - 0.5 * psi / (eta * eta) - e;
      double dfadv2 = (eta + Q) * (eta + Q);
      double dfadet = 2 * (eta + Q) * (v2 + psi / (eta * eta * eta));
      double dfbdv2 = n * (eta - 0.5 * w) + 0.5 * Q;
      double dfbdet = 1 - n * (1 - v2) + psi / (eta * eta * eta);
      double detj = dfadv2 * dfbdet - dfbdv2 * dfadet;
      v21 = v2 - (dfbdet * fa - dfadet * fb) / detj;
      eta1 = eta - (-dfbdv2 * fa + dfadv2 * fb) / detj;
      it++;
      if (v21 > 1.0) v21 = 0.5 * (v2 + 1.0);
      if (v21 < 0.0) v21 = 0.5 * v2;
      if (eta1 < 1.0) eta1 = 0.5 * (eta + 1.0);
      if (stop == DH_C2P_STOP_RESIDUAL) {
         if (fabs(fa) <= C2P_RES_TOL * dfadv2 && fabs(fb) <= C2P_RES_TOL * eta) {
            met = true;
            break;
         }
      } else {
         double err = (eta1 - eta) / eta;
         if (fabs(err) < prec && clean < 0) clean = (int)Nextra + 1;
         if (clean >= 0) clean--;
         if (clean == 0) {
            met = true;
            break;
         }
      }
      if (it == max_iter) break;
   }
   v2_io = v21;
   eta_io = eta1;
   return met;
}

// The residual rule's measure of a Newton start (v2, eta):
// max(|f_a| / (eta + Q)^2, |f_b| / eta), the functions of noble2d.
DHD double noble2d_residual(const C2pInv &v, double n, double v2, double eta)
{
   double Q = v.Q, psi = v.psi;
   double w = 1.0 / sqrt(1.0 - v2);
   double fa = (eta + Q) * (eta + Q) * v2 - psi * (2 * eta + Q) / (eta * eta) - v.s2;
   double fb = eta - n * (1 - v2) * (eta - w) + 0.5 * Q * (1 + v2) - 0.5 * psi / (eta * eta) - v.e;
   double ra = fabs(fa) / ((eta + Q) * (eta + Q)), rb = fabs(fb) / eta;
   return ra > rb ? ra : rb;
}

template <int NQ>
struct GRMHD {
   static constexpr int NC = 8;

   // bfield_scale_factor (grmhd2.c:899-909)
   static DHD double bscale_r(double r) { return 1.0 / r; }

   // prim2cons (grmhd2.c:88-180)
   static DHD void prim2cons(const dh_params &P, const double *prim, double *cons, double r,
                             double z, double dV)
   {
      double gl = P.gamma_law;
      GRMetric g = gr_metric(P, r, z);
      GRState s = gr_state(gl, g, r, prim);
      double rhoh = s.rho + gl / (gl - 1.0) * s.Pp + s.b2;
      double rhoe = s.Pp / (gl - 1.0);
      double jac = g.jac;
      cons[DDD] = jac * s.rho * s.u0 * dV;
      cons[SRR] = jac * (rhoh * s.u0 * s.l[0] - s.b0 * s.bd[0]) * dV;
      cons[LLL] = jac * (rhoh * s.u0 * s.l[1] - s.b0 * s.bd[1]) * dV;
      cons[SZZ] = jac * (rhoh * s.u0 * s.l[2] - s.b0 * s.bd[2]) * dV;
      cons[TAU] = jac * (-(rhoe + 0.5 * s.b2) * s.uU * s.u0 - (s.Pp + 0.5 * s.b2) * (s.uU * s.u0 + g.U[0]) -
                         s.rho * (s.uU + 1.0) * s.u0 + s.Ub * s.b0) * dV;
      cons[BRR] = g.sqrtgam * s.B[0] / r * dV;
      cons[BPP] = g.sqrtgam * s.B[1] * dV;
      cons[BZZ] = g.sqrtgam * s.B[2] * dV;
#pragma unroll
      for (int q = NC; q < NQ; ++q) cons[q] = prim[q] * cons[DDD];
   }

   // flux (grmhd2.c:272-381) through a face of unit normal n at radius r.
   static DHD void flux(const dh_params &P, const GRMetric &g, double r, const double *prim,
                        const double *n, double *F)
   {
      double gl = P.gamma_law;
      GRState s = gr_state(gl, g, r, prim);
      double v[3];
      for (int i = 0; i < 3; ++i) v[i] = s.u[i] / s.u0;
      double un = s.u[0] * n[0] + s.u[1] * n[1] + s.u[2] * n[2];
      double Bn = s.B[0] * n[0] + s.B[1] * n[1] + s.B[2] * n[2];
      double Un = g.U[1] * n[0] + g.U[2] * n[1] + g.U[3] * n[2];
      double vn = v[0] * n[0] + v[1] * n[1] + v[2] * n[2];
      double hn = n[0] + r * n[1] + n[2];
      double bn = (Bn + s.uB * un) / s.w;
      double rhoh = s.rho + gl / (gl - 1.0) * s.Pp + s.b2;
      double rhoe = s.Pp / (gl - 1.0);
      double Pt = s.Pp + 0.5 * s.b2;
      double jac = g.jac;
      F[DDD] = jac * hn * s.rho * un;
      F[SRR] = jac * hn * (rhoh * un * s.l[0] + Pt * n[0] - s.bd[0] * bn);
      F[LLL] = jac * hn * (rhoh * un * s.l[1] + Pt * n[1] - s.bd[1] * bn);
      F[SZZ] = jac * hn * (rhoh * un * s.l[2] + Pt * n[2] - s.bd[2] * bn);
      F[TAU] = jac * hn * (-(rhoe + 0.5 * s.b2) * s.uU * un - Pt * (s.uU * un + Un) -
                           s.rho * (s.uU + 1.0) * un + s.Ub * bn);
      F[BRR] = g.sqrtgam * hn * (s.B[0] * vn - Bn * v[0]) / r;
      F[BPP] = g.sqrtgam * hn * (s.B[1] * vn - Bn * v[1]);
      F[BZZ] = g.sqrtgam * hn * (s.B[2] * vn - Bn * v[2]);
#pragma unroll
      for (int q = NC; q < NQ; ++q) F[q] = prim[q] * F[DDD];
   }

   // getUstar (grmhd2.c:182-251): the HLLC star state on side k.
   static DHD void ustar(const dh_params &P, const GRMetric &g, double r, const double *prim,
                         double Sk, double Ss, const double *n, double *Us)
   {
      double gl = P.gamma_law;
      double rho = prim[RHO], Pp = prim[PPP];
      double l[3] = {prim[URR], prim[UPP], prim[UZZ]};
      double al = g.al;
      const double *be = g.be, *igam = g.igam, *U = g.U;
      double bn = n[0] * be[0] + n[1] * be[1] + n[2] * be[2];
      double ign = n[0] * igam[0] + n[1] * igam[4] + n[2] * igam[8];
      double Un = n[0] * U[1] + n[1] * U[2] + n[2] * U[3];
      double hn = n[0] + r * n[1] + n[2];
      double ss = Ss / hn, sk = Sk / hn;
      double uS[3], u2 = 0.0;
      for (int i = 0; i < 3; ++i) {
         uS[i] = 0.0;
         for (int j = 0; j < 3; ++j) uS[i] += igam[3 * i + j] * l[j];
         u2 += l[i] * uS[i];
      }
      double w = sqrt(1 + u2);
      double u0 = w / al;
      double l0 = -al * w + be[0] * l[0] + be[1] * l[1] + be[2] * l[2];
      double vn = al * (n[0] * uS[0] + n[1] * uS[1] + n[2] * uS[2]) / w - bn;
      double uU = l0 * U[0] + l[0] * U[1] + l[1] * U[2] + l[2] * U[3];
      double rhoh = rho + gl / (gl - 1.0) * Pp;
      double qE = rhoh * w * w * (vn - sk) + Pp * (sk + bn);
      double mn = rhoh * w * (n[0] * uS[0] + n[1] * uS[1] + n[2] * uS[2]);
      double qMn = mn * (vn - sk) + al * Pp * ign;
      double ssS = (ss + bn) / al;
      double skS = (sk + bn) / al;
      double Pstar = (qMn - ssS * qE) / (al * (ign - ssS * skS));
      double kappa = (vn - sk) / (ss - sk);
      double alpha1 = (Pp - Pstar) / (ss - sk);
      double alpha2 = (vn * Pp - ss * Pstar) / (ss - sk);
      double rhoe = Pp / (gl - 1.0);
      double tau = -rhoe * uU * u0 - Pp * (u0 * uU + U[0]) - rho * (uU + 1.0) * u0;
      double jac = g.jac;
      Us[DDD] = jac * rho * u0 * kappa;
      Us[SRR] = jac * (rhoh * u0 * l[0] * kappa + alpha1 * n[0]);
      Us[LLL] = jac * (rhoh * u0 * l[1] * kappa + alpha1 * n[1]);
      Us[SZZ] = jac * (rhoh * u0 * l[2] * kappa + alpha1 * n[2]);
      Us[TAU] = jac * (tau * kappa - Un * alpha1 + U[0] * alpha2);
      // the field components of the star state stay zero (grmhd2.c:182-251
      // leaves them unset; riemann.c:222-226 zeroes Ustr)
      Us[BRR] = Us[BPP] = Us[BZZ] = 0.0;
#pragma unroll
      for (int q = NC; q < NQ; ++q) Us[q] = prim[q] * Us[DDD];
   }

   // vel (grmhd2.c:542-746, EXACT_DISP 0): outer wave speeds Sl, Sr and the
   // contact speed Ss of the face.
   static DHD void vel(const dh_params &P, const GRMetric &g, double r, const double *prim1,
                       const double *prim2, const double *n, double *Sl, double *Sr, double *Ss)
   {
      double gl = P.gamma_law;
      double al = g.al;
      const double *be = g.be, *gam = g.gam, *igam = g.igam;
      double sqrtgam = r / (al * r);  // metric_jacobian / (al r)
      double rho1 = prim1[RHO], P1 = prim1[PPP];
      double l1[3] = {prim1[URR], prim1[UPP], prim1[UZZ]};
      double B1[3] = {prim1[BRR] / sqrtgam, prim1[BPP] / (r * sqrtgam), prim1[BZZ] / sqrtgam};
      double cs21 = gl * P1 / (rho1 + gl / (gl - 1.0) * P1);
      double rho2 = prim2[RHO], P2 = prim2[PPP];
      double l2[3] = {prim2[URR], prim2[UPP], prim2[UZZ]};
      double B2[3] = {prim2[BRR] / sqrtgam, prim2[BPP] / (r * sqrtgam), prim2[BZZ] / sqrtgam};
      double cs22 = gl * P2 / (rho2 + gl / (gl - 1.0) * P2);
      double u21 = 0.0, u22 = 0.0, uS1[3], uS2[3];
      for (int i = 0; i < 3; ++i) {
         uS1[i] = 0.0;
         uS2[i] = 0.0;
         for (int j = 0; j < 3; ++j) {
            uS1[i] += igam[3 * i + j] * l1[j];
            uS2[i] += igam[3 * i + j] * l2[j];
         }
         u21 += uS1[i] * l1[i];
         u22 += uS2[i] * l2[i];
      }
      double w1 = sqrt(1.0 + u21), w2 = sqrt(1.0 + u22);
      double v21 = u21 / (w1 * w1), v22 = u22 / (w2 * w2);
      double uB1 = l1[0] * B1[0] + l1[1] * B1[1] + l1[2] * B1[2];
      double uB2 = l2[0] * B2[0] + l2[1] * B2[1] + l2[2] * B2[2];
      double B21 = 0.0, B22 = 0.0;
      for (int i = 0; i < 3; ++i)
         for (int j = 0; j < 3; ++j) {
            B21 += gam[3 * i + j] * B1[i] * B1[j];
            B22 += gam[3 * i + j] * B2[i] * B2[j];
         }
      double b21 = (B21 + uB1 * uB1) / (w1 * w1);
      double b22 = (B22 + uB2 * uB2) / (w2 * w2);
      double va21 = b21 / (rho1 + gl / (gl - 1.0) * P1 + b21);
      double va22 = b22 / (rho2 + gl / (gl - 1.0) * P2 + b22);
      double cf21 = cs21 + va21 - cs21 * va21;
      double cf22 = cs22 + va22 - cs22 * va22;
      double vSn1 = (uS1[0] * n[0] + uS1[1] * n[1] + uS1[2] * n[2]) / w1;
      double vSn2 = (uS2[0] * n[0] + uS2[1] * n[1] + uS2[2] * n[2]) / w2;
      double bn = (be[0] * n[0] + be[1] * n[1] + be[2] * n[2]);
      double ign = igam[3 * 0 + 0] * n[0] + igam[3 * 1 + 1] * n[1] + igam[3 * 2 + 2] * n[2];
      double dv1 = sqrt(cf21 * (ign - vSn1 * vSn1 - cf21 * (ign * v21 - vSn1 * vSn1))) / w1;
      double dv2 = sqrt(cf22 * (ign - vSn2 * vSn2 - cf22 * (ign * v22 - vSn2 * vSn2))) / w2;
      double hn = n[0] + r * n[1] + n[2];
      double sl1 = hn * (al * (vSn1 * (1.0 - cf21) - dv1) / (1.0 - v21 * cf21) - bn);
      double sr1 = hn * (al * (vSn1 * (1.0 - cf21) + dv1) / (1.0 - v21 * cf21) - bn);
      double sl2 = hn * (al * (vSn2 * (1.0 - cf22) - dv2) / (1.0 - v22 * cf22) - bn);
      double sr2 = hn * (al * (vSn2 * (1.0 - cf22) + dv2) / (1.0 - v22 * cf22) - bn);
      *Sr = sr1 > sr2 ? sr1 : sr2;
      *Sl = sl1 < sl2 ? sl1 : sl2;

      // contact speed from the HLL state of energy and normal momentum
      double sL = *Sl / hn, sR = *Sr / hn;
      double rhohL = rho1 + gl / (gl - 1) * P1;
      double rhohR = rho2 + gl / (gl - 1) * P2;
      double vnL = al * vSn1 - bn, vnR = al * vSn2 - bn;
      double ML[3] = {rhohL * w1 * l1[0], rhohL * w1 * l1[1], rhohL * w1 * l1[2]};
      double MR[3] = {rhohR * w2 * l2[0], rhohR * w2 * l2[1], rhohR * w2 * l2[2]};
      double EL = rhohL * w1 * w1 - P1;
      double ER = rhohR * w2 * w2 - P2;
      double FML[3], FMR[3];
      for (int i = 0; i < 3; ++i) {
         FML[i] = ML[i] * vnL + n[i] * al * P1;
         FMR[i] = MR[i] * vnR + n[i] * al * P2;
      }
      double FEL = EL * vnL + P1 * (vnL + bn);
      double FER = ER * vnR + P2 * (vnR + bn);
      double UE = (sR * ER - sL * EL - FER + FEL) / (sR - sL);
      double FE = ((sR * FEL - sL * FER + sL * sR * (ER - EL)) / (sR - sL) - bn * UE) / al;
      double UM = 0.0, FM = 0.0;
      for (int i = 0; i < 3; ++i) {
         double UM_hll = (sR * MR[i] - sL * ML[i] - FMR[i] + FML[i]) / (sR - sL);
         double FM_hll = ((sR * FML[i] - sL * FMR[i] + sL * sR * (MR[i] - ML[i])) / (sR - sL) -
                          bn * UM_hll) / al;
         double igi = n[0] * igam[3 * i] + n[1] * igam[3 * i + 1] + n[2] * igam[3 * i + 2];
         UM += igi * UM_hll;
         FM += igi * FM_hll;
      }
      double A = FE;
      double B = -FM - ign * UE;
      double C = ign * UM;
      double sS;
      if (fabs(4 * A * C / (B * B)) < 1.0e-7)
         sS = -C / B * (1.0 + A * C / (B * B) + 2 * A * A * C * C / (B * B * B * B));
      else
         sS = (-B - sqrt(B * B - 4 * A * C)) / (2 * A);
      *Ss = hn * (al * sS - bn);
   }

   // get_Ustar_HLLD (hlld_grmhd.c:90-368): the relativistic HLLD fan solved
   // in the face's orthonormal frame, moved back to coordinate components.
   static DHD void riemann_hlld(const dh_params &P, const double *pL, const double *pR,
                                double r, double z, const double *n, double w, double *F,
                                double *U)
   {
      double gl = P.gamma_law;
      GRMetric g = gr_metric(P, r, z);
      double al = g.al, jac = g.jac, sqrtgam = g.sqrtgam;
      const double *be = g.be, *gam = g.gam, *igam = g.igam, *Uf = g.U;
      double rhoL = pL[RHO], PL = pL[PPP];
      double lL[3] = {pL[URR], pL[UPP], pL[UZZ]};
      double BL[3] = {pL[BRR] / sqrtgam, pL[BPP] / (r * sqrtgam), pL[BZZ] / sqrtgam};
      double rhoR = pR[RHO], PR = pR[PPP];
      double lR[3] = {pR[URR], pR[UPP], pR[UZZ]};
      double BR[3] = {pR[BRR] / sqrtgam, pR[BPP] / (r * sqrtgam), pR[BZZ] / sqrtgam};
      int xx = (n[0] > n[1] && n[0] > n[2]) ? 0 : (n[1] > n[2] ? 1 : 2);
      // the face's four-velocity
      if (xx == 1) w /= r;
      double be2 = 0.0, bex = 0.0;
      for (int i = 0; i < 3; ++i)
         for (int j = 0; j < 3; ++j) be2 += gam[3 * i + j] * be[i] * be[j];
      for (int i = 0; i < 3; ++i) bex += gam[3 * xx + i] * be[i];
      double uw[4] = {0.0, 0.0, 0.0, 0.0};
      uw[0] = 1.0 / sqrt(al * al - be2 - 2 * bex * w - gam[3 * xx + xx] * w * w);
      uw[xx + 1] = uw[0] * w;
      double e[16], ie[16];
      grh::tetrads(al, be, gam, igam, r * sqrtgam, xx, e, ie);
      // the states in the orthonormal frame (e_(t) is the normal observer)
      double uLo[4], uRo[4], BLo[3], BRo[3], uwo[4];
      for (int i = 0; i < 3; ++i) {
         uLo[i + 1] = 0.0;
         uRo[i + 1] = 0.0;
         BLo[i] = 0.0;
         BRo[i] = 0.0;
         for (int j = 0; j < 3; ++j) {
            uLo[i + 1] += e[4 * (i + 1) + j + 1] * lL[j];
            uRo[i + 1] += e[4 * (i + 1) + j + 1] * lR[j];
            BLo[i] += ie[4 * (i + 1) + j + 1] * BL[j];
            BRo[i] += ie[4 * (i + 1) + j + 1] * BR[j];
         }
      }
      uLo[0] = sqrt(1.0 + uLo[1] * uLo[1] + uLo[2] * uLo[2] + uLo[3] * uLo[3]);
      uRo[0] = sqrt(1.0 + uRo[1] * uRo[1] + uRo[2] * uRo[2] + uRo[3] * uRo[3]);
      for (int k = 0; k < 4; ++k) uwo[k] = ie[4 * k] * uw[0] + ie[4 * k + xx + 1] * uw[xx + 1];
      double Bx = 0.5 * (BLo[0] + BRo[0]);
      BLo[0] = Bx;
      BRo[0] = Bx;
      double wo = uwo[1] / uwo[0];
      // outer wave speeds in the orthonormal frame
      double rhohL = rhoL + gl / (gl - 1) * PL;
      double rhohR = rhoR + gl / (gl - 1) * PR;
      double cs2L = gl * PL / rhohL, cs2R = gl * PR / rhohR;
      double b0Lo = uLo[1] * BLo[0] + uLo[2] * BLo[1] + uLo[3] * BLo[2];
      double b0Ro = uRo[1] * BRo[0] + uRo[2] * BRo[1] + uRo[3] * BRo[2];
      double b2L = (BLo[0] * BLo[0] + BLo[1] * BLo[1] + BLo[2] * BLo[2] + b0Lo * b0Lo) / (uLo[0] * uLo[0]);
      double b2R = (BRo[0] * BRo[0] + BRo[1] * BRo[1] + BRo[2] * BRo[2] + b0Ro * b0Ro) / (uRo[0] * uRo[0]);
      double sL1, sL2, sR1, sR2;
      grh::fast_speeds(cs2L, uLo[0], uLo[1] / uLo[0], b2L, rhohL, 1.0, 0.0, 1.0, &sL1, &sL2);
      grh::fast_speeds(cs2R, uRo[0], uRo[1] / uRo[0], b2R, rhohR, 1.0, 0.0, 1.0, &sR1, &sR2);
      double sLo = sL1 < sR1 ? sL1 : sR1;
      double sRo = sL2 > sR2 ? sL2 : sR2;
      double ULo[8], FLo[8], URo[8], FRo[8], Uo[8], Fo[8];
      grh::state_UF(gl, rhoL, PL, uLo, BLo, ULo, FLo);
      grh::state_UF(gl, rhoR, PR, uRo, BRo, URo, FRo);
      double primLo[5] = {rhoL, PL, uLo[1], uLo[2], uLo[3]};
      double primRo[5] = {rhoR, PR, uRo[1], uRo[2], uRo[3]};
      int zone = grh::solve_fan(gl, sLo, sRo, Bx, ULo, FLo, URo, FRo, wo, primLo, primRo, Uo, Fo);
      // back to coordinate components: the energy-momentum covector and the
      // field vector of the fan state and flux
      using namespace grh;
      double Uc[4], Fc[4], UB[3], FB[3];
      for (int mu = 0; mu < 4; ++mu) {
         Uc[mu] = -ie[mu] * Uo[EN] + ie[4 + mu] * Uo[MX] + ie[8 + mu] * Uo[MY] + ie[12 + mu] * Uo[MZ];
         Fc[mu] = -ie[mu] * Fo[EN] + ie[4 + mu] * Fo[MX] + ie[8 + mu] * Fo[MY] + ie[12 + mu] * Fo[MZ];
      }
      for (int i = 0; i < 3; ++i) {
         UB[i] = e[4 + i + 1] * Uo[BX] + e[8 + i + 1] * Uo[BY] + e[12 + i + 1] * Uo[BZ];
         FB[i] = e[4 + i + 1] * Fo[BX] + e[8 + i + 1] * Fo[BY] + e[12 + i + 1] * Fo[BZ];
      }
      // Uc = {tau, S_r, S_phi, S_z} before the frame subtraction
      double hn = 1.0 * n[0] + r * n[1] + 1.0 * n[2];
      double t0 = e[0], t1 = e[4], x0 = e[xx + 1], x1 = e[4 + xx + 1];
      U[DDD] = jac * (t0 * Uo[DD] + t1 * Fo[DD]);
      U[SRR] = jac * (t0 * Uc[1] + t1 * Fc[1]);
      U[LLL] = jac * (t0 * Uc[2] + t1 * Fc[2]);
      U[SZZ] = jac * (t0 * Uc[3] + t1 * Fc[3]);
      U[TAU] = jac * (t0 * Uc[0] + t1 * Fc[0]);
      U[BRR] = sqrtgam * (t0 * UB[0] + t1 * FB[0]);
      U[BPP] = sqrtgam * (t0 * UB[1] + t1 * FB[1]);
      U[BZZ] = sqrtgam * (t0 * UB[2] + t1 * FB[2]);
      F[DDD] = jac * hn * (x0 * Uo[DD] + x1 * Fo[DD]);
      F[SRR] = jac * hn * (x0 * Uc[1] + x1 * Fc[1]);
      F[LLL] = jac * hn * (x0 * Uc[2] + x1 * Fc[2]);
      F[SZZ] = jac * hn * (x0 * Uc[3] + x1 * Fc[3]);
      F[TAU] = jac * hn * (x0 * Uc[0] + x1 * Fc[0]);
      F[BRR] = sqrtgam * hn * (x0 * UB[0] + x1 * FB[0]);
      F[BPP] = sqrtgam * hn * (x0 * UB[1] + x1 * FB[1]);
      F[BZZ] = sqrtgam * hn * (x0 * UB[2] + x1 * FB[2]);
      // the frame's energy and the rest mass out of tau
      U[TAU] = -Uf[0] * U[TAU] - Uf[1] * U[SRR] - Uf[2] * U[LLL] - Uf[3] * U[SZZ] - U[DDD];
      F[TAU] = -Uf[0] * F[TAU] - Uf[1] * F[SRR] - Uf[2] * F[LLL] - Uf[3] * F[SZZ] - F[DDD];
      U[BRR] /= r;
      F[BRR] /= r;
      const double *pk = (zone == 0 || zone == 1 || zone == 2) ? pL : pR;
#pragma unroll
      for (int q = NC; q < NQ; ++q) {
         U[q] = pk[q] * U[DDD];
         F[q] = pk[q] * F[DDD];
      }
   }

   // solve_riemann (riemann.c:216-292): HLL, HLLC or HLLD flux and fan
   // state at face speed w.  (r, z) is the face's position; r is also the
   // phi scale factor.
   static DHD void riemann(const dh_params &P, const double *pL, const double *pR, double r,
                           double z, const double *n, double w, double *F, double *U)
   {
      if (P.riemann_solver == RS_HLLD) {
         riemann_hlld(P, pL, pR, r, z, n, w, F, U);
         return;
      }
      GRMetric g = gr_metric(P, r, z);
      double Sl, Sr, Ss;
      vel(P, g, r, pL, pR, n, &Sl, &Sr, &Ss);
      if (w < Sl) {
         flux(P, g, r, pL, n, F);
         prim2cons(P, pL, U, r, z, 1.0);
      } else if (w > Sr) {
         flux(P, g, r, pR, n, F);
         prim2cons(P, pR, U, r, z, 1.0);
      } else if (P.riemann_solver == RS_HLL) {
         double Fl[NQ], Fr[NQ], Ul[NQ], Ur[NQ];
         double aL = Sr, aR = -Sl;
         prim2cons(P, pL, Ul, r, z, 1.0);
         prim2cons(P, pR, Ur, r, z, 1.0);
         flux(P, g, r, pL, n, Fl);
         flux(P, g, r, pR, n, Fr);
#pragma unroll
         for (int q = 0; q < NQ; ++q) {
            F[q] = (aL * Fl[q] + aR * Fr[q] + aL * aR * (Ul[q] - Ur[q])) / (aL + aR);
            U[q] = (aR * Ul[q] + aL * Ur[q] + Fl[q] - Fr[q]) / (aL + aR);
         }
      } else {
         const double *pk = w < Ss ? pL : pR;
         double Sk = w < Ss ? Sl : Sr;
         double Uk[NQ], Fk[NQ];
         prim2cons(P, pk, Uk, r, z, 1.0);
         ustar(P, g, r, pk, Sk, Ss, n, U);
         flux(P, g, r, pk, n, Fk);
#pragma unroll
         for (int q = 0; q < NQ; ++q) F[q] = Fk[q] + Sk * (U[q] - Uk[q]);
      }
   }

   // cons2prim (grmhd2.c:253-270) with the Noble et al. 2-D Newton solve
   // (grmhd2.c:1096-1314); prim holds the previous primitives on entry (the
   // initial guess) and the result on exit.  (r, z) is the cell centroid.
   static DHD void cons2prim(const dh_params &P, const double *cons_in, double *prim, double r,
                             double z, double dV)
   {
      cons2prim_guess(P, cons_in, prim, r, z, dV, nullptr, nullptr);
   }

   // cons2prim with a learned initial guess (v^2, eta) when guess is not
   // null and P.c2p_guess asks for it: a guess outside v^2 in [0, 1),
   // eta >= 1 (or not finite), or a solve from it that does not converge,
   // falls back to the previous primitives.  info, when not null, receives
   // the solve's record (C2P_INFO_*).
   static DHD void cons2prim_guess(const dh_params &P, const double *cons_in, double *prim,
                                   double r, double z, double dV, const float *guess,
                                   unsigned char *info)
   {
      double gl = P.gamma_law;
      double cons[NQ];
#pragma unroll
      for (int q = 0; q < NQ; ++q) cons[q] = cons_in[q] / dV;
      C2pInv v;
      c2p_invariants(gr_metric(P, r, z), r, cons, v);
      const GRMetric &g = v.g;
      double n = (gl - 1.0) / gl;

      // the plain guess: the previous v^2 and eta
      double u2 = 0.0;
      double l[3] = {prim[URR], prim[UPP], prim[UZZ]};
      for (int i = 0; i < 3; ++i)
         for (int j = 0; j < 3; ++j) u2 += g.igam[3 * i + j] * l[i] * l[j];
      double w = sqrt(1.0 + u2);
      double v2p = u2 / (1.0 + u2);
      double etap = w * (1.0 + gl / (gl - 1.0) * prim[PPP] / prim[RHO]);

      // the model's guess is taken where it is physical; with
      // DH_C2P_GUESS_BETTER only where it also starts closer to the
      // solution than the previous primitives (a smaller residual)
      bool learned = guess && P.c2p_guess != DH_C2P_GUESS_PREVIOUS && c2p_guess_ok(guess) &&
                     (P.c2p_guess != DH_C2P_GUESS_BETTER ||
                      noble2d_residual(v, n, (double)guess[0], (double)guess[1]) <
                          noble2d_residual(v, n, v2p, etap));
      double v2 = learned ? (double)guess[0] : v2p;
      double eta = learned ? (double)guess[1] : etap;
      int it = 0;
      bool ok = noble2d(v, n, P.c2p_stop, v2, eta, it);
      bool fell_back = false;
      if (learned && !ok) {
         int it2 = 0;
         v2 = v2p;
         eta = etap;
         ok = noble2d(v, n, P.c2p_stop, v2, eta, it2);
         it += it2;
         fell_back = true;
      }
      if (info) {
         int k = ok && it < C2P_INFO_ITERS ? it : C2P_INFO_ITERS;
         *info = (unsigned char)(k | (learned ? C2P_INFO_LEARNED : 0) |
                                 (fell_back || (guess && !learned) ? C2P_INFO_FALLBACK : 0));
      }
      w = 1.0 / sqrt(1 - v2);
      double u0 = w / g.al;
      double h = eta / w;
      double rho = v.D / (g.jac * u0);
      if (rho < P.rho_floor) rho = P.rho_floor;
      double Pp = n * rho * (h - 1.0);
      if (Pp < P.p_floor * rho) Pp = P.p_floor * rho;
      h = 1.0 + gl / (gl - 1.0) * Pp / rho;
      const double *gam = g.gam, *B = v.B;
      double Bd[3];
      for (int i = 0; i < 3; ++i)
         Bd[i] = gam[3 * i] * B[0] + gam[3 * i + 1] * B[1] + gam[3 * i + 2] * B[2];
      double sqrtgam = g.sqrtgam, D = v.D;
      for (int i = 0; i < 3; ++i)
         l[i] = (v.S[i] + sqrtgam * v.BS * Bd[i] / (D * h * w)) / (D * h + sqrtgam * v.B2 / w);
      prim[RHO] = rho;
      prim[URR] = l[0];
      prim[UPP] = l[1];
      prim[UZZ] = l[2];
      prim[PPP] = Pp;
      prim[BRR] = sqrtgam * B[0];
      prim[BPP] = sqrtgam * r * B[1];
      prim[BZZ] = sqrtgam * B[2];
#pragma unroll
      for (int q = NC; q < NQ; ++q) prim[q] = cons[q] / cons[DDD];
   }

   // source (grmhd2.c:383-500): 1/2 T^{mu nu} d_i g_{mu nu} in the momenta
   // and the frame term in the energy, at the cell's mid point.
   static DHD void source(const dh_params &P, const double *prim, double *cons, double rp,
                          double rm, double phip, double phim, double zp, double zm,
                          double dVdt, int polar, double Pmax)
   {
      double gl = P.gamma_law, M = P.metric_par2;
      double r = 0.5 * (rm + rp), z = gr_z(P, 0.5 * (zm + zp));
      GRMetric g = gr_metric(P, r, z);
      double al = g.al, jac = g.jac, sqrtgam = g.sqrtgam;
      const double *be = g.be, *gam = g.gam, *igam = g.igam, *U = g.U;
      double dU[16];
      gr_metric_frame_der_U(P, r, z, dU);
      double rho = prim[RHO], Pp = prim[PPP];
      double l[4] = {0.0, prim[URR], prim[UPP], prim[UZZ]};
      double B[3] = {prim[BRR] / sqrtgam, prim[BPP] / (r * sqrtgam), prim[BZZ] / sqrtgam};
      double ig[16];
      double ia2 = 1.0 / (al * al);
      ig[0] = -ia2;
      for (int mu = 0; mu < 3; ++mu) {
         ig[mu + 1] = be[mu] * ia2;
         ig[4 * (mu + 1)] = ig[mu + 1];
         for (int nu = 0; nu < 3; ++nu) ig[4 * (mu + 1) + nu + 1] = igam[3 * mu + nu] - be[mu] * be[nu] * ia2;
      }
      double igaml[3];
      for (int i = 0; i < 3; ++i)
         igaml[i] = igam[3 * i] * l[1] + igam[3 * i + 1] * l[2] + igam[3 * i + 2] * l[3];
      double u2 = l[1] * igaml[0] + l[2] * igaml[1] + l[3] * igaml[2];
      double w = sqrt(1.0 + u2);
      double u[4];
      u[0] = w / al;
      for (int i = 0; i < 3; ++i) u[i + 1] = igaml[i] - be[i] * u[0];
      l[0] = -al * w + be[0] * l[1] + be[1] * l[2] + be[2] * l[3];
      double uB = l[1] * B[0] + l[2] * B[1] + l[3] * B[2];
      double b[4], bd[4];
      b[0] = uB / al;
      for (int i = 0; i < 3; ++i) b[i + 1] = (B[i] + u[i + 1] * uB) / w;
      for (int i = 0; i < 3; ++i)
         bd[i + 1] = (gam[3 * i] * B[0] + gam[3 * i + 1] * B[1] + gam[3 * i + 2] * B[2] + l[i + 1] * uB) / w;
      double B2 = 0.0, beB = 0.0;
      for (int i = 0; i < 3; ++i)
         for (int j = 0; j < 3; ++j) {
            beB += gam[3 * i + j] * be[i] * B[j];
            B2 += gam[3 * i + j] * B[i] * B[j];
         }
      bd[0] = (beB + l[0] * uB) / w;
      double b2 = (B2 + uB * uB) / (w * w);
      double rhoh = rho + gl / (gl - 1.0) * Pp + b2;
      double Pt = Pp + 0.5 * b2;
      const double rcorr = 1.0, pcorr = 1.0;
      double Sk[3];
      for (int i = 0; i < 3; ++i) {
         Sk[i] = 0.0;
         if (gr_killing(i + 1)) continue;
         double dg[16];
         gr_metric_der_g(P, r, z, i + 1, dg);
         for (int mu = 0; mu < 4; ++mu)
            for (int nu = 0; nu < 4; ++nu)
               Sk[i] += (rhoh * u[mu] * u[nu] * rcorr * pcorr + ig[4 * mu + nu] * Pt -
                         b[mu] * b[nu] * rcorr * pcorr) * dg[4 * mu + nu];
         Sk[i] *= 0.5;
      }
      double S0 = -U[1] * Sk[0] - U[2] * Sk[1] - U[3] * Sk[2];
      for (int mu = 0; mu <= 3; ++mu)
         for (int nu = 0; nu <= 3; ++nu) {
            if (mu == nu)
               S0 += -(rhoh * u[mu] * l[nu] + Pt - b[mu] * bd[nu]) * dU[4 * mu + nu];
            else
               S0 += -(rhoh * u[mu] * l[nu] - b[mu] * bd[nu]) * dU[4 * mu + nu];
         }
      cons[SRR] += jac * Sk[0] * dVdt;
      cons[LLL] += jac * Sk[1] * dVdt;
      cons[SZZ] += jac * Sk[2] * dVdt;
      cons[TAU] += jac * S0 * dVdt;
      (void)phip; (void)phim; (void)polar; (void)Pmax;
   }

   static DHD void visc_source(const dh_params &, const double *, const double *,
                               const double *, const double *, double *, double, double)
   {
   }
   static DHD void visc_flux(const dh_params &, const double *, const double *,
                             const double *, const double *, double *, double,
                             const double *)
   {
   }

   // mindt (grmhd2.c:748-832) at the cell's mid point; w is the mean
   // angular velocity of the cell's two phi faces.
   static DHD double mindt(const dh_params &P, const double *prim, double wc, double rp,
                           double rm, double phip, double phim, double zp, double zm,
                           double Pmax)
   {
      double gl = P.gamma_law;
      double r = 0.5 * (rm + rp);
      GRMetric g = gr_metric(P, r, 0.5 * (zm + zp));
      double al = g.al;
      const double *be = g.be, *gam = g.gam, *igam = g.igam;
      double sqrtgam = r / (r * al);  // metric_jacobian / (r al)
      double rho = prim[RHO], Pp = prim[PPP];
      double l[3] = {prim[URR], prim[UPP], prim[UZZ]};
      double B[3] = {prim[BRR] / sqrtgam, prim[BPP] / (r * sqrtgam), prim[BZZ] / sqrtgam};
      double cs = sqrt(gl * Pp / (rho + gl / (gl - 1) * Pp));
      double uS[3], u2 = 0.0;
      for (int i = 0; i < 3; ++i) {
         uS[i] = 0.0;
         for (int j = 0; j < 3; ++j) uS[i] += igam[3 * i + j] * l[j];
         u2 += uS[i] * l[i];
      }
      double w = sqrt(1.0 + u2);
      double vS[3];
      for (int i = 0; i < 3; ++i) vS[i] = uS[i] / w;
      double v2 = u2 / (w * w);
      double uB = l[0] * B[0] + l[1] * B[1] + l[2] * B[2];
      double B2 = 0.0;
      for (int i = 0; i < 3; ++i)
         for (int j = 0; j < 3; ++j) B2 += gam[3 * i + j] * B[i] * B[j];
      double b2 = (B2 + uB * uB) / (w * w);
      double va2 = b2 / (rho + gl / (gl - 1.0) * Pp + b2);
      double cf = sqrt(cs * cs + va2 - cs * cs * va2);
      double sig = 1 - cf * cf;
      double dvr = cf * sqrt(igam[0] * (1 - cf * cf * v2) - sig * vS[0] * vS[0]) / w;
      double vrl = fabs(al * (vS[0] * sig - dvr) / (1 - v2 * cf * cf) - be[0]);
      double vrr = fabs(al * (vS[0] * sig + dvr) / (1 - v2 * cf * cf) - be[0]);
      double dvp = cf * sqrt(igam[4] * (1 - cf * cf * v2) - sig * vS[1] * vS[1]) / w;
      double vpl = fabs(r * (al * (vS[1] * sig - dvp) / (1 - v2 * cf * cf) - be[1] - wc));
      double vpr = fabs(r * (al * (vS[1] * sig + dvp) / (1 - v2 * cf * cf) - be[1] - wc));
      double dvz = cf * sqrt(igam[8] * (1 - cf * cf * v2) - sig * vS[2] * vS[2]) / w;
      double vzl = fabs(al * (vS[2] * sig - dvz) / (1 - v2 * cf * cf) - be[2]);
      double vzr = fabs(al * (vS[2] * sig + dvz) / (1 - v2 * cf * cf) - be[2]);
      double maxvr = vrr > vrl ? vrr : vrl;
      double maxvp = vpr > vpl ? vpr : vpl;
      double maxvz = vzr > vzl ? vzr : vzl;
      double dL0 = .5 * (rp + rm) * get_dp(phip, phim, Pmax);
      double dtr = (rp - rm) / maxvr;
      double dtp = dL0 / maxvp;
      double dtz = (zp - zm) / maxvz;
      double dt = dtr;
      dt = dt < dtp ? dt : dtp;
      dt = dt < dtz ? dt : dtz;
      return dt;
   }

   // get_omega (grmhd2.c:64-86): the coordinate angular velocity at (r, z).
   static DHD double omega(const dh_params &P, const double *prim, double r, double z)
   {
      GRMetric g = gr_metric(P, r, z);
      double l[3] = {prim[URR], prim[UPP], prim[UZZ]};
      double u2 = 0.0;
      for (int i = 0; i < 3; ++i)
         for (int j = 0; j < 3; ++j) u2 += g.igam[3 * i + j] * l[i] * l[j];
      double w = sqrt(1.0 + u2);
      return g.al * (g.igam[3] * l[0] + g.igam[4] * l[1] + g.igam[5] * l[2]) / w - g.be[1];
   }

   // flux_to_E (grmhd2.c:509-540): {E1, B1, E2, B2}; dim 0 phi, 1 r, 2 z.
   static DHD void flux_to_E(const double *F, const double *U, double r, int dim, double *emf)
   {
      if (dim == 0) {
         emf[0] = F[BRR] * r;
         emf[1] = U[BRR] * r * r;
         emf[2] = F[BZZ];
         emf[3] = U[BZZ] * r;
      } else if (dim == 1) {
         emf[0] = -F[BPP] * r;
         emf[1] = U[BRR] * r * r;
         emf[2] = 1.0 * F[BZZ];
         emf[3] = 0.0;
      } else {
         emf[0] = -F[BPP] * r;
         emf[1] = U[BZZ] * r;
         emf[2] = 1.0 * -F[BRR] * r;
         emf[3] = 0.0;
      }
   }
};

}  // namespace dh
