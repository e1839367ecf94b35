// dh_mhd.h -- Newtonian MHD on the moving cylindrical mesh (device side of the
// reference Hydro/mhd.c) with the HLLD Riemann solver (hlld.c).
//
// Primitive layout {rho, P, v_r, omega, v_z, B_r, B_phi, B_z, passives...};
// conserved {D, tau, S_r, L, S_z, B_r dV / r, B_phi dV / r, B_z dV, ...}
// (mhd.c:35-69).  The HLLD fan is solved in the orthonormal frame at the face
// (hlld.c:282-401), with the energy in the lab frame as the reference has it.
#pragma once
#include "dh_geom.h"

namespace dh {

template <int NQ>
struct MHD {
   static constexpr int NC = 8;

   // bfield_scale_factor (mhd.c:518-528)
   static DHD double bscale_r(double r) { return 1.0 / r; }

   // prim2cons (mhd.c:35-69)
   static DHD void prim2cons(const dh_params &P, const double *prim, double *cons,
                             double r, double /*z*/, double dV)
   {
      double rho = prim[RHO];
      double Pp = prim[PPP];
      double vr = prim[URR];
      double vp = prim[UPP] * r;
      double vz = prim[UZZ];
      double om = get_om(P, r);
      double vp_off = vp - om * r;
      double Br = prim[BRR], Bp = prim[BPP], Bz = prim[BZZ];
      double v2 = vr * vr + vp_off * vp_off + vz * vz;
      double B2 = Br * Br + Bp * Bp + Bz * Bz;
      double rhoe = Pp / (P.gamma_law - 1.);
      cons[DDD] = rho * dV;
      cons[TAU] = (.5 * rho * v2 + rhoe + .5 * B2) * dV;
      cons[SRR] = rho * vr * dV;
      cons[LLL] = r * rho * vp * dV;
      cons[SZZ] = rho * vz * dV;
      cons[BRR] = Br * dV * bscale_r(r);
      cons[BPP] = Bp * dV / r;
      cons[BZZ] = Bz * dV;
#pragma unroll
      for (int q = NC; q < NQ; ++q) cons[q] = prim[q] * cons[DDD];
   }

   // cons2prim (mhd.c:138-185)
   static DHD void cons2prim(const dh_params &P, const double *cons, double *prim,
                             double r, double /*z*/, double dV)
   {
      double rho = cons[DDD] / dV;
      if (rho < P.rho_floor) rho = P.rho_floor;
      double Sr = cons[SRR] / dV;
      double Sp = cons[LLL] / dV / r;
      double Sz = cons[SZZ] / dV;
      double E = cons[TAU] / dV;
      double om = get_om(P, r);
      double vr = Sr / rho;
      double vp = Sp / rho;
      double vp_off = vp - om * r;
      double vz = Sz / rho;
      double Br = cons[BRR] / (dV * bscale_r(r));
      double Bp = cons[BPP] / dV * r;
      double Bz = cons[BZZ] / dV;
      double B2 = Br * Br + Bp * Bp + Bz * Bz;
      double KE = .5 * (Sr * vr + rho * vp_off * vp_off + Sz * vz);
      double rhoe = E - KE - .5 * B2;
      double Pp = (P.gamma_law - 1.) * rhoe;
      if (Pp < P.p_floor * rho) Pp = P.p_floor * rho;
      if (P.isothermal) {
         double cs2 = get_cs2(P, r);
         Pp = cs2 * rho / P.gamma_law;
      }
      prim[RHO] = rho;
      prim[PPP] = Pp;
      prim[URR] = vr;
      prim[UPP] = vp / r;
      prim[UZZ] = vz;
      prim[BRR] = Br;
      prim[BPP] = Bp;
      prim[BZZ] = Bz;
#pragma unroll
      for (int q = NC; q < NQ; ++q) prim[q] = cons[q] / cons[DDD];
   }

   // source (mhd.c:225-271); polar = (NgRa == 0), mhd.c:21-22.
   static DHD void source(const dh_params &P, const double *prim, double *cons,
                          double rp, double rm, double phip, double phim, double zp,
                          double zm, double dVdt, int polar, double Pmax)
   {
      double dphi = get_dp(phip, phim, Pmax);
      double rho = prim[RHO];
      double Pp = prim[PPP];
      double r_1 = .5 * (rp + rm);
      double r2_3 = (rp * rp + rp * rm + rm * rm) / 3.;
      double vr = prim[URR];
      double omega = prim[UPP];
      double r = centroid_r(rp, rm);
      double Br = prim[BRR], Bp = prim[BPP], Bz = prim[BZZ];
      double B2 = Br * Br + Bp * Bp + Bz * Bz;
      double centrifugal;
      if (polar) {
         centrifugal = (rho * omega * omega * r2_3) / r_1 * sin(.5 * dphi) / (.5 * dphi);
         centrifugal -= Bp * Bp / r_1;
      } else {
         centrifugal = rho * omega * omega * r - Bp * Bp / r;
      }
      double press_bal = (Pp + .5 * B2) / r_1;
      cons[SRR] += dVdt * (centrifugal + press_bal);
      double om = get_om(P, r);
      double om1 = get_om1(P, r);
      cons[TAU] += dVdt * rho * vr * (om * om * r2_3 / r_1 - om1 * (omega - om) * r2_3);
      (void)zp; (void)zm;
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

   // mindt (mhd.c:452-483): fast-magnetosonic CFL time.
   static DHD double mindt(const dh_params &P, const double *prim, double w, double rp,
                           double rm, double phip, double phim, double zp, double zm,
                           double Pmax)
   {
      double r = centroid_r(rp, rm);
      double Pp = prim[PPP];
      double rho = prim[RHO];
      double vp = (prim[UPP] - w) * r;
      double vr = prim[URR];
      double vz = prim[UZZ];
      double cs = sqrt(P.gamma_law * Pp / rho);
      double Br = prim[BRR], Bp = prim[BPP], Bz = prim[BZZ];
      double b2 = (Br * Br + Bp * Bp + Bz * Bz) / rho;
      double cf = sqrt(cs * cs + b2);
      double maxvr = cf + fabs(vr);
      double maxvp = cf + fabs(vp);
      double maxvz = cf + fabs(vz);
      double dL0 = .5 * (rp + rm) * get_dp(phip, phim, Pmax);
      double dtr = (rp - rm) / maxvr;
      double dtp = dL0 / maxvp;
      double dtz = (zp - zm) / maxvz;
      double dt = dtr;
      if (dt > dtp) dt = dtp;
      if (dt > dtz) dt = dtz;
      return dt;
   }

   static DHD double omega(const dh_params &, const double *prim, double, double)
   {
      return prim[UPP];
   }

   // ---- HLLD (hlld.c) ----

   // get_velocities (hlld.c:20-70)
   static DHD void hd_velocities(double gam, const double *pL, const double *pR,
                                 const double *n, double *vel, double *Pt_star)
   {
      double rho_L = pL[RHO], P_L = pL[PPP];
      double u_L = pL[URR] * n[0] + pL[UPP] * n[1] + pL[UZZ] * n[2];
      double Bn_L = pL[BRR] * n[0] + pL[BPP] * n[1] + pL[BZZ] * n[2];
      double B2_L = pL[BRR] * pL[BRR] + pL[BPP] * pL[BPP] + pL[BZZ] * pL[BZZ];
      double Pt_L = P_L + .5 * B2_L;
      double rho_R = pR[RHO], P_R = pR[PPP];
      double u_R = pR[URR] * n[0] + pR[UPP] * n[1] + pR[UZZ] * n[2];
      double Bn_R = pR[BRR] * n[0] + pR[BPP] * n[1] + pR[BZZ] * n[2];
      double B2_R = pR[BRR] * pR[BRR] + pR[BPP] * pR[BPP] + pR[BZZ] * pR[BZZ];
      double Pt_R = P_R + .5 * B2_R;
      double Bn = .5 * (Bn_L + Bn_R);
      double aL = gam * P_L + B2_L, aR = gam * P_R + B2_R;
      double cf2_L = .5 * (aL + sqrt(fabs(aL * aL - 4. * gam * P_L * Bn_L * Bn_L))) / rho_L;
      double cf2_R = .5 * (aR + sqrt(fabs(aR * aR - 4. * gam * P_R * Bn_R * Bn_R))) / rho_R;
      double cf2 = cf2_L > cf2_R ? cf2_L : cf2_R;
      double umin = u_L < u_R ? u_L : u_R;
      double umax = u_L < u_R ? u_R : u_L;
      double S_R = umax + sqrt(fabs(cf2));
      double S_L = umin - sqrt(fabs(cf2));
      double S_M = ((S_R - u_R) * rho_R * u_R - (S_L - u_L) * rho_L * u_L - Pt_R + Pt_L) /
                   ((S_R - u_R) * rho_R - (S_L - u_L) * rho_L);
      double rhostar_L = rho_L * (S_L - u_L) / (S_L - S_M);
      double rhostar_R = rho_R * (S_R - u_R) / (S_R - S_M);
      vel[0] = S_L;
      vel[1] = S_M - fabs(Bn) / sqrt(rhostar_L);
      vel[2] = S_M;
      vel[3] = S_M + fabs(Bn) / sqrt(rhostar_R);
      vel[4] = S_R;
      double num = (S_R - u_R) * rho_R * Pt_L - (S_L - u_L) * rho_L * Pt_R +
                   rho_L * rho_R * (S_R - u_R) * (S_L - u_L) * (u_R - u_L);
      double denom = (S_R - u_R) * rho_R - (S_L - u_L) * rho_L;
      *Pt_star = num / denom;
   }

   // prim_to_cons (hlld.c:72-98)
   static DHD void hd_cons(double gam, const double *p, double *U)
   {
      double rho = p[RHO], Pp = p[PPP], vx = p[URR], vy = p[UPP], vz = p[UZZ];
      double Bx = p[BRR], By = p[BPP], Bz = p[BZZ];
      double v2 = vx * vx + vy * vy + vz * vz;
      double B2 = Bx * Bx + By * By + Bz * Bz;
      U[DDD] = rho;
      U[SRR] = rho * vx;
      U[LLL] = rho * vy;
      U[SZZ] = rho * vz;
      U[TAU] = (.5 * rho * v2 + Pp / (gam - 1.) + .5 * B2);
      U[BRR] = Bx;
      U[BPP] = By;
      U[BZZ] = Bz;
   }

   // get_flux (hlld.c:100-133)
   static DHD void hd_flux(double gam, const double *p, double *F, const double *n)
   {
      double rho = p[RHO], Pp = p[PPP], vx = p[URR], vy = p[UPP], vz = p[UZZ];
      double Bx = p[BRR], By = p[BPP], Bz = p[BZZ];
      double v2 = vx * vx + vy * vy + vz * vz;
      double vB = vx * Bx + vy * By + vz * Bz;
      double B2 = Bx * Bx + By * By + Bz * Bz;
      double vn = vx * n[0] + vy * n[1] + vz * n[2];
      double Bn = Bx * n[0] + By * n[1] + Bz * n[2];
      double Pt = Pp + .5 * B2;
      F[DDD] = rho * vn;
      F[SRR] = rho * vx * vn + Pt * n[0] - Bx * Bn;
      F[LLL] = rho * vy * vn + Pt * n[1] - By * Bn;
      F[SZZ] = rho * vz * vn + Pt * n[2] - Bz * Bn;
      F[TAU] = (.5 * rho * v2 + Pp / (gam - 1.) + Pp + B2) * vn - vB * Bn;
      F[BRR] = Bx * vn - vx * Bn;
      F[BPP] = By * vn - vy * Bn;
      F[BZZ] = Bz * vn - vz * Bn;
   }

   // get_single_star (hlld.c:135-193)
   static DHD void hd_single_star(double gam, const double *p, double *Us, double S_K,
                                  double S_M, double Pt_star, const double *n)
   {
      double rho = p[RHO], vx = p[URR], vy = p[UPP], vz = p[UZZ], Pp = p[PPP];
      double Bx = p[BRR], By = p[BPP], Bz = p[BZZ];
      double u_K = vx * n[0] + vy * n[1] + vz * n[2];
      double Bn = Bx * n[0] + By * n[1] + Bz * n[2];
      double vdotB = vx * Bx + vy * By + vz * Bz;
      double v2 = vx * vx + vy * vy + vz * vz;
      double B2 = Bx * Bx + By * By + Bz * Bz;
      double Pt = Pp + .5 * B2;
      double e = .5 * rho * v2 + Pp / (gam - 1.) + .5 * B2;
      double rhostar = rho * (S_K - u_K) / (S_K - S_M);
      double denom = rho * (S_K - u_K) * (S_K - S_M) - Bn * Bn;
      if (fabs(denom) < 1e-10) denom = 1e-10;  // hlld.c:12 TOL
      double sk2 = (S_K - u_K) * (S_K - u_K);
      double vx_s = vx - Bn * Bx * (S_M - u_K) / denom;
      double vy_s = vy - Bn * By * (S_M - u_K) / denom;
      double vz_s = vz - Bn * Bz * (S_M - u_K) / denom;
      double Bx_s = Bx * (rho * sk2 - Bn * Bn) / denom;
      double By_s = By * (rho * sk2 - Bn * Bn) / denom;
      double Bz_s = Bz * (rho * sk2 - Bn * Bn) / denom;
      double vn_s = vx_s * n[0] + vy_s * n[1] + vz_s * n[2];
      double Bn_s = Bx_s * n[0] + By_s * n[1] + Bz_s * n[2];
      vx_s += (S_M - vn_s) * n[0];
      vy_s += (S_M - vn_s) * n[1];
      vz_s += (S_M - vn_s) * n[2];
      Bx_s += (Bn - Bn_s) * n[0];
      By_s += (Bn - Bn_s) * n[1];
      Bz_s += (Bn - Bn_s) * n[2];
      double vBs = vx_s * Bx_s + vy_s * By_s + vz_s * Bz_s;
      double e_s = ((S_K - u_K) * e - Pt * u_K + Pt_star * S_M + Bn * (vdotB - vBs)) / (S_K - S_M);
      Us[DDD] = rhostar;
      Us[SRR] = rhostar * vx_s;
      Us[LLL] = rhostar * vy_s;
      Us[SZZ] = rhostar * vz_s;
      Us[TAU] = e_s;
      Us[BRR] = Bx_s;
      Us[BPP] = By_s;
      Us[BZZ] = Bz_s;
   }

   // get_double_star (hlld.c:195-280)
   static DHD void hd_double_star(const double *UsL, const double *UsR, double *Uss,
                                  const double *vel, const double *n, int LR)
   {
      double rho_L = UsL[DDD], rho_R = UsR[DDD];
      double rhostar_K = LR == 0 ? rho_L : rho_R;
      double rrh_L = sqrt(rho_L), rrh_R = sqrt(rho_R), rrh_K = sqrt(rhostar_K);
      double vx_L = UsL[SRR] / rho_L, vy_L = UsL[LLL] / rho_L, vz_L = UsL[SZZ] / rho_L;
      double vx_R = UsR[SRR] / rho_R, vy_R = UsR[LLL] / rho_R, vz_R = UsR[SZZ] / rho_R;
      double Bx_L = UsL[BRR], By_L = UsL[BPP], Bz_L = UsL[BZZ];
      double Bx_R = UsR[BRR], By_R = UsR[BPP], Bz_R = UsR[BZZ];
      double Bn_L = Bx_L * n[0] + By_L * n[1] + Bz_L * n[2];
      double Bn_R = Bx_R * n[0] + By_R * n[1] + Bz_R * n[2];
      double Bn = .5 * (Bn_L + Bn_R);
      double sB = Bn > 0.0 ? 1.0 : -1.0;
      double den = rrh_L + rrh_R;
      double vx = (rrh_L * vx_L + rrh_R * vx_R + (Bx_R - Bx_L) * sB) / den;
      double vy = (rrh_L * vy_L + rrh_R * vy_R + (By_R - By_L) * sB) / den;
      double vz = (rrh_L * vz_L + rrh_R * vz_R + (Bz_R - Bz_L) * sB) / den;
      double vn = vx * n[0] + vy * n[1] + vz * n[2];
      vx += (vel[2] - vn) * n[0];
      vy += (vel[2] - vn) * n[1];
      vz += (vel[2] - vn) * n[2];
      double Bx = (rrh_L * Bx_R + rrh_R * Bx_L + rrh_L * rrh_R * (vx_R - vx_L) * sB) / den;
      double By = (rrh_L * By_R + rrh_R * By_L + rrh_L * rrh_R * (vy_R - vy_L) * sB) / den;
      double Bz = (rrh_L * Bz_R + rrh_R * Bz_L + rrh_L * rrh_R * (vz_R - vz_L) * sB) / den;
      double Bnss = Bx * n[0] + By * n[1] + Bz * n[2];
      Bx += (Bn - Bnss) * n[0];
      By += (Bn - Bnss) * n[1];
      Bz += (Bn - Bnss) * n[2];
      double vB_ss = vx * Bx + vy * By + vz * Bz;
      double vBstar, e_star;
      if (LR == 0) {
         vBstar = vx_L * Bx_L + vy_L * By_L + vz_L * Bz_L;
         e_star = UsL[TAU];
      } else {
         vBstar = vx_R * Bx_R + vy_R * By_R + vz_R * Bz_R;
         e_star = UsR[TAU];
      }
      double pm = LR == 1 ? 1.0 : -1.0;
      Uss[DDD] = rhostar_K;
      Uss[SRR] = rhostar_K * vx;
      Uss[LLL] = rhostar_K * vy;
      Uss[SZZ] = rhostar_K * vz;
      Uss[TAU] = e_star + pm * rrh_K * (vBstar - vB_ss) * sB;
      Uss[BRR] = Bx;
      Uss[BPP] = By;
      Uss[BZZ] = Bz;
   }

   // get_Ustar_HLLD (hlld.c:282-401): flux and fan state at speed w.  hp is
   // the phi scale factor (r) at the face.
   static DHD void riemann(const dh_params &P, const double *primL, const double *primR,
                           double hp, double /*z*/, const double *n, double w, double *F,
                           double *U)
   {
      double gam = P.gamma_law;
      double pL[NQ], pR[NQ];
#pragma unroll
      for (int q = 0; q < NQ; ++q) { pL[q] = primL[q]; pR[q] = primR[q]; }
      pL[UPP] *= hp;
      pR[UPP] *= hp;
      double vel[5], Pt_star;
      hd_velocities(gam, pL, pR, n, vel, &Pt_star);
      if (w < vel[2]) {
         double UL[8];
         hd_flux(gam, pL, F, n);
         hd_cons(gam, pL, UL);
         if (w < vel[0]) {
            hd_cons(gam, pL, U);
         } else if (w < vel[1]) {
            hd_single_star(gam, pL, U, vel[0], vel[2], Pt_star, n);
#pragma unroll
            for (int q = 0; q < 8; ++q) F[q] += vel[0] * (U[q] - UL[q]);
         } else {
            double UsL[8], UsR[8];
            hd_single_star(gam, pL, UsL, vel[0], vel[2], Pt_star, n);
            hd_single_star(gam, pR, UsR, vel[4], vel[2], Pt_star, n);
            hd_double_star(UsL, UsR, U, vel, n, 0);
#pragma unroll
            for (int q = 0; q < 8; ++q) F[q] += vel[1] * (U[q] - UsL[q]) + vel[0] * (UsL[q] - UL[q]);
         }
#pragma unroll
         for (int q = 8; q < NQ; ++q) { U[q] = pL[q] * U[DDD]; F[q] = pL[q] * F[DDD]; }
      } else {
         double UR[8];
         hd_flux(gam, pR, F, n);
         hd_cons(gam, pR, UR);
         if (w > vel[4]) {
            hd_cons(gam, pR, U);
         } else if (w > vel[3]) {
            hd_single_star(gam, pR, U, vel[4], vel[2], Pt_star, n);
#pragma unroll
            for (int q = 0; q < 8; ++q) F[q] += vel[4] * (U[q] - UR[q]);
         } else {
            double UsL[8], UsR[8];
            hd_single_star(gam, pL, UsL, vel[0], vel[2], Pt_star, n);
            hd_single_star(gam, pR, UsR, vel[4], vel[2], Pt_star, n);
            hd_double_star(UsL, UsR, U, vel, n, 1);
#pragma unroll
            for (int q = 0; q < 8; ++q) F[q] += vel[3] * (U[q] - UsR[q]) + vel[4] * (UsR[q] - UR[q]);
         }
#pragma unroll
         for (int q = 8; q < NQ; ++q) { U[q] = pR[q] * U[DDD]; F[q] = pR[q] * F[DDD]; }
      }
      double brfac = 1.0 / hp;  // bfield_scale_factor(r, 0) with r = hp
      F[LLL] *= hp;
      F[BRR] *= brfac;
      F[BPP] *= 1.0 / hp;
      U[LLL] *= hp;
      U[BRR] *= brfac;
      U[BPP] *= 1.0 / hp;
   }

   // flux_to_E (mhd.c:293-316): EMF and normal field of a face solve, in
   // the order {E1, B1, E2, B2}; dim 0 phi, 1 r, 2 z.
   static DHD void flux_to_E(const double *F, const double *U, double r, int dim, double *emf)
   {
      double irfac = 1.0 / (1.0 / r);
      if (dim == 0) {
         emf[0] = F[BRR] * irfac;
         emf[1] = U[BRR] * irfac;
         emf[2] = -F[BZZ];
         emf[3] = U[BZZ];
      } else if (dim == 1) {
         emf[0] = -F[BPP] * r;
         emf[1] = U[BRR] * irfac;
         emf[2] = F[BZZ];
         emf[3] = 0.0;
      } else {
         emf[0] = F[BPP] * r;
         emf[1] = U[BZZ];
         emf[2] = -F[BRR] * irfac;
         emf[3] = 0.0;
      }
   }
};

}  // namespace dh
