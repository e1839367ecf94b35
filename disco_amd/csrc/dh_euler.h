// dh_euler.h -- Newtonian hydrodynamics on the moving cylindrical mesh
// (device side of the reference Hydro/euler.c).
//
// Primitive layout {rho, P, v_r, omega, v_z, passives...}; conserved layout
// {D, tau, S_r, L = r S_phi, S_z, passives...} (paul.h:4-5).  The total
// energy is measured in the frame rotating with get_om(r) (euler.c:45-56):
// this is the orbital-velocity subtraction that keeps Keplerian disks
// accurate, and it is why everything here runs in FP64.
#pragma once
#include "dh_geom.h"

namespace dh {

template <int NQ>
struct Euler {
   static constexpr int NC = 5;

   // prim2cons (euler.c:36-62)
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
      double v2 = vr * vr + vp_off * vp_off + vz * vz;
      double rhoe = Pp / (P.gamma_law - 1.);
      cons[DDD] = rho * dV;
      cons[TAU] = (.5 * rho * v2 + rhoe) * dV;
      cons[SRR] = rho * vr * dV;
      cons[LLL] = r * rho * vp * dV;
      cons[SZZ] = rho * vz * dV;
#pragma unroll
      for (int q = NC; q < NQ; ++q) cons[q] = prim[q] * cons[DDD];
   }

   // getUstar (euler.c:64-100): HLLC star state on side k.
   static DHD void getUstar(const dh_params &P, const double *prim, double *Us,
                            double r, double Sk, double Ss, const double *n)
   {
      double rho = prim[RHO];
      double vr = prim[URR];
      double vp = prim[UPP] * r;
      double vz = prim[UZZ];
      double Pp = prim[PPP];
      double om = get_om(P, r);
      double vp_off = vp - om * r;
      double v2 = vr * vr + vp_off * vp_off + vz * vz;
      double vn = vr * n[0] + vp * n[1] + vz * n[2];
      double vn_off = vn - om * r * n[1];
      double Ss_off = Ss + vn_off - vn;
      double rhoe = Pp / (P.gamma_law - 1.);
      double rhostar = rho * (Sk - vn) / (Sk - Ss);
      double Pstar = Pp * (Ss - vn) / (Sk - Ss);
      double Ue = rhoe * (Sk - vn) / (Sk - Ss);
      Us[DDD] = rhostar;
      Us[SRR] = rhostar * (vr + (Ss - vn) * n[0]);
      Us[LLL] = r * rhostar * (vp + (Ss - vn) * n[1]);
      Us[SZZ] = rhostar * (vz + (Ss - vn) * n[2]);
      Us[TAU] = .5 * rhostar * v2 + Ue + rhostar * Ss_off * (Ss - vn) + Pstar;
#pragma unroll
      for (int q = NC; q < NQ; ++q) Us[q] = prim[q] * Us[DDD];
   }

   // cons2prim (euler.c:102-139), with density/pressure floors and the
   // isothermal override.
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
      double KE = .5 * (Sr * vr + rho * vp_off * vp_off + Sz * vz);
      double rhoe = E - KE;
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
#pragma unroll
      for (int q = NC; q < NQ; ++q) prim[q] = cons[q] / cons[DDD];
   }

   // flux (euler.c:141-169) through a face of unit normal n.
   static DHD void flux(const dh_params &P, const double *prim, double *F, double r,
                        const double *n)
   {
      double rho = prim[RHO];
      double Pp = prim[PPP];
      double vr = prim[URR];
      double vp = prim[UPP] * r;
      double vz = prim[UZZ];
      double om = get_om(P, r);
      double vn = vr * n[0] + vp * n[1] + vz * n[2];
      double wn = om * r * n[1];
      double vp_off = vp - om * r;
      double rhoe = Pp / (P.gamma_law - 1.);
      double v2 = vr * vr + vp_off * vp_off + vz * vz;
      F[DDD] = rho * vn;
      F[SRR] = rho * vr * vn + Pp * n[0];
      F[LLL] = r * (rho * vp * vn + Pp * n[1]);
      F[SZZ] = rho * vz * vn + Pp * n[2];
      F[TAU] = (.5 * rho * v2 + rhoe + Pp) * vn - Pp * wn;
#pragma unroll
      for (int q = NC; q < NQ; ++q) F[q] = prim[q] * F[DDD];
   }

   // vel (euler.c:297-330): Davis wave-speed bounds and the HLLC contact speed.
   static DHD void vel(const dh_params &P, const double *p1, const double *p2,
                       double *Sl, double *Sr, double *Ss, const double *n, double r)
   {
      double g = P.gamma_law;
      double P1 = p1[PPP];
      double rho1 = p1[RHO];
      double vn1 = p1[URR] * n[0] + p1[UPP] * n[1] * r + p1[UZZ] * n[2];
      double cs1 = sqrt(g * P1 / rho1);
      double P2 = p2[PPP];
      double rho2 = p2[RHO];
      double vn2 = p2[URR] * n[0] + p2[UPP] * n[1] * r + p2[UZZ] * n[2];
      double cs2 = sqrt(g * P2 / rho2);
      *Ss = (P2 - P1 + rho1 * vn1 * (-cs1) - rho2 * vn2 * cs2) / (rho1 * (-cs1) - rho2 * cs2);
      *Sr = cs1 + vn1;
      *Sl = -cs1 + vn1;
      if (*Sr < cs2 + vn2) *Sr = cs2 + vn2;
      if (*Sl > -cs2 + vn2) *Sl = -cs2 + vn2;
   }

   // Kinematic viscosity: constant or alpha (euler.c:219-227).
   static DHD double nu(const dh_params &P, const double *prim, double r)
   {
      double v = P.viscosity;
      if (P.alpha_flag) {
         double c = sqrt(P.gamma_law * prim[PPP] / prim[RHO]);
         double h = c * pow(r, 1.5);
         v = P.viscosity * c * h;
      }
      return v;
   }

   // visc_flux (euler.c:217-259)
   static DHD void visc_flux(const dh_params &P, const double *prim, const double *gr,
                             const double *gp, const double *gz, double *F, double r,
                             const double *n)
   {
      double vnu = nu(P, prim, r);
      double rho = prim[RHO];
      double vr = prim[URR];
      double om_off = prim[UPP] - get_om(P, r);
      double vz = prim[UZZ];
      double divV_o_d = (gr[URR] + gp[UPP] + gz[UZZ] + vr / r) / 3.0;
      double srr = gr[URR] - divV_o_d;
      double spp = r * (r * gp[UPP] + vr - r * divV_o_d);
      double szz = gz[UZZ] - divV_o_d;
      double srp = 0.5 * (r * r * gr[UPP] + gp[URR]);
      double srz = 0.5 * (gr[UZZ] + gz[URR]);
      double spz = 0.5 * (gp[UZZ] + r * r * gz[UPP]);
      double nc0 = n[0], nc1 = n[1] / r, nc2 = n[2];
      double srn = srr * nc0 + srp * nc1 + srz * nc2;
      double spn = srp * nc0 + spp * nc1 + spz * nc2;
      double szn = srz * nc0 + spz * nc1 + szz * nc2;
      F[SRR] = -2 * vnu * rho * srn;
      F[LLL] = -2 * vnu * rho * spn;
      F[SZZ] = -2 * vnu * rho * szn;
      F[TAU] = -2 * vnu * rho * (vr * srn + om_off * spn + vz * szn);
   }

   // source (euler.c:171-215): geometric (centrifugal + pressure) and
   // energy-frame terms.  polar = NoBC_Rmin (euler.c:23-24).
   static DHD void source(const dh_params &P, const double *prim, double *cons,
                          double rp, double rm, double phip, double phim, double zp,
                          double zm, double dVdt, int polar, double Pmax)
   {
      double rho = prim[RHO];
      double Pp = prim[PPP];
      double r = centroid_r(rp, rm);
      double r_1 = .5 * (rp + rm);
      double vr = prim[URR];
      double omega = prim[UPP];
      double vz = prim[UZZ];
      double centrifugal = rho * omega * omega * r;
      if (polar) centrifugal *= polar_adjust(phip, phim, Pmax);
      double press_bal = Pp / r_1;
      cons[SRR] += dVdt * (centrifugal + press_bal);
      double om = get_om(P, r);
      double om_r = get_om1(P, r);
      double om_z = 0.0;  // get_om2 (omega.c:112-114)
      cons[TAU] += dVdt * rho * (r * vr * om * om - r * r * (omega - om) * (vr * om_r + vz * om_z));
      (void)zp; (void)zm;
   }

   // visc_source (euler.c:261-295)
   static DHD void visc_source(const dh_params &P, const double *prim, const double *gr,
                               const double *gp, const double *gz, double *cons,
                               double r, double dVdt)
   {
      double vnu = nu(P, prim, r);
      double rho = prim[RHO];
      double vr = prim[URR];
      double divV_o_d = (gr[URR] + gp[UPP] + gz[UZZ] + vr / r) / 3.0;
      double spp = (r * gp[UPP] + vr - r * divV_o_d) / (r * r * r);
      cons[SRR] += (-2 * r * rho * vnu * spp) * dVdt;
      double srp = 0.5 * (r * r * gr[UPP] + gp[URR]);
      double spz = 0.5 * (gp[UZZ] + r * r * gz[UPP]);
      double om_r = get_om1(P, r);
      double om_z = 0.0;
      cons[TAU] += (2 * rho * vnu * (srp * om_r + spz * om_z)) * dVdt;
   }

   // mindt (euler.c:332-380): CFL time of one cell.
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
      double maxvr = cs + fabs(vr);
      double maxvp = cs + fabs(vp);
      double maxvz = cs + fabs(vz);
      double dL0 = .5 * (rp + rm) * get_dp(phip, phim, Pmax);
      double dL1 = rp - rm;
      double dL2 = zp - zm;
      double dtr = dL1 / maxvr;
      double dtp = dL0 / maxvp;
      double dtz = dL2 / maxvz;
      double dt = dtr;
      if (dt > dtp) dt = dtp;
      if (dt > dtz) dt = dtz;
      if (P.visc_flag) {
         double dx = dL0;
         if (dx > dL1) dx = dL1;
         if (dx > dL2) dx = dL2;
         double dt_visc = 0.5 * dx * dx / nu(P, prim, r);
         if (dt > dt_visc) dt = dt_visc;
      }
      return dt;
   }

   // get_omega (euler.c:32-34): the cell's own angular velocity.
   static DHD double omega(const dh_params &, const double *prim, double, double)
   {
      return prim[UPP];
   }

   // solve_riemann (riemann.c:216-318), HLL/HLLC branch: the interface flux
   // F and the state U* it is associated with (the caller forms F - w U*).
   static DHD void riemann(const dh_params &P, const double *pL, const double *pR,
                           double r, double z, const double *n, double w, double *Flux,
                           double *Ustr)
   {
      double Sl, Sr, Ss;
      vel(P, pL, pR, &Sl, &Sr, &Ss, n, r);
      if (w < Sl) {
         flux(P, pL, Flux, r, n);
         prim2cons(P, pL, Ustr, r, z, 1.0);
      } else if (w > Sr) {
         flux(P, pR, Flux, r, n);
         prim2cons(P, pR, Ustr, r, z, 1.0);
      } else if (P.riemann_solver == RS_HLL) {
         double Fl[NQ], Fr[NQ], Ul[NQ], Ur[NQ];
         double aL = Sr, aR = -Sl;
         prim2cons(P, pL, Ul, r, z, 1.0);
         prim2cons(P, pR, Ur, r, z, 1.0);
         flux(P, pL, Fl, r, n);
         flux(P, pR, Fr, r, n);
#pragma unroll
         for (int q = 0; q < NQ; ++q) {
            Flux[q] = (aL * Fl[q] + aR * Fr[q] + aL * aR * (Ul[q] - Ur[q])) / (aL + aR);
            Ustr[q] = (aR * Ul[q] + aL * Ur[q] + Fl[q] - Fr[q]) / (aL + aR);
         }
      } else {
         double Uk[NQ], Fk[NQ];
         const double *pk = (w < Ss) ? pL : pR;
         double Sk = (w < Ss) ? Sl : Sr;
         prim2cons(P, pk, Uk, r, z, 1.0);
         getUstar(P, pk, Ustr, r, Sk, Ss, n);
         flux(P, pk, Fk, r, n);
#pragma unroll
         for (int q = 0; q < NQ; ++q) Flux[q] = Fk[q] + Sk * (Ustr[q] - Uk[q]);
      }
   }
};

}  // namespace dh
