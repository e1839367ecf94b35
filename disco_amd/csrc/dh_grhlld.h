// dh_grhlld.h -- the relativistic HLLD Riemann solver of the GRMHD module
// (device side of the reference hlld_grmhd.c): the face problem is moved
// into the orthonormal frame of the face's tetrad, solved there with the
// special-relativistic HLLD fan of Mignone, Ugliano & Bodo (2009), and the
// flux and fan state are moved back (hlld_grmhd.c:90-368).
//
// The total pressure of the fan is found by Newton's method from either the
// HLL state's pressure or the B = 0 estimate, with the reference's sequence
// of fallbacks (restarts near the singular pressures of each side); when
// none of them gives a physical fan the HLL state is used
// (hlld_grmhd.c:471-844).  The arithmetic keeps the reference's order so
// that the host build reproduces it bit for bit, including the cons2prim of
// the HLL state (hlld_grmhd.c:1364-1477).
#pragma once
#include "dh_geom.h"

namespace dh {
namespace grh {

// Components of the orthonormal-frame state (hlld_grmhd.c:15).
enum { DD = 0, EN = 1, MX = 2, MY = 3, MZ = 4, BX = 5, BY = 6, BZ = 7 };
constexpr double ACC = 1e-12;
constexpr int MAX_ITER = 10;

// calc_tetrads (hlld_grmhd.c:370-432): e[4 i + mu] = e_(i)^mu and
// ie[4 i + mu] = e^(i)_mu of the frame whose x axis is coordinate xx.
// sqrtgam is the full sqrt(det gamma).  The zz entry of e^(x) reads e_(x)
// one slot lower than the other two entries, as the reference does.
DHD void tetrads(double al, const double *be, const double *gam, const double *igam,
                 double sqrtgam, int xx, double *e, double *ie)
{
   int yy = (xx + 1) % 3, zz = (xx + 2) % 3;
   double ial = 1.0 / al;
   double xfac = 1.0 / sqrt(igam[3 * xx + xx]);
   e[0] = ial;
   e[xx + 1] = -be[xx] * ial;
   e[yy + 1] = -be[yy] * ial;
   e[zz + 1] = -be[zz] * ial;
   e[4] = 0.0;
   e[4 + xx + 1] = igam[3 * xx + xx] * xfac;
   e[4 + yy + 1] = igam[3 * xx + yy] * xfac;
   e[4 + zz + 1] = igam[3 * xx + zz] * xfac;
   e[8] = 0.0;
   e[8 + xx + 1] = 0.0;
   e[8 + yy + 1] = sqrt(gam[3 * zz + zz] / (igam[3 * xx + xx] * sqrtgam * sqrtgam));
   e[8 + zz + 1] = -gam[3 * yy + zz] / (sqrtgam * sqrt(igam[3 * xx + xx] * gam[3 * zz + zz]));
   e[12] = 0.0;
   e[12 + xx + 1] = 0.0;
   e[12 + yy + 1] = 0.0;
   e[12 + zz + 1] = 1.0 / sqrt(gam[3 * zz + zz]);

   ie[0] = al;
   ie[xx + 1] = 0.0;
   ie[yy + 1] = 0.0;
   ie[zz + 1] = 0.0;
   ie[4] = 0.0;
   ie[4 + xx + 1] = gam[3 * xx + xx] * e[4 + xx + 1] + gam[3 * xx + yy] * e[4 + yy + 1] +
                    gam[3 * xx + zz] * e[4 + zz + 1];
   ie[4 + yy + 1] = gam[3 * yy + xx] * e[4 + xx + 1] + gam[3 * yy + yy] * e[4 + yy + 1] +
                    gam[3 * yy + zz] * e[4 + zz + 1];
   ie[4 + zz + 1] = gam[3 * zz + xx] * e[4 + xx] + gam[3 * zz + yy] * e[4 + yy] +
                    gam[3 * zz + zz] * e[4 + zz];
   ie[8] = 0.0;
   ie[8 + xx + 1] = 0.0;
   ie[8 + yy + 1] = gam[3 * yy + yy] * e[8 + yy + 1] + gam[3 * yy + zz] * e[8 + zz + 1];
   ie[8 + zz + 1] = gam[3 * zz + yy] * e[8 + yy + 1] + gam[3 * zz + zz] * e[8 + zz + 1];
   ie[12] = 0.0;
   ie[12 + xx + 1] = 0.0;
   ie[12 + yy + 1] = 0.0;
   ie[12 + zz + 1] = gam[3 * zz + zz] * e[12 + zz + 1];
}

// calc_UF (hlld_grmhd.c:434-469): conserved state and x flux of a fluid
// with four-velocity u and field B in the orthonormal frame.
DHD void state_UF(double gl, double rho, double P, const double *u, const double *B, double *U,
                  double *F)
{
   double rhoh = rho + gl / (gl - 1) * P;
   double b[4];
   b[0] = u[1] * B[0] + u[2] * B[1] + u[3] * B[2];
   b[1] = (B[0] + b[0] * u[1]) / u[0];
   b[2] = (B[1] + b[0] * u[2]) / u[0];
   b[3] = (B[2] + b[0] * u[3]) / u[0];
   double b2 = -b[0] * b[0] + b[1] * b[1] + b[2] * b[2] + b[3] * b[3];
   U[DD] = rho * u[0];
   U[EN] = (rhoh + b2) * u[0] * u[0] - (P + 0.5 * b2) - b[0] * b[0];
   U[MX] = (rhoh + b2) * u[0] * u[1] - b[0] * b[1];
   U[MY] = (rhoh + b2) * u[0] * u[2] - b[0] * b[2];
   U[MZ] = (rhoh + b2) * u[0] * u[3] - b[0] * b[3];
   U[BX] = B[0];
   U[BY] = B[1];
   U[BZ] = B[2];
   F[DD] = rho * u[1];
   F[EN] = U[MX];
   F[MX] = (rhoh + b2) * u[1] * u[1] + (P + 0.5 * b2) - b[1] * b[1];
   F[MY] = (rhoh + b2) * u[2] * u[1] - b[2] * b[1];
   F[MZ] = (rhoh + b2) * u[3] * u[1] - b[3] * b[1];
   F[BX] = 0.0;
   F[BY] = (B[1] * u[1] - B[0] * u[2]) / u[0];
   F[BZ] = (B[2] * u[1] - B[0] * u[3]) / u[0];
}

// Outer fast speeds of calc_msfast (grmhd2.c:2263-2345), whose exact
// dispersion solve is switched off in the reference, leaving the estimate
// from the combined sound and Alfven speeds.
DHD void fast_speeds(double cs2, double w, double vi, double b2, double rhoh, double al,
                     double bei, double igamii, double *vL, double *vR)
{
   double viS = (vi + bei) / al;
   double v2 = 1.0 - 1.0 / (w * w);
   double va2 = b2 / (rhoh + b2);
   double cf2 = cs2 + va2 - cs2 * va2;
   double dv = sqrt(cf2 * (igamii - viS * viS - cf2 * (igamii * v2 - viS * viS))) / w;
   *vL = al * (viS * (1.0 - cf2) - dv) / (1.0 - v2 * cf2) - bei;
   *vR = al * (viS * (1.0 - cf2) + dv) / (1.0 - v2 * cf2) - bei;
}

// The fan state next to one outer wave of speed s at total pressure p: the
// velocity va, field Ba, total enthalpy w_a and K_a (MUB eqs. 21-43), with
// their p-derivatives when D is set (hlld_grmhd.c:877-972).
struct Side {
   double va[3], Ba[3], enth, Ka[3];
   double dva[3], dBa[3], denth, dKa[3];
};

template <bool D>
DHD void side_state(double p, double s, double Bx, const double *R, int LR, Side &a)
{
   // calc_va
   double A = R[MX] - s * R[EN] + p * (1.0 - s * s);
   double G = R[BY] * R[BY] + R[BZ] * R[BZ];
   double C = R[MY] * R[BY] + R[MZ] * R[BZ];
   double Q = -A - G + Bx * Bx * (1.0 - s * s);
   double iX = 1.0 / (Bx * (s * A * Bx + C) - (A + G) * (s * p + R[EN]));
   a.va[0] = iX * (Bx * (A * Bx + s * C) - (A + G) * (p + R[MX]));
   a.va[1] = iX * (Q * R[MY] + R[BY] * (C + Bx * (s * R[MX] - R[EN])));
   a.va[2] = iX * (Q * R[MZ] + R[BZ] * (C + Bx * (s * R[MX] - R[EN])));
   if (D) {
      double dA = 1.0 - s * s;
      double dQ = -dA;
      double diX = -iX * iX * (s * Bx * Bx * dA - s * (A + G) - dA * (s * p + R[EN]));
      a.dva[0] = diX * (Bx * (A * Bx + s * C) - (A + G) * (p + R[MX])) +
                 iX * (Bx * Bx * dA - dA * (p + R[MX]) - (A + G));
      a.dva[1] = diX * (Q * R[MY] + R[BY] * (C + Bx * (s * R[MX] - R[EN]))) + iX * (dQ * R[MY]);
      a.dva[2] = diX * (Q * R[MZ] + R[BZ] * (C + Bx * (s * R[MX] - R[EN]))) + iX * (dQ * R[MZ]);
   }
   // calc_Ba
   a.Ba[0] = Bx;
   a.Ba[1] = (R[BY] - Bx * a.va[1]) / (s - a.va[0]);
   a.Ba[2] = (R[BZ] - Bx * a.va[2]) / (s - a.va[0]);
   if (D) {
      double ids2 = 1.0 / ((s - a.va[0]) * (s - a.va[0]));
      a.dBa[0] = 0.0;
      a.dBa[1] = (-Bx * a.dva[1] * (s - a.va[0]) + (R[BY] - Bx * a.va[1]) * a.dva[0]) * ids2;
      a.dBa[2] = (-Bx * a.dva[2] * (s - a.va[0]) + (R[BZ] - Bx * a.va[2]) * a.dva[0]) * ids2;
   }
   // calc_entha
   a.enth = p + (R[EN] - a.va[0] * R[MX] - a.va[1] * R[MY] - a.va[2] * R[MZ]) / (s - a.va[0]);
   if (D)
      a.denth = 1.0 + ((-a.dva[0] * R[MX] - a.dva[1] * R[MY] - a.dva[2] * R[MZ]) * (s - a.va[0]) +
                       (R[EN] - a.va[0] * R[MX] - a.va[1] * R[MY] - a.va[2] * R[MZ]) * a.dva[0]) /
                      ((s - a.va[0]) * (s - a.va[0]));
   // calc_Ka
   int sgn = Bx > 0 ? LR : -LR;
   double eta = sgn * sqrt(a.enth);
   double denom = 1.0 / (s * p + R[EN] + Bx * eta);
   a.Ka[0] = (R[MX] + p + R[BX] * eta) * denom;
   a.Ka[1] = (R[MY] + R[BY] * eta) * denom;
   a.Ka[2] = (R[MZ] + R[BZ] * eta) * denom;
   if (D) {
      double deta = 0.5 / eta * a.denth;
      double ddenom = -denom * denom * (s + Bx * deta);
      a.dKa[0] = (1.0 + R[BX] * deta) * denom + (R[MX] + p + R[BX] * eta) * ddenom;
      a.dKa[1] = (R[BY] * deta) * denom + (R[MY] + R[BY] * eta) * ddenom;
      a.dKa[2] = (R[BZ] * deta) * denom + (R[MZ] + R[BZ] * eta) * ddenom;
   }
}

// cont_func (MUB eq. 48, hlld_grmhd.c:974-1051): the jump condition across
// the contact whose root in p is the fan's total pressure, and its slope.
DHD void contact_residual(double Bx, const Side &L, const Side &R, double *f, double *df)
{
   double saL = L.Ka[0], saR = R.Ka[0];
   double dsaL = L.dKa[0], dsaR = R.dKa[0];
   double Bch[3];
   Bch[0] = (saR - saL) * Bx;
   Bch[1] = (R.Ba[1] * (saR - R.va[0]) + Bx * R.va[1]) - (L.Ba[1] * (saL - L.va[0]) + Bx * L.va[1]);
   Bch[2] = (R.Ba[2] * (saR - R.va[0]) + Bx * R.va[2]) - (L.Ba[2] * (saL - L.va[0]) + Bx * L.va[2]);
   double KaL2 = L.Ka[0] * L.Ka[0] + L.Ka[1] * L.Ka[1] + L.Ka[2] * L.Ka[2];
   double KaR2 = R.Ka[0] * R.Ka[0] + R.Ka[1] * R.Ka[1] + R.Ka[2] * R.Ka[2];
   double KBL = L.Ka[0] * Bch[0] + L.Ka[1] * Bch[1] + L.Ka[2] * Bch[2];
   double KBR = R.Ka[0] * Bch[0] + R.Ka[1] * Bch[1] + R.Ka[2] * Bch[2];
   int sgn = Bx > 0 ? 1 : -1;
   double etaL = -sgn * sqrt(L.enth);
   double etaR = sgn * sqrt(R.enth);
   double delKx = (R.Ka[0] - L.Ka[0]);
   double YL = (1.0 - KaL2) / (etaL * delKx - KBL);
   double YR = (1.0 - KaR2) / (etaR * delKx - KBR);
   *f = (R.Ka[0] - L.Ka[0]) * (1.0 - Bx * (YR - YL));
   double dBch[3];
   dBch[0] = (dsaR - dsaL) * Bx;
   dBch[1] = (R.dBa[1] * (saR - R.va[0]) + R.Ba[1] * (dsaR - R.dva[0]) + Bx * R.dva[1]) -
             (L.dBa[1] * (saL - L.va[0]) + L.Ba[1] * (dsaL - L.dva[0]) + Bx * L.dva[1]);
   dBch[2] = (R.dBa[2] * (saR - R.va[0]) + R.Ba[2] * (dsaR - R.dva[0]) + Bx * R.dva[2]) -
             (L.dBa[2] * (saL - L.va[0]) + L.Ba[2] * (dsaL - L.dva[0]) + Bx * L.dva[2]);
   double dKaL2 = 2 * (L.Ka[0] * L.dKa[0] + L.Ka[1] * L.dKa[1] + L.Ka[2] * L.dKa[2]);
   double dKaR2 = 2 * (R.Ka[0] * R.dKa[0] + R.Ka[1] * R.dKa[1] + R.Ka[2] * R.dKa[2]);
   double dKBL = L.dKa[0] * Bch[0] + L.Ka[0] * dBch[0] + L.dKa[1] * Bch[1] + L.Ka[1] * dBch[1] +
                 L.dKa[2] * Bch[2] + L.Ka[2] * dBch[2];
   double dKBR = R.dKa[0] * Bch[0] + R.Ka[0] * dBch[0] + R.dKa[1] * Bch[1] + R.Ka[1] * dBch[1] +
                 R.dKa[2] * Bch[2] + R.Ka[2] * dBch[2];
   double detaL = 0.5 / etaL * L.denth;
   double detaR = 0.5 / etaR * R.denth;
   double ddelKx = R.dKa[0] - L.dKa[0];
   double dYL = (-dKaL2 * (etaL * delKx - KBL) - (1 - KaL2) * (detaL * delKx + etaL * ddelKx - dKBL)) /
                ((etaL * delKx - KBL) * (etaL * delKx - KBL));
   double dYR = (-dKaR2 * (etaR * delKx - KBR) - (1 - KaR2) * (detaR * delKx + etaR * ddelKx - dKBR)) /
                ((etaR * delKx - KBR) * (etaR * delKx - KBR));
   *df = ddelKx * (1.0 - Bx * (YR - YL)) + delKx * (-Bx * (dYR - dYL));
}

// calc_Ps (hlld_grmhd.c:1092-1179): Newton's method on the contact
// condition from the guess *Ps.
DHD void newton_pressure(const double *RL, const double *RR, double Bx, double sL, double sR,
                         double *Ps)
{
   double p1 = *Ps, p, dp;
   int i = 0;
   do {
      p = p1;
      Side L, R;
      side_state<true>(p, sL, Bx, RL, -1, L);
      side_state<true>(p, sR, Bx, RR, +1, R);
      double f, df;
      contact_residual(Bx, L, R, &f, &df);
      dp = -f / df;
      p1 = p + dp;
      if (p1 < 0.0) p1 = 0.5 * p;
      if (L.enth < 0.0 || R.enth < 0.0) p1 = 0.8 * p;
      i++;
   } while (fabs(dp) / p > ACC && i < MAX_ITER);
   *Ps = p1;
}

// The whole fan at total pressure p (calc_state, hlld_grmhd.c:1181-1201):
// both outer sides, the contact field Bc and the contact velocities.
struct Fan {
   Side L, R;
   double Bc[3], vcL[3], vcR[3];
};

DHD void contact_velocity(double Bx, const double *Ka, const double *Bc, double enth, int LR,
                          double *vc)
{
   int sgn = Bx > 0 ? LR : -LR;
   double eta = sgn * sqrt(enth);
   double K2 = Ka[0] * Ka[0] + Ka[1] * Ka[1] + Ka[2] * Ka[2];
   double KB = Ka[0] * Bx + Ka[1] * Bc[1] + Ka[2] * Bc[2];
   vc[0] = Ka[0] - Bx * (1.0 - K2) / (eta - KB);
   vc[1] = Ka[1] - Bc[1] * (1.0 - K2) / (eta - KB);
   vc[2] = Ka[2] - Bc[2] * (1.0 - K2) / (eta - KB);
}

DHD void fan_at(double p, const double *RL, const double *RR, double Bx, double sL, double sR,
                Fan &fn)
{
   side_state<false>(p, sL, Bx, RL, -1, fn.L);
   side_state<false>(p, sR, Bx, RR, +1, fn.R);
   double saL = fn.L.Ka[0], saR = fn.R.Ka[0];
   fn.Bc[0] = Bx;
   for (int k = 1; k < 3; ++k) {
      fn.Bc[k] = ((fn.R.Ba[k] * (saR - fn.R.va[0]) + Bx * fn.R.va[k]) -
                  (fn.L.Ba[k] * (saL - fn.L.va[0]) + Bx * fn.L.va[k])) /
                 (saR - saL);
      if (fn.Bc[k] != fn.Bc[k]) fn.Bc[k] = 0.0;
   }
   contact_velocity(Bx, fn.L.Ka, fn.Bc, fn.L.enth, -1, fn.vcL);
   contact_velocity(Bx, fn.R.Ka, fn.Bc, fn.R.enth, +1, fn.vcR);
}

// checkSolution (hlld_grmhd.c:846-875): a physical fan, ordered in speed.
DHD bool fan_ok(double Ps, double sL, double sR, const Fan &fn)
{
   const double *vaL = fn.L.va, *vaR = fn.R.va, *vcL = fn.vcL, *vcR = fn.vcR;
   bool ok = true;
   if (Ps != Ps) ok = false;
   if (fn.L.enth < Ps || fn.R.enth < Ps) ok = false;
   if (fabs(vaL[0]) > 1.0 || fabs(vaL[1]) > 1.0 || fabs(vaL[2]) > 1.0) ok = false;
   if (fabs(vaR[0]) > 1.0 || fabs(vaR[1]) > 1.0 || fabs(vaR[2]) > 1.0) ok = false;
   if (vaL[0] < sL || vaR[0] > sR) ok = false;
   if (fabs(vcL[0]) > 1.0 || fabs(vcL[1]) > 1.0 || fabs(vcL[2]) > 1.0) ok = false;
   if (fabs(vcR[0]) > 1.0 || fabs(vcR[1]) > 1.0 || fabs(vcR[2]) > 1.0) ok = false;
   if (vcL[0] < sL || vcR[0] > sR) ok = false;
   if (fn.L.Ka[0] < sL || fn.R.Ka[0] > sR) ok = false;
   return ok;
}

// One Newton solve from the guess Ps and the fan it gives.
DHD bool try_pressure(double *Ps, const double *RL, const double *RR, double Bx, double sL,
                      double sR, Fan &fn)
{
   newton_pressure(RL, RR, Bx, sL, sR, Ps);
   fan_at(*Ps, RL, RR, Bx, sL, sR, fn);
   return fan_ok(*Ps, sL, sR, fn);
}

// The third and fourth attempts (hlld_grmhd.c:538-666): restart just above
// and below the two pressures at which side R's wave becomes singular.
// first_sign is -1 (side L: roots in the order -, +) or +1 (side R: +, -).
DHD bool try_singular(double *Ps, const double *RS, double s, int first_sign,
                      const double *RL, const double *RR, double Bx, double sL, double sR, Fan &fn)
{
   double Ap = RS[MX] - s * RS[EN];
   double C = RS[MY] * RS[BY] + RS[MZ] * RS[BZ];
   double G = RS[BY] * RS[BY] + RS[BZ] * RS[BZ];
   double a = -s * (1.0 - s * s);
   double b = (s * Bx * Bx - RS[EN]) * (1.0 - s * s) - s * (Ap + G);
   double c = Bx * (s * Bx * Ap + C) - (Ap + G) * RS[EN];
   double desc = b * b - 4 * a * c;
   if (desc < 0) return false;
   double Psing1 = first_sign < 0 ? (-b - sqrt(b * b - 4 * a * c)) / (2 * a)
                                  : (-b + sqrt(b * b - 4 * a * c)) / (2 * a);
   double Psing2 = first_sign < 0 ? (-b + sqrt(b * b - 4 * a * c)) / (2 * a)
                                  : (-b - sqrt(b * b - 4 * a * c)) / (2 * a);
   double del = 1.0e-1;
   bool ok = false;
   if (Psing1 > 0.0) {
      *Ps = Psing1 * (1.0 + del);
      ok = try_pressure(Ps, RL, RR, Bx, sL, sR, fn);
      if (!ok) {
         *Ps = Psing1 * (1.0 - del);
         ok = try_pressure(Ps, RL, RR, Bx, sL, sR, fn);
      }
   }
   if (!ok && Psing2 > 0.0) {
      *Ps = Psing2 * (1.0 + del);
      ok = try_pressure(Ps, RL, RR, Bx, sL, sR, fn);
      if (!ok) {
         *Ps = Psing2 * (1.0 - del);
         ok = try_pressure(Ps, RL, RR, Bx, sL, sR, fn);
      }
   }
   return ok;
}

// cons2prim of the HLL state (hlld_grmhd.c:1364-1477): the same 2-D Newton
// scheme as the module's, in flat space, from the guess p (rho, P, u);
// returns the total pressure.
DHD double hll_total_pressure(double gl, const double *U, double *p)
{
   const double prec = 1.0e-10;
   const double max_iter = 30;
   const double Nextra = 2;
   double D = U[DD];
   double S[3] = {U[MX], U[MY], U[MZ]};
   double tau = U[EN];
   double B[3] = {U[BX], U[BY], U[BZ]};
   double s2 = (S[0] * S[0] + S[1] * S[1] + S[2] * S[2]) / (D * D);
   double B2 = B[0] * B[0] + B[1] * B[1] + B[2] * B[2];
   double BS = B[0] * S[0] + B[1] * S[1] + B[2] * S[2];
   double Q = B2 / D;
   double psi = BS * BS / (D * D * D);
   double e = tau / D;
   double n = (gl - 1.0) / gl;
   double u[3] = {p[URR], p[UPP], p[UZZ]};
   double u2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
   double w = sqrt(1.0 + u2);
   double v21 = u2 / (1.0 + u2);
   double eta1 = w * (1.0 + p[PPP] / (n * p[RHO]));
   double v2, eta;
   int i = 0, clean = -1;
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
      i++;
      if (v21 > 1.0) v21 = 0.5 * (v2 + 1.0);
      if (v21 < 0.0) v21 = 0.5 * v2;
      if (eta1 < 1.0) eta1 = 0.5 * (eta + 1.0);
      double err = (eta1 - eta) / eta;
      if (fabs(err) < prec && clean < 0) clean = (int)Nextra + 1;
      if (clean >= 0) clean--;
      if (clean == 0 || i == max_iter) break;
   }
   v2 = v21;
   eta = eta1;
   w = 1.0 / sqrt(1 - v2);
   double h = eta / w;
   double rho = D / w;
   double Pp = n * rho * (h - 1.0);
   h = 1.0 + gl / (gl - 1.0) * Pp / rho;
   for (int k = 0; k < 3; ++k) u[k] = (S[k] + BS * B[k] / (D * h * w)) / (D * h + B2 / w);
   double uB = u[0] * B[0] + u[1] * B[1] + u[2] * B[2];
   double b2 = (B2 + uB * uB) / (w * w);
   p[RHO] = rho;
   p[URR] = u[0];
   p[UPP] = u[1];
   p[UZZ] = u[2];
   p[PPP] = Pp;
   return Pp + 0.5 * b2;
}

// calc_Ua, calc_Uc (MUB eqs. 32-34, 50-52; hlld_grmhd.c:1203-1234).
DHD void outer_state(double p, double s, double Bx, const double *R, const Side &a, double *Ua)
{
   const double *va = a.va, *Ba = a.Ba;
   double vB = va[0] * Bx + va[1] * Ba[1] + va[2] * Ba[2];
   Ua[DD] = R[DD] / (s - va[0]);
   Ua[EN] = (R[EN] + p * va[0] - vB * Bx) / (s - va[0]);
   Ua[MX] = (Ua[EN] + p) * va[0] - vB * Bx;
   Ua[MY] = (Ua[EN] + p) * va[1] - vB * Ba[1];
   Ua[MZ] = (Ua[EN] + p) * va[2] - vB * Ba[2];
   Ua[BX] = Bx;
   Ua[BY] = Ba[1];
   Ua[BZ] = Ba[2];
}

DHD void contact_state(double p, double sa, double Bx, const double *va, const double *vc,
                       const double *Bc, const double *Ua, double *Uc)
{
   double vB = vc[0] * Bx + vc[1] * Bc[1] + vc[2] * Bc[2];
   Uc[DD] = Ua[DD] * (sa - va[0]) / (sa - vc[0]);
   Uc[EN] = (sa * Ua[EN] - Ua[MX] + p * vc[0] - vB * Bx) / (sa - vc[0]);
   Uc[MX] = (Uc[EN] + p) * vc[0] - vB * Bx;
   Uc[MY] = (Uc[EN] + p) * vc[1] - vB * Bc[1];
   Uc[MZ] = (Uc[EN] + p) * vc[2] - vB * Bc[2];
   Uc[BX] = Bx;
   Uc[BY] = Bc[1];
   Uc[BZ] = Bc[2];
}

// solve_HLLD_SR (hlld_grmhd.c:471-844): flux F and state U at speed w of the
// fan between UL and UR (orthonormal frame); returns the zone of w, 0..5 in
// the HLLD fan or -3..-1 when the HLL state stands in.
DHD int solve_fan(double gl, double sL, double sR, double Bx, const double *UL, const double *FL,
                  const double *UR, const double *FR, double w, const double *pL,
                  const double *pR, double *U, double *F)
{
   double RL[8], RR[8], Uhll[8], Fhll[8];
   for (int q = 0; q < 8; ++q) {
      RL[q] = sL * UL[q] - FL[q];
      RR[q] = sR * UR[q] - FR[q];
      Uhll[q] = (sR * UR[q] - sL * UL[q] + FL[q] - FR[q]) / (sR - sL);
      Fhll[q] = (sR * FL[q] - sL * FR[q] + sL * sR * (UR[q] - UL[q])) / (sR - sL);
   }
   double prim[5];
   for (int q = 0; q < 5; ++q) prim[q] = 0.5 * (pL[q] + pR[q]);
   double Pshll = hll_total_pressure(gl, Uhll, prim);
   double b = Uhll[EN] - Fhll[MX];
   double c = Uhll[MX] * Fhll[EN] - Fhll[MX] * Uhll[EN];
   double Ps0 = 0.5 * (-b + sqrt(b * b - 4 * c));
   double Ps = Bx * Bx < 0.1 * Pshll ? Ps0 : Pshll;
   bool mag = Bx * Bx > 1.0e-16 * Ps;
   Fan fn;
   bool ok;
   if (mag) {
      ok = try_pressure(&Ps, RL, RR, Bx, sL, sR, fn);
      if (!ok) {
         Ps = Bx * Bx < 0.1 * Pshll ? Pshll : Ps0;
         ok = try_pressure(&Ps, RL, RR, Bx, sL, sR, fn);
      }
      if (!ok) ok = try_singular(&Ps, RL, sL, -1, RL, RR, Bx, sL, sR, fn);
      if (!ok) ok = try_singular(&Ps, RR, sR, +1, RL, RR, Bx, sL, sR, fn);
   } else {
      fan_at(Ps, RL, RR, Bx, sL, sR, fn);
      ok = fan_ok(Ps, sL, sR, fn);
   }
   double saL = fn.L.Ka[0], saR = fn.R.Ka[0];
   double sc = 0.5 * (fn.vcL[0] + fn.vcR[0]);
   if (ok) {
      if (w < sL) {
         for (int q = 0; q < 8; ++q) { U[q] = UL[q]; F[q] = FL[q]; }
         return 0;
      }
      if (w < saL) {
         outer_state(Ps, sL, Bx, RL, fn.L, U);
         for (int q = 0; q < 8; ++q) F[q] = FL[q] + sL * (U[q] - UL[q]);
         return 1;
      }
      if (w < sc && mag) {
         double UaL[8];
         outer_state(Ps, sL, Bx, RL, fn.L, UaL);
         contact_state(Ps, saL, Bx, fn.L.va, fn.vcL, fn.Bc, UaL, U);
         for (int q = 0; q < 8; ++q) F[q] = FL[q] + sL * (UaL[q] - UL[q]) + saL * (U[q] - UaL[q]);
         return 2;
      }
      if (w < saR && mag) {
         double UaR[8];
         outer_state(Ps, sR, Bx, RR, fn.R, UaR);
         contact_state(Ps, saR, Bx, fn.R.va, fn.vcR, fn.Bc, UaR, U);
         for (int q = 0; q < 8; ++q) F[q] = FR[q] + sR * (UaR[q] - UR[q]) + saR * (U[q] - UaR[q]);
         return 3;
      }
      if (w < sR) {
         outer_state(Ps, sR, Bx, RR, fn.R, U);
         for (int q = 0; q < 8; ++q) F[q] = FR[q] + sR * (U[q] - UR[q]);
         return 4;
      }
      for (int q = 0; q < 8; ++q) { U[q] = UR[q]; F[q] = FR[q]; }
      return 5;
   }
   const double *Us = w < sL ? UL : w < sR ? Uhll : UR;
   const double *Fs = w < sL ? FL : w < sR ? Fhll : FR;
   for (int q = 0; q < 8; ++q) { U[q] = Us[q]; F[q] = Fs[q]; }
   return w < sL ? -3 : w < sR ? -2 : -1;
}

}  // namespace grh
}  // namespace dh
