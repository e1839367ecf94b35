// dh_geom.h -- device geometry of the cylindrical Disco mesh and the
// frame/mesh angular velocities.
//
// Cylindrical geometry follows Geometry/cylindrical.c:10-64 of the reference;
// the mesh and energy-frame angular velocities follow omega.c:43-140.  The
// expressions keep the reference's association so that results agree with
// the CPU path to rounding.
#pragma once
#include <hip/hip_runtime.h>
#include "disco_hip.h"

namespace dh {

enum { RHO = 0, PPP = 1, URR = 2, UPP = 3, UZZ = 4, BRR = 5, BPP = 6, BZZ = 7 };
enum { DDD = 0, TAU = 1, SRR = 2, LLL = 3, SZZ = 4 };
enum { RS_HLL = 0, RS_HLLC = 1, RS_HLLD = 2 };

#define DHD __host__ __device__ __forceinline__

// get_dp (cylindrical.c:10-15): phi difference wrapped into [0, phi_max].
DHD double get_dp(double phip, double phim, double Pmax)
{
   double dp = phip - phim;
   if (!std::isfinite(dp)) return dp;  // a non-finite value would never wrap
   if (fabs(dp) > 64.0 * Pmax) dp = fmod(dp, Pmax);  // bound the loops below
   while (dp < 0.0) dp += Pmax;
   while (dp > Pmax) dp -= Pmax;
   return dp;
}

// get_signed_dp (cylindrical.c:17-22): wrapped into [-phi_max/2, phi_max/2].
DHD double get_sdp(double phip, double phim, double Pmax)
{
   double dp = phip - phim;
   if (!std::isfinite(dp)) return dp;
   if (fabs(dp) > 64.0 * Pmax) dp = fmod(dp, Pmax);
   while (dp < -.5 * Pmax) dp += Pmax;
   while (dp > .5 * Pmax) dp -= Pmax;
   return dp;
}

// The wrapping loops of the reference on a signed difference: into
// [-Pmax/2, Pmax/2] (faces_alt.c:138-139 and friends).
DHD double wrap_half(double w, double Pmax)
{
   if (!std::isfinite(w)) return w;
   if (fabs(w) > 64.0 * Pmax) w = fmod(w, Pmax);
   while (w > .5 * Pmax) w -= Pmax;
   while (w < -.5 * Pmax) w += Pmax;
   return w;
}

// get_centroid (cylindrical.c:24-30): radial centroid of an annular cell.
DHD double centroid_r(double xp, double xm)
{
   return 2.0 * (xp * xp + xp * xm + xm * xm) / (3.0 * (xp + xm));
}
DHD double centroid_z(double xp, double xm) { return 0.5 * (xp + xm); }

// get_dV (cylindrical.c:56-64) for xp = {rp, phip, zp}, xm = {rm, phim, zm}.
DHD double cell_dV(double rp, double rm, double phip, double phim, double zp,
                   double zm, double Pmax)
{
   double r = .5 * (rp + rm);
   double dr = rp - rm;
   double dphi = get_dp(phip, phim, Pmax);
   double dz = zp - zm;
   return r * dr * dphi * dz;
}

// sin(dphi/2)/(dphi/2): geom_polar_vec_adjust (cylindrical.c:175-183).
DHD double polar_adjust(double phip, double phim, double Pmax)
{
   double dphi = get_dp(phip, phim, Pmax);
   return sin(0.5 * dphi) / (0.5 * dphi);
}

// mesh_om (omega.c:43-60)
DHD double mesh_om(const dh_params &P, double r)
{
   switch (P.exact_mesh_omega) {
   case 1: return 1.0;
   case 2: return pow(r, -1.5);
   case 3: {
      double n = 8.0;
      return 1. / pow(pow(r, 1.5 * n) + 1., 1. / n);
   }
   default: return 0.0;
   }
}

// get_om (omega.c:62-85): angular velocity of the energy frame.
DHD double get_om(const dh_params &P, double r)
{
   double O0 = P.rot_omega;
   switch (P.energy_omega) {
   case 1: return 1.0 - O0;
   case 2: return 1.0 / pow(r, 1.5) - O0;
   case 3: {
      double n = 8.0;
      return 1. / pow(pow(r, 1.5 * n) + 1., 1. / n) - O0;
   }
   case 4: return 10. * exp(-.5 * r * r) - O0;
   default: return 0.0 - O0;
   }
}

// get_om1 (omega.c:87-110): d(om)/dr.
DHD double get_om1(const dh_params &P, double r)
{
   switch (P.energy_omega) {
   case 1: return 0.0;
   case 2: return -1.5 / pow(r, 2.5);
   case 3: {
      double n = 8.0;
      return -1.5 * pow(r, -1 + 1.5 * n) / pow(pow(r, 1.5 * n) + 1., 1.0 + 1. / n);
   }
   case 4: return 10. * r * exp(-.5 * r * r);
   default: return 0.0;
   }
}

// get_cs2 (omega.c:116-140): isothermal sound speed squared.
DHD double get_cs2(const dh_params &P, double r)
{
   double Mach = P.mach;
   switch (P.cs2_profile) {
   case 1: return 1. / (Mach * Mach);
   case 2: {
      double nu = .5;
      return .5 / Mach / Mach / pow(r, 2. * nu);
   }
   case 3: {
      double r1 = P.init_par[2], H0 = P.init_par[4];
      return P.metric_par2 * H0 * H0 / (2 * r1 * r1 * r1);
   }
   case 4: {
      double v2 = P.metric_par2 / r;
      return v2 / (Mach * Mach);
   }
   default: return 1.0;
   }
}

}  // namespace dh
