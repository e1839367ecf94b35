/*
 * paul.h -- plugin ABI of the Disco host layer.
 *
 * User plugins (Initial/, Planets/, Boundary/ modules) are written against the
 * reference's paul.h (paul.h:1-257 of HangChelseaSu/Disco) and compile
 * unchanged against this header: the enums, compile-time sizes and the
 * members of param_list, domain, cell, face and planet keep the reference's
 * names and meaning.  Compile-time sizes come from -DNUM_C, -DNUM_N and
 * -DCT_MODE exactly as in the reference's Makefile.
 *
 * The struct layouts are the reference's (with USE_MPI=0), so the host
 * layer links into a program whose other objects were built against the
 * reference's own paul.h (tests/test_binding.py checks every offset).
 * What differs is ownership: the state lives on the GPU (libdisco_hip.so);
 * theCells is a host mirror filled by disco_sync_to_host(), and the host
 * layer keeps each domain's device context in a table of its own.
 */
#ifndef PAUL
#define PAUL

enum { RHO, PPP, URR, UPP, UZZ, BRR, BPP, BZZ };
enum { DDD, TAU, SRR, LLL, SZZ };
enum { PLPOINTMASS, PLPW, PLSURFACEGRAV };

enum {
   PROF_TOT, PROF_DT, PROF_TIMESTEP, PROF_OUTPUT, PROF_RECON, PROF_FLUX,
   PROF_CT, PROF_SOURCE, PROF_C2P, PROF_BOUND, PROF_EXCHANGE,
   PROF_RECON_R, PROF_RECON_P, PROF_RECON_Z,
   PROF_FLUX_R, PROF_FLUX_P, PROF_FLUX_Z,
   PROF_EXCH_NP_COUNT1, PROF_EXCH_NP_COMM1, PROF_EXCH_NP_COUNT2,
   PROF_EXCH_NP_COMM2, PROF_EXCH_NP_FIN,
   PROF_EXCH_PREP, PROF_EXCH_COMM, PROF_EXCH_FIN,
   NUM_PROF
};

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#ifndef NUM_C
#define NUM_C 5
#endif
#ifndef NUM_N
#define NUM_N 0
#endif
#ifndef CT_MODE
#define CT_MODE 0
#endif

#define NUM_Q (NUM_C + NUM_N)
#define NUM_G 2

#if CT_MODE == 1
#define NUM_EDGES 4
#define NUM_FACES 3
#define NUM_AZ_EDGES 0
#elif CT_MODE == 2
#define NUM_EDGES 8
#define NUM_FACES 5
#define NUM_AZ_EDGES 4
#else
#define NUM_EDGES 0
#define NUM_FACES 0
#define NUM_AZ_EDGES 0
#endif

struct param_list {
   double t_min, t_max;
   int Num_R, Num_Z;
   double aspect;
   int NumRepts, NumSnaps, NumChecks;
   int Out_LogTime;
   double rmin, rmax;
   double zmin, zmax;
   double phimax;
   int NoBC_Rmin, NoBC_Rmax, NoBC_Zmin, NoBC_Zmax;
   int LogZoning, R_Periodic, Z_Periodic;
   double LogRadius;
   double MaxShort, MaxLong;
   int Mesh_Motion, Riemann_Solver, Timestep;
   int Absorb_BC, Initial_Regrid, visc_flag, include_atmos;
   double CFL, PLM, maxDT;
   double Density_Floor, Pressure_Floor;
   int Exact_Mesh_Omega;
   double Exact_Mesh_Omega_Par;
   int Energy_Omega;
   double Energy_Omega_Par;
   int RotFrame;
   double RotOmega, RotD;
   double Adiabatic_Index;
   double viscosity;
   int isothermal_flag;
   int Cs2_Profile;
   double Cs2_Par;
   double Disk_Mach;
   double Mass_Ratio;
   double Eccentricity;
   double Drift_Rate, Drift_Exp;
   int grav2D;
   int alpha_flag;
   int restart_flag;
   int CT;
   int metricPar0;
   double metricPar1, metricPar2, metricPar3, metricPar4;
   int initPar0;
   double initPar1, initPar2, initPar3, initPar4, initPar5, initPar6, initPar7,
       initPar8;
   int noiseType;
   double noiseAbs;
   double noiseRel;
   int sinkType;
   double sinkPar1, sinkPar2, sinkPar3, sinkPar4;
   int nozzleType;
   double nozzlePar1, nozzlePar2, nozzlePar3, nozzlePar4;
};

struct diagnostic_avg {
   double *Qrz;
   double t_avg;
};

struct domain {
   struct cell **theCells;
   struct face *theFaces_1;
   struct face *theFaces_2;
   struct planet *thePlanets;
   struct profiler *prof;
   int *Np;
   int Nr, Nz, Ng;
   int NgRa, NgRb, NgZa, NgZb;
   int N0r, N0z, Nr_glob, Nz_glob, N0r_glob, N0z_glob;
   int N_ftracks_r;
   int N_ftracks_z;
   int Npl;
   double *r_jph;
   double *z_kph;
   double phi_max;
   int *fIndex_r;
   int *fIndex_z;
   time_t Wallt_init;
   int rank, size;
   int dim_rank[2];
   int dim_size[2];
   int left_rank[2];
   int right_rank[2];
   struct param_list theParList;
   int num_tools;
   struct diagnostic_avg theTools;
   double t;
   int count_steps;
   double t_init, t_fin;
   int nrpt;
   int N_rpt;
   int nsnp;
   int N_snp;
   int nchk;
   int N_chk;
   int final_step;
   int check_plz;
};

struct cell {
   double prim[NUM_Q];
   double cons[NUM_Q];
   double RKcons[NUM_Q];
   double gradr[NUM_Q];
   double gradp[NUM_Q];
   double gradz[NUM_Q];
   double piph;
   double dphi;
   double wiph;
   /* zero-length without constrained transport (a GNU C extension), as
    * in the reference, so the layout is the reference's for every CT_MODE */
   double E[NUM_EDGES];
   double B[NUM_EDGES];
   double E_phi[NUM_AZ_EDGES];
   double Phi[NUM_FACES];
   double RK_Phi[NUM_FACES];
   double tempDoub;
   int real;
};

struct face {
   struct cell *L;
   struct cell *R;
   double dxL;
   double dxR;
   double cm[3];
   double dphi;
   double dl;
   double dA;
   double E, B;
   int LRtype;
   int flip_flag;
};

struct planet {
   double r;
   double phi;
   double M;
   double omega;
   double vr;
   double RK_r;
   double RK_phi;
   double RK_M;
   double RK_omega;
   double RK_vr;
   double eps;
   double Fr;
   double Fp;
   int type;
};

struct profiler {
   clock_t ticks[NUM_PROF];
   clock_t elapsed_ticks[NUM_PROF];
};

#endif
