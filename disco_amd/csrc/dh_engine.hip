// dh_engine.hip -- MI355X (gfx950) executor of the timestep and the dh_* C ABI.
//
// The per-thread work of every kernel lives in dh_kernels.h and the launch
// sequence of a timestep in dh_pipeline.h; this file supplies what only the
// GPU has: the generic launch, four cooperative kernels (ring starts, face
// flips, the ghost-ring walk, the CFL reduction), HIP memory and streams,
// and per-kernel event timers.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "disco_hip.h"
#include "dh_pipeline.h"

using namespace dh;

static constexpr int TPB = 256;
static constexpr unsigned NXCD = 8;  // accelerator complex dies, each with its own L2
static inline unsigned nblocks(long n) { return (unsigned)((n + TPB - 1) / TPB); }
// Grid of the XCD-aware launches: rounded up to a whole number of blocks per XCD.
static inline unsigned nblocks_xcd(long n) { return (nblocks(n) + NXCD - 1) / NXCD * NXCD; }

// One thread per element: t = cell, face, face track or boundary line.
template <class K>
__global__ __launch_bounds__(TPB) void k_run(K k, long n)
{
   long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
   if (t < n) k(t);
}

// The same with the workgroups renumbered so that each XCD runs one
// contiguous eighth of the elements (workgroups are dispatched to the XCDs
// round-robin, block b to XCD b mod 8): the neighbouring annuli a kernel
// gathers from then stay in that XCD's L2.  Measured faster only for the
// kernels that declare xcd_contiguous (the r slopes and the face lists);
// streaming kernels run slower with it.
template <class K>
__global__ __launch_bounds__(TPB) void k_run_xcd(K k, long n)
{
   unsigned per = gridDim.x / NXCD;
   unsigned lb = (blockIdx.x % NXCD) * per + blockIdx.x / NXCD;
   long t = (long)lb * blockDim.x + threadIdx.x;
   if (t < n) k(t);
}

template <class K, class = void>
struct XcdContiguous {
   static constexpr bool value = false;
};
template <class K>
struct XcdContiguous<K, decltype((void)K::xcd_contiguous)> {
   static constexpr bool value = K::xcd_contiguous;
};

// The timestep's {dt, planets}, passed by value and stored by one thread,
// so that a recorded timestep that follows reads this step's values.
__global__ void k_set_step(StepArgs *dst, StepArgs sa) { *dst = sa; }

// Whole-buffer copy and clear in 8-byte words (kernels rather than the
// runtime's copy and memset, so that a recorded timestep holds kernel nodes
// only).
struct KCopyWords {
   unsigned long long *dst;
   const unsigned long long *src;
   DHD void operator()(long i) const { dst[i] = src[i]; }
};
struct KZeroWords {
   unsigned long long *p;
   DHD void operator()(long i) const { p[i] = 0ull; }
};

// Start cell of every annulus (faces_alt.c:59-78): one wavefront per
// annulus scans 64 cells at a time with ballots for the first cell in the
// first quadrant that follows a cell in the fourth.
// Block b of the launch (its TPB / 64 wavefronts take annuli 4 b .. 4 b + 3).
__device__ void ring_start_block(const Mesh &m, const double *piph, int *I0, unsigned blk)
{
   int a = blk * (TPB / 64) + threadIdx.x / 64;
   int lane = threadIdx.x & 63;
   if (a >= m.nann) return;
   int c0 = m.off[a], np = m.Np[a];
   double convert = 2. * M_PI / m.Pmax;
   bool seen4 = false;
   int found = -1;
   for (int base = 0; base < np; base += 64) {
      int i = base + lane;
      int quad = 0;
      if (i < np) {
         double sn, cs;
         sincos(piph[c0 + i] * convert, &sn, &cs);
         if (sn > 0. && cs > 0.) quad = 1;
         else if (sn < 0. && cs > 0.) quad = 4;
      }
      unsigned long long b4 = __ballot(quad == 4);
      unsigned long long b1 = __ballot(quad == 1);
      unsigned long long cand = 0;
      if (seen4) {
         cand = b1;
      } else if (b4) {
         int f4 = __ffsll((long long)b4) - 1;
         cand = f4 >= 63 ? 0ULL : (b1 & (~0ULL << (f4 + 1)));
      }
      if (cand) {
         found = base + __ffsll((long long)cand) - 1;
         break;
      }
      if (b4) seen4 = true;
   }
   if (lane == 0) I0[a] = found < 0 ? 0 : found;
}
__global__ __launch_bounds__(TPB) void k_ring_start(Mesh m, const double *piph, int *I0)
{
   ring_start_block(m, piph, I0, blockIdx.x);
}

// flip_fluxes (bfields.c:742-814): one wavefront per face track; the lanes
// read 64 flags at a time and lane 0 hands the flagged faces over in order.
__device__ void flips_block(const Mesh &m, int dim, const Faces &F, const signed char *flag,
                            double *Phi, double *rkPhi, unsigned blk)
{
   int JK = blk * (TPB / 64) + threadIdx.x / 64;
   int lane = threadIdx.x & 63;
   if (JK >= F.npairs || track_skipped(F.npairs, JK)) return;
   int base = F.fI[JK], N = F.fI[JK + 1] - base;
   for (int chunk = 0; chunk < N; chunk += 64) {
      int n = chunk + lane;
      unsigned long long bits = __ballot(n < N && flag[base + n] != 0);
      if (lane == 0) {
         while (bits) {
            int b = __ffsll((long long)bits) - 1;
            bits &= bits - 1;
            flip_one(dim, (size_t)m.ntot, F, base, N, chunk + b, Phi, rkPhi);
         }
      }
   }
}
__global__ __launch_bounds__(TPB) void k_flips(Mesh m, int dim, Faces F, const signed char *flag,
                                               double *Phi, double *rkPhi)
{
   flips_block(m, dim, F, flag, Phi, rkPhi, blockIdx.x);
}

// The ghost-ring walk (KBcScan), one workgroup per ring.  The walk's state
// on entering ghost cell i is the source cell holding the cell's lower edge;
// every thread guesses it for its cells by bisection (bc_guess), the first
// cell's by the walk's own start test (the lowest element that passes it),
// then checks each guess against one step of the walk from the guess before
// it.  A ring with any miss is walked again sequentially, so start[] is the
// sequential walk's in every case.
__device__ void bc_scan_block(const Mesh &m, int side, const double *piph, const double *dphi,
                              int *start, unsigned blk)
{
   __shared__ int s_first, s_miss;
   __shared__ double s_minv[TPB];
   __shared__ int s_mini[TPB];
   int lines = bc_lines(m, side);
   int t = blk;
   int line = t % lines, g = t / lines;
   int dst, src;
   bc_pair(m, side, line, g, &dst, &src);
   double pm = m.Pmax;
   int c0 = m.off[dst], s0 = m.off[src];
   int Np = m.Np[dst], Np1 = m.Np[src];
   int tid = threadIdx.x;
   if (tid == 0) {
      s_first = Np1;
      s_miss = 0;
   }
   __syncthreads();
   double phi0 = piph[c0] - dphi[c0];
   double mv = HUGE_VAL;
   int mi = 0;
   for (int k = tid; k < Np1; k += TPB) {
      if (bc_holds_start(piph, dphi, s0, k, phi0, pm)) atomicMin(&s_first, k);
      double v = piph[s0 + k];
      if (v < mv) { mv = v; mi = k; }
   }
   s_minv[tid] = mv;
   s_mini[tid] = mi;
   __syncthreads();
   for (int o = TPB / 2; o > 0; o >>= 1) {
      if (tid < o) {
         double v = s_minv[tid + o];
         int i = s_mini[tid + o];
         if (v < s_minv[tid] || (v == s_minv[tid] && i < s_mini[tid])) {
            s_minv[tid] = v;
            s_mini[tid] = i;
         }
      }
      __syncthreads();
   }
   int first = s_first < Np1 ? s_first : 0;
   int rot = s_mini[0];
   for (int i = tid; i < Np; i += TPB)
      start[c0 + i] = i == 0 ? first : bc_guess(piph, s0, Np1, rot, piph[c0 + i - 1], pm);
   __syncthreads();
   for (int i = tid; i < Np - 1; i += TPB)
      if (bc_walk_cell(piph, dphi, c0 + i, s0, Np1, start[c0 + i], pm) != start[c0 + i + 1])
         s_miss = 1;
   __syncthreads();
   if (s_miss && tid == 0) KBcScan{m, side, piph, dphi, start}(t);
}
__global__ __launch_bounds__(TPB) void k_bc_scan(Mesh m, int side, const double *piph,
                                                 const double *dphi, int *start)
{
   bc_scan_block(m, side, piph, dphi, start, blockIdx.x);
}

// add_diagnostics (analysis.c:52-106): one workgroup per annulus; each
// thread sums get_diagnostics x dV and dV over every TPB-th cell, each wave
// adds its lanes' sums and thread 0 the waves'.  (The host build sums the
// cells in order, as the reference does; the two differ by rounding only.)
template <int KIND>
__device__ void diag_block(const Mesh &m, const dh_params &P, const DiagPlanet &p1,
                           const double *prim, const double *piph, const double *dphi, double dt,
                           double *Qrz, unsigned blk)
{
   constexpr int NQD = KIND == DH_DIAG_EULER ? 8 : KIND == DH_DIAG_DISK2D ? 5 : 18;
   __shared__ double red[TPB / 64][NQD + 1];
   int a = blk;  // the whole workgroup's annulus
   int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
   if (a >= m.nann) return;
   size_t nt = (size_t)m.ntot;
   int j = a % m.Nr, k = a / m.Nr;
   double sum[NQD], vol = 0.0;
#pragma unroll
   for (int q = 0; q < NQD; ++q) sum[q] = 0.0;
   for (int i = tid; i < m.Np[a]; i += TPB) {
      long c = m.off[a] + i;
      double x[3], dV, p[BZZ + 1], Q[NQD];
      diag_geometry(m, piph, dphi, c, j, k, x, &dV);
#pragma unroll
      for (int q = 0; q < BZZ + 1; ++q)
         if (q < diag_fields(KIND)) p[q] = prim[q * nt + c];
      diag_cell(KIND, P, p1, x[0], x[1], x[2], p, Q);
#pragma unroll
      for (int q = 0; q < NQD; ++q) sum[q] += Q[q] * dV;
      vol += dV;
   }
   for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
      for (int q = 0; q < NQD; ++q) sum[q] += __shfl_down(sum[q], o, 64);
      vol += __shfl_down(vol, o, 64);
   }
   if (lane == 0) {
#pragma unroll
      for (int q = 0; q < NQD; ++q) red[wv][q] = sum[q];
      red[wv][NQD] = vol;
   }
   __syncthreads();
   if (tid == 0) {
      for (int w = 1; w < TPB / 64; ++w) {
#pragma unroll
         for (int q = 0; q < NQD; ++q) sum[q] += red[w][q];
         vol += red[w][NQD];
      }
#pragma unroll
      for (int q = 0; q < NQD; ++q) Qrz[(size_t)a * NQD + q] += sum[q] * dt / vol;
   }
}
template <int KIND>
__global__ __launch_bounds__(TPB) void k_diag(Mesh m, dh_params P, DiagPlanet p1,
                                              const double *prim, const double *piph,
                                              const double *dphi, double dt, double *Qrz)
{
   diag_block<KIND>(m, P, p1, prim, piph, dphi, dt, Qrz, blockIdx.x);
}

// report (report.c:7-202): one wavefront per real annulus; each lane sums
// the terms of every 64th cell, then the wave combines the lanes (the host
// build sums the cells in order; the two differ by rounding only).
__global__ __launch_bounds__(TPB) void k_report(Mesh m, dh_params P, DiagPlanet p0, DiagPlanet p1,
                                                int mhd, const double *prim, const double *piph,
                                                const double *dphi, double *R)
{
   int a = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
   int lane = threadIdx.x & 63;
   if (a >= m.nann) return;
   int j = a % m.Nr, k = a / m.Nr;
   if (!is_real(m, j, k)) return;
   double sum[RP_N], o[RP_N];
#pragma unroll
   for (int q = 0; q < RP_RHOMIN; ++q) sum[q] = 0.0;
   sum[RP_RHOMIN] = HUGE_VAL;
   for (int i = lane; i < m.Np[a]; i += 64) {
      report_cell(m, P, p0, p1, mhd, prim, piph, dphi, m.off[a] + i, j, k, o);
#pragma unroll
      for (int q = 0; q < RP_RHOMIN; ++q) sum[q] += o[q];
      sum[RP_RHOMIN] = fmin(sum[RP_RHOMIN], o[RP_RHOMIN]);
   }
   for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
      for (int q = 0; q < RP_RHOMIN; ++q) sum[q] += __shfl_down(sum[q], off, 64);
      sum[RP_RHOMIN] = fmin(sum[RP_RHOMIN], __shfl_down(sum[RP_RHOMIN], off, 64));
   }
   if (lane == 0)
#pragma unroll
      for (int q = 0; q < RP_N; ++q) R[(size_t)a * RP_N + q] = sum[q];
}

// CFL reduction over real cells (misc.c:39-84).  Positive doubles order like
// their bit patterns, so the minimum is an unsigned 64-bit atomic min, one
// per workgroup; the grid is capped (MINDT_BLOCKS) and each thread takes
// every stride-th cell, so that few atomics meet at the one address.
constexpr unsigned MINDT_BLOCKS = 2048;
// Block blk of nb blocks that share the cells (a grid-stride loop).
template <int NQ, class H>
__device__ void mindt_block(const Mesh &m, const dh_params &P, const double *prim,
                            const double *piph, const double *dphi, const double *wiph,
                            unsigned long long *out, unsigned blk, unsigned nb)
{
   __shared__ double red[TPB / 64];
   long stride = (long)nb * TPB;
   double dtc = HUGE_VAL;
   for (long c = (long)blk * TPB + threadIdx.x; c < m.ntot; c += stride) {
      double v = cell_mindt<NQ, H>(m, P, prim, piph, dphi, wiph, c);
      dtc = v < dtc ? v : dtc;
   }
   for (int o = 32; o > 0; o >>= 1) {
      double v = __shfl_down(dtc, o, 64);
      dtc = v < dtc ? v : dtc;
   }
   int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
   if (lane == 0) red[wv] = dtc;
   __syncthreads();
   if (threadIdx.x == 0) {
      double v = red[0];
      for (int w = 1; w < TPB / 64; ++w) v = red[w] < v ? red[w] : v;
      if (v < HUGE_VAL && v > 0.0) atomicMin(out, (unsigned long long)__double_as_longlong(v));
   }
}
template <int NQ, class H>
__global__ __launch_bounds__(TPB) void k_mindt(Mesh m, dh_params P, const double *prim,
                                               const double *piph, const double *dphi,
                                               const double *wiph, unsigned long long *out)
{
   mindt_block<NQ, H>(m, P, prim, piph, dphi, wiph, out, blockIdx.x, gridDim.x);
}

// ---------------------------------------------------------------------------
// The learned cons2prim guess: the MLP of dh_c2pnet.h on 64 cells per
// wavefront, its three layers as 16x16x16 bfloat16 MFMA products with float32
// sums.  A workgroup (4 wavefronts) takes 256 consecutive cells per tile;
// every lane forms its cell's features, and activations pass between the
// layers through LDS, where the A operands (16 rows of cells x 16 inputs)
// are read back in the register layout of the instruction: lane l holds row
// l % 16, inputs 4 (l / 16) .. 4 (l / 16) + 3.  The products' results
// arrive as lane l = column l % 16, rows 4 (l / 16) .. + 3.  Cells outside
// the real mesh run the same instructions on zero features and write NaN.
// ---------------------------------------------------------------------------

typedef short v4s_t __attribute__((ext_vector_type(4)));
typedef float v4f_t __attribute__((ext_vector_type(4)));

constexpr int C2P_XP = 20;  // LDS row pitch (bfloat16) of the features: 16 + 4 against bank conflicts
constexpr int C2P_HP = 68;  // of a hidden layer: 64 + 4

template <int NQ>
__global__ __launch_bounds__(TPB) void k_c2p_guess(Mesh m, dh_params P, const double *cons,
                                                   const double *piph, const double *dphi,
                                                   const float *prm, const v4s_t *frag,
                                                   float *guess, long ntiles)
{
   __shared__ __align__(16) unsigned short lx[TPB * C2P_XP];
   __shared__ __align__(16) unsigned short lh[TPB * C2P_HP];
   __shared__ float lo[TPB * C2PNET_OUT];
   int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
   int col = lane & 15, quad = lane >> 4;
   unsigned short *X = lx + wv * 64 * C2P_XP;
   unsigned short *H = lh + wv * 64 * C2P_HP;
   float *O = lo + wv * 64 * C2PNET_OUT;
   for (long t = blockIdx.x; t < ntiles; t += gridDim.x) {
      long c = t * TPB + threadIdx.x;
      float f[C2PNET_IN];
      bool have = false;
      double e = 0.0;
      if (c < m.ntot) {
         int a = m.cjk[c];
         int j = a % m.Nr, k = a / m.Nr;
         if (is_real(m, j, k)) {
            C2pInv v = cell_c2p_invariants<NQ>(m, P, cons, piph, dphi, c, j, k);
            c2p_features(v, P.gamma_law, f);
            e = v.e;
            have = true;
         }
      }
      for (int i = 0; i < C2PNET_IN; ++i)
         X[lane * C2P_XP + i] =
             have ? bf16_bits((f[i] - prm[C2PNET_IN_MU + i]) / prm[C2PNET_IN_SD + i]) : 0;
      for (int i = C2PNET_IN; i < 16; ++i) X[lane * C2P_XP + i] = 0;
      __syncthreads();
      // layer 1: 64 x 16 features times 16 x 64 weights
      for (int rb = 0; rb < 4; ++rb) {
         v4s_t av = *(const v4s_t *)(X + (16 * rb + col) * C2P_XP + 4 * quad);
         v4f_t acc[4];
         for (int nb = 0; nb < 4; ++nb) {
            acc[nb] = v4f_t{0.f, 0.f, 0.f, 0.f};
            acc[nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(
                av, frag[(C2PNET_FRAG_L1 + nb) * 64 + lane], acc[nb], 0, 0, 0);
         }
         for (int nb = 0; nb < 4; ++nb)
            for (int i = 0; i < 4; ++i) {
               float h = acc[nb][i] + prm[C2PNET_B1 + 16 * nb + col];
               H[(16 * rb + 4 * quad + i) * C2P_HP + 16 * nb + col] = bf16_bits(h > 0.f ? h : 0.f);
            }
      }
      __syncthreads();
      // layer 2: 64 x 64 times 64 x 64, one row block at a time, in place
      for (int rb = 0; rb < 4; ++rb) {
         v4f_t acc[4];
         for (int nb = 0; nb < 4; ++nb) acc[nb] = v4f_t{0.f, 0.f, 0.f, 0.f};
         for (int kb = 0; kb < 4; ++kb) {
            v4s_t av = *(const v4s_t *)(H + (16 * rb + col) * C2P_HP + 16 * kb + 4 * quad);
            for (int nb = 0; nb < 4; ++nb)
               acc[nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(
                   av, frag[(C2PNET_FRAG_L2 + 4 * kb + nb) * 64 + lane], acc[nb], 0, 0, 0);
         }
         __syncthreads();  // the row block's inputs are read before it is overwritten
         for (int nb = 0; nb < 4; ++nb)
            for (int i = 0; i < 4; ++i) {
               float h = acc[nb][i] + prm[C2PNET_B2 + 16 * nb + col];
               H[(16 * rb + 4 * quad + i) * C2P_HP + 16 * nb + col] = bf16_bits(h > 0.f ? h : 0.f);
            }
      }
      __syncthreads();
      // layer 3: 64 x 64 times 64 x 16 (two live columns)
      for (int rb = 0; rb < 4; ++rb) {
         v4f_t acc = v4f_t{0.f, 0.f, 0.f, 0.f};
         for (int kb = 0; kb < 4; ++kb) {
            v4s_t av = *(const v4s_t *)(H + (16 * rb + col) * C2P_HP + 16 * kb + 4 * quad);
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(
                av, frag[(C2PNET_FRAG_L3 + kb) * 64 + lane], acc, 0, 0, 0);
         }
         if (col < C2PNET_OUT)
            for (int i = 0; i < 4; ++i)
               O[(16 * rb + 4 * quad + i) * C2PNET_OUT + col] = acc[i] + prm[C2PNET_B3 + col];
      }
      __syncthreads();
      if (c < m.ntot) {
         if (have) {
            float tt[C2PNET_OUT], g[2];
            for (int o = 0; o < C2PNET_OUT; ++o)
               tt[o] = O[lane * C2PNET_OUT + o] * prm[C2PNET_OUT_SD + o] + prm[C2PNET_OUT_MU + o];
            c2p_targets_to_guess(tt, e, g);
            guess[2 * c] = g[0];
            guess[2 * c + 1] = g[1];
         } else {
            guess[2 * c] = guess[2 * c + 1] = __builtin_nanf("");
         }
      }
      __syncthreads();  // LDS is reused by the next tile
   }
}

// ---------------------------------------------------------------------------
// HIP executor
// ---------------------------------------------------------------------------

#define HIPCHK(x)                                                              \
   do {                                                                        \
      hipError_t e_ = (x);                                                     \
      if (e_ != hipSuccess)                                                    \
         return fail(DH_EHIP, "%s failed: %s", #x, hipGetErrorString(e_));     \
   } while (0)

#define LAUNCH_CHECK()                                                          \
   do {                                                                         \
      hipError_t e_ = hipGetLastError();                                        \
      if (e_ != hipSuccess)                                                     \
         return fail(DH_EHIP, "kernel launch failed: %s", hipGetErrorString(e_)); \
   } while (0)

struct TimerRow {
   std::string name;
   std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
   double ms = 0.0;
   long calls = 0;
};

// Launch variants of a kernel that runs over cells or faces (run_tuned):
// 0 storage order with the workgroups dealt to the XCDs round-robin (the
// hardware's order), 1 storage order with each XCD given a contiguous
// eighth (k_run_xcd), 2 + s the group order s of the kernel's OrderSet
// (strips of annuli, XCD-contiguous).  All give the same results.
struct TuneRow {
   std::string name;
   int nvar = 0, chosen = -1, dflt = 0;
   double ms[2 + MAX_ORDERS] = {0};
   int n[2 + MAX_ORDERS] = {0};
   int strip[MAX_ORDERS] = {0};
};
struct TunePending {
   int row, var;
   hipEvent_t e0, e1;
};

struct HipExec {
   hipStream_t stream = nullptr;
   bool timers = false;
   // Tuner (dh_autotune): while tune_left timesteps remain, a tunable
   // kernel runs in variant (tune_step mod its variant count) for all its
   // launches of the step, so that every variant is timed over whole steps
   // (the two substeps' launches of a kernel differ in cost, and variants
   // dealt out launch by launch would each meet only one of them); then
   // each kernel keeps its fastest variant.
   int tune_left = 0, tune_step = 0;
   std::vector<TuneRow> tunes;
   std::vector<TunePending> tpend;
   // Recorded timestep (dh_enable_graphs): replayed while the key it was
   // recorded under (the current face positions' buffer) is unchanged.
   bool graphs = false;
   hipGraphExec_t gexec = nullptr;
   const void *gkey = nullptr;
   std::vector<TimerRow> rows;
   std::vector<void *> allocs;
   unsigned long long *d_min = nullptr;

   // Brackets one launch with events when the timers are on.
   struct Bracket {
      HipExec &x;
      int row = -1;
      hipEvent_t e0{}, e1{};
      Bracket(HipExec &ex, const char *name) : x(ex)
      {
         if (!x.timers) return;
         for (size_t i = 0; i < x.rows.size(); ++i)
            if (x.rows[i].name == name) row = (int)i;
         if (row < 0) {
            x.rows.emplace_back();
            x.rows.back().name = name;
            row = (int)x.rows.size() - 1;
         }
         (void)hipEventCreate(&e0);
         (void)hipEventCreate(&e1);
         (void)hipEventRecord(e0, x.stream);
      }
      ~Bracket()
      {
         if (row < 0) return;
         (void)hipEventRecord(e1, x.stream);
         x.rows[row].pending.push_back({e0, e1});
      }
   };

   template <class K>
   int run(const char *name, const K &k, long n)
   {
      if (n <= 0) return DH_OK;
      {
         Bracket b(*this, name);
         if constexpr (XcdContiguous<K>::value)
            k_run_xcd<K><<<nblocks_xcd(n), TPB, 0, stream>>>(k, n);
         else
            k_run<K><<<nblocks(n), TPB, 0, stream>>>(k, n);
      }
      LAUNCH_CHECK();
      return DH_OK;
   }

   TuneRow &tune_row(const char *name, int nvar, const int *strips, int dflt)
   {
      for (auto &t : tunes)
         if (t.name == name) return t;
      tunes.emplace_back();
      TuneRow &t = tunes.back();
      t.name = name;
      t.nvar = nvar;
      t.dflt = dflt;
      for (int s = 0; s + 2 < nvar; ++s) t.strip[s] = strips[s];
      return t;
   }

   // A kernel over cells or faces in one of its launch variants (TuneRow).
   template <class K>
   int run_tuned(const char *name, const K &k, long n, const GroupOrder *ords, int nord,
                 const int *strips)
   {
      if (n <= 0) return DH_OK;
      int row = (int)(&tune_row(name, 2 + nord, strips, XcdContiguous<K>::value ? 1 : 0) -
                      tunes.data());
      TuneRow &t = tunes[row];
      int v = t.chosen >= 0 ? t.chosen : t.dflt;
      bool timing = tune_left > 0;
      if (timing) v = tune_step % t.nvar;
      hipEvent_t e0{}, e1{};
      if (timing) {
         HIPCHK(hipEventCreate(&e0));
         HIPCHK(hipEventCreate(&e1));
         HIPCHK(hipEventRecord(e0, stream));
      }
      {
         Bracket b(*this, name);
         if (v == 0)
            k_run<K><<<nblocks(n), TPB, 0, stream>>>(k, n);
         else if (v == 1)
            k_run_xcd<K><<<nblocks_xcd(n), TPB, 0, stream>>>(k, n);
         else
            k_run_xcd<Grouped<K>><<<nblocks_xcd(n), TPB, 0, stream>>>(Grouped<K>{k, ords[v - 2]}, n);
      }
      LAUNCH_CHECK();
      if (timing) {
         HIPCHK(hipEventRecord(e1, stream));
         tpend.push_back({row, v, e0, e1});
      }
      return DH_OK;
   }

   // After a tuned timestep: the launch times of its variants; on the last
   // tuned step, every kernel's choice.
   int tune_collect()
   {
      HIPCHK(hipStreamSynchronize(stream));
      for (auto &p : tpend) {
         float ms = 0.0f;
         HIPCHK(hipEventElapsedTime(&ms, p.e0, p.e1));
         tunes[p.row].ms[p.var] += ms;
         tunes[p.row].n[p.var] += 1;
         (void)hipEventDestroy(p.e0);
         (void)hipEventDestroy(p.e1);
      }
      tpend.clear();
      ++tune_step;
      if (--tune_left == 0)
         for (auto &t : tunes) {
            int best = -1;
            for (int v = 0; v < t.nvar; ++v)
               if (t.n[v] && (best < 0 || t.ms[v] / t.n[v] < t.ms[best] / t.n[best])) best = v;
            if (best >= 0) t.chosen = best;
         }
      return DH_OK;
   }

   int ring_start(const Mesh &m, const double *piph, int *I0)
   {
      {
         Bracket b(*this, "ring_start");
         k_ring_start<<<(unsigned)((m.nann + 3) / 4), TPB, 0, stream>>>(m, piph, I0);
      }
      LAUNCH_CHECK();
      return DH_OK;
   }

   int bc_scan(const Mesh &m, int side, const double *piph, const double *dphi, int *start)
   {
      long rings = (long)bc_lines(m, side) * bc_layers(m, side);
      if (rings <= 0) return DH_OK;
      {
         Bracket b(*this, "bc_scan");
         k_bc_scan<<<(unsigned)rings, TPB, 0, stream>>>(m, side, piph, dphi, start);
      }
      LAUNCH_CHECK();
      return DH_OK;
   }

   int flips(const Mesh &m, int dim, const Faces &F, const signed char *flag, double *Phi,
             double *rkPhi)
   {
      {
         Bracket b(*this, "flips");
         k_flips<<<(unsigned)((F.npairs + 3) / 4), TPB, 0, stream>>>(m, dim, F, flag, Phi, rkPhi);
      }
      LAUNCH_CHECK();
      return DH_OK;
   }

   int diag(const Mesh &m, const dh_params &P, int kind, const DiagPlanet &p1,
            const double *prim, const double *piph, const double *dphi, double dt, double *Qrz)
   {
      unsigned nb = (unsigned)m.nann;
      {
         Bracket b(*this, "diagnostics");
         if (kind == DH_DIAG_EULER)
            k_diag<DH_DIAG_EULER><<<nb, TPB, 0, stream>>>(m, P, p1, prim, piph, dphi, dt, Qrz);
         else if (kind == DH_DIAG_DISK2D)
            k_diag<DH_DIAG_DISK2D><<<nb, TPB, 0, stream>>>(m, P, p1, prim, piph, dphi, dt, Qrz);
         else
            k_diag<DH_DIAG_MHDMRI><<<nb, TPB, 0, stream>>>(m, P, p1, prim, piph, dphi, dt, Qrz);
      }
      LAUNCH_CHECK();
      return DH_OK;
   }

   int report(const Mesh &m, const dh_params &P, const DiagPlanet &p0, const DiagPlanet &p1,
              int mhd, const double *prim, const double *piph, const double *dphi, double *R)
   {
      unsigned nb = (unsigned)((m.nann + 3) / 4);
      {
         Bracket b(*this, "report");
         k_report<<<nb, TPB, 0, stream>>>(m, P, p0, p1, mhd, prim, piph, dphi, R);
      }
      LAUNCH_CHECK();
      return DH_OK;
   }

   template <int NQ, class H>
   int min_dt(const Mesh &m, const dh_params &P, const double *prim, const double *piph,
              const double *dphi, const double *wiph, double *out)
   {
      unsigned long long init = 0x7FF0000000000000ULL;  // +inf
      HIPCHK(hipMemcpyAsync(d_min, &init, sizeof init, hipMemcpyHostToDevice, stream));
      {
         Bracket b(*this, "mindt");
         unsigned nb = nblocks(m.ntot) < MINDT_BLOCKS ? nblocks(m.ntot) : MINDT_BLOCKS;
         k_mindt<NQ, H><<<nb, TPB, 0, stream>>>(m, P, prim, piph, dphi, wiph, d_min);
      }
      LAUNCH_CHECK();
      unsigned long long bits = 0;
      HIPCHK(hipMemcpyAsync(&bits, d_min, sizeof bits, hipMemcpyDeviceToHost, stream));
      HIPCHK(hipStreamSynchronize(stream));
      memcpy(out, &bits, sizeof bits);
      return DH_OK;
   }

   template <int NQ>
   int c2p_guess(const Mesh &m, const dh_params &P, const double *cons, const double *piph,
                 const double *dphi, const float *prm, const unsigned short *frag, float *guess)
   {
      long ntiles = (m.ntot + TPB - 1) / TPB;
      unsigned nb = (unsigned)(ntiles < 4096 ? ntiles : 4096);
      if (nb == 0) return DH_OK;
      {
         Bracket b(*this, "c2p_guess");
         k_c2p_guess<NQ><<<nb, TPB, 0, stream>>>(m, P, cons, piph, dphi, prm,
                                                  (const v4s_t *)frag, guess, ntiles);
      }
      LAUNCH_CHECK();
      return DH_OK;
   }

   int alloc(void **p, size_t bytes)
   {
      if (bytes == 0) bytes = 8;
      hipError_t e = hipMalloc(p, bytes);
      if (e != hipSuccess)
         return fail(DH_ENOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
      allocs.push_back(*p);
      return DH_OK;
   }
   // Frees one allocation of this executor (after the work queued on it).
   int release(void *p)
   {
      if (!p) return DH_OK;
      HIPCHK(hipStreamSynchronize(stream));
      for (size_t i = 0; i < allocs.size(); ++i)
         if (allocs[i] == p) {
            allocs.erase(allocs.begin() + (long)i);
            HIPCHK(hipFree(p));
            return DH_OK;
         }
      return fail(DH_EINVAL, "release of an unknown pointer");
   }
   int put(void *dst, const void *src, size_t bytes)
   {
      HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream));
      HIPCHK(hipStreamSynchronize(stream));
      return DH_OK;
   }
   int get(void *dst, const void *src, size_t bytes)
   {
      HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream));
      HIPCHK(hipStreamSynchronize(stream));
      return DH_OK;
   }
   static bool words(const void *p, size_t bytes)
   {
      return bytes % 8 == 0 && (uintptr_t)p % 8 == 0;
   }
   int copy(void *dst, const void *src, size_t bytes)
   {
      if (words(dst, bytes) && words(src, bytes))
         return run("copy", KCopyWords{(unsigned long long *)dst, (const unsigned long long *)src},
                    (long)(bytes / 8));
      HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, stream));
      return DH_OK;
   }
   int zero(void *p, size_t bytes)
   {
      if (words(p, bytes)) return run("zero", KZeroWords{(unsigned long long *)p}, (long)(bytes / 8));
      HIPCHK(hipMemsetAsync(p, 0, bytes, stream));
      return DH_OK;
   }
   int set_step(StepArgs *dst, const StepArgs &sa)
   {
      {
         Bracket b(*this, "set_step");
         k_set_step<<<1, 1, 0, stream>>>(dst, sa);
      }
      LAUNCH_CHECK();
      return DH_OK;
   }
   void drop_graph()
   {
      if (gexec) (void)hipGraphExecDestroy(gexec);
      gexec = nullptr;
      gkey = nullptr;
   }
   // Runs body, the launches of one timestep, eagerly; or, with graphs on
   // and a non-null key, records it once on the stream and replays the
   // recording from then on.  body must launch the same kernels with the
   // same arguments whenever key is the same.
   template <class F>
   int replayable(const void *key, F &&body)
   {
      if (tune_left > 0) {
         DH_TRY(body());
         return tune_collect();
      }
      if (!graphs || timers || !key) return body();
      if (!gexec || gkey != key) {
         drop_graph();
         HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
         int rc = body();
         hipGraph_t g = nullptr;
         hipError_t e = hipStreamEndCapture(stream, &g);
         if (rc || e != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            if (rc) return rc;
            return fail(DH_EHIP, "timestep capture failed: %s", hipGetErrorString(e));
         }
         e = hipGraphInstantiateWithFlags(&gexec, g, 0);
         (void)hipGraphDestroy(g);
         if (e != hipSuccess) {
            gexec = nullptr;
            return fail(DH_EHIP, "timestep graph instantiation failed: %s", hipGetErrorString(e));
         }
         gkey = key;
      }
      HIPCHK(hipGraphLaunch(gexec, stream));
      return DH_OK;
   }
   int sync()
   {
      HIPCHK(hipStreamSynchronize(stream));
      return DH_OK;
   }
   static constexpr bool merges_launches() { return true; }
};

struct dh_ctx {
   HipExec x;
   Buffers b;
};

static void release(dh_ctx *c)
{
   if (!c) return;
   if (c->x.stream) (void)hipStreamSynchronize(c->x.stream);
   c->x.drop_graph();
   for (auto &p : c->x.tpend) {
      (void)hipEventDestroy(p.e0);
      (void)hipEventDestroy(p.e1);
   }
   for (auto &row : c->x.rows)
      for (auto &pe : row.pending) {
         (void)hipEventDestroy(pe.first);
         (void)hipEventDestroy(pe.second);
      }
   for (void *p : c->x.allocs) (void)hipFree(p);
   if (c->x.stream) (void)hipStreamDestroy(c->x.stream);
   delete c;
}

template <int NQ, class H>
static int do_init(dh_ctx *c)
{
   Engine<HipExec> e{c->x, c->b};
   return e.template init<NQ, H>();
}
template <int NQ, class H>
static int do_timestep(dh_ctx *c, double dt)
{
   Engine<HipExec> e{c->x, c->b};
   return e.template timestep<NQ, H>(dt);
}
template <int NQ, class H>
static int do_mindt(dh_ctx *c, double *dt)
{
   Engine<HipExec> e{c->x, c->b};
   return e.template mindt<NQ, H>(dt);
}
template <int NQ, class H>
static int do_faces(dh_ctx *c)
{
   Engine<HipExec> e{c->x, c->b};
   return e.template faces<NQ, H>();
}

// ---------------------------------------------------------------------------
// Ensembles (dh_ens_*): independent meshes of one module selection, each
// with its own size, time step and planets, stepped together so that every
// kernel of their timestep is launched once for all of them.  A mesh's work
// in one launch is an "op" (a functor over its elements, or a cooperative
// kernel's per-block body); the grouped kernel k_group<Op> gives block b to
// the mesh whose range of the concatenated block table holds it.  The
// timestep of every mesh is recorded by the same launch sequence the single
// mesh runs (Engine::timestep on a recording executor), and the recorded
// ops of equal position and type become one launch.
// ---------------------------------------------------------------------------
template <class K>
struct RunOp {
   K k;
   long n;
   unsigned blocks() const { return nblocks(n); }
   __device__ void run(unsigned b) const
   {
      long t = (long)b * TPB + threadIdx.x;
      if (t < n) k(t);
   }
};
struct RingStartOp {
   Mesh m;
   const double *piph;
   int *I0;
   unsigned blocks() const { return (unsigned)((m.nann + 3) / 4); }
   __device__ void run(unsigned b) const { ring_start_block(m, piph, I0, b); }
};
struct FlipsOp {
   Mesh m;
   int dim;
   Faces F;
   const signed char *flag;
   double *Phi, *rkPhi;
   unsigned blocks() const { return (unsigned)((F.npairs + 3) / 4); }
   __device__ void run(unsigned b) const { flips_block(m, dim, F, flag, Phi, rkPhi, b); }
};
struct BcScanOp {
   Mesh m;
   int side;
   const double *piph, *dphi;
   int *start;
   unsigned blocks() const { return (unsigned)((long)bc_lines(m, side) * bc_layers(m, side)); }
   __device__ void run(unsigned b) const { bc_scan_block(m, side, piph, dphi, start, b); }
};
template <int NQ, class H>
struct MindtOp {
   Mesh m;
   dh_params P;
   const double *prim, *piph, *dphi, *wiph;
   unsigned long long *out;
   unsigned nb;
   unsigned blocks() const { return nb; }
   __device__ void run(unsigned b) const
   {
      mindt_block<NQ, H>(m, P, prim, piph, dphi, wiph, out, b, nb);
   }
};
// dt and planet 1 of a mesh's diagnostics, written per call
struct EnsDiagArgs {
   double dt;
   DiagPlanet p1;
};
template <int KIND>
struct DiagOp {
   Mesh m;
   dh_params P;
   const EnsDiagArgs *a;
   const double *prim, *piph, *dphi;
   double *Qrz;
   unsigned blocks() const { return (unsigned)m.nann; }
   __device__ void run(unsigned b) const
   {
      diag_block<KIND>(m, P, a->p1, prim, piph, dphi, a->dt, Qrz, b);
   }
};

// Block b of the concatenated table goes to mesh m with boff[m] <= b < boff[m + 1].
template <class Op>
__global__ __launch_bounds__(TPB) void k_group(const Op *ops, const unsigned *boff, int nm)
{
   unsigned b = blockIdx.x;
   int lo = 0, hi = nm;
   while (hi - lo > 1) {
      int mid = (lo + hi) >> 1;
      if (boff[mid] <= b) lo = mid;
      else hi = mid;
   }
   ops[lo].run(b - boff[lo]);
}

// Every mesh's {dt, planets} of the step, one thread per mesh.
__global__ void k_set_steps(StepArgs *const *dst, const StepArgs *src, int nm)
{
   int m = blockIdx.x * blockDim.x + threadIdx.x;
   if (m < nm) *dst[m] = src[m];
}
__global__ void k_fill_u64(unsigned long long *p, int n, unsigned long long v)
{
   int i = blockIdx.x * blockDim.x + threadIdx.x;
   if (i < n) p[i] = v;
}

// One recorded op of one mesh.
struct GroupOp {
   const void *tag;  // the op's type
   size_t size;
   std::vector<unsigned char> bytes;
   unsigned blocks;
   void (*launch)(const void *ops, const unsigned *boff, int nm, unsigned total, hipStream_t s);
};
template <class Op>
struct OpType {
   static const void *tag()
   {
      static const char t = 0;
      return &t;
   }
   static void launch(const void *ops, const unsigned *boff, int nm, unsigned total, hipStream_t s)
   {
      if (total) k_group<Op><<<total, TPB, 0, s>>>((const Op *)ops, boff, nm);
   }
};

// The executor that records a mesh's timestep as ops (see Engine::timestep
// for what a timestep asks of an executor).
struct GroupExec {
   std::vector<GroupOp> ops;
   StepArgs *step_dst = nullptr;
   StepArgs step{};
   template <class Op>
   int rec(const Op &op)
   {
      GroupOp g;
      g.tag = OpType<Op>::tag();
      g.size = sizeof(Op);
      g.bytes.resize(sizeof(Op));
      memcpy(g.bytes.data(), &op, sizeof(Op));
      g.blocks = op.blocks();
      g.launch = &OpType<Op>::launch;
      ops.push_back(std::move(g));
      return DH_OK;
   }
   template <class K>
   int run(const char *, const K &k, long n)
   {
      return n > 0 ? rec(RunOp<K>{k, n}) : DH_OK;
   }
   template <class K>
   int run_tuned(const char *name, const K &k, long n, const GroupOrder *, int, const int *)
   {
      return run(name, k, n);
   }
   int ring_start(const Mesh &m, const double *piph, int *I0)
   {
      return rec(RingStartOp{m, piph, I0});
   }
   int bc_scan(const Mesh &m, int side, const double *piph, const double *dphi, int *start)
   {
      if ((long)bc_lines(m, side) * bc_layers(m, side) <= 0) return DH_OK;
      return rec(BcScanOp{m, side, piph, dphi, start});
   }
   int flips(const Mesh &m, int dim, const Faces &F, const signed char *flag, double *Phi,
             double *rkPhi)
   {
      if (F.npairs <= 0) return DH_OK;
      return rec(FlipsOp{m, dim, F, flag, Phi, rkPhi});
   }
   template <int NQ>
   int c2p_guess(const Mesh &, const dh_params &, const double *, const double *, const double *,
                 const float *, const unsigned short *, float *)
   {
      return fail(DH_EINVAL, "the learned cons2prim guess does not run in an ensemble");
   }
   int copy(void *dst, const void *src, size_t bytes)
   {
      if (!HipExec::words(dst, bytes) || !HipExec::words(src, bytes))
         return fail(DH_EINVAL, "an ensemble copies whole 8-byte words only");
      return run("copy", KCopyWords{(unsigned long long *)dst, (const unsigned long long *)src},
                 (long)(bytes / 8));
   }
   int zero(void *p, size_t bytes)
   {
      if (!HipExec::words(p, bytes)) return fail(DH_EINVAL, "an ensemble clears whole 8-byte words only");
      return run("zero", KZeroWords{(unsigned long long *)p}, (long)(bytes / 8));
   }
   int set_step(StepArgs *dst, const StepArgs &sa)
   {
      step_dst = dst;
      step = sa;
      return DH_OK;
   }
   template <class F>
   int replayable(const void *, F &&body)
   {
      return body();
   }
   // Every member must record the same launch sequence.
   static constexpr bool merges_launches() { return false; }
};

// One launch of an ensemble: the ops of every mesh in device memory and
// their concatenated block table.
struct EnsLaunch {
   const void *tag = nullptr;
   void (*launch)(const void *, const unsigned *, int, unsigned, hipStream_t) = nullptr;
   std::vector<unsigned char> host;  // every mesh's op bytes, in mesh order
   unsigned total = 0;
   void *d_ops = nullptr;
   unsigned *d_boff = nullptr;
};

template <int NQ, class H>
static int ens_record_step(GroupExec *gx, Buffers *b, double dt)
{
   Engine<GroupExec> e{*gx, *b};
   return e.template timestep<NQ, H>(dt);
}
template <int NQ, class H>
static int ens_record_mindt(GroupExec *gx, Buffers *b, unsigned long long *out)
{
   unsigned nb = nblocks(b->m.ntot) < 512 ? nblocks(b->m.ntot) : 512;
   return gx->rec(MindtOp<NQ, H>{b->m, b->P, b->prim, b->piph, b->dphi, b->wiph, out, nb});
}

template <int KIND>
static int ens_record_diag(GroupExec *gx, Buffers *b, const EnsDiagArgs *a)
{
   return gx->rec(DiagOp<KIND>{b->m, b->P, a, b->prim, b->piph, b->dphi, b->diagQ});
}

struct dh_ens {
   std::vector<dh_ctx *> mem;
   std::vector<hipStream_t> own_stream;  // the members' streams before they joined
   int nm = 0;
   hipStream_t stream = nullptr;         // every member's work goes here while joined
   std::vector<EnsLaunch> plan;          // the timestep's launches
   bool graphs = false;
   hipGraphExec_t gexec = nullptr;
   StepArgs **d_dst = nullptr;           // each member's StepArgs
   StepArgs *d_step = nullptr, *h_step = nullptr;  // this step's values (h_step pinned)
   EnsLaunch mindt;
   unsigned long long *d_min = nullptr, *h_min = nullptr;
   EnsLaunch diag;
   int diag_kind = 0;
   EnsDiagArgs *d_diag = nullptr, *h_diag = nullptr;
   std::vector<void *> dev, pinned;
};

static void ens_free_launch(EnsLaunch &L)
{
   if (L.d_ops) (void)hipFree(L.d_ops);
   if (L.d_boff) (void)hipFree(L.d_boff);
   L = EnsLaunch();
}

static void ens_drop_graph(dh_ens *e)
{
   if (e->gexec) (void)hipGraphExecDestroy(e->gexec);
   e->gexec = nullptr;
}

static void ens_release(dh_ens *e)
{
   if (!e) return;
   if (e->stream) (void)hipStreamSynchronize(e->stream);
   ens_drop_graph(e);
   for (auto &L : e->plan) ens_free_launch(L);
   ens_free_launch(e->mindt);
   ens_free_launch(e->diag);
   for (void *p : e->dev) (void)hipFree(p);
   for (void *p : e->pinned) (void)hipHostFree(p);
   for (size_t m = 0; m < e->own_stream.size(); ++m) e->mem[m]->x.stream = e->own_stream[m];
   delete e;
}

// The ops of position i of every member's record as one launch (device
// copies of the ops and the block table).
// *changed tells whether the launch differs from what L held (then L is
// rebuilt; otherwise its device copies stay as they are).
static int ens_build_launch(const std::vector<GroupExec> &rec, size_t i, EnsLaunch &L,
                            bool *changed)
{
   int nm = (int)rec.size();
   const GroupOp &g0 = rec[0].ops[i];
   std::vector<unsigned char> host;
   std::vector<unsigned> boff(nm + 1, 0);
   for (int m = 0; m < nm; ++m) {
      const GroupOp &g = rec[m].ops[i];
      host.insert(host.end(), g.bytes.begin(), g.bytes.end());
      boff[m + 1] = boff[m] + g.blocks;
   }
   *changed = !(L.d_ops && L.tag == g0.tag && L.host == host && L.total == boff[nm]);
   if (!*changed) return DH_OK;
   ens_free_launch(L);
   L.tag = g0.tag;
   L.launch = g0.launch;
   L.host = std::move(host);
   L.total = boff[nm];
   HIPCHK(hipMalloc(&L.d_ops, L.host.size()));
   HIPCHK(hipMalloc((void **)&L.d_boff, sizeof(unsigned) * (nm + 1)));
   HIPCHK(hipMemcpy(L.d_ops, L.host.data(), L.host.size(), hipMemcpyHostToDevice));
   HIPCHK(hipMemcpy(L.d_boff, boff.data(), sizeof(unsigned) * (nm + 1), hipMemcpyHostToDevice));
   return DH_OK;
}

// The members' records agree op by op (same type at every position).
static int ens_check_records(const std::vector<GroupExec> &rec)
{
   for (size_t m = 1; m < rec.size(); ++m) {
      if (rec[m].ops.size() != rec[0].ops.size())
         return fail(DH_EINVAL, "ensemble member %zu launches %zu kernels in a step, member 0 %zu",
                     m, rec[m].ops.size(), rec[0].ops.size());
      for (size_t i = 0; i < rec[0].ops.size(); ++i)
         if (rec[m].ops[i].tag != rec[0].ops[i].tag)
            return fail(DH_EINVAL, "ensemble member %zu differs from member 0 at launch %zu", m, i);
   }
   return DH_OK;
}

static int ens_launch(dh_ens *e, const EnsLaunch &L)
{
   L.launch(L.d_ops, L.d_boff, e->nm, L.total, e->stream);
   LAUNCH_CHECK();
   return DH_OK;
}

// ---------------------------------------------------------------------------
// The learned guess's loss gradient without allocation (dh_c2p_grad_noalloc).
// The per-row activations and losses of dh_c2p_grad live in `scratch`, one
// file-scope device array: DH_C2P_NOALLOC_MAX_ROWS rows of C2PNET_ROW
// activations, then one loss per row.  Host code reaches it through the
// address hipGetSymbolAddress gives on each device (each GPU holds its own
// copy), looked up once per device with that device current.  Every row the
// kernels read is written first in the same call, so nothing needs clearing.
// ---------------------------------------------------------------------------
constexpr long C2P_NOALLOC_ROWS = DH_C2P_NOALLOC_MAX_ROWS;
constexpr size_t C2P_SCRATCH_FLOATS = (size_t)C2P_NOALLOC_ROWS * (C2PNET_ROW + 1);
static __device__ __align__(256) float scratch[C2P_SCRATCH_FLOATS];

constexpr int SCRATCH_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static float *g_scratch_ptr[SCRATCH_MAX_DEVICES];

// The device address of `scratch` on the current device.
static int scratch_of_current_device(float **out)
{
   int dev = -1;
   HIPCHK(hipGetDevice(&dev));
   if (dev < 0 || dev >= SCRATCH_MAX_DEVICES)
      return fail(DH_EINVAL, "device %d out of range for the scratch table", dev);
   std::lock_guard<std::mutex> lock(g_scratch_mu);
   if (!g_scratch_ptr[dev]) {
      void *p = nullptr;
      HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
      if (!p) return fail(DH_EHIP, "hipGetSymbolAddress(scratch) gave a null address");
      g_scratch_ptr[dev] = (float *)p;
   }
   *out = g_scratch_ptr[dev];
   return DH_OK;
}

// ---------------------------------------------------------------------------
// C ABI (include/disco_hip.h)
// ---------------------------------------------------------------------------

// The training entry points own a temporary executor (and, on the GPU, a
// stream) for the call.
static int with_exec(int (*fn)(dh_ctx *, void *), void *arg)
{
   int ndev = 0;
   if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(DH_ENODEV, "no GPU visible");
   dh_ctx *c = new dh_ctx();
   if (hipStreamCreate(&c->x.stream) != hipSuccess) {
      delete c;
      return fail(DH_EHIP, "hipStreamCreate failed");
   }
   int rc = fn(c, arg);
   if (!rc) rc = c->x.sync();
   release(c);
   return rc;
}

extern "C" {

const char *dh_last_error(void) { return g_err; }

int dh_destroy(dh_ctx *c)
{
   release(c);
   return DH_OK;
}

int dh_create(const dh_grid *g, const dh_params *par, dh_ctx **out)
{
   if (!g || !par || !out) return fail(DH_EINVAL, "null argument");
   *out = nullptr;
   int ndev = 0;
   if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
      return fail(DH_ENODEV, "no GPU visible");
   long ntot = 0;
   DH_TRY(validate(g, *par, &ntot));
   dh_ctx *c = new dh_ctx();
   if (hipStreamCreate(&c->x.stream) != hipSuccess) {
      delete c;
      return fail(DH_EHIP, "hipStreamCreate failed");
   }
   int rc = c->x.alloc((void **)&c->x.d_min, sizeof(unsigned long long));
   if (!rc) rc = create_buffers(c->x, c->b, g, *par);
   if (!rc) rc = c->x.sync();
   if (rc) {
      release(c);
      return rc;
   }
   *out = c;
   return DH_OK;
}

int dh_sizes(const dh_ctx *c, long *ntot, int *nq, int *nf)
{
   if (!c) return fail(DH_EINVAL, "null context");
   if (ntot) *ntot = c->b.m.ntot;
   if (nq) *nq = c->b.nq;
   if (nf) *nf = c->b.nf;
   return DH_OK;
}

int dh_upload_state(dh_ctx *c, const double *prim, const double *cons, const double *piph,
                    const double *dphi, const double *wiph, const double *Phi)
{
   if (!c) return fail(DH_EINVAL, "null context");
   return upload_state(c->x, c->b, prim, cons, piph, dphi, wiph, Phi);
}

int dh_download_state(dh_ctx *c, double *prim, double *cons, double *piph, double *dphi,
                      double *wiph, double *Phi)
{
   if (!c) return fail(DH_EINVAL, "null context");
   return download_state(c->x, c->b, prim, cons, piph, dphi, wiph, Phi);
}

int dh_init_cells(dh_ctx *c)
{
   if (!c) return fail(DH_EINVAL, "null context");
   return DH_DISPATCH(c->b, do_init, c);
}

int dh_face_count(dh_ctx *c, int dim, long *nfaces)
{
   if (!c || !nfaces || (dim != 1 && dim != 2)) return fail(DH_EINVAL, "bad argument");
   *nfaces = c->b.F[dim - 1].n;
   return DH_OK;
}

int dh_get_faces(dh_ctx *c, int dim, int *L, int *R, int *LRtype, double *cm, double *dA)
{
   if (!c || (dim != 1 && dim != 2)) return fail(DH_EINVAL, "bad argument");
   if (c->b.F[dim - 1].n == 0) return DH_OK;
   DH_TRY(DH_DISPATCH(c->b, do_faces, c));
   return export_faces(c->x, c->b, dim, L, R, LRtype, cm, dA);
}

int dh_set_planets(dh_ctx *c, int substep_, const double *pl, int npl)
{
   if (!c) return fail(DH_EINVAL, "null context");
   return set_planets(c->b, substep_, pl, npl);
}

int dh_getmindt(dh_ctx *c, double *dt)
{
   if (!c || !dt) return fail(DH_EINVAL, "null argument");
   return DH_DISPATCH(c->b, do_mindt, c, dt);
}

int dh_timestep(dh_ctx *c, double dt)
{
   if (!c) return fail(DH_EINVAL, "null context");
   if (!(dt > 0.0)) return fail(DH_EINVAL, "dt must be positive");
   return DH_DISPATCH(c->b, do_timestep, c, dt);
}

int dh_sync(dh_ctx *c)
{
   if (!c) return fail(DH_EINVAL, "null context");
   return c->x.sync();
}

int dh_add_diagnostics(dh_ctx *c, int kind, double dt, const double *planets, int npl)
{
   if (!c) return fail(DH_EINVAL, "null context");
   return add_diagnostics(c->x, c->b, kind, dt, planets, npl);
}

int dh_get_diagnostics(dh_ctx *c, double *Qrz)
{
   if (!c || !Qrz) return fail(DH_EINVAL, "null argument");
   return get_diagnostics(c->x, c->b, Qrz);
}

int dh_zero_diagnostics(dh_ctx *c)
{
   if (!c) return fail(DH_EINVAL, "null context");
   return zero_diagnostics(c->x, c->b);
}

int dh_avg_diagnostics(dh_ctx *c, double t_avg)
{
   if (!c) return fail(DH_EINVAL, "null context");
   return avg_diagnostics(c->x, c->b, t_avg);
}

int dh_set_c2p_model(dh_ctx *c, const float *params, long n)
{
   if (!c) return fail(DH_EINVAL, "null context");
   return set_c2p_model(c->x, c->b, params, n);
}

int dh_c2p_harvest(dh_ctx *c, float *feat, float *targ)
{
   if (!c) return fail(DH_EINVAL, "null context");
   return c2p_harvest(c->x, c->b, feat, targ);
}

int dh_c2p_predict(dh_ctx *c, float *guess)
{
   if (!c) return fail(DH_EINVAL, "null context");
   return c2p_predict(c->x, c->b, guess);
}


struct TrainArgs {
   const float *X, *Y;
   long n;
   int epochs, batch;
   float lr;
   unsigned long long seed;
   float *params, *loss;
};

struct GradArgs {
   const float *params, *X, *Y;
   long n;
   float *grad, *loss;
};

int dh_c2p_train(const float *X, const float *Y, long n, int epochs, int batch, float lr,
                 unsigned long long seed, float *params, float *loss)
{
   TrainArgs a{X, Y, n, epochs, batch, lr, seed, params, loss};
   return with_exec([](dh_ctx *c, void *p) {
      TrainArgs *t = (TrainArgs *)p;
      return c2p_train(c->x, t->X, t->Y, t->n, t->epochs, t->batch, t->lr, t->seed, t->params,
                       t->loss);
   }, &a);
}

int dh_c2p_grad(const float *params, const float *X, const float *Y, long n, float *grad,
                float *loss)
{
   GradArgs a{params, X, Y, n, grad, loss};
   return with_exec([](dh_ctx *c, void *p) {
      GradArgs *t = (GradArgs *)p;
      return c2p_grad(c->x, t->params, t->X, t->Y, t->n, t->grad, t->loss);
   }, &a);
}

int dh_c2p_grad_noalloc(const float *params, const float *X, const float *Y, long n, float *grad,
                        float *loss, void *stream)
{
   if (!params || !X || !Y || !grad || !loss) return fail(DH_EINVAL, "null argument");
   if (n < 1 || n > C2P_NOALLOC_ROWS)
      return fail(DH_EINVAL, "dh_c2p_grad_noalloc takes 1 to %ld rows (got %ld)",
                  C2P_NOALLOC_ROWS, n);
   float *s = nullptr;
   DH_TRY(scratch_of_current_device(&s));
   hipStream_t st = (hipStream_t)stream;
   MlpBatch B{params, X, Y, nullptr, (int)n, s, s + (size_t)n * C2PNET_ROW};
   k_run<KMlpRow><<<nblocks(n), TPB, 0, st>>>(KMlpRow{B}, n);
   LAUNCH_CHECK();
   k_run<KMlpGrad><<<nblocks(C2PNET_NTRAIN), TPB, 0, st>>>(KMlpGrad{B, grad + C2PNET_TRAIN0},
                                                           C2PNET_NTRAIN);
   LAUNCH_CHECK();
   long nframe = C2PNET_NPARAM - C2PNET_NTRAIN;
   k_run<KGradFrame><<<nblocks(nframe), TPB, 0, st>>>(KGradFrame{grad}, nframe);
   LAUNCH_CHECK();
   k_run<KLossMean><<<1, TPB, 0, st>>>(KLossMean{B.loss, n, loss}, 1);
   LAUNCH_CHECK();
   return DH_OK;
}

int dh_ens_create(dh_ctx *const *members, int n, dh_ens **out)
{
   if (!members || n < 1 || !out) return fail(DH_EINVAL, "bad arguments");
   for (int m = 0; m < n; ++m) {
      if (!members[m]) return fail(DH_EINVAL, "null member %d", m);
      for (int k = 0; k < m; ++k)
         if (members[k] == members[m]) return fail(DH_EINVAL, "member %d appears twice", m);
      const Buffers &a = members[0]->b, &b = members[m]->b;
      if (b.P.hydro != a.P.hydro || b.nq != a.nq || b.nf != a.nf)
         return fail(DH_EINVAL, "member %d has another hydro module or field count", m);
   }
   dh_ens *e = new dh_ens();
   e->mem.assign(members, members + n);
   e->nm = n;
   e->stream = members[0]->x.stream;
   for (int m = 0; m < n; ++m) {
      dh_ctx *c = members[m];
      e->own_stream.push_back(c->x.stream);
      if (hipStreamSynchronize(c->x.stream) != hipSuccess) {
         ens_release(e);
         return fail(DH_EHIP, "member %d: stream synchronisation failed", m);
      }
      c->x.drop_graph();
      c->x.stream = e->stream;
   }
   std::vector<StepArgs *> dst(n);
   for (int m = 0; m < n; ++m) dst[m] = members[m]->b.step;
   void *p = nullptr;
   auto dev = [&](void **q, size_t bytes) -> int {
      if (hipMalloc(q, bytes) != hipSuccess) return fail(DH_ENOMEM, "hipMalloc(%zu) failed", bytes);
      e->dev.push_back(*q);
      return DH_OK;
   };
   auto pin = [&](void **q, size_t bytes) -> int {
      if (hipHostMalloc(q, bytes, 0) != hipSuccess)
         return fail(DH_ENOMEM, "hipHostMalloc(%zu) failed", bytes);
      e->pinned.push_back(*q);
      return DH_OK;
   };
   int rc = dev(&p, sizeof(StepArgs *) * n);
   e->d_dst = (StepArgs **)p;
   if (!rc) rc = dev((void **)&e->d_step, sizeof(StepArgs) * n);
   if (!rc) rc = pin((void **)&e->h_step, sizeof(StepArgs) * n);
   if (!rc) rc = dev((void **)&e->d_min, sizeof(unsigned long long) * n);
   if (!rc) rc = pin((void **)&e->h_min, sizeof(unsigned long long) * n);
   if (!rc) rc = dev((void **)&e->d_diag, sizeof(EnsDiagArgs) * n);
   if (!rc) rc = pin((void **)&e->h_diag, sizeof(EnsDiagArgs) * n);
   if (!rc && hipMemcpy(e->d_dst, dst.data(), sizeof(StepArgs *) * n, hipMemcpyHostToDevice) !=
                  hipSuccess)
      rc = fail(DH_EHIP, "hipMemcpy of the step table failed");
   if (rc) {
      ens_release(e);
      return rc;
   }
   *out = e;
   return DH_OK;
}

int dh_ens_destroy(dh_ens *e)
{
   ens_release(e);
   return DH_OK;
}

int dh_ens_enable_graphs(dh_ens *e, int on)
{
   if (!e) return fail(DH_EINVAL, "null ensemble");
   e->graphs = on != 0;
   if (!e->graphs) ens_drop_graph(e);
   return DH_OK;
}

int dh_ens_sync(dh_ens *e)
{
   if (!e) return fail(DH_EINVAL, "null ensemble");
   HIPCHK(hipStreamSynchronize(e->stream));
   return DH_OK;
}

int dh_ens_getmindt(dh_ens *e, double *dts)
{
   if (!e || !dts) return fail(DH_EINVAL, "null argument");
   std::vector<GroupExec> rec(e->nm);
   for (int m = 0; m < e->nm; ++m) {
      Buffers &b = e->mem[m]->b;
      DH_TRY(DH_DISPATCH(b, ens_record_mindt, &rec[m], &b, e->d_min + m));
   }
   bool changed = false;
   DH_TRY(ens_build_launch(rec, 0, e->mindt, &changed));
   k_fill_u64<<<(e->nm + 63) / 64, 64, 0, e->stream>>>(e->d_min, e->nm, 0x7FF0000000000000ULL);
   LAUNCH_CHECK();
   DH_TRY(ens_launch(e, e->mindt));
   HIPCHK(hipMemcpyAsync(e->h_min, e->d_min, sizeof(unsigned long long) * e->nm,
                         hipMemcpyDeviceToHost, e->stream));
   HIPCHK(hipStreamSynchronize(e->stream));
   for (int m = 0; m < e->nm; ++m) {
      const dh_params &P = e->mem[m]->b.P;
      double v;
      memcpy(&v, &e->h_min[m], sizeof v);
      double d0 = P.max_dt / P.cfl;  // as Engine::mindt
      if (d0 <= 0.0) d0 = 1.0e100;
      if (v < d0) d0 = v;
      dts[m] = d0 * P.cfl;
   }
   return DH_OK;
}

int dh_ens_timestep(dh_ens *e, const double *dts)
{
   if (!e || !dts) return fail(DH_EINVAL, "null argument");
   for (int m = 0; m < e->nm; ++m)
      if (!(dts[m] > 0.0)) return fail(DH_EINVAL, "dt of member %d must be positive", m);
   std::vector<GroupExec> rec(e->nm);
   for (int m = 0; m < e->nm; ++m) {
      Buffers &b = e->mem[m]->b;
      DH_TRY(DH_DISPATCH(b, ens_record_step, &rec[m], &b, dts[m]));
      if (rec[m].step_dst != b.step) return fail(DH_EINVAL, "member %d recorded no step arguments", m);
   }
   DH_TRY(ens_check_records(rec));
   size_t nops = rec[0].ops.size();
   bool same = e->plan.size() == nops;
   while (e->plan.size() > nops) {
      ens_free_launch(e->plan.back());
      e->plan.pop_back();
   }
   e->plan.resize(nops);
   for (size_t i = 0; i < nops; ++i) {
      bool changed = false;
      DH_TRY(ens_build_launch(rec, i, e->plan[i], &changed));
      if (changed) same = false;
   }
   if (!same) ens_drop_graph(e);
   // the previous step's upload has left the pinned buffer (the stream
   // runs in order; getmindt and sync wait for it)
   HIPCHK(hipStreamSynchronize(e->stream));
   for (int m = 0; m < e->nm; ++m) e->h_step[m] = rec[m].step;
   HIPCHK(hipMemcpyAsync(e->d_step, e->h_step, sizeof(StepArgs) * e->nm, hipMemcpyHostToDevice,
                         e->stream));
   k_set_steps<<<(e->nm + 63) / 64, 64, 0, e->stream>>>(e->d_dst, e->d_step, e->nm);
   LAUNCH_CHECK();
   if (e->graphs && same) {
      if (!e->gexec) {
         HIPCHK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
         int rc = DH_OK;
         for (const EnsLaunch &L : e->plan)
            if (!rc) rc = ens_launch(e, L);
         hipGraph_t g = nullptr;
         hipError_t er = hipStreamEndCapture(e->stream, &g);
         if (rc || er != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return rc ? rc : fail(DH_EHIP, "ensemble capture failed: %s", hipGetErrorString(er));
         }
         er = hipGraphInstantiateWithFlags(&e->gexec, g, 0);
         (void)hipGraphDestroy(g);
         if (er != hipSuccess) {
            e->gexec = nullptr;
            return fail(DH_EHIP, "ensemble graph instantiation failed: %s", hipGetErrorString(er));
         }
      }
      HIPCHK(hipGraphLaunch(e->gexec, e->stream));
      return DH_OK;
   }
   for (const EnsLaunch &L : e->plan) DH_TRY(ens_launch(e, L));
   return DH_OK;
}

int dh_ens_add_diagnostics(dh_ens *e, int kind, const double *dts, const double *planets, int npl)
{
   if (!e || !dts) return fail(DH_EINVAL, "null argument");
   if (!diag_count(kind)) return fail(DH_EINVAL, "unknown diagnostics kind %d", kind);
   if (npl < 0 || npl > DH_MAX_PLANETS) return fail(DH_EINVAL, "npl %d out of range", npl);
   if (npl > 1 && !planets) return fail(DH_EINVAL, "null planet list");
   HIPCHK(hipStreamSynchronize(e->stream));  // h_diag is free again
   std::vector<GroupExec> rec(e->nm);
   for (int m = 0; m < e->nm; ++m) {
      dh_ctx *c = e->mem[m];
      Buffers &b = c->b;
      if (kind == DH_DIAG_MHDMRI && !b.mhd)
         return fail(DH_EINVAL, "diag_mhdmri needs the MHD fields (NUM_C 8)");
      size_t bytes = sizeof(double) * (size_t)b.m.nann * DH_DIAG_MAXQ;
      if (!b.diagQ) {
         DH_TRY(c->x.alloc((void **)&b.diagQ, bytes));
         DH_TRY(c->x.zero(b.diagQ, bytes));
      } else if (b.diag_kind != kind) {
         DH_TRY(c->x.zero(b.diagQ, bytes));
      }
      b.diag_kind = kind;
      EnsDiagArgs &a = e->h_diag[m];
      a.dt = dts[m];
      a.p1 = DiagPlanet{};
      a.p1.on = npl > 1;
      if (a.p1.on)
         for (int q = 0; q < DH_PLANET_STRIDE; ++q)
            a.p1.v[q] = planets[(size_t)m * DH_MAX_PLANETS * DH_PLANET_STRIDE + DH_PLANET_STRIDE + q];
      const EnsDiagArgs *da = e->d_diag + m;
      int rc = kind == DH_DIAG_EULER    ? ens_record_diag<DH_DIAG_EULER>(&rec[m], &b, da)
               : kind == DH_DIAG_DISK2D ? ens_record_diag<DH_DIAG_DISK2D>(&rec[m], &b, da)
                                        : ens_record_diag<DH_DIAG_MHDMRI>(&rec[m], &b, da);
      DH_TRY(rc);
   }
   bool changed = false;
   DH_TRY(ens_build_launch(rec, 0, e->diag, &changed));
   e->diag_kind = kind;
   HIPCHK(hipMemcpyAsync(e->d_diag, e->h_diag, sizeof(EnsDiagArgs) * e->nm, hipMemcpyHostToDevice,
                         e->stream));
   return ens_launch(e, e->diag);
}

int dh_c2p_stats(dh_ctx *c, long long *out)
{
   if (!c) return fail(DH_EINVAL, "null context");
   return c2p_stats(c->x, c->b, out);
}

int dh_c2p_info(dh_ctx *c, unsigned char *info)
{
   if (!c) return fail(DH_EINVAL, "null context");
   return c2p_info(c->x, c->b, info);
}

int dh_enable_timers(dh_ctx *c, int on)
{
   if (!c) return fail(DH_EINVAL, "null context");
   c->x.timers = on != 0;
   return DH_OK;
}

int dh_report(dh_ctx *c, const double *planets, int npl, double *out)
{
   if (!c || !out) return fail(DH_EINVAL, "null argument");
   return report(c->x, c->b, planets, npl, out);
}

int dh_enable_graphs(dh_ctx *c, int on)
{
   if (!c) return fail(DH_EINVAL, "null context");
   c->x.graphs = on != 0;
   if (!c->x.graphs) {
      HIPCHK(hipStreamSynchronize(c->x.stream));
      c->x.drop_graph();
   }
   return DH_OK;
}

int dh_autotune(dh_ctx *c, int steps)
{
   if (!c) return fail(DH_EINVAL, "null context");
   if (steps < 0) return fail(DH_EINVAL, "negative step count");
   c->x.drop_graph();
   c->x.tune_left = steps;
   c->x.tune_step = 0;
   if (steps > 0)
      for (auto &t : c->x.tunes) {
         for (int v = 0; v < t.nvar; ++v) t.ms[v] = 0.0, t.n[v] = 0;
      }
   return DH_OK;
}

int dh_read_tuning(dh_ctx *c, int cap, char (*names)[48], int *variant, int *strip,
                   double *ms_chosen, double *ms_default, int *nrows)
{
   if (!c || cap < 0 || !nrows) return fail(DH_EINVAL, "bad arguments");
   int n = 0;
   for (auto &t : c->x.tunes) {
      if (n >= cap) break;
      int v = t.chosen >= 0 ? t.chosen : t.dflt;
      if (names) {
         snprintf(names[n], 48, "%s", t.name.c_str());
      }
      if (variant) variant[n] = v;
      if (strip) strip[n] = v >= 2 ? t.strip[v - 2] : 0;
      if (ms_chosen) ms_chosen[n] = t.n[v] ? t.ms[v] / t.n[v] : -1.0;
      if (ms_default) ms_default[n] = t.n[t.dflt] ? t.ms[t.dflt] / t.n[t.dflt] : -1.0;
      ++n;
   }
   *nrows = n;
   return DH_OK;
}

int dh_read_tuning_sweep(dh_ctx *c, const char *name, int cap, int *strip, double *ms,
                         int *nvar)
{
   if (!c || !name || cap < 0 || !nvar) return fail(DH_EINVAL, "bad arguments");
   *nvar = 0;
   for (auto &t : c->x.tunes) {
      if (t.name != name) continue;
      *nvar = t.nvar;
      for (int v = 0; v < t.nvar && v < cap; ++v) {
         if (strip) strip[v] = v >= 2 ? t.strip[v - 2] : 0;
         if (ms) ms[v] = t.n[v] ? t.ms[v] / t.n[v] : -1.0;
      }
      break;
   }
   return DH_OK;
}

int dh_read_timers(dh_ctx *c, int cap, char (*names)[48], double *ms, long *calls, int *nrows)
{
   if (!c) return fail(DH_EINVAL, "null context");
   DH_TRY(c->x.sync());
   int n = 0;
   for (auto &row : c->x.rows) {
      for (auto &pe : row.pending) {
         float t = 0.f;
         (void)hipEventElapsedTime(&t, pe.first, pe.second);
         row.ms += t;
         row.calls++;
         (void)hipEventDestroy(pe.first);
         (void)hipEventDestroy(pe.second);
      }
      row.pending.clear();
      if (n < cap) {
         snprintf(names[n], 48, "%s", row.name.c_str());
         ms[n] = row.ms;
         calls[n] = row.calls;
      }
      n++;
   }
   if (nrows) *nrows = n < cap ? n : cap;
   return DH_OK;
}

}  // extern "C"
