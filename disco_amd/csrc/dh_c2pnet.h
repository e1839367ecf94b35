// dh_c2pnet.h -- the learned initial guess of grmhd2's cons2prim.
//
// noble2d (grmhd2.c:1096-1314) solves two equations for (v^2, eta = W h)
// whose coefficients are four invariants of the conserved state (s2, e, Q,
// psi; C2pInv) and the adiabatic index.  A small MLP, 8 -> 64 -> 64 -> 2
// with ReLU, maps features of those invariants to the solution, and the
// Newton solve starts there instead of at the previous primitives.  This has
// no counterpart in the reference; cons2prim falls back to the reference's
// guess wherever the model's output is unphysical or the solve from it does
// not converge, so the model can change the iteration count, not the answer.
//
// Arithmetic: weights and activations in bfloat16, sums in float32 (MFMA on
// the device, k_c2p_guess; the host build rounds the same way and sums in
// order, so the two differ only by float32 summation order).
#pragma once
#include <math.h>

#include "dh_grmhd.h"

namespace dh {

constexpr int C2PNET_IN = 8, C2PNET_H = 64, C2PNET_OUT = 2;

// The float32 parameter block of dh_set_c2p_model, in this order:
//   in_mu[8] in_sd[8]  W1[8][64] b1[64]  W2[64][64] b2[64]  W3[64][2] b3[2]
//   out_mu[2] out_sd[2]
// W[k][j] multiplies input k into output j.  A layer input is normalised as
// (f - in_mu) / in_sd and the output denormalised as y * out_sd + out_mu.
enum {
   C2PNET_IN_MU = 0,
   C2PNET_IN_SD = C2PNET_IN_MU + C2PNET_IN,
   C2PNET_W1 = C2PNET_IN_SD + C2PNET_IN,
   C2PNET_B1 = C2PNET_W1 + C2PNET_IN * C2PNET_H,
   C2PNET_W2 = C2PNET_B1 + C2PNET_H,
   C2PNET_B2 = C2PNET_W2 + C2PNET_H * C2PNET_H,
   C2PNET_W3 = C2PNET_B2 + C2PNET_H,
   C2PNET_B3 = C2PNET_W3 + C2PNET_H * C2PNET_OUT,
   C2PNET_OUT_MU = C2PNET_B3 + C2PNET_OUT,
   C2PNET_OUT_SD = C2PNET_OUT_MU + C2PNET_OUT,
   C2PNET_NPARAM = C2PNET_OUT_SD + C2PNET_OUT
};

// Model features of a state: scale-free combinations of the invariants
// (in the hydrodynamic limit eta ~ e and v^2 ~ s2 / e^2) and their
// logarithms.  A state with e <= 0 gives non-finite features, hence a
// rejected guess.
DHD void c2p_features(const C2pInv &v, double gamma_law, float *f)
{
   double e = v.e, s2 = v.s2, Q = v.Q, psi = v.psi;
   f[0] = (float)log(e);
   f[1] = (float)(s2 / (e * e));
   f[2] = (float)(Q / e);
   f[3] = (float)(psi / (e * e * e));
   f[4] = (float)((gamma_law - 1.0) / gamma_law);
   f[5] = (float)log1p(s2);
   f[6] = (float)log1p(Q);
   f[7] = (float)log1p(psi);
}

// The model's targets: t0 = v^2, t1 = log(eta / e).
DHD void c2p_targets_to_guess(const float *t, double e, float *g)
{
   g[0] = t[0];
   g[1] = (float)(e * exp((double)t[1]));
}

// float32 -> bfloat16 with round-to-nearest-even (finite inputs).
DHD unsigned short bf16_bits(float x)
{
   union {
      float f;
      unsigned u;
   } b;
   b.f = x;
   unsigned r = b.u + 0x7fffu + ((b.u >> 16) & 1u);
   return (unsigned short)(r >> 16);
}

DHD float bf16_value(unsigned short h)
{
   union {
      float f;
      unsigned u;
   } b;
   b.u = (unsigned)h << 16;
   return b.f;
}

DHD float bf16_round(float x) { return bf16_value(bf16_bits(x)); }

// The weights as bfloat16 B operands of the 16x16x16 MFMA products of
// k_c2p_guess.  Fragment (kb, nb) of a layer holds, for lane l and item i,
// W[16 kb + 4 (l / 16) + i][16 nb + l % 16] (zero outside the layer): the B
// operand's register layout.  Layer 1 (K 8 padded to 16, N 64): fragments
// 0..3 (nb); layer 2 (K 64, N 64): 4 + 4 kb + nb; layer 3 (K 64, N 2
// padded to 16): 20 + kb.
constexpr int C2PNET_FRAG_L1 = 0, C2PNET_FRAG_L2 = 4, C2PNET_FRAG_L3 = 20, C2PNET_NFRAGS = 24;
constexpr int C2PNET_NFRAG = C2PNET_NFRAGS * 64 * 4;  // bfloat16 values

inline void c2pnet_fragments(const float *prm, unsigned short *frag)
{
   auto fill = [&](int id, int kb, int nb, int off, int K, int N) {
      for (int l = 0; l < 64; ++l)
         for (int i = 0; i < 4; ++i) {
            int k = 16 * kb + 4 * (l / 16) + i, j = 16 * nb + l % 16;
            float w = (k < K && j < N) ? prm[off + k * N + j] : 0.0f;
            frag[(id * 64 + l) * 4 + i] = bf16_bits(w);
         }
   };
   for (int nb = 0; nb < 4; ++nb) fill(C2PNET_FRAG_L1 + nb, 0, nb, C2PNET_W1, C2PNET_IN, C2PNET_H);
   for (int kb = 0; kb < 4; ++kb)
      for (int nb = 0; nb < 4; ++nb)
         fill(C2PNET_FRAG_L2 + 4 * kb + nb, kb, nb, C2PNET_W2, C2PNET_H, C2PNET_H);
   for (int kb = 0; kb < 4; ++kb) fill(C2PNET_FRAG_L3 + kb, kb, 0, C2PNET_W3, C2PNET_H, C2PNET_OUT);
}

// The model on one state, rounding as the device does: normalised
// features, weights and hidden activations in bfloat16; sums, biases and
// the output in float32.  f: features, t: targets.
inline void c2pnet_eval_host(const float *prm, const float *f, float *t)
{
   float x[C2PNET_IN], h1[C2PNET_H], h2[C2PNET_H];
   for (int k = 0; k < C2PNET_IN; ++k)
      x[k] = bf16_round((f[k] - prm[C2PNET_IN_MU + k]) / prm[C2PNET_IN_SD + k]);
   for (int j = 0; j < C2PNET_H; ++j) {
      float s = 0.0f;
      for (int k = 0; k < C2PNET_IN; ++k) s += x[k] * bf16_round(prm[C2PNET_W1 + k * C2PNET_H + j]);
      s += prm[C2PNET_B1 + j];
      h1[j] = bf16_round(s > 0.0f ? s : 0.0f);
   }
   for (int j = 0; j < C2PNET_H; ++j) {
      float s = 0.0f;
      for (int k = 0; k < C2PNET_H; ++k) s += h1[k] * bf16_round(prm[C2PNET_W2 + k * C2PNET_H + j]);
      s += prm[C2PNET_B2 + j];
      h2[j] = bf16_round(s > 0.0f ? s : 0.0f);
   }
   for (int j = 0; j < C2PNET_OUT; ++j) {
      float s = 0.0f;
      for (int k = 0; k < C2PNET_H; ++k) s += h2[k] * bf16_round(prm[C2PNET_W3 + k * C2PNET_OUT + j]);
      s += prm[C2PNET_B3 + j];
      t[j] = s * prm[C2PNET_OUT_SD + j] + prm[C2PNET_OUT_MU + j];
   }
}

}  // namespace dh
