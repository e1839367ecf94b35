// dh_c2ptrain.h -- training of the learned cons2prim guess (dh_c2pnet.h) on
// the device: the MLP's forward and backward passes, the parameter
// gradients and Adam, in float32.
//
// Loss: the mean over a batch of nb rows and the two outputs of the squared
// error between the model's output and the normalised targets,
// L = 1 / (2 nb) sum_b sum_o (y_bo - t_bo)^2, t = (Y - out_mu) / out_sd; the
// inputs are normalised with in_mu, in_sd.  Per batch:
//   KMlpRow   one thread per row: forward (z1, h1 = relu z1, z2, h2, y) and
//             backward (D3 = dL/dy, D2 = dL/dz2, D1 = dL/dz1), stored per row;
//   KMlpGrad  one thread per trainable parameter: its gradient summed over
//             the rows in row order (dW[k][j] = sum_b A[b][k] D[b][j], A the
//             layer's input), so the sum is deterministic without atomics;
//   KAdam     one thread per trainable parameter (Kingma & Ba 2015, with
//             bias correction).
// ReLU's derivative at 0 is taken as 0.
#pragma once
#include <math.h>

#include "dh_c2pnet.h"

namespace dh {

// The trainable parameters are the contiguous block [C2PNET_W1, C2PNET_OUT_MU)
// of the parameter block; the normalisation constants around it are fixed.
constexpr int C2PNET_TRAIN0 = C2PNET_W1, C2PNET_NTRAIN = C2PNET_OUT_MU - C2PNET_W1;
constexpr int C2PNET_ROW = 4 * C2PNET_H + C2PNET_OUT;  // stored floats per batch row

struct MlpBatch {
   const float *prm;  // parameter block
   const float *X;    // features [n][8]
   const float *Y;    // targets [n][2]
   const int *idx;    // rows of this batch: sample indices (null: rows 0..nb-1)
   int nb;            // rows in the batch
   float *act;        // per row: H1[64] H2[64] D1[64] D2[64] D3[2]
   float *loss;       // per row: squared error summed over the outputs
};

struct KMlpRow {
   MlpBatch B;
   DHD void operator()(long b) const
   {
      const float *p = B.prm;
      int i = B.idx ? B.idx[b] : (int)b;
      float x[C2PNET_IN], h1[C2PNET_H], h2[C2PNET_H];
      for (int k = 0; k < C2PNET_IN; ++k)
         x[k] = (B.X[(size_t)i * C2PNET_IN + k] - p[C2PNET_IN_MU + k]) / p[C2PNET_IN_SD + k];
      for (int j = 0; j < C2PNET_H; ++j) {
         float s = p[C2PNET_B1 + j];
         for (int k = 0; k < C2PNET_IN; ++k) s += x[k] * p[C2PNET_W1 + k * C2PNET_H + j];
         h1[j] = s > 0.0f ? s : 0.0f;
      }
      for (int j = 0; j < C2PNET_H; ++j) {
         float s = p[C2PNET_B2 + j];
         for (int k = 0; k < C2PNET_H; ++k) s += h1[k] * p[C2PNET_W2 + k * C2PNET_H + j];
         h2[j] = s > 0.0f ? s : 0.0f;
      }
      float *a = B.act + (size_t)b * C2PNET_ROW;
      float *H1 = a, *H2 = a + C2PNET_H, *D1 = a + 2 * C2PNET_H, *D2 = a + 3 * C2PNET_H,
            *D3 = a + 4 * C2PNET_H;
      for (int j = 0; j < C2PNET_H; ++j) {
         H1[j] = h1[j];
         H2[j] = h2[j];
      }
      float d3[C2PNET_OUT], l = 0.0f;
      for (int o = 0; o < C2PNET_OUT; ++o) {
         float y = p[C2PNET_B3 + o];
         for (int k = 0; k < C2PNET_H; ++k) y += h2[k] * p[C2PNET_W3 + k * C2PNET_OUT + o];
         float t = (B.Y[(size_t)i * C2PNET_OUT + o] - p[C2PNET_OUT_MU + o]) / p[C2PNET_OUT_SD + o];
         float e = y - t;
         l += e * e;
         d3[o] = 2.0f * e / (float)(C2PNET_OUT * B.nb);
         D3[o] = d3[o];
      }
      B.loss[b] = l;
      // h2 becomes dL/dz2 in place
      for (int j = 0; j < C2PNET_H; ++j) {
         float s = 0.0f;
         for (int o = 0; o < C2PNET_OUT; ++o) s += d3[o] * p[C2PNET_W3 + j * C2PNET_OUT + o];
         h2[j] = h2[j] > 0.0f ? s : 0.0f;
         D2[j] = h2[j];
      }
      for (int k = 0; k < C2PNET_H; ++k) {
         float s = 0.0f;
         for (int j = 0; j < C2PNET_H; ++j) s += h2[j] * p[C2PNET_W2 + k * C2PNET_H + j];
         D1[k] = h1[k] > 0.0f ? s : 0.0f;
      }
   }
};

// The gradient of trainable parameter t (offset C2PNET_TRAIN0 + t of the
// block) over the batch.
struct KMlpGrad {
   MlpBatch B;
   float *grad;  // [C2PNET_NTRAIN]
   DHD void operator()(long t) const
   {
      const float *p = B.prm;
      int q = C2PNET_TRAIN0 + (int)t;
      int in = -1, col = 0, dof = 0, aof = 0;  // input k (-1: bias), output j, D and A row offsets
      if (q < C2PNET_B1) {
         in = (q - C2PNET_W1) / C2PNET_H;
         col = (q - C2PNET_W1) % C2PNET_H;
         dof = 2 * C2PNET_H;
         aof = -1;  // the normalised features
      } else if (q < C2PNET_W2) {
         col = q - C2PNET_B1;
         dof = 2 * C2PNET_H;
      } else if (q < C2PNET_B2) {
         in = (q - C2PNET_W2) / C2PNET_H;
         col = (q - C2PNET_W2) % C2PNET_H;
         dof = 3 * C2PNET_H;
         aof = 0;
      } else if (q < C2PNET_W3) {
         col = q - C2PNET_B2;
         dof = 3 * C2PNET_H;
      } else if (q < C2PNET_B3) {
         in = (q - C2PNET_W3) / C2PNET_OUT;
         col = (q - C2PNET_W3) % C2PNET_OUT;
         dof = 4 * C2PNET_H;
         aof = C2PNET_H;
      } else {
         col = q - C2PNET_B3;
         dof = 4 * C2PNET_H;
      }
      float s = 0.0f;
      for (int b = 0; b < B.nb; ++b) {
         const float *a = B.act + (size_t)b * C2PNET_ROW;
         float d = a[dof + col];
         if (in < 0) {
            s += d;
         } else if (aof < 0) {
            int i = B.idx ? B.idx[b] : b;
            float x = (B.X[(size_t)i * C2PNET_IN + in] - p[C2PNET_IN_MU + in]) / p[C2PNET_IN_SD + in];
            s += x * d;
         } else {
            s += a[aof + in] * d;
         }
      }
      grad[t] = s;
   }
};

struct KAdam {
   float *prm, *m, *v;  // prm: the parameter block; m, v: [C2PNET_NTRAIN]
   const float *grad;
   float lr, b1, b2, eps, c1, c2;  // c1 = 1 - b1^step, c2 = 1 - b2^step
   DHD void operator()(long t) const
   {
      float g = grad[t];
      float mt = b1 * m[t] + (1.0f - b1) * g;
      float vt = b2 * v[t] + (1.0f - b2) * g * g;
      m[t] = mt;
      v[t] = vt;
      prm[C2PNET_TRAIN0 + t] -= lr * (mt / c1) / (sqrtf(vt / c2) + eps);
   }
};

// The entries of a gradient block outside the trainable range (the
// normalisation constants), cleared: thread q < C2PNET_TRAIN0 takes entry q,
// the others the entries after the trainable range.
struct KGradFrame {
   float *grad;  // [C2PNET_NPARAM]
   DHD void operator()(long q) const
   {
      grad[q < C2PNET_TRAIN0 ? q : q + C2PNET_NTRAIN] = 0.0f;
   }
};

// The mean loss of n rows, L = sum of the rows' squared errors / (2 n), summed
// in double in row order (one thread).
struct KLossMean {
   const float *loss;
   long n;
   float *out;
   DHD void operator()(long) const
   {
      double s = 0.0;
      for (long b = 0; b < n; ++b) s += loss[b];
      *out = (float)(s / (2.0 * n));
   }
};

// Sum of a batch's row losses into one slot (a single thread, in order).
struct KLossSum {
   const float *loss;
   int nb;
   float *out;
   DHD void operator()(long) const
   {
      float s = 0.0f;
      for (int b = 0; b < nb; ++b) s += loss[b];
      *out += s;
   }
};

}  // namespace dh
