// BF16 MFMA kernels of the per-cell collision surrogate network.
//
// Inference keeps the weights in registers as MFMA A operands and the
// samples as the N dimension, so the activations pass from layer to layer
// in registers (k_mlp_forward).  Training: one wavefront owns a tile of 32
// samples (two 16-row MFMA tiles) and runs all three layers on it, staging
// activations in LDS between layers so that each layer reads its A operand
// in the MFMA row layout.  The instruction is
// v_mfma_f32_16x16x32_bf16: per lane l, A holds row (l & 15), k = 8 (l >> 4)
// + 0..7; B holds column (l & 15) with the same k; D holds column (l & 15),
// rows 4 (l >> 4) + 0..3.  Biases enter through the FP32 accumulator.
//
// Training: the same tile computes the forward pass, the MSE error and the
// backward pass; weight gradients are MFMA products with the samples as the
// contraction dimension, accumulated in registers over the wavefront's tiles
// and written as per-block partial sums, which a second kernel adds in a
// fixed order (deterministic gradients).  Adam updates the FP32 master copy.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <random>
#include <vector>

#include "pcp_mlp.hpp"

namespace pcp {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int WAVE = 64;

__device__ __forceinline__ f32x4 mfma(const bf16x8 &a, const bf16x8 &b, const f32x4 &c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// LDS tile of one wavefront: rows = samples
struct MlpTile {
  float x[MLP_TILE][MLP_INP];
  float h1[MLP_TILE][MLP_HID];
  float h2[MLP_TILE][MLP_HID];
};
struct MlpTrainTile {
  MlpTile f;
  float dy[MLP_TILE][MLP_OUTP];
  float dz2[MLP_TILE][MLP_HID];
  float dz1[MLP_TILE][MLP_HID];
};

// A operand from an LDS matrix M[rows][ld]: row r0 + (l & 15), columns
// c0 + 8 (l >> 4) + i
template <int LD>
__device__ __forceinline__ bf16x8 a_rows(const float (*M)[LD], int r0, int c0) {
  const int l = threadIdx.x & 63;
  const float *p = &M[r0 + (l & 15)][c0 + 8 * (l >> 4)];
  bf16x8 a;
#pragma unroll
  for (int i = 0; i < 8; ++i) a[i] = (__bf16)p[i];
  return a;
}
// A operand from the transpose of an LDS matrix: element (row m, k) = M[k][m]
// with m = m0 + (l & 15), k = k0 + 8 (l >> 4) + i
template <int LD>
__device__ __forceinline__ bf16x8 a_cols(const float (*M)[LD], int m0, int k0) {
  const int l = threadIdx.x & 63;
  bf16x8 a;
#pragma unroll
  for (int i = 0; i < 8; ++i) a[i] = (__bf16)M[k0 + 8 * (l >> 4) + i][m0 + (l & 15)];
  return a;
}
// B operand (k, n) = W[n][k] from a row-major global weight matrix W[.][LD]:
// n = n0 + (l & 15), k = k0 + 8 (l >> 4) + i; rows >= nmax read as zero
template <int LD>
__device__ __forceinline__ bf16x8 b_wrows(const float *W, int n0, int k0, int nmax) {
  const int l = threadIdx.x & 63;
  const int n = n0 + (l & 15);
  bf16x8 b;
#pragma unroll
  for (int i = 0; i < 8; ++i) b[i] = (__bf16)(n < nmax ? W[n * LD + k0 + 8 * (l >> 4) + i] : 0.f);
  return b;
}
// B operand (k, n) = W[k][n] (no transpose): k = k0 + 8 (l >> 4) + i, rows
// k >= kmax read as zero
template <int LD>
__device__ __forceinline__ bf16x8 b_wcols(const float *W, int n0, int k0, int kmax) {
  const int l = threadIdx.x & 63;
  const int n = n0 + (l & 15);
  bf16x8 b;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    int k = k0 + 8 * (l >> 4) + i;
    b[i] = (__bf16)(k < kmax ? W[k * LD + n] : 0.f);
  }
  return b;
}
// B operand (k = sample, n) from an LDS matrix M[sample][ld]
template <int LD>
__device__ __forceinline__ bf16x8 b_lds(const float (*M)[LD], int n0, int k0) {
  const int l = threadIdx.x & 63;
  bf16x8 b;
#pragma unroll
  for (int i = 0; i < 8; ++i) b[i] = (__bf16)M[k0 + 8 * (l >> 4) + i][n0 + (l & 15)];
  return b;
}

__device__ __forceinline__ f32x4 splat(float v) { return f32x4{v, v, v, v}; }

// forward pass of the tile in LDS (x filled); writes h1, h2 and returns the
// outputs of the two row tiles (D layout, columns = outputs)
__device__ void tile_forward(const float *P, MlpTile &T, f32x4 y[2]) {
  const int l = threadIdx.x & 63;
  const float *W1 = P + MLP_OFF_W1, *b1 = P + MLP_OFF_B1;
  const float *W2 = P + MLP_OFF_W2, *b2 = P + MLP_OFF_B2;
  const float *W3 = P + MLP_OFF_W3, *b3 = P + MLP_OFF_B3;
  // layer 1: Z1 = X W1^T + b1
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    const bf16x8 b = b_wrows<MLP_INP>(W1, 16 * nt, 0, MLP_HID);
    const float bias = b1[16 * nt + (l & 15)];
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
      f32x4 d = mfma(a_rows<MLP_INP>(T.x, 16 * mt, 0), b, splat(bias));
#pragma unroll
      for (int i = 0; i < 4; ++i) T.h1[16 * mt + 4 * (l >> 4) + i][16 * nt + (l & 15)] = tanhf(d[i]);
    }
  }
  __syncthreads();
  // layer 2: Z2 = H1 W2^T + b2
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    const bf16x8 b0 = b_wrows<MLP_HID>(W2, 16 * nt, 0, MLP_HID);
    const bf16x8 b1v = b_wrows<MLP_HID>(W2, 16 * nt, 32, MLP_HID);
    const float bias = b2[16 * nt + (l & 15)];
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
      f32x4 d = mfma(a_rows<MLP_HID>(T.h1, 16 * mt, 0), b0, splat(bias));
      d = mfma(a_rows<MLP_HID>(T.h1, 16 * mt, 32), b1v, d);
#pragma unroll
      for (int i = 0; i < 4; ++i) T.h2[16 * mt + 4 * (l >> 4) + i][16 * nt + (l & 15)] = tanhf(d[i]);
    }
  }
  __syncthreads();
  // layer 3: Y = H2 W3^T + b3 (16 padded outputs, 4 used)
  {
    const bf16x8 b0 = b_wrows<MLP_HID>(W3, 0, 0, MLP_OUTP);
    const bf16x8 b1v = b_wrows<MLP_HID>(W3, 0, 32, MLP_OUTP);
    const float bias = b3[l & 15];
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
      f32x4 d = mfma(a_rows<MLP_HID>(T.h2, 16 * mt, 0), b0, splat(bias));
      y[mt] = mfma(a_rows<MLP_HID>(T.h2, 16 * mt, 32), b1v, d);
    }
  }
}

__device__ __forceinline__ void load_x(MlpTile &T, const float *X, const long *idx, long base, long n) {
  for (int t = threadIdx.x; t < MLP_TILE * MLP_INP; t += WAVE) {
    const int s = t / MLP_INP, k = t % MLP_INP;
    const long g = base + s;
    float v = 0.f;
    if (g < n && k < MLP_IN) v = X[(idx ? idx[g] : g) * MLP_XSTRIDE + k];
    T.x[s][k] = v;
  }
  __syncthreads();
}

// Inference, weights resident: each wavefront loads the three weight
// matrices once, as MFMA A operands (rows = a layer's outputs), and runs
// tiles of 16 samples as the N dimension, Z^T = W X^T + b.  A layer's D tiles
// (lane l: rows 4 g + 0..3 of each 16-row tile, g = l >> 4, column = sample
// l & 15) are the next layer's B operand without leaving the registers: K
// position 8 g + i of K-step ks stands for hidden unit hperm(ks, g, i), and
// the next layer's weight columns are loaded in that order.
__device__ __forceinline__ int hperm(int ks, int g, int i) { return 16 * (2 * ks + (i >> 2)) + 4 * g + (i & 3); }

// tanh(x) = 1 - 2 / (exp(2 x) + 1) with the hardware exponential and
// reciprocal: exact limits, absolute error about 1e-7 (below the BF16
// rounding the next layer applies)
__device__ __forceinline__ float tanh_fast(float x) {
  return 1.f - 2.f * __builtin_amdgcn_rcpf(__expf(2.f * x) + 1.f);
}

constexpr int MLP_FWD_BLOCK = 256;

__global__ void __launch_bounds__(MLP_FWD_BLOCK) k_mlp_forward(const float *P, const float *X, long n, float *Y) {
  const int l = threadIdx.x & 63, g = l >> 4, r = l & 15;
  const float *W1 = P + MLP_OFF_W1, *b1 = P + MLP_OFF_B1;
  const float *W2 = P + MLP_OFF_W2, *b2 = P + MLP_OFF_B2;
  const float *W3 = P + MLP_OFF_W3, *b3 = P + MLP_OFF_B3;
  bf16x8 a1[4], a2[4][2], a3[2];
  f32x4 c1[4], c2[4], c3;
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) {
#pragma unroll
    for (int i = 0; i < 8; ++i) a1[mt][i] = (__bf16)W1[(16 * mt + r) * MLP_INP + 8 * g + i];
#pragma unroll
    for (int ks = 0; ks < 2; ++ks)
#pragma unroll
      for (int i = 0; i < 8; ++i) a2[mt][ks][i] = (__bf16)W2[(16 * mt + r) * MLP_HID + hperm(ks, g, i)];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      c1[mt][i] = b1[16 * mt + 4 * g + i];
      c2[mt][i] = b2[16 * mt + 4 * g + i];
    }
  }
#pragma unroll
  for (int ks = 0; ks < 2; ++ks)
#pragma unroll
    for (int i = 0; i < 8; ++i) a3[ks][i] = (__bf16)W3[r * MLP_HID + hperm(ks, g, i)];
#pragma unroll
  for (int i = 0; i < 4; ++i) c3[i] = b3[4 * g + i];

  const long wave = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const long nwaves = ((long)gridDim.x * blockDim.x) >> 6;
  for (long t0 = wave * 16; t0 < n; t0 += nwaves * 16) {
    const long s = t0 + r;
    // X^T: lanes of group 0 hold the sample's inputs (padded with zeros)
    bf16x8 xb;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      float v = 0.f;
      if (g == 0 && i < MLP_IN && s < n) v = X[s * MLP_XSTRIDE + i];
      xb[i] = (__bf16)v;
    }
    f32x4 h[4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      h[mt] = mfma(a1[mt], xb, c1[mt]);
#pragma unroll
      for (int i = 0; i < 4; ++i) h[mt][i] = tanh_fast(h[mt][i]);
    }
    bf16x8 hb[2];
#pragma unroll
    for (int ks = 0; ks < 2; ++ks)
#pragma unroll
      for (int i = 0; i < 8; ++i) hb[ks][i] = (__bf16)h[2 * ks + (i >> 2)][i & 3];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      h[mt] = mfma(a2[mt][1], hb[1], mfma(a2[mt][0], hb[0], c2[mt]));
#pragma unroll
      for (int i = 0; i < 4; ++i) h[mt][i] = tanh_fast(h[mt][i]);
    }
#pragma unroll
    for (int ks = 0; ks < 2; ++ks)
#pragma unroll
      for (int i = 0; i < 8; ++i) hb[ks][i] = (__bf16)h[2 * ks + (i >> 2)][i & 3];
    const f32x4 y = mfma(a3[1], hb[1], mfma(a3[0], hb[0], c3));
    // outputs 0..3 are rows 0..3: lane group 0
    if (g == 0 && s < n) {
#pragma unroll
      for (int i = 0; i < MLP_OUT; ++i) Y[s * MLP_YSTRIDE + i] = y[i];
    }
  }
}

// forward + backward of MSE loss sum (y - t)^2 / n_total over the samples
// idx[0 .. n) (or 0 .. n when idx is null); per-block partial gradients
__global__ void __launch_bounds__(WAVE) k_mlp_train(const float *P, const float *X, const float *Tg, const long *idx,
                                                    long n, float inv_n, float *partials) {
  __shared__ MlpTrainTile S;
  const int l = threadIdx.x;
  const float *W2 = P + MLP_OFF_W2, *W3 = P + MLP_OFF_W3;
  f32x4 gW1[4], gW2[4][4], gW3[4];
  float gb1 = 0.f, gb2 = 0.f, gb3 = 0.f, loss = 0.f;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    gW1[i] = splat(0.f);
    gW3[i] = splat(0.f);
#pragma unroll
    for (int j = 0; j < 4; ++j) gW2[i][j] = splat(0.f);
  }
  for (long base = (long)blockIdx.x * MLP_TILE; base < n; base += (long)gridDim.x * MLP_TILE) {
    load_x(S.f, X, idx, base, n);
    f32x4 y[2];
    tile_forward(P, S.f, y);
    // output error (D layout: column = output o, rows = samples)
    const int o = l & 15;
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int s = 16 * mt + 4 * (l >> 4) + i;
        const long g = base + s;
        float e = 0.f;
        if (o < MLP_OUT && g < n) {
          e = y[mt][i] - Tg[(idx ? idx[g] : g) * MLP_YSTRIDE + o];
          loss += e * e * inv_n;
        }
        S.dy[s][o] = 2.f * e * inv_n;
      }
    __syncthreads();
    // dW3 (o x unit) += dY^T H2, db3
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) gW3[nt] = mfma(a_cols<MLP_OUTP>(S.dy, 0, 0), b_lds<MLP_HID>(S.f.h2, 16 * nt, 0), gW3[nt]);
    if (l < MLP_OUTP)
      for (int s = 0; s < MLP_TILE; ++s) gb3 += S.dy[s][l];
    // dZ2 = (dY W3) * (1 - H2^2)
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
      const bf16x8 b = b_wcols<MLP_HID>(W3, 16 * nt, 0, MLP_OUTP);
#pragma unroll
      for (int mt = 0; mt < 2; ++mt) {
        // A: rows = samples, k = outputs (16 used of 32)
        bf16x8 a;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          const int k = 8 * (l >> 4) + i;
          a[i] = (__bf16)(k < MLP_OUTP ? S.dy[16 * mt + (l & 15)][k] : 0.f);
        }
        f32x4 d = mfma(a, b, splat(0.f));
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int s = 16 * mt + 4 * (l >> 4) + i, u = 16 * nt + (l & 15);
          const float h = S.f.h2[s][u];
          S.dz2[s][u] = d[i] * (1.f - h * h);
        }
      }
    }
    __syncthreads();
    // dW2 (unit_out x unit_in) += dZ2^T H1, db2
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      const bf16x8 a = a_cols<MLP_HID>(S.dz2, 16 * mt, 0);
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) gW2[mt][nt] = mfma(a, b_lds<MLP_HID>(S.f.h1, 16 * nt, 0), gW2[mt][nt]);
    }
    for (int s = 0; s < MLP_TILE; ++s) gb2 += S.dz2[s][l];
    // dZ1 = (dZ2 W2) * (1 - H1^2)
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
      const bf16x8 b0 = b_wcols<MLP_HID>(W2, 16 * nt, 0, MLP_HID);
      const bf16x8 b1v = b_wcols<MLP_HID>(W2, 16 * nt, 32, MLP_HID);
#pragma unroll
      for (int mt = 0; mt < 2; ++mt) {
        f32x4 d = mfma(a_rows<MLP_HID>(S.dz2, 16 * mt, 0), b0, splat(0.f));
        d = mfma(a_rows<MLP_HID>(S.dz2, 16 * mt, 32), b1v, d);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int s = 16 * mt + 4 * (l >> 4) + i, u = 16 * nt + (l & 15);
          const float h = S.f.h1[s][u];
          S.dz1[s][u] = d[i] * (1.f - h * h);
        }
      }
    }
    __syncthreads();
    // dW1 (unit x input) += dZ1^T X (inputs 0..15), db1
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) gW1[mt] = mfma(a_cols<MLP_HID>(S.dz1, 16 * mt, 0), b_lds<MLP_INP>(S.f.x, 0, 0), gW1[mt]);
    for (int s = 0; s < MLP_TILE; ++s) gb1 += S.dz1[s][l];
    __syncthreads();
  }
  // partial sums of this block
  float *out = partials + (long)blockIdx.x * (MLP_NPARAM + 1);
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int u = 16 * mt + 4 * (l >> 4) + i;
      out[MLP_OFF_W1 + u * MLP_INP + (l & 15)] = gW1[mt][i];
      out[MLP_OFF_W1 + u * MLP_INP + 16 + (l & 15)] = 0.f;
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) out[MLP_OFF_W2 + u * MLP_HID + 16 * nt + (l & 15)] = gW2[mt][nt][i];
    }
#pragma unroll
  for (int nt = 0; nt < 4; ++nt)
#pragma unroll
    for (int i = 0; i < 4; ++i) out[MLP_OFF_W3 + (4 * (l >> 4) + i) * MLP_HID + 16 * nt + (l & 15)] = gW3[nt][i];
  out[MLP_OFF_B1 + l] = gb1;
  out[MLP_OFF_B2 + l] = gb2;
  if (l < MLP_OUTP) out[MLP_OFF_B3 + l] = gb3;
  // loss: wavefront sum
  for (int off = 32; off > 0; off >>= 1) loss += __shfl_xor(loss, off, 64);
  if (l == 0) out[MLP_NPARAM] = loss;
}

// grads[p] = sum over blocks of partials[b][p], in block order
__global__ void k_mlp_reduce(const float *partials, int blocks, float *grads, float *loss) {
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p <= MLP_NPARAM; p += gridDim.x * blockDim.x) {
    float s = 0.f;
    for (int b = 0; b < blocks; ++b) s += partials[(long)b * (MLP_NPARAM + 1) + p];
    if (p < MLP_NPARAM) grads[p] = s;
    else if (loss) *loss = s;
  }
}

// Adam (Kingma & Ba, 2015) on the FP32 master copy; padded entries of W1
// (inputs >= 6) and of W3 / b3 (outputs >= 4) stay zero
__global__ void k_mlp_adam(float *P, const float *G, float *M, float *V, float lr, float b1c, float b2c) {
  const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < MLP_NPARAM; p += gridDim.x * blockDim.x) {
    bool pad = false;
    if (p >= MLP_OFF_W1 && p < MLP_OFF_B1) pad = (p - MLP_OFF_W1) % MLP_INP >= MLP_IN;
    if (p >= MLP_OFF_W3 && p < MLP_OFF_B3) pad = (p - MLP_OFF_W3) / MLP_HID >= MLP_OUT;
    if (p >= MLP_OFF_B3) pad = p - MLP_OFF_B3 >= MLP_OUT;
    if (pad) { P[p] = 0.f; continue; }
    const float g = G[p];
    const float m = beta1 * M[p] + (1.f - beta1) * g;
    const float v = beta2 * V[p] + (1.f - beta2) * g * g;
    M[p] = m;
    V[p] = v;
    P[p] -= lr * (m / b1c) / (sqrtf(v / b2c) + eps);
  }
}

// sums for the normalisation: features (6), squared features (6), squared
// targets (4)
__global__ void k_mlp_norm_sums(const float *X, const float *T, long n, double *acc) {
  double a[16];
  for (int k = 0; k < 16; ++k) a[k] = 0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    for (int k = 0; k < MLP_IN; ++k) {
      const double x = X[i * MLP_XSTRIDE + k];
      a[k] += x;
      a[MLP_IN + k] += x * x;
    }
    for (int k = 0; k < MLP_OUT; ++k) {
      const double t = T[i * MLP_YSTRIDE + k];
      a[2 * MLP_IN + k] += t * t;
    }
  }
  for (int k = 0; k < 16; ++k) {
    double v = a[k];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(&acc[k], v);
  }
}

__global__ void k_mlp_normalise(const float *X, const float *T, long n, SurrNorm nm, float *Xn, float *Tn) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    for (int k = 0; k < MLP_XSTRIDE; ++k)
      Xn[i * MLP_XSTRIDE + k] = k < MLP_IN ? (X[i * MLP_XSTRIDE + k] - nm.mean[k]) / nm.std[k] : 0.f;
    for (int k = 0; k < MLP_YSTRIDE; ++k) Tn[i * MLP_YSTRIDE + k] = T[i * MLP_YSTRIDE + k] / nm.tscale[k];
  }
}

}  // namespace

int mlp_alloc(pcp_ctx *c) {
  if (c->mlp) return 0;
  MlpState *m = new MlpState();
  c->mlp = m;
  m->partial_blocks = c->n_cu * 2;
  if (hipMalloc(&m->params, MLP_NPARAM * sizeof(float)) != hipSuccess ||
      hipMalloc(&m->grads, MLP_NPARAM * sizeof(float)) != hipSuccess ||
      hipMalloc(&m->m, MLP_NPARAM * sizeof(float)) != hipSuccess ||
      hipMalloc(&m->v, MLP_NPARAM * sizeof(float)) != hipSuccess ||
      hipMalloc(&m->partials, (long)m->partial_blocks * (MLP_NPARAM + 1) * sizeof(float)) != hipSuccess)
    return -1;
  if (hipMemset(m->params, 0, MLP_NPARAM * sizeof(float)) != hipSuccess ||
      hipMemset(m->m, 0, MLP_NPARAM * sizeof(float)) != hipSuccess ||
      hipMemset(m->v, 0, MLP_NPARAM * sizeof(float)) != hipSuccess)
    return -1;
  for (int i = 0; i < MLP_IN; ++i) { m->feat_mean[i] = 0.f; m->feat_std[i] = 1.f; }
  for (int i = 0; i < MLP_OUT; ++i) m->targ_scale[i] = 1.f;
  return 0;
}

void mlp_free(pcp_ctx *c) {
  MlpState *m = c->mlp;
  if (!m) return;
  (void)hipFree(m->params); (void)hipFree(m->grads); (void)hipFree(m->m); (void)hipFree(m->v);
  (void)hipFree(m->partials); (void)hipFree(m->X); (void)hipFree(m->T);
  (void)hipFree(m->xb); (void)hipFree(m->tb); (void)hipFree(m->yb);
  (void)hipFree(m->Xn); (void)hipFree(m->Tn); (void)hipFree(m->idx);
  delete m;
  c->mlp = nullptr;
}

void mlp_forward(pcp_ctx *c, const float *X, long n, float *Y) {
  if (n <= 0) return;
  // 16 samples per wavefront trip; enough wavefronts to fill every SIMD
  // several times over, each loading the weights once
  const long waves = (n + 15) / 16;
  const long per_block = MLP_FWD_BLOCK / WAVE;
  unsigned grid = (unsigned)std::max<long>(1, std::min<long>((waves + per_block - 1) / per_block, (long)c->n_cu * 16));
  hipLaunchKernelGGL(k_mlp_forward, dim3(grid), dim3(MLP_FWD_BLOCK), 0, c->stream, c->mlp->params, X, n, Y);
}

void mlp_grad_idx(pcp_ctx *c, const float *X, const float *T, const long *idx, long n, float *loss_dev) {
  MlpState *m = c->mlp;
  long tiles = (n + MLP_TILE - 1) / MLP_TILE;
  int blocks = (int)std::max<long>(1, std::min<long>(tiles, m->partial_blocks));
  hipLaunchKernelGGL(k_mlp_train, dim3(blocks), dim3(WAVE), 0, c->stream, m->params, X, T, idx, n, 1.f / (float)n,
                     m->partials);
  hipLaunchKernelGGL(k_mlp_reduce, dim3((MLP_NPARAM + 256) / 256), dim3(256), 0, c->stream, m->partials, blocks,
                     m->grads, loss_dev);
}

void mlp_grad(pcp_ctx *c, const float *X, const float *T, long n, float *loss_dev) {
  mlp_grad_idx(c, X, T, nullptr, n, loss_dev);
}

void mlp_adam(pcp_ctx *c, float lr) {
  MlpState *m = c->mlp;
  m->adam_t += 1;
  const float b1c = 1.f - powf(0.9f, (float)m->adam_t), b2c = 1.f - powf(0.999f, (float)m->adam_t);
  hipLaunchKernelGGL(k_mlp_adam, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, c->stream, m->params, m->grads, m->m,
                     m->v, lr, b1c, b2c);
}

int mlp_init(pcp_ctx *c, uint64_t seed) {
  MlpState *m = c->mlp;
  std::vector<float> P(MLP_NPARAM, 0.f);
  std::mt19937_64 gen(seed);
  auto fill = [&](int off, int rows, int rows_used, int cols, int cols_used, int fan_in, int fan_out) {
    std::uniform_real_distribution<float> U(-std::sqrt(6.f / (fan_in + fan_out)), std::sqrt(6.f / (fan_in + fan_out)));
    for (int r = 0; r < rows; ++r)
      for (int k = 0; k < cols; ++k) P[off + r * cols + k] = (r < rows_used && k < cols_used) ? U(gen) : 0.f;
  };
  fill(MLP_OFF_W1, MLP_HID, MLP_HID, MLP_INP, MLP_IN, MLP_IN, MLP_HID);
  fill(MLP_OFF_W2, MLP_HID, MLP_HID, MLP_HID, MLP_HID, MLP_HID, MLP_HID);
  fill(MLP_OFF_W3, MLP_OUTP, MLP_OUT, MLP_HID, MLP_HID, MLP_HID, MLP_OUT);
  if (hipMemcpy(m->params, P.data(), MLP_NPARAM * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(m->m, 0, MLP_NPARAM * sizeof(float)) != hipSuccess ||
      hipMemset(m->v, 0, MLP_NPARAM * sizeof(float)) != hipSuccess)
    return -1;
  m->adam_t = 0;
  m->inited = 1;
  return 0;
}

int mlp_reserve(pcp_ctx *c, long rows) {
  MlpState *m = c->mlp;
  if (m->n + rows <= m->cap) return 0;
  long cap = std::max(m->n + rows, 2 * m->cap);
  float *X = nullptr, *T = nullptr;
  if (hipMalloc(&X, cap * MLP_XSTRIDE * sizeof(float)) != hipSuccess ||
      hipMalloc(&T, cap * MLP_YSTRIDE * sizeof(float)) != hipSuccess)
    return -1;
  if (m->n > 0) {
    if (hipMemcpy(X, m->X, m->n * MLP_XSTRIDE * sizeof(float), hipMemcpyDeviceToDevice) != hipSuccess ||
        hipMemcpy(T, m->T, m->n * MLP_YSTRIDE * sizeof(float), hipMemcpyDeviceToDevice) != hipSuccess)
      return -1;
  }
  (void)hipFree(m->X);
  (void)hipFree(m->T);
  m->X = X;
  m->T = T;
  m->cap = cap;
  return 0;
}

int mlp_fit(pcp_ctx *c, int epochs, long batch, float lr, uint64_t seed, float *loss) {
  MlpState *m = c->mlp;
  const long n = m->n;
  if (n < MLP_TILE || batch < 1) return -2;
  if (!m->inited && mlp_init(c, seed) < 0) return -1;
  if (m->ncap < n) {
    (void)hipFree(m->Xn); (void)hipFree(m->Tn); (void)hipFree(m->idx);
    m->Xn = m->Tn = nullptr;
    m->idx = nullptr;
    if (hipMalloc(&m->Xn, n * MLP_XSTRIDE * sizeof(float)) != hipSuccess ||
        hipMalloc(&m->Tn, n * MLP_YSTRIDE * sizeof(float)) != hipSuccess ||
        hipMalloc(&m->idx, n * sizeof(long)) != hipSuccess)
      return -1;
    m->ncap = n;
  }
  // normalisation from the training set
  double *acc = nullptr;
  if (hipMalloc(&acc, 16 * sizeof(double)) != hipSuccess) return -1;
  (void)hipMemsetAsync(acc, 0, 16 * sizeof(double), c->stream);
  unsigned g = (unsigned)std::min<long>((n + 255) / 256, (long)c->n_cu * 8);
  hipLaunchKernelGGL(k_mlp_norm_sums, dim3(g), dim3(256), 0, c->stream, m->X, m->T, n, acc);
  double h[16];
  if (hipMemcpyAsync(h, acc, sizeof h, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    return -1;
  (void)hipFree(acc);
  SurrNorm nm;
  for (int k = 0; k < MLP_IN; ++k) {
    const double mean = h[k] / n, var = h[MLP_IN + k] / n - mean * mean;
    m->feat_mean[k] = nm.mean[k] = (float)mean;
    m->feat_std[k] = nm.std[k] = (float)std::sqrt(std::max(var, 1e-8));
  }
  // one scale for all targets (the largest rms): an output that is mostly
  // cell-level noise (the ions' relative changes) then weighs little in the
  // loss instead of competing with the electrons' for the hidden units
  double tmax = 1e-20;
  for (int k = 0; k < MLP_OUT; ++k) tmax = std::max(tmax, h[2 * MLP_IN + k] / n);
  for (int k = 0; k < MLP_OUT; ++k) m->targ_scale[k] = nm.tscale[k] = (float)std::sqrt(tmax);
  hipLaunchKernelGGL(k_mlp_normalise, dim3(g), dim3(256), 0, c->stream, m->X, m->T, n, nm, m->Xn, m->Tn);
  // mini-batch Adam over shuffled passes
  const long nb = (n + batch - 1) / batch;
  float *ld = nullptr;
  if (hipMalloc(&ld, nb * sizeof(float)) != hipSuccess) return -1;
  std::vector<long> perm(n);
  std::mt19937_64 gen(seed);
  std::vector<float> hl(nb);
  for (int e = 0; e < epochs; ++e) {
    // cosine annealing from lr to lr / 20 over the passes
    const float lr_e = lr * (0.05f + 0.95f * 0.5f * (1.f + std::cos(3.14159265f * (float)e / (float)epochs)));
    for (long i = 0; i < n; ++i) perm[i] = i;
    std::shuffle(perm.begin(), perm.end(), gen);
    if (hipMemcpyAsync(m->idx, perm.data(), n * sizeof(long), hipMemcpyHostToDevice, c->stream) != hipSuccess)
      return -1;
    for (long b = 0; b < nb; ++b) {
      const long b0 = b * batch, cnt = std::min(batch, n - b0);
      mlp_grad_idx(c, m->Xn, m->Tn, m->idx + b0, cnt, ld + b);
      mlp_adam(c, lr_e);
    }
    // the host permutation buffer is reused by the next pass
    if (hipStreamSynchronize(c->stream) != hipSuccess) return -1;
  }
  if (hipMemcpy(hl.data(), ld, nb * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) return -1;
  (void)hipFree(ld);
  double s = 0;
  for (float v : hl) s += v;
  if (loss) *loss = (float)(s / nb);
  m->trained = 1;
  return 0;
}

}  // namespace pcp
