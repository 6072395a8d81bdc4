// Per-cell collision surrogate network (defined in pcp_mlp.hip): a
// 6 -> 64 -> 64 -> 4 perceptron with tanh hidden layers, evaluated and
// trained with hand-written BF16 MFMA kernels (v_mfma_f32_16x16x32_bf16,
// FP32 accumulation).  No counterpart in the reference.
//
// Parameter block (FP32 master copy, one contiguous array):
//   W1 [64][32]  (inputs padded from 6 to 32 with zero columns)
//   b1 [64]
//   W2 [64][64]
//   b2 [64]
//   W3 [16][64]  (outputs padded from 4 to 16 with zero rows)
//   b3 [16]
// Kernels convert weights to BF16 when they load them; activations and
// back-propagated errors are rounded to BF16 where they enter an MFMA.
#pragma once

#include "pcp_internal.hpp"

namespace pcp {

constexpr int MLP_IN = 6, MLP_HID = 64, MLP_OUT = 4;
constexpr int MLP_INP = 32, MLP_OUTP = 16;     // padded widths
constexpr int MLP_OFF_W1 = 0;
constexpr int MLP_OFF_B1 = MLP_OFF_W1 + MLP_HID * MLP_INP;
constexpr int MLP_OFF_W2 = MLP_OFF_B1 + MLP_HID;
constexpr int MLP_OFF_B2 = MLP_OFF_W2 + MLP_HID * MLP_HID;
constexpr int MLP_OFF_W3 = MLP_OFF_B2 + MLP_HID;
constexpr int MLP_OFF_B3 = MLP_OFF_W3 + MLP_OUTP * MLP_HID;
constexpr int MLP_NPARAM = MLP_OFF_B3 + MLP_OUTP;
constexpr int MLP_TILE = 32;                    // samples per wavefront tile
constexpr int MLP_XSTRIDE = 8;                  // floats per input row in memory
constexpr int MLP_YSTRIDE = 4;                  // floats per output row in memory

// feature normalisation and target scale of a fitted network
struct SurrNorm {
  float mean[MLP_IN], std[MLP_IN], tscale[MLP_OUT];
};

struct MlpState {
  float *params = nullptr, *grads = nullptr, *m = nullptr, *v = nullptr;
  float *partials = nullptr;       // [blocks][MLP_NPARAM + 1] (last: loss)
  int partial_blocks = 0;
  long adam_t = 0;
  // training set on the device: X [n][8], T [n][4]
  float *X = nullptr, *T = nullptr;
  long n = 0, cap = 0;
  // scratch for host-given batches and the normalised training set
  float *xb = nullptr, *tb = nullptr, *yb = nullptr;
  long bcap = 0;
  float *Xn = nullptr, *Tn = nullptr;
  long *idx = nullptr;
  long ncap = 0;
  // feature / target normalisation (applied by the callers)
  float feat_mean[MLP_IN], feat_std[MLP_IN], targ_scale[MLP_OUT];
  int trained = 0;
  int inited = 0;
};

int mlp_alloc(pcp_ctx *c);
void mlp_free(pcp_ctx *c);
// Y[n][4] = net(X[n][8]) (first 6 columns used), device pointers
void mlp_forward(pcp_ctx *c, const float *X, long n, float *Y);
// loss = sum over samples and outputs of (y - t)^2 / n; gradients into
// state.grads (deterministic block partials + reduction)
void mlp_grad(pcp_ctx *c, const float *X, const float *T, long n, float *loss_dev);
void mlp_adam(pcp_ctx *c, float lr);
// gradients over the samples idx[0 .. n) of X / T
void mlp_grad_idx(pcp_ctx *c, const float *X, const float *T, const long *idx, long n, float *loss_dev);
// Glorot-uniform weights, zero biases and Adam state (host RNG, seeded)
int mlp_init(pcp_ctx *c, uint64_t seed);
// room for `rows` more training rows
int mlp_reserve(pcp_ctx *c, long rows);
// normalise the training set (feature mean / std, one target scale: the
// largest rms of the four outputs), then
// `epochs` passes of shuffled mini-batches with Adam, the step size
// annealed from lr to lr / 20 along a cosine; *loss = mean batch loss of the
// last pass (normalised units)
int mlp_fit(pcp_ctx *c, int epochs, long batch, float lr, uint64_t seed, float *loss);

}  // namespace pcp
