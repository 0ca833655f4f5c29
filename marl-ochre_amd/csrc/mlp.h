// mlp.h — the surrogate network's sizes, parameter layout and device state.
#pragma once

#include <stddef.h>
#include <stdint.h>

struct marl_ctx;

constexpr int kIn = 80;      // 16 internal + 12 environment + light + 16 x 3 reaction params + 3 topology classes
constexpr int kHidden = 256;
constexpr int kOut = 28;     // 16 internal deltas + 12 exchange fluxes

// Flat FP32 parameter vector: for each layer l, W_l [in x out] row-major, then b_l [out].
struct MlpLayout {
  int in[4], out[4];
  size_t w[4], b[4], total;
};
MlpLayout mlp_layout();

struct MlpState {
  float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;  // params, grads, Adam moments
  uint16_t* packed = nullptr;                                     // bf16 MFMA fragment images
  uint16_t* pk[4] = {nullptr, nullptr, nullptr, nullptr};
  float* norm = nullptr;  // feature mean[kIn], std[kIn], target mean[kOut], std[kOut]
  float* ws = nullptr;    // split-K partial products of the weight gradients
  int step = 0;
  int batch = 0;
  float *x = nullptr, *t = nullptr, *y = nullptr, *dy = nullptr;
  float* z[3] = {nullptr, nullptr, nullptr};
  float* h[3] = {nullptr, nullptr, nullptr};
  float *dh = nullptr, *dz = nullptr, *loss_part = nullptr;
  float* dzpart = nullptr;  // column partials of dZ per block of rows (bias gradients)
  uint16_t* xb = nullptr;
  float* yinf = nullptr;
  size_t infer_cap = 0;
};

int mlp_alloc(marl_ctx* ctx, MlpState& s, int batch);
void mlp_free(MlpState& s);
int mlp_forward_train(marl_ctx* ctx, MlpState& s, int M);
int mlp_backward(marl_ctx* ctx, MlpState& s, int M);
int mlp_loss(marl_ctx* ctx, MlpState& s, int M, float* loss_host);
int mlp_adam(marl_ctx* ctx, MlpState& s, float lr);
int mlp_pack(marl_ctx* ctx, MlpState& s);
int mlp_infer(marl_ctx* ctx, MlpState& s, int M, const float* xin, float* yout);
int mlp_infer_bf16(marl_ctx* ctx, MlpState& s, int M, const uint16_t* xb, float* yout);
