// cnn_net.hip -- the learned initial-guess network behind include/aldic.h:
// parameters, activations and training scratch on the device, the forward
// pass (implicit-GEMM convolutions on the matrix cores), the backward pass
// and Adam, and the patch -> network -> IC-GN starting point pipeline.
//
// Network (input: 48 x 48 x 2 NHWC bf16, reference and deformed patch, each
// normalised to zero mean and unit variance):
//   conv3x3  2 -> 32, stride 2   48 -> 24
//   conv3x3 32 -> 32, stride 2   24 -> 12
//   conv3x3 32 -> 64, stride 2   12 -> 6
//   conv3x3 64 -> 64, stride 1    6 -> 6
//   conv3x3 64 -> 64, stride 2    6 -> 3
//   conv3x3 64 -> 64, stride 1    3 -> 3
//   fc 576 -> 128, fc 128 -> 6
// ReLU after every layer but the last.  Outputs are the normalised warp
// parameters [u_x v_x u_y v_y du dv] / [sf sf sf sf su su].
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "aldic_internal.h"
#include "aldic_kernels.h"

namespace {

constexpr int NLAYER = 8;
constexpr int NCONV = 6;
constexpr int PATCH = 48;
constexpr float SCALE_F = 0.05f;  // |u_x| .. |v_y| of the training range
constexpr float SCALE_U = 4.0f;   // displacement unit of the outputs (px)

struct LayerDef {
  int cin, cout, stride, hin;  // conv: input side hin; fc: hin = 0
};
constexpr LayerDef LAYERS[NLAYER] = {
    {2, 32, 2, 48}, {32, 32, 2, 24}, {32, 64, 2, 12}, {64, 64, 1, 6},
    {64, 64, 2, 6}, {64, 64, 1, 3},  {576, 128, 0, 0}, {128, 6, 0, 0},
};

// the fused front end (cnn.hip: front_kernel) is written for this geometry
static_assert(LAYERS[0].cin == 2 && LAYERS[0].cout == 32 && LAYERS[0].stride == 2 &&
                  LAYERS[0].hin == 48 && LAYERS[1].cin == 32 && LAYERS[1].cout == 32 &&
                  LAYERS[1].stride == 2 && LAYERS[1].hin == 24,
              "front_kernel geometry");

struct Layer {
  int cin, cout, stride, H, W, Ho, Wo;
  int K, Kp;  // reduction length and its padding to a multiple of 16
  int64_t w_off, b_off;
  bool conv;
};

int64_t up(int64_t v, int64_t a) { return (v + a - 1) / a * a; }

// Geometry and parameter offsets of every layer; returns the parameter count.
int64_t build_layers(Layer* L) {
  int64_t off = 0;
  for (int l = 0; l < NLAYER; l++) {
    const LayerDef& d = LAYERS[l];
    Layer& y = L[l];
    y.cin = d.cin;
    y.cout = d.cout;
    y.conv = l < NCONV;
    y.stride = d.stride;
    if (y.conv) {
      y.H = y.W = d.hin;
      y.Ho = y.Wo = (d.hin - 1) / d.stride + 1;
      y.K = 9 * d.cin;
    } else {
      y.H = y.W = y.Ho = y.Wo = 1;
      y.K = d.cin;
    }
    y.Kp = (int)up(y.K, 16);
    y.w_off = off;
    off = up(off + (int64_t)y.cout * y.Kp, 64);
    y.b_off = off;
    off = up(off + y.cout, 64);
  }
  return off;
}

}  // namespace

struct aldic_cnn {
  aldic_ctx* ctx = nullptr;
  int device = 0;  // destroy may run after the context is gone
  int bmax = 0;
  Layer L[NLAYER];
  int64_t nparam = 0;
  int64_t step = 0;
  // parameters (FP32 master, bf16 copy in the same layout), gradients, Adam
  float* wf = nullptr;
  void* wb = nullptr;
  float* gw = nullptr;
  float* am = nullptr;
  float* av = nullptr;
  // activations: act[l] = output of layer l (bf16), out = FP32 network output
  void* act[NLAYER] = {nullptr};
  float* out = nullptr;
  // training scratch
  void* col = nullptr;     // im2col of a layer input, bf16
  void* colT = nullptr;    // its transpose
  void* dYT = nullptr;     // transposed output gradient, bf16
  void* dYb = nullptr;     // output gradient, bf16
  float* dYf = nullptr;    // output gradient, FP32
  float* dcol = nullptr;   // input-column gradient, FP32
  float* dX = nullptr;     // input gradient, FP32
  void* WT = nullptr;      // transposed weights, bf16
  void* WT2 = nullptr;     // transposed fc2 weights [128][16], columns 6..15 zero
  float* part = nullptr;   // split-K partial sums
  int64_t part_cap = 0;
  float* dout = nullptr;   // dloss/dout, FP32 [bmax x 16]
  void* doutb = nullptr;   // bf16 copy
  float* loss = nullptr;
  float* lscale = nullptr;
};

namespace {

// Zeroed device buffer.  hipMemset runs on the null stream, which the
// context's non-blocking stream does not wait for: finish it before any
// kernel on that stream can touch the buffer.
int cnn_alloc(void** p, size_t bytes) {
  HIP_TRY(hipMalloc(p, std::max<size_t>(bytes, 256)));
  HIP_TRY(hipMemset(*p, 0, std::max<size_t>(bytes, 256)));
  HIP_TRY(hipDeviceSynchronize());
  return ALDIC_OK;
}

template <typename T>
int cnn_alloc_t(T** p, size_t count, size_t elem) {
  void* v = nullptr;
  RC_TRY(cnn_alloc(&v, count * elem));
  *p = static_cast<T*>(v);
  return ALDIC_OK;
}

void cnn_free(aldic_cnn* net) {
  void* ptrs[] = {net->wf, net->wb, net->gw, net->am, net->av, net->out, net->col, net->colT,
                  net->dYT, net->dYb, net->dYf, net->dcol, net->dX, net->WT, net->WT2, net->part,
                  net->dout, net->doutb, net->loss, net->lscale};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (int l = 0; l < NLAYER; l++)
    if (net->act[l]) (void)hipFree(net->act[l]);
}

inline const uint16_t* bfp(const void* base, int64_t off) {
  return static_cast<const uint16_t*>(base) + off;
}
inline uint16_t* bfp(void* base, int64_t off) { return static_cast<uint16_t*>(base) + off; }

// Forward pass over B inputs from layer l0 on (l0 > 0: the input is
// act[l0 - 1]); activations stay in net->act for backward.
int forward(aldic_cnn* net, const void* x, int B, int l0 = 0) {
  hipStream_t s = net->ctx->stream;
  const void* in = l0 == 0 ? x : net->act[l0 - 1];
  for (int l = l0; l < NCONV; l++) {
    const Layer& y = net->L[l];
    HIP_TRY(aldic::cnn_conv(in, B, y.H, y.W, y.cin, y.stride, y.Ho, y.Wo, bfp(net->wb, y.w_off),
                            y.Kp, y.cout, net->wf + y.b_off, 1, net->act[l], nullptr, s));
    in = net->act[l];
  }
  const Layer& f1 = net->L[NCONV];
  HIP_TRY(aldic::cnn_dense(in, B, f1.Kp, bfp(net->wb, f1.w_off), f1.cout, net->wf + f1.b_off, 1,
                           net->act[NCONV], s));
  const Layer& f2 = net->L[NCONV + 1];
  HIP_TRY(aldic::cnn_gemm(net->act[NCONV], f2.Kp, bfp(net->wb, f2.w_off), f2.Kp, B, f2.cout,
                          f2.Kp, 1, net->out, f2.cout, net->wf + f2.b_off, 0, nullptr, 0, s));
  return ALDIC_OK;
}

// C[M x N] (FP32, ldc = N) = A[M x K] * Bt[N x K]^T with the reduction over
// the batch split across workgroups when it is long.
int wgrad(aldic_cnn* net, const void* A, int lda, const void* Bt, int ldb, int M, int N, int K,
          float* C) {
  hipStream_t s = net->ctx->stream;
  int splits = std::min(64, std::max(1, K / 512));
  while (splits > 1 && (int64_t)splits * M * N > net->part_cap) splits--;
  if (splits == 1) {
    HIP_TRY(aldic::cnn_gemm(A, lda, Bt, ldb, M, N, K, 1, C, N, nullptr, 0, nullptr, 0, s));
    return ALDIC_OK;
  }
  // the launcher may merge chunks; it reports nothing, so recompute the count
  const int ksteps = K / 16, per = (ksteps + splits - 1) / splits;
  const int used = (ksteps + per - 1) / per;
  HIP_TRY(aldic::cnn_gemm(A, lda, Bt, ldb, M, N, K, splits, net->part, N, nullptr, 0, nullptr, 0,
                          s));
  HIP_TRY(aldic::cnn_splitk_reduce(net->part, used, M, N, N, nullptr, 0, C, N, nullptr, 0, s));
  return ALDIC_OK;
}

// Backward pass of the loss whose output gradient is in net->dout (FP32,
// row stride 16, columns 6..15 zero); fills net->gw.
int backward(aldic_cnn* net, const void* x, int B) {
  hipStream_t s = net->ctx->stream;
  const Layer& f2 = net->L[NCONV + 1];
  const Layer& f1 = net->L[NCONV];
  // ---- fc2: out = h W2^T + b2 ----
  HIP_TRY(aldic::cnn_to_bf16(net->dout, (int64_t)B * 16, net->doutb, s));
  HIP_TRY(aldic::cnn_transpose_f32(net->dout, B, f2.cout, 16, net->dYT, B, s));   // [6][B]
  HIP_TRY(aldic::cnn_transpose_bf16(net->act[NCONV], B, f2.Kp, f2.Kp, net->colT, B, s));  // [128][B]
  RC_TRY(wgrad(net, net->dYT, B, net->colT, B, f2.cout, f2.Kp, B, net->gw + f2.w_off));
  HIP_TRY(aldic::cnn_colsum(net->dout, B, f2.cout, 16, net->gw + f2.b_off, s));
  // dh = dout W2: W2^T is [128][16] with columns 6..15 zero
  HIP_TRY(aldic::cnn_transpose_bf16(bfp(net->wb, f2.w_off), f2.cout, f2.Kp, f2.Kp, net->WT2, 16, s));
  HIP_TRY(aldic::cnn_gemm(net->doutb, 16, net->WT2, 16, B, f2.Kp, 16, 1, net->dYf, f2.Kp, nullptr,
                          0, nullptr, 0, s));
  HIP_TRY(aldic::cnn_relu_back(net->dYf, net->act[NCONV], (int64_t)B * f1.cout, net->dYb,
                               net->dYf, s));
  // ---- fc1: h = relu(a6 W1^T + b1) ----
  HIP_TRY(aldic::cnn_transpose_bf16(net->dYb, B, f1.cout, f1.cout, net->dYT, B, s));  // [128][B]
  HIP_TRY(aldic::cnn_transpose_bf16(net->act[NCONV - 1], B, f1.Kp, f1.Kp, net->colT, B, s));
  RC_TRY(wgrad(net, net->dYT, B, net->colT, B, f1.cout, f1.Kp, B, net->gw + f1.w_off));
  HIP_TRY(aldic::cnn_colsum(net->dYf, B, f1.cout, f1.cout, net->gw + f1.b_off, s));
  // da6 = dh W1, W1^T is [576][128]
  HIP_TRY(aldic::cnn_transpose_bf16(bfp(net->wb, f1.w_off), f1.cout, f1.Kp, f1.Kp, net->WT,
                                    f1.cout, s));
  HIP_TRY(aldic::cnn_gemm(net->dYb, f1.cout, net->WT, f1.cout, B, f1.Kp, f1.cout, 1, net->dYf,
                          f1.Kp, nullptr, 0, nullptr, 0, s));
  HIP_TRY(aldic::cnn_relu_back(net->dYf, net->act[NCONV - 1], (int64_t)B * f1.Kp, net->dYb,
                               net->dYf, s));
  // ---- convolutions, last to first; dYb/dYf hold the layer's output gradient ----
  for (int l = NCONV - 1; l >= 0; l--) {
    const Layer& y = net->L[l];
    const void* in = l == 0 ? x : net->act[l - 1];
    const int rows = B * y.Ho * y.Wo;
    HIP_TRY(aldic::cnn_colsum(net->dYf, rows, y.cout, y.cout, net->gw + y.b_off, s));
    HIP_TRY(aldic::cnn_im2col(in, B, y.H, y.W, y.cin, y.stride, y.Ho, y.Wo, y.Kp, net->col, s));
    HIP_TRY(aldic::cnn_transpose_bf16(net->col, rows, y.Kp, y.Kp, net->colT, rows, s));
    HIP_TRY(aldic::cnn_transpose_bf16(net->dYb, rows, y.cout, y.cout, net->dYT, rows, s));
    RC_TRY(wgrad(net, net->dYT, rows, net->colT, rows, y.cout, y.Kp, rows, net->gw + y.w_off));
    if (l == 0) break;
    // dcol = dY W, W^T is [Kp][cout]
    HIP_TRY(aldic::cnn_transpose_bf16(bfp(net->wb, y.w_off), y.cout, y.Kp, y.Kp, net->WT, y.cout,
                                      s));
    HIP_TRY(aldic::cnn_gemm(net->dYb, y.cout, net->WT, y.cout, rows, y.Kp, y.cout, 1, net->dcol,
                            y.Kp, nullptr, 0, nullptr, 0, s));
    HIP_TRY(aldic::cnn_col2im(net->dcol, B, y.H, y.W, y.cin, y.stride, y.Ho, y.Wo, y.Kp, net->dX,
                              s));
    HIP_TRY(aldic::cnn_relu_back(net->dX, net->act[l - 1], (int64_t)B * y.H * y.W * y.cin,
                                 net->dYb, net->dYf, s));
  }
  return ALDIC_OK;
}

// Buffers that only the layer-by-layer forward (layer-0 activation) and
// training (im2col, transposes, gradients) use.
int ensure_layer0(aldic_cnn* net) {
  if (net->act[0]) return ALDIC_OK;
  const Layer& y = net->L[0];
  return cnn_alloc(&net->act[0], (size_t)net->bmax * y.Ho * y.Wo * y.cout * 2);
}

int ensure_training(aldic_cnn* net) {
  if (net->col) return ALDIC_OK;
  const int64_t mb = net->bmax;
  int64_t max_act = 0, max_col = 0, max_wt = 0, max_dx = 0;
  for (int l = 0; l < NLAYER; l++) {
    const Layer& y = net->L[l];
    const int64_t rows = mb * y.Ho * y.Wo;
    max_act = std::max(max_act, rows * y.cout);
    max_col = std::max(max_col, rows * y.Kp);
    max_wt = std::max<int64_t>(max_wt, (int64_t)y.Kp * std::max(y.cout, 16));
    max_dx = std::max(max_dx, mb * y.H * y.W * y.cin);
  }
  const size_t f4 = sizeof(float), b2 = 2;
  RC_TRY(cnn_alloc_t(&net->colT, max_col, b2));
  RC_TRY(cnn_alloc_t(&net->dYT, max_act, b2));
  RC_TRY(cnn_alloc_t(&net->dYb, std::max(max_act, max_dx), b2));
  RC_TRY(cnn_alloc_t(&net->dYf, std::max(max_act, max_dx), f4));
  RC_TRY(cnn_alloc_t(&net->dcol, max_col, f4));
  RC_TRY(cnn_alloc_t(&net->dX, max_dx, f4));
  RC_TRY(cnn_alloc_t(&net->WT, max_wt, b2));
  RC_TRY(cnn_alloc_t(&net->WT2, (size_t)net->L[NCONV + 1].Kp * 16, b2));
  RC_TRY(cnn_alloc_t(&net->part, net->part_cap, f4));
  RC_TRY(cnn_alloc_t(&net->dout, mb * 16, f4));
  RC_TRY(cnn_alloc_t(&net->doutb, mb * 16, b2));
  RC_TRY(cnn_alloc_t(&net->col, max_col, b2));  // last: marks the set complete
  return ALDIC_OK;
}

int check_net(const aldic_cnn* net) {
  if (!net || !net->ctx) return aldic::set_error(ALDIC_EINVAL, "NULL network");
  HIP_TRY(hipSetDevice(net->ctx->device));
  return ALDIC_OK;
}

int check_batch(const aldic_cnn* net, int64_t B) {
  if (B < 1 || B > net->bmax)
    return aldic::set_error(ALDIC_EINVAL, "batch must be in [1, max_batch]");
  return ALDIC_OK;
}

}  // namespace

extern "C" int aldic_cnn_create(aldic_ctx* ctx, int32_t max_batch, aldic_cnn** out) {
  if (!ctx || !out) return aldic::set_error(ALDIC_EINVAL, "cnn_create: NULL argument");
  if (max_batch < 16 || max_batch % 16)
    return aldic::set_error(ALDIC_EINVAL, "cnn_create: max_batch must be a positive multiple of 16");
  HIP_TRY(hipSetDevice(ctx->device));
  auto* net = new aldic_cnn();
  net->ctx = ctx;
  net->device = ctx->device;
  net->bmax = max_batch;
  const int64_t off = build_layers(net->L);
  // the flattened 3 x 3 x 64 output of the last convolution feeds fc1
  const Layer& c6 = net->L[NCONV - 1];
  if (c6.Ho * c6.Wo * c6.cout != net->L[NCONV].K) {
    delete net;
    return aldic::set_error(ALDIC_EINVAL, "cnn_create: inconsistent layer table");
  }
  net->nparam = off;
  net->part_cap = 64LL * 64 * 1024;
  const size_t f4 = sizeof(float), b2 = 2;
  int rc = ALDIC_OK;
  auto A = [&](auto** p, size_t count, size_t elem) {
    if (rc == ALDIC_OK) rc = cnn_alloc_t(p, count, elem);
  };
  A(&net->wf, off, f4);
  A(&net->wb, off, b2);
  A(&net->gw, off, f4);
  A(&net->am, off, f4);
  A(&net->av, off, f4);
  // activations from layer 1 on (the fused front end writes layer 1); the
  // layer-0 activation, the patch buffer and the training scratch are
  // allocated on first use, so an inference-only network stays small
  for (int l = 1; l < NLAYER; l++)
    A(&net->act[l], (size_t)max_batch * net->L[l].Ho * net->L[l].Wo * net->L[l].cout, b2);
  A(&net->out, (size_t)max_batch * 6, f4);
  A(&net->loss, 1, f4);
  A(&net->lscale, 6, f4);
  if (rc == ALDIC_OK) {
    const float ones[6] = {1, 1, 1, 1, 1, 1};
    if (hipMemcpy(net->lscale, ones, sizeof(ones), hipMemcpyHostToDevice) != hipSuccess)
      rc = aldic::set_error(ALDIC_EDEVICE, "cnn_create: upload failed");
  }
  if (rc != ALDIC_OK) {
    cnn_free(net);
    delete net;
    return rc;
  }
  *out = net;
  return ALDIC_OK;
}

extern "C" int aldic_cnn_destroy(aldic_cnn* net) {
  if (!net) return ALDIC_OK;
  (void)hipSetDevice(net->device);
  (void)hipDeviceSynchronize();
  cnn_free(net);
  delete net;
  return ALDIC_OK;
}

extern "C" int aldic_cnn_num_params(const aldic_cnn* net, int64_t* n) {
  if (!n) return aldic::set_error(ALDIC_EINVAL, "cnn_num_params: NULL argument");
  Layer L[NLAYER];
  *n = net ? net->nparam : build_layers(L);
  return ALDIC_OK;
}

extern "C" int aldic_cnn_layout(const aldic_cnn* net, int32_t layer, int64_t info[9]) {
  if (!info) return aldic::set_error(ALDIC_EINVAL, "cnn_layout: NULL argument");
  if (layer < 0 || layer >= NLAYER) return aldic::set_error(ALDIC_EINVAL, "cnn_layout: no such layer");
  Layer L[NLAYER];
  if (!net) build_layers(L);
  const Layer& y = net ? net->L[layer] : L[layer];
  const int64_t v[9] = {y.w_off, y.b_off, y.cout, y.K, y.Kp, y.cin, y.stride, y.H, y.conv ? 1 : 0};
  std::memcpy(info, v, sizeof(v));
  return ALDIC_OK;
}

extern "C" int aldic_cnn_scales(double* sf, double* su) {
  if (sf) *sf = SCALE_F;
  if (su) *su = SCALE_U;
  return ALDIC_OK;
}

extern "C" int aldic_cnn_set_params(aldic_cnn* net, const float* host, int64_t n) {
  RC_TRY(check_net(net));
  if (!host || n != net->nparam) return aldic::set_error(ALDIC_EINVAL, "cnn_set_params: size");
  hipStream_t s = net->ctx->stream;
  HIP_TRY(hipMemcpyAsync(net->wf, host, n * sizeof(float), hipMemcpyHostToDevice, s));
  HIP_TRY(aldic::cnn_to_bf16(net->wf, n, net->wb, s));
  HIP_TRY(hipMemsetAsync(net->am, 0, n * sizeof(float), s));
  HIP_TRY(hipMemsetAsync(net->av, 0, n * sizeof(float), s));
  HIP_TRY(hipStreamSynchronize(s));
  net->step = 0;
  return ALDIC_OK;
}

extern "C" int aldic_cnn_get_params(aldic_cnn* net, float* host, int64_t n) {
  RC_TRY(check_net(net));
  if (!host || n != net->nparam) return aldic::set_error(ALDIC_EINVAL, "cnn_get_params: size");
  HIP_TRY(hipMemcpyAsync(host, net->wf, n * sizeof(float), hipMemcpyDeviceToHost, net->ctx->stream));
  HIP_TRY(hipStreamSynchronize(net->ctx->stream));
  return ALDIC_OK;
}

extern "C" int aldic_cnn_get_grads(aldic_cnn* net, float* host, int64_t n) {
  RC_TRY(check_net(net));
  if (!host || n != net->nparam) return aldic::set_error(ALDIC_EINVAL, "cnn_get_grads: size");
  HIP_TRY(hipMemcpyAsync(host, net->gw, n * sizeof(float), hipMemcpyDeviceToHost, net->ctx->stream));
  HIP_TRY(hipStreamSynchronize(net->ctx->stream));
  return ALDIC_OK;
}

extern "C" int aldic_cnn_set_loss_weights(aldic_cnn* net, const float w[6]) {
  RC_TRY(check_net(net));
  if (!w) return aldic::set_error(ALDIC_EINVAL, "cnn_set_loss_weights: NULL");
  HIP_TRY(hipMemcpyAsync(net->lscale, w, 6 * sizeof(float), hipMemcpyHostToDevice,
                         net->ctx->stream));
  HIP_TRY(hipStreamSynchronize(net->ctx->stream));
  return ALDIC_OK;
}

extern "C" int aldic_cnn_forward_dev(aldic_cnn* net, const void* d_x, int64_t B, float* d_out) {
  RC_TRY(check_net(net));
  RC_TRY(check_batch(net, B));
  if (!d_x) return aldic::set_error(ALDIC_EINVAL, "cnn_forward_dev: NULL input");
  RC_TRY(ensure_layer0(net));
  RC_TRY(forward(net, d_x, (int)B));
  if (d_out)
    HIP_TRY(hipMemcpyAsync(d_out, net->out, B * 6 * sizeof(float), hipMemcpyDeviceToDevice,
                           net->ctx->stream));
  return ALDIC_OK;
}

extern "C" int aldic_cnn_train_step_dev(aldic_cnn* net, const void* d_x, const float* d_labels,
                                        int64_t B, float lr, int32_t update, float* d_loss) {
  RC_TRY(check_net(net));
  RC_TRY(check_batch(net, B));
  if (B % 16) return aldic::set_error(ALDIC_EINVAL, "cnn_train_step_dev: batch must be a multiple of 16");
  if (!d_x || !d_labels) return aldic::set_error(ALDIC_EINVAL, "cnn_train_step_dev: NULL input");
  hipStream_t s = net->ctx->stream;
  RC_TRY(ensure_layer0(net));
  RC_TRY(ensure_training(net));
  RC_TRY(forward(net, d_x, (int)B));
  HIP_TRY(aldic::cnn_mse(net->out, 6, d_labels, net->lscale, (int)B, net->dout, 16, net->loss, s));
  RC_TRY(backward(net, d_x, (int)B));
  if (update) {
    net->step++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f / (1.0f - std::pow(b1, (float)net->step));
    const float c2 = 1.0f / (1.0f - std::pow(b2, (float)net->step));
    HIP_TRY(aldic::cnn_adam(net->wf, net->gw, net->am, net->av, net->nparam, lr, b1, b2, 1e-8f, c1,
                            c2, 0.0f, nullptr, net->wb, s));
  }
  if (d_loss) HIP_TRY(hipMemcpyAsync(d_loss, net->loss, sizeof(float), hipMemcpyDeviceToDevice, s));
  return ALDIC_OK;
}

extern "C" int aldic_cnn_patches_dev(aldic_ctx* ctx, const aldic_images* im,
                                     const double* d_coords, int64_t n, const float* d_prior,
                                     void* d_out) {
  if (!ctx || !im || !im->fm || !im->g) return aldic::set_error(ALDIC_EINVAL, "cnn_patches_dev: bad images");
  if (n == 0) return ALDIC_OK;
  if (!d_coords || !d_out || n < 0) return aldic::set_error(ALDIC_EINVAL, "cnn_patches_dev: NULL array");
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(aldic::cnn_patches(im->fm, im->g, (int)im->M, (int)im->N, d_coords, (int)n, 0, (int)n,
                             d_prior, d_out, ctx->stream));
  return ALDIC_OK;
}

extern "C" int aldic_cnn_predict_dev(aldic_cnn* net, const aldic_images* im,
                                     const double* d_coords, int64_t n, const float* d_prior,
                                     float* d_U0, float* d_F0) {
  RC_TRY(check_net(net));
  if (!im || !im->fm || !im->g) return aldic::set_error(ALDIC_EINVAL, "cnn_predict_dev: bad images");
  if (n == 0) return ALDIC_OK;
  if (!d_coords || !d_U0 || !d_F0 || n < 0)
    return aldic::set_error(ALDIC_EINVAL, "cnn_predict_dev: NULL array");
  hipStream_t s = net->ctx->stream;
  StageTimer t(net->ctx, ALDIC_STAGE_CNN);
  const Layer& c1 = net->L[0];
  const Layer& c2 = net->L[1];
  for (int64_t first = 0; first < n; first += net->bmax) {
    const int cnt = (int)std::min<int64_t>(net->bmax, n - first);
    // patches + the first two convolutions fused, activations in LDS
    HIP_TRY(aldic::cnn_front(im->fm, im->g, (int)im->M, (int)im->N, d_coords, (int)n, (int)first,
                             cnt, d_prior, bfp(net->wb, c1.w_off), net->wf + c1.b_off,
                             bfp(net->wb, c2.w_off), net->wf + c2.b_off, net->act[1], s));
    RC_TRY(forward(net, nullptr, cnt, 2));
    HIP_TRY(aldic::cnn_predict(net->out, 6, (int)first, cnt, d_prior, SCALE_F, SCALE_U, d_U0,
                               d_F0, s));
  }
  return ALDIC_OK;
}

extern "C" int aldic_speckle_mosaic_dev(aldic_ctx* ctx, int64_t M, int64_t N, int32_t tile,
                                        const float* d_warps, uint32_t seed, uint32_t noise_seed,
                                        double density, double radius, double scale, double noise,
                                        double top, float* d_img) {
  if (!ctx || !d_img || M < 1 || N < 1 || tile < 2 || M > 0x7fffffff / 2 || N > 0x7fffffff / 2)
    return aldic::set_error(ALDIC_EINVAL, "speckle_mosaic_dev: bad argument");
  if (d_warps && (M < tile || N < tile))
    return aldic::set_error(ALDIC_EINVAL, "speckle_mosaic_dev: image smaller than one tile");
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(aldic::cnn_mosaic((int)M, (int)N, tile, d_warps, (int)(M / tile), seed, noise_seed,
                            (float)density, (float)radius, (float)scale, (float)noise, (float)top,
                            d_img, ctx->stream));
  return ALDIC_OK;
}
