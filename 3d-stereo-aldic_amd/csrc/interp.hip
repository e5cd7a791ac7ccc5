// interp.hip -- ba_interp2_spline(F, X, Y) on the GPU, double and single,
// with the MEX's semantics (ba_interp2_spline.cpp:235-334):
//   * X is the dim-2 and Y the dim-1 coordinate, 1-based; dx = frac(Y),
//     dy = frac(X) (the "Zach revised" swap, :246-248);
//   * the 4x4 cell starts at 0-based column floor(X)-2 and row floor(Y)-2
//     and is clamped to the image (access(), :60-66);
//   * value = sum_ij a_ij dx^i dy^j, a_ij = rot90(BC*Q*BC',2): each a_ij is a
//     single-precision literal scale times an integer combination of the
//     16 samples (:310-325), so the double path multiplies by the same
//     float-rounded constants as the MEX;
//   * P channels (F is M x N x P), output ND x P.
// One thread per point; the 16 gathers hit L2/MALL for the scattered
// patterns this entry point serves (it is the drop-in for the MEX, not the
// IC-GN hot loop, which uses the separable LDS form in icgn.hip).
#include "aldic_common.h"
#include "aldic_kernels.h"

namespace {

// 56 * w_k(t) = sum_p W56[k][p] t^p, taps k = -1..2 (rows of BC' times 56,
// func/BicubicBsplineBuild.m:13-16).
__constant__ int c_W56[4][4] = {{0, -33, 57, -24}, {56, -3, -117, 64}, {0, 45, 75, -64}, {0, -9, -15, 24}};
// Divisor that turns W56[r][i]*W56[c][j] into the MEX's integer weights of
// a_ij; its magnitude / 3136 is the MEX's literal scale, its sign the sign
// in front of that literal.
__constant__ int c_GDIV[4][4] = {{3136, -168, 168, 448}, {-168, 9, -9, 24}, {168, -9, 9, -24}, {448, 24, -24, 64}};

__device__ __forceinline__ float scale_literal(int g) {
  const int a = g < 0 ? -g : g;
  float s;
  switch (a) {
    case 64: s = 1.0f / 49; break;
    case 24: s = 3.0f / 392; break;
    case 448: s = 1.0f / 7; break;
    case 9: s = (1.0f / 3136) * 9; break;
    case 168: s = 3.0f / 56; break;
    default: s = 1.0f; break;
  }
  return g < 0 ? -s : s;
}

__device__ __forceinline__ int64_t clampl(int64_t v, int64_t lo, int64_t hi) {
  return v < lo ? lo : (v > hi ? hi : v);
}

template <typename T>
__device__ __forceinline__ T tfloor(T v);
template <>
__device__ __forceinline__ double tfloor<double>(double v) { return floor(v); }
template <>
__device__ __forceinline__ float tfloor<float>(float v) { return floorf(v); }

template <typename T>
__global__ __launch_bounds__(256) void interp_kernel(const T* __restrict__ F, int64_t M, int64_t N,
                                                     int64_t P, const T* __restrict__ X,
                                                     const T* __restrict__ Y, int64_t ND,
                                                     T* __restrict__ Z) {
  const int64_t LO = M * N;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ND;
       i += (int64_t)gridDim.x * blockDim.x) {
    const T x = X[i], y = Y[i];
    const T xf = tfloor(x), yf = tfloor(y);
    const T dx = y - yf, dy = x - xf;
    const T dxx = dx * dx, dyy = dy * dy;
    const T dxp[4] = {T(1), dx, dxx, dxx * dx};
    const T dyp[4] = {T(1), dy, dyy, dyy * dy};
    // int(x_floor-1): non-finite or huge values are clamped (their result
    // is NaN through dx, dy in the MEX as well)
    double xb = (double)xf - 1.0, yb = (double)yf - 1.0;
    xb = (xb > -1e9) ? (xb < 1e9 ? xb : 1e9) : -1e9;
    yb = (yb > -1e9) ? (yb < 1e9 ? yb : 1e9) : -1e9;
    const int64_t xi = (int64_t)xb, yi = (int64_t)yb;
    int64_t idx[4][4];
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
      for (int c = 0; c < 4; c++)
        idx[r][c] = M * clampl(xi - 1 + c, 0, N - 1) + clampl(yi - 1 + r, 0, M - 1);
    for (int64_t ch = 0; ch < P; ch++) {
      T Q[4][4];
#pragma unroll
      for (int r = 0; r < 4; r++)
#pragma unroll
        for (int c = 0; c < 4; c++) Q[r][c] = F[idx[r][c] + ch * LO];
      T out = T(0);
#pragma unroll
      for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) {
          T coef;
          if (a == 0 && b == 0) {
            coef = Q[1][1];
          } else {
            T s = T(0);
#pragma unroll
            for (int r = 0; r < 4; r++)
#pragma unroll
              for (int c = 0; c < 4; c++)
                s += (T)(c_W56[r][a] * c_W56[c][b] / c_GDIV[a][b]) * Q[r][c];
            coef = (T)scale_literal(c_GDIV[a][b]) * s;
          }
          const T term = coef * dxp[a] * dyp[b];
          out = (a == 0 && b == 0) ? term : out + term;
        }
      Z[i + ch * ND] = out;
    }
  }
}

int grid_for(int64_t n) {
  int64_t g = (n + 255) / 256;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return (int)g;
}

}  // namespace

namespace aldic {
hipError_t launch_interp_f64(const double* F, int64_t M, int64_t N, int64_t P, const double* X,
                             const double* Y, int64_t ND, double* Z, hipStream_t s) {
  if (ND <= 0) return hipSuccess;
  interp_kernel<double><<<grid_for(ND), 256, 0, s>>>(F, M, N, P, X, Y, ND, Z);
  return hipGetLastError();
}
hipError_t launch_interp_f32(const float* F, int64_t M, int64_t N, int64_t P, const float* X,
                             const float* Y, int64_t ND, float* Z, hipStream_t s) {
  if (ND <= 0) return hipSuccess;
  interp_kernel<float><<<grid_for(ND), 256, 0, s>>>(F, M, N, P, X, Y, ND, Z);
  return hipGetLastError();
}
}  // namespace aldic
