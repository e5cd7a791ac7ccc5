// aldic_common.h -- shared device helpers for the gfx950 kernels of the ALDIC
// local matching path.  Wave64 everywhere: reductions use DPP row operations
// and the GFX9 row broadcasts, never 32-lane warp idioms.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define ALDIC_WAVE 64

// Two FP16 values in one 32-bit register.
typedef _Float16 half2_t __attribute__((ext_vector_type(2)));

// a * h + c with h the low (fma_lo) or high (fma_hi) FP16 half of a packed
// register, as one mixed-precision FMA (v_fma_mix_f32: the half is widened
// exactly and the FMA rounds once, as an FP32 FMA would).  Written as an
// instruction so that the compiler keeps the half packed instead of
// holding a widened FP32 copy in a register of its own.
__device__ __forceinline__ float fma_lo(float a, half2_t h, float c) {
  float d;
  __asm__("v_fma_mix_f32 %0, %1, %2, %3 op_sel_hi:[0,1,0]" : "=v"(d) : "v"(a), "v"(h), "v"(c));
  return d;
}
__device__ __forceinline__ float fma_hi(float a, half2_t h, float c) {
  float d;
  __asm__("v_fma_mix_f32 %0, %1, %2, %3 op_sel:[0,1,0] op_sel_hi:[0,1,0]"
          : "=v"(d) : "v"(a), "v"(h), "v"(c));
  return d;
}

// Status codes shared with the reference (funICGNQuadtree.m:425-444).
#define ST_ERR (-1)      // MATLAB would throw inside the subset
#define ST_NOCONV 101    // not converged / NaN / warped out of the image
#define ST_MASKED 102    // >= 40% of the subset masked
#define ST_EMPTY 103     // sum |f| < 1e-6
#define ST_DEFER (-100)  // internal: fast kernel handed the subset on

// Sum over the 64 lanes of a wave; every lane receives the total.
__device__ __forceinline__ float wave_sum(float v) {
  v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0xB1, 0xF, 0xF, false));
  v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x4E, 0xF, 0xF, false));
  v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x124, 0xF, 0xF, false));
  v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x128, 0xF, 0xF, false));
  v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x142, 0xA, 0xF, false));
  v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x143, 0xC, 0xF, false));
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 63));
}

__device__ __forceinline__ int wave_sum_i(int v) {
  v += __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x124, 0xF, 0xF, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x128, 0xF, 0xF, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xA, 0xF, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xC, 0xF, false);
  return __builtin_amdgcn_readlane(v, 63);
}

// A value that is the same on every lane, moved to a scalar register so
// wave-uniform state (warp parameters, Cholesky factor) does not occupy
// vector registers.
__device__ __forceinline__ float uni(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
}

// True on every lane if the predicate holds on any lane of the wave.
__device__ __forceinline__ bool wave_any(bool p) { return __ballot(p) != 0ull; }

// MATLAB round(): half away from zero.
__device__ __forceinline__ int64_t mround(double v) { return (int64_t)round(v); }

// Separable form of the reference's local bicubic kernel: the MEX's 16
// coefficients a_ij = rot90(BC*Q*BC',2) (ba_interp2_spline.cpp:310-325,
// func/BicubicBsplineBuild.m:13-16) factor as sum_rc Q[r][c] w_r(dx) w_c(dy)
// with 56*w_k(t) the cubics below (taps -1, 0, +1, +2).
// The same weights times 56.  With m = t (1 - t) the cubics factor as
// 56 w_-1 = m (24 t - 33), 56 w_+2 = m (-24 t - 9), 56 w_+1 = 56 t +
// m (64 t - 11) and 56 w_0 = 56 + 42 m - 56 w_+1 (the four sum to 56): ten
// FMA-class operations for the four instead of twelve by Horner.  A value
// built from them is 3136 times the interpolant, which the zero-normalised
// residual does not see.
__device__ __forceinline__ void spline_weights56(float t, float w[4]) {
  const float m = fmaf(-t, t, t);
  w[0] = m * fmaf(24.0f, t, -33.0f);
  w[3] = m * fmaf(-24.0f, t, -9.0f);
  w[2] = fmaf(m, fmaf(64.0f, t, -11.0f), 56.0f * t);
  w[1] = fmaf(42.0f, m, 56.0f) - w[2];
}

__device__ __forceinline__ void spline_weights(float t, float w[4]) {
  const float s = 1.0f / 56.0f;
  w[0] = t * (-33.0f + t * (57.0f - 24.0f * t)) * s;
  w[1] = (56.0f + t * (-3.0f + t * (-117.0f + 64.0f * t))) * s;
  w[2] = t * (45.0f + t * (75.0f - 64.0f * t)) * s;
  w[3] = t * (-9.0f + t * (-15.0f + 24.0f * t)) * s;
}

// The interpolating cubic B-spline (an option beside the reference's
// kernel, evaluated on prefiltered coefficients): 6 times the basis at the
// taps -1, 0, +1, +2 of the cell, (1-t)^3, 3t^3 - 6t^2 + 4,
// -3t^3 + 3t^2 + 3t + 1, t^3.  A value built from them is 36 times the
// spline.
__device__ __forceinline__ void bspline_weights6(float t, float w[4]) {
  const float u = 1.0f - t;
  const float t2 = t * t, t3 = t2 * t;
  w[0] = u * u * u;
  w[1] = fmaf(3.0f, t3, fmaf(-6.0f, t2, 4.0f));
  w[2] = fmaf(-3.0f, t3, fmaf(3.0f, t2, fmaf(3.0f, t, 1.0f)));
  w[3] = t3;
}
