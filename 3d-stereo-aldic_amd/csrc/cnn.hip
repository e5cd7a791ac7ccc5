// cnn.hip -- the learned initial guess: a small CNN that maps a 48 x 48
// reference/deformed patch pair to the six first-order warp parameters
// [u_x v_x u_y v_y u v] that IC-GN starts from (no counterpart in the
// reference; it replaces the integer search as the starting point).
//
// Every layer is a GEMM on the matrix cores: convolutions are im2col (NHWC
// activations, K = 9*Cin padded to a multiple of 16) times the weight matrix
// W[Cout][K]; the GEMM uses v_mfma_f32_32x32x16_bf16 (BF16 operands, FP32
// accumulation) with both operands K-contiguous so every lane fetches its
// 8-element fragment with one 16-byte load.  Backward passes reuse the same
// GEMM on transposed operands (explicit transposes), split over K when the
// reduction runs over the batch.
#include <hip/hip_bf16.h>

#include <algorithm>

#include "aldic_common.h"
#include "aldic_kernels.h"

namespace {

typedef __bf16 bf16;
typedef bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ bf16 to_bf16(float v) { return (bf16)v; }
__device__ __forceinline__ float from_bf16(bf16 v) { return (float)v; }

// Elements 4 q0 .. 4 q0 + 3 of an MFMA 32x32 accumulator plus their biases,
// optional ReLU, as four bf16 in one 8-byte word (the rounding of to_bf16).
__device__ __forceinline__ uint2 bias_relu_pack4(const floatx16& acc, int q0, float4 b, bool relu) {
  float v[4] = {acc[4 * q0] + b.x, acc[4 * q0 + 1] + b.y, acc[4 * q0 + 2] + b.z,
                acc[4 * q0 + 3] + b.w};
  if (relu) {
#pragma unroll
    for (int k = 0; k < 4; k++) v[k] = fmaxf(v[k], 0.0f);
  }
  uint32_t h[4];
#pragma unroll
  for (int k = 0; k < 4; k++) h[k] = __builtin_bit_cast(uint16_t, to_bf16(v[k]));
  return make_uint2(h[0] | (h[1] << 16), h[2] | (h[3] << 16));
}

// ---------------------------------------------------------------------------
// GEMM  C[M x N] = A[M x K] * Bt[N x K]^T
//   Bt row stride ldb (elements); K % 16 == 0; ldb a multiple of 8.  One
//   wave computes 32 rows x 32*NB columns with NB MFMA 32x32x16
//   accumulators; a 256-thread block stacks 4 waves along M (128 rows).
//   blockIdx.z splits K into chunks of kchunk (multiple of 16).
//   Fragment layout of v_mfma_f32_32x32x16_bf16 (wave64):
//     A: lane l holds A[l % 32][8*(l/32) .. 8*(l/32)+7]
//     B: lane l holds B[8*(l/32) .. +7][l % 32]  (= Bt[l % 32][same k])
//     D: element i of lane l is D[8*(i/4) + 4*(l/32) + i%4][l % 32]
//   A comes from one of three loaders (AMODE):
//     0  a dense row-major matrix, row stride lda (multiple of 8);
//     1  implicit im2col of a 3x3 / pad 1 convolution over NHWC bf16
//        activations with Cin a multiple of 8: row (n, oy, ox), column
//        (ky*3 + kx)*Cin + ci, so 8 consecutive columns are 8 channels of
//        one tap -- one 16-byte load or zero padding;
//     2  the same with Cin = 2: 8 columns are 4 taps x 2 channels.
//   EPI 0: write FP32 partial sums to C + z*M*ldc;
//   EPI 1: add bias[n], optional ReLU, write bf16 Cb and/or FP32 C.
// ---------------------------------------------------------------------------
struct ConvGeom {
  int H, W, Cin, stride, Ho, Wo;
};

template <int AMODE>
struct ALoader {
  const bf16* base;
  int n, iy0, ix0;  // conv: batch index, top-left input coordinate of the row
  bool ok;
  __device__ void init(const bf16* A, int lda, int row, bool rowok, const ConvGeom& g, int kh) {
    ok = rowok;
    if (AMODE == 0) {
      base = A + (int64_t)row * lda + kh;
    } else {
      const int hw = g.Ho * g.Wo;
      n = row / hw;
      const int rem = row - n * hw;
      const int oy = rem / g.Wo, ox = rem - oy * g.Wo;
      iy0 = oy * g.stride - 1;
      ix0 = ox * g.stride - 1;
      base = A + (int64_t)n * g.H * g.W * g.Cin;
    }
  }
  __device__ bf16x8 load(int k, int kh, const ConvGeom& g) const {
    bf16x8 v = {0, 0, 0, 0, 0, 0, 0, 0};
    if (!ok) return v;
    if (AMODE == 0) {
      v = *(const bf16x8*)(base + k);
    } else if (AMODE == 1) {
      const int kk = k + kh;
      const int tap = kk / g.Cin, ci = kk - tap * g.Cin;
      const int ky = tap / 3, kx = tap - 3 * ky;
      const int iy = iy0 + ky, ix = ix0 + kx;
      if (tap < 9 && iy >= 0 && iy < g.H && ix >= 0 && ix < g.W)
        v = *(const bf16x8*)(base + ((int64_t)iy * g.W + ix) * g.Cin + ci);
    } else {
      const int tap0 = (k + kh) / 2;
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int tap = tap0 + q;
        const int ky = tap / 3, kx = tap - 3 * ky;
        const int iy = iy0 + ky, ix = ix0 + kx;
        if (tap < 9 && iy >= 0 && iy < g.H && ix >= 0 && ix < g.W) {
          const bf16* src = base + ((int64_t)iy * g.W + ix) * 2;
          v[2 * q] = src[0];
          v[2 * q + 1] = src[1];
        }
      }
    }
    return v;
  }
};

template <int NB, int EPI, int AMODE>
__global__ __launch_bounds__(256) void gemm_kernel(const bf16* __restrict__ A, int lda,
                                                   const bf16* __restrict__ Bt, int ldb, int M,
                                                   int N, int K, int kchunk, float* __restrict__ C,
                                                   int ldc, const float* __restrict__ bias,
                                                   int relu, bf16* __restrict__ Cb, int ldcb,
                                                   ConvGeom geom) {
  const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
  const int m0 = blockIdx.x * 128 + wave * 32;
  const int n0 = blockIdx.y * (32 * NB);
  const int kbeg = blockIdx.z * kchunk;
  const int kend = min(K, kbeg + kchunk);
  const int r = lane % 32, kh = 8 * (lane / 32);
  floatx16 acc[NB];
#pragma unroll
  for (int j = 0; j < NB; j++)
#pragma unroll
    for (int i = 0; i < 16; i++) acc[j][i] = 0.0f;
  ALoader<AMODE> al;
  al.init(A, lda, m0 + r, (m0 + r) < M, geom, kh);
  const bf16* bp[NB];
  bool bcol[NB];
#pragma unroll
  for (int j = 0; j < NB; j++) {
    bcol[j] = (n0 + 32 * j + r) < N;
    bp[j] = Bt + (int64_t)(n0 + 32 * j + r) * ldb + kh;
  }
  const bf16x8 zero = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int k = kbeg; k < kend; k += 16) {
    const bf16x8 a = al.load(k, kh, geom);
#pragma unroll
    for (int j = 0; j < NB; j++) {
      const bf16x8 b = bcol[j] ? *(const bf16x8*)(bp[j] + k) : zero;
      acc[j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[j], 0, 0, 0);
    }
  }
#pragma unroll
  for (int j = 0; j < NB; j++) {
    const int col = n0 + 32 * j + r;
    if (col >= N) continue;
    const float bv = (EPI == 1 && bias) ? bias[col] : 0.0f;
#pragma unroll
    for (int i = 0; i < 16; i++) {
      const int row = m0 + 8 * (i / 4) + 4 * (lane / 32) + (i % 4);
      if (row >= M) continue;
      float v = acc[j][i];
      if (EPI == 0) {
        C[(int64_t)blockIdx.z * M * ldc + (int64_t)row * ldc + col] = v;
      } else {
        v += bv;
        if (relu) v = fmaxf(v, 0.0f);
        if (C) C[(int64_t)row * ldc + col] = v;
        if (Cb) Cb[(int64_t)row * ldcb + col] = to_bf16(v);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Inference-shaped GEMM: the whole weight matrix Bt[N][K] (N = 32*NB) is
// staged once per workgroup in LDS (rows padded by 16 bytes), and the
// workgroup loops over 256-row tiles of A.  Each wave computes 64 rows x N
// (two MFMA row tiles share every B fragment); the next A fragments are
// loaded while the current ones feed the matrix cores.  Epilogue: bias,
// optional ReLU, bf16 store (row stride N).
// ---------------------------------------------------------------------------
template <int NB, int AMODE>
__global__ __launch_bounds__(256) void gemm_lds_kernel(const bf16* __restrict__ A, int lda,
                                                       const bf16* __restrict__ Bt, int M, int K,
                                                       const float* __restrict__ bias, int relu,
                                                       bf16* __restrict__ Cb, int ldc,
                                                       ConvGeom geom, int ntiles) {
  extern __shared__ bf16x8 sB[];
  constexpr int N = 32 * NB;
  const int kq = K / 8, rs = kq + 1;
  for (int e = threadIdx.x; e < N * kq; e += 256) {
    const int row = e / kq, q = e - row * kq;
    sB[row * rs + q] = *(const bf16x8*)(Bt + (int64_t)row * K + q * 8);
  }
  __syncthreads();
  const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
  const int r = lane % 32, kh = 8 * (lane / 32), hq = lane / 32;
  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int m0 = tile * 256 + wave * 64;
    ALoader<AMODE> a0, a1;
    a0.init(A, lda, m0 + r, (m0 + r) < M, geom, kh);
    a1.init(A, lda, m0 + 32 + r, (m0 + 32 + r) < M, geom, kh);
    floatx16 acc0[NB], acc1[NB];
#pragma unroll
    for (int j = 0; j < NB; j++)
#pragma unroll
      for (int i = 0; i < 16; i++) {
        acc0[j][i] = 0.0f;
        acc1[j][i] = 0.0f;
      }
    bf16x8 x0 = a0.load(0, kh, geom), x1 = a1.load(0, kh, geom);
    for (int k = 0; k < K; k += 16) {
      bf16x8 n0 = x0, n1 = x1;
      if (k + 16 < K) {
        n0 = a0.load(k + 16, kh, geom);
        n1 = a1.load(k + 16, kh, geom);
      }
#pragma unroll
      for (int j = 0; j < NB; j++) {
        const bf16x8 b = sB[(32 * j + r) * rs + k / 8 + hq];
        acc0[j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(b, x0, acc0[j], 0, 0, 0);
        acc1[j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(b, x1, acc1[j], 0, 0, 0);
      }
      x0 = n0;
      x1 = n1;
    }
    // computed transposed (weights as the A operand): element i of lane
    // (r, hq) is column 32 j + 8 (i / 4) + 4 hq + i % 4 of row m0 + 32 h + r,
    // so runs of 4 columns go out as 8-byte words (ldc % 4 == 0)
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const int row = m0 + 32 * h + r;
      if (row >= M) continue;
#pragma unroll
      for (int j = 0; j < NB; j++)
#pragma unroll
        for (int q0 = 0; q0 < 4; q0++) {
          const int col = 32 * j + 8 * q0 + 4 * hq;
          const float4 b = bias ? *(const float4*)(bias + col) : make_float4(0, 0, 0, 0);
          *(uint2*)(Cb + (int64_t)row * ldc + col) =
              bias_relu_pack4(h ? acc1[j] : acc0[j], q0, b, relu != 0);
        }
    }
  }
}

// ---------------------------------------------------------------------------
// 3x3 / pad 1 convolution as an implicit GEMM with compile-time geometry
// of the reduction: CIN input channels (2, 32 or 64), K = 9*CIN (32 for
// CIN = 2, zero padded), N = 32*NB output channels.  Weights live in LDS
// for the whole workgroup; per 256-row tile every lane computes the nine
// tap offsets of its two rows once, and the fully unrolled k-loop has no
// index arithmetic left, so the compiler can issue the A loads of many
// k-steps ahead of the matrix-core work that consumes them.
// ---------------------------------------------------------------------------
template <int NB, int CIN>
__global__ __launch_bounds__(256) void conv_lds_kernel(const bf16* __restrict__ x,
                                                       const bf16* __restrict__ Wt, int M,
                                                       const float* __restrict__ bias, int relu,
                                                       bf16* __restrict__ out, ConvGeom g,
                                                       int ntiles) {
  constexpr int N = 32 * NB;
  constexpr int K = CIN == 2 ? 32 : 9 * CIN;
  constexpr int KQ = K / 8, RS = KQ + 1;
  constexpr int STEPS = K / 16;
  __shared__ bf16x8 sB[N * RS];
  for (int e = threadIdx.x; e < N * KQ; e += 256) {
    const int row = e / KQ, q = e - row * KQ;
    sB[row * RS + q] = *(const bf16x8*)(Wt + (int64_t)row * K + q * 8);
  }
  __syncthreads();
  const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
  const int r = lane % 32, hq = lane / 32, kh = 8 * hq;
  const int hw = g.Ho * g.Wo;
  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int m0 = tile * 256 + wave * 64;
    const bf16* base[2];
    int off[2][9];
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const int row = m0 + 32 * h + r;
      const bool ok = row < M;
      const int n = ok ? row / hw : 0;
      const int rem = row - n * hw;
      const int oy = rem / g.Wo, ox = rem - (rem / g.Wo) * g.Wo;
      base[h] = x + (int64_t)n * g.H * g.W * CIN;
#pragma unroll
      for (int t = 0; t < 9; t++) {
        const int iy = oy * g.stride - 1 + t / 3, ix = ox * g.stride - 1 + t % 3;
        off[h][t] = (ok && iy >= 0 && iy < g.H && ix >= 0 && ix < g.W)
                        ? (iy * g.W + ix) * CIN : -1;
      }
    }
    floatx16 acc[2][NB];
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
      for (int j = 0; j < NB; j++)
#pragma unroll
        for (int i = 0; i < 16; i++) acc[h][j][i] = 0.0f;
    if (CIN == 2) {
#pragma unroll
      for (int st = 0; st < STEPS; st++) {
        bf16x8 a[2];
#pragma unroll
        for (int h = 0; h < 2; h++) {
          bf16x8 v = {0, 0, 0, 0, 0, 0, 0, 0};
          // columns 16*st + kh .. +7 are taps 8*st + 4*hq .. +3, two channels each
#pragma unroll
          for (int q = 0; q < 4; q++) {
            const int t0 = 8 * st + q, t1 = 8 * st + 4 + q;
            const int o = hq ? (t1 < 9 ? off[h][t1 < 9 ? t1 : 0] : -1)
                             : (t0 < 9 ? off[h][t0 < 9 ? t0 : 0] : -1);
            if (o >= 0) {
              const uint32_t w2 = *(const uint32_t*)(base[h] + o);
              v[2 * q] = __builtin_bit_cast(bf16, (uint16_t)(w2 & 0xffffu));
              v[2 * q + 1] = __builtin_bit_cast(bf16, (uint16_t)(w2 >> 16));
            }
          }
          a[h] = v;
        }
#pragma unroll
        for (int j = 0; j < NB; j++) {
          const bf16x8 b = sB[(32 * j + r) * RS + 2 * st + hq];
          acc[0][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(b, a[0], acc[0][j], 0, 0, 0);
          acc[1][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(b, a[1], acc[1][j], 0, 0, 0);
        }
      }
    } else {
      // one tap = SPT k-steps; the fragments of tap t+1 are in flight while
      // tap t feeds the matrix cores
      constexpr int SPT = CIN == 2 ? 1 : CIN / 16;
      bf16x8 cur[2][SPT], nxt[2][SPT];
      auto load_tap = [&](int t, bf16x8 (&dst)[2][SPT]) {
#pragma unroll
        for (int h = 0; h < 2; h++) {
          const int o = off[h][t];
#pragma unroll
          for (int c = 0; c < SPT; c++) {
            bf16x8 v = {0, 0, 0, 0, 0, 0, 0, 0};
            if (o >= 0) v = *(const bf16x8*)(base[h] + o + c * 16 + kh);
            dst[h][c] = v;
          }
        }
      };
      load_tap(0, cur);
#pragma unroll
      for (int t = 0; t < 9; t++) {
        if (t + 1 < 9) load_tap(t + 1, nxt);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int c = 0; c < SPT; c++) {
          const int st = t * SPT + c;
#pragma unroll
          for (int j = 0; j < NB; j++) {
            const bf16x8 b = sB[(32 * j + r) * RS + 2 * st + hq];
            acc[0][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(b, cur[0][c], acc[0][j], 0, 0, 0);
            acc[1][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(b, cur[1][c], acc[1][j], 0, 0, 0);
          }
        }
        __builtin_amdgcn_sched_barrier(0);
        if (t + 1 < 9) {
#pragma unroll
          for (int h = 0; h < 2; h++)
#pragma unroll
            for (int c = 0; c < SPT; c++) cur[h][c] = nxt[h][c];
        }
      }
    }
    // computed transposed (weights as the A operand): element i of lane
    // (r, hq) is channel 32 j + 8 (i / 4) + 4 hq + i % 4 of output row
    // m0 + 32 h + r, so runs of 4 channels go out as 8-byte words
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const int row = m0 + 32 * h + r;
      if (row >= M) continue;
#pragma unroll
      for (int j = 0; j < NB; j++)
#pragma unroll
        for (int q0 = 0; q0 < 4; q0++) {
          const int ch = 32 * j + 8 * q0 + 4 * hq;
          const float4 b = bias ? *(const float4*)(bias + ch) : make_float4(0, 0, 0, 0);
          *(uint2*)(out + (int64_t)row * N + ch) = bias_relu_pack4(acc[h][j], q0, b, relu != 0);
        }
    }
  }
}

// sum of split-K partials (+ bias, ReLU) -> FP32 and/or bf16
__global__ void splitk_reduce_kernel(const float* __restrict__ P, int splits, int M, int N, int ldp,
                                     const float* __restrict__ bias, int relu,
                                     float* __restrict__ C, int ldc, bf16* __restrict__ Cb,
                                     int ldcb) {
  const int64_t total = (int64_t)M * N;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int row = (int)(e / N), col = (int)(e % N);
    float v = 0.0f;
    for (int z = 0; z < splits; z++) v += P[(int64_t)z * M * ldp + (int64_t)row * ldp + col];
    if (bias) v += bias[col];
    if (relu) v = fmaxf(v, 0.0f);
    if (C) C[(int64_t)row * ldc + col] = v;
    if (Cb) Cb[(int64_t)row * ldcb + col] = to_bf16(v);
  }
}

// ---------------------------------------------------------------------------
// im2col / col2im for 3x3 convolutions, pad 1, NHWC bf16 activations.
// col row (n, oy, ox), column (ky*3 + kx)*Cin + ci, zero padded to Kp.
// ---------------------------------------------------------------------------
__global__ void im2col_kernel(const bf16* __restrict__ x, int B, int H, int W, int Cin, int stride,
                              int Ho, int Wo, int Kp, bf16* __restrict__ col) {
  const int64_t total = (int64_t)B * Ho * Wo * Kp;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int kk = (int)(e % Kp);
    const int64_t m = e / Kp;
    const int ox = (int)(m % Wo), oy = (int)((m / Wo) % Ho), n = (int)(m / ((int64_t)Wo * Ho));
    bf16 v = to_bf16(0.0f);
    if (kk < 9 * Cin) {
      const int ci = kk % Cin, t = kk / Cin, kx = t % 3, ky = t / 3;
      const int iy = oy * stride + ky - 1, ix = ox * stride + kx - 1;
      if (iy >= 0 && iy < H && ix >= 0 && ix < W)
        v = x[(((int64_t)n * H + iy) * W + ix) * Cin + ci];
    }
    col[e] = v;
  }
}

// dx(n, y, x, ci) = sum over the (ky, kx) taps that read it of dcol
__global__ void col2im_kernel(const float* __restrict__ dcol, int B, int H, int W, int Cin,
                              int stride, int Ho, int Wo, int Kp, float* __restrict__ dx) {
  const int64_t total = (int64_t)B * H * W * Cin;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int ci = (int)(e % Cin);
    const int64_t p = e / Cin;
    const int x = (int)(p % W), y = (int)((p / W) % H), n = (int)(p / ((int64_t)W * H));
    float s = 0.0f;
    for (int ky = 0; ky < 3; ky++) {
      const int ty = y + 1 - ky;
      if (ty < 0 || ty % stride) continue;
      const int oy = ty / stride;
      if (oy >= Ho) continue;
      for (int kx = 0; kx < 3; kx++) {
        const int tx = x + 1 - kx;
        if (tx < 0 || tx % stride) continue;
        const int ox = tx / stride;
        if (ox >= Wo) continue;
        s += dcol[(((int64_t)n * Ho + oy) * Wo + ox) * Kp + (ky * 3 + kx) * Cin + ci];
      }
    }
    dx[e] = s;
  }
}

// out[c][r] = in[r][c] (bf16 -> bf16 or FP32 -> bf16), 32 x 32 LDS tiles
template <typename TI>
__global__ void transpose_kernel(const TI* __restrict__ in, int R, int Cc, int ldi,
                                 bf16* __restrict__ out, int ldo) {
  __shared__ float t[32][33];
  const int r0 = blockIdx.y * 32, c0 = blockIdx.x * 32;
  const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;  // 32 x 8
  for (int i = ty; i < 32; i += 8) {
    const int r = r0 + i, c = c0 + tx;
    t[i][tx] = (r < R && c < Cc) ? (float)in[(int64_t)r * ldi + c] : 0.0f;
  }
  __syncthreads();
  for (int i = ty; i < 32; i += 8) {
    const int c = c0 + i, r = r0 + tx;
    if (c < Cc && r < R) out[(int64_t)c * ldo + r] = to_bf16(t[tx][i]);
  }
}

// g = g * (y > 0) with y the layer's post-ReLU output; FP32 in, bf16 out
__global__ void relu_back_kernel(const float* __restrict__ g, const bf16* __restrict__ y,
                                 int64_t n, bf16* __restrict__ gb, float* __restrict__ gf) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    const float v = from_bf16(y[e]) > 0.0f ? g[e] : 0.0f;
    gb[e] = to_bf16(v);
    if (gf) gf[e] = v;
  }
}

// column sums of an M x N FP32 matrix (bias gradients)
__global__ void colsum_kernel(const float* __restrict__ g, int M, int ldg, float* __restrict__ out) {
  const int col = blockIdx.x;
  float s = 0.0f;
  for (int r = threadIdx.x; r < M; r += blockDim.x) s += g[(int64_t)r * ldg + col];
  s = wave_sum(s);
  __shared__ float red[4];
  if (threadIdx.x % 64 == 0) red[threadIdx.x / 64] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[col] = red[0] + red[1] + red[2] + red[3];
}

__global__ void f32_to_bf16_kernel(const float* __restrict__ in, int64_t n, bf16* __restrict__ out) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x)
    out[e] = to_bf16(in[e]);
}

// Adam on FP32 master weights; refreshes the bf16 copy.  mask (may be
// null) zeroes the update of padding entries.
__global__ void adam_kernel(float* __restrict__ w, const float* __restrict__ g,
                            float* __restrict__ m, float* __restrict__ v, int64_t n, float lr,
                            float b1, float b2, float eps, float c1, float c2, float wd,
                            const float* __restrict__ mask, bf16* __restrict__ wb) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    const float gg = g[e] + wd * w[e];
    const float mm = b1 * m[e] + (1.0f - b1) * gg;
    const float vv = b2 * v[e] + (1.0f - b2) * gg * gg;
    m[e] = mm;
    v[e] = vv;
    float upd = lr * (mm * c1) / (sqrtf(vv * c2) + eps);
    if (mask) upd *= mask[e];
    const float nw = w[e] - upd;
    w[e] = nw;
    if (wb) wb[e] = to_bf16(nw);
  }
}

// loss = mean over the batch of sum_k s_k (pred_k - label_k)^2; d = dloss/dpred
__global__ void mse_kernel(const float* __restrict__ pred, int ldp, const float* __restrict__ lab,
                           const float* __restrict__ scale, int B, float* __restrict__ d, int ldd,
                           float* __restrict__ loss) {
  float s = 0.0f;
  for (int i = threadIdx.x; i < B * 6; i += blockDim.x) {
    const int b = i / 6, k = i % 6;
    const float e = pred[(int64_t)b * ldp + k] - lab[b * 6 + k];
    s += scale[k] * e * e;
    d[(int64_t)b * ldd + k] = 2.0f * scale[k] * e / (float)B;
  }
  s = wave_sum(s);
  __shared__ float red[16];
  if (threadIdx.x % 64 == 0) red[threadIdx.x / 64] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    float t = 0.0f;
    for (int k = 0; k < (int)(blockDim.x / 64); k++) t += red[k];
    *loss = t / (float)B;
  }
}

// ---------------------------------------------------------------------------
// Patch pairs: for each node, the 48 x 48 reference patch centred on the
// node and the deformed patch centred on node + shift (integer prior),
// each normalised to zero mean and unit variance, as NHWC bf16 [B,48,48,2].
// Image pixel (i, j) 1-based; patch pixel (py, px) = image
// (c - 24 + px, c' - 24 + py) along (dim 1, dim 2).  One 256-thread
// workgroup per node; the moments are summed in a fixed order (9 values
// per thread, wave sums, then the 4 waves in order) that the fused front
// end repeats, so both produce the same patches.
// ---------------------------------------------------------------------------
constexpr int PATCH = 48;

struct PatchMoments {
  float mf, sf, mg, sg;
};

// the 36 pixels of each image a thread owns: p = tid + 256 k
__device__ __forceinline__ void patch_raw(const float* __restrict__ f, const float* __restrict__ g,
                                          int M, int N, int cx, int cy, int sxs, int sys,
                                          float* vf, float* vg) {
  constexpr int NP = PATCH * PATCH;
  const int tid = threadIdx.x;
#pragma unroll
  for (int k = 0; k < NP / 256; k++) {
    const int p = tid + 256 * k;
    const int px = p % PATCH, py = p / PATCH;
    const int i = min(max(cx - PATCH / 2 + px, 1), M), j = min(max(cy - PATCH / 2 + py, 1), N);
    vf[k] = f[(int64_t)(i - 1) + (int64_t)(j - 1) * M];
    const int i2 = min(max(cx + sxs - PATCH / 2 + px, 1), M);
    const int j2 = min(max(cy + sys - PATCH / 2 + py, 1), N);
    vg[k] = g[(int64_t)(i2 - 1) + (int64_t)(j2 - 1) * M];
  }
}

// The same pixels when the whole 48 x 48 window lies inside the image (no
// clamp): one load per pixel from a per-node base and a per-thread byte
// offset fixed for the launch.
__device__ __forceinline__ bool patch_inside(int M, int N, int cx, int cy) {
  return cx - PATCH / 2 >= 1 && cx + PATCH / 2 - 1 <= M && cy - PATCH / 2 >= 1 &&
         cy + PATCH / 2 - 1 <= N;
}

__device__ __forceinline__ void patch_load(const float* __restrict__ f, const float* __restrict__ g,
                                           int M, int N, int cx, int cy, int sxs, int sys,
                                           float* vf, float* vg) {
  constexpr int NP = PATCH * PATCH;
  if (patch_inside(M, N, cx, cy) && patch_inside(M, N, cx + sxs, cy + sys)) {
    const char* bf = reinterpret_cast<const char*>(
        f + ((int64_t)(cy - PATCH / 2 - 1) * M + (cx - PATCH / 2 - 1)));
    const char* bg = reinterpret_cast<const char*>(
        g + ((int64_t)(cy + sys - PATCH / 2 - 1) * M + (cx + sxs - PATCH / 2 - 1)));
#pragma unroll
    for (int k = 0; k < NP / 256; k++) {
      const uint32_t p = threadIdx.x + 256 * k, py = p / PATCH, px = p - PATCH * py;
      const uint32_t off = (py * (uint32_t)M + px) * 4u;
      vf[k] = *reinterpret_cast<const float*>(bf + off);
      vg[k] = *reinterpret_cast<const float*>(bg + off);
    }
  } else {
    patch_raw(f, g, M, N, cx, cy, sxs, sys, vf, vg);
  }
}

__device__ __forceinline__ PatchMoments patch_moments(const float* vf, const float* vg,
                                                      float* red) {
  constexpr int NP = PATCH * PATCH;
  const int tid = threadIdx.x, lane = tid % 64, wave = tid / 64;
  float st[4] = {0, 0, 0, 0};
#pragma unroll
  for (int k = 0; k < NP / 256; k++) {
    st[0] += vf[k];
    st[1] += vf[k] * vf[k];
    st[2] += vg[k];
    st[3] += vg[k] * vg[k];
  }
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const float v = wave_sum(st[q]);
    if (lane == 0) red[q * 4 + wave] = v;
  }
  __syncthreads();
  float S[4];
#pragma unroll
  for (int q = 0; q < 4; q++) S[q] = red[q * 4] + red[q * 4 + 1] + red[q * 4 + 2] + red[q * 4 + 3];
  __syncthreads();  // red is reused by the next node
  PatchMoments m;
  m.mf = S[0] / NP;
  m.mg = S[2] / NP;
  m.sf = rsqrtf(fmaxf(S[1] / NP - m.mf * m.mf, 1e-12f));
  m.sg = rsqrtf(fmaxf(S[3] / NP - m.mg * m.mg, 1e-12f));
  return m;
}

__device__ __forceinline__ uint32_t patch_pair(float vf, float vg, const PatchMoments& m) {
  const uint32_t lo = __builtin_bit_cast(uint16_t, to_bf16((vf - m.mf) * m.sf));
  const uint32_t hi = __builtin_bit_cast(uint16_t, to_bf16((vg - m.mg) * m.sg));
  return lo | (hi << 16);
}

__device__ __forceinline__ void node_shift(const double* coords, int n, const float* prior,
                                           int node, int* cx, int* cy, int* sxs, int* sys) {
  *cx = (int)mround(coords[node]);
  *cy = (int)mround(coords[node + n]);
  *sxs = prior ? (int)rintf(prior[2 * node]) : 0;
  *sys = prior ? (int)rintf(prior[2 * node + 1]) : 0;
}

__global__ __launch_bounds__(256) void patch_kernel(const float* __restrict__ f,
                                                    const float* __restrict__ g, int M, int N,
                                                    const double* __restrict__ coords, int n,
                                                    int first, int count,
                                                    const float* __restrict__ prior,
                                                    bf16* __restrict__ out) {
  constexpr int NP = PATCH * PATCH;
  __shared__ float red[16];
  for (int local = blockIdx.x; local < count; local += gridDim.x) {
    int cx, cy, sxs, sys;
    node_shift(coords, n, prior, first + local, &cx, &cy, &sxs, &sys);
    float vf[NP / 256], vg[NP / 256];
    patch_raw(f, g, M, N, cx, cy, sxs, sys, vf, vg);
    const PatchMoments m = patch_moments(vf, vg, red);
    // one 4-byte store per pixel: (reference, deformed) channel pair
    uint32_t* o = reinterpret_cast<uint32_t*>(out + (int64_t)local * NP * 2);
#pragma unroll
    for (int k = 0; k < NP / 256; k++) o[threadIdx.x + 256 * k] = patch_pair(vf[k], vg[k], m);
  }
}

// ---------------------------------------------------------------------------
// Network output -> IC-GN starting point.  out[b][0..5] are the normalised
// [u_x v_x u_y v_y du dv]; U0 = round(prior) + du*su, F0 = F*sf.
// ---------------------------------------------------------------------------
__global__ void predict_kernel(const float* __restrict__ out, int ldo, int first, int count,
                               const float* __restrict__ prior, float sf, float su,
                               float* __restrict__ U0, float* __restrict__ F0) {
  const int local = blockIdx.x * blockDim.x + threadIdx.x;
  if (local >= count) return;
  const float* y = out + (int64_t)local * ldo;
  const int b = first + local;
  const float px = prior ? rintf(prior[2 * b]) : 0.0f;
  const float py = prior ? rintf(prior[2 * b + 1]) : 0.0f;
  U0[2 * b] = px + y[4] * su;
  U0[2 * b + 1] = py + y[5] * su;
#pragma unroll
  for (int k = 0; k < 4; k++) F0[4 * b + k] = y[k] * sf;
}

// ---------------------------------------------------------------------------
// Training data: a mosaic of square tiles of seeded speckle.  Spots are
// generated per 8 x 8 px cell from a counter-based hash (Poisson count with
// mean density*64, uniform positions), so any pixel can be rendered without
// a spot list.  The reference image samples the spot field at X = x; in the
// deformed image every tile has its own affine displacement
// u(X) = t + A (X - c) about its centre node c, and pixel x samples the
// field at X = c + (I + A)^-1 (x - c - t).  Grey levels: 20 + 215*min(I/2,1)
// scaled by `scale`, plus Gaussian noise, rounded and clipped to [0, top].
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t mix32(uint32_t h) {
  h ^= h >> 16;
  h *= 0x7feb352du;
  h ^= h >> 15;
  h *= 0x846ca68bu;
  h ^= h >> 16;
  return h;
}
__device__ __forceinline__ float hash_unit(uint32_t a, uint32_t b, uint32_t c) {
  const uint32_t h = mix32(a * 0x9E3779B1u ^ mix32(b + 0x632BE5ABu * c + 0x85EBCA77u));
  return ((h >> 8) + 0.5f) * (1.0f / 16777216.0f);
}

constexpr int CELL = 8;

__device__ float spot_field(float X, float Y, uint32_t seed, float lambda, float inv_r2,
                            float reach) {
  const int cx0 = (int)floorf((X - reach) / CELL), cx1 = (int)floorf((X + reach) / CELL);
  const int cy0 = (int)floorf((Y - reach) / CELL), cy1 = (int)floorf((Y + reach) / CELL);
  const float e0 = __expf(-lambda);
  float acc = 0.0f;
  for (int cy = cy0; cy <= cy1; cy++)
    for (int cx = cx0; cx <= cx1; cx++) {
      const uint32_t cell = (uint32_t)(cx + 4096) * 16411u + (uint32_t)(cy + 4096);
      // Poisson count by inversion of the CDF
      const float uc = hash_unit(seed, cell, 0);
      int cnt = 0;
      float pk = e0, cdf = e0;
      while (uc > cdf && cnt < 12) {
        cnt++;
        pk *= lambda / cnt;
        cdf += pk;
      }
      for (int k = 0; k < cnt; k++) {
        const float sx = (cx + hash_unit(seed, cell, 2 * k + 1)) * CELL;
        const float sy = (cy + hash_unit(seed, cell, 2 * k + 2)) * CELL;
        const float dx = X - sx, dy = Y - sy;
        const float d2 = dx * dx + dy * dy;
        if (d2 < reach * reach) acc += __expf(-d2 * inv_r2);
      }
    }
  return acc;
}

__global__ void mosaic_kernel(int M, int N, int tile, const float* __restrict__ warps,
                              int ntx, uint32_t seed, uint32_t noise_seed, float density,
                              float radius, float scale, float noise, float top,
                              float* __restrict__ img) {
  const int64_t total = (int64_t)M * N;
  const float lambda = density * CELL * CELL;
  const float inv_r2 = 1.0f / (radius * radius);
  const float reach = 4.0f * radius;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(e % M) + 1, j = (int)(e / M) + 1;  // 1-based pixel
    float X = (float)i, Y = (float)j;
    if (warps) {
      // pixels past the last whole tile belong to the last tile
      const int tx = min((i - 1) / tile, ntx - 1), ty = min((j - 1) / tile, N / tile - 1);
      const float* w = warps + 6 * ((int64_t)ty * ntx + tx);
      const float cx = (float)(tx * tile + tile / 2), cy = (float)(ty * tile + tile / 2);
      // w = [u_x v_x u_y v_y u v]; (I + A) = [1+u_x u_y; v_x 1+v_y]
      const float a = 1.0f + w[0], b = w[2], c = w[1], d = 1.0f + w[3];
      const float det = a * d - b * c;
      const float rx = X - cx - w[4], ry = Y - cy - w[5];
      X = cx + (d * rx - b * ry) / det;
      Y = cy + (-c * rx + a * ry) / det;
    }
    const float I = spot_field(X, Y, seed, lambda, inv_r2, reach);
    float gl = (20.0f + 215.0f * fminf(I * 0.5f, 1.0f)) * scale;
    // Box-Muller noise
    const float u1 = hash_unit(noise_seed, (uint32_t)e, 7), u2 = hash_unit(noise_seed, (uint32_t)e, 9);
    gl += noise * scale * sqrtf(-2.0f * __logf(u1)) * __cosf(6.2831853f * u2);
    img[e] = fminf(fmaxf(rintf(gl), 0.0f), top);
  }
}

// ---------------------------------------------------------------------------
// Inference front end, fused: per node, the normalised 48 x 48 patch pair
// (as patch_kernel), conv 2->32 stride 2 (48 -> 24) and conv 32->32
// stride 2 (24 -> 12), with the patch and the 24 x 24 x 32 activation kept
// in LDS.  The matrix-core arithmetic is the same as the layer-by-layer
// path (same fragments, same k order, same bias/ReLU/bf16 rounding), so the
// 12 x 12 x 32 output is identical to it.  One workgroup per node at a
// time (persistent over nodes).
// ---------------------------------------------------------------------------
constexpr int FR_L1C = 40;  // padded channel stride of the L1 activation in LDS

__global__ __launch_bounds__(256, 2) void front_kernel(const float* __restrict__ f,
                                                    const float* __restrict__ g, int M, int N,
                                                    const double* __restrict__ coords, int n,
                                                    int first, int count,
                                                    const float* __restrict__ prior,
                                                    const bf16* __restrict__ W1,
                                                    const float* __restrict__ b1,
                                                    const bf16* __restrict__ W2,
                                                    const float* __restrict__ b2,
                                                    bf16* __restrict__ out) {
  constexpr int NP = PATCH * PATCH;
  constexpr int PP = PATCH + 2;  // the patch with a zero border, so taps need no bounds test
  constexpr int K1 = 32, K2 = 288;
  constexpr int RS1 = K1 / 8 + 1, RS2 = K2 / 8 + 1;
  __shared__ uint32_t sx[PP * PP];             // patch: (ref, def) bf16 pairs
  __shared__ bf16 sa1[24 * 24 * FR_L1C];       // L1 output, NHWC, padded channels
  __shared__ bf16x8 sW1[32 * RS1];
  __shared__ bf16x8 sW2[32 * RS2];
  __shared__ float4 sbias[16];  // b1 then b2, 4 channels per entry
  __shared__ float red[16];
  const int tid = threadIdx.x, lane = tid % 64, wave = tid / 64;
  const int r = lane % 32, hq = lane / 32, kh = 8 * hq;
  for (int e = tid; e < 32 * (K1 / 8); e += 256) {
    const int row = e / (K1 / 8), q = e % (K1 / 8);
    sW1[row * RS1 + q] = *(const bf16x8*)(W1 + row * K1 + q * 8);
  }
  for (int e = tid; e < 32 * (K2 / 8); e += 256) {
    const int row = e / (K2 / 8), q = e % (K2 / 8);
    sW2[row * RS2 + q] = *(const bf16x8*)(W2 + row * K2 + q * 8);
  }
  if (tid < 16) sbias[tid] = tid < 8 ? ((const float4*)b1)[tid] : ((const float4*)b2)[tid - 8];
  for (int e = tid; e < PP * PP; e += 256) sx[e] = 0u;  // the border stays zero
  // L1 taps of this lane: t = 8 st2 + 4 hq + q, (ky, kx) = (t / 3, t % 3), as
  // an offset into the padded patch; t >= 9 is the zero padding of K
  int toff0[4];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int t = 4 * hq + q;
    toff0[q] = (t / 3) * PP + t % 3;
  }
  const int toff8 = 2 * PP + 2;  // t = 8, used by the hq = 0 half only
  __syncthreads();
  // the pixels of the next node are loaded while this node's convolutions run
  float vf[NP / 256], vg[NP / 256];
  if (blockIdx.x < count) {
    int cx, cy, sxs, sys;
    node_shift(coords, n, prior, first + blockIdx.x, &cx, &cy, &sxs, &sys);
    patch_load(f, g, M, N, cx, cy, sxs, sys, vf, vg);
  }
  for (int local = blockIdx.x; local < count; local += gridDim.x) {
    // ---- patch pair, normalised (patch_kernel's arithmetic) ----
    const PatchMoments m = patch_moments(vf, vg, red);
#pragma unroll
    for (int k = 0; k < NP / 256; k++) {
      const uint32_t p = tid + 256 * k, py = p / PATCH, px = p - PATCH * py;
      sx[(py + 1) * PP + px + 1] = patch_pair(vf[k], vg[k], m);
    }
    __syncthreads();
    if (local + (int)gridDim.x < count) {
      int cx, cy, sxs, sys;
      node_shift(coords, n, prior, first + local + gridDim.x, &cx, &cy, &sxs, &sys);
      patch_load(f, g, M, N, cx, cy, sxs, sys, vf, vg);
    }
    // ---- L1: 576 rows (24 x 24), K 32 (taps 0..8 x 2 channels, padded), N 32 ----
    // Computed transposed (weights as the A operand, the patch as B), so a
    // lane ends with 16 channels of one output pixel, in runs of 4 that go
    // to LDS as 8-byte words; the products and their k order are those of
    // the untransposed layer.
    for (int tile = wave; tile < 18; tile += 4) {
      const int row = tile * 32 + r;
      const int oy = row / 24, ox = row % 24;
      // tap (ky, kx) of output (oy, ox) reads patch pixel (2 oy - 1 + ky, 2 ox - 1 + kx),
      // padded index (2 oy + ky) * PP + 2 ox + kx
      const int rb = 2 * oy * PP + 2 * ox;
      floatx16 acc;
#pragma unroll
      for (int i = 0; i < 16; i++) acc[i] = 0.0f;
      {
        // taps 0..7: every element is a (ref, def) pair already in bf16
        uint32_t w[4];
#pragma unroll
        for (int q = 0; q < 4; q++) w[q] = sx[rb + toff0[q]];
        const bf16x8 v = __builtin_bit_cast(bf16x8, w);
        const bf16x8 b = sW1[r * RS1 + hq];
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(b, v, acc, 0, 0, 0);
      }
      {
        // tap 8 (first pair of the hq = 0 half), zero padding elsewhere
        uint32_t w[4] = {0u, 0u, 0u, 0u};
        const uint32_t t8 = sx[rb + toff8];
        w[0] = hq == 0 ? t8 : 0u;
        const bf16x8 v = __builtin_bit_cast(bf16x8, w);
        const bf16x8 b = sW1[r * RS1 + 2 + hq];
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(b, v, acc, 0, 0, 0);
      }
      // element i: channel 8 (i / 4) + 4 hq + i % 4 of pixel `row`
#pragma unroll
      for (int q0 = 0; q0 < 4; q0++) {
        const int ch = 8 * q0 + 4 * hq;
        *(uint2*)(sa1 + row * FR_L1C + ch) = bias_relu_pack4(acc, q0, sbias[ch / 4], true);
      }
    }
    __syncthreads();
    // ---- L2: 144 rows (12 x 12), K 288 (tap-major, 32 channels), N 32 ----
    // transposed as L1: a lane ends with 16 channels of one output pixel
    for (int tile = wave; tile < 5; tile += 4) {
      const int row = tile * 32 + r;
      const bool ok = row < 144;
      const int oy = ok ? row / 12 : 0, ox = ok ? row % 12 : 0;
      floatx16 acc;
#pragma unroll
      for (int i = 0; i < 16; i++) acc[i] = 0.0f;
#pragma unroll
      for (int t = 0; t < 9; t++) {
        const int iy = oy * 2 - 1 + t / 3, ix = ox * 2 - 1 + t % 3;
        const bool in = ok && iy >= 0 && iy < 24 && ix >= 0 && ix < 24;
#pragma unroll
        for (int c = 0; c < 2; c++) {
          bf16x8 v = {0, 0, 0, 0, 0, 0, 0, 0};
          if (in) v = *(const bf16x8*)(sa1 + (iy * 24 + ix) * FR_L1C + c * 16 + kh);
          const bf16x8 b = sW2[r * RS2 + 2 * (2 * t + c) + hq];
          acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(b, v, acc, 0, 0, 0);
        }
      }
      if (ok) {
        bf16* o = out + ((int64_t)local * 144 + row) * 32;
#pragma unroll
        for (int q0 = 0; q0 < 4; q0++) {
          const int ch = 8 * q0 + 4 * hq;
          *(uint2*)(o + ch) = bias_relu_pack4(acc, q0, sbias[8 + ch / 4], true);
        }
      }
    }
    __syncthreads();
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

template <int AMODE>
static void launch_gemm(dim3 grid, int NB, bool fused, const bf16* a, int lda, const bf16* b,
                        int ldb, int M, int N, int K, int kchunk, float* C, int ldc,
                        const float* bias, int relu, bf16* cb, int ldcb, ConvGeom g,
                        hipStream_t s) {
#define LAUNCH(NBV, E)                                                                         \
  gemm_kernel<NBV, E, AMODE><<<grid, 256, 0, s>>>(a, lda, b, ldb, M, N, K, kchunk, C, ldc, bias, \
                                                  relu, cb, ldcb, g)
  if (fused) {
    if (NB == 1) LAUNCH(1, 1);
    else if (NB == 2) LAUNCH(2, 1);
    else LAUNCH(4, 1);
  } else {
    if (NB == 1) LAUNCH(1, 0);
    else if (NB == 2) LAUNCH(2, 0);
    else LAUNCH(4, 0);
  }
#undef LAUNCH
}

static hipError_t gemm_any(int amode, const void* A, int lda, const void* Bt, int ldb, int M,
                           int N, int K, int splits, float* C, int ldc, const float* bias,
                           int relu, void* Cb, int ldcb, ConvGeom g, hipStream_t s) {
  if (K % 16 || ldb % 8 || splits < 1 || (amode == 0 && lda % 8)) return hipErrorInvalidValue;
  if (amode == 1 && g.Cin % 8) return hipErrorInvalidValue;
  if (amode == 2 && g.Cin != 2) return hipErrorInvalidValue;
  const int ksteps = K / 16;
  const int per = (ksteps + splits - 1) / splits;
  const int kchunk = per * 16;
  splits = (ksteps + per - 1) / per;
  const int NB = N <= 32 ? 1 : (N <= 64 ? 2 : 4);
  dim3 grid((M + 127) / 128, (N + 32 * NB - 1) / (32 * NB), splits);
  const bool fused = splits == 1 && (Cb != nullptr || bias != nullptr || relu);
  const bf16* a = (const bf16*)A;
  const bf16* b = (const bf16*)Bt;
  if (amode == 0)
    launch_gemm<0>(grid, NB, fused, a, lda, b, ldb, M, N, K, kchunk, C, ldc, bias, relu,
                   (bf16*)Cb, ldcb, g, s);
  else if (amode == 1)
    launch_gemm<1>(grid, NB, fused, a, lda, b, ldb, M, N, K, kchunk, C, ldc, bias, relu,
                   (bf16*)Cb, ldcb, g, s);
  else
    launch_gemm<2>(grid, NB, fused, a, lda, b, ldb, M, N, K, kchunk, C, ldc, bias, relu,
                   (bf16*)Cb, ldcb, g, s);
  return hipGetLastError();
}

hipError_t cnn_gemm(const void* A, int lda, const void* Bt, int ldb, int M, int N, int K,
                    int splits, float* C, int ldc, const float* bias, int relu, void* Cb, int ldcb,
                    hipStream_t s) {
  return gemm_any(0, A, lda, Bt, ldb, M, N, K, splits, C, ldc, bias, relu, Cb, ldcb,
                  ConvGeom{0, 0, 0, 0, 0, 0}, s);
}

template <int NB, int AMODE>
static hipError_t launch_lds(const bf16* A, int lda, const bf16* Bt, int M, int K,
                             const float* bias, int relu, bf16* Cb, int ldc, ConvGeom g,
                             hipStream_t s) {
  const size_t smem = (size_t)32 * NB * (K / 8 + 1) * 16;
  if (smem > 160 * 1024) return hipErrorInvalidValue;
  // 8-byte output words: the row stride, the column block and the bias
  // block must keep 4-element alignment
  if (ldc % 4 != 0 || (reinterpret_cast<uintptr_t>(Cb) & 7) != 0 ||
      (bias && (reinterpret_cast<uintptr_t>(bias) & 15) != 0))
    return hipErrorInvalidValue;
  static bool attr = false;  // one attribute call per instantiation
  if (!attr) {
    hipError_t e = hipFuncSetAttribute((const void*)gemm_lds_kernel<NB, AMODE>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
    if (e != hipSuccess) return e;
    attr = true;
  }
  const int ntiles = (M + 255) / 256;
  const int per_cu = smem > 80 * 1024 ? 1 : (smem > 40 * 1024 ? 2 : 4);
  const int grid = std::max(1, std::min(ntiles, 256 * per_cu * 2));
  gemm_lds_kernel<NB, AMODE><<<grid, 256, smem, s>>>(A, lda, Bt, M, K, bias, relu, Cb, ldc, g,
                                                      ntiles);
  return hipGetLastError();
}

template <int NB, int CIN>
static hipError_t launch_conv(const bf16* x, const bf16* Wt, int M, const float* bias, int relu,
                              bf16* out, ConvGeom g, hipStream_t s) {
  constexpr int K = CIN == 2 ? 32 : 9 * CIN;
  constexpr size_t smem = (size_t)32 * NB * (K / 8 + 1) * 16;
  // 8-byte output words: the output and bias blocks keep 4-element alignment
  if ((reinterpret_cast<uintptr_t>(out) & 7) != 0 ||
      (bias && (reinterpret_cast<uintptr_t>(bias) & 15) != 0))
    return hipErrorInvalidValue;
  const int ntiles = (M + 255) / 256;
  const int per_cu = smem > 80 * 1024 ? 1 : (smem > 40 * 1024 ? 2 : 4);
  const int grid = std::max(1, std::min(ntiles, 256 * per_cu * 2));
  conv_lds_kernel<NB, CIN><<<grid, 256, 0, s>>>(x, Wt, M, bias, relu, out, g, ntiles);
  return hipGetLastError();
}

hipError_t cnn_conv(const void* x, int B, int H, int W, int Cin, int stride, int Ho, int Wo,
                    const void* Wt, int Kp, int Cout, const float* bias, int relu, void* y,
                    float* yf, hipStream_t s) {
  const int M = B * Ho * Wo;
  const ConvGeom g{H, W, Cin, stride, Ho, Wo};
  const bool tab = (Cin == 2 && Kp == 32) || (Cin == 32 && Kp == 288) || (Cin == 64 && Kp == 576);
  if (!yf && tab && (Cout == 32 || Cout == 64)) {
    const bf16* a = (const bf16*)x;
    const bf16* w = (const bf16*)Wt;
    bf16* o = (bf16*)y;
    if (Cin == 2)
      return Cout == 32 ? launch_conv<1, 2>(a, w, M, bias, relu, o, g, s)
                        : launch_conv<2, 2>(a, w, M, bias, relu, o, g, s);
    if (Cin == 32)
      return Cout == 32 ? launch_conv<1, 32>(a, w, M, bias, relu, o, g, s)
                        : launch_conv<2, 32>(a, w, M, bias, relu, o, g, s);
    return Cout == 32 ? launch_conv<1, 64>(a, w, M, bias, relu, o, g, s)
                      : launch_conv<2, 64>(a, w, M, bias, relu, o, g, s);
  }
  return gemm_any(Cin == 2 ? 2 : 1, x, 0, Wt, Kp, M, Cout, Kp, 1, yf, Cout, bias, relu, y, Cout,
                  g, s);
}

hipError_t cnn_dense(const void* x, int B, int K, const void* Wt, int Cout, const float* bias,
                     int relu, void* y, hipStream_t s) {
  // two 64-column halves: a workgroup then stages half the weights (75 KB
  // at K = 576), so it can share a CU with the IC-GN kernels of another
  // stream instead of waiting for one whose LDS is empty
  if (Cout == 128 && K % 16 == 0 && (size_t)64 * (K / 8 + 1) * 16 <= 80 * 1024) {
    for (int h = 0; h < 2; h++) {
      const hipError_t e = launch_lds<2, 0>((const bf16*)x, K, (const bf16*)Wt + (int64_t)64 * h * K,
                                            B, K, bias ? bias + 64 * h : nullptr, relu,
                                            (bf16*)y + 64 * h, 128, ConvGeom{0, 0, 0, 0, 0, 0}, s);
      if (e != hipSuccess) return e;
    }
    return hipSuccess;
  }
  return gemm_any(0, x, K, Wt, K, B, Cout, K, 1, nullptr, Cout, bias, relu, y, Cout,
                  ConvGeom{0, 0, 0, 0, 0, 0}, s);
}

hipError_t cnn_splitk_reduce(const float* P, int splits, int M, int N, int ldp, const float* bias,
                             int relu, float* C, int ldc, void* Cb, int ldcb, hipStream_t s) {
  splitk_reduce_kernel<<<grid_for((int64_t)M * N), 256, 0, s>>>(P, splits, M, N, ldp, bias, relu, C,
                                                                 ldc, (bf16*)Cb, ldcb);
  return hipGetLastError();
}

hipError_t cnn_im2col(const void* x, int B, int H, int W, int Cin, int stride, int Ho, int Wo,
                      int Kp, void* col, hipStream_t s) {
  im2col_kernel<<<grid_for((int64_t)B * Ho * Wo * Kp), 256, 0, s>>>(
      (const bf16*)x, B, H, W, Cin, stride, Ho, Wo, Kp, (bf16*)col);
  return hipGetLastError();
}

hipError_t cnn_col2im(const float* dcol, int B, int H, int W, int Cin, int stride, int Ho, int Wo,
                      int Kp, float* dx, hipStream_t s) {
  col2im_kernel<<<grid_for((int64_t)B * H * W * Cin), 256, 0, s>>>(dcol, B, H, W, Cin, stride, Ho,
                                                                    Wo, Kp, dx);
  return hipGetLastError();
}

hipError_t cnn_transpose_bf16(const void* in, int R, int Cc, int ldi, void* out, int ldo,
                              hipStream_t s) {
  dim3 grid((Cc + 31) / 32, (R + 31) / 32);
  transpose_kernel<bf16><<<grid, 256, 0, s>>>((const bf16*)in, R, Cc, ldi, (bf16*)out, ldo);
  return hipGetLastError();
}

hipError_t cnn_transpose_f32(const float* in, int R, int Cc, int ldi, void* out, int ldo,
                             hipStream_t s) {
  dim3 grid((Cc + 31) / 32, (R + 31) / 32);
  transpose_kernel<float><<<grid, 256, 0, s>>>(in, R, Cc, ldi, (bf16*)out, ldo);
  return hipGetLastError();
}

hipError_t cnn_relu_back(const float* g, const void* y, int64_t n, void* gb, float* gf,
                         hipStream_t s) {
  relu_back_kernel<<<grid_for(n), 256, 0, s>>>(g, (const bf16*)y, n, (bf16*)gb, gf);
  return hipGetLastError();
}

hipError_t cnn_colsum(const float* g, int M, int N, int ldg, float* out, hipStream_t s) {
  colsum_kernel<<<N, 256, 0, s>>>(g, M, ldg, out);
  return hipGetLastError();
}

hipError_t cnn_to_bf16(const float* in, int64_t n, void* out, hipStream_t s) {
  f32_to_bf16_kernel<<<grid_for(n), 256, 0, s>>>(in, n, (bf16*)out);
  return hipGetLastError();
}

hipError_t cnn_adam(float* w, const float* g, float* m, float* v, int64_t n, float lr, float b1,
                    float b2, float eps, float c1, float c2, float wd, const float* mask,
                    void* wb, hipStream_t s) {
  adam_kernel<<<grid_for(n), 256, 0, s>>>(w, g, m, v, n, lr, b1, b2, eps, c1, c2, wd, mask,
                                           (bf16*)wb);
  return hipGetLastError();
}

hipError_t cnn_mse(const float* pred, int ldp, const float* lab, const float* scale, int B,
                   float* d, int ldd, float* loss, hipStream_t s) {
  mse_kernel<<<1, 1024, 0, s>>>(pred, ldp, lab, scale, B, d, ldd, loss);
  return hipGetLastError();
}

hipError_t cnn_patches(const float* f, const float* g, int M, int N, const double* coords, int n,
                       int first, int count, const float* prior, void* out, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  if (first < 0 || first + count > n) return hipErrorInvalidValue;
  patch_kernel<<<std::min(count, 4096), 256, 0, s>>>(f, g, M, N, coords, n, first, count, prior,
                                                      (bf16*)out);
  return hipGetLastError();
}

hipError_t cnn_front(const float* f, const float* g, int M, int N, const double* coords, int n,
                     int first, int count, const float* prior, const void* W1, const float* b1,
                     const void* W2, const float* b2, void* out, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  if (first < 0 || first + count > n) return hipErrorInvalidValue;
  // biases read as float4, outputs written as 8-byte words
  if ((reinterpret_cast<uintptr_t>(b1) & 15) != 0 || (reinterpret_cast<uintptr_t>(b2) & 15) != 0 ||
      (reinterpret_cast<uintptr_t>(out) & 7) != 0)
    return hipErrorInvalidValue;
  front_kernel<<<std::min(count, 2048), 256, 0, s>>>(f, g, M, N, coords, n, first, count, prior,
                                                      (const bf16*)W1, b1, (const bf16*)W2, b2,
                                                      (bf16*)out);
  return hipGetLastError();
}

hipError_t cnn_predict(const float* out, int ldo, int first, int count, const float* prior,
                       float sf, float su, float* U0, float* F0, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  predict_kernel<<<(count + 255) / 256, 256, 0, s>>>(out, ldo, first, count, prior, sf, su, U0,
                                                     F0);
  return hipGetLastError();
}

hipError_t cnn_mosaic(int M, int N, int tile, const float* warps, int ntx, uint32_t seed,
                      uint32_t noise_seed, float density, float radius, float scale, float noise,
                      float top, float* img, hipStream_t s) {
  mosaic_kernel<<<grid_for((int64_t)M * N), 256, 0, s>>>(M, N, tile, warps, ntx, seed, noise_seed,
                                                          density, radius, scale, noise, top, img);
  return hipGetLastError();
}

}  // namespace aldic
