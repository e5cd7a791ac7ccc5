// search.hip -- integer search of ALDIC (func/funIntegerSearch.m:100-203):
// per point, the ZNCC of the (ws+1)^2 template against the search window
// (normxcorr2, textbook zero-padded form), restricted to the lag window the
// reference crops (:166), then findpeak (func/findpeak.m).
//
// Three launches, each spread over the whole chip whatever the number of
// points:
//   prep   one workgroup per point: window clamps, the template with its
//          mean removed (stored transposed and zero padded so the tile
//          kernel reads it with scalar loads), its norm;
//   tiles  one workgroup per (point, lag tile): a tile is 64 dim-1 lags
//          (lanes) x 32 dim-2 lags (4 waves x 8 registers).  The window
//          slice is staged in LDS; each lane slides an 8-value register
//          window along dim 2 so one LDS load feeds 8 FMAs, and the template
//          value comes from a scalar register.  The normxcorr2 denominators
//          are exact box sums from LDS column sums (an FP32 summed-area
//          table loses 1e-3 relative on the sum of squares).  Each tile
//          keeps its first maximum in column-major order;
//   final  one wave per point: the first maximum over the tiles, then the
//          3x3 quadratic fit of findpeak.m (FP64, rounded to 1e-3), its nine
//          ZNCC values recomputed at the peak's neighbours.
// Points are the reference's grid (funIntegerSearch.m:106-107), a device
// list of some of its points, or a device list of nodes whose templates are
// centred on the nodes; list counts are read on the device.
#include <algorithm>

#include "aldic_common.h"
#include "aldic_kernels.h"

namespace {

constexpr int KT1 = 64;  // dim-1 lags per tile (lanes)
constexpr int KT2 = 32;  // dim-2 lags per tile (4 waves x 8)
constexpr int RB = 8;    // register block along dim 2

struct PointPrm {
  int ii, jj;      // template origin (1-based)
  int L, R, U, D;  // clamped search region
  int valid;       // 0: constant template (normxcorr2 throws) or template off the image
  float denT;      // || T - mean(T) ||
};

struct SearchArgs {
  const float* f;  // M x N reference (normalised)
  const float* g;  // M x N deformed
  int M, N;
  int ws, sp;      // template side s = ws + 1, padded row length sp
  int reg[4];      // L R U D
  int ntiles, nt1; // lag tiles per point (max over points), along dim 1
  int fft;         // lag tiles are FFT blocks (search_fft_kernel) instead of direct tiles
  // grid mode
  int nX, nY, gx0, gx1, gy0, gy1, step;
  // list mode (list != nullptr): nodes; grid-list mode (glist != nullptr):
  // grid point indices ci + cj*nX.  Either count is lcount.
  const int* list;
  const int* glist;
  const int* lcount;
  const double* ncoords;
  int nn;
  float* U0;
  float* F0;
  // scratch
  PointPrm* prm;
  float* tmpl;  // per point s x sp, transposed: tmpl[a*sp + b] = T(a, b) - mean
  float* tbv;   // per (point, tile) best value
  int* tbi;     // and its lag index
  // grid outputs
  double *x, *y, *u, *v, *ccmax;
  int *ipr, *ipc;
};

__device__ __forceinline__ int num_points(const SearchArgs& A) {
  return (A.list || A.glist) ? *A.lcount : A.nX * A.nY;
}

// grid index of point t (grid modes)
__device__ __forceinline__ int grid_index(const SearchArgs& A, int t) {
  return A.glist ? A.glist[t] : t;
}

// Template origin and per-point clamps of funIntegerSearch.m:139-152 (the
// right/down clamps use the ROI end, as the reference does).
__device__ bool point_params(const SearchArgs& A, int t, int* ii, int* jj, int* L, int* R, int* U,
                             int* D) {
  int gx1, gy1;
  if (A.list) {
    const int node = A.list[t];
    const int h = A.ws / 2;
    *ii = (int)mround(A.ncoords[node]) - h;
    *jj = (int)mround(A.ncoords[node + A.nn]) - h;
    gx1 = A.M;
    gy1 = A.N;
    if (*ii < 1 || *jj < 1 || *ii + A.ws > A.M || *jj + A.ws > A.N) return false;
  } else {
    const int gi = grid_index(A, t);
    const int ci = gi % A.nX, cj = gi / A.nX;
    *ii = A.gx0 + ci * A.step;
    *jj = A.gy0 + cj * A.step;
    gx1 = A.gx1;
    gy1 = A.gy1;
  }
  int l = A.reg[0], r = A.reg[1], u = A.reg[2], d = A.reg[3];
  if (*ii - l < 1) l = *ii - 1;
  if (*jj - u < 1) u = *jj - 1;
  if (*ii + A.ws + r > A.M) r = gx1 - *ii - A.ws - 1;
  if (*jj + A.ws + d > A.N) d = gy1 - *jj - A.ws - 1;
  *L = l < 0 ? 0 : l;
  *R = r < 0 ? 0 : r;
  *U = u < 0 ? 0 : u;
  *D = d < 0 ? 0 : d;
  return true;
}

// ---------------------------------------------------------------------------
// prep: one 256-thread workgroup per point
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void search_prep_kernel(SearchArgs A) {
  __shared__ float red[8];
  const int tid = threadIdx.x, lane = tid % 64, w = tid / 64;
  const int total = num_points(A);
  const int s = A.ws + 1, n = s * s;
  for (int t = blockIdx.x; t < total; t += gridDim.x) {
    PointPrm P{};
    const bool inside = point_params(A, t, &P.ii, &P.jj, &P.L, &P.R, &P.U, &P.D);
    float st = 0.0f, diff = 0.0f;
    if (inside) {
      const float t0 = A.f[(int64_t)(P.ii - 1) + (int64_t)(P.jj - 1) * A.M];
      for (int p = tid; p < n; p += 256) {
        const float v = A.f[(int64_t)(P.ii - 1 + p % s) + (int64_t)(P.jj - 1 + p / s) * A.M];
        st += v;
        diff += (v != t0) ? 1.0f : 0.0f;
      }
    }
    st = wave_sum(st);
    diff = wave_sum(diff);
    if (lane == 0) {
      red[w * 2] = st;
      red[w * 2 + 1] = diff;
    }
    __syncthreads();
    st = red[0] + red[2] + red[4] + red[6];
    diff = red[1] + red[3] + red[5] + red[7];
    __syncthreads();
    const float tmean = st / (float)n;
    float* T = A.tmpl + (int64_t)t * s * A.sp;
    float ss = 0.0f;
    // transposed and padded: T[a*sp + b], zeros for b >= s
    for (int e = tid; e < s * A.sp; e += 256) {
      const int a = e / A.sp, b = e - a * A.sp;
      float d = 0.0f;
      if (inside && b < s) {
        d = A.f[(int64_t)(P.ii - 1 + a) + (int64_t)(P.jj - 1 + b) * A.M] - tmean;
        ss += d * d;
      }
      T[e] = d;
    }
    ss = wave_sum(ss);
    if (lane == 0) red[w] = ss;
    __syncthreads();
    if (tid == 0) {
      P.denT = sqrtf(red[0] + red[1] + red[2] + red[3]);
      P.valid = (inside && diff != 0.0f) ? 1 : 0;
      A.prm[t] = P;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// tiles: one 256-thread workgroup per (point, lag tile)
// ---------------------------------------------------------------------------
// NT < RB template taps of the last, partial block of a template column:
// the FMAs of the full block with the zero taps past s left out (the same
// sums: a zero tap only adds +0).
template <int NT>
__device__ __forceinline__ void tail_fmas(const float* Ta, const float (&win)[2 * RB],
                                          float (&num)[RB]) {
#pragma unroll
  for (int k = 0; k < NT; k++) {
    const float tv = Ta[k];
#pragma unroll
    for (int r = 0; r < RB; r++) num[r] = fmaf(tv, win[k + r], num[r]);
  }
}

// Window sums of a lag tile, shared by the direct and the FFT tiles:
// CB(i, q) = sum_{b<s} sub(i, q+b) for q = qa .. qa + CB_SEG - 1 (sub(i, j)
// at sub[j * ld + i]), by a running sum restarted every CB_SEG lags; CB is
// stored with row length SW.
constexpr int CB_SEG = 16;

__device__ __forceinline__ void col_box_sums(const float* sub, int ld, int SW, int s, int qa,
                                             int i, float* CB, float* CB2) {
  float c1 = 0.0f, c2 = 0.0f;
  for (int b = 0; b < s; b++) {
    const float v = sub[(qa + b) * ld + i];
    c1 += v;
    c2 = fmaf(v, v, c2);
  }
  CB[qa * SW + i] = c1;
  CB2[qa * SW + i] = c2;
  for (int q = qa + 1; q < qa + CB_SEG; q++) {
    const float vo = sub[(q - 1) * ld + i], vn = sub[(q - 1 + s) * ld + i];
    c1 += vn - vo;
    c2 += fmaf(vn, vn, -(vo * vo));
    CB[q * SW + i] = c1;
    CB2[q * SW + i] = c2;
  }
}

__device__ __forceinline__ float den_of(float sa, float sa2, float inv_n, float denT) {
  const float varA = sa2 - sa * sa * inv_n;
  return (varA > 1e-6f * sa2 && varA > 0.0f) ? denT * sqrtf(varA) : 0.0f;
}

// normxcorr2's denominator at tile lag (p, q) from the column sums:
// denT * sqrt(var) of the s x s window, or 0 for a window of (numerically)
// zero variance (its coefficient is 0)
__device__ __forceinline__ float lag_den(const float* CB, const float* CB2, int SW, int s, int q,
                                         int p, float inv_n, float denT) {
  float sa = 0.0f, sa2 = 0.0f;
  for (int a2 = 0; a2 < s; a2++) {
    sa += CB[q * SW + p + a2];
    sa2 += CB2[q * SW + p + a2];
  }
  return den_of(sa, sa2, inv_n, denT);
}

// the coefficient; one outside [-1, 1] beyond FP32 round-off is 0
__device__ __forceinline__ float zncc_of(float num, float den) {
  float gam = 0.0f;
  if (den > 0.0f) {
    gam = num / den;
    if (fabsf(gam) - 1.0f > 1e-5f) gam = 0.0f;
  }
  return gam;
}

__global__ __launch_bounds__(256) void search_tile_kernel(SearchArgs A,
                                                          const float* __restrict__ tmpl,
                                                          const float* __restrict__ g) {
  extern __shared__ float smem[];
  __shared__ float bestv[256];
  __shared__ int besti[256];
  const int tid = threadIdx.x, lane = tid % 64, w = tid / 64;
  const int total = num_points(A) * A.ntiles;
  const int s = A.ws + 1;
  const int M = A.M;
  const int SW = KT1 + s - 1;       // slice rows (dim 1)
  const int SC = KT2 + s - 1 + 16;  // slice columns (dim 2), 16 spare zero columns
  float* sub = smem;                // SW*SC
  float* CB = sub + SW * SC;        // KT2*SW
  float* CB2 = CB + KT2 * SW;       // KT2*SW
  const float inv_n = 1.0f / (float)(s * s);
  for (int item = blockIdx.x; item < total; item += gridDim.x) {
    const int t = item / A.ntiles, tile = item - t * A.ntiles;
    const PointPrm P = A.prm[t];
    const int K1 = P.L + P.R + 3, K2 = P.U + P.D + 3;
    const int W1 = s + P.L + P.R, W2 = s + P.U + P.D;
    const int p0 = (tile % A.nt1) * KT1, q0 = (tile / A.nt1) * KT2;
    float my_best = -1e30f;
    int my_idx = 0x7fffffff;
    if (P.valid && p0 < K1 && q0 < K2) {
      const int g_r0 = P.ii - P.L - 1, g_c0 = P.jj - P.U - 1;  // 0-based window (0,0)
      // slice: window rows p0-1 .. p0-1+SW-1, columns q0-1 .. q0-1+SC-1
      for (int e = tid; e < SW * SC; e += 256) {
        const int i = e % SW, j = e / SW;
        const int wr = p0 - 1 + i, wc = q0 - 1 + j;
        float v = 0.0f;
        if (wr >= 0 && wr < W1 && wc >= 0 && wc < W2 && j < KT2 + s - 1)
          v = g[(int64_t)(g_r0 + wr) + (int64_t)(g_c0 + wc) * M];
        sub[e] = v;
      }
      __syncthreads();
      // column sums along dim 2: CB(i, q) = sum_{b<s} sub(i, q+b)
      for (int task = tid; task < SW * 2; task += 256) {
        const int i = task % SW, half = task / SW;
        col_box_sums(sub, SW, SW, s, half * CB_SEG, i, CB, CB2);
      }
      // numerator: lane -> dim-1 lag p0+lane, wave w -> dim-2 lags q0+8w..+7
      const float* T = tmpl + (int64_t)t * s * A.sp;
      float num[RB];
#pragma unroll
      for (int r = 0; r < RB; r++) num[r] = 0.0f;
      // a wave whose dim-2 lags all lie past the lag window (the last tile
      // column) has nothing to sum; it still meets the barrier below
      const bool busy = q0 + w * RB < K2;
      const int sfull = s - s % RB, stail = s % RB;
      for (int a = 0; a < s && busy; a++) {
        const float* colp = sub + (w * RB) * SW + lane + a;
        const float* Ta = T + a * A.sp;
        float win[2 * RB];
#pragma unroll
        for (int k = 0; k < RB; k++) win[k] = colp[k * SW];
        for (int b0 = 0; b0 < sfull; b0 += RB) {
#pragma unroll
          for (int k = 0; k < RB; k++) win[RB + k] = colp[(b0 + RB + k) * SW];
#pragma unroll
          for (int k = 0; k < RB; k++) {
            const float tv = Ta[b0 + k];  // uniform: scalar load
#pragma unroll
            for (int r = 0; r < RB; r++) num[r] = fmaf(tv, win[k + r], num[r]);
          }
#pragma unroll
          for (int k = 0; k < RB; k++) win[k] = win[RB + k];
        }
        if (stail) {
#pragma unroll
          for (int k = 0; k < RB; k++) win[RB + k] = colp[(sfull + RB + k) * SW];
          const float* Tt = Ta + sfull;
          static_assert(RB == 8, "the tail cases cover 1 .. RB - 1 taps");
          switch (stail) {
            case 1: tail_fmas<1>(Tt, win, num); break;
            case 2: tail_fmas<2>(Tt, win, num); break;
            case 3: tail_fmas<3>(Tt, win, num); break;
            case 4: tail_fmas<4>(Tt, win, num); break;
            case 5: tail_fmas<5>(Tt, win, num); break;
            case 6: tail_fmas<6>(Tt, win, num); break;
            default: tail_fmas<7>(Tt, win, num); break;
          }
        }
      }
      __syncthreads();  // CB ready
      const int p = p0 + lane;
#pragma unroll
      for (int r = 0; r < RB; r++) {
        const int q = q0 + w * RB + r;
        const float gam = zncc_of(num[r], lag_den(CB, CB2, SW, s, w * RB + r, lane, inv_n, P.denT));
        if (p < K1 && q < K2) {
          const int idx = p + q * K1;
          if (gam > my_best || (gam == my_best && idx < my_idx)) {
            my_best = gam;
            my_idx = idx;
          }
        }
      }
    }
    bestv[tid] = my_best;
    besti[tid] = my_idx;
    __syncthreads();
    if (tid < 64) {
      float bv = bestv[tid];
      int bi = besti[tid];
      for (int k = tid + 64; k < 256; k += 64) {
        if (bestv[k] > bv || (bestv[k] == bv && besti[k] < bi)) {
          bv = bestv[k];
          bi = besti[k];
        }
      }
      // wave reduction with the same rule
      for (int off = 32; off > 0; off >>= 1) {
        const float ov = __shfl_down(bv, off);
        const int oi = __shfl_down(bi, off);
        if (ov > bv || (ov == bv && oi < bi)) {
          bv = ov;
          bi = oi;
        }
      }
      if (tid == 0) {
        A.tbv[item] = bv;
        A.tbi[item] = bi;
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// FFT tiles: one 576-thread workgroup per (point, 64 x 64 lag block).
//
// The block's numerators are one circular cross-correlation of size
// 128 x 128: the window region R (rows p0-1 .. p0+126, columns q0-1 ..
// q0+126, zero outside the clamped window) against the mean-removed
// template T zero-padded to 128 x 128; c(p, q) = sum_ab T(a, b) R(p+a, q+b)
// = IFFT2(FFT2(R) . conj(FFT2(T))) has no wrap-around for p, q < 64 when
// s <= 65.  Everything stays in LDS:
//   1. R is staged (real, 128 x 128) and the block's window sums are taken
//      from it exactly as the direct tiles take them (same helpers, same
//      order), so both paths divide by the same denominators;
//   2. dim-1 transforms of R, two real columns packed into one complex
//      128-point line and separated afterwards (65 bins kept);
//   3. dim-1 transforms of T the same way (ceil(s/2) packed lines);
//   4. per dim-1 bin (65 lines): dim-2 transforms of both, the product
//      FFT(R) conj(FFT(T)), its inverse dim-2 transform;
//   5. inverse dim-1 transforms, two real outputs per complex line;
//   6. coefficients and the block's first maximum in column-major order.
// A 128-point line transform is done by 8 lanes of one wave: a 16-point DFT
// in registers per lane (input stride 8), twiddles, one exchange through
// the line's own LDS storage, two 8-point DFTs per lane (in-place in the
// line: the 8 lanes are one wave, whose LDS accesses run in program order).
// Forward transforms use W = exp(-2 pi i / 128), the inverse ones conj(FFT
// (conj x)); the 128^2 scale is taken out at the end (a power of two).
// ---------------------------------------------------------------------------
constexpr int FFT_P = 128;                      // transform size per axis
constexpr int FFT_B = 64;                       // lags per block and axis
constexpr int FFT_SMAX = FFT_P - FFT_B + 1;     // largest template side (65)
constexpr int FFT_BINS = FFT_P / 2 + 1;         // dim-1 bins kept (65)
// Line pitches: the 8 line groups of a wave sit 16 floats apart modulo 64
// LDS banks (a real column pair and a spectrum line both span 272 floats).
constexpr int FFT_CS = FFT_P + 8;               // real column pitch in X (floats)
constexpr int FFT_LP = FFT_P + 8;               // complex line pitch of the spectrum
constexpr int FFT_THREADS = 576;                // 72 groups of 8 lanes >= 65 lines
constexpr int FFT_CBP = FFT_P + 1;              // window-sum row pitch (>= 64 + s - 1)
// floats, a multiple of 4 so that Y starts 16-byte aligned
constexpr int FFT_XF = (FFT_CS * 2 * ((FFT_SMAX + 1) / 2) + 2 * FFT_BINS * FFT_SMAX + 3) / 4 * 4;
constexpr int FFT_YF = 2 * FFT_BINS * FFT_LP;   // floats
static_assert(FFT_XF >= FFT_CS * FFT_P, "X holds the real region");
static_assert(FFT_YF >= 2 * FFT_B * FFT_CBP, "Y holds the window sums");

__device__ __forceinline__ float2 c_add(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 c_sub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 c_mul(float2 a, float2 b) {
  return make_float2(fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ float2 c_mulconj(float2 a, float2 b) {  // a * conj(b)
  return make_float2(fmaf(a.x, b.x, a.y * b.y), fmaf(a.y, b.x, -a.x * b.y));
}
__device__ __forceinline__ float2 c_conj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 c_mi(float2 a) { return make_float2(a.y, -a.x); }  // -i a

// forward 4-point DFT in place
__device__ __forceinline__ void dft4(float2& a0, float2& a1, float2& a2, float2& a3) {
  const float2 t0 = c_add(a0, a2), t1 = c_sub(a0, a2), t2 = c_add(a1, a3), t3 = c_mi(c_sub(a1, a3));
  a0 = c_add(t0, t2);
  a2 = c_sub(t0, t2);
  a1 = c_add(t1, t3);
  a3 = c_sub(t1, t3);
}

// forward 8-point DFT in place, natural order
__device__ __forceinline__ void dft8(float2 (&v)[8]) {
  constexpr float R = 0.70710678118654752f;
  dft4(v[0], v[2], v[4], v[6]);  // E[k] at v[0], v[2], v[4], v[6]
  dft4(v[1], v[3], v[5], v[7]);  // O[k] at v[1], v[3], v[5], v[7]
  const float2 o1 = c_mul(v[3], make_float2(R, -R));
  const float2 o2 = c_mi(v[5]);
  const float2 o3 = c_mul(v[7], make_float2(-R, -R));
  const float2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6], o0 = v[1];
  v[0] = c_add(e0, o0);
  v[4] = c_sub(e0, o0);
  v[1] = c_add(e1, o1);
  v[5] = c_sub(e1, o1);
  v[2] = c_add(e2, o2);
  v[6] = c_sub(e2, o2);
  v[3] = c_add(e3, o3);
  v[7] = c_sub(e3, o3);
}

// forward 16-point DFT, x[n] in v[n]; X[k1 + 4 k2] returned in v[4 k1 + k2]
__device__ __forceinline__ void dft16(float2 (&v)[16]) {
  constexpr float C1 = 0.92387953251128676f, S1 = 0.38268343236508977f;
  constexpr float R = 0.70710678118654752f;
#pragma unroll
  for (int n2 = 0; n2 < 4; n2++) dft4(v[n2], v[4 + n2], v[8 + n2], v[12 + n2]);
  // v[n2 + 4 k1] *= W16^(n2 k1)
  v[1 + 4 * 1] = c_mul(v[1 + 4 * 1], make_float2(C1, -S1));
  v[1 + 4 * 2] = c_mul(v[1 + 4 * 2], make_float2(R, -R));
  v[1 + 4 * 3] = c_mul(v[1 + 4 * 3], make_float2(S1, -C1));
  v[2 + 4 * 1] = c_mul(v[2 + 4 * 1], make_float2(R, -R));
  v[2 + 4 * 2] = c_mi(v[2 + 4 * 2]);
  v[2 + 4 * 3] = c_mul(v[2 + 4 * 3], make_float2(-R, -R));
  v[3 + 4 * 1] = c_mul(v[3 + 4 * 1], make_float2(S1, -C1));
  v[3 + 4 * 2] = c_mul(v[3 + 4 * 2], make_float2(-R, -R));
  v[3 + 4 * 3] = c_mul(v[3 + 4 * 3], make_float2(-C1, S1));
#pragma unroll
  for (int k1 = 0; k1 < 4; k1++) dft4(v[4 * k1], v[4 * k1 + 1], v[4 * k1 + 2], v[4 * k1 + 3]);
}

// slot of y[k1][n] (k1 < 16, n < 8) in a line's exchange: rows of 8, the
// neighbouring slots of rows 4..7 and 12..15 swapped, so that the 8 lanes
// reading one column (k1 = g + 8j) share fewer banks
__device__ __forceinline__ int ex_slot(int k1, int n) { return 8 * k1 + (n ^ ((k1 >> 2) & 1)); }

// One 128-point forward transform by the 8 lanes g = 0..7 of a group.  In:
// v[n1] = x[8 n1 + g].  ex: the line's 128 complex LDS slots (overwritten).
// Out: o[8 j + k2] = X[(g + 8 j) + 16 k2], j = 0, 1, k2 = 0..7.
__device__ __forceinline__ void fft128(float2 (&v)[16], float2* ex, int g, const float2* tw,
                                       float2 (&o)[16]) {
  dft16(v);
  // y[k1] = X16[k1] W128^(g k1) -> ex[8 k1 + g]; X16[k1 + 4 k2] is in v[4 k1 + k2]
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 4; b++) {
      const int k1 = a + 4 * b;
      ex[ex_slot(k1, g)] = c_mul(v[4 * a + b], tw[g * k1]);
    }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
  for (int j = 0; j < 2; j++) {
    float2 z[8];
#pragma unroll
    for (int n2 = 0; n2 < 8; n2++) z[n2] = ex[ex_slot(g + 8 * j, n2)];
    dft8(z);
#pragma unroll
    for (int k2 = 0; k2 < 8; k2++) o[8 * j + k2] = z[k2];
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// output slot m of fft128 -> frequency index
__device__ __forceinline__ int fft_out_index(int g, int m) { return (g + 8 * (m / 8)) + 16 * (m % 8); }

// Dim-1 transforms of real columns held in X (column c at X[c * FFT_CS + i]):
// the pair (2l, 2l+1) is one complex line, transformed in its own storage;
// the spectra are separated and bins 0..64 written to out[k * ld + 2l (+1)].
// Columns >= ncols read as zero.
__device__ __forceinline__ void dim1_real_pair(float* X, int l, int g, int ncols, float2* out,
                                               int ld, const float2* tw) {
  float* ca = X + (2 * l) * FFT_CS;
  float* cb = ca + FFT_CS;
  const bool hasb = 2 * l + 1 < ncols;
  float2 v[16], o[16];
#pragma unroll
  for (int n1 = 0; n1 < 16; n1++) {
    const int i = 8 * n1 + g;
    v[n1] = make_float2(ca[i], hasb ? cb[i] : 0.0f);
  }
  float2* ex = reinterpret_cast<float2*>(ca);  // the pair's first 256 floats
  fft128(v, ex, g, tw, o);
#pragma unroll
  for (int m = 0; m < 16; m++) ex[fft_out_index(g, m)] = o[m];
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  // A[k] = (Z[k] + conj Z[-k]) / 2, B[k] = -i (Z[k] - conj Z[-k]) / 2
  for (int k = g; k < FFT_BINS; k += 8) {
    const float2 zk = ex[k], zn = c_conj(ex[(FFT_P - k) & (FFT_P - 1)]);
    const float2 A = make_float2(0.5f * (zk.x + zn.x), 0.5f * (zk.y + zn.y));
    const float2 B = c_mi(make_float2(0.5f * (zk.x - zn.x), 0.5f * (zk.y - zn.y)));
    out[k * ld + 2 * l] = A;
    if (hasb) out[k * ld + 2 * l + 1] = B;
  }
}

__global__ __launch_bounds__(FFT_THREADS) void search_fft_kernel(SearchArgs A,
                                                                 const float* __restrict__ tmpl,
                                                                 const float* __restrict__ g) {
  extern __shared__ float4 smem4[];  // 16-byte aligned
  __shared__ float bestv[FFT_THREADS];
  __shared__ int besti[FFT_THREADS];
  float* smem = reinterpret_cast<float*>(smem4);
  float* X = smem;                                          // FFT_XF floats
  float2* Y = reinterpret_cast<float2*>(smem + FFT_XF);    // FFT_BINS x FFT_LP
  float2* tw = Y + FFT_BINS * FFT_LP;                       // W128^m, m < 128
  float* den = reinterpret_cast<float*>(tw + FFT_P);        // per lag of the block
  const int tid = threadIdx.x, grp = tid / 8, gl = tid % 8;
  for (int m = tid; m < FFT_P; m += FFT_THREADS) {
    double sn, cs;
    sincospi(2.0 * m / FFT_P, &sn, &cs);
    tw[m] = make_float2((float)cs, (float)-sn);
  }
  const int total = num_points(A) * A.ntiles;
  const int s = A.ws + 1;
  const int M = A.M;
  const int SW = FFT_B + s - 1;  // region rows the window sums use
  const float inv_n = 1.0f / (float)(s * s);
  for (int item = blockIdx.x; item < total; item += gridDim.x) {
    const int t = item / A.ntiles, tile = item - t * A.ntiles;
    const PointPrm P = A.prm[t];
    const int K1 = P.L + P.R + 3, K2 = P.U + P.D + 3;
    const int W1 = s + P.L + P.R, W2 = s + P.U + P.D;
    const int p0 = (tile % A.nt1) * FFT_B, q0 = (tile / A.nt1) * FFT_B;
    float my_best = -1e30f;
    int my_idx = 0x7fffffff;
    __syncthreads();  // the previous item's reads of X, Y and best are done
    if (P.valid && p0 < K1 && q0 < K2) {
      const int g_r0 = P.ii - P.L - 1, g_c0 = P.jj - P.U - 1;  // 0-based window (0,0)
      // 1. region R(i, j) = window(p0 - 1 + i, q0 - 1 + j) at X[j * FFT_CS + i];
      //    the columns past the block's reach are zero (as in the direct tiles)
      for (int e = tid; e < FFT_P * FFT_P; e += FFT_THREADS) {
        const int i = e % FFT_P, j = e / FFT_P;
        const int wr = p0 - 1 + i, wc = q0 - 1 + j;
        float v = 0.0f;
        if (wr >= 0 && wr < W1 && wc >= 0 && wc < W2 && j < FFT_B + s - 1)
          v = g[(int64_t)(g_r0 + wr) + (int64_t)(g_c0 + wc) * M];
        X[j * FFT_CS + i] = v;
      }
      __syncthreads();
      // window sums, rows of FFT_CBP floats (odd: the sliding pass below
      // reads 8 rows at once)
      float* CB = reinterpret_cast<float*>(Y);
      float* CB2 = CB + FFT_B * FFT_CBP;
      for (int task = tid; task < SW * (FFT_B / CB_SEG); task += FFT_THREADS) {
        const int i = task % SW, seg = task / SW;
        col_box_sums(X, FFT_CS, FFT_CBP, s, seg * CB_SEG, i, CB, CB2);
      }
      __syncthreads();
      // denominators by window sums slid along dim 1: task (q, c) takes
      // lags p = 8c .. 8c + 7 of column q
      for (int task = tid; task < FFT_B * (FFT_B / 8); task += FFT_THREADS) {
        const int c = task % (FFT_B / 8), q = task / (FFT_B / 8), pa = 8 * c;
        const float* r1 = CB + q * FFT_CBP;
        const float* r2 = CB2 + q * FFT_CBP;
        float sa = 0.0f, sa2 = 0.0f;
        for (int a2 = 0; a2 < s; a2++) {
          sa += r1[pa + a2];
          sa2 += r2[pa + a2];
        }
        for (int k = 0;; k++) {
          den[q * FFT_B + pa + k] = den_of(sa, sa2, inv_n, P.denT);
          if (k == 7) break;
          sa += r1[pa + k + s] - r1[pa + k];
          sa2 += r2[pa + k + s] - r2[pa + k];
        }
      }
      __syncthreads();
      // 2. dim-1 transforms of R -> Y[k][j]
      if (grp < FFT_P / 2) dim1_real_pair(X, grp, gl, FFT_P, Y, FFT_LP, tw);
      __syncthreads();
      // 3. the template, T(a, b) at X[b * FFT_CS + a] (a < s, zero below), and
      //    its dim-1 transforms -> Tp[k * s + b] after the pair columns
      const int npair = (s + 1) / 2;
      float2* Tp = reinterpret_cast<float2*>(X + FFT_CS * 2 * npair);
      const float* T = tmpl + (int64_t)t * s * A.sp;
      for (int e = tid; e < FFT_P * s; e += FFT_THREADS) {
        const int a = e % FFT_P, b = e / FFT_P;
        X[b * FFT_CS + a] = a < s ? T[a * A.sp + b] : 0.0f;
      }
      __syncthreads();
      if (grp < npair) dim1_real_pair(X, grp, gl, s, Tp, s, tw);
      __syncthreads();
      // 4. per dim-1 bin k: dim-2 transforms, product, inverse dim-2 transform
      if (grp < FFT_BINS) {
        const int k = grp;
        float2* line = Y + k * FFT_LP;
        float2 vr[16], vt[16], oR[16], oT[16];
#pragma unroll
        for (int n1 = 0; n1 < 16; n1++) {
          const int j = 8 * n1 + gl;
          vr[n1] = line[j];
          vt[n1] = j < s ? Tp[k * s + j] : make_float2(0.0f, 0.0f);
        }
        fft128(vt, line, gl, tw, oT);
        fft128(vr, line, gl, tw, oR);
        // conj of the product into the line in natural order, for the
        // inverse transform conj(FFT(conj x))
#pragma unroll
        for (int m = 0; m < 16; m++) line[fft_out_index(gl, m)] = c_conj(c_mulconj(oR[m], oT[m]));
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
        for (int n1 = 0; n1 < 16; n1++) vr[n1] = line[8 * n1 + gl];
        fft128(vr, line, gl, tw, oR);
#pragma unroll
        for (int m = 0; m < 16; m++) line[fft_out_index(gl, m)] = c_conj(oR[m]);
      }
      __syncthreads();
      // 5. inverse dim-1 transforms: columns (2l, 2l+1) from bins 0..64 of
      //    their Hermitian spectra, real parts to X[2l * FFT_CS + i],
      //    imaginary parts to X[(2l+1) * FFT_CS + i]
      if (grp < FFT_P / 2) {
        const int l = grp;
        float2 v[16], o[16];
#pragma unroll
        for (int n1 = 0; n1 < 16; n1++) {
          const int k = 8 * n1 + gl;
          float2 a, b;
          if (k < FFT_BINS) {
            a = Y[k * FFT_LP + 2 * l];
            b = Y[k * FFT_LP + 2 * l + 1];
          } else {
            a = c_conj(Y[(FFT_P - k) * FFT_LP + 2 * l]);
            b = c_conj(Y[(FFT_P - k) * FFT_LP + 2 * l + 1]);
          }
          // z = a + i b, transformed as conj(z)
          v[n1] = make_float2(a.x - b.y, -(a.y + b.x));
        }
        float* ca = X + (2 * l) * FFT_CS;
        fft128(v, reinterpret_cast<float2*>(ca), gl, tw, o);
#pragma unroll
        for (int m = 0; m < 16; m++) {
          const int i = fft_out_index(gl, m);
          ca[i] = o[m].x;             // real part of conj(o)
          ca[FFT_CS + i] = -o[m].y;   // imaginary part of conj(o)
        }
      }
      __syncthreads();
      // 6. coefficients and the block's first maximum
      constexpr float SCALE = 1.0f / (float)(FFT_P * FFT_P);
      for (int l = tid; l < FFT_B * FFT_B; l += FFT_THREADS) {
        const int p = l % FFT_B, q = l / FFT_B;
        const float gam = zncc_of(X[q * FFT_CS + p] * SCALE, den[l]);
        if (p0 + p < K1 && q0 + q < K2) {
          const int idx = (p0 + p) + (q0 + q) * K1;
          if (gam > my_best || (gam == my_best && idx < my_idx)) {
            my_best = gam;
            my_idx = idx;
          }
        }
      }
    }
    bestv[tid] = my_best;
    besti[tid] = my_idx;
    __syncthreads();
    if (tid < 64) {
      float bv = bestv[tid];
      int bi = besti[tid];
      for (int k = tid + 64; k < FFT_THREADS; k += 64) {
        if (bestv[k] > bv || (bestv[k] == bv && besti[k] < bi)) {
          bv = bestv[k];
          bi = besti[k];
        }
      }
      for (int off = 32; off > 0; off >>= 1) {
        const float ov = __shfl_down(bv, off);
        const int oi = __shfl_down(bi, off);
        if (ov > bv || (ov == bv && oi < bi)) {
          bv = ov;
          bi = oi;
        }
      }
      if (tid == 0) {
        A.tbv[item] = bv;
        A.tbi[item] = bi;
      }
    }
  }
}

// ZNCC at one lag (p, q) of a point, the formula of the tile kernel:
// window(wr, wc) = g(g_r0 + wr, g_c0 + wc) inside W1 x W2, zero outside.
__device__ float zncc_at(const SearchArgs& A, const float* T, const PointPrm& P, int p, int q,
                         int lane) {
  const int s = A.ws + 1, n = s * s;
  const int W1 = s + P.L + P.R, W2 = s + P.U + P.D;
  const int g_r0 = P.ii - P.L - 1, g_c0 = P.jj - P.U - 1;
  float num = 0.0f, sa = 0.0f, sa2 = 0.0f;
  for (int e = lane; e < n; e += 64) {
    const int a = e % s, b = e / s;
    const int wr = p - 1 + a, wc = q - 1 + b;
    float v = 0.0f;
    if (wr >= 0 && wr < W1 && wc >= 0 && wc < W2)
      v = A.g[(int64_t)(g_r0 + wr) + (int64_t)(g_c0 + wc) * A.M];
    num += T[a * A.sp + b] * v;
    sa += v;
    sa2 += v * v;
  }
  num = wave_sum(num);
  sa = wave_sum(sa);
  sa2 = wave_sum(sa2);
  const float varA = sa2 - sa * sa / (float)n;
  float gam = 0.0f;
  if (varA > 1e-6f * sa2 && varA > 0.0f) {
    gam = num / (P.denT * sqrtf(varA));
    if (fabsf(gam) - 1.0f > 1e-5f) gam = 0.0f;
  }
  return gam;
}

// ---------------------------------------------------------------------------
// final: one wave per point
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void search_final_kernel(SearchArgs A) {
  const int lane = threadIdx.x % 64;
  const int total = num_points(A);
  // uniform trip count per wave: every lane of a wave has the same t
  for (int t = blockIdx.x * 4 + threadIdx.x / 64; t < total; t += gridDim.x * 4) {
    const PointPrm P = A.prm[t];
    const int K1 = P.L + P.R + 3, K2 = P.U + P.D + 3;
    const int node = A.list ? A.list[t] : -1;
    const int gi = A.list ? 0 : grid_index(A, t);
    const int o = A.list ? t : (gi / A.nX) + (gi % A.nX) * A.nY;  // nY x nX column-major
    if (!P.valid) {
      // normxcorr2 throws on a constant template -> NaN (funIntegerSearch.m:171-174);
      // a listed node keeps its starting point
      if (!A.list && lane == 0) {
        const double nan = __builtin_nan("");
        A.u[o] = nan;
        A.v[o] = nan;
        A.ccmax[o] = nan;
        A.ipr[o] = 0;
        A.ipc[o] = 0;
        A.x[o] = (double)(P.ii + P.ii + A.ws) / 2.0;
        A.y[o] = (double)(P.jj + P.jj + A.ws) / 2.0;
      }
      continue;
    }
    // first maximum in column-major order over the tiles
    float bv = -1e30f;
    int bi = 0x7fffffff;
    for (int k = 0; k < A.ntiles; k++) {
      const float v = A.tbv[(int64_t)t * A.ntiles + k];
      const int i = A.tbi[(int64_t)t * A.ntiles + k];
      if (v > bv || (v == bv && i < bi)) {
        bv = v;
        bi = i;
      }
    }
    const int yp = bi % K1 + 1, xp = bi / K1 + 1;  // 1-based row, column
    double xpeak = xp, ypeak = yp, maxf = bv;
    if (!(xp == 1 || xp == K2 || yp == 1 || yp == K1)) {
      const float* T = A.tmpl + (int64_t)t * (A.ws + 1) * A.sp;
      double uu[3][3];
      for (int dc = -1; dc <= 1; dc++)
        for (int dr = -1; dr <= 1; dr++)
          uu[dc + 1][dr + 1] = zncc_at(A, T, P, yp - 1 + dr, xp - 1 + dc, lane);
      // 3x3 least squares of u = A1 + A2 x + A3 y + A4 xy + A5 x^2 + A6 y^2
      // (x: column offset, y: row offset), findpeak.m:37-65
      double su = 0, sxu = 0, syu = 0, sxyu = 0, sx2u = 0, sy2u = 0;
      for (int dc = -1; dc <= 1; dc++)
        for (int dr = -1; dr <= 1; dr++) {
          const double v = uu[dc + 1][dr + 1];
          su += v;
          sxu += dc * v;
          syu += dr * v;
          sxyu += dc * dr * v;
          sx2u += dc * dc * v;
          sy2u += dr * dr * v;
        }
      double Ac[6];
      Ac[1] = sxu / 6.0;
      Ac[2] = syu / 6.0;
      Ac[3] = sxyu / 4.0;
      // [9 6 6; 6 6 4; 6 4 6] [A1 A5 A6]' = [su sx2u sy2u]', closed form
      Ac[0] = (5.0 * su - 3.0 * sx2u - 3.0 * sy2u) / 9.0;
      Ac[4] = (-3.0 * su + 4.5 * sx2u) / 9.0;
      Ac[5] = (-3.0 * su + 4.5 * sy2u) / 9.0;
      const double den = Ac[3] * Ac[3] - 4 * Ac[4] * Ac[5];
      double xo = (-Ac[2] * Ac[3] + 2 * Ac[5] * Ac[1]) / den;
      double yo = -1 / den * (Ac[3] * Ac[1] - 2 * Ac[4] * Ac[2]);
      if (!(fabs(xo) > 1 || fabs(yo) > 1)) {
        xo = round(1000 * xo) / 1000;
        yo = round(1000 * yo) / 1000;
        xpeak = xp + xo;
        ypeak = yp + yo;
        maxf = Ac[0] + xo * Ac[1] + yo * Ac[2] + xo * yo * Ac[3] + xo * xo * Ac[4] +
               yo * yo * Ac[5];
      }
    }
    if (lane == 0) {
      if (A.list) {
        A.U0[2 * node] = (float)(ypeak - 2 - (double)P.L);
        A.U0[2 * node + 1] = (float)(xpeak - 2 - (double)P.U);
        if (A.F0)
          for (int k = 0; k < 4; k++) A.F0[4 * node + k] = 0.0f;
      } else {
        A.u[o] = ypeak - 2 - (double)P.L;
        A.v[o] = xpeak - 2 - (double)P.U;
        A.ccmax[o] = maxf;
        A.ipr[o] = yp;
        A.ipc[o] = xp;
        A.x[o] = (double)(P.ii + P.ii + A.ws) / 2.0;
        A.y[o] = (double)(P.jj + P.jj + A.ws) / 2.0;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// q-factors (funIntegerSearch.m:162-163, compute_qFactor :211-264): one
// 256-thread workgroup per point over the FULL normxcorr2 map, (W1 + s - 1)
// x (W2 + s - 1) lags, template origins -(s-1) .. W-1 in the window.  Pass
// 1 finds the map's min and max; pass 2 recomputes each coefficient (the
// same arithmetic, so the same value) and accumulates sum (c - min)^2 and a
// 30-bin histogram with the bin edges of MATLAB's histcounts.  The window
// and the template are staged in LDS when they fit (WIN_LDS), else the
// window is read from the image.
// ---------------------------------------------------------------------------
constexpr int QBINS = 30;

__device__ __forceinline__ double wave_sum_d(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

__device__ __forceinline__ double eps_of(double x) {
  x = fabs(x);
  if (x == 0.0) return 4.9406564584124654e-324;
  int e;
  frexp(x, &e);
  return ldexp(1.0, e - 53);
}

// histcounts(x, 30) bin width and right edge for data in [0, xmax] (its
// bin picker with the number of bins given; the left edge is 0 here);
// returns false for nearly constant data (one bin holds everything)
__device__ bool q_bins(double xmax, double* bw, double* right) {
  const double xscale = xmax, xrange = xmax;
  if (!(xrange > fmax(sqrt(eps_of(xscale)), 2.2250738585072014e-308))) return false;
  const double ll = xrange / QBINS, ul = xrange / (QBINS - 1);
  const double q10 = pow(10.0, floor(log10(ul - ll)));
  *bw = q10 * ceil(ll / q10);
  *right = fmax(QBINS * *bw, xmax);
  return true;
}

template <bool WIN_LDS>
__device__ float q_coef(const SearchArgs& A, const float* Tl, const float* win, int s, int W1,
                        int W2, int g_r0, int g_c0, int o1, int o2, float denT) {
  const int a0 = max(0, -o1), a1 = min(s, W1 - o1);
  const int b0 = max(0, -o2), b1 = min(s, W2 - o2);
  float num = 0.0f, sa = 0.0f, sa2 = 0.0f;
  for (int b = b0; b < b1; b++) {
    for (int a = a0; a < a1; a++) {
      const int wr = o1 + a, wc = o2 + b;
      const float v = WIN_LDS ? win[wr + wc * W1]
                              : A.g[(int64_t)(g_r0 + wr) + (int64_t)(g_c0 + wc) * A.M];
      num = fmaf(Tl[a + b * s], v, num);
      sa += v;
      sa2 = fmaf(v, v, sa2);
    }
  }
  // the tile kernel's rule: a window of (numerically) zero variance, or a
  // coefficient outside [-1, 1] beyond FP32 round-off, gives 0
  const float varA = sa2 - sa * sa / (float)(s * s);
  float gam = 0.0f;
  if (varA > 1e-6f * sa2 && varA > 0.0f) {
    gam = num / (denT * sqrtf(varA));
    if (fabsf(gam) - 1.0f > 1e-5f) gam = 0.0f;
  }
  return gam;
}

template <bool WIN_LDS>
__global__ __launch_bounds__(256) void search_qfactor_kernel(SearchArgs A, double* q) {
  extern __shared__ float qsm[];
  __shared__ float rmin[4], rmax[4];
  __shared__ double rsum[4];
  __shared__ int hist[QBINS];
  const int tid = threadIdx.x, lane = tid % 64, w = tid / 64;
  const int total = num_points(A);
  const int s = A.ws + 1;
  for (int t = blockIdx.x; t < total; t += gridDim.x) {
    const PointPrm P = A.prm[t];
    if (!P.valid) {  // normxcorr2 throws: [Inf Inf] (funIntegerSearch.m:173)
      if (tid == 0) q[2 * t] = q[2 * t + 1] = __builtin_inf();
      continue;
    }
    const int W1 = s + P.L + P.R, W2 = s + P.U + P.D;
    const int F1 = W1 + s - 1, nF = F1 * (W2 + s - 1);
    const int g_r0 = P.ii - P.L - 1, g_c0 = P.jj - P.U - 1;
    float* Tl = qsm;        // s x s, Tl[a + b*s] = T(a, b) - mean
    float* win = Tl + s * s;  // W1 x W2 (WIN_LDS)
    const float* T = A.tmpl + (int64_t)t * s * A.sp;
    for (int e = tid; e < s * s; e += 256) Tl[e] = T[(e % s) * A.sp + e / s];
    if (WIN_LDS)
      for (int e = tid; e < W1 * W2; e += 256)
        win[e] = A.g[(int64_t)(g_r0 + e % W1) + (int64_t)(g_c0 + e / W1) * A.M];
    if (tid < QBINS) hist[tid] = 0;
    __syncthreads();
    float mn = 1e30f, mx = -1e30f;
    for (int e = tid; e < nF; e += 256) {
      const float c = q_coef<WIN_LDS>(A, Tl, win, s, W1, W2, g_r0, g_c0, e % F1 - (s - 1),
                                      e / F1 - (s - 1), P.denT);
      mn = fminf(mn, c);
      mx = fmaxf(mx, c);
    }
    for (int off = 32; off > 0; off >>= 1) {
      mn = fminf(mn, __shfl_xor(mn, off));
      mx = fmaxf(mx, __shfl_xor(mx, off));
    }
    if (lane == 0) {
      rmin[w] = mn;
      rmax[w] = mx;
    }
    __syncthreads();
    mn = fminf(fminf(rmin[0], rmin[1]), fminf(rmin[2], rmin[3]));
    mx = fmaxf(fmaxf(rmax[0], rmax[1]), fmaxf(rmax[2], rmax[3]));
    const double cmin = mn, peak = (double)mx - (double)mn;
    double bw = 0.0, right = 0.0;
    const bool spread = q_bins(peak, &bw, &right);
    double e2 = 0.0;
    for (int e = tid; e < nF; e += 256) {
      const float c = q_coef<WIN_LDS>(A, Tl, win, s, W1, W2, g_r0, g_c0, e % F1 - (s - 1),
                                      e / F1 - (s - 1), P.denT);
      const double v = (double)c - cmin;
      e2 += v * v;
      int b = 0;
      if (spread) {
        // bin j holds edges(j) <= v < edges(j+1), edges(j) = j bw, the last
        // bin also its right edge
        b = min(max((int)floor(v / bw), 0), QBINS - 1);
        if (b > 0 && v < b * bw) b--;
        if (b < QBINS - 1 && v >= (b + 1) * bw) b++;
      }
      atomicAdd(&hist[b], 1);
    }
    e2 = wave_sum_d(e2);
    if (lane == 0) rsum[w] = e2;
    __syncthreads();
    if (tid == 0) {
      const double E2 = rsum[0] + rsum[1] + rsum[2] + rsum[3];
      double ent = 0.0;
      for (int j = 0; j < QBINS; j++)
        if (hist[j] > 0) {
          const double p = (double)hist[j] / (double)nF;
          ent += p * log(1.0 / p);
        }
      q[2 * t] = peak * peak / (E2 / (double)nF);
      q[2 * t + 1] = 1.0 / ent;
    }
    __syncthreads();
  }
}

void tile_counts(const int reg[4], int* nt1, int* ntiles) {
  const int K1 = reg[0] + reg[1] + 3, K2 = reg[2] + reg[3] + 3;
  *nt1 = (K1 + KT1 - 1) / KT1;
  *ntiles = *nt1 * ((K2 + KT2 - 1) / KT2);
}

// FFT blocks per point (never more than the direct tiles: the scratch is
// sized for those)
void fft_tile_counts(const int reg[4], int* nt1, int* ntiles) {
  const int K1 = reg[0] + reg[1] + 3, K2 = reg[2] + reg[3] + 3;
  *nt1 = (K1 + FFT_B - 1) / FFT_B;
  *ntiles = *nt1 * ((K2 + FFT_B - 1) / FFT_B);
}

constexpr size_t FFT_SMEM = sizeof(float) * ((size_t)FFT_XF + FFT_YF + 2 * FFT_P + FFT_B * FFT_B);

// The FFT tiles take a lag window when asked for and the template fits.
// Method 2 (automatic) takes them where they are the faster ones: a whole
// grid of at least FFT_AUTO_POINTS points with lag windows of at least
// 32 x 32 (measured on MI355X: 2.0x faster than the direct tiles at
// 103 x 103 lags with 33 px templates, 7.6x with 65 px, 2.3x slower at
// 23 x 23).  A block holds most of a CU's LDS, so a small or listed search
// that runs beside other kernels (a fallback, a sparse prior) keeps the
// direct tiles, which share CUs.
constexpr int FFT_AUTO_POINTS = 1024;

bool use_fft(int method, int ws, const int reg[4], int points, bool listed) {
  if (ws + 1 > FFT_SMAX) return false;
  if (method == 1) return true;
  return method == 2 && !listed && points >= FFT_AUTO_POINTS && reg[0] + reg[1] + 3 >= 32 &&
         reg[2] + reg[3] + 3 >= 32;
}

int pad_len(int ws) { return (ws + 1 + RB - 1) / RB * RB; }

size_t up256(size_t v) { return (v + 255) / 256 * 256; }

void carve(SearchArgs& A, void* scratch, int maxpts) {
  char* p = static_cast<char*>(scratch);
  A.prm = reinterpret_cast<PointPrm*>(p);
  p += up256((size_t)maxpts * sizeof(PointPrm));
  A.tmpl = reinterpret_cast<float*>(p);
  p += up256((size_t)maxpts * (A.ws + 1) * A.sp * 4);
  A.tbv = reinterpret_cast<float*>(p);
  p += up256((size_t)maxpts * A.ntiles * 4);
  A.tbi = reinterpret_cast<int*>(p);
}

// the q-factor kernel's LDS: the template, plus the largest clamped window
// when that fits in QWIN_LDS_MAX bytes
constexpr size_t QWIN_LDS_MAX = 150 * 1024;

hipError_t run(SearchArgs A, int maxpts, hipStream_t s, double* q = nullptr) {
  const size_t smem = aldic::search_smem_bytes(A.ws);
  if (smem > 160 * 1024) return hipErrorInvalidValue;
  if (A.fft) {
    fft_tile_counts(A.reg, &A.nt1, &A.ntiles);
    const hipError_t e = hipFuncSetAttribute((const void*)search_fft_kernel,
                                             hipFuncAttributeMaxDynamicSharedMemorySize,
                                             (int)FFT_SMEM);
    if (e != hipSuccess) return e;
  } else if (smem > 64 * 1024) {
    // dynamic LDS beyond the 64 KB default is asked for, per launch size
    const hipError_t e = hipFuncSetAttribute((const void*)search_tile_kernel,
                                             hipFuncAttributeMaxDynamicSharedMemorySize,
                                             (int)smem);
    if (e != hipSuccess) return e;
  }
  if (maxpts <= 0) return hipSuccess;
  search_prep_kernel<<<std::min(maxpts, 2048), 256, 0, s>>>(A);
  if (q) {
    const int sd = A.ws + 1;
    const size_t tb = sizeof(float) * sd * sd;
    const size_t wb = sizeof(float) * (size_t)(sd + A.reg[0] + A.reg[1]) *
                      (size_t)(sd + A.reg[2] + A.reg[3]);
    const int grid = std::min(maxpts, 4096);
    const bool lds_win = tb + wb <= QWIN_LDS_MAX;
    const size_t qsm = lds_win ? tb + wb : tb;
    if (qsm > QWIN_LDS_MAX) return hipErrorInvalidValue;
    if (qsm > 64 * 1024) {
      const hipError_t e = hipFuncSetAttribute(
          lds_win ? (const void*)search_qfactor_kernel<true>
                  : (const void*)search_qfactor_kernel<false>,
          hipFuncAttributeMaxDynamicSharedMemorySize, (int)qsm);
      if (e != hipSuccess) return e;
    }
    if (lds_win)
      search_qfactor_kernel<true><<<grid, 256, qsm, s>>>(A, q);
    else
      search_qfactor_kernel<false><<<grid, 256, qsm, s>>>(A, q);
  }
  const int64_t items = (int64_t)maxpts * A.ntiles;
  if (A.fft)
    search_fft_kernel<<<(int)std::min<int64_t>(items, 2048), FFT_THREADS, FFT_SMEM, s>>>(
        A, A.tmpl, A.g);
  else
    search_tile_kernel<<<(int)std::min<int64_t>(items, 4096), 256, smem, s>>>(A, A.tmpl, A.g);
  search_final_kernel<<<std::min((maxpts + 3) / 4, 1024), 256, 0, s>>>(A);
  return hipGetLastError();
}

}  // namespace

namespace aldic {

size_t search_smem_bytes(int ws) {
  const int s = ws + 1;
  const int SW = KT1 + s - 1, SC = KT2 + s - 1 + 16;
  return sizeof(float) * ((size_t)SW * SC + 2 * (size_t)KT2 * SW);
}

size_t search_scratch_bytes(int maxpts, int ws, const int reg[4]) {
  int nt1, nt;
  tile_counts(reg, &nt1, &nt);
  return up256((size_t)maxpts * sizeof(PointPrm)) +
         up256((size_t)maxpts * (ws + 1) * pad_len(ws) * 4) +
         2 * up256((size_t)maxpts * nt * 4);
}

hipError_t launch_search_nodes(const float* f, const float* g, int M, int N,
                               const double* coords, int n, const int* list, const int* count,
                               int max_count, int ws, const int reg[4], void* scratch, float* U0,
                               float* F0, hipStream_t s, int method) {
  SearchArgs A{};
  A.f = f;
  A.g = g;
  A.M = M;
  A.N = N;
  A.ws = ws;
  A.sp = pad_len(ws);
  for (int k = 0; k < 4; k++) A.reg[k] = reg[k];
  tile_counts(reg, &A.nt1, &A.ntiles);
  A.fft = use_fft(method, ws, reg, max_count, true) ? 1 : 0;
  A.list = list;
  A.lcount = count;
  A.ncoords = coords;
  A.nn = n;
  A.U0 = U0;
  A.F0 = F0;
  carve(A, scratch, max_count);
  return run(A, max_count, s);
}

hipError_t launch_search_points(const float* f, const float* g, int M, int N, int nX, int nY,
                                const int gridx[2], const int gridy[2], int ws, int step,
                                const int reg[4], const int* points, const int* count,
                                void* scratch, double* x, double* y, double* u, double* v,
                                double* ccmax, int* ipr, int* ipc, hipStream_t s, int method) {
  SearchArgs A{};
  A.f = f;
  A.g = g;
  A.M = M;
  A.N = N;
  A.nX = nX;
  A.nY = nY;
  A.gx0 = gridx[0];
  A.gx1 = gridx[1];
  A.gy0 = gridy[0];
  A.gy1 = gridy[1];
  A.ws = ws;
  A.sp = pad_len(ws);
  A.step = step;
  for (int k = 0; k < 4; k++) A.reg[k] = reg[k];
  tile_counts(reg, &A.nt1, &A.ntiles);
  A.fft = use_fft(method, ws, reg, nX * nY, true) ? 1 : 0;
  A.glist = points;
  A.lcount = count;
  A.x = x;
  A.y = y;
  A.u = u;
  A.v = v;
  A.ccmax = ccmax;
  A.ipr = ipr;
  A.ipc = ipc;
  carve(A, scratch, nX * nY);
  return run(A, nX * nY, s);
}

hipError_t launch_search(const float* f, const float* g, int M, int N, int nX, int nY,
                         const int gridx[2], const int gridy[2], int ws, int step,
                         const int reg[4], void* scratch, double* x, double* y, double* u,
                         double* v, double* ccmax, int* ipr, int* ipc, hipStream_t s,
                         double* q, int method) {
  SearchArgs A{};
  A.f = f;
  A.g = g;
  A.M = M;
  A.N = N;
  A.nX = nX;
  A.nY = nY;
  A.gx0 = gridx[0];
  A.gx1 = gridx[1];
  A.gy0 = gridy[0];
  A.gy1 = gridy[1];
  A.ws = ws;
  A.sp = pad_len(ws);
  A.step = step;
  for (int k = 0; k < 4; k++) A.reg[k] = reg[k];
  tile_counts(reg, &A.nt1, &A.ntiles);
  A.fft = use_fft(method, ws, reg, nX * nY, false) ? 1 : 0;
  A.x = x;
  A.y = y;
  A.u = u;
  A.v = v;
  A.ccmax = ccmax;
  A.ipr = ipr;
  A.ipc = ipc;
  carve(A, scratch, nX * nY);
  return run(A, nX * nY, s, q);
}

}  // namespace aldic
