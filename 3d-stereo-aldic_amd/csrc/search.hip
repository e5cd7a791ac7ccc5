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
        const int qa = half * (KT2 / 2);
        float c1 = 0.0f, c2 = 0.0f;
        for (int b = 0; b < s; b++) {
          const float v = sub[(qa + b) * SW + i];
          c1 += v;
          c2 += v * v;
        }
        CB[qa * SW + i] = c1;
        CB2[qa * SW + i] = c2;
        for (int q = qa + 1; q < qa + KT2 / 2; q++) {
          const float vo = sub[(q - 1) * SW + i], vn = sub[(q - 1 + s) * SW + i];
          c1 += vn - vo;
          c2 += vn * vn - vo * vo;
          CB[q * SW + i] = c1;
          CB2[q * SW + i] = c2;
        }
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
        float sa = 0.0f, sa2 = 0.0f;
        const int qq = w * RB + r;
        for (int a2 = 0; a2 < s; a2++) {
          sa += CB[qq * SW + lane + a2];
          sa2 += CB2[qq * SW + lane + a2];
        }
        // normxcorr2 denominator; a window of (numerically) zero variance
        // gives 0, and so does a coefficient outside [-1, 1] beyond FP32
        // round-off
        const float varA = sa2 - sa * sa * inv_n;
        float gam = 0.0f;
        if (varA > 1e-6f * sa2 && varA > 0.0f) {
          gam = num[r] / (P.denT * sqrtf(varA));
          if (fabsf(gam) - 1.0f > 1e-5f) gam = 0.0f;
        }
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

hipError_t run(const SearchArgs& A, int maxpts, hipStream_t s, double* q = nullptr) {
  const size_t smem = aldic::search_smem_bytes(A.ws);
  if (smem > 160 * 1024) return hipErrorInvalidValue;
  // dynamic LDS beyond the 64 KB default is asked for, per launch size
  if (smem > 64 * 1024) {
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
                               float* F0, hipStream_t s) {
  SearchArgs A{};
  A.f = f;
  A.g = g;
  A.M = M;
  A.N = N;
  A.ws = ws;
  A.sp = pad_len(ws);
  for (int k = 0; k < 4; k++) A.reg[k] = reg[k];
  tile_counts(reg, &A.nt1, &A.ntiles);
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
                                double* ccmax, int* ipr, int* ipc, hipStream_t s) {
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
                         double* q) {
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
