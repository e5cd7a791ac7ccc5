// search.hip -- whole-field integer search of ALDIC (func/funIntegerSearch.m:
// 100-203): per grid point, the ZNCC of the (ws+1)^2 template against the
// search window (normxcorr2, textbook zero-padded form), restricted to the
// lag window the reference crops (:166), then findpeak (func/findpeak.m).
//
// One 256-thread workgroup per grid point.  The lag window is processed in
// tiles of 64 (dim-1 lags, one per lane) x 32 (dim-2 lags, 8 per lane over
// 4 waves).  Per tile the zero-padded window slice is staged in LDS; each
// lane slides an 8-value register window along dim 2 so one LDS load feeds
// 8 FMAs, and the template value is an LDS broadcast.  The local sums of
// the window and its square (the normxcorr2 denominator) are exact box sums
// computed in LDS from column sums, not a summed-area table (FP32 SATs lose
// 1e-3 relative on the sum of squares).  ZNCC values go to a per-point map
// in global memory; the first maximum in column-major order and the 3x3
// quadratic fit (FP64, rounded to 1e-3 as findpeak.m:58-59) finish on the
// device.
#include "aldic_common.h"
#include "aldic_kernels.h"

namespace {

constexpr int KT1 = 64;  // dim-1 lags per tile (lanes)
constexpr int KT2 = 32;  // dim-2 lags per tile (4 waves x 8)
constexpr int RB = 8;    // register block along dim 2

struct SearchArgs {
  const float* f;  // M x N reference (normalised)
  const float* g;  // M x N deformed
  int M, N;
  int nX, nY;
  int gx0, gx1, gy0, gy1;  // ROI (1-based)
  int ws, step;
  int reg[4];  // L R U D
  float* gmap;  // per-point lag maps, stride kmax
  int64_t kmax;
  double *x, *y, *u, *v, *ccmax;
  int *ipr, *ipc;
  // node-list mode (list != nullptr): templates centred on the listed
  // nodes, count read on the device, results written as IC-GN starting
  // points U0 (and F0 = 0) of those nodes
  const int* list;
  const int* lcount;
  const double* ncoords;
  int nn;
  float* U0;
  float* F0;
};

__device__ void search_point_params(const SearchArgs& A, int t, int* ii, int* jj, int* L, int* R,
                                    int* U, int* D) {
  const int ci = t % A.nX, cj = t / A.nX;
  *ii = A.gx0 + ci * A.step;
  *jj = A.gy0 + cj * A.step;
  int l = A.reg[0], r = A.reg[1], u = A.reg[2], d = A.reg[3];
  // per-point clamps, funIntegerSearch.m:139-152 (right/down use the ROI end)
  if (*ii - l < 1) l = *ii - 1;
  if (*jj - u < 1) u = *jj - 1;
  if (*ii + A.ws + r > A.M) r = A.gx1 - *ii - A.ws - 1;
  if (*jj + A.ws + d > A.N) d = A.gy1 - *jj - A.ws - 1;
  *L = l < 0 ? 0 : l;
  *R = r < 0 ? 0 : r;
  *U = u < 0 ? 0 : u;
  *D = d < 0 ? 0 : d;
}

// Template origin and clamped region of a listed node: template centred on
// the rounded node, clamps as funIntegerSearch.m:139-152 with the whole
// image as the ROI.  False when the template leaves the image.
__device__ bool search_node_params(const SearchArgs& A, int node, int* ii, int* jj, int* L, int* R,
                                   int* U, int* D) {
  const int h = A.ws / 2;
  *ii = (int)mround(A.ncoords[node]) - h;
  *jj = (int)mround(A.ncoords[node + A.nn]) - h;
  if (*ii < 1 || *jj < 1 || *ii + A.ws > A.M || *jj + A.ws > A.N) return false;
  int l = A.reg[0], r = A.reg[1], u = A.reg[2], d = A.reg[3];
  if (*ii - l < 1) l = *ii - 1;
  if (*jj - u < 1) u = *jj - 1;
  if (*ii + A.ws + r > A.M) r = A.M - *ii - A.ws - 1;
  if (*jj + A.ws + d > A.N) d = A.N - *jj - A.ws - 1;
  *L = l < 0 ? 0 : l;
  *R = r < 0 ? 0 : r;
  *U = u < 0 ? 0 : u;
  *D = d < 0 ? 0 : d;
  return true;
}

__device__ void search_one(const SearchArgs& A, int t, int slot, float* smem, float* red,
                           float* bestv, int* besti) {
  const int tid = threadIdx.x, lane = tid % 64, w = tid / 64;
  int ii, jj, L, R, U, D;
  int node = -1;
  if (A.list) {
    node = A.list[t];
    if (!search_node_params(A, node, &ii, &jj, &L, &R, &U, &D)) return;  // block-uniform
  } else {
    search_point_params(A, t, &ii, &jj, &L, &R, &U, &D);
  }
  const int s = A.ws + 1;
  const int n = s * s;
  const int W1 = s + L + R, W2 = s + U + D;
  const int K1 = W1 - s + 3, K2 = W2 - s + 3;
  const int M = A.M;
  // u(cj, ci), nY x nX column-major (grid mode only)
  const int o = A.list ? t : (t / A.nX) + (t % A.nX) * A.nY;

  // LDS: template (mean removed), window slice, column sums
  const int SW = KT1 + s - 1;        // slice rows (dim 1)
  const int SC = KT2 + s - 1 + 16;   // slice columns (dim 2), 16 spare zero columns
  float* Tp = smem;                  // s*s
  float* sub = Tp + n;               // SW*SC
  float* CB = sub + SW * SC;         // KT2*SW
  float* CB2 = CB + KT2 * SW;        // KT2*SW

  // ---- template C = f(ii:ii+ws, jj:jj+ws) ----
  float st = 0.0f, diff = 0.0f;
  const float t0 = A.f[(int64_t)(ii - 1) + (int64_t)(jj - 1) * M];
  for (int p = tid; p < n; p += 256) {
    const float v = A.f[(int64_t)(ii - 1 + p % s) + (int64_t)(jj - 1 + p / s) * M];
    Tp[p] = v;
    st += v;
    diff += (v != t0) ? 1.0f : 0.0f;
  }
  float sums[2] = {wave_sum(st), wave_sum(diff)};
  if (lane == 0) {
    red[w * 2] = sums[0];
    red[w * 2 + 1] = sums[1];
  }
  __syncthreads();
  st = red[0] + red[2] + red[4] + red[6];
  diff = red[1] + red[3] + red[5] + red[7];
  __syncthreads();
  if (diff == 0.0f) {
    // normxcorr2 throws on a constant template -> NaN (funIntegerSearch.m:171-174);
    // a listed node keeps its starting point
    if (tid == 0 && !A.list) {
      const double nan = __builtin_nan("");
      A.u[o] = nan;
      A.v[o] = nan;
      A.ccmax[o] = nan;
      A.ipr[o] = 0;
      A.ipc[o] = 0;
      A.x[o] = (double)(ii + ii + A.ws) / 2.0;
      A.y[o] = (double)(jj + jj + A.ws) / 2.0;
    }
    return;
  }
  const float tmean = st / (float)n;
  float ss = 0.0f;
  for (int p = tid; p < n; p += 256) {
    const float d = Tp[p] - tmean;
    Tp[p] = d;
    ss += d * d;
  }
  ss = wave_sum(ss);
  if (lane == 0) red[w] = ss;
  __syncthreads();
  const float denT = sqrtf(red[0] + red[1] + red[2] + red[3]);
  const float inv_n = 1.0f / (float)n;
  float* gmap = A.gmap + (int64_t)slot * A.kmax;

  float my_best = -1e30f;
  int my_idx = 0x7fffffff;
  const int g_r0 = ii - L - 1, g_c0 = jj - U - 1;  // 0-based image index of window (0,0)

  for (int q0 = 0; q0 < K2; q0 += KT2) {
    for (int p0 = 0; p0 < K1; p0 += KT1) {
      __syncthreads();
      // slice rows: window rows p0-1 .. p0-1+SW-1; columns q0-1 .. q0-1+SC-1
      for (int e = tid; e < SW * SC; e += 256) {
        const int i = e % SW, j = e / SW;
        const int wr = p0 - 1 + i, wc = q0 - 1 + j;
        float v = 0.0f;
        if (wr >= 0 && wr < W1 && wc >= 0 && wc < W2 && j < KT2 + s - 1)
          v = A.g[(int64_t)(g_r0 + wr) + (int64_t)(g_c0 + wc) * M];
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
      float num[RB];
#pragma unroll
      for (int r = 0; r < RB; r++) num[r] = 0.0f;
      for (int a = 0; a < s; a++) {
        const float* colp = sub + (w * RB) * SW + lane + a;
        float win[2 * RB];
#pragma unroll
        for (int k = 0; k < RB; k++) win[k] = colp[k * SW];
        for (int b0 = 0; b0 < s; b0 += RB) {
#pragma unroll
          for (int k = 0; k < RB; k++) win[RB + k] = colp[(b0 + RB + k) * SW];
#pragma unroll
          for (int k = 0; k < RB; k++) {
            if (b0 + k < s) {
              const float tv = Tp[a + (b0 + k) * s];
#pragma unroll
              for (int r = 0; r < RB; r++) num[r] += tv * win[k + r];
            }
          }
#pragma unroll
          for (int k = 0; k < RB; k++) win[k] = win[RB + k];
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
          gam = num[r] / (denT * sqrtf(varA));
          if (fabsf(gam) - 1.0f > 1e-5f) gam = 0.0f;
        }
        if (p < K1 && q < K2) {
          const int idx = p + q * K1;
          gmap[idx] = gam;
          if (gam > my_best || (gam == my_best && idx < my_idx)) {
            my_best = gam;
            my_idx = idx;
          }
        }
      }
    }
  }
  bestv[tid] = my_best;
  besti[tid] = my_idx;
  __syncthreads();
  if (tid == 0) {
    float bv = bestv[0];
    int bi = besti[0];
    for (int k = 1; k < 256; k++) {
      if (bestv[k] > bv || (bestv[k] == bv && besti[k] < bi)) {
        bv = bestv[k];
        bi = besti[k];
      }
    }
    const int yp = bi % K1 + 1, xp = bi / K1 + 1;  // 1-based row, column
    double xpeak = xp, ypeak = yp, maxf = bv;
    if (!(xp == 1 || xp == K2 || yp == 1 || yp == K1)) {
      // 3x3 least squares of u = A1 + A2 x + A3 y + A4 xy + A5 x^2 + A6 y^2
      // (x: column offset, y: row offset), findpeak.m:37-65
      double su = 0, sxu = 0, syu = 0, sxyu = 0, sx2u = 0, sy2u = 0;
      for (int dc = -1; dc <= 1; dc++)
        for (int dr = -1; dr <= 1; dr++) {
          const double uu = gmap[(yp - 1 + dr) + (xp - 1 + dc) * K1];
          su += uu;
          sxu += dc * uu;
          syu += dr * uu;
          sxyu += dc * dr * uu;
          sx2u += dc * dc * uu;
          sy2u += dr * dr * uu;
        }
      double Ac[6];
      Ac[1] = sxu / 6.0;
      Ac[2] = syu / 6.0;
      Ac[3] = sxyu / 4.0;
      // [9 6 6; 6 6 4; 6 4 6] [A1 A5 A6]' = [su sx2u sy2u]'
      // closed form of that normal-equation block
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
        maxf = Ac[0] + xo * Ac[1] + yo * Ac[2] + xo * yo * Ac[3] + xo * xo * Ac[4] + yo * yo * Ac[5];
      }
    }
    if (A.list) {
      A.U0[2 * node] = (float)(ypeak - 2 - (double)L);
      A.U0[2 * node + 1] = (float)(xpeak - 2 - (double)U);
      if (A.F0)
        for (int k = 0; k < 4; k++) A.F0[4 * node + k] = 0.0f;
    } else {
      A.u[o] = ypeak - 2 - (double)L;
      A.v[o] = xpeak - 2 - (double)U;
      A.ccmax[o] = maxf;
      A.ipr[o] = yp;
      A.ipc[o] = xp;
      A.x[o] = (double)(ii + ii + A.ws) / 2.0;
      A.y[o] = (double)(jj + jj + A.ws) / 2.0;
    }
  }
}

__global__ __launch_bounds__(256) void search_kernel(SearchArgs A) {
  extern __shared__ float smem[];
  __shared__ float red[8 * 4];
  __shared__ float bestv[256];
  __shared__ int besti[256];
  if (A.list) {
    const int total = *A.lcount;
    for (int t = blockIdx.x; t < total; t += gridDim.x) {
      search_one(A, t, blockIdx.x, smem, red, bestv, besti);
      __syncthreads();
    }
  } else {
    search_one(A, blockIdx.x, blockIdx.x, smem, red, bestv, besti);
  }
}

}  // namespace

namespace aldic {

size_t search_smem_bytes(int ws) {
  const int s = ws + 1;
  const int SW = KT1 + s - 1, SC = KT2 + s - 1 + 16;
  return sizeof(float) * ((size_t)s * s + (size_t)SW * SC + 2 * (size_t)KT2 * SW);
}

static hipError_t search_attr(int ws, size_t* smem) {
  *smem = search_smem_bytes(ws);
  if (*smem > 160 * 1024) return hipErrorInvalidValue;
  if (*smem > 64 * 1024)
    return hipFuncSetAttribute((const void*)search_kernel,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)*smem);
  return hipSuccess;
}

hipError_t launch_search_nodes(const float* f, const float* g, int M, int N,
                               const double* coords, int n, const int* list, const int* count,
                               int grid, int ws, const int reg[4], float* gmap, int64_t kmax,
                               float* U0, float* F0, hipStream_t s) {
  SearchArgs A{};
  A.f = f;
  A.g = g;
  A.M = M;
  A.N = N;
  A.ws = ws;
  for (int k = 0; k < 4; k++) A.reg[k] = reg[k];
  A.gmap = gmap;
  A.kmax = kmax;
  A.list = list;
  A.lcount = count;
  A.ncoords = coords;
  A.nn = n;
  A.U0 = U0;
  A.F0 = F0;
  size_t smem = 0;
  hipError_t e = search_attr(ws, &smem);
  if (e != hipSuccess) return e;
  if (grid <= 0) return hipSuccess;
  search_kernel<<<grid, 256, smem, s>>>(A);
  return hipGetLastError();
}

hipError_t launch_search(const float* f, const float* g, int M, int N, int nX, int nY,
                         const int gridx[2], const int gridy[2], int ws, int step,
                         const int reg[4], float* gmap, int64_t kmax, double* x, double* y,
                         double* u, double* v, double* ccmax, int* ipr, int* ipc,
                         hipStream_t s) {
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
  A.step = step;
  for (int k = 0; k < 4; k++) A.reg[k] = reg[k];
  A.gmap = gmap;
  A.kmax = kmax;
  A.x = x;
  A.y = y;
  A.u = u;
  A.v = v;
  A.ccmax = ccmax;
  A.ipr = ipr;
  A.ipc = ipc;
  size_t smem = 0;
  hipError_t e = search_attr(ws, &smem);
  if (e != hipSuccess) return e;
  if (nX * nY == 0) return hipSuccess;
  search_kernel<<<nX * nY, 256, smem, s>>>(A);
  return hipGetLastError();
}

}  // namespace aldic
