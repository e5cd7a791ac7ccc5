// prep.hip -- image planes consumed by the solvers.
//   * FP64 MATLAB arrays -> FP32 planes f.*mask, g, mask
//   * Df (funImgGradient.m:49-78): 7-tap 6th-order central differences
//     [-1/60 3/20 -3/4 0 3/4 -3/20 1/60] along dim 1 (DfDx) and dim 2 (DfDy)
//     times the mask, evaluated in FP64 and stored either on the full grid
//     (FP32, zero on the 3-pixel border, for the solvers) or as the
//     reference's cropped (M-6) x (N-6) FP64 arrays.
// Threads run along dim 1 (image x), the contiguous axis, so every load of
// the stencil is coalesced.
#include <algorithm>

#include "aldic_common.h"
#include "aldic_kernels.h"

namespace {

__global__ __launch_bounds__(256) void planes_kernel(const double* __restrict__ ref,
                                                     const double* __restrict__ def,
                                                     const double* __restrict__ mask, int64_t MN,
                                                     float* __restrict__ fm, float* __restrict__ g,
                                                     float* __restrict__ maskf) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < MN;
       k += (int64_t)gridDim.x * blockDim.x) {
    const double m = mask ? mask[k] : 1.0;
    if (fm) fm[k] = (float)(ref[k] * m);
    if (g) g[k] = (float)def[k];
    if (maskf) maskf[k] = (float)m;
  }
}

template <typename TI, typename TO>
__global__ __launch_bounds__(256) void gradient_kernel(const TI* __restrict__ img,
                                                       const TI* __restrict__ mask, int64_t M,
                                                       int64_t N, TO* __restrict__ gx,
                                                       TO* __restrict__ gy, int full_grid) {
  // full_grid: outputs are M x N with zero border; else (M-6) x (N-6)
  const int64_t total = full_grid ? M * N : (M - 6) * (N - 6);
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < total;
       k += (int64_t)gridDim.x * blockDim.x) {
    int64_t i, j;
    if (full_grid) {
      i = k % M;
      j = k / M;
      if (i < 3 || i >= M - 3 || j < 3 || j >= N - 3) {
        gx[k] = TO(0);
        gy[k] = TO(0);
        continue;
      }
    } else {
      i = k % (M - 6) + 3;
      j = k / (M - 6) + 3;
    }
    const TI* c = img + i + j * M;
    const double dxv = (-1.0 / 60) * (double)c[-3] + (3.0 / 20) * (double)c[-2] +
                       (-3.0 / 4) * (double)c[-1] + (3.0 / 4) * (double)c[1] +
                       (-3.0 / 20) * (double)c[2] + (1.0 / 60) * (double)c[3];
    const double dyv = (-1.0 / 60) * (double)c[-3 * M] + (3.0 / 20) * (double)c[-2 * M] +
                       (-3.0 / 4) * (double)c[-M] + (3.0 / 4) * (double)c[M] +
                       (-3.0 / 20) * (double)c[2 * M] + (1.0 / 60) * (double)c[3 * M];
    const double m = mask ? (double)mask[i + j * M] : 1.0;
    gx[k] = (TO)(dxv * m);
    gy[k] = (TO)(dyv * m);
  }
}

__global__ __launch_bounds__(256) void place_df_kernel(const double* __restrict__ DfDx,
                                                       const double* __restrict__ DfDy, int64_t M,
                                                       int64_t N, float* __restrict__ fx,
                                                       float* __restrict__ fy) {
  const int64_t MN = M * N;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < MN;
       k += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = k % M, j = k / M;
    if (i < 3 || i >= M - 3 || j < 3 || j >= N - 3) {
      fx[k] = 0.0f;
      fy[k] = 0.0f;
    } else {
      const int64_t c = (i - 3) + (j - 3) * (M - 6);
      fx[k] = (float)DfDx[c];
      fy[k] = (float)DfDy[c];
    }
  }
}

// Sum and sum of squares of a raw 8- or 16-bit image over the ROI
// [x0, x1] x [y0, y1] (0-based, inclusive), exact in 64-bit integers (the
// order of the atomic additions does not matter).
template <class T>
__global__ __launch_bounds__(256) void raw_sums_kernel(const T* __restrict__ raw, int64_t M,
                                                       int64_t x0, int64_t x1, int64_t y0,
                                                       int64_t y1,
                                                       unsigned long long* __restrict__ sums) {
  const int64_t w = x1 - x0 + 1, cnt = w * (y1 - y0 + 1);
  unsigned long long s = 0, q = 0;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < cnt;
       k += (int64_t)gridDim.x * blockDim.x) {
    const unsigned long long v = raw[(x0 + k % w) + (y0 + k / w) * M];
    s += v;
    q += v * v;
  }
  for (int o = ALDIC_WAVE / 2; o > 0; o >>= 1) {
    s += __shfl_xor(s, o);
    q += __shfl_xor(q, o);
  }
  // one atomic pair per block
  __shared__ unsigned long long red[2][256 / ALDIC_WAVE];
  const int wv = threadIdx.x / ALDIC_WAVE;
  if (threadIdx.x % ALDIC_WAVE == 0) {
    red[0][wv] = s;
    red[1][wv] = q;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < (int)(blockDim.x / ALDIC_WAVE); k++) {
      s += red[0][k];
      q += red[1][k];
    }
    atomicAdd(&sums[0], s);
    atomicAdd(&sums[1], q);
  }
}

template <class T>
__global__ __launch_bounds__(256) void raw_normalize_kernel(const T* __restrict__ raw, int64_t n,
                                                            double mean, double sd,
                                                            double* __restrict__ out) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * blockDim.x)
    out[k] = ((double)raw[k] - mean) / sd;
}

__global__ void zero_u64_kernel(unsigned long long* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// ---------------------------------------------------------------------------
// Interpolating cubic B-spline coefficients (not the reference's interpolant,
// whose MEX evaluates a local 4 x 4 kernel on raw pixels): c with
// (c[k-1] + 4 c[k] + c[k+1]) / 6 = s[k] along each line, on the whole-sample
// mirror extension of the image, by the separable causal / anticausal
// recursion with pole z = sqrt(3) - 2 and gain 6 per axis (Unser, Aldroubi,
// Eden, IEEE Trans. Signal Process. 41(2), 1993).
//
// One workgroup per 64 x 64 output tile.  It loads the tile with a 16-pixel
// apron on every side (mirror-reflected at the image border, coalesced along
// dim 1) into LDS, runs the recursion along dim 1 on all 96 columns of the
// apron tile, then along dim 2 on the 64 core rows, and stores the core.  A
// recursion started 16 samples before (and ended 16 after) the core, from a
// zero state, differs from the exact infinite one by at most |z|^16 = 7e-10
// of the signal's scale; no pass over a whole line is needed, so every tile
// is independent and the image is read about 2.25 times (from L2) and
// written once.
//
// Zero rule: a coefficient whose pixel lies in a 4 x 4 window of all-zero
// samples inside the image is set to zero, so that the spline vanishes
// exactly on the cells whose 16 samples are zero, as the reference's kernel
// does there (a zeroed, masked region of ImgDef is what its g == 0 test,
// funICGNQuadtree.m:166-170, detects).  The flags come from the raw apron
// tile before the recursion: all-zero runs of 4 along dim 1, then windows
// of 4 such runs along dim 2, then each core pixel ORs the 16 windows that
// cover it.
// ---------------------------------------------------------------------------
constexpr int BS_CORE = 64, BS_APRON = 16, BS_T = BS_CORE + 2 * BS_APRON, BS_PITCH = BS_T + 1;
constexpr int BS_ZS = BS_CORE + 3;  // window starts per axis: core offsets -3 .. 63
constexpr int BS_ZC = BS_CORE + 6;  // columns the runs are needed on: -3 .. 66

__device__ __forceinline__ int reflect_index(int k, int n) {
  if (n == 1) return 0;
  const int P = 2 * n - 2;
  k %= P;
  if (k < 0) k += P;
  return k < n ? k : P - k;
}

// The recursion along one line of the apron tile (BS_T samples STRIDE
// apart in LDS), held in registers; only the core's coefficients are
// written back (the other pass and the store read nothing else).
template <int STRIDE>
__device__ __forceinline__ void bspline_line(float* line) {
  const float z = -0.26794919243112270f;  // sqrt(3) - 2
  float v[BS_T];
#pragma unroll
  for (int e = 0; e < BS_T; e++) v[e] = line[e * STRIDE];
  float c = 0.0f;
#pragma unroll
  for (int e = 0; e < BS_T; e++) {
    c = fmaf(z, c, v[e]);
    v[e] = c;
  }
  c = 0.0f;
#pragma unroll
  for (int e = BS_T - 1; e >= 0; e--) {
    c = z * (c - v[e]);
    v[e] = 6.0f * c;
  }
#pragma unroll
  for (int e = BS_APRON; e < BS_APRON + BS_CORE; e++) line[e * STRIDE] = v[e];
}

__global__ __launch_bounds__(256) void bspline_prefilter_kernel(const float* __restrict__ in,
                                                                float* __restrict__ out, int M,
                                                                int N) {
  __shared__ float tile[BS_T * BS_PITCH];  // tile[j * BS_PITCH + i], i along dim 1
  __shared__ uint8_t run0[BS_ZC * BS_ZS];   // [column][start]: 4 zero samples down dim 1
  __shared__ uint8_t win0[BS_ZS * BS_ZS];   // [start col][start row]: 4 x 4 zero window
  const int i0 = blockIdx.x * BS_CORE, j0 = blockIdx.y * BS_CORE;
  for (int t = threadIdx.x; t < BS_T * BS_T; t += 256) {
    const int i = t % BS_T, j = t / BS_T;
    const int gi = reflect_index(i0 + i - BS_APRON, M), gj = reflect_index(j0 + j - BS_APRON, N);
    tile[j * BS_PITCH + i] = in[(int64_t)gi + (int64_t)gj * M];
  }
  __syncthreads();
  // zero flags (windows must lie inside the image: the apron's reflected
  // samples do not count)
  for (int t = threadIdx.x; t < BS_ZC * BS_ZS; t += 256) {
    const int a = t % BS_ZS, c = t / BS_ZS;  // start row i0 - 3 + a, column j0 - 3 + c
    const int gi = i0 - 3 + a;
    const float* q = tile + (BS_APRON - 3 + c) * BS_PITCH + BS_APRON - 3 + a;
    run0[t] = gi >= 0 && gi + 3 < M && q[0] == 0.0f && q[1] == 0.0f && q[2] == 0.0f &&
              q[3] == 0.0f;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < BS_ZS * BS_ZS; t += 256) {
    const int a = t % BS_ZS, b = t / BS_ZS;  // start row i0 - 3 + a, column j0 - 3 + b
    const int gj = j0 - 3 + b;
    win0[t] = gj >= 0 && gj + 3 < N && run0[b * BS_ZS + a] && run0[(b + 1) * BS_ZS + a] &&
              run0[(b + 2) * BS_ZS + a] && run0[(b + 3) * BS_ZS + a];
  }
  __syncthreads();
  // each thread keeps the flags of the core pixels it stores (the same
  // t -> (i, j) map as the store below)
  uint32_t kill = 0;
  for (int t = threadIdx.x, k = 0; t < BS_CORE * BS_CORE; t += 256, k++) {
    const int i = t % BS_CORE, j = t / BS_CORE;  // covered by starts a in i..i+3, b in j..j+3
    bool z = false;
#pragma unroll
    for (int db = 0; db < 4; db++)
#pragma unroll
      for (int da = 0; da < 4; da++) z |= win0[(j + db) * BS_ZS + i + da] != 0;
    kill |= (uint32_t)z << k;
  }
  // dim 1: one column of the apron tile per thread; dim 2: one core row
  if (threadIdx.x < BS_T) bspline_line<1>(tile + threadIdx.x * BS_PITCH);
  __syncthreads();
  if (threadIdx.x < BS_CORE) bspline_line<BS_PITCH>(tile + BS_APRON + threadIdx.x);
  __syncthreads();
  for (int t = threadIdx.x, k = 0; t < BS_CORE * BS_CORE; t += 256, k++) {
    const int i = t % BS_CORE, j = t / BS_CORE;
    if (i0 + i < M && j0 + j < N)
      out[(int64_t)(i0 + i) + (int64_t)(j0 + j) * M] =
          (kill >> k) & 1u ? 0.0f : tile[(j + BS_APRON) * BS_PITCH + i + BS_APRON];
  }
}

int grid_for(int64_t n) {
  int64_t g = (n + 255) / 256;
  if (g > 16384) g = 16384;
  if (g < 1) g = 1;
  return (int)g;
}

}  // namespace

namespace aldic {

hipError_t launch_raw_sums(const void* raw, int bits, int64_t M, int64_t x0, int64_t x1,
                           int64_t y0, int64_t y1, unsigned long long* sums, hipStream_t s) {
  zero_u64_kernel<<<1, 64, 0, s>>>(sums, 2);
  const int g = std::min(grid_for((x1 - x0 + 1) * (y1 - y0 + 1)), 1024);
  if (bits == 8)
    raw_sums_kernel<unsigned char><<<g, 256, 0, s>>>(static_cast<const unsigned char*>(raw), M,
                                                     x0, x1, y0, y1, sums);
  else
    raw_sums_kernel<unsigned short><<<g, 256, 0, s>>>(static_cast<const unsigned short*>(raw), M,
                                                      x0, x1, y0, y1, sums);
  return hipGetLastError();
}

hipError_t launch_raw_normalize(const void* raw, int bits, int64_t n, double mean, double sd,
                                double* out, hipStream_t s) {
  if (bits == 8)
    raw_normalize_kernel<unsigned char><<<grid_for(n), 256, 0, s>>>(
        static_cast<const unsigned char*>(raw), n, mean, sd, out);
  else
    raw_normalize_kernel<unsigned short><<<grid_for(n), 256, 0, s>>>(
        static_cast<const unsigned short*>(raw), n, mean, sd, out);
  return hipGetLastError();
}

hipError_t launch_prep_planes(const double* ref, const double* def, const double* mask, int64_t M,
                              int64_t N, float* fm, float* g, float* maskf, hipStream_t s) {
  planes_kernel<<<grid_for(M * N), 256, 0, s>>>(ref, def, mask, M * N, fm, g, maskf);
  return hipGetLastError();
}

hipError_t launch_gradient(const double* ref, const double* mask, int64_t M, int64_t N, float* fx,
                           float* fy, hipStream_t s) {
  gradient_kernel<double, float><<<grid_for(M * N), 256, 0, s>>>(ref, mask, M, N, fx, fy, 1);
  return hipGetLastError();
}

hipError_t launch_gradient_f32(const float* ref, const float* mask, int64_t M, int64_t N, float* fx,
                               float* fy, hipStream_t s) {
  gradient_kernel<float, float><<<grid_for(M * N), 256, 0, s>>>(ref, mask, M, N, fx, fy, 1);
  return hipGetLastError();
}

hipError_t launch_gradient_out_f64(const double* ref, const double* mask, int64_t M, int64_t N,
                                   double* DfDx, double* DfDy, hipStream_t s) {
  gradient_kernel<double, double><<<grid_for((M - 6) * (N - 6)), 256, 0, s>>>(ref, mask, M, N, DfDx,
                                                                              DfDy, 0);
  return hipGetLastError();
}

hipError_t launch_bspline_prefilter(const float* in, float* out, int64_t M, int64_t N,
                                   hipStream_t s) {
  const dim3 grid((unsigned)((M + BS_CORE - 1) / BS_CORE), (unsigned)((N + BS_CORE - 1) / BS_CORE));
  bspline_prefilter_kernel<<<grid, 256, 0, s>>>(in, out, (int)M, (int)N);
  return hipGetLastError();
}

hipError_t launch_place_df(const double* DfDx, const double* DfDy, int64_t M, int64_t N, float* fx,
                           float* fy, hipStream_t s) {
  place_df_kernel<<<grid_for(M * N), 256, 0, s>>>(DfDx, DfDy, M, N, fx, fy);
  return hipGetLastError();
}

}  // namespace aldic

// ---------------------------------------------------------------------------
// Search grid -> node initial guess.  The integer search yields u, v on the
// nY x nX grid of node centres x(ci) = x0 + ci*step (funIntegerSearch.m:
// 175-176); a node list that is not that grid (quadtree nodes, per-node
// subset sizes) takes the bilinear interpolant of the grid, held constant
// beyond its outer nodes.  Output is the interleaved U0 (FP32) the solvers
// read.
// ---------------------------------------------------------------------------
namespace {
// Grid cell (ix, iy) and weights (tx, ty) of node k: bilinear between the
// four grid nodes around it, constant beyond the outer ones.
__device__ __forceinline__ void grid_cell(const double* coords, int n, int k, int nX, int nY,
                                          double x0, double y0, double step, int* ix, int* iy,
                                          double* tx, double* ty) {
  double fx = (coords[k] - x0) / step, fy = (coords[k + n] - y0) / step;
  fx = fmin(fmax(fx, 0.0), (double)(nX - 1));
  fy = fmin(fmax(fy, 0.0), (double)(nY - 1));
  *ix = min((int)floor(fx), max(nX - 2, 0));
  *iy = min((int)floor(fy), max(nY - 2, 0));
  *tx = fx - *ix;
  *ty = fy - *iy;
}

// list (may be null): the nodes to fill, count read on the device; F0
// (may be null) of those nodes is set to zero
__global__ __launch_bounds__(256) void grid_to_nodes_kernel(const double* __restrict__ gu,
                                                            const double* __restrict__ gv,
                                                            int nX, int nY, double x0, double y0,
                                                            double step,
                                                            const double* __restrict__ coords,
                                                            int n, const int* __restrict__ list,
                                                            const int* __restrict__ count,
                                                            float* __restrict__ U0,
                                                            float* __restrict__ F0) {
  const int total = list ? *count : n;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const int k = list ? list[i] : i;
  int ix, iy;
  double tx, ty;
  grid_cell(coords, n, k, nX, nY, x0, y0, step, &ix, &iy, &tx, &ty);
  const int ix1 = min(ix + 1, nX - 1), iy1 = min(iy + 1, nY - 1);
  // grids are nY x nX column-major: (cj, ci) -> cj + ci*nY
  auto at = [&](const double* a, int ci, int cj) { return a[cj + (int64_t)ci * nY]; };
  const double u = (1 - tx) * (1 - ty) * at(gu, ix, iy) + tx * (1 - ty) * at(gu, ix1, iy) +
                   (1 - tx) * ty * at(gu, ix, iy1) + tx * ty * at(gu, ix1, iy1);
  const double v = (1 - tx) * (1 - ty) * at(gv, ix, iy) + tx * (1 - ty) * at(gv, ix1, iy) +
                   (1 - tx) * ty * at(gv, ix, iy1) + tx * ty * at(gv, ix1, iy1);
  U0[2 * k] = (float)u;
  U0[2 * k + 1] = (float)v;
  if (F0)
    for (int q = 0; q < 4; q++) F0[4 * k + q] = 0.0f;
}

// A warp field held at the points of a regular grid (point t = ci + cj*nX
// at (x0 + ci*step, y0 + cj*step); Ug 2 and Fg 4 floats per point, the
// solvers' interleaved layout) -> the bilinear interpolant at every node,
// held constant beyond the outer points.
__global__ __launch_bounds__(256) void field_to_nodes_kernel(const float* __restrict__ Ug,
                                                             const float* __restrict__ Fg,
                                                             int nX, int nY, double x0, double y0,
                                                             double step,
                                                             const double* __restrict__ coords,
                                                             int n, float* __restrict__ U0,
                                                             float* __restrict__ F0) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  int ix, iy;
  double tx, ty;
  grid_cell(coords, n, k, nX, nY, x0, y0, step, &ix, &iy, &tx, &ty);
  const int ix1 = min(ix + 1, nX - 1), iy1 = min(iy + 1, nY - 1);
  const int64_t t00 = ix + (int64_t)iy * nX, t10 = ix1 + (int64_t)iy * nX;
  const int64_t t01 = ix + (int64_t)iy1 * nX, t11 = ix1 + (int64_t)iy1 * nX;
  const double w00 = (1 - tx) * (1 - ty), w10 = tx * (1 - ty), w01 = (1 - tx) * ty, w11 = tx * ty;
  for (int q = 0; q < 2; q++)
    U0[2 * k + q] = (float)(w00 * Ug[2 * t00 + q] + w10 * Ug[2 * t10 + q] +
                            w01 * Ug[2 * t01 + q] + w11 * Ug[2 * t11 + q]);
  if (F0 && Fg)
    for (int q = 0; q < 4; q++)
      F0[4 * k + q] = (float)(w00 * Fg[4 * t00 + q] + w10 * Fg[4 * t10 + q] +
                              w01 * Fg[4 * t01 + q] + w11 * Fg[4 * t11 + q]);
}

// flags of the grid nodes the listed nodes interpolate from
__global__ void mark_cells_kernel(int nX, int nY, double x0, double y0, double step,
                                  const double* __restrict__ coords, int n,
                                  const int* __restrict__ list, const int* __restrict__ count,
                                  int* __restrict__ flags) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= *count) return;
  int ix, iy;
  double tx, ty;
  grid_cell(coords, n, list[i], nX, nY, x0, y0, step, &ix, &iy, &tx, &ty);
  const int ix1 = min(ix + 1, nX - 1), iy1 = min(iy + 1, nY - 1);
  // grid point index t = ci + cj*nX (the search's point order)
  flags[ix + iy * nX] = 1;
  flags[ix1 + iy * nX] = 1;
  flags[ix + iy1 * nX] = 1;
  flags[ix1 + iy1 * nX] = 1;
}

__global__ void compact_flags_kernel(const int* __restrict__ flags, int m, int* __restrict__ out,
                                     int* __restrict__ count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m && flags[i]) out[atomicAdd(count, 1)] = i;
}

__global__ void fill_int_kernel(int* p, int n, int v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
}  // namespace

namespace aldic {
hipError_t launch_grid_to_nodes(const double* gu, const double* gv, int nX, int nY, double x0,
                                double y0, double step, const double* coords, int n, float* U0,
                                hipStream_t s) {
  if (n <= 0) return hipSuccess;
  grid_to_nodes_kernel<<<(n + 255) / 256, 256, 0, s>>>(gu, gv, nX, nY, x0, y0, step, coords, n,
                                                       nullptr, nullptr, U0, nullptr);
  return hipGetLastError();
}

hipError_t launch_field_to_nodes(const float* Ug, const float* Fg, int nX, int nY, double x0,
                                 double y0, double step, const double* coords, int n, float* U0,
                                 float* F0, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  field_to_nodes_kernel<<<(n + 255) / 256, 256, 0, s>>>(Ug, Fg, nX, nY, x0, y0, step, coords, n,
                                                        U0, F0);
  return hipGetLastError();
}

hipError_t launch_grid_to_listed_nodes(const double* gu, const double* gv, int nX, int nY,
                                       double x0, double y0, double step, const double* coords,
                                       int n, const int* list, const int* count, float* U0,
                                       float* F0, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  grid_to_nodes_kernel<<<(n + 255) / 256, 256, 0, s>>>(gu, gv, nX, nY, x0, y0, step, coords, n,
                                                       list, count, U0, F0);
  return hipGetLastError();
}

hipError_t launch_mark_cells(int nX, int nY, double x0, double y0, double step,
                             const double* coords, int n, const int* list, const int* count,
                             int* flags, int* points, int* npoints, hipStream_t s) {
  const int m = nX * nY;
  fill_int_kernel<<<(m + 255) / 256, 256, 0, s>>>(flags, m, 0);
  fill_int_kernel<<<1, 64, 0, s>>>(npoints, 1, 0);
  if (n > 0)
    mark_cells_kernel<<<(n + 255) / 256, 256, 0, s>>>(nX, nY, x0, y0, step, coords, n, list,
                                                      count, flags);
  compact_flags_kernel<<<(m + 255) / 256, 256, 0, s>>>(flags, m, points, npoints);
  return hipGetLastError();
}
}  // namespace aldic
