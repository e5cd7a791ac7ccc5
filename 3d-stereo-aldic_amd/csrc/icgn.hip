// icgn.hip -- gfx950 kernels for the per-subset solvers of ALDIC's local
// matching path:
//   * first-order IC-GN  (func_quadtree/funICGNQuadtree.m:1-454)
//   * ALM subproblem 1   (func_quadtree/funICGN_Subpb1quadtree.m:1-476)
//
// Two kernels per solver share one body each:
//   fast_kernel<MODE,NW,PPL,TS>   NW waves per subset (NW=1: one wave per
//       subset, 4 subsets per workgroup), PPL subset pixels per lane held in
//       VGPRs (f, DfDx, DfDy, offsets), the deformed image staged in an LDS
//       tile of TS x TS floats that is reloaded only when the warped subset
//       leaves it, 4x4 separable spline gathers from that tile, all sums by
//       wave (DPP) reductions in FP32, 6x6 Cholesky in registers.  It runs
//       the reference's unmasked branch.  The moment a zero shows up in f
//       or g (the reference then takes its mask branch, :68-88 and
//       :166-226) or the warp outgrows the tile, it hands the subset to
//   generic_kernel<MODE>   one 256-thread workgroup per subset, any size,
//       full mask semantics (window growth, bwselect flood fills in LDS,
//       per-iteration re-sizing), gathers straight from the global image.
// Both restart nothing the other has written: a handed-on subset is solved
// from scratch by the generic kernel, so the result is the reference's
// whatever path produced it.
#include <algorithm>
#include <type_traits>

#include "aldic_common.h"
#include "aldic_kernels.h"

namespace {

constexpr int MODE_ICGN = 0;
constexpr int MODE_SUBPB1 = 1;

// ---------------------------------------------------------------------------
// Small dense algebra (uniform across the lanes of a subset)
// ---------------------------------------------------------------------------

// Packed lower-triangular Cholesky of a 6x6 SPD matrix stored as the 21
// entries H[r][c], r >= c, column by column (the reference's HGlobal order).
// Returns false when a pivot is not positive.
__device__ __forceinline__ int tri(int r, int c) {
  // index of (r, c), r >= c, in column-major lower-triangular packing
  return c * 6 - (c * (c - 1)) / 2 + (r - c);
}

__device__ __forceinline__ bool chol6(const float h[21], float L[21]) {
  // straight-line: a non-positive pivot clears ok and is replaced by 1 so
  // the remaining columns stay finite (the caller drops L when !ok)
  bool ok = true;
#pragma unroll
  for (int j = 0; j < 6; j++) {
    float d = h[tri(j, j)];
#pragma unroll
    for (int k = 0; k < j; k++) d -= L[tri(j, k)] * L[tri(j, k)];
    ok = ok && (d > 0.0f);
    d = (d > 0.0f) ? d : 1.0f;
    // hardware square root and reciprocal (1 ulp): no denormal rescaling
    // or division fix-up sequences in the uniform setup
    const float ljj = __builtin_amdgcn_sqrtf(d);
    L[tri(j, j)] = ljj;
    const float inv = __builtin_amdgcn_rcpf(ljj);
#pragma unroll
    for (int i = j + 1; i < 6; i++) {
      float s = h[tri(i, j)];
#pragma unroll
      for (int k = 0; k < j; k++) s -= L[tri(i, k)] * L[tri(j, k)];
      L[tri(i, j)] = s * inv;
    }
  }
  return ok;
}

// Forward and back substitution; the diagonal is used through its
// hardware reciprocal (v_rcp_f32, 1 ulp), computed once per solve.
__device__ __forceinline__ void chol6_solve(const float L[21], const float b[6], float x[6]) {
  float y[6], id[6];
#pragma unroll
  for (int i = 0; i < 6; i++) id[i] = __builtin_amdgcn_rcpf(L[tri(i, i)]);
#pragma unroll
  for (int i = 0; i < 6; i++) {
    float s = b[i];
#pragma unroll
    for (int k = 0; k < i; k++) s -= L[tri(i, k)] * y[k];
    y[i] = s * id[i];
  }
#pragma unroll
  for (int i = 5; i >= 0; i--) {
    float s = y[i];
#pragma unroll
    for (int k = i + 1; k < 6; k++) s -= L[tri(k, i)] * x[k];
    x[i] = s * id[i];
  }
}

// LU with partial pivoting for the rare indefinite H (MATLAB's backslash
// falls back to it when Cholesky fails).
// Every index is static (the pivot swap is a select), so the matrix stays in
// registers.
__device__ __forceinline__ void lu6_solve(const float h[21], const float b[6], float x[6]) {
  float A[6][6], bb[6];
#pragma unroll
  for (int r = 0; r < 6; r++) {
    bb[r] = b[r];
#pragma unroll
    for (int c = 0; c < 6; c++) A[r][c] = (r >= c) ? h[tri(r, c)] : h[tri(c, r)];
  }
#pragma unroll
  for (int k = 0; k < 6; k++) {
    int p = k;
    float mx = fabsf(A[k][k]);
#pragma unroll
    for (int i = k + 1; i < 6; i++) {
      const bool take = fabsf(A[i][k]) > mx;
      mx = take ? fabsf(A[i][k]) : mx;
      p = take ? i : p;
    }
#pragma unroll
    for (int i = k + 1; i < 6; i++) {
      const bool sw = (i == p);
#pragma unroll
      for (int j = 0; j < 6; j++) {
        const float t = A[k][j];
        A[k][j] = sw ? A[i][j] : t;
        A[i][j] = sw ? t : A[i][j];
      }
      const float t = bb[k];
      bb[k] = sw ? bb[i] : t;
      bb[i] = sw ? t : bb[i];
    }
#pragma unroll
    for (int i = k + 1; i < 6; i++) {
      const float f = A[i][k] / A[k][k];
#pragma unroll
      for (int j = k; j < 6; j++) A[i][j] -= f * A[k][j];
      bb[i] -= f * bb[k];
    }
  }
#pragma unroll
  for (int i = 5; i >= 0; i--) {
    float s = bb[i];
#pragma unroll
    for (int j = i + 1; j < 6; j++) s -= A[i][j] * x[j];
    x[i] = s / A[i][i];
  }
}

// 2x2 solve of the Subpb1 step (funICGN_Subpb1quadtree.m:393-396).
__device__ __forceinline__ void solve2(const float H[4], float b0, float b1, float* x0,
                                       float* x1) {
  const float a = H[0], b = H[1], c = H[2], d = H[3];
  if (a > 0.0f) {
    const float l11 = sqrtf(a), l21 = c / l11, t = d - l21 * l21;
    if (t > 0.0f) {
      const float l22 = sqrtf(t);
      const float y0 = b0 / l11, y1 = (b1 - l21 * y0) / l22;
      *x1 = y1 / l22;
      *x0 = (y0 - l21 * *x1) / l11;
      return;
    }
  }
  const float det = a * d - b * c;
  *x0 = (d * b0 - b * b1) / det;
  *x1 = (a * b1 - c * b0) / det;
}

// Inverse-compositional update W(P) o W(dP)^-1 (funICGNQuadtree.m:396-414).
// Returns false when det(I + dP_F) == 0.
__device__ __forceinline__ bool ic_update(float P[6], const float d[6]) {
  const float det = (1.0f + d[0]) * (1.0f + d[3]) - d[1] * d[2];
  if (det == 0.0f) return false;
  const float id = __builtin_amdgcn_rcpf(det);
  const float t1 = (-d[0] - d[0] * d[3] + d[1] * d[2]) * id;
  const float t2 = -d[1] * id;
  const float t3 = -d[2] * id;
  const float t4 = (-d[3] - d[0] * d[3] + d[1] * d[2]) * id;
  const float t5 = (-d[4] - d[3] * d[4] + d[2] * d[5]) * id;
  const float t6 = (-d[5] - d[0] * d[5] + d[1] * d[4]) * id;
  const float m11 = 1.0f + P[0], m12 = P[2], m13 = P[4];
  const float m21 = P[1], m22 = 1.0f + P[3], m23 = P[5];
  P[0] = (m11 * (1.0f + t1) + m12 * t2) - 1.0f;
  P[1] = m21 * (1.0f + t1) + m22 * t2;
  P[2] = m11 * t3 + m12 * (1.0f + t4);
  P[3] = (m21 * t3 + m22 * (1.0f + t4)) - 1.0f;
  P[4] = m11 * t5 + m12 * t6 + m13;
  P[5] = m21 * t5 + m22 * t6 + m23;
  return true;
}

// The step-size stopping rule: the update dP = [du/dx dv/dx du/dy dv/dy du
// dv] moves the subset centre by (du, dv) and a point (dx, dy) from it by
// (du/dx dx + du/dy dy, dv/dx dx + dv/dy dy); the rule bounds the root sum
// of squares of the centre shift and the gradient terms at the half widths
// (hx, hy): sqrt(du^2 + dv^2 + (du/dx^2 + dv/dx^2) hx^2 + (du/dy^2 +
// dv/dy^2) hy^2) < tol_dp px.
//
// rule 1: converged when that measure m_k of the k-th update is below tol_dp.
// rule 2 (a-posteriori): also converged when the iteration contracts (m_k <=
// m_{k-1} / 2) and the distance to its fixed point that the observed ratio q
// = m_k / m_{k-1} predicts, m_k q / (1 - q), is below tol_dp, i.e. m_k^2 <
// tol_dp (m_{k-1} - m_k).  *m_prev carries m_{k-1} (negative before the
// first update).
__device__ __forceinline__ bool dp_small(int rule, const float d[6], float hx, float hy,
                                         float tol_dp, float* m_prev) {
  const float m2 = d[4] * d[4] + d[5] * d[5] + (d[0] * d[0] + d[1] * d[1]) * hx * hx +
                   (d[2] * d[2] + d[3] * d[3]) * hy * hy;
  if (m2 < tol_dp * tol_dp) return true;
  if (rule != 2) return false;
  const float m = __builtin_amdgcn_sqrtf(m2), mp = *m_prev;
  *m_prev = m;
  return mp > 0.0f && 2.0f * m <= mp && m2 < tol_dp * (mp - m);
}

// ---------------------------------------------------------------------------
// Group reductions: NW waves cooperate on one subset.
// ---------------------------------------------------------------------------
// NW > 1: each value is summed over 8-lane groups (three DPP adds: a quad
// and the quad four lanes along its 16-lane row), the NW * 8 partial sums
// (lanes 0 and 8 of every row) go to LDS, the first K lanes of the block
// add them up and every lane reads the K totals back: three vector
// operations per value and wave instead of the eleven of a full wave sum.
// red holds group_red_floats<NW>() floats.
constexpr int GROUP_KMAX = 34;
template <int NW>
constexpr int group_red_floats() {
  return NW * 8 * GROUP_KMAX + GROUP_KMAX;
}

// The fast kernel's group sums alternate between two such areas (rb
// toggles at every call, across the subsets a group walks), so that a sum
// never writes the area the previous one is still being read from and its
// guarding barrier can go: the area of call i is written again by call
// i + 2 only, after the barriers of call i + 1.
template <int NW>
constexpr int fast_red_floats() {
  return 2 * group_red_floats<NW>();
}

// Lanes 0-3 of each 16-lane row end with the sum of lanes 0-7, lanes 8-11
// with that of lanes 8-15 (every DPP source lane is valid: bound_ctrl is
// moot and lets the add take the DPP operand directly).
__device__ __forceinline__ float half_row_sum(float v) {
  v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0xB1, 0xF, 0xF, true));
  v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x4E, 0xF, 0xF, true));
  v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x124, 0xF, 0xF, true));
  // the sum is complete here, with every lane active: otherwise the last
  // add sinks into the writers' branch and can no longer take the DPP
  // operand directly
  __asm__ volatile("" : "+v"(v));
  return v;
}

template <int NW, int K, bool GUARD = true>
__device__ __forceinline__ void group_sum(float (&v)[K], float* red) {
  static_assert(K <= GROUP_KMAX, "group_sum holds at most GROUP_KMAX values");
  if constexpr (NW == 1) {
#pragma unroll
    for (int k = 0; k < K; k++) v[k] = wave_sum(v[k]);
  } else {
    constexpr int R = NW * 8;  // 8-lane partial sums of the group
    const int t = threadIdx.x % (ALDIC_WAVE * NW);
#pragma unroll
    for (int k = 0; k < K; k++) v[k] = half_row_sum(v[k]);
    if constexpr (GUARD) __syncthreads();  // the previous sum's totals have been read
    if ((t & 7) == 0) {
#pragma unroll
      for (int k = 0; k < K; k++) red[(t >> 3) * K + k] = v[k];
    }
    __syncthreads();
    if (t < K) {
      float s = 0.0f;
#pragma unroll
      for (int q = 0; q < R; q++) s += red[q * K + t];
      red[R * K + t] = s;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; k++) v[k] = red[R * K + k];
  }
}

template <int NW, int K>
__device__ __forceinline__ void group_sum_db(float (&v)[K], float* red, int& rb) {
  if constexpr (NW == 1) {
    group_sum<1, K>(v, red);
  } else {
    rb ^= 1;
    group_sum<NW, K, false>(v, red + rb * group_red_floats<NW>());
  }
}

// ---------------------------------------------------------------------------
// Per-subset output writers
// ---------------------------------------------------------------------------
// The Hessian is passed by reference with a flag (never as a pointer that
// may be null) so the private array keeps living in registers.
__device__ __forceinline__ void write_icgn(const SolveArgs& a, int node, int st, const float (&P)[6],
                                           const float (&h21)[21], bool has_h, float zn) {
  a.conv[node] = st;
  const bool err = (st == ST_ERR);
  const float nan = __builtin_nanf("");
  a.U[2 * node] = err ? nan : P[4];
  a.U[2 * node + 1] = err ? nan : P[5];
  if (a.F) {
#pragma unroll
    for (int k = 0; k < 4; k++) a.F[4 * node + k] = err ? nan : P[k];
  }
  if (a.H21) {
#pragma unroll
    for (int k = 0; k < 21; k++) a.H21[(int64_t)k * a.n + node] = (err || !has_h) ? 0.0f : h21[k];
  }
  if (a.zncc) a.zncc[node] = zn;
}

// Every output but H; H is zeroed when zero_h (the fast kernel writes a
// subset's H at setup and zeroes it again only when the subset throws).
__device__ __forceinline__ void write_icgn_noh(const SolveArgs& a, int node, int st,
                                               const float (&P)[6], float zn,
                                               bool zero_h = true) {
  a.conv[node] = st;
  const bool err = (st == ST_ERR);
  const float nan = __builtin_nanf("");
  a.U[2 * node] = err ? nan : P[4];
  a.U[2 * node + 1] = err ? nan : P[5];
  if (a.F) {
#pragma unroll
    for (int k = 0; k < 4; k++) a.F[4 * node + k] = err ? nan : P[k];
  }
  if (a.H21 && zero_h) {
#pragma unroll
    for (int k = 0; k < 21; k++) a.H21[(int64_t)k * a.n + node] = 0.0f;
  }
  if (a.zncc) a.zncc[node] = zn;
}

__device__ __forceinline__ void write_subpb1(const SolveArgs& a, int node, int st,
                                             const float P[6]) {
  // A throw leaves ConvItPerEle at its initial 0 (Subpb1Quadtree.m:96-99).
  a.conv[node] = (st == ST_ERR) ? 0 : st;
  const float nan = __builtin_nanf("");
  a.U[2 * node] = (st == ST_ERR) ? nan : P[4];
  a.U[2 * node + 1] = (st == ST_ERR) ? nan : P[5];
}

__device__ __forceinline__ void node_window(const SolveArgs& a, int node, int* x0, int* y0,
                                            int* wx, int* wy) {
  *x0 = (int)mround(a.coords[node]);
  *y0 = (int)mround(a.coords[node + a.n]);
  if (a.ws_per_node) {
    *wx = a.wsx[node];
    *wy = a.wsy ? a.wsy[node] : a.wsx[node];
  } else {
    *wx = a.wsx[0];
    *wy = a.wsy ? a.wsy[0] : a.wsx[0];
  }
}

// ---------------------------------------------------------------------------
// Fast kernel (unmasked branch)
// ---------------------------------------------------------------------------
// Block shape: one wave per subset packs 4 subsets in a 256-thread block;
// several waves per subset take the whole block (barriers stay inside one
// subset's control flow).
template <int NW>
constexpr int fast_block() {
  return NW == 1 ? 256 : ALDIC_WAVE * NW;
}

// One subset (list entry slot) of fast_kernel; its wave group returns from
// it together.  tileb is the group's LDS tile slot (TS * TS + TS + 1
// floats): the tile starts TS + 1 floats into it, so that the gather's
// (-1, -1) tap offset is a non-negative base address.
// FS > 0: every subset of the launch is FS x FS (the plan checked it), so
// the pixel layout, the partial slot and the divisions are compile-time.
template <int MODE, int NW, int PPL, int TS, int FS, bool WALK, int BS>
__device__ __forceinline__ void fast_subset(const SolveArgs& a, int slot, float* tileb,
                                            float* red, int& rb) {
  constexpr int NT = ALDIC_WAVE * NW;
  // IC-GN evaluates b in one pass over the gathered g (no g kept per pixel)
  // in the 65 px class's launch shapes; smaller subsets keep the two-pass
  // residual, whose rounding their gradients need for 1e-5 agreement
  constexpr bool OP = MODE == MODE_ICGN && NT * PPL >= 2048;
  // Column layout of the 65 x 65 subset on four waves (COL): slot k < 16 of
  // lane l in wave w holds the pixel (dx, dy) = (l - 32, 4 (k - 8) + w), so
  // a lane's dx is fixed and its dy steps by 4 from slot to slot; slot 16
  // holds row dy = 32 (wave 0) or column dx = 32 (waves 1-2, 65 pixels).
  // Sums of a lane's pixels then factor into dx times sums over the slot
  // index, and the warp advances by one FMA per slot.  The partial slot is
  // valid for tid < 129, as in the raster layout (p = tid + 256 k).
  constexpr bool COL = OP && FS == 65 && NT == 256 && PPL == 17;
  constexpr int KM = COL ? 16 : 0;  // slots in the column layout
  int tid = threadIdx.x % NT;
  // in a list walk the lane index is made opaque to the optimiser, so that
  // nothing derived from it is hoisted out of the walk and kept live across
  // the whole solve (the one-entry kernel keeps its known range)
  if constexpr (WALK) __asm__ volatile("" : "+v"(tid));
  // every value of one subset is uniform across its waves: keep it scalar
  const int node = __builtin_amdgcn_readfirstlane(a.list[slot]);
  float* tile = tileb + TS + 1;
  const int M = a.M, N = a.N;
  if (M < TS || N < TS) {  // the tile would not fit in the image
    if (tid == 0) {
      const int q = atomicAdd(a.defer_cnt, 1);
      a.defer_list[q] = node;
    }
    return;
  }

  int x0, y0, wx, wy;
  node_window(a, node, &x0, &y0, &wx, &wy);
  x0 = __builtin_amdgcn_readfirstlane(x0);
  y0 = __builtin_amdgcn_readfirstlane(y0);
  wx = __builtin_amdgcn_readfirstlane(wx);
  wy = __builtin_amdgcn_readfirstlane(wy);
  if (FS > 0 && (wx != FS - 1 || wy != FS - 1)) {  // not the size the launch is built for
    if (tid == 0) {
      const int q = atomicAdd(a.defer_cnt, 1);
      a.defer_list[q] = node;
    }
    return;
  }
  const int hx = FS > 0 ? (FS - 1) / 2 : wx / 2, hy = FS > 0 ? (FS - 1) / 2 : wy / 2;
  const int S1 = FS > 0 ? FS : wx + 1, S2 = FS > 0 ? FS : wy + 1;
  const int NS = S1 * S2;
  if (FS == 0 && (NS > NT * PPL || S1 + 8 > TS || S2 + 8 > TS)) {
    // a window larger than this launch shape holds (a winsize array that
    // is not the one the plan was built from): the generic kernel takes it
    if (tid == 0) {
      const int q = atomicAdd(a.defer_cnt, 1);
      a.defer_list[q] = node;
    }
    return;
  }

  float P[6];
  if (MODE == MODE_ICGN) {
    // P0 = [0 0 0 0 U0] (funICGNQuadtree.m:41), or [F0 U0] when a first
    // guess of the gradients is supplied (learned initial guess)
#pragma unroll
    for (int k = 0; k < 4; k++) P[k] = a.FOld ? uni(a.FOld[4 * node + k]) : 0.0f;
    P[4] = uni(a.U0[2 * node]);
    P[5] = uni(a.U0[2 * node + 1]);
  } else {
#pragma unroll
    for (int k = 0; k < 4; k++) P[k] = a.FOld[4 * node + k];
    P[4] = a.UOld[2 * node];
    P[5] = a.UOld[2 * node + 1];
  }
  // Window inside the image (ImgRef rows 1..M) and inside Df.DfDx (rows
  // 4..M-3); outside the latter the reference throws.
  const bool in_img = x0 - hx >= 1 && x0 + hx <= M && y0 - hy >= 1 && y0 + hy <= N;
  const bool in_df = x0 - hx >= 4 && x0 + hx <= M - 3 && y0 - hy >= 4 && y0 + hy <= N - 3;

  // ---- load the reference subset into registers ----
  // Pixel p = tid + NT*k.  Slots k < kfull hold a pixel on every lane, so
  // their sums need no lane mask; the slots of the last, partial round are
  // masked and their offsets set to the centre so every gather stays inside
  // the tile.
  // All loads of the subset are issued before any of them is used (one
  // memory latency for the whole subset): each slot reads at a 32-bit byte
  // offset from the window's first pixel, which an empty slot reads too and
  // then discards.  A window that leaves the image is read from the
  // nearest position inside it: such a subset only needs its hole flag,
  // which is set for it whatever it reads, and the throw of :146-147.
  const int kfull = NS / NT;
  // pd holds the pixel's offsets from the centre (dx, dy) as two halves
  // of one register: integers of magnitude <= 32 are exact in FP16, and the
  // mixed-precision FMA reads either half as an FP32 operand
  float pf[PPL], pgx[PPL], pgy[PPL], pg[PPL];
  half2_t pd[PPL];
  int zf = 0, nanf = 0, tinyf = 0;
  float fmin = 3.0e38f;  // the smallest |f| of the lane's pixels
  // column layout: the lane's column offset and the wave's row phase
  const int lane = tid & (ALDIC_WAVE - 1);
  const int wv = __builtin_amdgcn_readfirstlane(tid / ALDIC_WAVE);
  {
    uint32_t boff[PPL];
#pragma unroll
    for (int k = 0; k < PPL; k++) {
      if (k < KM) {
        boff[k] = 4u * (uint32_t)(lane + (4 * k + wv) * M);
        continue;
      }
      int aa, bb;
      bool v;
      if constexpr (COL) {  // the edge slot
        const int j = (wv - 1) * ALDIC_WAVE + lane;
        v = wv == 0 || j < S2;
        aa = wv == 0 ? lane : S1 - 1;
        bb = wv == 0 ? S2 - 1 : j;
      } else {
        const int p = tid + NT * k;
        aa = p % S1;
        bb = p / S1;
        v = p < NS;
      }
      pd[k] = __builtin_bit_cast(half2_t, __builtin_amdgcn_cvt_pkrtz(v ? (float)(aa - hx) : 0.0f,
                                                                     v ? (float)(bb - hy) : 0.0f));
      boff[k] = v ? 4u * (uint32_t)(aa + bb * M) : 0u;
    }
    const int wx0 = min(max(x0 - hx - 1, 0), M - S1), wy0 = min(max(y0 - hy - 1, 0), N - S2);
    const int64_t w0 = (int64_t)wx0 + (int64_t)wy0 * M;
    const char *fb = (const char*)(a.fm + w0), *xb = (const char*)(a.fx + w0),
               *yb = (const char*)(a.fy + w0);
#pragma unroll
    for (int k = 0; k < PPL; k++) {
      pf[k] = *(const float*)(fb + boff[k]);
      pgx[k] = *(const float*)(xb + boff[k]);
      pgy[k] = *(const float*)(yb + boff[k]);
    }
  }
  // ---- the deformed image's tile ----
  int tx0 = -1000000, ty0 = -1000000;  // tile origin (0-based image index)
  // corner offsets of the subset for the tile bounding box
  const float cdx[2] = {(float)(-hx), (float)(hx)};
  const float cdy[2] = {(float)(-hy), (float)(hy)};
  const float lo_u = 3.0f - (float)x0, hi_u = (float)(M - 2 - x0);
  const float lo_v = 3.0f - (float)y0, hi_v = (float)(N - 2 - y0);
  // The warped subset's bounding box: 0 inside the image and the tile's
  // reach (lox..hiy set), 1 outside the image (the bounds test of
  // :154-157), 2 a NaN warp, 3 larger than the tile.  The warp (:149-152) is
  // affine, so its extremes over the subset are at the four corners.
  // The box's corner extremes (umin, umax, vmin, vmax) of the last test.
  float umin = 0.0f, umax = 0.0f, vmin = 0.0f, vmax = 0.0f;
  auto warp_box = [&](int& lox, int& hix, int& loy, int& hiy) -> int {
    const float A11 = 1.0f + P[0], A12 = P[2], A21 = P[1], A22 = 1.0f + P[3];
    umin = 1e30f, umax = -1e30f, vmin = 1e30f, vmax = -1e30f;
    bool nanw = false;
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
      for (int j = 0; j < 2; j++) {
        const float cu = (A11 * cdx[i] + A12 * cdy[j]) + P[4];
        const float cv = (A21 * cdx[i] + A22 * cdy[j]) + P[5];
        nanw |= (cu != cu) | (cv != cv);
        umin = fminf(umin, cu);
        umax = fmaxf(umax, cu);
        vmin = fminf(vmin, cv);
        vmax = fmaxf(vmax, cv);
      }
    if (!nanw && (umin < lo_u || umax > hi_u || vmin < lo_v || vmax > hi_v)) return 1;
    if (nanw) return 2;
    lox = (int)floorf((float)(x0 - 1) + umin) - 2;
    hix = (int)floorf((float)(x0 - 1) + umax) + 3;
    loy = (int)floorf((float)(y0 - 1) + vmin) - 2;
    hiy = (int)floorf((float)(y0 - 1) + vmax) + 3;
    return (hix - lox + 1 > TS || hiy - loy + 1 > TS) ? 3 : 0;
  };
  // Stage the TS x TS tile at (tx0, ty0) into LDS (the callers order it
  // against the gathers).
  auto stage_tile = [&](auto one_group) {
    // thread t < CG * TS owns tile row t % TS and the columns t / TS +
    // CG k: each load is the launch-uniform column base plus the thread's
    // fixed offset, each LDS store a fixed offset from the thread's slot,
    // so the staging needs no per-element index arithmetic.  The loads go
    // out in groups of TG before their stores (one memory latency a group:
    // one group at setup, where the registers are free, up to two in the
    // loop), at 32-bit byte offsets from the tile's first pixel (a column
    // of the tile is at most TS * M * 4 bytes from it).
    constexpr int CG = NT / TS;                // column groups
    constexpr int KC = (TS + CG - 1) / CG;     // columns per thread
    constexpr int TG = (decltype(one_group)::value || KC <= 16) ? KC : (KC + 1) / 2;
    static_assert(CG >= 1, "a tile row per thread");
    if (tid < CG * TS) {
      const int rr = tid % TS, cg = tid / TS;
      const char* gb = (const char*)(a.g + (int64_t)tx0 + (int64_t)ty0 * M);
      const uint32_t off0 = 4u * (uint32_t)(rr + cg * M), cstep = 4u * (uint32_t)(CG * M);
      float* tt = tile + rr + cg * TS;
#pragma unroll 1
      for (int k0 = 0; k0 < KC; k0 += TG) {
        float tv[TG];
#pragma unroll
        for (int k = 0; k < TG; k++) {
          const bool in = k0 + k < KC && cg + CG * (k0 + k) < TS;
          tv[k] = in ? *(const float*)(gb + (off0 + (uint32_t)(k0 + k) * cstep)) : 0.0f;
        }
#pragma unroll
        for (int k = 0; k < TG; k++) {
          const bool in = k0 + k < KC && cg + CG * (k0 + k) < TS;
          if (in) tt[CG * (k0 + k) * TS] = tv[k];
        }
      }
    }
  };
  // The first warp's tile is staged now, so that its loads overlap the
  // setup below; the loop stages again only when the warp leaves it.
  // Nothing reads the tile before the setup's group sums, whose barriers
  // order these stores before the first gather.  In the column layout the
  // tile's values stay in registers until the Hessian's sum (DEFER): the
  // setup's sums then wait for the subset's loads only, and the tile's
  // loads complete behind their arithmetic.
  constexpr bool DEFER = COL;
  constexpr int TCG = NT / TS, TKC = (TS + TCG - 1) / TCG;  // the staging's shape
  float tv0[DEFER ? TKC : 1];
  bool held = false;
  auto store_held = [&]() {
    if (DEFER && held && tid < TCG * TS) {
      float* tt = tile + tid % TS + (tid / TS) * TS;
#pragma unroll
      for (int k = 0; k < TKC; k++)
        if (tid / TS + TCG * k < TS) tt[TCG * k * TS] = tv0[k];
    }
  };
  {
    int lox = 0, hix = 0, loy = 0, hiy = 0;
    const int box0 = warp_box(lox, hix, loy, hiy);
    if constexpr (DEFER) {
      // The loads go out on every lane and every path, from the clamped
      // origin (whether or not the tile is kept), so that no branch joins
      // paths with different loads in flight and the sums' waits count
      // the subset's loads only.
      const int ox = min(max((lox + hix) / 2 - TS / 2, 0), M - TS);
      const int oy = min(max((loy + hiy) / 2 - TS / 2, 0), N - TS);
      held = box0 == 0;
      if (held) {
        tx0 = ox;
        ty0 = oy;
      }
      __builtin_amdgcn_sched_barrier(0);  // after the subset's loads
      // (the lanes past the staging's, tid >= TCG * TS, read the last
      // column group's pixels and store nothing)
      static_assert(TCG - 1 + TCG * (TKC - 1) < TS, "every column group's columns lie in the tile");
      const int rr = tid % TS, cg = min(tid / TS, TCG - 1);
      const char* gb = (const char*)(a.g + (int64_t)ox + (int64_t)oy * M);
      const uint32_t off0 = 4u * (uint32_t)(rr + cg * M), cstep = 4u * (uint32_t)(TCG * M);
#pragma unroll
      for (int k = 0; k < TKC; k++) tv0[k] = *(const float*)(gb + (off0 + (uint32_t)k * cstep));
    } else if (box0 == 0) {
      tx0 = min(max((lox + hix) / 2 - TS / 2, 0), M - TS);
      ty0 = min(max((loy + hiy) / 2 - TS / 2, 0), N - TS);
      stage_tile(std::true_type{});  // one group: its loads fly with the subset's
      if constexpr (NW == 1) __builtin_amdgcn_wave_barrier();
    }
  }
  // empty slots: f and its gradients 0; the smallest |f| of the lane
#pragma unroll
  for (int k = 0; k < PPL; k++) {
    if (k >= kfull && tid + NT * k >= NS) pf[k] = pgx[k] = pgy[k] = 0.0f;
    if (k < kfull)
      fmin = fminf(fmin, fabsf(pf[k]));
    else
      fmin = fminf(fmin, tid + NT * k < NS ? fabsf(pf[k]) : 3.0e38f);
  }

  // the first moments of f (:119-120, :434: every pixel counts; empty
  // slots hold 0) and whether any pixel has |f| <= 1e-10, in one sum.  A NaN
  // pixel makes sum |f| NaN (fminf passes over it); zeros, the reference's
  // mask, are tiny, so they are counted only when a tiny pixel exists.
  float sumf, sabs;
  {
    float v[3] = {fmin <= 1e-10f ? 1.0f : 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int k = 0; k < PPL; k++) {
      v[1] += pf[k];
      v[2] += fabsf(pf[k]);
    }
    group_sum_db<NW, 3>(v, red, rb);
    tinyf = v[0] > 0.0f;
    sumf = v[1];
    sabs = v[2];
    nanf = sabs != sabs;
    if (tinyf) {  // uniform over the group
      float z[1] = {0.0f};
#pragma unroll
      for (int k = 0; k < PPL; k++)
        z[0] += ((k < kfull || tid + NT * k < NS) && pf[k] == 0.0f) ? 1.0f : 0.0f;
      group_sum_db<NW, 1>(z, red, rb);
      zf = (int)z[0];
    }
  }
  // markCoordHoleStrainOrNot (LocalICGNQuadtree.m:75-86)
  if (MODE == MODE_ICGN && a.hole && tid == 0)
    a.hole[node] = (!in_img || nanf || (float)zf > 0.6f * (float)S1 * (float)S2) ? 1 : 0;
  if (!in_df) {
    if (tid == 0) {
      if (MODE == MODE_ICGN)
        write_icgn_noh(a, node, ST_ERR, P, __builtin_nanf(""));
      else
        write_subpb1(a, node, ST_ERR, P);
    }
    return;
  }
  if (nanf) {  // logical(NaN) throws
    if (tid == 0) {
      if (MODE == MODE_ICGN)
        write_icgn_noh(a, node, ST_ERR, P, __builtin_nanf(""));
      else
        write_subpb1(a, node, ST_ERR, P);
    }
    return;
  }
  if ((float)zf >= 0.4f * (float)S1 * (float)S2) {
    // >= 40% masked: status 102, P stays P0, H = 0 (:440-444)
    if (tid == 0) {
      if (MODE == MODE_ICGN) {
        write_icgn_noh(a, node, ST_MASKED, P, __builtin_nanf(""));
      } else {
        write_subpb1(a, node, ST_MASKED, P);
      }
    }
    return;
  }
  if (tinyf) {
    // a zero in f is the reference's mask branch; a value of magnitude
    // <= 1e-10 leaves the moments' pixel set (:119-120): either way the
    // generic kernel, which carries both rules, takes the subset
    if (tid == 0) {
      const int q = atomicAdd(a.defer_cnt, 1);
      a.defer_list[q] = node;
    }
    return;
  }

  // ---- second moment of f (:119-120) and the Hessian, in one sum ----
  const float meanf = sumf / (float)NS;
  float sqf = 0.0f;
#pragma unroll
  for (int k = 0; k < PPL; k++) {
    const float d = pf[k] - meanf;
    if (k < kfull)
      sqf += d * d;
    else
      sqf += (tid + NT * k < NS) ? d * d : 0.0f;
  }
  float h21[21];
  float L[21];
  float H2[4] = {0, 0, 0, 0};
  float SF[6], S6[6];
  float bottomf;
  if (MODE == MODE_ICGN) {
    constexpr int NH = OP ? 34 : 22;
    float hv[NH];
#pragma unroll
    for (int k = 0; k < NH; k++) hv[k] = 0.0f;
    hv[21] = sqf;
    if constexpr (COL) {
      // sd = (fx dx, fy dx, fx dy, fy dy, fx, fy) = g_i q_a with g = (fx,
      // fy), q = (dx, dy, 1): each H entry is a moment of a gradient product
      // g_i g_j over q_a q_b.  Over the column slots dx is the lane's and dy
      // = 4 kk + w (kk = k - 8), so the lane sums each product, its kk- and
      // kk^2-weighted sums (A), the same of g_i (f - mean f) and of g_i (E,
      // G), and forms the moments once.
      float A[9], E[4], G[4];
#pragma unroll
      for (int j = 0; j < 9; j++) A[j] = 0.0f;
#pragma unroll
      for (int j = 0; j < 4; j++) E[j] = G[j] = 0.0f;
#pragma unroll
      for (int k = 0; k < KM; k++) {
        const float kk = (float)(k - 8), k2 = kk * kk;
        const float gx = pgx[k], gy = pgy[k];
        const float gg[3] = {gx * gx, gx * gy, gy * gy};
#pragma unroll
        for (int j = 0; j < 3; j++) {
          A[3 * j] += gg[j];
          A[3 * j + 1] = fmaf(kk, gg[j], A[3 * j + 1]);
          A[3 * j + 2] = fmaf(k2, gg[j], A[3 * j + 2]);
        }
        const float d = pf[k] - meanf;
        const float ex = gx * d, ey = gy * d;
        E[0] += ex;
        E[1] = fmaf(kk, ex, E[1]);
        E[2] += ey;
        E[3] = fmaf(kk, ey, E[3]);
        G[0] += gx;
        G[1] = fmaf(kk, gx, G[1]);
        G[2] += gy;
        G[3] = fmaf(kk, gy, G[3]);
      }
      const float dx = (float)(lane - hx), w = (float)wv;
      // moments over q: index 0 dx^2, 1 dx dy, 2 dy^2, 3 dx, 4 dy, 5 1
      float mom[3][6];
#pragma unroll
      for (int j = 0; j < 3; j++) {
        const float a0 = A[3 * j], a1 = A[3 * j + 1], a2 = A[3 * j + 2];
        const float my = fmaf(4.0f, a1, w * a0);                          // sum gg dy
        mom[j][2] = fmaf(16.0f, a2, w * fmaf(8.0f, a1, w * a0));          // sum gg dy^2
        mom[j][3] = dx * a0;
        mom[j][0] = dx * mom[j][3];
        mom[j][1] = dx * my;
        mom[j][4] = my;
        mom[j][5] = a0;
      }
      // the q pair of sd indices a, b (0 dx, 1 dy, 2 one)
      constexpr int QQ[3][3] = {{0, 1, 3}, {1, 2, 4}, {3, 4, 5}};
      int t = 0;
#pragma unroll
      for (int c = 0; c < 6; c++)
#pragma unroll
        for (int r = c; r < 6; r++, t++) hv[t] = mom[(r & 1) + (c & 1)][QQ[r >> 1][c >> 1]];
#pragma unroll
      for (int i = 0; i < 2; i++) {
        const float e0 = E[2 * i], e1 = E[2 * i + 1], g0 = G[2 * i], g1 = G[2 * i + 1];
        hv[22 + i] = dx * e0;
        hv[24 + i] = fmaf(4.0f, e1, w * e0);
        hv[26 + i] = e0;
        hv[28 + i] = dx * g0;
        hv[30 + i] = fmaf(4.0f, g1, w * g0);
        hv[32 + i] = g0;
      }
    }
#pragma unroll
    for (int k = KM; k < PPL; k++) {
      const float sd[6] = {fma_lo(pgx[k], pd[k], 0.0f), fma_lo(pgy[k], pd[k], 0.0f),
                           fma_hi(pgx[k], pd[k], 0.0f), fma_hi(pgy[k], pd[k], 0.0f), pgx[k],
                           pgy[k]};
      int t = 0;
#pragma unroll
      for (int c = 0; c < 6; c++)
#pragma unroll
        for (int r = c; r < 6; r++) hv[t++] += sd[r] * sd[c];
      if constexpr (OP) {
        const float d = pf[k] - meanf;  // empty slots: sd = 0
#pragma unroll
        for (int j = 0; j < 6; j++) {
          hv[22 + j] = fmaf(sd[j], d, hv[22 + j]);
          hv[28 + j] += sd[j];
        }
      }
    }
    store_held();  // ordered before the first gather by this sum's barriers
    group_sum_db<NW, NH>(hv, red, rb);
#pragma unroll
    for (int k = 0; k < 21; k++) h21[k] = uni(hv[k]);
    // H goes out now (a later throw overwrites it with zeros), so that its
    // 21 values are not held in registers through the iterations
    if (a.H21 && tid == 0) {
#pragma unroll
      for (int k = 0; k < 21; k++) a.H21[(int64_t)k * a.n + node] = h21[k];
    }
    bottomf = (NS > 1) ? __builtin_amdgcn_sqrtf(hv[21]) : 0.0f;
    if constexpr (OP) {
      const float ibf = __builtin_amdgcn_rcpf(bottomf);
#pragma unroll
      for (int j = 0; j < 6; j++) {
        SF[j] = uni(hv[22 + j] * ibf);
        S6[j] = uni(hv[28 + j]);
      }
    }
    const bool spd = chol6(h21, L);
    if (wave_any(!spd)) {
      // an indefinite H takes MATLAB's LU fallback: the generic kernel
      // carries it, and the fast kernel keeps no LU factors in registers
      if (tid == 0) {
        const int q = atomicAdd(a.defer_cnt, 1);
        a.defer_list[q] = node;
      }
      return;
    }
#pragma unroll
    for (int k = 0; k < 21; k++) L[k] = uni(L[k]);
  } else {
    float v[4] = {0.0f, 0.0f, 0.0f, sqf};
#pragma unroll
    for (int k = 0; k < PPL; k++) {
      v[0] += pgx[k] * pgx[k];
      v[1] += pgx[k] * pgy[k];
      v[2] += pgy[k] * pgy[k];
    }
    group_sum_db<NW, 4>(v, red, rb);
    bottomf = (NS > 1) ? __builtin_amdgcn_sqrtf(v[3]) : 0.0f;
    const float s = 2.0f / (bottomf * bottomf);
    H2[0] = v[0] * s + a.mu;
    H2[1] = v[1] * s;
    H2[2] = v[1] * s;
    H2[3] = v[2] * s + a.mu;
  }
  // fc = (f - meanf) / bottomf, the reference half of the residual
  {
    const float ibf = __builtin_amdgcn_rcpf(bottomf);
#pragma unroll
    for (int k = 0; k < PPL; k++) {
      const float fc = (pf[k] - meanf) * ibf;
      pf[k] = (k < kfull || tid + NT * k < NS) ? fc : 0.0f;
    }
  }

  // ---- Gauss-Newton loop (:135-422) ----
  const float tol = a.tol;
  const int maxit = a.maxit > 0 ? a.maxit : 100;
  float normNew = 1.0f, normAbs = 1.0f, normInit = 0.0f;
  int step = 0;
  // step-size rule (IC-GN only): converged once the applied update moves
  // the subset by less than a.tol_dp px (centre and corners, see dp_small)
  const bool dp_rule = MODE == MODE_ICGN && a.stop_rule >= 1;
  bool dp_conv = false;
  float m_prev = -1.0f;
  float last_zncc = __builtin_nanf("");
  bool err = false;

  // The bounding-box test is skipped while the corners cannot have left
  // the tile or the image: slack is how far they may move from where the
  // last test found them (at the warp Pb), and an affine warp's corners
  // move by at most |du| + |dA11| hx + |dA12| hy (and the same for v).
  // A non-finite warp never passes the comparison, so it is always tested.
  float slack = -1.0f, Pb[6] = {0, 0, 0, 0, 0, 0};
  while (step <= maxit && (dp_rule || (normNew > tol && normAbs > tol))) {
    step++;
    const float A11 = 1.0f + P[0], A12 = P[2], A21 = P[1], A22 = 1.0f + P[3];
    const float mu = fabsf(P[4] - Pb[4]) + fabsf(P[0] - Pb[0]) * (float)hx + fabsf(P[2] - Pb[2]) * (float)hy;
    const float mv = fabsf(P[5] - Pb[5]) + fabsf(P[1] - Pb[1]) * (float)hx + fabsf(P[3] - Pb[3]) * (float)hy;
    // (uniform: P, Pb and slack are the same on every lane)
    if (!__builtin_amdgcn_readfirstlane(fmaxf(mu, mv) < slack)) {
      int lox, hix, loy, hiy;
      const int box = warp_box(lox, hix, loy, hiy);
      if (box == 1) {
        normNew = 1e6f;
        break;
      }
      if (box == 2) {  // interpolating at NaN gives NaN g and logical(NaN) throws
        err = true;
        break;
      }
      if (box == 3) {
        // the warp outgrew the tile: the generic kernel takes the subset
        if (tid == 0) {
          const int q = atomicAdd(a.defer_cnt, 1);
          a.defer_list[q] = node;
        }
        return;
      }
      if (lox < tx0 || hix > tx0 + TS - 1 || loy < ty0 || hiy > ty0 + TS - 1) {
        tx0 = min(max((lox + hix) / 2 - TS / 2, 0), M - TS);
        ty0 = min(max((loy + hiy) / 2 - TS / 2, 0), N - TS);
        if constexpr (NW > 1) __syncthreads();
        stage_tile(std::false_type{});
        if constexpr (NW > 1)
          __syncthreads();
        else
          __builtin_amdgcn_wave_barrier();
      }
      // the room left: the image's bounds test (lo_u ... hi_v) and the
      // tile's reach (floor(x0 - 1 + u) - 2 >= tx0 holds once u >= tx0 -
      // x0 + 4; floor(x0 - 1 + u) + 3 <= tx0 + TS - 1 once u <= tx0 + TS -
      // 3 - x0), less a margin for the rounding of the corners
      const float su = fminf(umin - fmaxf(lo_u, (float)(tx0 - x0 + 4)),
                             fminf(hi_u, (float)(tx0 + TS - 3 - x0)) - umax);
      const float sv = fminf(vmin - fmaxf(lo_v, (float)(ty0 - y0 + 4)),
                             fminf(hi_v, (float)(ty0 + TS - 3 - y0)) - vmax);
      slack = uni(fminf(su, sv) - 1e-3f);
#pragma unroll
      for (int k = 0; k < 6; k++) Pb[k] = uni(P[k]);
    }
    // gather + interpolate.  The value is kept 3136 times the interpolant
    // (unscaled 56x weights on both axes; 36x for the B-spline option): the
    // residual below only sees (g - mean g) / sqrt(sum (g - mean g)^2),
    // which is scale free, and the magnitude test of the reference
    // (|g| <= 1e-10, :272-273) is scaled with it.
    constexpr float TINY = 1e-10f * (BS ? 36.0f : 3136.0f);
    const float cx = (float)(x0 - 1 - tx0) + P[4], cy = (float)(y0 - 1 - ty0) + P[5];
    float s1[2] = {0.0f, 0.0f};  // sum, sum of squares
    float gmin = 3.0e38f;        // smallest |g| of the lane's pixels
    float bg[7] = {0, 0, 0, 0, 0, 0, 0};  // IC-GN: sum sd_j g (6), sum fc g
    // column slots: the lane's first position and the step of 4 rows, and
    // sum g fx, sum g fy and their kk-weighted sums (kk = k - 8)
    float bx = 0.0f, by = 0.0f, s4x = 0.0f, s4y = 0.0f;
    float cs[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    if constexpr (COL) {
      const float dx = (float)(lane - hx), w = (float)wv;
      bx = fmaf(A11, dx, fmaf(A12, w, cx));
      by = fmaf(A21, dx, fmaf(A22, w, cy));
      s4x = 4.0f * A12;
      s4y = 4.0f * A22;
      // per-lane copies: with the slot constant in a scalar register, the
      // per-slot FMA takes them as its vector operand
      __asm__ volatile("" : "+v"(s4x), "+v"(s4y));
    }
#pragma unroll
    for (int k = 0; k < PPL; k++) {
      float tx, ty;
      if (k < KM) {
        tx = fmaf((float)(k - 8), s4x, bx);
        ty = fmaf((float)(k - 8), s4y, by);
      } else {
        tx = fma_lo(A11, pd[k], fma_hi(A12, pd[k], cx));
        ty = fma_lo(A21, pd[k], fma_hi(A22, pd[k], cy));
      }
      // tx, ty >= 2 inside the tile (the bounding-box test above leaves one
      // pixel to spare), where fract(t) = t - floor(t) exactly and the
      // truncating conversion is the floor; the 4 x 4 cell starts one pixel
      // up and left of it
      const float dxf = __builtin_amdgcn_fractf(tx), dyf = __builtin_amdgcn_fractf(ty);
      const int ix = (int)tx, iy = (int)ty;
      float wr[4], wc[4];
      if constexpr (BS) {
        bspline_weights6(dxf, wr);
        bspline_weights6(dyf, wc);
      } else {
        spline_weights56(dxf, wr);
        spline_weights56(dyf, wc);
      }
      const float* base = tileb + __umul24(iy, TS) + ix;
      // The 16 taps are read into registers of their own before the
      // combine.  Read inside the combine's expressions, the compiler has
      // given several instances (the B-spline option, the persistent
      // launch) a minimum-register order in which every read waits for the
      // one before it (one destination pair for all the loop's reads; the
      // B-spline instance took 1.49 instead of 1.07 ms per evaluation);
      // this form keeps the eight two-tap reads of a pixel in flight
      // together in every instance.
      float tp[16];
#pragma unroll
      for (int c = 0; c < 4; c++)
#pragma unroll
        for (int r = 0; r < 4; r++) tp[4 * c + r] = base[c * TS + r];
      float acc = 0.0f;
#pragma unroll
      for (int c = 0; c < 4; c++) {
        const float* row = tp + 4 * c;
        const float rsum = wr[0] * row[0] + wr[1] * row[1] + wr[2] * row[2] + wr[3] * row[3];
        acc = fmaf(wc[c], rsum, acc);
      }
      // k < kfull is uniform (a scalar branch); the partial slots mask by
      // selects, never by a divergent branch
      if (k < kfull) {
        s1[0] += acc;
        s1[1] = fmaf(acc, acc, s1[1]);
        gmin = fminf(gmin, fabsf(acc));
      } else {
        const bool in = tid + NT * k < NS;
        const float av = in ? acc : 0.0f;
        s1[0] += av;
        s1[1] = fmaf(av, av, s1[1]);
        gmin = fminf(gmin, in ? fabsf(acc) : 3.0e38f);
      }
      if (k < KM) {
        const float kk = (float)(k - 8);
        const float t = acc * pgx[k], u = acc * pgy[k];
        cs[0] += t;
        cs[1] = fmaf(kk, t, cs[1]);
        cs[2] += u;
        cs[3] = fmaf(kk, u, cs[3]);
        bg[6] = fmaf(pf[k], acc, bg[6]);
      } else if constexpr (OP) {
        // empty slots carry zero gradients and fc: they add nothing
        const float t = acc * pgx[k], u = acc * pgy[k];
        bg[0] = fma_lo(t, pd[k], bg[0]);
        bg[1] = fma_lo(u, pd[k], bg[1]);
        bg[2] = fma_hi(t, pd[k], bg[2]);
        bg[3] = fma_hi(u, pd[k], bg[3]);
        bg[4] += t;
        bg[5] += u;
        bg[6] = fmaf(pf[k], acc, bg[6]);
      } else {
        pg[k] = acc;
      }
    }
    if constexpr (COL) {
      // sum over the column slots of sd_j g: dx sum g fx, dx sum g fy,
      // sum g fx dy = 4 (sum kk g fx) + w sum g fx, the same for fy, sum g fx,
      // sum g fy
      const float dx = (float)(lane - hx), w = (float)wv;
      bg[0] = fmaf(dx, cs[0], bg[0]);
      bg[1] = fmaf(dx, cs[2], bg[1]);
      bg[2] += fmaf(4.0f, cs[1], w * cs[0]);
      bg[3] += fmaf(4.0f, cs[3], w * cs[2]);
      bg[4] += cs[0];
      bg[5] += cs[2];
    }
    int tiny;
    if constexpr (OP) {
      float v[10] = {s1[0], s1[1], gmin <= TINY ? 1.0f : 0.0f,
                     bg[0], bg[1], bg[2], bg[3], bg[4], bg[5], bg[6]};
      group_sum_db<NW, 10>(v, red, rb);
      s1[0] = v[0];
      s1[1] = v[1];
      tiny = (int)v[2];
#pragma unroll
      for (int j = 0; j < 7; j++) bg[j] = v[3 + j];
    } else {
      float v[3] = {s1[0], s1[1], gmin <= TINY ? 1.0f : 0.0f};
      group_sum_db<NW, 3>(v, red, rb);
      s1[0] = v[0];
      s1[1] = v[1];
      tiny = (int)v[2];
    }
    if (tiny) {
      // a zero in g is the reference's mask branch (:166-170); a tiny value
      // leaves the moments' pixel set: the generic kernel takes the subset
      if (tid == 0) {
        const int q = atomicAdd(a.defer_cnt, 1);
        a.defer_list[q] = node;
      }
      return;
    }
    // compile-time size: the mean by a constant reciprocal; the norm and
    // its reciprocal by the hardware instructions (1 ulp)
    const float meang = FS > 0 ? s1[0] * (1.0f / (float)(FS * FS)) : s1[0] / (float)NS;
    const float varsum = fmaxf(fmaf(-(float)NS * meang, meang, s1[1]), 0.0f);
    const float bottomg = (NS > 1) ? __builtin_amdgcn_sqrtf(varsum) : 0.0f;
    const float ig = __builtin_amdgcn_rcpf(bottomg);
    const float c2 = meang * ig;
    // residual r = (f - mean f)/bottom f - (g - mean g)/bottom g and b
    // (:365-374)
    float bnorm;
    float bvec[6];
    if constexpr (OP) {
      // one pass: b_j = bottomf (SF_j + c2 S_j - ig sum sd_j g) with the
      // subset constants SF_j = sum sd_j fc and S_j = sum sd_j; with gc =
      // ig (g - mean g), sum gc = 0 and sum gc^2 = 1, and as sum fc = 0 and
      // sum fc^2 = 1, sum r^2 = 2 - 2 ig sum fc g
      float nb = 0.0f;
#pragma unroll
      for (int k = 0; k < 6; k++) {
        bvec[k] = bottomf * fmaf(-ig, bg[k], fmaf(c2, S6[k], SF[k]));
        nb += bvec[k] * bvec[k];
      }
      bnorm = __builtin_amdgcn_sqrtf(nb);
      last_zncc = ig * bg[6];
    } else if constexpr (MODE == MODE_ICGN) {
      // two passes: r per pixel from the kept g; empty slots carry zero
      // gradients and contribute to b nothing, only their r^2 is masked
      float v[7] = {0, 0, 0, 0, 0, 0, 0};
#pragma unroll
      for (int k = 0; k < PPL; k++) {
        const float r = fmaf(-ig, pg[k], pf[k] + c2);
        const float rx = r * pgx[k], ry = r * pgy[k];
        v[0] = fma_lo(rx, pd[k], v[0]);
        v[1] = fma_lo(ry, pd[k], v[1]);
        v[2] = fma_hi(rx, pd[k], v[2]);
        v[3] = fma_hi(ry, pd[k], v[3]);
        v[4] += rx;
        v[5] += ry;
        if (k < kfull)
          v[6] = fmaf(r, r, v[6]);
        else
          v[6] += (tid + NT * k < NS) ? r * r : 0.0f;
      }
      group_sum_db<NW, 7>(v, red, rb);
      float nb = 0.0f;
#pragma unroll
      for (int k = 0; k < 6; k++) {
        bvec[k] = bottomf * v[k];
        nb += bvec[k] * bvec[k];
      }
      bnorm = __builtin_amdgcn_sqrtf(nb);
      last_zncc = 1.0f - 0.5f * v[6];
    } else {
      float v[3] = {0, 0, 0};
#pragma unroll
      for (int k = 0; k < PPL; k++) {
        const float r = fmaf(-ig, pg[k], pf[k] + c2);
        v[0] = fmaf(r, pgx[k], v[0]);
        v[1] = fmaf(r, pgy[k], v[1]);
        if (k < kfull)
          v[2] = fmaf(r, r, v[2]);
        else
          v[2] += (tid + NT * k < NS) ? r * r : 0.0f;
      }
      group_sum_db<NW, 3>(v, red, rb);
      const float s = 2.0f / (bottomf * bottomf);
      bvec[0] = bottomf * v[0] * s + a.mu * (P[4] - a.UOld[2 * node] - a.vdual[2 * node]);
      bvec[1] = bottomf * v[1] * s + a.mu * (P[5] - a.UOld[2 * node + 1] - a.vdual[2 * node + 1]);
      bnorm = sqrtf(bvec[0] * bvec[0] + bvec[1] * bvec[1]);
      last_zncc = 1.0f - 0.5f * v[2];
    }
    // convergence rule (:376-389)
    normNew = bnorm;
    normAbs = bnorm;
    if (step == 1) normInit = bnorm;
    normNew = (normInit > tol) ? normNew * __builtin_amdgcn_rcpf(normInit) : 0.0f;
    if (!dp_rule && (normNew < tol || normAbs < tol)) break;
    float dP[6];
    if (MODE == MODE_ICGN) {
      float nb[6];
#pragma unroll
      for (int k = 0; k < 6; k++) nb[k] = -bvec[k];
      chol6_solve(L, nb, dP);
    } else {
      const float dmax = fmaxf(H2[0], H2[3]);
      const float tH[4] = {H2[0] + 1e-3f * dmax, H2[1], H2[2], H2[3] + 1e-3f * dmax};
      dP[0] = dP[1] = dP[2] = dP[3] = 0.0f;
      solve2(tH, -bvec[0], -bvec[1], &dP[4], &dP[5]);
    }
    if (!ic_update(P, dP)) break;
#pragma unroll
    for (int k = 0; k < 6; k++) P[k] = uni(P[k]);
    if (dp_rule && dp_small(a.stop_rule, dP, (float)hx, (float)hy, a.tol_dp, &m_prev)) {
      dp_conv = true;
      break;
    }
  }
  int st;
  if (err) {
    st = ST_ERR;
  } else {
    st = step;
    if (dp_rule ? !dp_conv : !(normNew < tol || normAbs < tol)) st = 101;
    if (normNew != normNew) st = 101;
    if (sabs < 1e-6f) st = 103;
  }
  if (tid == 0) {
    if (MODE == MODE_ICGN)
      write_icgn_noh(a, node, st, P, last_zncc, st == ST_ERR);
    else
      write_subpb1(a, node, st, P);
  }
}

// One subset per wave group (SUBS groups per block).  WALK: a list counted
// on the device (a.dcnt) and launched at a bounded grid, which each group
// walks with a grid stride; otherwise the grid covers the list and every
// group takes one entry (a separate instance: the loop costs the one-entry
// kernel registers; a walking multi-wave one is held to two waves per
// SIMD).  The one-entry instance of the 65 px class at four waves per
// subset and 17 pixels per lane is held to three waves per SIMD (168
// registers: three subsets per CU); its column layout needs about 150.  At
// four waves per SIMD (128 registers, a few spills) the setup was faster
// and an evaluation slower, for the same time at 2.4 evaluations.
template <int NW, int PPL, bool WALK>
constexpr int fast_waves_per_simd() {
  return (NW == 4 && PPL == 17 && !WALK) ? 3 : (WALK && NW > 1 ? 2 : 1);
}

template <int MODE, int NW, int PPL, int TS, int FS, bool WALK, int BS>
__global__ __launch_bounds__(fast_block<NW>(), (fast_waves_per_simd<NW, PPL, WALK>())) void fast_kernel(
    SolveArgs a) {
  constexpr int NT = ALDIC_WAVE * NW;
  constexpr int SUBS = fast_block<NW>() / NT;
  static_assert(NW == 1 || SUBS == 1, "multi-wave subsets own their block");
  __shared__ float tiles[SUBS][TS * TS + TS + 1];
  __shared__ float red[fast_red_floats<NW>()];
  if (a.skip && *a.skip) return;  // uniform: the whole launch is skipped
  const int sub = threadIdx.x / NT;
  int rb = 0;  // the group's reduction area (alternates per sum)
  // a device count is clamped to the list's capacity a.cnt
  const int cnt = a.dcnt ? min(*a.dcnt, a.cnt) : a.cnt;
  // uniform per subset group (SUBS == 1 when NW > 1); the group's barriers
  // and the reduction's barrier keep one entry's LDS use apart from the next
  if constexpr (WALK) {
    for (int slot = blockIdx.x * SUBS + sub; slot < cnt; slot += gridDim.x * SUBS)
      fast_subset<MODE, NW, PPL, TS, FS, true, BS>(a, slot, tiles[sub], red, rb);
  } else {
    const int slot = blockIdx.x * SUBS + sub;
    if (slot < cnt) fast_subset<MODE, NW, PPL, TS, FS, false, BS>(a, slot, tiles[sub], red, rb);
  }
}

// ---------------------------------------------------------------------------
// Persistent kernel: every fast size class in one launch.  The grid is what
// stays resident (three 256-thread blocks per CU); a size-sorted work queue
// feeds it: blocks take whole-block items (65 px, then 33 px subsets, four
// waves each) from one counter while any are left, then each wave takes
// single-wave items (17 px, then 11 px subsets) from a second counter, so
// the small subsets fill the tail of the large ones.  Each subset runs the
// same per-subset code as the per-class launches (bit-identical results)
// and leaves its loop when it converges; masked and odd subsets are handed
// on to the workgroup kernel as there.
// ---------------------------------------------------------------------------
constexpr int PK_TS3 = 84, PK_TS2 = 48, PK_TS1 = 32, PK_TS0 = 24;
constexpr int PK_TILE = PK_TS3 * PK_TS3 + PK_TS3 + 1;  // floats: the largest block tile
static_assert(4 * (PK_TS1 * PK_TS1 + PK_TS1 + 1) <= PK_TILE, "four wave tiles fit the block's");
static_assert(PK_TS2 * PK_TS2 + PK_TS2 + 1 <= PK_TILE, "the 33 px block tile fits");

template <int MODE, int BS>
__global__ __launch_bounds__(256, aldic::PERSIST_BLOCKS_PER_CU) void persistent_kernel(aldic::PersistArgs P) {
  __shared__ float tile[PK_TILE];
  __shared__ float red[fast_red_floats<4>()];
  __shared__ int item_s;
  int rb = 0;  // the block's reduction area (alternates per sum)
  if (P.a.skip && *P.a.skip) return;
  const int n3 = P.cnt[3], n2 = P.cnt[2], n1 = P.cnt[1], n0 = P.cnt[0];
  SolveArgs a = P.a;
  // whole-block items: classes 3 and 2
  for (;;) {
    if (threadIdx.x == 0) item_s = atomicAdd(&P.queue[0], 1);
    __syncthreads();
    const int it = item_s;
    __syncthreads();
    if (it >= n3 + n2) break;
    if (it < n3) {
      a.list = P.list[3];
      if (P.full[3])
        fast_subset<MODE, 4, 17, PK_TS3, 65, true, BS>(a, it, tile, red, rb);
      else
        fast_subset<MODE, 4, 17, PK_TS3, 0, true, BS>(a, it, tile, red, rb);
    } else {
      a.list = P.list[2];
      if (P.full[2])
        fast_subset<MODE, 4, 5, PK_TS2, 33, true, BS>(a, it - n3, tile, red, rb);
      else
        fast_subset<MODE, 4, 5, PK_TS2, 0, true, BS>(a, it - n3, tile, red, rb);
    }
    __syncthreads();  // the block's tile is free for the next item
  }
  // single-wave items: classes 1 and 0, a tile per wave
  const int w = threadIdx.x / ALDIC_WAVE;
  float* wt = tile + w * (PK_TS1 * PK_TS1 + PK_TS1 + 1);
  for (;;) {
    int it = 0;
    if (threadIdx.x % ALDIC_WAVE == 0) it = atomicAdd(&P.queue[1], 1);
    it = __builtin_amdgcn_readfirstlane(__shfl(it, 0));
    if (it >= n1 + n0) break;
    if (it < n1) {
      a.list = P.list[1];
      if (P.full[1])
        fast_subset<MODE, 1, 5, PK_TS1, 17, true, BS>(a, it, wt, red, rb);
      else
        fast_subset<MODE, 1, 5, PK_TS1, 0, true, BS>(a, it, wt, red, rb);
    } else {
      a.list = P.list[0];
      if (P.full[0])
        fast_subset<MODE, 1, 2, PK_TS0, 11, true, BS>(a, it - n1, wt, red, rb);
      else
        fast_subset<MODE, 1, 2, PK_TS0, 0, true, BS>(a, it - n1, wt, red, rb);
    }
  }
}

// ---------------------------------------------------------------------------
// Generic kernel: any size, full mask semantics.  One workgroup per subset.
// ---------------------------------------------------------------------------
constexpr int GEN_THREADS = 256;
constexpr int GEN_MAXPIX = 8192;  // largest (grown) window, e.g. 90 x 90

template <int K>
__device__ __forceinline__ void block_sum(float (&v)[K], float* red) {
  group_sum<GEN_THREADS / ALDIC_WAVE, K>(v, red);
}

// bwselect(BW, c, r, conn) on the S1 x S2 window held in LDS: out = the
// conn-connected component of BW containing (r, c) (1-based).  Rounds of
// line sweeps: each thread owns a line along dim 1 (then along dim 2) and
// walks it both ways, setting every BW pixel whose predecessor in the walk
// is set or that touches a set pixel of the two neighbouring lines (4: the
// same position; 8: also the diagonal ones), until a round changes
// nothing.  A whole run of a line fills in one walk, so a round advances
// the fill around one bend of the component.  Every pixel set is in BW and
// touches a set pixel, so reading a neighbouring line while its owner
// writes it only delays the fill; each pixel is written by one thread per
// phase.
__device__ void lds_bwselect(const uint8_t* bw, uint8_t* out, int S1, int S2, int r, int c,
                             int conn) {
  const int NS = S1 * S2;
  for (int p = threadIdx.x; p < NS; p += GEN_THREADS) out[p] = 0;
  __syncthreads();
  if (threadIdx.x == 0 && r >= 1 && r <= S1 && c >= 1 && c <= S2) {
    const int s = (r - 1) + (c - 1) * S1;
    out[s] = bw[s] ? 1 : 0;
  }
  __syncthreads();
  const bool diag = conn == 8;
  bool again = true;
  while (again) {
    int changed = 0;
    // lines along dim 1: column j, pixels i + j S1, neighbours in columns j +- 1
    for (int j = threadIdx.x; j < S2; j += GEN_THREADS) {
      for (int dir = 0; dir < 2; dir++) {
        int prev = 0;
        for (int k = 0; k < S1; k++) {
          const int i = dir == 0 ? k : S1 - 1 - k;
          const int q = i + j * S1;
          if (!bw[q]) {
            prev = 0;
            continue;
          }
          int v = out[q] | prev;
          if (!v) {
#pragma unroll
            for (int dj = -1; dj <= 1; dj += 2) {
              const int jj = j + dj;
              if (jj < 0 || jj >= S2) continue;
              const uint8_t* o = out + jj * S1;
              v |= o[i];
              if (diag) v |= (i > 0 ? o[i - 1] : 0) | (i + 1 < S1 ? o[i + 1] : 0);
            }
          }
          if (v && !out[q]) {
            out[q] = 1;
            changed = 1;
          }
          prev = v;
        }
      }
    }
    __syncthreads();
    // lines along dim 2: row i, pixels i + j S1, neighbours in rows i +- 1
    for (int i = threadIdx.x; i < S1; i += GEN_THREADS) {
      for (int dir = 0; dir < 2; dir++) {
        int prev = 0;
        for (int k = 0; k < S2; k++) {
          const int j = dir == 0 ? k : S2 - 1 - k;
          const int q = i + j * S1;
          if (!bw[q]) {
            prev = 0;
            continue;
          }
          int v = out[q] | prev;
          if (!v) {
#pragma unroll
            for (int di = -1; di <= 1; di += 2) {
              const int ii = i + di;
              if (ii < 0 || ii >= S1) continue;
              v |= out[ii + j * S1];
              if (diag)
                v |= (j > 0 ? out[ii + (j - 1) * S1] : 0) | (j + 1 < S2 ? out[ii + (j + 1) * S1] : 0);
            }
          }
          if (v && !out[q]) {
            out[q] = 1;
            changed = 1;
          }
          prev = v;
        }
      }
    }
    again = __syncthreads_or(changed) != 0;
  }
}

template <int BS>
__device__ __forceinline__ float interp_global(const float* g, int M, float u22, float v22) {
  // u22, v22 are 1-based (dim-1, dim-2) coordinates inside [3, M-2] x [3, N-2];
  // BS: g holds B-spline coefficients and the weights are the basis'
  const float fu = floorf(u22), fv = floorf(v22);
  float wr[4], wc[4];
  if constexpr (BS) {
    bspline_weights6(u22 - fu, wr);
    bspline_weights6(v22 - fv, wc);
#pragma unroll
    for (int k = 0; k < 4; k++) {
      wr[k] *= 1.0f / 6.0f;
      wc[k] *= 1.0f / 6.0f;
    }
  } else {
    spline_weights(u22 - fu, wr);
    spline_weights(v22 - fv, wc);
  }
  const int ix = (int)fu - 2, iy = (int)fv - 2;  // 0-based first tap
  float acc = 0.0f;
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const float* row = g + (int64_t)(iy + c) * M + ix;
    acc += wc[c] * (wr[0] * row[0] + wr[1] * row[1] + wr[2] * row[2] + wr[3] * row[3]);
  }
  return acc;
}

__device__ __forceinline__ int grow_ws(float A, float B, int w) {
  // 2*max(ceil(sqrt(A/B)*w/2)); -1 for a non-finite size (colon throws)
  const float v = 2.0f * ceilf(sqrtf(A / B) * (float)w * 0.5f);
  if (!(v >= 0.0f) || v > 1e5f) return -1;
  return (int)v;
}

template <int MODE, int BS>
__global__ __launch_bounds__(GEN_THREADS) void generic_kernel(SolveArgs a, const int* list,
                                                              const int* count, int fixed_count) {
  __shared__ float sg[GEN_MAXPIX];
  __shared__ uint8_t sbwf[GEN_MAXPIX];  // bwselect of the f mask (setup)
  __shared__ uint8_t sbwg[GEN_MAXPIX];  // current g mask
  __shared__ uint8_t stmp[GEN_MAXPIX];  // logical image fed to bwselect
  __shared__ uint8_t sfsel[GEN_MAXPIX]; // Subpb1: f selection after its mask block
  __shared__ uint8_t sbwf8[GEN_MAXPIX]; // 8-connected component of the f mask (on first use)
  __shared__ float red[group_red_floats<GEN_THREADS / ALDIC_WAVE>()];
  if (a.skip && *a.skip) return;  // uniform: the whole launch is skipped
  const int total = count ? min(*count, fixed_count) : fixed_count;
  const int M = a.M, N = a.N;
  const float nan = __builtin_nanf("");
  for (int it = blockIdx.x; it < total; it += gridDim.x) {
    const int node = list[it];
    int x0, y0, wx, wy;
    node_window(a, node, &x0, &y0, &wx, &wy);
    const int wx0 = wx;
    float P[6];
    float UOld0 = 0.0f, UOld1 = 0.0f, vd0 = 0.0f, vd1 = 0.0f;
    if (MODE == MODE_ICGN) {
      for (int k = 0; k < 4; k++) P[k] = a.FOld ? a.FOld[4 * node + k] : 0.0f;
      P[4] = a.U0[2 * node];
      P[5] = a.U0[2 * node + 1];
    } else {
      for (int k = 0; k < 4; k++) P[k] = a.FOld[4 * node + k];
      UOld0 = a.UOld[2 * node];
      UOld1 = a.UOld[2 * node + 1];
      vd0 = a.vdual[2 * node];
      vd1 = a.vdual[2 * node + 1];
      P[4] = UOld0;
      P[5] = UOld1;
    }
    float h21[21];
    for (int k = 0; k < 21; k++) h21[k] = 0.0f;
    int st = ST_ERR;
    float last_zncc = nan;
    // window helpers
    auto win_ok = [&](int ax, int ay) {
      return x0 - ax / 2 >= 4 && x0 + ax / 2 <= M - 3 && y0 - ay / 2 >= 4 && y0 + ay / 2 <= N - 3;
    };
    if (MODE == MODE_ICGN && a.hole) {
      // markCoordHoleStrainOrNot on the nominal window (LocalICGNQuadtree.m:75-86)
      const bool in_img = x0 - wx / 2 >= 1 && x0 + wx / 2 <= M && y0 - wy / 2 >= 1 && y0 + wy / 2 <= N;
      float v[2] = {0.0f, 0.0f};
      if (in_img) {
        const int S1 = wx + 1, NS = (wx + 1) * (wy + 1);
        for (int p = threadIdx.x; p < NS; p += GEN_THREADS) {
          const float f = a.fm[(int64_t)(x0 - wx / 2 - 1 + p % S1) + (int64_t)(y0 - wy / 2 - 1 + p / S1) * M];
          v[0] += (f == 0.0f) ? 1.0f : 0.0f;
          v[1] += (f != f) ? 1.0f : 0.0f;
        }
      }
      block_sum<2>(v, red);
      if (threadIdx.x == 0)
        a.hole[node] = (!in_img || v[1] > 0.0f || v[0] > 0.6f * (float)(wx + 1) * (float)(wy + 1)) ? 1 : 0;
    }
    do {
      if (!win_ok(wx, wy)) break;
      int S1 = wx + 1, S2 = wy + 1;
      int NS = S1 * S2;
      if (NS > GEN_MAXPIX) break;  // outside what this build supports: reported as a throw
      // zeros and NaN in tempf = ImgRef .* mask
      float zc = 0.0f, nanc = 0.0f;
      for (int p = threadIdx.x; p < NS; p += GEN_THREADS) {
        const int i = x0 - wx / 2 - 1 + p % S1, j = y0 - wy / 2 - 1 + p / S1;
        const float v = a.fm[(int64_t)i + (int64_t)j * M];
        zc += (v == 0.0f);
        nanc += (v != v);
      }
      {
        float v[2] = {zc, nanc};
        block_sum<2>(v, red);
        zc = v[0];
        nanc = v[1];
      }
      if (nanc > 0.0f) break;
      const float c0 = zc;
      const float area0 = (float)S1 * (float)S2;
      if (c0 >= 0.4f * area0) {
        st = ST_MASKED;
        if (MODE == MODE_ICGN) P[0] = P[1] = P[2] = P[3] = 0.0f;
        break;
      }
      if (c0 > 0.0f) {
        if (MODE == MODE_ICGN) {
          const int wn = grow_ws(area0, area0 - c0, wx);
          if (wn < 0) break;
          wx = wy = wn;
        } else {
          const int wxn = grow_ws(area0, area0 - c0, wx);
          if (wxn < 0) break;
          const float area1 = (float)(wxn + 1) * (float)(wy + 1);
          const int wyn = grow_ws(area1, area1 - c0, wy);
          if (wyn < 0) break;
          wx = wxn;
          wy = wyn;
        }
        if (!win_ok(wx, wy)) break;
        S1 = wx + 1;
        S2 = wy + 1;
        NS = S1 * S2;
        if (NS > GEN_MAXPIX) break;
      }
      const int hx = wx / 2, hy = wy / 2;
      const int ix0 = x0 - hx - 1, iy0 = y0 - hy - 1;  // 0-based window origin
      // seeds: ICGN (f, f); Subpb1 row from winsize_y, column from winsize_x
      const int seed_r = (MODE == MODE_ICGN) ? (wx + 1) / 2 : (wy + 1) / 2;
      const int seed_c = (wx + 1) / 2;
      // bwselect(logical(tempfImgMask), seed, 4)
      for (int p = threadIdx.x; p < NS; p += GEN_THREADS) {
        const int64_t idx = (int64_t)(ix0 + p % S1) + (int64_t)(iy0 + p / S1) * M;
        stmp[p] = a.mask[idx] != 0.0f;
      }
      __syncthreads();
      lds_bwselect(stmp, sbwf, S1, S2, seed_r, seed_c, 4);
      __syncthreads();
      // f statistics and H over the selected pixels
      auto fval = [&](int p, const uint8_t* sel) {
        const int64_t idx = (int64_t)(ix0 + p % S1) + (int64_t)(iy0 + p / S1) * M;
        return sel[p] ? a.fm[idx] : 0.0f;
      };
      float meanf, bottomf, sabs;
      auto f_moments = [&](const uint8_t* sel) {
        float v[3] = {0, 0, 0};
        for (int p = threadIdx.x; p < NS; p += GEN_THREADS) {
          const float f = fval(p, sel);
          const bool use = fabsf(f) > 1e-10f;
          v[0] += use ? 1.0f : 0.0f;
          v[1] += use ? f : 0.0f;
          v[2] += fabsf(f);
        }
        block_sum<3>(v, red);
        meanf = v[1] / v[0];
        sabs = v[2];
        float w[1] = {0.0f};
        for (int p = threadIdx.x; p < NS; p += GEN_THREADS) {
          const float f = fval(p, sel);
          const float d = f - meanf;
          w[0] += (fabsf(f) > 1e-10f) ? d * d : 0.0f;
        }
        block_sum<1>(w, red);
        bottomf = (v[0] > 1.5f) ? sqrtf(w[0]) : (v[0] > 0.5f ? 0.0f : nan);
      };
      float H2[4];
      auto hessian = [&](const uint8_t* sel) {
        if (MODE == MODE_ICGN) {
          for (int k = 0; k < 21; k++) h21[k] = 0.0f;
          for (int p = threadIdx.x; p < NS; p += GEN_THREADS) {
            const int64_t idx = (int64_t)(ix0 + p % S1) + (int64_t)(iy0 + p / S1) * M;
            const float m = sel[p] ? 1.0f : 0.0f;
            const float gx = a.fx[idx] * m, gy = a.fy[idx] * m;
            const float dxm = (float)(p % S1 - hx), dym = (float)(p / S1 - hy);
            const float sd[6] = {dxm * gx, dxm * gy, dym * gx, dym * gy, gx, gy};
            int t = 0;
            for (int c = 0; c < 6; c++)
              for (int r = c; r < 6; r++) h21[t++] += sd[r] * sd[c];
          }
          block_sum<21>(h21, red);
        } else {
          float v[3] = {0, 0, 0};
          for (int p = threadIdx.x; p < NS; p += GEN_THREADS) {
            const int64_t idx = (int64_t)(ix0 + p % S1) + (int64_t)(iy0 + p / S1) * M;
            const float m = sel[p] ? 1.0f : 0.0f;
            const float gx = a.fx[idx] * m, gy = a.fy[idx] * m;
            v[0] += gx * gx;
            v[1] += gx * gy;
            v[2] += gy * gy;
          }
          block_sum<3>(v, red);
          const float s = 2.0f / (bottomf * bottomf);
          H2[0] = v[0] * s + a.mu;
          H2[1] = H2[2] = v[1] * s;
          H2[3] = v[2] * s + a.mu;
        }
      };
      f_moments(sbwf);
      hessian(sbwf);
      float L[21];
      bool spd = (MODE == MODE_ICGN) ? chol6(h21, L) : true;
      // the f selection used in the residual: sbwf (ICGN always; Subpb1 until
      // its mask block replaces it with the g mask)
      const uint8_t* fsel = sbwf;
      const float tol = a.tol;
      const int maxit = a.maxit > 0 ? a.maxit : 100;
      float normNew = 1.0f, normAbs = 1.0f, normInit = 0.0f;
      int step = 0;
      const bool dp_rule = MODE == MODE_ICGN && a.stop_rule >= 1;
      bool dp_conv = false;
      float m_prev = -1.0f;
      bool have_bwg = false, err = false, defer_oob = false, have_f8 = false;
      float sum_bwg = 0.0f;
      const int ws0 = wx0;
      int cur_w = wx;
      while (step <= maxit && (dp_rule || (normNew > tol && normAbs > tol))) {
        step++;
        if (MODE == MODE_ICGN && step > 1 && c0 > 0.0f) {
          // :139-145 -- a change of size throws at the next elementwise op
          if (!have_bwg || !(sum_bwg > 0.0f)) {
            err = true;
            break;
          }
          const int wn = grow_ws((float)(ws0 + 1) * (float)(ws0 + 1), sum_bwg, cur_w);
          if (wn < 0) {
            err = true;
            break;
          }
          if (wn != cur_w) {
            const int h2 = wn / 2;
            if (x0 - h2 < 1 || x0 + h2 > M || y0 - h2 < 1 || y0 + h2 > N) {
              err = true;
              break;
            }
            float o = 0.0f;
            const int n2 = (wn + 1) * (wn + 1);
            for (int p = threadIdx.x; p < n2; p += GEN_THREADS) {
              const float dxm = (float)(p % (wn + 1) - h2), dym = (float)(p / (wn + 1) - h2);
              const float u = ((1.0f + P[0]) * dxm + P[2] * dym) + P[4];
              const float v = (P[1] * dxm + (1.0f + P[3]) * dym) + P[5];
              o += ((u < 3.0f - (float)x0) | (u > (float)(M - 2 - x0)) | (v < 3.0f - (float)y0) |
                    (v > (float)(N - 2 - y0)))
                       ? 1.0f
                       : 0.0f;
            }
            float vv[1] = {o};
            block_sum<1>(vv, red);
            if (vv[0] > 0.0f) {
              defer_oob = true;  // break with status 101
              normNew = 1e6f;
              break;
            }
            err = true;
            break;
          }
        }
        // warp and bounds
        float oob = 0.0f, nanw = 0.0f, zg = 0.0f;
        for (int p = threadIdx.x; p < NS; p += GEN_THREADS) {
          const float dxm = (float)(p % S1 - hx), dym = (float)(p / S1 - hy);
          const float u = ((1.0f + P[0]) * dxm + P[2] * dym) + P[4];
          const float v = (P[1] * dxm + (1.0f + P[3]) * dym) + P[5];
          const bool o = (u < 3.0f - (float)x0) | (u > (float)(M - 2 - x0)) |
                         (v < 3.0f - (float)y0) | (v > (float)(N - 2 - y0));
          const bool nn = (u != u) | (v != v);
          oob += o ? 1.0f : 0.0f;
          nanw += nn ? 1.0f : 0.0f;
          float gv = 0.0f;
          if (!o && !nn) gv = interp_global<BS>(a.g, M, (float)x0 + u, (float)y0 + v);
          sg[p] = gv;
          zg += (gv == 0.0f) ? 1.0f : 0.0f;
        }
        {
          float v[3] = {oob, nanw, zg};
          block_sum<3>(v, red);
          oob = v[0];
          nanw = v[1];
          zg = v[2];
        }
        if (oob > 0.0f) {
          normNew = 1e6f;
          break;
        }
        if (nanw > 0.0f) {
          err = true;
          break;
        }
        if (c0 > 0.0f || zg > 0.0f) {
          // :170-214 (ICGN) / :158-192 (Subpb1): the component of g ~= 0 at
          // the centre; where it is the whole window, the 8-connected
          // component of the f mask instead.  With no zero in g that
          // component is the whole window without any fill, and the f
          // mask's component does not change over the iterations: it is
          // computed once per subset.
          bool allone = true;
          if (zg > 0.0f) {
            for (int p = threadIdx.x; p < NS; p += GEN_THREADS) stmp[p] = sg[p] != 0.0f;
            __syncthreads();
            lds_bwselect(stmp, sbwg, S1, S2, seed_r, seed_c, MODE == MODE_ICGN ? 8 : 4);
            __syncthreads();
            float nz = 0.0f;
            for (int p = threadIdx.x; p < NS; p += GEN_THREADS) nz += sbwg[p] ? 0.0f : 1.0f;
            float vv[1] = {nz};
            block_sum<1>(vv, red);
            allone = vv[0] == 0.0f;
          }
          if (allone) {
            if (!have_f8) {
              for (int p = threadIdx.x; p < NS; p += GEN_THREADS) {
                const int64_t idx = (int64_t)(ix0 + p % S1) + (int64_t)(iy0 + p / S1) * M;
                stmp[p] = a.mask[idx] != 0.0f;
              }
              __syncthreads();
              lds_bwselect(stmp, sbwf8, S1, S2, seed_r, seed_c, 8);
              __syncthreads();
              have_f8 = true;
            }
            for (int p = threadIdx.x; p < NS; p += GEN_THREADS) sbwg[p] = sbwf8[p];
            __syncthreads();
          }
          float sb = 0.0f;
          for (int p = threadIdx.x; p < NS; p += GEN_THREADS) {
            sg[p] = sbwg[p] ? sg[p] : 0.0f;
            sb += sbwg[p] ? 1.0f : 0.0f;
          }
          float sv[1] = {sb};
          block_sum<1>(sv, red);
          sum_bwg = sv[0];
          have_bwg = true;
          if (MODE == MODE_SUBPB1) {
            // f, DfDx, DfDy and H are rebuilt on the g mask (:173-192);
            // tempf keeps the image mask as well
            for (int p = threadIdx.x; p < NS; p += GEN_THREADS) {
              const int64_t idx = (int64_t)(ix0 + p % S1) + (int64_t)(iy0 + p / S1) * M;
              sfsel[p] = sbwg[p] && (a.mask[idx] != 0.0f);
            }
            __syncthreads();
            f_moments(sfsel);
            hessian(sbwg);
            fsel = sfsel;
          }
          // IC-GN recomputes the f moments here over the same selection
          // (sbwf): the values of the setup, kept
          __syncthreads();
        }
        // g moments over |g| > 1e-10 (two passes as in the reference)
        float gm[2] = {0, 0};
        for (int p = threadIdx.x; p < NS; p += GEN_THREADS) {
          const bool use = fabsf(sg[p]) > 1e-10f;
          gm[0] += use ? 1.0f : 0.0f;
          gm[1] += use ? sg[p] : 0.0f;
        }
        block_sum<2>(gm, red);
        const float meang = gm[1] / gm[0];
        float gv[1] = {0.0f};
        for (int p = threadIdx.x; p < NS; p += GEN_THREADS) {
          const float d = sg[p] - meang;
          gv[0] += (fabsf(sg[p]) > 1e-10f) ? d * d : 0.0f;
        }
        block_sum<1>(gv, red);
        const float bottomg = (gm[0] > 1.5f) ? sqrtf(gv[0]) : (gm[0] > 0.5f ? 0.0f : nan);
        // b
        float bvec[6];
        float bnorm;
        const uint8_t* dsel = (MODE == MODE_SUBPB1 && have_bwg) ? sbwg : sbwf;
        if (MODE == MODE_ICGN) {
          float v[7] = {0, 0, 0, 0, 0, 0, 0};
          for (int p = threadIdx.x; p < NS; p += GEN_THREADS) {
            const int64_t idx = (int64_t)(ix0 + p % S1) + (int64_t)(iy0 + p / S1) * M;
            const float m = dsel[p] ? 1.0f : 0.0f;
            const float f = fval(p, fsel);
            const float gx = a.fx[idx] * m, gy = a.fy[idx] * m;
            const float r = (f - meanf) / bottomf - (sg[p] - meang) / bottomg;
            const float dxm = (float)(p % S1 - hx), dym = (float)(p / S1 - hy);
            v[0] += dxm * gx * r;
            v[1] += dxm * gy * r;
            v[2] += dym * gx * r;
            v[3] += dym * gy * r;
            v[4] += gx * r;
            v[5] += gy * r;
            v[6] += r * r;
          }
          block_sum<7>(v, red);
          float nb = 0.0f;
          for (int k = 0; k < 6; k++) {
            bvec[k] = bottomf * v[k];
            nb += bvec[k] * bvec[k];
          }
          bnorm = sqrtf(nb);
          last_zncc = 1.0f - 0.5f * v[6];
        } else {
          float v[3] = {0, 0, 0};
          for (int p = threadIdx.x; p < NS; p += GEN_THREADS) {
            const int64_t idx = (int64_t)(ix0 + p % S1) + (int64_t)(iy0 + p / S1) * M;
            const float m = dsel[p] ? 1.0f : 0.0f;
            const float f = fval(p, fsel);
            const float gx = a.fx[idx] * m, gy = a.fy[idx] * m;
            const float r = (f - meanf) / bottomf - (sg[p] - meang) / bottomg;
            v[0] += gx * r;
            v[1] += gy * r;
            v[2] += r * r;
          }
          block_sum<3>(v, red);
          const float s = 2.0f / (bottomf * bottomf);
          bvec[0] = bottomf * v[0] * s + a.mu * (P[4] - UOld0 - vd0);
          bvec[1] = bottomf * v[1] * s + a.mu * (P[5] - UOld1 - vd1);
          bnorm = sqrtf(bvec[0] * bvec[0] + bvec[1] * bvec[1]);
          last_zncc = 1.0f - 0.5f * v[2];
        }
        normNew = bnorm;
        normAbs = bnorm;
        if (step == 1) normInit = bnorm;
        normNew = (normInit > tol) ? normNew / normInit : 0.0f;
        if (!dp_rule && (normNew < tol || normAbs < tol)) break;
        float dP[6];
        if (MODE == MODE_ICGN) {
          float nb[6];
          for (int k = 0; k < 6; k++) nb[k] = -bvec[k];
          if (spd)
            chol6_solve(L, nb, dP);
          else
            lu6_solve(h21, nb, dP);
        } else {
          const float dmax = (H2[0] != H2[0]) ? H2[3] : ((H2[3] != H2[3]) ? H2[0] : fmaxf(H2[0], H2[3]));
          const float tH[4] = {H2[0] + 1e-3f * dmax, H2[1], H2[2], H2[3] + 1e-3f * dmax};
          dP[0] = dP[1] = dP[2] = dP[3] = 0.0f;
          solve2(tH, -bvec[0], -bvec[1], &dP[4], &dP[5]);
        }
        if (!ic_update(P, dP)) break;
        __syncthreads();
        if (dp_rule && dp_small(a.stop_rule, dP, 0.5f * (float)(S1 - 1), 0.5f * (float)(S2 - 1),
                                a.tol_dp, &m_prev)) {
          dp_conv = true;
          break;
        }
      }
      (void)defer_oob;
      if (err) {
        st = ST_ERR;
        break;
      }
      st = step;
      if (dp_rule ? !dp_conv : !(normNew < tol || normAbs < tol)) st = 101;
      if (normNew != normNew) st = 101;
      // sum |tempf| with the current f selection
      float sa[1] = {0.0f};
      for (int p = threadIdx.x; p < NS; p += GEN_THREADS) sa[0] += fabsf(fval(p, fsel));
      block_sum<1>(sa, red);
      if (sa[0] < 1e-6f) st = 103;
      (void)sabs;
    } while (0);
    if (threadIdx.x == 0) {
      if (MODE == MODE_ICGN)
        write_icgn(a, node, st, P, h21, st != ST_MASKED, st == ST_MASKED ? nan : last_zncc);
      else
        write_subpb1(a, node, st, P);
    }
    __syncthreads();
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// Host-side launchers (declared in aldic_kernels.h)
// ---------------------------------------------------------------------------
namespace aldic {

// Size classes of the fast kernel (largest subset side, LDS tile side) and,
// per class, the launch shapes the autotuner may choose from (waves per
// subset NW, pixels per lane PPL).  Every shape must hold a full square
// subset of its class's largest side (NW * 64 * PPL >= side^2); every tile
// has room for that side, the 4x4 cell and a few pixels of warp.
constexpr int CLS_SIDE[4] = {11, 17, 33, 65};
constexpr int CLS_TS[4] = {24, 32, 48, 84};
constexpr int NVAR = 3;
constexpr int LIST_GRID = 1024;  // blocks of a launch over a device-counted list
constexpr int VAR_NW[4][NVAR] = {{1, 1, 1}, {1, 2, 2}, {1, 2, 4}, {4, 8, 5}};
constexpr int VAR_PPL[4][NVAR] = {{2, 2, 2}, {5, 3, 3}, {18, 9, 5}, {17, 9, 14}};

constexpr bool shapes_hold() {
  for (int c = 0; c < 4; c++) {
    if (CLS_TS[c] < CLS_SIDE[c] + 8) return false;
    for (int v = 0; v < NVAR; v++)
      if (VAR_NW[c][v] * 64 * VAR_PPL[c][v] < CLS_SIDE[c] * CLS_SIDE[c]) return false;
  }
  return true;
}
static_assert(shapes_hold(), "a launch shape cannot hold its class's largest subset");

int size_class(int S1, int S2) {
  const int s = S1 > S2 ? S1 : S2;
  for (int c = 0; c < 4; c++)
    if (s <= CLS_SIDE[c]) return c;
  return 4;  // generic kernel only
}

int num_variants() { return NVAR; }

int class_side(int cls) { return cls >= 0 && cls < 4 ? CLS_SIDE[cls] : 0; }

template <int MODE, int C, int V, int BS>
static void launch_shape(const SolveArgs& a, hipStream_t s) {
  constexpr int NW = VAR_NW[C][V];
  constexpr int block = fast_block<NW>();
  constexpr int subs = block / (ALDIC_WAVE * NW);
  // a list counted on the device is walked by a bounded grid (a few
  // resident rounds), not launched at its full capacity: the idle blocks
  // of a capacity launch would take slots from the kernels beside it
  const int grid = (a.cnt + subs - 1) / subs;
  if (a.dcnt) {
    const int g = std::min(grid, LIST_GRID);
    if (a.fixed_side)
      fast_kernel<MODE, NW, VAR_PPL[C][V], CLS_TS[C], CLS_SIDE[C], true, BS><<<g, block, 0, s>>>(a);
    else
      fast_kernel<MODE, NW, VAR_PPL[C][V], CLS_TS[C], 0, true, BS><<<g, block, 0, s>>>(a);
  } else if (a.fixed_side) {
    fast_kernel<MODE, NW, VAR_PPL[C][V], CLS_TS[C], CLS_SIDE[C], false, BS><<<grid, block, 0, s>>>(a);
  } else {
    fast_kernel<MODE, NW, VAR_PPL[C][V], CLS_TS[C], 0, false, BS><<<grid, block, 0, s>>>(a);
  }
}

template <int MODE, int C, int BS>
static void launch_variant(int var, const SolveArgs& a, hipStream_t s) {
  switch (var) {
    case 1:
      launch_shape<MODE, C, 1, BS>(a, s);
      break;
    case 2:
      launch_shape<MODE, C, 2, BS>(a, s);
      break;
    default:
      launch_shape<MODE, C, 0, BS>(a, s);
      break;
  }
}

template <int MODE, int C>
static void launch_class(int var, const SolveArgs& a, hipStream_t s) {
  if (a.basis)
    launch_variant<MODE, C, 1>(var, a, s);
  else
    launch_variant<MODE, C, 0>(var, a, s);
}

template <int MODE>
static hipError_t launch_fast_mode(int cls, int var, const SolveArgs& a, hipStream_t s) {
  if (a.cnt <= 0) return hipSuccess;
  switch (cls) {
    case 0:
      launch_class<MODE, 0>(var, a, s);
      break;
    case 1:
      launch_class<MODE, 1>(var, a, s);
      break;
    case 2:
      launch_class<MODE, 2>(var, a, s);
      break;
    case 3:
      launch_class<MODE, 3>(var, a, s);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_fast(int mode, int cls, int var, const SolveArgs& a, hipStream_t s) {
  return mode == 0 ? launch_fast_mode<MODE_ICGN>(cls, var, a, s)
                   : launch_fast_mode<MODE_SUBPB1>(cls, var, a, s);
}

hipError_t launch_persistent(int mode, const PersistArgs& P, int grid, hipStream_t s) {
  static_assert(CLS_TS[0] == PK_TS0 && CLS_TS[1] == PK_TS1 && CLS_TS[2] == PK_TS2 &&
                    CLS_TS[3] == PK_TS3,
                "the persistent kernel's tiles are the classes'");
  if (grid <= 0) return hipSuccess;
  if (mode == 0 && P.a.basis)
    persistent_kernel<MODE_ICGN, 1><<<grid, 256, 0, s>>>(P);
  else if (mode == 0)
    persistent_kernel<MODE_ICGN, 0><<<grid, 256, 0, s>>>(P);
  else if (P.a.basis)
    persistent_kernel<MODE_SUBPB1, 1><<<grid, 256, 0, s>>>(P);
  else
    persistent_kernel<MODE_SUBPB1, 0><<<grid, 256, 0, s>>>(P);
  return hipGetLastError();
}

hipError_t launch_generic(int mode, const SolveArgs& a, const int* list, const int* count,
                          int fixed_count, int grid, hipStream_t s) {
  if (mode == 0 && a.basis)
    generic_kernel<MODE_ICGN, 1><<<grid, GEN_THREADS, 0, s>>>(a, list, count, fixed_count);
  else if (mode == 0)
    generic_kernel<MODE_ICGN, 0><<<grid, GEN_THREADS, 0, s>>>(a, list, count, fixed_count);
  else if (a.basis)
    generic_kernel<MODE_SUBPB1, 1><<<grid, GEN_THREADS, 0, s>>>(a, list, count, fixed_count);
  else
    generic_kernel<MODE_SUBPB1, 0><<<grid, GEN_THREADS, 0, s>>>(a, list, count, fixed_count);
  return hipGetLastError();
}

int generic_max_pixels() { return GEN_MAXPIX; }

}  // namespace aldic

namespace {
__global__ void bucket_kernel(SolveArgs a, const int* list, const int* count, int max_count,
                              aldic::ClassLists out) {
  const int total = min(*count, max_count);
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
    const int node = list[i];
    int x0, y0, wx, wy;
    node_window(a, node, &x0, &y0, &wx, &wy);
    const int side = max(wx, wy) + 1;
    int c = 4;
    for (int k = 3; k >= 0; k--)
      if (side <= aldic::CLS_SIDE[k]) c = k;
    int q = c < 4 ? atomicAdd(&out.counts[c], 1) : out.cap[c];
    if (q >= out.cap[c]) {  // class list full (or not allocated): workgroup kernel
      c = 4;
      q = atomicAdd(&out.counts[4], 1);
      if (q >= out.cap[4]) {
        atomicAdd(&out.counts[5], 1);
        continue;
      }
    }
    out.list[c][q] = node;
  }
}
}  // namespace

__global__ void aldic::zero_ints_kernel(int* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

namespace aldic {

hipError_t launch_bucket(const SolveArgs& a, const int* list, const int* count, int max_count,
                         ClassLists out, hipStream_t s) {
  zero_ints_kernel<<<1, 64, 0, s>>>(out.counts, 6);
  const int grid = std::max(1, std::min((max_count + 255) / 256, 1024));
  bucket_kernel<<<grid, 256, 0, s>>>(a, list, count, max_count, out);
  return hipGetLastError();
}

}  // namespace aldic
