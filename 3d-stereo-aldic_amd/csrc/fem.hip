// fem.hip -- the global step of the ALM loop, Subpb2Quadtree
// (func_quadtree/Subpb2Quadtree.m:1-397): one finite-element solve on the
// quadtree Q4 mesh with hanging nodes.
//
// The reference assembles, per element, 16 x 16 matrices
//   A_e = sum_gp det(J) w_xi w_eta ((beta + alpha) DN' DN + mu N' N)
// and right-hand sides
//   b_e = sum_gp det(J) w_xi w_eta (beta diag(DN' F_e') + mu N' N U_e + alpha DN' DN U_e)
// (:211-214; its udual and vdual are multiplied by zero, :42-44), fixes the
// Dirichlet nodes to U, and solves the free rows directly (:353-355).
// Neither matrix couples the x and y components and both component blocks
// are the same, so the solve here is one scalar matrix (rows = nodes)
// applied to the two components at once:
//   A_ij = sum_gp w ((beta + alpha) grad N_i . grad N_j + mu N_i N_j).
// Everything is FP64.  Element work runs one thread per element; the
// element blocks are summed into the CSR values, and the element vectors
// into the nodal right-hand side, by gathers in a fixed order, so a call
// is bit-reproducible.  The free system is solved by conjugate gradients
// with a Jacobi preconditioner, started from U, to a relative residual of
// 1e-13 per component; the iterations run in chunks whose kernels stop
// doing work once a device flag says the solve has converged, so the host
// reads the state once per chunk.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "aldic_common.h"
#include "aldic_internal.h"

namespace aldic {
int set_error(int code, const std::string& msg);
}

struct aldic_fem {
  aldic_ctx* ctx = nullptr;
  int device = 0;
  int64_t n = 0;   // nodes
  int64_t ne = 0;  // elements
  int gauss = 2;
  // host copies for argument checks and the Neumann term
  std::vector<double> h_neumann;  // nn x 4 (node1, node2, n_x, n_y), column-major
  int64_t nn = 0;
  int64_t nfree = 0, ndir = 0;
  // device
  double* xy = nullptr;        // 2n: x then y
  int* elem = nullptr;         // ne x 8 (element-major), -1 = no midpoint node
  int* rowptr = nullptr;       // n + 1
  int* col = nullptr;          // nnz
  double* val = nullptr;       // nnz
  int64_t nnz = 0;
  int* slot_ptr = nullptr;     // nnz + 1
  int* slot_src = nullptr;     // element block entries, e * 64 + a * 8 + b
  int* node_ptr = nullptr;     // n + 1
  int* node_src = nullptr;     // element vector entries, e * 8 + a
  unsigned char* kind = nullptr;  // per node: 0 not solved, 1 free, 2 Dirichlet
  double* neumann = nullptr;   // nn x 4 on the device
  double* Ae = nullptr;        // ne x 64
  double* be = nullptr;        // ne x 16
  double* b = nullptr;         // 2n, nodal right-hand side
  double* x = nullptr;         // 2n
  double* r = nullptr;
  double* z = nullptr;
  double* p = nullptr;
  double* q = nullptr;
  double* dinv = nullptr;      // n
  double* part = nullptr;      // per-block partial sums
  double* sc = nullptr;        // scalars (see cg kernels)
  int* ctr = nullptr;          // [0] finishing-block counter, [1] done flag, [2] iterations
  int nblk = 0;
  // a chunk of CG iterations recorded once as a graph and replayed; rebuilt
  // when the iteration limit it was recorded with changes
  hipGraphExec_t cg_exec = nullptr;
  int32_t cg_exec_maxit = -1;
};

namespace {

constexpr int FEM_T = 256;

// Gauss points and weights of Subpb2Quadtree.m:49-64, literals as written
__device__ __forceinline__ void gauss_rule(int order, int k, double* pt, double* wt) {
  switch (order) {
    case 2: {
      const double g = 1.0 / sqrt(3.0);
      *pt = k == 0 ? -g : g;
      *wt = 1.0;
      break;
    }
    case 3: {
      const double g = sqrt(3.0 / 5.0);
      *pt = k == 0 ? 0.0 : (k == 1 ? g : -g);
      *wt = k == 0 ? 8.0 / 9.0 : 5.0 / 9.0;
      break;
    }
    case 4: {
      const double P[4] = {0.339981, -0.339981, 0.861136, -0.861136};
      const double W[4] = {0.652145, 0.652145, 0.347855, 0.347855};
      *pt = P[k];
      *wt = W[k];
      break;
    }
    default: {
      const double P[5] = {0.0, 0.538469, -0.538469, 0.90618, -0.90618};
      const double W[5] = {0.568889, 0.478629, 0.478629, 0.236927, 0.236927};
      *pt = P[k];
      *wt = W[k];
      break;
    }
  }
}

// MATLAB sign(): -1, 0 or 1
__device__ __forceinline__ double msign(double v) { return v > 0.0 ? 1.0 : (v < 0.0 ? -1.0 : 0.0); }

// Shape functions of the element with hanging midpoint nodes 5..8 on edges
// 2-3, 3-4, 4-1, 1-2 (present where d[k] = 1) and their derivatives in
// (xi, eta), Subpb2Quadtree.m:170-180 and :325-390.
__device__ __forceinline__ void shape(double xi, double eta, const double d[4], double N[8],
                                      double Nxi[8], double Neta[8]) {
  const double ax = fabs(xi), ae = fabs(eta);
  N[4] = d[0] * 0.5 * (1.0 + xi) * (1.0 - ae);
  N[5] = d[1] * 0.5 * (1.0 + eta) * (1.0 - ax);
  N[6] = d[2] * 0.5 * (1.0 - xi) * (1.0 - ae);
  N[7] = d[3] * 0.5 * (1.0 - eta) * (1.0 - ax);
  N[0] = (1.0 - xi) * (1.0 - eta) * 0.25 - 0.5 * (N[6] + N[7]);
  N[1] = (1.0 + xi) * (1.0 - eta) * 0.25 - 0.5 * (N[7] + N[4]);
  N[2] = (1.0 + xi) * (1.0 + eta) * 0.25 - 0.5 * (N[4] + N[5]);
  N[3] = (1.0 - xi) * (1.0 + eta) * 0.25 - 0.5 * (N[5] + N[6]);
  const double se = msign(-eta), sx = msign(-xi);
  Nxi[4] = d[0] * 0.5 * (1.0 - ae);
  Neta[4] = d[0] * 0.5 * (1.0 + xi) * se;
  Nxi[5] = d[1] * 0.5 * (1.0 + eta) * sx;
  Neta[5] = d[1] * 0.5 * (1.0 - ax);
  Nxi[6] = d[2] * 0.5 * (-1.0) * (1.0 - ae);
  Neta[6] = d[2] * 0.5 * (1.0 - xi) * se;
  Nxi[7] = d[3] * 0.5 * (1.0 - eta) * sx;
  Neta[7] = d[3] * 0.5 * (-1.0) * (1.0 - ax);
  Nxi[0] = -0.25 * (1.0 - eta) - 0.5 * (Nxi[6] + Nxi[7]);
  Neta[0] = -0.25 * (1.0 - xi) - 0.5 * (Neta[6] + Neta[7]);
  Nxi[1] = 0.25 * (1.0 - eta) - 0.5 * (Nxi[7] + Nxi[4]);
  Neta[1] = -0.25 * (1.0 + xi) - 0.5 * (Neta[7] + Neta[4]);
  Nxi[2] = 0.25 * (1.0 + eta) - 0.5 * (Nxi[4] + Nxi[5]);
  Neta[2] = 0.25 * (1.0 + xi) - 0.5 * (Neta[4] + Neta[5]);
  Nxi[3] = -0.25 * (1.0 + eta) - 0.5 * (Nxi[5] + Nxi[6]);
  Neta[3] = 0.25 * (1.0 - xi) - 0.5 * (Neta[5] + Neta[6]);
}

// One thread per element: the scalar 8 x 8 block (row-major) and the two
// 8-vectors (interleaved x, y) of the right-hand side.
__global__ void element_kernel(int64_t ne, int64_t n, const int* __restrict__ elem,
                               const double* __restrict__ xy, int order, double beta, double mu,
                               double alpha, const double* __restrict__ U,
                               const double* __restrict__ F, double* __restrict__ Ae,
                               double* __restrict__ be) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  int nd[8];
  double px[8], py[8], ux[8], uy[8], f11[8], f21[8], f12[8], f22[8], d[4];
#pragma unroll
  for (int a = 0; a < 8; a++) {
    nd[a] = elem[e * 8 + a];
    const bool has = nd[a] >= 0;
    px[a] = has ? xy[nd[a]] : 0.0;
    py[a] = has ? xy[n + nd[a]] : 0.0;
    ux[a] = has ? U[2 * nd[a]] : 0.0;
    uy[a] = has ? U[2 * nd[a] + 1] : 0.0;
    f11[a] = has ? F[4 * nd[a]] : 0.0;
    f21[a] = has ? F[4 * nd[a] + 1] : 0.0;
    f12[a] = has ? F[4 * nd[a] + 2] : 0.0;
    f22[a] = has ? F[4 * nd[a] + 3] : 0.0;
  }
#pragma unroll
  for (int k = 0; k < 4; k++) d[k] = nd[4 + k] >= 0 ? 1.0 : 0.0;
  double A[64], bx[8], by[8];
#pragma unroll
  for (int k = 0; k < 64; k++) A[k] = 0.0;
#pragma unroll
  for (int k = 0; k < 8; k++) bx[k] = by[k] = 0.0;
  const double ba = beta + alpha;
  for (int gj = 0; gj < order; gj++) {
    for (int gi = 0; gi < order; gi++) {
      // ndgrid(ksietaList, ksietaList)(:): xi runs fastest (:147-150)
      double xi, eta, wxi, weta;
      gauss_rule(order, gi, &xi, &wxi);
      gauss_rule(order, gj, &eta, &weta);
      double N[8], Nxi[8], Neta[8];
      shape(xi, eta, d, N, Nxi, Neta);
      double J11 = 0, J12 = 0, J21 = 0, J22 = 0;
#pragma unroll
      for (int a = 0; a < 8; a++) {
        J11 += Nxi[a] * px[a];
        J12 += Nxi[a] * py[a];
        J21 += Neta[a] * px[a];
        J22 += Neta[a] * py[a];
      }
      const double det = J11 * J22 - J12 * J21;
      const double id = 1.0 / det;
      double Nx[8], Ny[8];
#pragma unroll
      for (int a = 0; a < 8; a++) {
        Nx[a] = id * J22 * Nxi[a] + id * (-J12) * Neta[a];
        Ny[a] = id * (-J21) * Nxi[a] + id * J11 * Neta[a];
      }
      const double w = det * wxi * weta;
      double NU = 0, NV = 0, GUx = 0, GUy = 0, GVx = 0, GVy = 0;
#pragma unroll
      for (int c = 0; c < 8; c++) {
        NU += N[c] * ux[c];
        NV += N[c] * uy[c];
        GUx += Nx[c] * ux[c];
        GUy += Ny[c] * ux[c];
        GVx += Nx[c] * uy[c];
        GVy += Ny[c] * uy[c];
      }
#pragma unroll
      for (int a = 0; a < 8; a++) {
#pragma unroll
        for (int c = 0; c < 8; c++)
          A[a * 8 + c] += w * (ba * (Nx[a] * Nx[c] + Ny[a] * Ny[c]) + mu * (N[a] * N[c]));
        bx[a] += w * (beta * (Nx[a] * f11[a] + Ny[a] * f12[a]) + mu * N[a] * NU +
                      alpha * (Nx[a] * GUx + Ny[a] * GUy));
        by[a] += w * (beta * (Nx[a] * f21[a] + Ny[a] * f22[a]) + mu * N[a] * NV +
                      alpha * (Nx[a] * GVx + Ny[a] * GVy));
      }
    }
  }
#pragma unroll
  for (int k = 0; k < 64; k++) Ae[e * 64 + k] = A[k];
#pragma unroll
  for (int a = 0; a < 8; a++) {
    be[e * 16 + 2 * a] = bx[a];
    be[e * 16 + 2 * a + 1] = by[a];
  }
}

// CSR values and the nodal right-hand side as ordered sums of the element
// contributions; the Jacobi inverse of the diagonal on the way.
__global__ void gather_kernel(int64_t nnz, int64_t n, const int* __restrict__ slot_ptr,
                              const int* __restrict__ slot_src, const double* __restrict__ Ae,
                              double* __restrict__ val) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nnz) return;
  double v = 0.0;
  for (int k = slot_ptr[s]; k < slot_ptr[s + 1]; k++) v += Ae[slot_src[k]];
  val[s] = v;
}

__global__ void rhs_kernel(int64_t n, const int* __restrict__ node_ptr,
                           const int* __restrict__ node_src, const double* __restrict__ be,
                           double* __restrict__ b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double vx = 0.0, vy = 0.0;
  for (int k = node_ptr[i]; k < node_ptr[i + 1]; k++) {
    const int s = node_src[k];
    const int e = s / 8, a = s % 8;
    vx += be[(int64_t)e * 16 + 2 * a];
    vy += be[(int64_t)e * 16 + 2 * a + 1];
  }
  b[2 * i] = vx;
  b[2 * i + 1] = vy;
}

// Neumann edges in the reference's order (Subpb2Quadtree.m:336-341), one
// thread: BCForce = -F / winstepsize.
__global__ void neumann_kernel(int64_t nn, const double* __restrict__ nm,
                               const double* __restrict__ xy, int64_t n,
                               const double* __restrict__ F, double winstep,
                               double* __restrict__ b) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  for (int64_t t = 0; t < nn; t++) {
    const int i1 = (int)nm[t] - 1, i2 = (int)nm[nn + t] - 1;
    const double n3 = nm[2 * nn + t], n4 = nm[3 * nn + t];
    const double dx = xy[i1] - xy[i2], dy = xy[n + i1] - xy[n + i2];
    const double L = 0.5 * sqrt(dx * dx + dy * dy);
    const int nodes[2] = {i1, i2};
    double addx[2], addy[2];
    const double sc = -1.0 / winstep;
    for (int k = 0; k < 2; k++) {
      const int j = nodes[k];
      const double b11 = sc * F[4 * j], b21 = sc * F[4 * j + 1];
      const double b12 = sc * F[4 * j + 2], b22 = sc * F[4 * j + 3];
      addx[k] = L * (b11 * n3 + b12 * n4);
      addy[k] = L * (b21 * n3 + b22 * n4);
    }
    for (int k = 0; k < 2; k++) b[2 * nodes[k]] += addx[k];
    for (int k = 0; k < 2; k++) b[2 * nodes[k] + 1] += addy[k];
  }
}

// b - A * Uhat on the free rows (Uhat = U on the Dirichlet nodes, 0
// elsewhere), Jacobi inverse, CG start x = U, r = rhs - A x.
__global__ void setup_kernel(int64_t n, const int* __restrict__ rowptr, const int* __restrict__ col,
                             const double* __restrict__ val, const unsigned char* __restrict__ kind,
                             const double* __restrict__ U, const double* __restrict__ b,
                             double* __restrict__ x, double* __restrict__ r,
                             double* __restrict__ dinv) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (kind[i] != 1) {
    x[2 * i] = x[2 * i + 1] = 0.0;
    r[2 * i] = r[2 * i + 1] = 0.0;
    dinv[i] = 0.0;
    return;
  }
  double rx = b[2 * i], ry = b[2 * i + 1], diag = 0.0;
  for (int k = rowptr[i]; k < rowptr[i + 1]; k++) {
    const int j = col[k];
    const double a = val[k];
    if (j == i) diag = a;
    if (kind[j] == 2 || kind[j] == 1) {  // Dirichlet part of b - A Uhat, then - A x0
      rx -= a * U[2 * j];
      ry -= a * U[2 * j + 1];
    }
  }
  x[2 * i] = U[2 * i];
  x[2 * i + 1] = U[2 * i + 1];
  r[2 * i] = rx;
  r[2 * i + 1] = ry;
  dinv[i] = 1.0 / diag;
}

// Block sums of NV values, then the last block to finish adds the block
// sums into out: thread t takes blocks t, t + FEM_T, ... in order, then the
// same shuffle/LDS tree as the block sums.  The order is fixed, whatever
// the order the blocks finish in, so the result is deterministic.
template <int NV>
__device__ bool reduce_store(double (&v)[NV], double* part, int nblk, int* counter, double* out) {
  __shared__ double sh[NV][FEM_T / ALDIC_WAVE];
  __shared__ bool last;
  const int lane = threadIdx.x % ALDIC_WAVE, wv = threadIdx.x / ALDIC_WAVE;
#pragma unroll
  for (int k = 0; k < NV; k++) {
    double s = v[k];
    for (int o = ALDIC_WAVE / 2; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0) sh[k][wv] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; k++) {
      double s = 0.0;
      for (int w = 0; w < FEM_T / ALDIC_WAVE; w++) s += sh[k][w];
      part[(int64_t)k * nblk + blockIdx.x] = s;
    }
    __threadfence();
    last = atomicAdd(counter, 1) == nblk - 1;
  }
  __syncthreads();
  if (last) {
    __threadfence();
    double t[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) {
      t[k] = 0.0;
      for (int b = threadIdx.x; b < nblk; b += FEM_T) t[k] += part[(int64_t)k * nblk + b];
      for (int o = ALDIC_WAVE / 2; o > 0; o >>= 1) t[k] += __shfl_xor(t[k], o);
    }
    __syncthreads();  // every wave is done reading sh from the block sums
    if (lane == 0) {
#pragma unroll
      for (int k = 0; k < NV; k++) sh[k][wv] = t[k];
    }
    __syncthreads();
    if (threadIdx.x < NV) {
      double s = 0.0;
      for (int w = 0; w < FEM_T / ALDIC_WAVE; w++) s += sh[threadIdx.x][w];
      out[threadIdx.x] = s;
    }
    if (threadIdx.x == 0) *counter = 0;
  }
  return last;
}

// sc layout: (r . z) of even iterations at [0,1] with (r . r) after it at
// [2,3], of odd iterations at [10,11] with (r . r) at [12,13]; (p . q) at
// [4,5]; ||rhs||^2 at [6,7].  Iteration k reads its (r . z) from slot k & 1
// and leaves the next one in the other slot, so no kernel overwrites a
// scalar that another block of the same launch may still read.
__device__ __forceinline__ double* rz_slot(double* sc, int s) { return sc + (s ? 10 : 0); }

__global__ void cg_init_kernel(int64_t n, const unsigned char* __restrict__ kind,
                               const double* __restrict__ r, const double* __restrict__ dinv,
                               double* __restrict__ z, double* __restrict__ p,
                               const double* __restrict__ b, const double* __restrict__ U,
                               double* part, int nblk, int* ctr, double* sc) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double v[4] = {0, 0, 0, 0};
  if (i < n && kind[i] == 1) {
    const double zx = r[2 * i] * dinv[i], zy = r[2 * i + 1] * dinv[i];
    z[2 * i] = p[2 * i] = zx;
    z[2 * i + 1] = p[2 * i + 1] = zy;
    v[0] = r[2 * i] * zx;
    v[1] = r[2 * i + 1] * zy;
    v[2] = r[2 * i] * r[2 * i];
    v[3] = r[2 * i + 1] * r[2 * i + 1];
  } else if (i < n) {
    z[2 * i] = p[2 * i] = z[2 * i + 1] = p[2 * i + 1] = 0.0;
  }
  reduce_store<4>(v, part, nblk, ctr, sc);
}

// ||b_free - A_fd U_d||^2 per component: the scale of the stopping test
__global__ void cg_norm_kernel(int64_t n, const int* __restrict__ rowptr,
                               const int* __restrict__ col, const double* __restrict__ val,
                               const unsigned char* __restrict__ kind, const double* __restrict__ U,
                               const double* __restrict__ b, double* part, int nblk, int* ctr,
                               double* sc) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double v[2] = {0, 0};
  if (i < n && kind[i] == 1) {
    double rx = b[2 * i], ry = b[2 * i + 1];
    for (int k = rowptr[i]; k < rowptr[i + 1]; k++) {
      const int j = col[k];
      if (kind[j] == 2) {
        rx -= val[k] * U[2 * j];
        ry -= val[k] * U[2 * j + 1];
      }
    }
    v[0] = rx * rx;
    v[1] = ry * ry;
  }
  reduce_store<2>(v, part, nblk, ctr, sc + 6);
}

// q = A p on the free rows and (p . q) per component
__global__ void cg_spmv_kernel(int64_t n, const int* __restrict__ rowptr,
                               const int* __restrict__ col, const double* __restrict__ val,
                               const unsigned char* __restrict__ kind,
                               const double* __restrict__ p, double* __restrict__ q, double* part,
                               int nblk, int* ctr, double* sc) {
  if (ctr[1]) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double v[2] = {0, 0};
  if (i < n && kind[i] == 1) {
    double qx = 0.0, qy = 0.0;
    const double2* p2 = reinterpret_cast<const double2*>(p);
    int k = rowptr[i];
    const int ke = rowptr[i + 1];
    // four entries at a time so their loads are in flight together; the
    // sums keep the entry order
    for (; k + 4 <= ke; k += 4) {
      int j[4];
      double a[4];
      double2 pj[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        j[u] = col[k + u];
        a[u] = val[k + u];
      }
#pragma unroll
      for (int u = 0; u < 4; u++) pj[u] = p2[j[u]];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        qx += a[u] * pj[u].x;  // p is zero off the free rows
        qy += a[u] * pj[u].y;
      }
    }
    for (; k < ke; k++) {
      const double2 pj = p2[col[k]];
      const double a = val[k];
      qx += a * pj.x;
      qy += a * pj.y;
    }
    q[2 * i] = qx;
    q[2 * i + 1] = qy;
    v[0] = p[2 * i] * qx;
    v[1] = p[2 * i + 1] * qy;
  }
  reduce_store<2>(v, part, nblk, ctr, sc + 4);
}

// x += a p, r -= a q, z = D^-1 r, then (r . z) and (r . r) into the other
// slot; the block that finishes the sums also counts the iteration and sets
// the done flag (converged, iteration limit, or NaN)
__global__ void cg_update_kernel(int64_t n, const unsigned char* __restrict__ kind,
                                 double* __restrict__ x, double* __restrict__ r,
                                 double* __restrict__ z, const double* __restrict__ p,
                                 const double* __restrict__ q, const double* __restrict__ dinv,
                                 double* part, int nblk, int* ctr, double* sc, int par,
                                 double tol2, int max_iter) {
  if (ctr[1]) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const double* rz = rz_slot(sc, par);
  double* rzn = rz_slot(sc, par ^ 1);
  double v[4] = {0, 0, 0, 0};
  if (i < n && kind[i] == 1) {
#pragma unroll
    for (int c = 0; c < 2; c++) {
      const double a = sc[4 + c] != 0.0 ? rz[c] / sc[4 + c] : 0.0;
      x[2 * i + c] += a * p[2 * i + c];
      const double rr = r[2 * i + c] - a * q[2 * i + c];
      r[2 * i + c] = rr;
      const double zz = rr * dinv[i];
      z[2 * i + c] = zz;
      v[c] = rr * zz;
      v[2 + c] = rr * rr;
    }
  }
  if (reduce_store<4>(v, part, nblk, ctr, rzn)) {
    __syncthreads();
    if (threadIdx.x == 0) {
      const int it = ctr[2] + 1;
      ctr[2] = it;
      const double rx = rzn[2], ry = rzn[3];
      const bool conv = rx <= tol2 * sc[6] && ry <= tol2 * sc[7];
      if (conv || it >= max_iter || !(rx == rx) || !(ry == ry)) ctr[1] = 1;
    }
  }
}

// p = z + (rz_new / rz) p, rz_new from the slot the update filled
__global__ void cg_dir_kernel(int64_t n, const unsigned char* __restrict__ kind,
                              const double* __restrict__ z, double* __restrict__ p,
                              const double* sc, const int* ctr, int par) {
  if (ctr[1]) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const double* rz = sc + (par ? 10 : 0);
  const double* rzn = sc + (par ? 0 : 10);
  const double bx = rzn[0] != 0.0 && rz[0] != 0.0 ? rzn[0] / rz[0] : 0.0;
  const double by = rzn[1] != 0.0 && rz[1] != 0.0 ? rzn[1] / rz[1] : 0.0;
  if (i < n && kind[i] == 1) {
    p[2 * i] = z[2 * i] + bx * p[2 * i];
    p[2 * i + 1] = z[2 * i + 1] + by * p[2 * i + 1];
  }
}

// The whole CG iteration in one workgroup, for meshes of up to
// CG_SMALL_N nodes: the chunked launches above cost a few microseconds of
// dispatch per kernel (three per iteration, and the asynchronous solve
// launches every chunk its limit allows), while one CU runs such a mesh's
// SpMV in about as long.  Same iteration as cg_spmv/update/dir (the
// Jacobi-preconditioned CG on the free rows, both components, stopped by
// the same per-component test after each update, or the limit, or a NaN);
// every block sum is a fixed tree, so the result is deterministic.  Starts
// from what cg_init left (x, r, z, p and the (r . z), (r . r) and norm
// scalars); leaves the iteration count in ctr[2] and the done flag set.
constexpr int CG_BT = 1024;        // threads of the one workgroup
constexpr int64_t CG_SMALL_N = 8192;

template <int NV>
__device__ __forceinline__ void block_sum_cg(double (&v)[NV], double (*sh)[CG_BT / ALDIC_WAVE]) {
  const int lane = threadIdx.x % ALDIC_WAVE, wv = threadIdx.x / ALDIC_WAVE;
#pragma unroll
  for (int k = 0; k < NV; k++) {
    double t = v[k];
    for (int o = ALDIC_WAVE / 2; o > 0; o >>= 1) t += __shfl_xor(t, o);
    if (lane == 0) sh[k][wv] = t;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; k++) {
    double t = 0.0;
    for (int w = 0; w < CG_BT / ALDIC_WAVE; w++) t += sh[k][w];
    v[k] = t;
  }
  __syncthreads();  // every thread has read sh before the next sum writes it
}

__global__ __launch_bounds__(CG_BT) void cg_block_kernel(
    int64_t n, const int* __restrict__ rowptr, const int* __restrict__ col,
    const double* __restrict__ val, const unsigned char* __restrict__ kind,
    double* __restrict__ x, double* __restrict__ r, double* __restrict__ z, double* p,
    double* __restrict__ q, const double* __restrict__ dinv, int* ctr, double* sc, double tol2,
    int max_iter) {
  __shared__ double sh[4][CG_BT / ALDIC_WAVE];
  if (ctr[1]) return;  // a skipped solve
  double rz[2] = {sc[0], sc[1]};
  const double nx = sc[6], ny = sc[7];
  const double2* p2 = reinterpret_cast<const double2*>(p);
  int it = 0;
  while (true) {
    // q = A p and (p . q) per component
    double pq[2] = {0.0, 0.0};
    for (int64_t i = threadIdx.x; i < n; i += CG_BT) {
      if (kind[i] != 1) continue;
      double qx = 0.0, qy = 0.0;
      for (int k = rowptr[i]; k < rowptr[i + 1]; k++) {
        const double2 pj = p2[col[k]];  // p is zero off the free rows
        const double a = val[k];
        qx += a * pj.x;
        qy += a * pj.y;
      }
      q[2 * i] = qx;
      q[2 * i + 1] = qy;
      pq[0] += p[2 * i] * qx;
      pq[1] += p[2 * i + 1] * qy;
    }
    block_sum_cg<2>(pq, sh);
    // x += a p, r -= a q, z = D^-1 r, (r . z) and (r . r)
    const double al[2] = {pq[0] != 0.0 ? rz[0] / pq[0] : 0.0, pq[1] != 0.0 ? rz[1] / pq[1] : 0.0};
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (int64_t i = threadIdx.x; i < n; i += CG_BT) {
      if (kind[i] != 1) continue;
#pragma unroll
      for (int c = 0; c < 2; c++) {
        x[2 * i + c] += al[c] * p[2 * i + c];
        const double rr = r[2 * i + c] - al[c] * q[2 * i + c];
        r[2 * i + c] = rr;
        const double zz = rr * dinv[i];
        z[2 * i + c] = zz;
        v[c] += rr * zz;
        v[2 + c] += rr * rr;
      }
    }
    block_sum_cg<4>(v, sh);
    it++;
    const bool conv = v[2] <= tol2 * nx && v[3] <= tol2 * ny;
    if (conv || it >= max_iter || !(v[2] == v[2]) || !(v[3] == v[3])) break;
    // p = z + (rz_new / rz) p
    const double bx = v[0] != 0.0 && rz[0] != 0.0 ? v[0] / rz[0] : 0.0;
    const double by = v[1] != 0.0 && rz[1] != 0.0 ? v[1] / rz[1] : 0.0;
    for (int64_t i = threadIdx.x; i < n; i += CG_BT) {
      if (kind[i] != 1) continue;
      p[2 * i] = z[2 * i] + bx * p[2 * i];
      p[2 * i + 1] = z[2 * i + 1] + by * p[2 * i + 1];
    }
    rz[0] = v[0];
    rz[1] = v[1];
    __syncthreads();  // the next SpMV reads every row's new p
  }
  if (threadIdx.x == 0) {
    ctr[2] = it;
    ctr[1] = 1;
  }
}

// the iteration count of the last solve, for the caller's device buffer
__global__ void iters_out_kernel(const int* ctr, int* out) { *out = ctr[2]; }

__global__ void fill_ints(int* p, int n, int v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Uhat: the solution on free nodes, U on Dirichlet nodes, 0 elsewhere
__global__ void out_kernel(int64_t n, const unsigned char* __restrict__ kind,
                           const double* __restrict__ x, const double* __restrict__ U,
                           double* __restrict__ Uhat) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int k = kind[i];
  Uhat[2 * i] = k == 1 ? x[2 * i] : (k == 2 ? U[2 * i] : 0.0);
  Uhat[2 * i + 1] = k == 1 ? x[2 * i + 1] : (k == 2 ? U[2 * i + 1] : 0.0);
}

int fail(int code, const std::string& m) { return aldic::set_error(code, m); }

template <class T>
int dev_upload(T** d, const std::vector<T>& h) {
  *d = nullptr;
  HIP_TRY(hipMalloc(d, sizeof(T) * std::max<size_t>(h.size(), 1)));
  if (!h.empty()) HIP_TRY(hipMemcpy(*d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
  return ALDIC_OK;
}

template <class T>
int dev_alloc(T** d, size_t count) {
  *d = nullptr;
  HIP_TRY(hipMalloc(d, sizeof(T) * std::max<size_t>(count, 1)));
  return ALDIC_OK;
}

void fem_free(aldic_fem* f) {
  void* ptrs[] = {f->xy, f->elem, f->rowptr, f->col, f->val, f->slot_ptr, f->slot_src,
                  f->node_ptr, f->node_src, f->kind, f->neumann, f->Ae, f->be, f->b, f->x,
                  f->r, f->z, f->p, f->q, f->dinv, f->part, f->sc, f->ctr};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (f->cg_exec) (void)hipGraphExecDestroy(f->cg_exec);
  f->cg_exec = nullptr;
}

constexpr int CG_CHUNK = 16;  // even: a chunk starts on an even iteration

void launch_cg_chunk(aldic_fem* f, hipStream_t s, double tol2, int max_iter) {
  const int64_t n = f->n;
  for (int k = 0; k < CG_CHUNK; k++) {
    const int par = k & 1;
    cg_spmv_kernel<<<f->nblk, FEM_T, 0, s>>>(n, f->rowptr, f->col, f->val, f->kind, f->p, f->q,
                                             f->part, f->nblk, f->ctr, f->sc);
    cg_update_kernel<<<f->nblk, FEM_T, 0, s>>>(n, f->kind, f->x, f->r, f->z, f->p, f->q, f->dinv,
                                               f->part, f->nblk, f->ctr, f->sc, par, tol2,
                                               max_iter);
    cg_dir_kernel<<<f->nblk, FEM_T, 0, s>>>(n, f->kind, f->z, f->p, f->sc, f->ctr, par);
  }
}

void launch_cg_block(aldic_fem* f, hipStream_t s, double tol2, int max_iter) {
  cg_block_kernel<<<1, CG_BT, 0, s>>>(f->n, f->rowptr, f->col, f->val, f->kind, f->x, f->r, f->z,
                                       f->p, f->q, f->dinv, f->ctr, f->sc, tol2, max_iter);
}

// The chunk as a graph: recorded on the caller's stream the first time (or
// when the iteration limit changes), replayed after.  Returns false when it
// cannot be recorded (the stream is already being captured, say); the
// caller then launches the chunk directly.
bool cg_chunk_graph(aldic_fem* f, hipStream_t s, double tol2, int max_iter) {
  if (f->cg_exec && f->cg_exec_maxit == max_iter) return true;
  if (f->cg_exec) {
    (void)hipGraphExecDestroy(f->cg_exec);
    f->cg_exec = nullptr;
  }
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &st) != hipSuccess || st != hipStreamCaptureStatusNone) return false;
  if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) return false;
  launch_cg_chunk(f, s, tol2, max_iter);
  hipGraph_t g = nullptr;
  if (hipStreamEndCapture(s, &g) != hipSuccess || !g) {
    (void)hipGetLastError();
    return false;
  }
  hipGraphExec_t ex = nullptr;
  const bool ok = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0) == hipSuccess;
  (void)hipGraphDestroy(g);
  if (!ok) {
    (void)hipGetLastError();
    return false;
  }
  f->cg_exec = ex;
  f->cg_exec_maxit = max_iter;
  return true;
}

inline int blocks(int64_t n) { return (int)std::max<int64_t>(1, (n + FEM_T - 1) / FEM_T); }

}  // namespace

extern "C" int aldic_fem_create(aldic_ctx* ctx, const double* coords, int64_t n,
                                const int32_t* elements, int64_t ne, const int32_t* dirichlet,
                                int64_t nd, const double* neumann, int64_t nn, int32_t gauss_order,
                                aldic_fem** out) {
  if (!ctx || !out || n < 1 || ne < 1 || !coords || !elements || nd < 0 || nn < 0 ||
      (nd > 0 && !dirichlet) || (nn > 0 && !neumann))
    return fail(ALDIC_EINVAL, "fem_create: bad argument");
  if (gauss_order < 2 || gauss_order > 5)
    return fail(ALDIC_EINVAL, "fem_create: GaussPtOrder must be 2, 3, 4 or 5");
  if (n > 0x3fffffff || ne * 64 > 0x7fffffff)
    return fail(ALDIC_EINVAL, "fem_create: mesh too large");
  *out = nullptr;
  // ---- host: connectivity ----
  std::vector<int> el(ne * 8);
  std::vector<char> seen(n + 1, 0);  // index 0 stands for "no node" (0 in elementsFEM)
  for (int64_t e = 0; e < ne; e++)
    for (int a = 0; a < 8; a++) {
      const int32_t v = elements[e + a * ne];  // ne x 8, column-major, 1-based
      if (v < 0 || v > n || (a < 4 && v == 0))
        return fail(ALDIC_EINVAL, "fem_create: element node index out of range");
      el[e * 8 + a] = v - 1;
      seen[v] = 1;
    }
  // coordsIndexInvolved = unique(elementsFEM(:,1:8)) without its first entry
  // (the 0 of a missing midpoint; the smallest node when there is none),
  // Subpb2Quadtree.m:314-318
  std::vector<unsigned char> kind(n, 0);
  bool dropped = false;
  for (int64_t v = 0; v <= n; v++) {
    if (!seen[v]) continue;
    if (!dropped) {
      dropped = true;
      continue;
    }
    kind[v - 1] = 1;
  }
  for (int64_t k = 0; k < nd; k++) {
    const int32_t v = dirichlet[k];
    if (v < 1 || v > n) return fail(ALDIC_EINVAL, "fem_create: Dirichlet node out of range");
    kind[v - 1] = 2;
  }
  for (int64_t k = 0; k < nn; k++) {
    const double a = neumann[k], b = neumann[nn + k];
    if (!(a >= 1 && a <= n && b >= 1 && b <= n))
      return fail(ALDIC_EINVAL, "fem_create: Neumann node out of range");
  }
  // scalar sparsity: pairs of present nodes of each element (a missing
  // midpoint has N = 0 and contributes nothing)
  std::vector<int64_t> keys;
  keys.reserve(ne * 64);
  for (int64_t e = 0; e < ne; e++)
    for (int a = 0; a < 8; a++) {
      const int i = el[e * 8 + a];
      if (i < 0) continue;
      for (int c = 0; c < 8; c++) {
        const int j = el[e * 8 + c];
        if (j < 0) continue;
        keys.push_back((int64_t)i * n + j);
      }
    }
  std::vector<int64_t> uk(keys);
  std::sort(uk.begin(), uk.end());
  uk.erase(std::unique(uk.begin(), uk.end()), uk.end());
  const int64_t nnz = (int64_t)uk.size();
  std::vector<int> rowptr(n + 1, 0), col(nnz);
  for (int64_t s = 0; s < nnz; s++) {
    rowptr[uk[s] / n + 1]++;
    col[s] = (int)(uk[s] % n);
  }
  for (int64_t i = 0; i < n; i++) rowptr[i + 1] += rowptr[i];
  // contributions per slot in element order
  std::vector<int> slot_cnt(nnz + 1, 0), slot_src(keys.size());
  std::vector<int> node_cnt(n + 1, 0), node_src;
  std::vector<int64_t> kpos(keys.size());
  for (size_t k = 0; k < keys.size(); k++) {
    kpos[k] = std::lower_bound(uk.begin(), uk.end(), keys[k]) - uk.begin();
    slot_cnt[kpos[k] + 1]++;
  }
  for (int64_t s = 0; s < nnz; s++) slot_cnt[s + 1] += slot_cnt[s];
  {
    std::vector<int> fillp(slot_cnt.begin(), slot_cnt.end() - 1);
    size_t k = 0;
    for (int64_t e = 0; e < ne; e++)
      for (int a = 0; a < 8; a++) {
        if (el[e * 8 + a] < 0) continue;
        for (int c = 0; c < 8; c++) {
          if (el[e * 8 + c] < 0) continue;
          slot_src[fillp[kpos[k]]++] = (int)(e * 64 + a * 8 + c);
          k++;
        }
      }
  }
  for (int64_t e = 0; e < ne; e++)
    for (int a = 0; a < 8; a++)
      if (el[e * 8 + a] >= 0) node_cnt[el[e * 8 + a] + 1]++;
  for (int64_t i = 0; i < n; i++) node_cnt[i + 1] += node_cnt[i];
  node_src.resize(node_cnt[n]);
  {
    std::vector<int> fillp(node_cnt.begin(), node_cnt.end() - 1);
    for (int64_t e = 0; e < ne; e++)
      for (int a = 0; a < 8; a++)
        if (el[e * 8 + a] >= 0) node_src[fillp[el[e * 8 + a]]++] = (int)(e * 8 + a);
  }
  std::vector<double> xy(2 * n);
  for (int64_t i = 0; i < n; i++) {
    xy[i] = coords[i];
    xy[n + i] = coords[n + i];
  }
  // ---- device ----
  auto* f = new aldic_fem();
  f->ctx = ctx;
  f->device = ctx->device;
  f->n = n;
  f->ne = ne;
  f->gauss = gauss_order;
  f->nnz = nnz;
  f->nn = nn;
  f->h_neumann.assign(neumann ? neumann : (const double*)nullptr,
                      neumann ? neumann + 4 * nn : (const double*)nullptr);
  for (int64_t i = 0; i < n; i++) {
    f->nfree += kind[i] == 1;
    f->ndir += kind[i] == 2;
  }
  f->nblk = blocks(n);
  int rc = ALDIC_OK;
  auto step = [&](int r) {
    if (rc == ALDIC_OK) rc = r;
  };
  if (hipSetDevice(ctx->device) != hipSuccess) step(fail(ALDIC_EDEVICE, "fem_create: hipSetDevice"));
  step(dev_upload(&f->xy, xy));
  step(dev_upload(&f->elem, el));
  step(dev_upload(&f->rowptr, rowptr));
  step(dev_upload(&f->col, col));
  step(dev_alloc(&f->val, nnz));
  step(dev_upload(&f->slot_ptr, slot_cnt));
  step(dev_upload(&f->slot_src, slot_src));
  step(dev_upload(&f->node_ptr, node_cnt));
  step(dev_upload(&f->node_src, node_src));
  step(dev_upload(&f->kind, kind));
  step(dev_upload(&f->neumann, f->h_neumann));
  step(dev_alloc(&f->Ae, (size_t)ne * 64));
  step(dev_alloc(&f->be, (size_t)ne * 16));
  step(dev_alloc(&f->b, 2 * n));
  step(dev_alloc(&f->x, 2 * n));
  step(dev_alloc(&f->r, 2 * n));
  step(dev_alloc(&f->z, 2 * n));
  step(dev_alloc(&f->p, 2 * n));
  step(dev_alloc(&f->q, 2 * n));
  step(dev_alloc(&f->dinv, n));
  step(dev_alloc(&f->part, (size_t)4 * f->nblk));
  step(dev_alloc(&f->sc, 16));
  step(dev_alloc(&f->ctr, 4));
  if (rc == ALDIC_OK) {
    fill_ints<<<1, 64, 0, ctx->stream>>>(f->ctr, 4, 0);
    if (hipStreamSynchronize(ctx->stream) != hipSuccess)
      rc = fail(ALDIC_EDEVICE, "fem_create: initialising counters failed");
  }
  if (rc != ALDIC_OK) {
    fem_free(f);
    delete f;
    return rc;
  }
  *out = f;
  return ALDIC_OK;
}

extern "C" int aldic_fem_destroy(aldic_fem* f) {
  if (!f) return ALDIC_OK;
  (void)hipSetDevice(f->device);
  (void)hipDeviceSynchronize();
  fem_free(f);
  delete f;
  return ALDIC_OK;
}

extern "C" int aldic_fem_info(const aldic_fem* f, int64_t* nnz, int64_t* nfree, int64_t* ndir) {
  if (!f) return fail(ALDIC_EINVAL, "fem_info: NULL mesh");
  if (nnz) *nnz = f->nnz;
  if (nfree) *nfree = f->nfree;
  if (ndir) *ndir = f->ndir;
  return ALDIC_OK;
}

namespace {

// Assembly, right-hand side, Jacobi inverse and the CG start on stream s.
void subpb2_setup(aldic_fem* f, double beta, double mu, double alpha, double winstepsize,
                  const double* d_U, const double* d_F, hipStream_t s) {
  const int64_t n = f->n;
  element_kernel<<<blocks(f->ne), FEM_T, 0, s>>>(f->ne, n, f->elem, f->xy, f->gauss, beta, mu,
                                                 alpha, d_U, d_F, f->Ae, f->be);
  gather_kernel<<<blocks(f->nnz), FEM_T, 0, s>>>(f->nnz, n, f->slot_ptr, f->slot_src, f->Ae,
                                                 f->val);
  rhs_kernel<<<blocks(n), FEM_T, 0, s>>>(n, f->node_ptr, f->node_src, f->be, f->b);
  if (f->nn > 0)
    neumann_kernel<<<1, 1, 0, s>>>(f->nn, f->neumann, f->xy, n, d_F, winstepsize, f->b);
  setup_kernel<<<f->nblk, FEM_T, 0, s>>>(n, f->rowptr, f->col, f->val, f->kind, d_U, f->b, f->x,
                                         f->r, f->dinv);
  fill_ints<<<1, 64, 0, s>>>(f->ctr, 4, 0);
  cg_norm_kernel<<<f->nblk, FEM_T, 0, s>>>(n, f->rowptr, f->col, f->val, f->kind, d_U, f->b,
                                           f->part, f->nblk, f->ctr, f->sc);
  cg_init_kernel<<<f->nblk, FEM_T, 0, s>>>(n, f->kind, f->r, f->dinv, f->z, f->p, f->b, d_U,
                                           f->part, f->nblk, f->ctr, f->sc);
}

int subpb2_check(aldic_ctx* ctx, aldic_fem* f, double winstepsize, const double* d_U,
                 const double* d_F, double* d_Uhat, const char* who) {
  if (!ctx || !f || !d_U || !d_F || !d_Uhat)
    return fail(ALDIC_EINVAL, std::string(who) + ": NULL argument");
  if (f->nn > 0 && !(winstepsize != 0.0))
    return fail(ALDIC_EINVAL, std::string(who) + ": winstepsize must be nonzero with Neumann edges");
  return ALDIC_OK;
}

constexpr double CG_TOL2 = 1e-26;  // (1e-13)^2, relative, per component

}  // namespace

extern "C" int aldic_subpb2_dev(aldic_ctx* ctx, aldic_fem* f, double beta, double mu, double alpha,
                                double winstepsize, const double* d_U, const double* d_F,
                                double* d_Uhat, int32_t max_iter, int32_t* iterations) {
  RC_TRY(subpb2_check(ctx, f, winstepsize, d_U, d_F, d_Uhat, "subpb2_dev"));
  if (max_iter <= 0) max_iter = (int32_t)std::min<int64_t>(std::max<int64_t>(2 * f->n, 100), 1 << 20);
  HIP_TRY(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  StageTimer t(ctx, ALDIC_STAGE_SUBPB2);
  subpb2_setup(f, beta, mu, alpha, winstepsize, d_U, d_F, s);
  const bool small = f->n <= CG_SMALL_N;
  if (small) launch_cg_block(f, s, CG_TOL2, max_iter);
  const bool graph = !small && cg_chunk_graph(f, s, CG_TOL2, max_iter);
  int done = 0, iters = 0;
  while (!done) {
    if (graph)
      HIP_TRY(hipGraphLaunch(f->cg_exec, s));
    else if (!small)
      launch_cg_chunk(f, s, CG_TOL2, max_iter);
    int h[4];
    HIP_TRY(hipMemcpyAsync(h, f->ctr, sizeof(h), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    done = h[1];
    iters = h[2];
  }
  out_kernel<<<f->nblk, FEM_T, 0, s>>>(f->n, f->kind, f->x, d_U, d_Uhat);
  HIP_TRY(hipGetLastError());
  if (iterations) *iterations = iters;
  return ALDIC_OK;
}

namespace {
// a skipped solve: the CG's done flag is set before its first chunk
__global__ void skip_cg_kernel(const int* skip, int* ctr) {
  if (*skip) ctr[1] = 1;
}
}  // namespace

int64_t aldic::fem_nodes(const aldic_fem* f) { return f ? f->n : -1; }

extern "C" int aldic_subpb2_dev_async(aldic_ctx* ctx, aldic_fem* f, double beta, double mu,
                                      double alpha, double winstepsize, const double* d_U,
                                      const double* d_F, double* d_Uhat, int32_t max_iter,
                                      int32_t* d_iterations) {
  return aldic::subpb2_async(ctx, f, beta, mu, alpha, winstepsize, d_U, d_F, d_Uhat, max_iter,
                             d_iterations, nullptr);
}

int aldic::subpb2_async(aldic_ctx* ctx, aldic_fem* f, double beta, double mu, double alpha,
                        double winstepsize, const double* d_U, const double* d_F,
                        double* d_Uhat, int32_t max_iter, int32_t* d_iterations,
                        const int* d_skip) {
  RC_TRY(subpb2_check(ctx, f, winstepsize, d_U, d_F, d_Uhat, "subpb2_dev_async"));
  if (max_iter <= 0 || max_iter > (1 << 16))
    return fail(ALDIC_EINVAL, "subpb2_dev_async: max_iter must be in 1 .. 65536");
  HIP_TRY(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  HIP_TRY(hipStreamIsCapturing(s, &cs));
  const bool capturing = cs != hipStreamCaptureStatusNone;
  StageTimer t(ctx, ALDIC_STAGE_SUBPB2);
  subpb2_setup(f, beta, mu, alpha, winstepsize, d_U, d_F, s);
  if (d_skip) skip_cg_kernel<<<1, 1, 0, s>>>(d_skip, f->ctr);
  // every chunk the limit may need; those after convergence return at once.
  // Under a caller's capture the chunks are recorded as plain launches (the
  // mesh's own chunk graph is for eager calls).
  if (f->n <= CG_SMALL_N) {
    launch_cg_block(f, s, CG_TOL2, max_iter);
  } else {
    const bool graph = !capturing && cg_chunk_graph(f, s, CG_TOL2, max_iter);
    for (int done = 0; done < max_iter; done += CG_CHUNK) {
      if (graph)
        HIP_TRY(hipGraphLaunch(f->cg_exec, s));
      else
        launch_cg_chunk(f, s, CG_TOL2, max_iter);
    }
  }
  out_kernel<<<f->nblk, FEM_T, 0, s>>>(f->n, f->kind, f->x, d_U, d_Uhat);
  if (d_iterations) iters_out_kernel<<<1, 1, 0, s>>>(f->ctr, d_iterations);
  HIP_TRY(hipGetLastError());
  return ALDIC_OK;
}

extern "C" int aldic_subpb2(aldic_ctx* ctx, aldic_fem* f, double beta, double mu, double alpha,
                            double winstepsize, const double* U, const double* F, double* Uhat,
                            int32_t* iterations) {
  if (!ctx || !f || !U || !F || !Uhat) return fail(ALDIC_EINVAL, "subpb2: NULL argument");
  HIP_TRY(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  const int64_t n = f->n;
  double *dU = nullptr, *dF = nullptr, *dO = nullptr;
  int rc = ALDIC_OK;
  if (hipMalloc(&dU, sizeof(double) * 2 * n) != hipSuccess ||
      hipMalloc(&dF, sizeof(double) * 4 * n) != hipSuccess ||
      hipMalloc(&dO, sizeof(double) * 2 * n) != hipSuccess) {
    rc = fail(ALDIC_ENOMEM, "subpb2: device allocation failed");
  }
  if (rc == ALDIC_OK &&
      (hipMemcpyAsync(dU, U, sizeof(double) * 2 * n, hipMemcpyHostToDevice, s) != hipSuccess ||
       hipMemcpyAsync(dF, F, sizeof(double) * 4 * n, hipMemcpyHostToDevice, s) != hipSuccess))
    rc = fail(ALDIC_EDEVICE, "subpb2: upload failed");
  if (rc == ALDIC_OK) rc = aldic_subpb2_dev(ctx, f, beta, mu, alpha, winstepsize, dU, dF, dO, 0,
                                            iterations);
  if (rc == ALDIC_OK &&
      (hipMemcpyAsync(Uhat, dO, sizeof(double) * 2 * n, hipMemcpyDeviceToHost, s) != hipSuccess ||
       hipStreamSynchronize(s) != hipSuccess))
    rc = fail(ALDIC_EDEVICE, "subpb2: download failed");
  (void)hipStreamSynchronize(s);
  if (dU) (void)hipFree(dU);
  if (dF) (void)hipFree(dF);
  if (dO) (void)hipFree(dO);
  return rc;
}
