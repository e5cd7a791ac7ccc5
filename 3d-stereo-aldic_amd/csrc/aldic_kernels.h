// aldic_kernels.h -- internal interface between the C-ABI layer (capi.hip) and
// the kernels.  Not part of the public header.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

// Arguments of the per-subset solvers (IC-GN and Subpb1).  Images are M x N
// column-major FP32 planes with the first dimension along image x, all of
// the same shape: fm = ImgRef .* mask, fx/fy = Df.DfDx/Df.DfDy placed on the
// full grid (pixel (i,j) holds DfDx(i-3,j-3), zero on the 3-pixel border),
// g = ImgDef, mask = ImgRefMask.
struct SolveArgs {
  const float* fm;
  const float* fx;
  const float* fy;
  const float* g;       // ImgDef, or its B-spline coefficients when basis is 1
  const float* mask;
  int M, N;
  int basis;            // 0: the reference's local kernel on ImgDef; 1: the
                        // interpolating cubic B-spline on its coefficients
  int n;                // number of nodes (stride of H21)
  const int* list;      // node indices handled by this launch
  int cnt;              // length of list
  const double* coords; // n x 2 column-major (coordinatesFEM), 1-based
  const int* wsx;       // winsize (IC-GN) or winsize_List(:,1) (Subpb1)
  const int* wsy;       // winsize_List(:,2) for Subpb1, nullptr for IC-GN
  int ws_per_node;      // 0: wsx[0] (and wsy[0]) for every node
  const float* U0;      // IC-GN: 2n interleaved initial guess
  const float* UOld;    // Subpb1: 2n
  const float* FOld;    // Subpb1: 4n
  const float* vdual;   // Subpb1: 2n
  float mu;
  float tol;
  int maxit;            // evaluation cap k <= maxit (0: the reference's 100)
  int stop_rule;        // IC-GN: 0 the reference's relative ||b|| rule, 1 the step-size rule
  float tol_dp;         // step-size rule threshold (px)
  const int* dcnt;      // fast kernels: list length read on the device (cnt = capacity)
  int fixed_side;       // fast kernels: every listed subset has the class's full side
  float* U;             // 2n
  float* F;             // 4n (IC-GN), may be null
  float* H21;           // n x 21 column-major (IC-GN), may be null
  int* conv;            // n
  float* zncc;          // n (IC-GN), may be null
  uint8_t* hole;        // n (IC-GN markCoordHoleStrain flags), may be null
  int* defer_cnt;       // subsets handed from the fast to the generic kernel
  int* defer_list;
  const int* skip;      // may be null: nonzero on the device -> the launch does nothing
};

namespace aldic {
// The persistent launch over the fast size classes: per class c < 4 its
// node list, count and whether every subset has the class's full side;
// queue: two zeroed ints (the block and the wave work counters).
struct PersistArgs {
  SolveArgs a;
  const int* list[4];
  int cnt[4];
  int full[4];
  int* queue;
};
// 256-thread blocks of the persistent launch resident per CU (its launch
// bound): the grid is this many blocks per CU, or fewer when there is less work
constexpr int PERSIST_BLOCKS_PER_CU = 3;
hipError_t launch_persistent(int mode, const PersistArgs& P, int grid, hipStream_t s);
int size_class(int S1, int S2);
int class_side(int cls);  // the largest subset side of a fast-kernel class
hipError_t launch_fast(int mode, int cls, int var, const SolveArgs& a, hipStream_t s);
int num_variants();
// The workgroup kernel over list[0 .. n): n = fixed_count, or the device
// count *count (when count is not null) clamped to the capacity fixed_count.
hipError_t launch_generic(int mode, const SolveArgs& a, const int* list, const int* count,
                          int fixed_count, int grid, hipStream_t s);
int generic_max_pixels();
// Sort the nodes list[0 .. min(*count, max_count)) into per-size-class
// lists (appended at counts[c]; counts must start at zero).  List c holds
// at most cap[c] entries: a node that finds its class list full goes to
// list 4 (the workgroup kernel, any size), and one that finds list 4 full
// is counted in counts[5] and dropped.  Readers clamp counts[c] to cap[c].
struct ClassLists {
  int* list[5];
  int cap[5];
  int* counts;  // 6 ints
};
hipError_t launch_bucket(const SolveArgs& a, const int* list, const int* count, int max_count,
                         ClassLists out, hipStream_t s);
__global__ void zero_ints_kernel(int* p, int n);

// image preparation (prep.hip)
// raw 8/16-bit images: exact sums over a 0-based inclusive ROI (sums[0..1]
// = sum, sum of squares), then (raw - mean) / sd over all n pixels in FP64
hipError_t launch_raw_sums(const void* raw, int bits, int64_t M, int64_t x0, int64_t x1,
                           int64_t y0, int64_t y1, unsigned long long* sums, hipStream_t s);
hipError_t launch_raw_normalize(const void* raw, int bits, int64_t n, double mean, double sd,
                                double* out, hipStream_t s);
// interpolating cubic B-spline coefficients of an M x N FP32 plane (mirror
// boundaries); in and out must not overlap
hipError_t launch_bspline_prefilter(const float* in, float* out, int64_t M, int64_t N,
                                   hipStream_t s);
hipError_t launch_prep_planes(const double* ref, const double* def, const double* mask,
                              int64_t M, int64_t N, float* fm, float* g, float* maskf,
                              hipStream_t s);
hipError_t launch_gradient(const double* ref, const double* mask, int64_t M, int64_t N,
                           float* fx, float* fy, hipStream_t s);
hipError_t launch_place_df(const double* DfDx, const double* DfDy, int64_t M, int64_t N,
                           float* fx, float* fy, hipStream_t s);
hipError_t launch_gradient_f32(const float* ref, const float* mask, int64_t M, int64_t N,
                               float* fx, float* fy, hipStream_t s);
hipError_t launch_gradient_out_f64(const double* ref, const double* mask, int64_t M, int64_t N,
                                   double* DfDx, double* DfDy, hipStream_t s);

// integer search (search.hip); scratch of search_scratch_bytes(points, ...)
size_t search_smem_bytes(int ws);
size_t search_scratch_bytes(int maxpts, int ws, const int reg[4]);
// q (may be null): raw (PCE, PPE) per grid point, 2 doubles, point order
// ci + cj * nX (compute_qFactor of the full normxcorr2 map).  method: 0
// direct lag tiles, 1 FFT lag blocks where the template side is <= 65
// (direct otherwise), 2 FFT blocks where they are the faster ones; the
// sub-pixel fit is direct either way.
hipError_t launch_search(const float* f, const float* g, int M, int N, int nX, int nY,
                         const int gridx[2], const int gridy[2], int ws, int step,
                         const int reg[4], void* scratch, double* x, double* y, double* u,
                         double* v, double* ccmax, int* ipr, int* ipc, hipStream_t s,
                         double* q = nullptr, int method = 0);
hipError_t launch_search_nodes(const float* f, const float* g, int M, int N,
                               const double* coords, int n, const int* list, const int* count,
                               int max_count, int ws, const int reg[4], void* scratch, float* U0,
                               float* F0, hipStream_t s, int method = 0);

// grid search over a device list of grid points (indices ci + cj*nX)
hipError_t launch_search_points(const float* f, const float* g, int M, int N, int nX, int nY,
                                const int gridx[2], const int gridy[2], int ws, int step,
                                const int reg[4], const int* points, const int* count,
                                void* scratch, double* x, double* y, double* u, double* v,
                                double* ccmax, int* ipr, int* ipc, hipStream_t s,
                                int method = 0);
// a warp field on a regular grid (point ci + cj*nX; U 2 and F 4 floats per
// point) -> the bilinear interpolant at every node (F0 may be null)
hipError_t launch_field_to_nodes(const float* Ug, const float* Fg, int nX, int nY, double x0,
                                 double y0, double step, const double* coords, int n, float* U0,
                                 float* F0, hipStream_t s);
hipError_t launch_grid_to_listed_nodes(const double* gu, const double* gv, int nX, int nY,
                                       double x0, double y0, double step, const double* coords,
                                       int n, const int* list, const int* count, float* U0,
                                       float* F0, hipStream_t s);
hipError_t launch_mark_cells(int nX, int nY, double x0, double y0, double step,
                             const double* coords, int n, const int* list, const int* count,
                             int* flags, int* points, int* npoints, hipStream_t s);

// search grid -> node initial guess (prep.hip)
hipError_t launch_grid_to_nodes(const double* gu, const double* gv, int nX, int nY, double x0,
                                double y0, double step, const double* coords, int n, float* U0,
                                hipStream_t s);

// interpolation (interp.hip)
hipError_t launch_interp_f64(const double* F, int64_t M, int64_t N, int64_t P, const double* X,
                             const double* Y, int64_t ND, double* Z, hipStream_t s);
hipError_t launch_interp_f32(const float* F, int64_t M, int64_t N, int64_t P, const float* X,
                             const float* Y, int64_t ND, float* Z, hipStream_t s);

// learned initial guess (cnn.hip); activations and weights are bf16 (void*)
hipError_t cnn_gemm(const void* A, int lda, const void* Bt, int ldb, int M, int N, int K,
                    int splits, float* C, int ldc, const float* bias, int relu, void* Cb, int ldcb,
                    hipStream_t s);
hipError_t cnn_conv(const void* x, int B, int H, int W, int Cin, int stride, int Ho, int Wo,
                    const void* Wt, int Kp, int Cout, const float* bias, int relu, void* y,
                    float* yf, hipStream_t s);
hipError_t cnn_dense(const void* x, int B, int K, const void* Wt, int Cout, const float* bias,
                     int relu, void* y, hipStream_t s);
hipError_t cnn_splitk_reduce(const float* P, int splits, int M, int N, int ldp, const float* bias,
                             int relu, float* C, int ldc, void* Cb, int ldcb, hipStream_t s);
hipError_t cnn_im2col(const void* x, int B, int H, int W, int Cin, int stride, int Ho, int Wo,
                      int Kp, void* col, hipStream_t s);
hipError_t cnn_col2im(const float* dcol, int B, int H, int W, int Cin, int stride, int Ho, int Wo,
                      int Kp, float* dx, hipStream_t s);
hipError_t cnn_transpose_bf16(const void* in, int R, int Cc, int ldi, void* out, int ldo,
                              hipStream_t s);
hipError_t cnn_transpose_f32(const float* in, int R, int Cc, int ldi, void* out, int ldo,
                             hipStream_t s);
hipError_t cnn_relu_back(const float* g, const void* y, int64_t n, void* gb, float* gf,
                         hipStream_t s);
hipError_t cnn_colsum(const float* g, int M, int N, int ldg, float* out, hipStream_t s);
hipError_t cnn_to_bf16(const float* in, int64_t n, void* out, hipStream_t s);
hipError_t cnn_adam(float* w, const float* g, float* m, float* v, int64_t n, float lr, float b1,
                    float b2, float eps, float c1, float c2, float wd, const float* mask,
                    void* wb, hipStream_t s);
hipError_t cnn_mse(const float* pred, int ldp, const float* lab, const float* scale, int B,
                   float* d, int ldd, float* loss, hipStream_t s);
hipError_t cnn_patches(const float* f, const float* g, int M, int N, const double* coords, int n,
                       int first, int count, const float* prior, void* out, hipStream_t s);
hipError_t cnn_front(const float* f, const float* g, int M, int N, const double* coords, int n,
                     int first, int count, const float* prior, const void* W1, const float* b1,
                     const void* W2, const float* b2, void* out, hipStream_t s);
hipError_t cnn_predict(const float* out, int ldo, int first, int count, const float* prior,
                       float sf, float su, float* U0, float* F0, hipStream_t s);
hipError_t cnn_mosaic(int M, int N, int tile, const float* warps, int ntx, uint32_t seed,
                      uint32_t noise_seed, float density, float radius, float scale, float noise,
                      float top, float* img, hipStream_t s);
}  // namespace aldic
