/*
 * aldic_oracle.c -- FP64 CPU restatement of ALDIC's local matching path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library and the timed
 * CPU baseline of bench.py.  The product path never calls it.
 *
 * Each function cites the reference lines it restates.  What pins each one
 * (the reference ships no golden vectors for this path, and its MEX needs
 * MATLAB's mex.h, so it cannot be built here):
 *   orc_interp_spline_*   closed form A = rot90(BC*Q*BC',2) with BC from
 *                         func/BicubicBsplineBuild.m:9-16, recomputed
 *                         independently in tests; integer nodes return Q11;
 *                         partition of unity; ramp error t(2t-1)(t-1)/14;
 *                         clamp-to-edge against an explicit padded image.
 *   orc_img_gradient      exact on polynomials of degree <= 6.
 *   orc_normxcorr2_crop   brute-force Pearson correlation per lag; affine
 *                         invariance; exact integer shifts.  The zero-
 *                         variance tolerance of MATLAB's builtin is
 *                         parity unpinned (normxcorr2 is not vendored).
 *   orc_findpeak          exact on sampled quadratics; edge and |offset|>1
 *                         rules on constructed inputs.
 *   orc_integer_search    integer shifts of speckle recover the shift; grid
 *                         layout and clamps on constructed cases.
 *   orc_icgn_subset       zero motion converges at step 1; integer
 *                         translation is recovered exactly; affine warps of
 *                         an analytic image; H symmetric PD; status codes
 *                         102/103/101/-1 forced by construction; intensity
 *                         affine invariance.
 *   orc_subpb1_subset     mu -> 0 reproduces the translation-only IC-GN;
 *                         large mu pins U to UOld + vdual.
 *   orc_bspline_prefilter against scipy.ndimage.spline_filter (order 3,
 *                         mirror); (c[k-1] + 4c[k] + c[k+1])/6 = s[k]; the
 *                         zero rule by brute force on an image with a hole.
 *   orc_interp_bspline_d  against scipy.ndimage.map_coordinates (order 3,
 *                         mirror) inside the image; reproduces samples at
 *                         nodes and linear ramps exactly.
 *   orc_bwselect          brute force on small random masks (components by
 *                         union-find), 4- vs 8-connectivity cases.  MATLAB's
 *                         bwselect itself is not vendored: parity unpinned
 *                         beyond its documented definition.
 */
#include "aldic_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* Bicubic "B-spline" of ba_interp2_spline.cpp                               */
/* ------------------------------------------------------------------------ */

/* 56 * w_k(t) = sum_p W56[k][p] t^p for the taps k = -1,0,1,2 of the local
 * kernel (rows of BC' in func/BicubicBsplineBuild.m:13-16 times 56). */
static const int W56[4][4] = {
    {0, -33, 57, -24},
    {56, -3, -117, 64},
    {0, 45, 75, -64},
    {0, -9, -15, 24},
};

/* The MEX writes each polynomial coefficient a_ij as a single-precision
 * literal scale times an integer combination of the 16 samples
 * (ba_interp2_spline.cpp:310-325).  GDIV[i][j] is the divisor that turns
 * W56[r][i]*W56[c][j] into those integers; the sign is folded into it.  The
 * scale literal is |GDIV|/3136 in the rounded form the MEX uses, so the
 * double path sees the same float-rounded constants. */
static const int GDIV[4][4] = {
    {3136, -168, 168, 448},
    {-168, 9, -9, 24},
    {168, -9, 9, -24},
    {448, 24, -24, 64},
};

static float scale_literal(int g) {
  int a = g < 0 ? -g : g;
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

static int KCOEF[4][4][4][4]; /* [i][j][r][c] */
static float SCALE[4][4];
static int coef_ready = 0;

static void coef_init(void) {
  if (coef_ready) return;
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      SCALE[i][j] = scale_literal(GDIV[i][j]);
      for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++)
          KCOEF[i][j][r][c] = W56[r][i] * W56[c][j] / GDIV[i][j];
    }
  coef_ready = 1;
}

static inline int64_t clampi(int64_t v, int64_t lo, int64_t hi) {
  return v < lo ? lo : (v > hi ? hi : v);
}

/* ba_interp2_spline.cpp:60-66 access(): clamp column x to [0,N-1] and row y
 * to [0,M-1]; column-major offset. */
static inline int64_t acc(int64_t M, int64_t N, int64_t x, int64_t y) {
  return M * clampi(x, 0, N - 1) + clampi(y, 0, M - 1);
}

/* Gather Q[r][c] = F(row yb-1+r, col xb-1+c), ba_interp2_spline.cpp:68-130
 * (the fast path :109-129 equals the clamped path away from the border). */
static inline void cell_index(int64_t M, int64_t N, double X, double Y, int64_t idx[4][4]) {
  double xf = floor(X), yf = floor(Y);
  /* int(x_floor-1) of the MEX; non-finite or huge values are clamped here
   * (the MEX result is NaN for them anyway through dx, dy). */
  double xb = xf - 1, yb = yf - 1;
  if (!(xb > -1e9)) xb = -1e9;
  if (!(xb < 1e9)) xb = 1e9;
  if (!(yb > -1e9)) yb = -1e9;
  if (!(yb < 1e9)) yb = 1e9;
  int64_t xi = (int64_t)xb, yi = (int64_t)yb;
  for (int r = 0; r < 4; r++)
    for (int c = 0; c < 4; c++) idx[r][c] = acc(M, N, xi - 1 + c, yi - 1 + r);
}

void orc_interp_spline_d(const double* F, int64_t M, int64_t N, int64_t P, const double* X,
                         const double* Y, int64_t ND, double* Z) {
  coef_init();
  const int64_t LO = M * N;
  for (int64_t i = 0; i < ND; i++) {
    const double x = X[i], y = Y[i];
    /* "Zach revised" swap, ba_interp2_spline.cpp:246-248 */
    const double dx = y - floor(y);
    const double dy = x - floor(x);
    const double dxp[4] = {1.0, dx, dx * dx, dx * dx * dx};
    const double dyp[4] = {1.0, dy, dy * dy, dy * dy * dy};
    int64_t idx[4][4];
    cell_index(M, N, x, y, idx);
    for (int64_t ch = 0; ch < P; ch++) {
      double Q[4][4];
      for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) Q[r][c] = F[idx[r][c] + ch * LO];
      double out = 0.0;
      int first = 1;
      for (int a = 0; a < 4; a++)
        for (int b = 0; b < 4; b++) {
          double coef;
          if (a == 0 && b == 0) {
            coef = Q[1][1];
          } else {
            double s = 0.0;
            for (int r = 0; r < 4; r++)
              for (int c = 0; c < 4; c++) s += (double)KCOEF[a][b][r][c] * Q[r][c];
            coef = (double)SCALE[a][b] * s;
          }
          double term = coef * dxp[a] * dyp[b];
          if (first) {
            out = term;
            first = 0;
          } else {
            out += term;
          }
        }
      Z[i + ch * ND] = out;
    }
  }
}

void orc_interp_spline_f(const float* F, int64_t M, int64_t N, int64_t P, const float* X,
                         const float* Y, int64_t ND, float* Z) {
  coef_init();
  const int64_t LO = M * N;
  for (int64_t i = 0; i < ND; i++) {
    const float x = X[i], y = Y[i];
    const float dx = y - floorf(y);
    const float dy = x - floorf(x);
    const float dxx = dx * dx, dyy = dy * dy;
    const float dxp[4] = {1.0f, dx, dxx, dxx * dx};
    const float dyp[4] = {1.0f, dy, dyy, dyy * dy};
    int64_t idx[4][4];
    cell_index(M, N, (double)x, (double)y, idx);
    for (int64_t ch = 0; ch < P; ch++) {
      float Q[4][4];
      for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) Q[r][c] = F[idx[r][c] + ch * LO];
      float out = 0.0f;
      int first = 1;
      for (int a = 0; a < 4; a++)
        for (int b = 0; b < 4; b++) {
          float coef;
          if (a == 0 && b == 0) {
            coef = Q[1][1];
          } else {
            float s = 0.0f;
            for (int r = 0; r < 4; r++)
              for (int c = 0; c < 4; c++) s += (float)KCOEF[a][b][r][c] * Q[r][c];
            coef = SCALE[a][b] * s;
          }
          float term = coef * dxp[a] * dyp[b];
          if (first) {
            out = term;
            first = 0;
          } else {
            out += term;
          }
        }
      Z[i + ch * ND] = out;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Optional interpolant: the interpolating cubic B-spline (global prefilter)  */
/* ------------------------------------------------------------------------ */

/* Not the reference's interpolant: its MEX evaluates the local kernel above
 * (ba_interp2_spline.cpp:310-325), whose 4-point collocation inverse
 * approximates the B-spline coefficients cell by cell.  This is the exact
 * interpolating cubic B-spline: coefficients c with (c[k-1] + 4 c[k] +
 * c[k+1]) / 6 = s[k] on the whole-sample mirror extension of each line
 * (s[-k] = s[k], s[n-1+k] = s[n-1-k]), by the causal / anticausal recursion
 * with pole z = sqrt(3) - 2 and gain 6 (Unser, Aldroubi, Eden, IEEE Trans.
 * Signal Process. 41(2), 1993; Unser, IEEE Signal Process. Mag. 16(6), 1999),
 * with the exact initial values of that extension: the causal start sums one
 * period 2n-2 and divides by 1 - z^(2n-2). */
static void bspline_line(double* s, int64_t n, int64_t stride) {
  if (n < 2) return;
  const double z = sqrt(3.0) - 2.0;
  const int64_t P = 2 * n - 2;
  double zk = 1.0, c0 = 0.0;
  for (int64_t k = 0; k < P; k++) {
    c0 += zk * s[(k < n ? k : P - k) * stride];
    zk *= z;
  }
  c0 /= 1.0 - zk;
  for (int64_t k = 0; k < n; k++) s[k * stride] *= 6.0;
  s[0] = 6.0 * c0;
  for (int64_t k = 1; k < n; k++) s[k * stride] += z * s[(k - 1) * stride];
  s[(n - 1) * stride] =
      (z / (z * z - 1.0)) * (s[(n - 1) * stride] + z * s[(n - 2) * stride]);
  for (int64_t k = n - 2; k >= 0; k--) s[k * stride] = z * (s[(k + 1) * stride] - s[k * stride]);
}

/* Then every coefficient whose pixel lies in a 4 x 4 window of all-zero
 * samples inside the image is set to zero: the spline then vanishes exactly
 * on the cells whose 16 samples are all zero, where the reference's kernel
 * does (a masked, zeroed region of ImgDef, which funICGNQuadtree.m:166-170
 * detects by g == 0), instead of decaying as 0.268^d into it. */
void orc_bspline_prefilter(const double* in, int64_t M, int64_t N, double* out) {
  double* c = (double*)malloc(sizeof(double) * (size_t)(M * N));
  memcpy(c, in, sizeof(double) * (size_t)(M * N));
  for (int64_t j = 0; j < N; j++) bspline_line(c + j * M, M, 1);
  for (int64_t i = 0; i < M; i++) bspline_line(c + i, N, M);
  for (int64_t j = 0; j < N; j++)
    for (int64_t i = 0; i < M; i++) {
      int kill = 0;
      for (int64_t b = j - 3; b <= j && !kill; b++)
        for (int64_t a = i - 3; a <= i && !kill; a++) {
          if (a < 0 || b < 0 || a + 3 >= M || b + 3 >= N) continue;
          int all0 = 1;
          for (int64_t q = 0; q < 4 && all0; q++)
            for (int64_t p = 0; p < 4 && all0; p++) all0 = in[(a + p) + (b + q) * M] == 0.0;
          kill = all0;
        }
      out[i + j * M] = kill ? 0.0 : c[i + j * M];
    }
  free(c);
}

/* The spline at (X, Y) (X along dim 2, Y along dim 1, 1-based) from its
 * coefficients: the same 4 x 4 cell and edge clamp as the MEX, weighted by
 * the cubic B-spline basis (1/6)[(1-t)^3, 3t^3-6t^2+4, -3t^3+3t^2+3t+1, t^3]
 * at the fractional parts. */
static inline void bspline_weights(double t, double w[4]) {
  const double u = 1.0 - t;
  w[0] = u * u * u / 6.0;
  w[1] = (3.0 * t * t * t - 6.0 * t * t + 4.0) / 6.0;
  w[2] = (-3.0 * t * t * t + 3.0 * t * t + 3.0 * t + 1.0) / 6.0;
  w[3] = t * t * t / 6.0;
}

void orc_interp_bspline_d(const double* C, int64_t M, int64_t N, const double* X, const double* Y,
                          int64_t ND, double* Z) {
  for (int64_t i = 0; i < ND; i++) {
    double wr[4], wc[4];
    bspline_weights(Y[i] - floor(Y[i]), wr);
    bspline_weights(X[i] - floor(X[i]), wc);
    int64_t idx[4][4];
    cell_index(M, N, X[i], Y[i], idx);
    double out = 0.0;
    for (int c = 0; c < 4; c++) {
      double col = 0.0;
      for (int r = 0; r < 4; r++) col += wr[r] * C[idx[r][c]];
      out += wc[c] * col;
    }
    Z[i] = out;
  }
}

/* g at the warped points: the reference's kernel on ImgDef, or the spline
 * on ImgDefCoef when the images carry it. */
static void interp_def(const orc_images* im, const double* X, const double* Y, int64_t n,
                       double* g) {
  if (im->ImgDefCoef)
    orc_interp_bspline_d(im->ImgDefCoef, im->M, im->N, X, Y, n, g);
  else
    orc_interp_spline_d(im->ImgDef, im->M, im->N, 1, X, Y, n, g);
}


/* ------------------------------------------------------------------------ */
/* funImgGradient.m:49-78 and funNormalizeImg.m:36-40                        */
/* ------------------------------------------------------------------------ */

void orc_img_gradient(const double* img, const double* mask, int64_t M, int64_t N,
                      double* DfDx, double* DfDy) {
  /* imfilter correlation with [-1/60 3/20 -3/4 0 3/4 -3/20 1/60], centre
   * tap 4; the crop of 3 removes every zero-padded output. */
  static const double h[7] = {-1.0 / 60, 3.0 / 20, -3.0 / 4, 0.0, 3.0 / 4, -3.0 / 20, 1.0 / 60};
  const int64_t Mc = M - 6, Nc = N - 6;
  for (int64_t b = 0; b < Nc; b++)
    for (int64_t a = 0; a < Mc; a++) {
      const int64_t i = a + 3, j = b + 3; /* 0-based pixel */
      double gx = 0.0, gy = 0.0;
      for (int k = 0; k < 7; k++) {
        gx += h[k] * img[(i + k - 3) + j * M];
        gy += h[k] * img[i + (j + k - 3) * M];
      }
      const double m = mask ? mask[i + j * M] : 1.0;
      DfDx[a + b * Mc] = gx * m;
      DfDy[a + b * Mc] = gy * m;
    }
}

void orc_normalize(const double* img, int64_t M, int64_t N, int64_t gx0, int64_t gx1,
                   int64_t gy0, int64_t gy1, double* out) {
  if (gx0 < 1) gx0 = 1;
  if (gx1 > M) gx1 = M;
  if (gy0 < 1) gy0 = 1;
  if (gy1 > N) gy1 = N;
  double s = 0.0;
  int64_t n = 0;
  for (int64_t j = gy0 - 1; j < gy1; j++)
    for (int64_t i = gx0 - 1; i < gx1; i++) {
      s += img[i + j * M];
      n++;
    }
  const double mean = s / (double)n;
  double ss = 0.0;
  for (int64_t j = gy0 - 1; j < gy1; j++)
    for (int64_t i = gx0 - 1; i < gx1; i++) {
      const double d = img[i + j * M] - mean;
      ss += d * d;
    }
  const double sd = sqrt(ss / (double)(n - 1));
  for (int64_t k = 0; k < M * N; k++) out[k] = (img[k] - mean) / sd;
}

/* ------------------------------------------------------------------------ */
/* bwselect (Image Processing Toolbox, not vendored): connected component   */
/* of the foreground containing the seed.                                    */
/* ------------------------------------------------------------------------ */

void orc_bwselect(const uint8_t* BW, int64_t m1, int64_t m2, int64_t r, int64_t c, int conn,
                  uint8_t* out) {
  const int64_t n = m1 * m2;
  memset(out, 0, (size_t)n);
  if (r < 1 || r > m1 || c < 1 || c > m2) return;
  const int64_t s = (r - 1) + (c - 1) * m1;
  if (!BW[s]) return;
  int64_t* stack = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
  int64_t top = 0;
  stack[top++] = s;
  out[s] = 1;
  while (top > 0) {
    const int64_t p = stack[--top];
    const int64_t pr = p % m1, pc = p / m1;
    for (int dc = -1; dc <= 1; dc++)
      for (int dr = -1; dr <= 1; dr++) {
        if (dr == 0 && dc == 0) continue;
        if (conn == 4 && dr != 0 && dc != 0) continue;
        const int64_t qr = pr + dr, qc = pc + dc;
        if (qr < 0 || qr >= m1 || qc < 0 || qc >= m2) continue;
        const int64_t q = qr + qc * m1;
        if (BW[q] && !out[q]) {
          out[q] = 1;
          stack[top++] = q;
        }
      }
  }
  free(stack);
}

/* ------------------------------------------------------------------------ */
/* Small dense helpers                                                       */
/* ------------------------------------------------------------------------ */

/* x = A \ b for a symmetric n x n system (n <= 6): Cholesky when A is
 * positive definite, else LU with partial pivoting (the solver MATLAB's
 * backslash selects for such a matrix). */
static void solve_sym(int n, const double* Ain, const double* b, double* x) {
  double A[36], L[36];
  memcpy(A, Ain, sizeof(double) * n * n);
  int spd = 1;
  memset(L, 0, sizeof(L));
  for (int j = 0; j < n && spd; j++) {
    double d = A[j * n + j];
    for (int k = 0; k < j; k++) d -= L[j * n + k] * L[j * n + k];
    if (!(d > 0.0)) {
      spd = 0;
      break;
    }
    L[j * n + j] = sqrt(d);
    for (int i = j + 1; i < n; i++) {
      double s = A[i * n + j];
      for (int k = 0; k < j; k++) s -= L[i * n + k] * L[j * n + k];
      L[i * n + j] = s / L[j * n + j];
    }
  }
  if (spd) {
    double y[6];
    for (int i = 0; i < n; i++) {
      double s = b[i];
      for (int k = 0; k < i; k++) s -= L[i * n + k] * y[k];
      y[i] = s / L[i * n + i];
    }
    for (int i = n - 1; i >= 0; i--) {
      double s = y[i];
      for (int k = i + 1; k < n; k++) s -= L[k * n + i] * x[k];
      x[i] = s / L[i * n + i];
    }
    return;
  }
  double bb[6];
  memcpy(bb, b, sizeof(double) * n);
  for (int k = 0; k < n; k++) {
    int p = k;
    double mx = fabs(A[k * n + k]);
    for (int i = k + 1; i < n; i++)
      if (fabs(A[i * n + k]) > mx) {
        mx = fabs(A[i * n + k]);
        p = i;
      }
    if (p != k) {
      for (int j = 0; j < n; j++) {
        double t = A[k * n + j];
        A[k * n + j] = A[p * n + j];
        A[p * n + j] = t;
      }
      double t = bb[k];
      bb[k] = bb[p];
      bb[p] = t;
    }
    for (int i = k + 1; i < n; i++) {
      const double f = A[i * n + k] / A[k * n + k];
      for (int j = k; j < n; j++) A[i * n + j] -= f * A[k * n + j];
      bb[i] -= f * bb[k];
    }
  }
  for (int i = n - 1; i >= 0; i--) {
    double s = bb[i];
    for (int j = i + 1; j < n; j++) s -= A[i * n + j] * x[j];
    x[i] = s / A[i * n + i];
  }
}

/* mean and sqrt((n-1)*var) over the entries with |v| > 1e-10
 * (funICGNQuadtree.m:119-120, :272-273). */
static void masked_moments(const double* v, int64_t n, double* mean, double* bottom) {
  double s = 0.0;
  int64_t k = 0;
  for (int64_t i = 0; i < n; i++)
    if (fabs(v[i]) > 1e-10) {
      s += v[i];
      k++;
    }
  if (k == 0) {
    *mean = NAN;
    *bottom = NAN;
    return;
  }
  const double m = s / (double)k;
  double ss = 0.0;
  for (int64_t i = 0; i < n; i++)
    if (fabs(v[i]) > 1e-10) {
      const double d = v[i] - m;
      ss += d * d;
    }
  const double var = (k > 1) ? ss / (double)(k - 1) : 0.0;
  *mean = m;
  *bottom = sqrt((double)(k - 1) * var);
}

/* Inverse-compositional update W(P) o W(dP)^-1, funICGNQuadtree.m:396-414.
 * Returns 0 when det == 0 (the reference breaks out of the loop). */
static int ic_update(double P[6], const double d[6]) {
  const double det = (1 + d[0]) * (1 + d[3]) - d[1] * d[2];
  if (det == 0.0) return 0;
  const double t1 = (-d[0] - d[0] * d[3] + d[1] * d[2]) / det;
  const double t2 = -d[1] / det;
  const double t3 = -d[2] / det;
  const double t4 = (-d[3] - d[0] * d[3] + d[1] * d[2]) / det;
  const double t5 = (-d[4] - d[3] * d[4] + d[2] * d[5]) / det;
  const double t6 = (-d[5] - d[0] * d[5] + d[1] * d[4]) / det;
  const double m11 = 1 + P[0], m12 = P[2], m13 = P[4];
  const double m21 = P[1], m22 = 1 + P[3], m23 = P[5];
  const double n11 = 1 + t1, n12 = t3, n13 = t5;
  const double n21 = t2, n22 = 1 + t4, n23 = t6;
  P[0] = (m11 * n11 + m12 * n21) - 1;
  P[1] = m21 * n11 + m22 * n21;
  P[2] = m11 * n12 + m12 * n22;
  P[3] = (m21 * n12 + m22 * n22) - 1;
  P[4] = m11 * n13 + m12 * n23 + m13;
  P[5] = m21 * n13 + m22 * n23 + m23;
  return 1;
}

/* A rectangular subset window [x0-hx, x0+hx] x [y0-hy, y0+hy] (1-based). */
typedef struct {
  int64_t i0, j0; /* 1-based first row / column */
  int S1, S2;     /* rows (dim 1) and columns (dim 2) */
} win_t;

/* Indexing ImgRef/mask needs rows 1..M; indexing Df.DfDx(x-3) needs rows
 * 4..M-3 (funICGNQuadtree.m:52-56).  Out of range -> MATLAB error. */
static int win_ok(const orc_images* im, const win_t* w) {
  return w->i0 >= 4 && w->i0 + w->S1 - 1 <= im->M - 3 && w->j0 >= 4 &&
         w->j0 + w->S2 - 1 <= im->N - 3;
}
static int win_in_image(const orc_images* im, const win_t* w) {
  return w->i0 >= 1 && w->i0 + w->S1 - 1 <= im->M && w->j0 >= 1 && w->j0 + w->S2 - 1 <= im->N;
}

/* tempfImgMask, tempf = ImgRef .* mask, DfDx, DfDy over the window. */
static void win_extract(const orc_images* im, const win_t* w, double* fmask, double* f,
                        double* fx, double* fy) {
  const int64_t M = im->M, Mc = im->M - 6;
  for (int b = 0; b < w->S2; b++)
    for (int a = 0; a < w->S1; a++) {
      const int64_t i = w->i0 - 1 + a, j = w->j0 - 1 + b;
      const int64_t k = a + (int64_t)b * w->S1;
      const double m = im->mask ? im->mask[i + j * M] : 1.0;
      fmask[k] = m;
      f[k] = im->ImgRef[i + j * M] * m;
      fx[k] = im->DfDx[(i - 3) + (j - 3) * Mc];
      fy[k] = im->DfDy[(i - 3) + (j - 3) * Mc];
    }
}

static int64_t count_zero(const double* v, int64_t n) {
  int64_t c = 0;
  for (int64_t i = 0; i < n; i++) c += (v[i] == 0.0);
  return c;
}
static int any_nan(const double* v, int64_t n) {
  for (int64_t i = 0; i < n; i++)
    if (isnan(v[i])) return 1;
  return 0;
}

/* MATLAB's winsize = 2*max(ceil(sqrt(A/B)*w/2)); returns -1 for a
 * non-finite size (the colon operator then throws). */
static int grow_winsize(double A, double B, int w) {
  const double v = 2.0 * ceil(sqrt(A / B) * (double)w / 2.0);
  if (!isfinite(v) || v > 1e6 || v < 0) return -1;
  return (int)v;
}

/* ------------------------------------------------------------------------ */
/* funICGNQuadtree.m                                                          */
/* ------------------------------------------------------------------------ */

int orc_icgn_subset(const orc_images* im, int64_t x0, int64_t y0, int winsize, double tol,
                    const double U0[2], double U[2], double F[4], double H21[21],
                    double* zncc) {
  return orc_icgn_subset_ex(im, x0, y0, winsize, tol, 100, 0, 0.0, U0, NULL, U, F, H21, zncc);
}

/* The loop of :135 runs while k <= maxIterNum (k counts evaluations, so at
 * most maxIterNum + 1 of them; maxit >= 0 is that cap, the reference's is
 * 100); the status codes 101/102/103 are the reference's (maxIterNum + 1,
 * + 2, + 3 with maxIterNum = 100) whatever cap is passed.  F0 (NULL =
 * zeros) starts P at [F0 U0] instead of [0 0 0 0 U0] (:41). */
int orc_icgn_subset_ex(const orc_images* im, int64_t x0, int64_t y0, int winsize, double tol,
                       int maxit, int stop_rule, double tol_dp, const double U0[2],
                       const double F0[4], double U[2], double F[4], double H21[21],
                       double* zncc) {
  const int maxIterNum = maxit;
  const int winsize0 = winsize;
  double P[6] = {0, 0, 0, 0, U0[0], U0[1]};
  if (F0)
    for (int k = 0; k < 4; k++) P[k] = F0[k];
  double H[36];
  memset(H, 0, sizeof(H));
  for (int k = 0; k < 21; k++) H21[k] = 0.0;
  *zncc = NAN;
  U[0] = U[1] = NAN;
  F[0] = F[1] = F[2] = F[3] = NAN;

  win_t w = {x0 - winsize / 2, y0 - winsize / 2, winsize + 1, winsize + 1};
  if (!win_ok(im, &w)) return ORC_ERR;
  int S = winsize + 1;
  int64_t NS = (int64_t)S * S;
  /* Buffers big enough for the enlarged window (<= 1.3x per side). */
  const int Smax = 2 * (winsize + 1) + 8;
  const int64_t NMAX = (int64_t)Smax * Smax;
  double* fmask = (double*)malloc(sizeof(double) * NMAX);
  double* f = (double*)malloc(sizeof(double) * NMAX);
  double* fx = (double*)malloc(sizeof(double) * NMAX);
  double* fy = (double*)malloc(sizeof(double) * NMAX);
  double* g = (double*)malloc(sizeof(double) * NMAX);
  double* X = (double*)malloc(sizeof(double) * NMAX);
  double* Y = (double*)malloc(sizeof(double) * NMAX);
  double* r = (double*)malloc(sizeof(double) * NMAX);
  uint8_t* bw = (uint8_t*)malloc(NMAX);
  uint8_t* bwg = (uint8_t*)malloc(NMAX);
  uint8_t* lg = (uint8_t*)malloc(NMAX);
  int status = ORC_ERR;
  win_extract(im, &w, fmask, f, fx, fy);
  if (any_nan(f, NS)) goto done; /* logical(NaN) throws */
  const int64_t c0 = count_zero(f, NS);

  if ((double)c0 < 0.4 * (double)S * (double)S) {
    if (c0 > 0) {
      const int wn = grow_winsize((double)S * S, (double)S * S - (double)c0, winsize);
      if (wn < 0 || (wn + 1) > Smax) goto done;
      winsize = wn;
      S = winsize + 1;
      NS = (int64_t)S * S;
      w.i0 = x0 - winsize / 2;
      w.j0 = y0 - winsize / 2;
      w.S1 = w.S2 = S;
      if (!win_ok(im, &w)) goto done;
      win_extract(im, &w, fmask, f, fx, fy);
    }
    /* bwselect(logical(tempfImgMask), floor((ws+1)/2), floor((ws+1)/2), 4) */
    {
      const int64_t seed = (winsize + 1) / 2;
      for (int64_t k = 0; k < NS; k++) lg[k] = (fmask[k] != 0.0);
      orc_bwselect(lg, S, S, seed, seed, 4, bw);
      for (int64_t k = 0; k < NS; k++) {
        f[k] *= (double)bw[k];
        fx[k] *= (double)bw[k];
        fy[k] *= (double)bw[k];
      }
    }
    const int h = winsize / 2;
    /* H = sum SD' SD, SD = [dx*fx dx*fy dy*fx dy*fy fx fy] (:91-103) */
    for (int b = 0; b < S; b++)
      for (int a = 0; a < S; a++) {
        const int64_t k = a + (int64_t)b * S;
        const double dxm = (double)(a - h), dym = (double)(b - h);
        const double sd[6] = {dxm * fx[k], dxm * fy[k], dym * fx[k], dym * fy[k], fx[k], fy[k]};
        for (int p = 0; p < 6; p++)
          for (int q = 0; q < 6; q++) H[p * 6 + q] += sd[p] * sd[q];
      }
    double meanf, bottomf;
    masked_moments(f, NS, &meanf, &bottomf);

    double normNew = 1.0, normAbs = 1.0, normInit = 0.0;
    int step = 0;
    const int dp_rule = (stop_rule == 1 || stop_rule == 2);
    int dp_conv = 0;
    double m_prev = -1.0; /* measure of the previous update (rule 2) */
    int have_bwg = 0;
    int64_t sum_bwg = 0;
    double last_zncc = NAN;
    while (step <= maxIterNum && (dp_rule || (normNew > tol && normAbs > tol))) {
      step++;
      if (step > 1 && c0 > 0) {
        /* :139-145: the window is re-sized from the last g mask, but f,
         * DfDx, DfDy and H keep their size; a change of size makes the
         * next elementwise operation throw. */
        if (!have_bwg || sum_bwg == 0) goto done;
        const int wn = grow_winsize((double)(winsize0 + 1) * (winsize0 + 1), (double)sum_bwg,
                                    winsize);
        if (wn < 0) goto done;
        if (wn != winsize) {
          win_t w2 = {x0 - wn / 2, y0 - wn / 2, wn + 1, wn + 1};
          if (!win_in_image(im, &w2)) goto done;
          /* u22/v22 at the new size; out of image -> break (status 101) */
          const int h2 = wn / 2;
          int oob = 0;
          for (int b = 0; b <= wn && !oob; b++)
            for (int a = 0; a <= wn; a++) {
              const double dxm = (double)(a - h2), dym = (double)(b - h2);
              const double u22 = ((1 + P[0]) * dxm + P[2] * dym) + ((double)x0 + P[4]);
              const double v22 = (P[1] * dxm + (1 + P[3]) * dym) + ((double)y0 + P[5]);
              if (u22 < 3 || u22 > (double)im->M - 2 || v22 < 3 || v22 > (double)im->N - 2) {
                oob = 1;
                break;
              }
            }
          if (oob) {
            normNew = 1e6;
            break;
          }
          goto done; /* size mismatch -> MATLAB error */
        }
      }
      /* warp (:149-152) */
      int oob = 0;
      for (int b = 0; b < S; b++)
        for (int a = 0; a < S; a++) {
          const int64_t k = a + (int64_t)b * S;
          const double dxm = (double)(a - h), dym = (double)(b - h);
          const double u22 = ((1 + P[0]) * dxm + P[2] * dym) + ((double)x0 + P[4]);
          const double v22 = (P[1] * dxm + (1 + P[3]) * dym) + ((double)y0 + P[5]);
          if (u22 < 3 || u22 > (double)im->M - 2 || v22 < 3 || v22 > (double)im->N - 2) oob = 1;
          Y[k] = u22;
          X[k] = v22;
        }
      if (oob) {
        normNew = 1e6;
        break;
      }
      interp_def(im, X, Y, NS, g);
      if (any_nan(g, NS)) goto done; /* logical(NaN) throws (:166) */
      const int64_t cg = count_zero(g, NS);
      if (c0 > 0 || cg > 0) {
        /* :170-214 */
        const int64_t seed = (winsize + 1) / 2;
        for (int64_t k = 0; k < NS; k++) lg[k] = (g[k] != 0.0);
        orc_bwselect(lg, S, S, seed, seed, 8, bwg);
        int allone = 1;
        for (int64_t k = 0; k < NS; k++)
          if (!bwg[k]) {
            allone = 0;
            break;
          }
        if (allone) {
          for (int64_t k = 0; k < NS; k++) lg[k] = (fmask[k] != 0.0);
          orc_bwselect(lg, S, S, seed, seed, 8, bwg);
        }
        sum_bwg = 0;
        for (int64_t k = 0; k < NS; k++) {
          g[k] *= (double)bwg[k];
          sum_bwg += bwg[k];
        }
        have_bwg = 1;
        masked_moments(f, NS, &meanf, &bottomf);
      }
      double meang, bottomg;
      masked_moments(g, NS, &meang, &bottomg);
      /* b (:365-374) */
      double b2[6] = {0, 0, 0, 0, 0, 0};
      double rr = 0.0;
      for (int bb = 0; bb < S; bb++)
        for (int a = 0; a < S; a++) {
          const int64_t k = a + (int64_t)bb * S;
          const double dxm = (double)(a - h), dym = (double)(bb - h);
          const double rk = (f[k] - meanf) / bottomf - (g[k] - meang) / bottomg;
          r[k] = rk;
          rr += rk * rk;
          b2[0] += dxm * fx[k] * rk;
          b2[1] += dxm * fy[k] * rk;
          b2[2] += dym * fx[k] * rk;
          b2[3] += dym * fy[k] * rk;
          b2[4] += fx[k] * rk;
          b2[5] += fy[k] * rk;
        }
      last_zncc = 1.0 - 0.5 * rr;
      double bvec[6], nb = 0.0;
      for (int p = 0; p < 6; p++) {
        bvec[p] = bottomf * b2[p];
        nb += bvec[p] * bvec[p];
      }
      nb = sqrt(nb);
      normNew = nb;
      normAbs = nb;
      if (step == 1) normInit = nb;
      if (normInit > tol)
        normNew = normNew / normInit;
      else
        normNew = 0.0;
      if (!dp_rule && (normNew < tol || normAbs < tol)) break;
      double dP[6], nbv[6];
      for (int p = 0; p < 6; p++) nbv[p] = -bvec[p];
      solve_sym(6, H, nbv, dP);
      if (!ic_update(P, dP)) break;
      if (dp_rule) {
        /* the step-size rule (no counterpart in the reference): the update
         * just applied moves the centre by (dP5, dP6) and the half-width
         * points by the gradient terms */
        const double hh = (double)(winsize / 2);
        const double m = dP[4] * dP[4] + dP[5] * dP[5] + (dP[0] * dP[0] + dP[1] * dP[1]) * hh * hh +
                         (dP[2] * dP[2] + dP[3] * dP[3]) * hh * hh;
        if (m < tol_dp * tol_dp) {
          dp_conv = 1;
          break;
        }
        if (stop_rule == 2) {
          /* the a-posteriori form: the iteration contracts by q = m_k /
           * m_{k-1} <= 1/2 and the distance to its fixed point that q
           * predicts, m_k q / (1 - q), is below tol_dp */
          const double mk = sqrt(m), mp = m_prev;
          m_prev = mk;
          if (mp > 0.0 && 2.0 * mk <= mp && m < tol_dp * (mp - mk)) {
            dp_conv = 1;
            break;
          }
        }
      }
    }
    status = step;
    if (dp_rule ? !dp_conv : !(normNew < tol || normAbs < tol)) status = ORC_NOCONV;
    if (isnan(normNew)) status = ORC_NOCONV;
    {
      double sa = 0.0;
      for (int64_t k = 0; k < NS; k++) sa += fabs(f[k]);
      if (sa < 1e-6) status = ORC_EMPTY;
    }
    *zncc = last_zncc;
  } else {
    memset(H, 0, sizeof(H));
    status = ORC_MASKED;
  }
  U[0] = P[4];
  U[1] = P[5];
  F[0] = P[0];
  F[1] = P[1];
  F[2] = P[2];
  F[3] = P[3];
  {
    /* HGlobal = H([1:6 8:12 15:18 22:24 29:30 36]) (column-major) */
    int k = 0;
    for (int col = 0; col < 6; col++)
      for (int row = col; row < 6; row++) H21[k++] = H[row * 6 + col];
  }
done:
  if (status == ORC_ERR) {
    U[0] = U[1] = NAN;
    F[0] = F[1] = F[2] = F[3] = NAN;
    for (int k = 0; k < 21; k++) H21[k] = 0.0;
    *zncc = NAN;
  }
  free(fmask);
  free(f);
  free(fx);
  free(fy);
  free(g);
  free(X);
  free(Y);
  free(r);
  free(bw);
  free(bwg);
  free(lg);
  return status;
}

/* ------------------------------------------------------------------------ */
/* funICGN_Subpb1quadtree.m                                                   */
/* ------------------------------------------------------------------------ */

static void h2x2_from(const double* fx, const double* fy, int64_t n, double Hm[4]) {
  double a = 0, b = 0, c = 0;
  for (int64_t k = 0; k < n; k++) {
    a += fx[k] * fx[k];
    b += fx[k] * fy[k];
    c += fy[k] * fy[k];
  }
  Hm[0] = a;
  Hm[1] = b;
  Hm[2] = b;
  Hm[3] = c;
}

int orc_subpb1_subset(const orc_images* im, int64_t x0, int64_t y0, int wx, int wy, double mu,
                      const double udual[4], const double vdual[2], const double UOld[2],
                      const double FOld[4], double tol, double U[2]) {
  (void)udual; /* read but unused by the reference's first-order update */
  const int maxIterNum = 100;
  double P[6] = {FOld[0], FOld[1], FOld[2], FOld[3], UOld[0], UOld[1]};
  U[0] = U[1] = NAN;
  win_t w = {x0 - wx / 2, y0 - wy / 2, wx + 1, wy + 1};
  if (!win_ok(im, &w)) return ORC_ERR;
  const int Smax1 = 2 * (wx + 1) + 8, Smax2 = 2 * (wy + 1) + 8;
  const int64_t NMAX = (int64_t)Smax1 * Smax2;
  double* fmask = (double*)malloc(sizeof(double) * NMAX);
  double* f = (double*)malloc(sizeof(double) * NMAX);
  double* fx = (double*)malloc(sizeof(double) * NMAX);
  double* fy = (double*)malloc(sizeof(double) * NMAX);
  double* g = (double*)malloc(sizeof(double) * NMAX);
  double* X = (double*)malloc(sizeof(double) * NMAX);
  double* Y = (double*)malloc(sizeof(double) * NMAX);
  uint8_t* bw = (uint8_t*)malloc(NMAX);
  uint8_t* bwg = (uint8_t*)malloc(NMAX);
  uint8_t* lg = (uint8_t*)malloc(NMAX);
  int status = ORC_ERR;
  int64_t NS = (int64_t)w.S1 * w.S2;
  win_extract(im, &w, fmask, f, fx, fy);
  if (any_nan(f, NS)) goto done;
  const int64_t c0 = count_zero(f, NS);
  const double area0 = (double)(wx + 1) * (double)(wy + 1);
  if ((double)c0 < 0.4 * area0) {
    if (c0 > 0) {
      /* :78-79 -- winsize_y uses the already-updated winsize_x */
      const int wxn = grow_winsize(area0, area0 - (double)c0, wx);
      if (wxn < 0) goto done;
      const double area1 = (double)(wxn + 1) * (double)(wy + 1);
      const int wyn = grow_winsize(area1, area1 - (double)c0, wy);
      if (wyn < 0 || wxn + 1 > Smax1 || wyn + 1 > Smax2) goto done;
      wx = wxn;
      wy = wyn;
      w.i0 = x0 - wx / 2;
      w.j0 = y0 - wy / 2;
      w.S1 = wx + 1;
      w.S2 = wy + 1;
      NS = (int64_t)w.S1 * w.S2;
      if (!win_ok(im, &w)) goto done;
      win_extract(im, &w, fmask, f, fx, fy);
    }
    /* bwselect(BW, c=floor((wx+1)/2), r=floor((wy+1)/2), 4): the column
     * index comes from winsize_x and the row index from winsize_y. */
    const int64_t seed_c = (wx + 1) / 2, seed_r = (wy + 1) / 2;
    for (int64_t k = 0; k < NS; k++) lg[k] = (fmask[k] != 0.0);
    orc_bwselect(lg, w.S1, w.S2, seed_r, seed_c, 4, bw);
    for (int64_t k = 0; k < NS; k++) {
      f[k] *= (double)bw[k];
      fx[k] *= (double)bw[k];
      fy[k] *= (double)bw[k];
    }
    double Hm[4];
    h2x2_from(fx, fy, NS, Hm);
    double meanf, bottomf;
    masked_moments(f, NS, &meanf, &bottomf);
    double H2[4];
    for (int k = 0; k < 4; k++) H2[k] = Hm[k] * 2 / (bottomf * bottomf);
    H2[0] += mu;
    H2[3] += mu;

    const int hx = wx / 2, hy = wy / 2;
    double normNew = 1.0, normAbs = 1.0, normInit = 0.0;
    int step = 0;
    while (step <= 100 && normNew > tol && normAbs > tol) {
      step++;
      int oob = 0;
      for (int b = 0; b < w.S2; b++)
        for (int a = 0; a < w.S1; a++) {
          const int64_t k = a + (int64_t)b * w.S1;
          const double dxm = (double)(a - hx), dym = (double)(b - hy);
          const double u22 = ((1 + P[0]) * dxm + P[2] * dym) + ((double)x0 + P[4]);
          const double v22 = (P[1] * dxm + (1 + P[3]) * dym) + ((double)y0 + P[5]);
          if (u22 < 3 || u22 > (double)im->M - 2 || v22 < 3 || v22 > (double)im->N - 2) oob = 1;
          Y[k] = u22;
          X[k] = v22;
        }
      if (oob) {
        normNew = 1e6;
        break;
      }
      interp_def(im, X, Y, NS, g);
      if (any_nan(g, NS)) goto done;
      const int64_t cg = count_zero(g, NS);
      if (c0 > 0 || cg > 0) {
        /* :158-192 */
        for (int64_t k = 0; k < NS; k++) lg[k] = (g[k] != 0.0);
        orc_bwselect(lg, w.S1, w.S2, seed_r, seed_c, 4, bwg);
        int allone = 1;
        for (int64_t k = 0; k < NS; k++)
          if (!bwg[k]) {
            allone = 0;
            break;
          }
        if (allone) {
          for (int64_t k = 0; k < NS; k++) lg[k] = (fmask[k] != 0.0);
          orc_bwselect(lg, w.S1, w.S2, seed_r, seed_c, 8, bwg);
        }
        for (int64_t k = 0; k < NS; k++) g[k] *= (double)bwg[k];
        {
          const int64_t Mi = im->M, Mc = im->M - 6;
          for (int b = 0; b < w.S2; b++)
            for (int a = 0; a < w.S1; a++) {
              const int64_t i = w.i0 - 1 + a, j = w.j0 - 1 + b;
              const int64_t k = a + (int64_t)b * w.S1;
              f[k] = im->ImgRef[i + j * Mi] * fmask[k] * (double)bwg[k];
              fx[k] = im->DfDx[(i - 3) + (j - 3) * Mc] * (double)bwg[k];
              fy[k] = im->DfDy[(i - 3) + (j - 3) * Mc] * (double)bwg[k];
            }
        }
        h2x2_from(fx, fy, NS, Hm);
        masked_moments(f, NS, &meanf, &bottomf);
        for (int k = 0; k < 4; k++) H2[k] = Hm[k] * 2 / (bottomf * bottomf);
        H2[0] += mu;
        H2[3] += mu;
      }
      double meang, bottomg;
      masked_moments(g, NS, &meang, &bottomg);
      double b2[2] = {0, 0};
      for (int64_t k = 0; k < NS; k++) {
        const double rk = (f[k] - meanf) / bottomf - (g[k] - meang) / bottomg;
        b2[0] += fx[k] * rk;
        b2[1] += fy[k] * rk;
      }
      const double bb0 = bottomf * b2[0], bb1 = bottomf * b2[1];
      const double tb0 = bb0 * 2 / (bottomf * bottomf) + mu * (P[4] - UOld[0] - vdual[0]);
      const double tb1 = bb1 * 2 / (bottomf * bottomf) + mu * (P[5] - UOld[1] - vdual[1]);
      const double dmax = H2[0] > H2[3] ? H2[0] : H2[3];
      double tH[4] = {H2[0] + 1e-3 * dmax, H2[1], H2[2], H2[3] + 1e-3 * dmax};
      if (isnan(H2[0]) || isnan(H2[3])) {
        /* max() ignores NaN unless both are NaN */
        const double m2 = isnan(H2[0]) ? H2[3] : H2[0];
        tH[0] = H2[0] + 1e-3 * m2;
        tH[3] = H2[3] + 1e-3 * m2;
      }
      double d56[2];
      const double nb[2] = {-tb0, -tb1};
      solve_sym(2, tH, nb, d56);
      const double nrm = sqrt(tb0 * tb0 + tb1 * tb1);
      normNew = nrm;
      normAbs = nrm;
      if (step == 1) normInit = nrm;
      if (normInit > tol)
        normNew = normNew / normInit;
      else
        normNew = 0.0;
      if (normNew < tol || normAbs < tol) break;
      const double dP[6] = {0, 0, 0, 0, d56[0], d56[1]};
      if (!ic_update(P, dP)) break;
    }
    status = step;
    if (!(normNew < tol || normAbs < tol)) status = maxIterNum + 1;
    if (isnan(normNew)) status = maxIterNum + 1;
    {
      double sa = 0.0;
      for (int64_t k = 0; k < NS; k++) sa += fabs(f[k]);
      if (sa < 1e-6) status = maxIterNum + 3;
    }
  } else {
    status = 102;
  }
  U[0] = P[4];
  U[1] = P[5];
done:
  if (status == ORC_ERR) U[0] = U[1] = NAN;
  free(fmask);
  free(f);
  free(fx);
  free(fy);
  free(g);
  free(X);
  free(Y);
  free(bw);
  free(bwg);
  free(lg);
  return status;
}

/* ------------------------------------------------------------------------ */
/* Drivers: LocalICGNQuadtree.m:67-163, Subpb1Quadtree.m:71-165               */
/* ------------------------------------------------------------------------ */

static int64_t mround(double v) { return (int64_t)round(v); } /* half away from zero */

void orc_local_icgn(const orc_images* im, const double* coords, const double* U0, int64_t n,
                    const int32_t* winsize, int winsize_per_node, double tol, int nthreads,
                    double* U, double* F, double* H, int32_t* convIt, uint8_t* markHole,
                    double* zncc) {
  orc_local_icgn_ex(im, coords, U0, NULL, n, winsize, winsize_per_node, tol, 100, 0, 0.0,
                    nthreads, U, F, H, convIt, markHole, zncc);
}

void orc_local_icgn_ex(const orc_images* im, const double* coords, const double* U0,
                       const double* F0, int64_t n, const int32_t* winsize,
                       int winsize_per_node, double tol, int maxit, int stop_rule,
                       double tol_dp, int nthreads, double* U, double* F, double* H,
                       int32_t* convIt, uint8_t* markHole, double* zncc) {
  coef_init();
#ifdef _OPENMP
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
#endif
  for (int64_t t = 0; t < n; t++) {
    const int ws = winsize_per_node ? winsize[t] : winsize[0];
    const int64_t x0 = mround(coords[t]), y0 = mround(coords[t + n]);
    /* hole flag (:75-86) */
    {
      win_t w = {x0 - ws / 2, y0 - ws / 2, ws + 1, ws + 1};
      if (!win_in_image(im, &w)) {
        markHole[t] = 1;
      } else {
        int64_t c = 0;
        int bad = 0;
        for (int b = 0; b < w.S2; b++)
          for (int a = 0; a < w.S1; a++) {
            const int64_t i = w.i0 - 1 + a, j = w.j0 - 1 + b;
            const double m = im->mask ? im->mask[i + j * im->M] : 1.0;
            const double v = im->ImgRef[i + j * im->M] * m;
            if (isnan(v)) bad = 1;
            c += (v == 0.0);
          }
        markHole[t] = (bad || (double)c > 0.6 * (double)(ws + 1) * (double)(ws + 1)) ? 1 : 0;
      }
    }
    double Ut[2], Ft[4], Ht[21], zt;
    const double u0[2] = {U0[2 * t], U0[2 * t + 1]};
    const int st = orc_icgn_subset_ex(im, x0, y0, ws, tol, maxit, stop_rule, tol_dp, u0,
                                      F0 ? F0 + 4 * t : NULL, Ut, Ft, Ht, &zt);
    convIt[t] = st;
    U[2 * t] = Ut[0];
    U[2 * t + 1] = Ut[1];
    for (int k = 0; k < 4; k++) F[4 * t + k] = Ft[k];
    for (int k = 0; k < 21; k++) H[t + k * n] = Ht[k];
    zncc[t] = zt;
  }
}

void orc_subpb1(const orc_images* im, const double* coords, const double* UOld,
                const double* FOld, const double* udual, const double* vdual, int64_t n,
                const int32_t* winsize_xy, double mu, double tol, int nthreads, double* U,
                int32_t* convIt) {
  coef_init();
#ifdef _OPENMP
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
#endif
  for (int64_t t = 0; t < n; t++) {
    const int64_t x0 = mround(coords[t]), y0 = mround(coords[t + n]);
    double Ut[2];
    const int st = orc_subpb1_subset(im, x0, y0, winsize_xy[t], winsize_xy[t + n], mu,
                                     udual + 4 * t, vdual + 2 * t, UOld + 2 * t, FOld + 4 * t,
                                     tol, Ut);
    /* A throw leaves ConvItPerEle at its initial 0 (Subpb1Quadtree.m:96-99). */
    convIt[t] = (st == ORC_ERR) ? 0 : st;
    U[2 * t] = Ut[0];
    U[2 * t + 1] = Ut[1];
  }
}

/* ------------------------------------------------------------------------ */
/* normxcorr2 crop, findpeak, funIntegerSearchWholeField                      */
/* ------------------------------------------------------------------------ */

static double eps_of(double x) {
  x = fabs(x);
  if (x == 0.0) return DBL_MIN * DBL_EPSILON;
  int e;
  frexp(x, &e);
  return ldexp(1.0, e - 53);
}

/* normxcorr2 (Lewis 1995, zero-padded) at the lags whose template origin
 * in A is (o1 + p, o2 + q), p < K1, q < K2 (0-based); coefficients whose
 * denominator is at most sqrt(eps) of the largest one over these lags, or
 * which leave [-1, 1] by more than sqrt(eps), are 0. */
static int ncc_lags(const double* T, int64_t s1, int64_t s2, const double* A, int64_t W1,
                    int64_t W2, int64_t o1, int64_t K1, int64_t o2, int64_t K2, double* out) {
  const int64_t n = s1 * s2;
  double st = 0.0;
  for (int64_t k = 0; k < n; k++) st += T[k];
  const double tm = st / (double)n;
  double ss = 0.0;
  int allsame = 1;
  for (int64_t k = 0; k < n; k++) {
    const double d = T[k] - tm;
    ss += d * d;
    if (T[k] != T[0]) allsame = 0;
  }
  if (allsame) return -1; /* "The values of TEMPLATE cannot all be the same." */
  const double denT = sqrt(ss);
  double* den = (double*)malloc(sizeof(double) * K1 * K2);
  double* num = (double*)malloc(sizeof(double) * K1 * K2);
  double dmax = 0.0;
  for (int64_t q = 0; q < K2; q++)
    for (int64_t p = 0; p < K1; p++) {
      const int64_t t1 = o1 + p, t2 = o2 + q; /* template top-left in A */
      double sa = 0.0, sa2 = 0.0, nm = 0.0;
      for (int64_t b = 0; b < s2; b++) {
        const int64_t j = t2 + b;
        if (j < 0 || j >= W2) continue;
        for (int64_t a = 0; a < s1; a++) {
          const int64_t i = t1 + a;
          if (i < 0 || i >= W1) continue;
          const double v = A[i + j * W1];
          sa += v;
          sa2 += v * v;
          nm += (T[a + b * s1] - tm) * v;
        }
      }
      double dl = sa2 - sa * sa / (double)n;
      if (dl < 0) dl = 0;
      const double d = denT * sqrt(dl);
      den[p + q * K1] = d;
      num[p + q * K1] = nm;
      if (d > dmax) dmax = d;
    }
  const double tolv = sqrt(eps_of(dmax));
  const double back = sqrt(DBL_EPSILON);
  for (int64_t k = 0; k < K1 * K2; k++) {
    double c = 0.0;
    if (den[k] > tolv) c = num[k] / den[k];
    if (fabs(c) - 1 > back) c = 0.0;
    out[k] = c;
  }
  free(den);
  free(num);
  return 0;
}

int orc_normxcorr2_crop(const double* T, int64_t s1, int64_t s2, const double* A, int64_t W1,
                        int64_t W2, double* out) {
  return ncc_lags(T, s1, s2, A, W1, W2, -1, W1 - s1 + 3, -1, W2 - s2 + 3, out);
}

int orc_normxcorr2_full(const double* T, int64_t s1, int64_t s2, const double* A, int64_t W1,
                        int64_t W2, double* out) {
  return ncc_lags(T, s1, s2, A, W1, W2, -(s1 - 1), W1 + s1 - 1, -(s2 - 1), W2 + s2 - 1, out);
}

/* Bin edges of MATLAB's histcounts(x, nbins) for data in [xmin, xmax]: the
 * "nice" edges of its bin picker with the number of bins given (a left edge
 * on a multiple of a rounded raw width, then the width rounded up at the
 * decimal place of the slack between (xmax - left)/nbins and
 * (xmax - left)/(nbins - 1)), or a unit-wide cover for nearly constant data.
 * Parity unpinned beyond the cases tests/test_oracle.py constructs:
 * MATLAB's bin picker is not vendored. */
static void hist_edges(double xmin, double xmax, int nbins, double* edges) {
  const double xscale = fmax(fabs(xmin), fabs(xmax));
  const double xrange = xmax - xmin;
  double raw = xrange / nbins;
  raw = fmax(raw, eps_of(xscale));
  double left, right, bw;
  if (xrange > fmax(sqrt(eps_of(xscale)), DBL_MIN)) {
    const double p10 = pow(10.0, floor(log10(raw)));
    const double rel = raw / p10;
    bw = p10 * floor(rel);
    left = fmax(fmin(bw * floor(xmin / bw), xmin), -DBL_MAX);
    if (nbins > 1) {
      const double ll = (xmax - left) / nbins;
      const double ul = (xmax - left) / (nbins - 1);
      const double q10 = pow(10.0, floor(log10(ul - ll)));
      bw = q10 * ceil(ll / q10);
    }
    right = fmin(fmax(left + nbins * bw, xmax), DBL_MAX);
  } else {
    const double brange = fmax(1.0, ceil(nbins * eps_of(xscale)));
    left = floor(2.0 * (xmin - brange / 4.0)) / 2.0;
    right = ceil(2.0 * (xmax + brange / 4.0)) / 2.0;
    bw = (right - left) / nbins;
  }
  edges[0] = left;
  for (int k = 1; k < nbins; k++) edges[k] = left + k * bw;
  edges[nbins] = right;
}

void orc_hist_edges(double xmin, double xmax, int nbins, double* edges) {
  hist_edges(xmin, xmax, nbins, edges);
}

/* compute_qFactor of funIntegerSearch.m:211-264 on one correlation map:
 * cc_min = cc - min(cc), peak = max(cc_min); PCE = peak^2 / mean(cc_min^2);
 * PPE = 1 / entropy of histcounts(cc_min, 30). */
void orc_qfactor(const double* cc, int64_t n, double q[2]) {
  double mn = INFINITY, mx = -INFINITY;
  for (int64_t k = 0; k < n; k++) {
    mn = fmin(mn, cc[k]);
    mx = fmax(mx, cc[k]);
  }
  const double peak = mx - mn;
  double e2 = 0.0;
  for (int64_t k = 0; k < n; k++) e2 += (cc[k] - mn) * (cc[k] - mn);
  q[0] = (peak * peak) / (e2 / (double)n);
  double edges[31];
  hist_edges(0.0, peak, 30, edges);
  int64_t cnt[30] = {0};
  int64_t tot = 0;
  for (int64_t k = 0; k < n; k++) {
    const double v = cc[k] - mn;
    if (v < edges[0] || v > edges[30]) continue;
    int b = 29;
    for (int j = 1; j <= 30; j++)
      if (v < edges[j]) {
        b = j - 1;
        break;
      }
    cnt[b]++;
    tot++;
  }
  double ent = 0.0;
  for (int j = 0; j < 30; j++)
    if (cnt[j] > 0) {
      const double p = (double)cnt[j] / (double)tot;
      ent += p * log(1.0 / p);
    }
  q[1] = 1.0 / ent;
}

void orc_findpeak(const double* f, int64_t m, int64_t n, double* xpeak, double* ypeak,
                  double* max_f, int64_t* ipr, int64_t* ipc) {
  /* [max_f, imax] = max(f(:)): first maximum in column-major order */
  int64_t imax = 0;
  double mx = -INFINITY;
  int found = 0;
  for (int64_t k = 0; k < m * n; k++) {
    if (isnan(f[k])) continue;
    if (!found || f[k] > mx) {
      mx = f[k];
      imax = k;
      found = 1;
    }
  }
  const int64_t yp = imax % m + 1, xp = imax / m + 1;
  *ipr = yp;
  *ipc = xp;
  *xpeak = (double)xp;
  *ypeak = (double)yp;
  *max_f = mx;
  if (xp == 1 || xp == n || yp == 1 || yp == m) return;
  /* 3x3 quadratic least squares, u = A1 + A2 x + A3 y + A4 xy + A5 x^2 +
   * A6 y^2 over x (column offset), y (row offset) in {-1,0,1}.  Normal
   * equations of the constant design (findpeak.m:37-46). */
  double su = 0, sxu = 0, syu = 0, sxyu = 0, sx2u = 0, sy2u = 0;
  for (int dc = -1; dc <= 1; dc++)
    for (int dr = -1; dr <= 1; dr++) {
      const double u = f[(yp - 1 + dr) + (xp - 1 + dc) * m];
      su += u;
      sxu += dc * u;
      syu += dr * u;
      sxyu += dc * dr * u;
      sx2u += dc * dc * u;
      sy2u += dr * dr * u;
    }
  double A[6];
  A[1] = sxu / 6.0;
  A[2] = syu / 6.0;
  A[3] = sxyu / 4.0;
  {
    /* [9 6 6; 6 6 4; 6 4 6] [A1 A5 A6]' = [su sx2u sy2u]' */
    const double Nm[9] = {9, 6, 6, 6, 6, 4, 6, 4, 6};
    const double rhs[3] = {su, sx2u, sy2u};
    double sol[3];
    solve_sym(3, Nm, rhs, sol);
    A[0] = sol[0];
    A[4] = sol[1];
    A[5] = sol[2];
  }
  const double den = A[3] * A[3] - 4 * A[4] * A[5];
  double xo = (-A[2] * A[3] + 2 * A[5] * A[1]) / den;
  double yo = -1 / den * (A[3] * A[1] - 2 * A[4] * A[2]);
  if (fabs(xo) > 1 || fabs(yo) > 1) return;
  xo = round(1000 * xo) / 1000;
  yo = round(1000 * yo) / 1000;
  *xpeak = xp + xo;
  *ypeak = yp + yo;
  *max_f = A[0] + xo * A[1] + yo * A[2] + xo * yo * A[3] + xo * xo * A[4] + yo * yo * A[5];
}

int orc_integer_search(const double* f, const double* g, int64_t M, int64_t N,
                       const int64_t gridx[2], const int64_t gridy[2], int ws, int step,
                       const int64_t region[4], int nthreads, double* x, double* y, double* u,
                       double* v, double* ccmax, int32_t* ipeak_row, int32_t* ipeak_col,
                       int64_t* nX_out, int64_t* nY_out) {
  return orc_integer_search_q(f, g, M, N, gridx, gridy, ws, step, region, nthreads, x, y, u, v,
                              ccmax, ipeak_row, ipeak_col, NULL, nX_out, nY_out);
}

int orc_integer_search_q(const double* f, const double* g, int64_t M, int64_t N,
                         const int64_t gridx[2], const int64_t gridy[2], int ws, int step,
                         const int64_t region[4], int nthreads, double* x, double* y, double* u,
                         double* v, double* ccmax, int32_t* ipeak_row, int32_t* ipeak_col,
                         double* qfactors, int64_t* nX_out, int64_t* nY_out) {
  /* XList = gridx(1):step:gridx(2)-ws-step (funIntegerSearch.m:106-107) */
  int64_t nX = 0, nY = 0;
  for (int64_t i = gridx[0]; i <= gridx[1] - ws - step; i += step) nX++;
  for (int64_t j = gridy[0]; j <= gridy[1] - ws - step; j += step) nY++;
  *nX_out = nX;
  *nY_out = nY;
  if (!x) return 0; /* size query */
  const int64_t tot = nX * nY;
#ifdef _OPENMP
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
#endif
  for (int64_t t = 0; t < tot; t++) {
    const int64_t ci = t % nX, cj = t / nX; /* 0-based */
    const int64_t ii = gridx[0] + ci * step, jj = gridy[0] + cj * step;
    int64_t L = region[0], R = region[1], Uu = region[2], Dd = region[3];
    if (ii - L < 1) L = ii - 1;
    if (jj - Uu < 1) Uu = jj - 1;
    if (ii + ws + R > M) R = gridx[1] - ii - ws - 1;
    if (jj + ws + Dd > N) Dd = gridy[1] - jj - ws - 1;
    if (L < 0) L = 0;
    if (R < 0) R = 0;
    if (Uu < 0) Uu = 0;
    if (Dd < 0) Dd = 0;
    const int64_t s = ws + 1;
    const int64_t W1 = s + L + R, W2 = s + Uu + Dd;
    double* C = (double*)malloc(sizeof(double) * s * s);
    double* Dw = (double*)malloc(sizeof(double) * W1 * W2);
    for (int64_t b = 0; b < s; b++)
      for (int64_t a = 0; a < s; a++) C[a + b * s] = f[(ii - 1 + a) + (jj - 1 + b) * M];
    for (int64_t b = 0; b < W2; b++)
      for (int64_t a = 0; a < W1; a++)
        Dw[a + b * W1] = g[(ii - L - 1 + a) + (jj - Uu - 1 + b) * M];
    const int64_t K1 = W1 - s + 3, K2 = W2 - s + 3;
    double* cc = (double*)malloc(sizeof(double) * K1 * K2);
    const int64_t o = cj + ci * nY; /* u(cj, ci), nY x nX column-major */
    if (qfactors) {
      /* q-factors of the full normxcorr2 map (:162-163); [Inf Inf] where
       * normxcorr2 throws (:173) */
      const int64_t F1 = W1 + s - 1, F2 = W2 + s - 1;
      double* full = (double*)malloc(sizeof(double) * F1 * F2);
      double q[2] = {INFINITY, INFINITY};
      if (orc_normxcorr2_full(C, s, s, Dw, W1, W2, full) == 0) orc_qfactor(full, F1 * F2, q);
      qfactors[t] = q[0];
      qfactors[t + tot] = q[1];
      free(full);
    }
    if (orc_normxcorr2_crop(C, s, s, Dw, W1, W2, cc) != 0) {
      u[o] = NAN;
      v[o] = NAN;
      ccmax[o] = NAN;
      ipeak_row[o] = 0;
      ipeak_col[o] = 0;
    } else {
      double xp, yp, mf;
      int64_t pr, pc;
      orc_findpeak(cc, K1, K2, &xp, &yp, &mf, &pr, &pc);
      u[o] = yp - 2 - (double)L;
      v[o] = xp - 2 - (double)Uu;
      ccmax[o] = mf;
      ipeak_row[o] = (int32_t)pr;
      ipeak_col[o] = (int32_t)pc;
    }
    x[o] = (double)(ii + ii + ws) / 2.0;
    y[o] = (double)(jj + jj + ws) / 2.0;
    free(C);
    free(Dw);
    free(cc);
  }
  if (qfactors && tot > 0) {
    /* qf_ = q - min(q); cc.qfactors = qf_ / max(qf_) per column, NaN
     * ignored by min and max (:184-187) */
    for (int k = 0; k < 2; k++) {
      double* qc = qfactors + k * tot;
      double mn = NAN, mx = NAN;
      for (int64_t t = 0; t < tot; t++)
        if (!isnan(qc[t]) && (isnan(mn) || qc[t] < mn)) mn = qc[t];
      for (int64_t t = 0; t < tot; t++) qc[t] -= mn;
      for (int64_t t = 0; t < tot; t++)
        if (!isnan(qc[t]) && (isnan(mx) || qc[t] > mx)) mx = qc[t];
      for (int64_t t = 0; t < tot; t++) qc[t] /= mx;
    }
  }
  return 0;
}
