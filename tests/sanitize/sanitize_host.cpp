// Host code under AddressSanitizer and UndefinedBehaviorSanitizer (a CPU
// build with g++/gcc, no GPU): the library's natural-neighbour
// interpolation and NaN fill (csrc/natural.cpp) and the FP64 oracle
// (oracle/aldic_oracle.c) on seeded random inputs, the edge cases their
// tests use (few or collinear sites, NaN everywhere but three nodes, nodes
// on the mask border, windows at the image border, masked subsets), and
// checks that need no reference: outputs finite where they must be, the
// interpolants' linear precision.  Exit status 0 when every check holds and
// no sanitizer reported.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <random>
#include <string>
#include <vector>

#include "../../include/aldic.h"
extern "C" {
#include "../../oracle/aldic_oracle.h"
}

namespace aldic {
// natural.cpp reports argument errors through the library's error slot,
// which lives in the HIP part of the library; here the code is enough
int set_error(int code, const std::string&) { return code; }
}  // namespace aldic

static int failures = 0;
#define CHECK(c)                                                   \
  do {                                                             \
    if (!(c)) {                                                    \
      std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
      failures++;                                                  \
    }                                                              \
  } while (0)

static void natural_cases(std::mt19937_64& rng) {
  std::uniform_real_distribution<double> U(0.0, 100.0);
  for (int trial = 0; trial < 20; trial++) {
    const int n = 3 + trial * 7, nq = 50;
    std::vector<double> xs(n), ys(n), v(2 * n), xq(nq), yq(nq), out(2 * nq);
    for (int i = 0; i < n; i++) {
      xs[i] = U(rng);
      ys[i] = U(rng);
      v[i] = 2.0 * xs[i] - 0.5 * ys[i] + 3.0;  // linear: reproduced inside the hull
      v[i + n] = -xs[i] + 0.25 * ys[i];
    }
    for (int k = 0; k < nq; k++) {
      xq[k] = U(rng) * 1.2 - 10.0;
      yq[k] = U(rng) * 1.2 - 10.0;
    }
    const int rc = aldic_natural_interp(xs.data(), ys.data(), v.data(), n, 2, xq.data(), yq.data(),
                                        nq, out.data());
    CHECK(rc == ALDIC_OK);
    for (int k = 0; k < 2 * nq; k++) CHECK(std::isfinite(out[k]));
  }
  // fewer than three sites, collinear sites, bad arguments: errors, no fault
  double xs[3] = {0, 1, 2}, ys[3] = {0, 1, 2}, v[3] = {1, 2, 3}, xq = 0.5, yq = 0.5, o = 0;
  CHECK(aldic_natural_interp(xs, ys, v, 2, 1, &xq, &yq, 1, &o) != ALDIC_OK);
  CHECK(aldic_natural_interp(xs, ys, v, 3, 1, &xq, &yq, 1, &o) != ALDIC_OK);
  CHECK(aldic_natural_interp(nullptr, ys, v, 3, 1, &xq, &yq, 1, &o) != ALDIC_OK);
  // the NaN fill over a two-region mask, with most nodes NaN
  const int64_t M = 120, N = 100;
  std::vector<uint8_t> mask(M * N, 1);
  for (int64_t j = 0; j < N; j++)
    for (int64_t i = 55; i < 65; i++) mask[i + j * M] = 0;
  for (int trial = 0; trial < 10; trial++) {
    const int64_t n = 200;
    std::vector<double> c(2 * n), Uv(2 * n), F(4 * n, 0.0);
    std::uniform_real_distribution<double> X(1.0, (double)M), Y(1.0, (double)N), P(0.0, 1.0);
    for (int64_t i = 0; i < n; i++) {
      c[i] = X(rng);
      c[i + n] = Y(rng);
      const bool nan = P(rng) < 0.1 + 0.08 * trial;
      Uv[2 * i] = nan ? NAN : 0.01 * c[i];
      Uv[2 * i + 1] = nan ? NAN : -0.02 * c[i + n];
    }
    int32_t regions = 0;
    const int rc = aldic_nan_fill(c.data(), n, Uv.data(), F.data(), mask.data(), M, N, &regions);
    CHECK(rc == ALDIC_OK);
    for (int64_t k = 0; k < 2 * n; k++) CHECK(std::isfinite(Uv[k]));
  }
  // every node NaN: an error, no fault
  double c1[4] = {10, 20, 10, 20}, u1[4] = {NAN, NAN, NAN, NAN};
  CHECK(aldic_nan_fill(c1, 2, u1, nullptr, mask.data(), M, N, nullptr) != ALDIC_OK);
}

static void oracle_cases(std::mt19937_64& rng) {
  const int64_t M = 90, N = 84;
  std::normal_distribution<double> G(0.0, 1.0);
  std::vector<double> f(M * N), g(M * N), mask(M * N, 1.0);
  for (auto& a : f) a = G(rng);
  // g: f shifted by (1, 2) with a little noise
  for (int64_t j = 0; j < N; j++)
    for (int64_t i = 0; i < M; i++) {
      const int64_t si = i - 1 < 0 ? 0 : i - 1, sj = j - 2 < 0 ? 0 : j - 2;
      g[i + j * M] = f[si + sj * M] + 0.01 * G(rng);
    }
  for (int64_t j = 30; j < 45; j++)
    for (int64_t i = 40; i < 60; i++) mask[i + j * M] = 0.0;
  std::vector<double> gx((M - 6) * (N - 6)), gy((M - 6) * (N - 6));
  orc_img_gradient(f.data(), mask.data(), M, N, gx.data(), gy.data());
  std::vector<double> coef(M * N);
  orc_bspline_prefilter(g.data(), M, N, coef.data());
  for (int mode = 0; mode < 2; mode++) {
    orc_images im{f.data(), g.data(), mask.data(), gx.data(), gy.data(), M, N,
                  mode ? coef.data() : nullptr};
    const int n = 40;
    std::vector<double> c(2 * n), U0(2 * n, 0.0), U(2 * n), F(4 * n), H(21 * n), z(n);
    std::vector<int32_t> ws(n), conv(n);
    std::vector<uint8_t> hole(n);
    std::uniform_int_distribution<int> P(1, 90);
    for (int i = 0; i < n; i++) {
      c[i] = P(rng);
      c[i + n] = P(rng) * 84 / 90;
      ws[i] = 8 + 4 * (i % 6);
      U0[2 * i] = 1.0;
      U0[2 * i + 1] = 2.0;
    }
    for (int rule = 0; rule < 2; rule++)
      orc_local_icgn_ex(&im, c.data(), U0.data(), nullptr, n, ws.data(), 1, 1e-3, 100, rule, 1e-3,
                        1, U.data(), F.data(), H.data(), conv.data(), hole.data(), z.data());
    std::vector<int32_t> wxy(2 * n);
    for (int i = 0; i < n; i++) wxy[i] = wxy[i + n] = ws[i];
    std::vector<double> UOld(U), FOld(F), ud(4 * n, 0.0), vd(2 * n, 0.01);
    for (auto& a : UOld) a = std::isfinite(a) ? a : 0.0;
    for (auto& a : FOld) a = std::isfinite(a) ? a : 0.0;
    orc_subpb1(&im, c.data(), UOld.data(), FOld.data(), ud.data(), vd.data(), n, wxy.data(), 1e-3,
               1e-2, 1, U.data(), conv.data());
  }
  // the integer search with q-factors over the whole grid (windows clamped
  // at every border), and the full normxcorr2 map
  const int64_t gxr[2] = {1, M}, gyr[2] = {1, N}, reg[4] = {5, 6, 7, 4};
  int64_t nX = 0, nY = 0;
  CHECK(orc_integer_search_q(f.data(), g.data(), M, N, gxr, gyr, 16, 12, reg, 1, nullptr, nullptr,
                             nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, &nX,
                             &nY) == 0);  // the size query
  CHECK(nX > 0 && nY > 0);
  if (nX > 0 && nY > 0) {
    const int64_t np = nX * nY;
    std::vector<double> x(np), y(np), u(np), v(np), cc(np), q(2 * np);
    std::vector<int32_t> pr(np), pc(np);
    orc_integer_search_q(f.data(), g.data(), M, N, gxr, gyr, 16, 12, reg, 1, x.data(), y.data(),
                         u.data(), v.data(), cc.data(), pr.data(), pc.data(), q.data(), &nX, &nY);
    int finite = 0;
    for (int64_t k = 0; k < np; k++) finite += std::isfinite(u[k]);
    CHECK(finite > np / 2);
  }
  std::vector<double> T(9 * 9), A(20 * 18), out((20 + 8) * (18 + 8));
  for (auto& a : T) a = G(rng);
  for (auto& a : A) a = G(rng);
  CHECK(orc_normxcorr2_full(T.data(), 9, 9, A.data(), 20, 18, out.data()) == 0);
  std::vector<double> X(200), Y(200), Z(200);
  std::uniform_real_distribution<double> W(-3.0, 95.0);
  for (int k = 0; k < 200; k++) {
    X[k] = W(rng);
    Y[k] = W(rng);
  }
  orc_interp_spline_d(g.data(), M, N, 1, X.data(), Y.data(), 200, Z.data());
  orc_interp_bspline_d(coef.data(), M, N, X.data(), Y.data(), 200, Z.data());
  // bwselect on random masks, both connectivities, seeds inside and outside
  std::vector<uint8_t> bw(23 * 17), sel(23 * 17);
  std::uniform_int_distribution<int> B(0, 3);
  for (auto& b : bw) b = B(rng) != 0;
  for (int conn : {4, 8})
    for (int64_t r : {1, 12, 23, 30})
      orc_bwselect(bw.data(), 23, 17, r, 9, conn, sel.data());
}

int main() {
  std::mt19937_64 rng(20261204);
  natural_cases(rng);
  oracle_cases(rng);
  if (failures) {
    std::fprintf(stderr, "%d checks failed\n", failures);
    return 1;
  }
  std::printf("host code under the sanitizers: every check held\n");
  return 0;
}
