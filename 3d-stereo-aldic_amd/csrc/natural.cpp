// natural.cpp -- host post-processing of the local-step drivers: natural-
// neighbour (Sibson) interpolation with nearest-neighbour extrapolation, the
// interpolant the reference builds with scatteredInterpolant(x, y, v,
// 'natural', 'nearest'), and the NaN fill that uses it
// (func_quadtree/LocalICGNQuadtree.m:200-247 and :290-325,
// func_quadtree/Subpb1Quadtree.m:189-254).
//
// Interpolant.  For a query q strictly inside the convex hull of the sites,
// the value is sum_i w_i v_i with Sibson's weights: w_i is the area that
// q's Voronoi cell (in the diagram of the sites plus q) takes from the cell
// of site i, over the area of q's cell.  q's cell is built by clipping a box
// with the bisector half-planes of the sites around q, visited bucket ring
// by bucket ring outwards until the sites left are at least twice the
// cell's circumradius away, beyond which no site can cut the cell (near the
// hull the cell reaches far out; if the box still bounds it after every
// site, the box grows).  The area taken from site i is q's cell clipped
// to the points nearer to i than to every other natural neighbour (a point
// of q's cell lies in the old cell of one of them).  On the hull boundary
// the weights reduce to linear interpolation between the two ends of the
// hull edge; outside the hull the value is that of the nearest site (the
// first site in input order on a tie).  A query equal to a site takes its
// value.  Duplicate sites are merged and their values averaged.
//
// Degenerate input (fewer than three sites, or all sites on one line) is
// refused with ALDIC_EINVAL, as the reference's interpolant construction
// fails there.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <string>
#include <thread>
#include <vector>

#include "../../include/aldic.h"

namespace aldic {
int set_error(int code, const std::string& msg);
}

namespace {

struct Vtx {
  double x, y;
  int lab;  // site whose bisector carries the edge from this vertex to the next (-1: box)
};

// Clip convex polygon P (CCW) to {p : a*px + b*py <= c}; the new edge along
// the line gets label lab.
void clip(const std::vector<Vtx>& P, double a, double b, double c, int lab, std::vector<Vtx>& out) {
  out.clear();
  const size_t n = P.size();
  if (n == 0) return;
  for (size_t k = 0; k < n; k++) {
    const Vtx& p0 = P[k];
    const Vtx& p1 = P[(k + 1) % n];
    const double d0 = a * p0.x + b * p0.y - c;
    const double d1 = a * p1.x + b * p1.y - c;
    const bool in0 = d0 <= 0.0, in1 = d1 <= 0.0;
    if (in0) out.push_back(p0);
    if (in0 != in1) {
      const double t = d0 / (d0 - d1);
      Vtx m{p0.x + t * (p1.x - p0.x), p0.y + t * (p1.y - p0.y), in0 ? lab : p0.lab};
      // leaving: the new edge runs along the clip line; entering: it keeps
      // the label of the edge being cut
      out.push_back(m);
    }
  }
}

double area(const std::vector<Vtx>& P) {
  double s = 0.0;
  const size_t n = P.size();
  for (size_t k = 0; k < n; k++) {
    const Vtx& p = P[k];
    const Vtx& q = P[(k + 1) % n];
    s += p.x * q.y - q.x * p.y;
  }
  return 0.5 * s;
}

// Working storage of one query's Sibson evaluation; one per thread.
struct Scratch {
  std::vector<int> nn;
  std::vector<Vtx> poly, tmp, tmp2;
};

class Interp {
 public:
  // xs, ys: n sites; vals: n x nv column-major
  int build(const double* xs, const double* ys, const double* vals, int64_t n, int nv) {
    nv_ = nv;
    std::vector<int64_t> ord;
    ord.reserve(n);
    for (int64_t i = 0; i < n; i++) {
      if (!std::isfinite(xs[i]) || !std::isfinite(ys[i]))
        return aldic::set_error(ALDIC_EINVAL, "natural_interp: non-finite site coordinate");
      ord.push_back(i);
    }
    std::stable_sort(ord.begin(), ord.end(), [&](int64_t a, int64_t b) {
      return xs[a] < xs[b] || (xs[a] == xs[b] && ys[a] < ys[b]);
    });
    x_.clear();
    y_.clear();
    v_.clear();
    first_.clear();
    for (size_t k = 0; k < ord.size();) {
      size_t e = k;
      int64_t lo = ord[k];
      while (e < ord.size() && xs[ord[e]] == xs[ord[k]] && ys[ord[e]] == ys[ord[k]]) {
        lo = std::min(lo, ord[e]);
        e++;
      }
      x_.push_back(xs[ord[k]]);
      y_.push_back(ys[ord[k]]);
      for (int c = 0; c < nv; c++) {
        double s = 0.0;
        for (size_t j = k; j < e; j++) s += vals[ord[j] + (int64_t)c * n];
        v_.push_back(s / (double)(e - k));
      }
      first_.push_back(lo);
      k = e;
    }
    const int m = (int)x_.size();
    if (m < 3) return aldic::set_error(ALDIC_EINVAL, "natural_interp: fewer than 3 distinct sites");
    // convex hull (Andrew's monotone chain over the x-then-y sorted sites)
    hull_.clear();
    std::vector<int> H(2 * m);
    int k = 0;
    for (int i = 0; i < m; i++) {
      while (k >= 2 && cross(H[k - 2], H[k - 1], x_[i], y_[i]) <= 0.0L) k--;
      H[k++] = i;
    }
    for (int i = m - 2, t = k + 1; i >= 0; i--) {
      while (k >= t && cross(H[k - 2], H[k - 1], x_[i], y_[i]) <= 0.0L) k--;
      H[k++] = i;
    }
    hull_.assign(H.begin(), H.begin() + std::max(k - 1, 0));
    if (hull_.size() < 3)
      return aldic::set_error(ALDIC_EINVAL, "natural_interp: the sites are collinear");
    // bucket grid
    xmin_ = *std::min_element(x_.begin(), x_.end());
    xmax_ = *std::max_element(x_.begin(), x_.end());
    ymin_ = *std::min_element(y_.begin(), y_.end());
    ymax_ = *std::max_element(y_.begin(), y_.end());
    const double w = xmax_ - xmin_, h = ymax_ - ymin_;
    cell_ = std::sqrt(std::max(w * h, 1e-300) / m) * 1.5;
    cell_ = std::max(cell_, 1e-9 * std::max(w, h));
    gnx_ = std::max(1, std::min(4096, (int)(w / cell_) + 1));
    gny_ = std::max(1, std::min(4096, (int)(h / cell_) + 1));
    cellx_ = w / gnx_ > 0 ? w / gnx_ : cell_;
    celly_ = h / gny_ > 0 ? h / gny_ : cell_;
    cstart_.assign((size_t)gnx_ * gny_ + 1, 0);
    std::vector<int> cid(m);
    for (int i = 0; i < m; i++) {
      cid[i] = cell_of(x_[i], y_[i]);
      cstart_[cid[i] + 1]++;
    }
    for (size_t c = 0; c + 1 < cstart_.size(); c++) cstart_[c + 1] += cstart_[c];
    cidx_.assign(m, 0);
    std::vector<int> fill(cstart_.begin(), cstart_.end() - 1);
    for (int i = 0; i < m; i++) cidx_[fill[cid[i]]++] = i;
    scale_ = std::max(w, h);
    return ALDIC_OK;
  }

  // out: nv values
  void eval(double qx, double qy, double* out) { eval(qx, qy, out, own_); }

  // Queries k = 0..nq-1 at (qx[k], qy[k]); value c of query k goes to
  // out[k * os + c * cs].  Queries are independent, so they are split over
  // threads when there are enough of them; the values do not depend on the
  // split.
  void eval_many(const double* qx, const double* qy, int64_t nq, double* out, int64_t os,
                 int64_t cs) {
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int64_t nt = std::min<int64_t>(std::min<int64_t>(hw, 16), nq / 128);
    auto work = [&](int64_t k0, int64_t k1, Scratch& S) {
      double o[16];
      std::vector<double> ob(nv_ > 16 ? nv_ : 0);
      double* op = nv_ > 16 ? ob.data() : o;
      for (int64_t k = k0; k < k1; k++) {
        eval(qx[k], qy[k], op, S);
        for (int c = 0; c < nv_; c++) out[k * os + c * cs] = op[c];
      }
    };
    if (nt <= 1) {
      work(0, nq, own_);
      return;
    }
    std::vector<Scratch> S(nt);
    std::vector<std::thread> th;
    th.reserve(nt);
    for (int64_t t = 0; t < nt; t++)
      th.emplace_back(work, nq * t / nt, nq * (t + 1) / nt, std::ref(S[t]));
    for (auto& x : th) x.join();
  }

  void eval(double qx, double qy, double* out, Scratch& S) const {
    if (!std::isfinite(qx) || !std::isfinite(qy)) {
      for (int c = 0; c < nv_; c++) out[c] = std::nan("");
      return;
    }
    const int hit = coincident(qx, qy);
    if (hit >= 0) {
      for (int c = 0; c < nv_; c++) out[c] = v_[(size_t)hit * nv_ + c];
      return;
    }
    int ea = -1, eb = -1;
    double et = 0.0;
    const int where = locate(qx, qy, &ea, &eb, &et);
    if (where < 0) {  // outside: nearest site
      const int s = nearest(qx, qy);
      for (int c = 0; c < nv_; c++) out[c] = v_[(size_t)s * nv_ + c];
      return;
    }
    if (where == 1 && sibson(qx, qy, out, S)) return;
    if (where == 0) {  // on the hull boundary
      for (int c = 0; c < nv_; c++)
        out[c] = (1.0 - et) * v_[(size_t)ea * nv_ + c] + et * v_[(size_t)eb * nv_ + c];
      return;
    }
    // inside by the hull test but with an unbounded or empty cell (rounding
    // at the hull): the nearest site
    const int s = nearest(qx, qy);
    for (int c = 0; c < nv_; c++) out[c] = v_[(size_t)s * nv_ + c];
  }

 private:
  long double cross(int a, int b, double px, double py) const {
    return ((long double)x_[b] - x_[a]) * ((long double)py - y_[a]) -
           ((long double)y_[b] - y_[a]) * ((long double)px - x_[a]);
  }
  int cell_of(double px, double py) const {
    int cx = (int)std::floor((px - xmin_) / cellx_), cy = (int)std::floor((py - ymin_) / celly_);
    cx = std::min(std::max(cx, 0), gnx_ - 1);
    cy = std::min(std::max(cy, 0), gny_ - 1);
    return cx + cy * gnx_;
  }
  template <class F>
  void for_cells(double px, double py, double r, F f) const {
    int cx0 = (int)std::floor((px - r - xmin_) / cellx_), cx1 = (int)std::floor((px + r - xmin_) / cellx_);
    int cy0 = (int)std::floor((py - r - ymin_) / celly_), cy1 = (int)std::floor((py + r - ymin_) / celly_);
    cx0 = std::max(cx0, 0);
    cy0 = std::max(cy0, 0);
    cx1 = std::min(cx1, gnx_ - 1);
    cy1 = std::min(cy1, gny_ - 1);
    for (int cy = cy0; cy <= cy1; cy++)
      for (int cx = cx0; cx <= cx1; cx++) {
        const int c = cx + cy * gnx_;
        for (int k = cstart_[c]; k < cstart_[c + 1]; k++) f(cidx_[k]);
      }
  }
  int coincident(double qx, double qy) const {
    if (qx < xmin_ || qx > xmax_ || qy < ymin_ || qy > ymax_) return -1;
    const int c = cell_of(qx, qy);
    for (int k = cstart_[c]; k < cstart_[c + 1]; k++) {
      const int i = cidx_[k];
      if (x_[i] == qx && y_[i] == qy) return i;
    }
    return -1;
  }
  int nearest(double qx, double qy) const {
    // rings of growing radius until the best distance is inside the ring
    double r = std::max(cellx_, celly_);
    const double far = 4.0 * (scale_ + std::fabs(qx - xmin_) + std::fabs(qy - ymin_)) + r;
    for (;;) {
      int best = -1;
      double bd = 0.0;
      for_cells(qx, qy, r, [&](int i) {
        const double dx = x_[i] - qx, dy = y_[i] - qy, d = dx * dx + dy * dy;
        if (best < 0 || d < bd || (d == bd && first_[i] < first_[best])) {
          best = i;
          bd = d;
        }
      });
      if (best >= 0 && std::sqrt(bd) <= r) return best;
      if (r > far) return best;
      r *= 2.0;
    }
  }
  // 1 strictly inside, 0 on the boundary (edge ea->eb at parameter et), -1 outside
  int locate(double qx, double qy, int* ea, int* eb, double* et) const {
    const size_t h = hull_.size();
    int on = -1;
    for (size_t k = 0; k < h; k++) {
      const int a = hull_[k], b = hull_[(k + 1) % h];
      const long double c = cross(a, b, qx, qy);
      const long double len = std::hypot((long double)x_[b] - x_[a], (long double)y_[b] - y_[a]);
      const long double tol = 1e-12L * len * (long double)scale_;
      if (c < -tol) return -1;
      if (c <= tol) on = (int)k;
    }
    if (on < 0) return 1;
    const int a = hull_[on], b = hull_[(on + 1) % h];
    const double dx = x_[b] - x_[a], dy = y_[b] - y_[a];
    double t = ((qx - x_[a]) * dx + (qy - y_[a]) * dy) / (dx * dx + dy * dy);
    t = std::min(std::max(t, 0.0), 1.0);
    *ea = a;
    *eb = b;
    *et = t;
    return 0;
  }
  bool sibson(double qx, double qy, double* out, Scratch& S) const {
    // q's cell, clipped from a box around q: the buckets are visited in
    // square rings around q's bucket, and the walk stops once every site
    // left is at least twice the cell's circumradius from q (such a site's
    // bisector cannot reach the cell).  A site whose bisector misses the
    // current cell is skipped without touching the polygon.  If the box
    // still bounds the cell when every site has been seen, q is on the hull
    // in effect and the box grows.
    const double rall = std::hypot(std::max(qx - xmin_, xmax_ - qx), std::max(qy - ymin_, ymax_ - qy));
    double B = 2.0 * (scale_ + rall);
    std::vector<Vtx>& V = S.poly;
    std::vector<Vtx>& T = S.tmp;
    std::vector<int>& nn_ = S.nn;
    std::vector<Vtx>& tmp2_ = S.tmp2;
    const int cx = std::min(std::max((int)std::floor((qx - xmin_) / cellx_), 0), gnx_ - 1);
    const int cy = std::min(std::max((int)std::floor((qy - ymin_) / celly_), 0), gny_ - 1);
    const int tmax = std::max(std::max(cx, gnx_ - 1 - cx), std::max(cy, gny_ - 1 - cy));
    for (;;) {
      V = {{-B, -B, -1}, {B, -B, -1}, {B, B, -1}, {-B, B, -1}};
      double R2 = 2.0 * B * B;
      auto visit = [&](int c) {
        for (int k = cstart_[c]; k < cstart_[c + 1]; k++) {
          const int s = cidx_[k];
          // nearer to q (the origin) than to s: s . p <= |s|^2 / 2
          const double sx = x_[s] - qx, sy = y_[s] - qy;
          const double h = 0.5 * (sx * sx + sy * sy);
          if (h * h > R2 * (sx * sx + sy * sy)) continue;  // |s| > 2R
          bool cuts = false;
          for (const Vtx& v : V) cuts |= sx * v.x + sy * v.y - h > 0.0;
          if (!cuts) continue;
          clip(V, sx, sy, h, s, T);
          V.swap(T);
          R2 = 0.0;
          for (const Vtx& v : V) R2 = std::max(R2, v.x * v.x + v.y * v.y);
        }
      };
      for (int t = 0; t <= tmax; t++) {
        if (t > 0) {
          // sites outside rings 0..t-1 are at least this far from q
          const double gx = std::min(qx - (xmin_ + (cx - t + 1) * cellx_),
                                     (xmin_ + (cx + t) * cellx_) - qx);
          const double gy = std::min(qy - (ymin_ + (cy - t + 1) * celly_),
                                     (ymin_ + (cy + t) * celly_) - qy);
          const double g = std::min(gx, gy) - 1e-9 * (cellx_ + celly_);
          if (g > 0.0 && g * g > 4.0 * R2) break;
        }
        const int x0 = cx - t, x1 = cx + t, y0 = cy - t, y1 = cy + t;
        for (int yy = std::max(y0, 0); yy <= std::min(y1, gny_ - 1); yy++) {
          if (yy == y0 || yy == y1) {
            for (int xx = std::max(x0, 0); xx <= std::min(x1, gnx_ - 1); xx++) visit(xx + yy * gnx_);
          } else {
            if (x0 >= 0) visit(x0 + yy * gnx_);
            if (x1 < gnx_ && x1 != x0) visit(x1 + yy * gnx_);
          }
        }
      }
      bool boxed = false;
      const double etol = 1e-12 * B;
      for (size_t k = 0; k < V.size(); k++) {
        const Vtx& p = V[k];
        const Vtx& q = V[(k + 1) % V.size()];
        if (p.lab < 0 && std::hypot(q.x - p.x, q.y - p.y) > etol) boxed = true;
      }
      if (!boxed) break;
      if (B > 1e8 * (scale_ + rall)) return false;  // unbounded: q is on the hull in effect
      B *= 16.0;
    }
    // natural neighbours: sites carrying an edge of positive length
    nn_.clear();
    const double ltol = 1e-12 * std::sqrt(std::max(area(V), 0.0)) + 1e-300;
    for (size_t k = 0; k < V.size(); k++) {
      const Vtx& p = V[k];
      const Vtx& q = V[(k + 1) % V.size()];
      if (p.lab >= 0 && std::hypot(q.x - p.x, q.y - p.y) > ltol &&
          std::find(nn_.begin(), nn_.end(), p.lab) == nn_.end())
        nn_.push_back(p.lab);
    }
    double wsum = 0.0;
    for (int c = 0; c < nv_; c++) out[c] = 0.0;
    for (int i : nn_) {
      T = V;
      const double ix = x_[i] - qx, iy = y_[i] - qy;
      for (int j : nn_) {
        if (j == i) continue;
        // nearer to i than to j: (j - i) . p <= (|j|^2 - |i|^2) / 2
        const double jx = x_[j] - qx, jy = y_[j] - qy;
        clip(T, jx - ix, jy - iy, 0.5 * ((jx * jx + jy * jy) - (ix * ix + iy * iy)), -2, tmp2_);
        T.swap(tmp2_);
        if (T.empty()) break;
      }
      const double a = T.empty() ? 0.0 : std::max(area(T), 0.0);
      wsum += a;
      for (int c = 0; c < nv_; c++) out[c] += a * v_[(size_t)i * nv_ + c];
    }
    if (!(wsum > 0.0)) return false;
    for (int c = 0; c < nv_; c++) out[c] /= wsum;
    return true;
  }

  int nv_ = 1;
  std::vector<double> x_, y_, v_;
  std::vector<int64_t> first_;
  std::vector<int> hull_;
  double xmin_ = 0, xmax_ = 0, ymin_ = 0, ymax_ = 0, cell_ = 1, cellx_ = 1, celly_ = 1, scale_ = 1;
  int gnx_ = 1, gny_ = 1;
  std::vector<int> cstart_, cidx_;
  Scratch own_;
};

// MATLAB round(): half away from zero
inline int64_t mround(double v) { return (int64_t)(v < 0 ? std::ceil(v - 0.5) : std::floor(v + 0.5)); }

// True if some byte of w is zero.
inline bool has_zero_byte(uint64_t w) {
  return ((w - 0x0101010101010101ULL) & ~w & 0x8080808080808080ULL) != 0;
}

// 8-connected components of mask != 0 (M x N, column-major), reported as a
// component id per node pixel: lab[k] for pixel (pi[k], pj[k]) (0-based),
// 0 where the pixel is off the mask, ids 1..K in order of first appearance
// among the nodes.  Returns K.  Only the runs of each column are stored: a
// run is joined (union-find) to every run of the next column that overlaps
// it or touches it diagonally, which is 8-connectivity.
int64_t label_nodes(const uint8_t* mask, int64_t M, int64_t N, const int64_t* pi,
                    const int64_t* pj, int64_t n, int64_t* lab) {
  std::vector<int64_t> cs((size_t)N + 1), ra, rb;
  for (int64_t j = 0; j < N; j++) {
    cs[j] = (int64_t)ra.size();
    const uint8_t* col = mask + j * M;
    int64_t i = 0;
    while (i < M) {
      uint64_t w;
      while (i + 8 <= M && (std::memcpy(&w, col + i, 8), w == 0)) i += 8;
      while (i < M && !col[i]) i++;
      if (i >= M) break;
      const int64_t a = i;
      while (i + 8 <= M && (std::memcpy(&w, col + i, 8), !has_zero_byte(w))) i += 8;
      while (i < M && col[i]) i++;
      ra.push_back(a);
      rb.push_back(i - 1);
    }
  }
  cs[N] = (int64_t)ra.size();
  const int64_t R = cs[N];
  std::vector<int64_t> par((size_t)R);
  for (int64_t r = 0; r < R; r++) par[r] = r;
  auto find = [&](int64_t x) {
    while (par[x] != x) {
      par[x] = par[par[x]];
      x = par[x];
    }
    return x;
  };
  for (int64_t j = 0; j + 1 < N; j++) {
    int64_t p = cs[j], q = cs[j + 1];
    const int64_t pe = cs[j + 1], qe = cs[j + 2];
    while (p < pe && q < qe) {
      if (rb[p] + 1 < ra[q]) {
        p++;
      } else if (rb[q] + 1 < ra[p]) {
        q++;
      } else {
        const int64_t x = find(p), y = find(q);
        if (x != y) par[std::max(x, y)] = std::min(x, y);
        if (rb[p] < rb[q])
          p++;
        else
          q++;
      }
    }
  }
  std::vector<int64_t> id((size_t)R, 0);
  int64_t K = 0;
  for (int64_t k = 0; k < n; k++) {
    const int64_t i = pi[k], j = pj[k];
    const auto b0 = ra.begin() + cs[j], b1 = ra.begin() + cs[j + 1];
    const auto it = std::upper_bound(b0, b1, i);
    lab[k] = 0;
    if (it == b0) continue;
    const int64_t r = (int64_t)(it - ra.begin()) - 1;
    if (rb[r] < i) continue;
    const int64_t root = find(r);
    if (!id[root]) id[root] = ++K;
    lab[k] = id[root];
  }
  return K;
}

}  // namespace

extern "C" int aldic_natural_interp(const double* xs, const double* ys, const double* vals,
                                    int64_t n, int32_t nv, const double* xq, const double* yq,
                                    int64_t nq, double* out) {
  if (n < 0 || nq < 0 || nv < 1 || (n > 0 && (!xs || !ys || !vals)) ||
      (nq > 0 && (!xq || !yq || !out)))
    return aldic::set_error(ALDIC_EINVAL, "natural_interp: bad argument");
  Interp I;
  const int rc = I.build(xs, ys, vals, n, nv);
  if (rc != ALDIC_OK) return rc;
  I.eval_many(xq, yq, nq, out, 1, nq);
  return ALDIC_OK;
}

extern "C" int aldic_nan_fill(const double* coords, int64_t n, double* U, const double* F,
                              const uint8_t* mask, int64_t M, int64_t N, int32_t* regions_filled) {
  if (n < 0 || (n > 0 && (!coords || !U)) || (mask && (M < 1 || N < 1)))
    return aldic::set_error(ALDIC_EINVAL, "nan_fill: bad argument");
  if (regions_filled) *regions_filled = 0;
  const double* cx = coords;
  const double* cy = coords + n;
  auto isnan_u = [&](int64_t i) { return std::isnan(U[2 * i]); };
  std::vector<int64_t> nanidx, good;
  for (int64_t i = 0; i < n; i++) (isnan_u(i) ? nanidx : good).push_back(i);
  // ---- per connected mask region (LocalICGNQuadtree.m:203-247) ----
  if (!nanidx.empty() && mask) {
    std::vector<int64_t> pi(n), pj(n), node_lab(n);
    for (int64_t i = 0; i < n; i++) {
      const int64_t rx = mround(cx[i]), ry = mround(cy[i]);
      if (rx < 1 || rx > M || ry < 1 || ry > N)
        return aldic::set_error(ALDIC_EINVAL, "nan_fill: a node lies outside the image");
      pi[i] = rx - 1;
      pj[i] = ry - 1;
    }
    const int64_t K = label_nodes(mask, M, N, pi.data(), pj.data(), n, node_lab.data());
    std::vector<std::vector<int64_t>> src(K + 1), qry(K + 1);
    for (int64_t i : good) src[node_lab[i]].push_back(i);
    for (int64_t i : nanidx) qry[node_lab[i]].push_back(i);
    std::vector<double> sx, sy, sv, qx, qy, qo;
    for (int64_t r = 1; r <= K; r++) {
      if (src[r].size() < 3 || qry[r].empty()) continue;
      const int64_t ns = (int64_t)src[r].size(), nq = (int64_t)qry[r].size();
      sx.resize(ns);
      sy.resize(ns);
      sv.resize(2 * ns);
      for (int64_t k = 0; k < ns; k++) {
        const int64_t i = src[r][k];
        sx[k] = cx[i];
        sy[k] = cy[i];
        sv[k] = U[2 * i];
        sv[k + ns] = U[2 * i + 1];
      }
      qx.resize(nq);
      qy.resize(nq);
      qo.resize(2 * nq);
      for (int64_t k = 0; k < nq; k++) {
        qx[k] = cx[qry[r][k]];
        qy[k] = cy[qry[r][k]];
      }
      Interp I;
      if (I.build(sx.data(), sy.data(), sv.data(), ns, 2) != ALDIC_OK) continue;  // the region's catch
      I.eval_many(qx.data(), qy.data(), nq, qo.data(), 2, 1);
      for (int64_t k = 0; k < nq; k++) {
        U[2 * qry[r][k]] = qo[2 * k];
        U[2 * qry[r][k] + 1] = qo[2 * k + 1];
      }
      if (regions_filled) (*regions_filled)++;
    }
  }
  // ---- global pass over every node (LocalICGNQuadtree.m:290-325) ----
  nanidx.clear();
  good.clear();
  for (int64_t i = 0; i < n; i++) (isnan_u(i) ? nanidx : good).push_back(i);
  bool fnan = false;
  if (F)
    for (int64_t i = 0; i < n && !fnan; i++) fnan = std::isnan(F[4 * i]);
  if (nanidx.empty() && !fnan) return ALDIC_OK;
  if (good.empty()) return aldic::set_error(ALDIC_EINVAL, "nan_fill: every node is NaN");
  const int64_t ns = (int64_t)good.size();
  std::vector<double> sx(ns), sy(ns), sv(2 * ns);
  for (int64_t k = 0; k < ns; k++) {
    const int64_t i = good[k];
    sx[k] = cx[i];
    sy[k] = cy[i];
    sv[k] = U[2 * i];
    sv[k + ns] = U[2 * i + 1];
  }
  Interp I;
  const int rc = I.build(sx.data(), sy.data(), sv.data(), ns, 2);
  if (rc != ALDIC_OK) return rc;
  I.eval_many(cx, cy, n, U, 2, 1);
  return ALDIC_OK;
}
