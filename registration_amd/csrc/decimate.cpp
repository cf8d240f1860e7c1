// Triangle-mesh decimation by quadric-error edge collapse (Garland & Heckbert, "Surface
// Simplification Using Quadric Error Metrics", SIGGRAPH 1997), host code.
//
// The reference simplifies a template before registering it when subsampleTargetSize > 0
// (py-lddmm/surfaceMatching.py:131-133 -> surfaces.py:244-259, VTK's vtkQuadricDecimation
// with target reduction 1 - target / #points); without VTK it raises.  This is this
// library's own implementation of the published algorithm:
//   * every vertex carries the sum of the area-weighted plane quadrics of its faces, plus,
//     on boundary edges, a heavily weighted plane through the edge perpendicular to its face
//     (keeps open borders in place);
//   * an edge's cost is the quadric error of its optimal position (the minimiser of
//     v^T (Q1 + Q2) v; the best of the end points and the midpoint when that system is
//     singular);
//   * edges are collapsed cheapest first (ties broken by vertex ids, so the result is
//     deterministic) while more than `target` vertices remain, skipping collapses that would
//     make the mesh non-manifold (link condition) or flip or degenerate a face.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <iterator>
#include <queue>
#include <vector>

#include "lddmm.h"

namespace {

struct Quadric {
  // symmetric 4x4: a00 a01 a02 a03 a11 a12 a13 a22 a23 a33
  double q[10];
  void clear() { std::memset(q, 0, sizeof(q)); }
  void add_plane(double a, double b, double c, double d, double w) {
    q[0] += w * a * a; q[1] += w * a * b; q[2] += w * a * c; q[3] += w * a * d;
    q[4] += w * b * b; q[5] += w * b * c; q[6] += w * b * d;
    q[7] += w * c * c; q[8] += w * c * d;
    q[9] += w * d * d;
  }
  void add(const Quadric &o) {
    for (int i = 0; i < 10; ++i) q[i] += o.q[i];
  }
  double eval(const double v[3]) const {
    const double x = v[0], y = v[1], z = v[2];
    return q[0] * x * x + 2 * q[1] * x * y + 2 * q[2] * x * z + 2 * q[3] * x + q[4] * y * y +
           2 * q[5] * y * z + 2 * q[6] * y + q[7] * z * z + 2 * q[8] * z + q[9];
  }
};

struct Cand {
  double cost;
  int32_t a, b;          // a < b
  uint32_t sa, sb;       // vertex stamps when the candidate was made
  double pos[3];
  bool operator<(const Cand &o) const {  // priority_queue pops the largest: invert
    if (cost != o.cost) return cost > o.cost;
    if (a != o.a) return a > o.a;
    return b > o.b;
  }
};

void sub3(const double *a, const double *b, double *o) {
  o[0] = a[0] - b[0]; o[1] = a[1] - b[1]; o[2] = a[2] - b[2];
}
void cross3(const double *a, const double *b, double *o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}
double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

class Decimator {
 public:
  Decimator(const double *V, int64_t nv, const int32_t *F, int64_t nf)
      : nv_(nv), nf_(nf), pos_(V, V + 3 * nv), faces_(F, F + 3 * nf), face_alive_(nf, 1),
        vert_alive_(nv, 1), stamp_(nv, 0), vf_(nv), quad_(nv) {
    for (int64_t f = 0; f < nf; ++f)
      for (int k = 0; k < 3; ++k) vf_[faces_[3 * f + k]].push_back((int32_t)f);
    for (auto &q : quad_) q.clear();
    double scale2 = 0;  // squared bounding-box diagonal, for the singular-system test
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t v = 0; v < nv; ++v)
      for (int c = 0; c < 3; ++c) {
        lo[c] = std::fmin(lo[c], pos_[3 * v + c]);
        hi[c] = std::fmax(hi[c], pos_[3 * v + c]);
      }
    for (int c = 0; c < 3; ++c) scale2 += (hi[c] - lo[c]) * (hi[c] - lo[c]);
    scale2_ = scale2 > 0 ? scale2 : 1.0;
    for (int64_t f = 0; f < nf; ++f) {
      double n[3], e1[3], e2[3];
      const double *p0 = &pos_[3 * faces_[3 * f]], *p1 = &pos_[3 * faces_[3 * f + 1]],
                   *p2 = &pos_[3 * faces_[3 * f + 2]];
      sub3(p1, p0, e1);
      sub3(p2, p0, e2);
      cross3(e1, e2, n);
      const double len = std::sqrt(dot3(n, n));
      if (!(len > 0)) continue;
      const double a = n[0] / len, b = n[1] / len, c = n[2] / len, d = -(a * p0[0] + b * p0[1] + c * p0[2]);
      for (int k = 0; k < 3; ++k) quad_[faces_[3 * f + k]].add_plane(a, b, c, d, 0.5 * len);
    }
    add_boundary_quadrics();
  }

  void run(int64_t target) {
    int64_t alive = nv_;
    for (int64_t v = 0; v < nv_; ++v) push_edges_of((int32_t)v, true);
    while (alive > target && !heap_.empty()) {
      Cand c = heap_.top();
      heap_.pop();
      if (!vert_alive_[c.a] || !vert_alive_[c.b] || stamp_[c.a] != c.sa || stamp_[c.b] != c.sb)
        continue;
      if (!collapse(c)) continue;
      --alive;
    }
  }

  void output(double *Vout, int32_t *Fout, int64_t *nv_out, int64_t *nf_out) const {
    std::vector<int32_t> idx(nv_, -1);
    int64_t k = 0;
    for (int64_t v = 0; v < nv_; ++v)
      if (vert_alive_[v] && !vf_[v].empty()) {
        idx[v] = (int32_t)k;
        for (int c = 0; c < 3; ++c) Vout[3 * k + c] = pos_[3 * v + c];
        ++k;
      }
    *nv_out = k;
    int64_t m = 0;
    for (int64_t f = 0; f < nf_; ++f)
      if (face_alive_[f]) {
        for (int c = 0; c < 3; ++c) Fout[3 * m + c] = idx[faces_[3 * f + c]];
        ++m;
      }
    *nf_out = m;
  }

 private:
  int64_t nv_, nf_;
  std::vector<double> pos_;
  std::vector<int32_t> faces_;
  std::vector<uint8_t> face_alive_, vert_alive_;
  std::vector<uint32_t> stamp_;
  std::vector<std::vector<int32_t>> vf_;
  std::vector<Quadric> quad_;
  std::priority_queue<Cand> heap_;
  double scale2_;

  // Neighbours of v (distinct, ascending) from its live faces.
  void neighbours(int32_t v, std::vector<int32_t> &out) const {
    out.clear();
    for (int32_t f : vf_[v])
      for (int k = 0; k < 3; ++k) {
        const int32_t w = faces_[3 * f + k];
        if (w != v) out.push_back(w);
      }
    std::sort(out.begin(), out.end());
    out.erase(std::unique(out.begin(), out.end()), out.end());
  }

  void add_boundary_quadrics() {
    // An edge used by one face is a boundary edge: count uses per (min, max) pair.
    std::vector<std::pair<int64_t, int32_t>> edges;  // key, face
    edges.reserve(3 * nf_);
    for (int64_t f = 0; f < nf_; ++f)
      for (int k = 0; k < 3; ++k) {
        int32_t a = faces_[3 * f + k], b = faces_[3 * f + (k + 1) % 3];
        if (a > b) std::swap(a, b);
        edges.push_back({(int64_t)a * nv_ + b, (int32_t)f});
      }
    std::sort(edges.begin(), edges.end());
    for (size_t i = 0; i < edges.size();) {
      size_t j = i;
      while (j < edges.size() && edges[j].first == edges[i].first) ++j;
      if (j - i == 1) {
        const int32_t a = (int32_t)(edges[i].first / nv_), b = (int32_t)(edges[i].first % nv_);
        const int64_t f = edges[i].second;
        double e[3], e1[3], e2[3], n[3], m[3];
        sub3(&pos_[3 * b], &pos_[3 * a], e);
        sub3(&pos_[3 * faces_[3 * f + 1]], &pos_[3 * faces_[3 * f]], e1);
        sub3(&pos_[3 * faces_[3 * f + 2]], &pos_[3 * faces_[3 * f]], e2);
        cross3(e1, e2, n);
        cross3(e, n, m);  // in the face's plane, perpendicular to the edge
        const double len = std::sqrt(dot3(m, m));
        if (len > 0) {
          const double A = m[0] / len, B = m[1] / len, C = m[2] / len;
          const double D = -(A * pos_[3 * a] + B * pos_[3 * a + 1] + C * pos_[3 * a + 2]);
          const double w = 1000.0 * dot3(e, e);
          quad_[a].add_plane(A, B, C, D, w);
          quad_[b].add_plane(A, B, C, D, w);
        }
      }
      i = j;
    }
  }

  // Optimal position and its cost for collapsing edge (a, b).
  void place(int32_t a, int32_t b, double pos[3], double *cost) const {
    Quadric q = quad_[a];
    q.add(quad_[b]);
    const double *Q = q.q;
    // Solve [[q0 q1 q2],[q1 q4 q5],[q2 q5 q7]] x = -[q3 q6 q8] by Cramer's rule.
    const double a00 = Q[0], a01 = Q[1], a02 = Q[2], a11 = Q[4], a12 = Q[5], a22 = Q[7];
    const double b0 = -Q[3], b1 = -Q[6], b2 = -Q[8];
    const double c00 = a11 * a22 - a12 * a12, c01 = a02 * a12 - a01 * a22, c02 = a01 * a12 - a02 * a11;
    const double det = a00 * c00 + a01 * c01 + a02 * c02;
    const double tr = a00 + a11 + a22;
    bool ok = std::fabs(det) > 1e-10 * tr * tr * tr && tr > 0;
    if (ok) {
      const double c11 = a00 * a22 - a02 * a02, c12 = a01 * a02 - a00 * a12, c22 = a00 * a11 - a01 * a01;
      pos[0] = (c00 * b0 + c01 * b1 + c02 * b2) / det;
      pos[1] = (c01 * b0 + c11 * b1 + c12 * b2) / det;
      pos[2] = (c02 * b0 + c12 * b1 + c22 * b2) / det;
      // keep the optimum near the edge (a far one means an ill-posed system)
      double mid[3], dv[3], e[3];
      for (int c = 0; c < 3; ++c) mid[c] = 0.5 * (pos_[3 * a + c] + pos_[3 * b + c]);
      sub3(pos, mid, dv);
      sub3(&pos_[3 * b], &pos_[3 * a], e);
      ok = dot3(dv, dv) <= 4.0 * dot3(e, e) + 1e-30 * scale2_;
    }
    if (ok) {
      *cost = q.eval(pos);
      return;
    }
    double best = INFINITY;
    const double *pa = &pos_[3 * a], *pb = &pos_[3 * b];
    for (int k = 0; k < 3; ++k) {
      double p[3];
      for (int c = 0; c < 3; ++c) p[c] = k == 0 ? pa[c] : (k == 1 ? pb[c] : 0.5 * (pa[c] + pb[c]));
      const double e = q.eval(p);
      if (e < best) {
        best = e;
        std::memcpy(pos, p, sizeof(p));
      }
    }
    *cost = best;
  }

  void push_edges_of(int32_t v, bool only_higher) {
    std::vector<int32_t> nb;
    neighbours(v, nb);
    for (int32_t w : nb) {
      if (only_higher && w < v) continue;
      Cand c;
      c.a = std::min(v, w);
      c.b = std::max(v, w);
      c.sa = stamp_[c.a];
      c.sb = stamp_[c.b];
      place(c.a, c.b, c.pos, &c.cost);
      c.cost = std::fmax(c.cost, 0.0);
      heap_.push(c);
    }
  }

  bool collapse(const Cand &c) {
    const int32_t a = c.a, b = c.b;
    // Link condition: common neighbours of a and b = the third vertices of the faces on ab.
    std::vector<int32_t> na, nb2, common;
    neighbours(a, na);
    neighbours(b, nb2);
    std::set_intersection(na.begin(), na.end(), nb2.begin(), nb2.end(), std::back_inserter(common));
    int shared = 0;
    for (int32_t f : vf_[a]) {
      const int32_t *t = &faces_[3 * f];
      if (t[0] == b || t[1] == b || t[2] == b) ++shared;
    }
    if (shared < 1 || (int)common.size() != shared) return false;
    // No face around a or b (other than those on ab) may flip or degenerate.
    for (int side = 0; side < 2; ++side) {
      const int32_t v = side ? b : a, o = side ? a : b;
      for (int32_t f : vf_[v]) {
        const int32_t *t = &faces_[3 * f];
        if (t[0] == o || t[1] == o || t[2] == o) continue;
        double p[3][3], q[3][3];
        for (int k = 0; k < 3; ++k)
          for (int cc = 0; cc < 3; ++cc) {
            p[k][cc] = pos_[3 * t[k] + cc];
            q[k][cc] = t[k] == v ? c.pos[cc] : p[k][cc];
          }
        double e1[3], e2[3], n0[3], n1[3];
        sub3(p[1], p[0], e1);
        sub3(p[2], p[0], e2);
        cross3(e1, e2, n0);
        sub3(q[1], q[0], e1);
        sub3(q[2], q[0], e2);
        cross3(e1, e2, n1);
        const double l0 = std::sqrt(dot3(n0, n0)), l1 = std::sqrt(dot3(n1, n1));
        if (!(l1 > 1e-12 * scale2_)) return false;            // degenerate
        if (dot3(n0, n1) < 0.2 * l0 * l1) return false;        // flipped or folded
      }
    }
    // Collapse b into a at the optimal position.
    for (int32_t f : vf_[b]) {
      int32_t *t = &faces_[3 * f];
      if (t[0] == a || t[1] == a || t[2] == a) {
        face_alive_[f] = 0;  // the faces on ab disappear
        for (int k = 0; k < 3; ++k) {
          const int32_t w = t[k];
          if (w == b) continue;
          auto &l = vf_[w];
          l.erase(std::remove(l.begin(), l.end(), f), l.end());
        }
      } else {
        for (int k = 0; k < 3; ++k)
          if (t[k] == b) t[k] = a;
        vf_[a].push_back(f);
      }
    }
    vf_[b].clear();
    vert_alive_[b] = 0;
    std::memcpy(&pos_[3 * a], c.pos, sizeof(c.pos));
    quad_[a].add(quad_[b]);
    // Only the edges at a change cost (its quadric and position did); the others' stay
    // valid and are re-checked when popped.
    ++stamp_[a];
    ++stamp_[b];
    push_edges_of(a, false);
    return true;
  }
};

}  // namespace

extern "C" int lddmm_decimate_surface(const double *vertices, int64_t nv, const int32_t *faces,
                                      int64_t nf, int64_t target, double *out_vertices,
                                      int32_t *out_faces, int64_t *nv_out, int64_t *nf_out) {
  if (!vertices || !faces || !out_vertices || !out_faces || !nv_out || !nf_out || nv < 3 ||
      nf < 1 || target < 3 || nv > INT32_MAX)
    return LDDMM_EINVAL;
  for (int64_t i = 0; i < 3 * nf; ++i)
    if (faces[i] < 0 || faces[i] >= nv) return LDDMM_EINVAL;
  for (int64_t i = 0; i < 3 * nv; ++i)
    if (!std::isfinite(vertices[i])) return LDDMM_EINVAL;
  Decimator d(vertices, nv, faces, nf);
  d.run(target);
  d.output(out_vertices, out_faces, nv_out, nf_out);
  return LDDMM_OK;
}
