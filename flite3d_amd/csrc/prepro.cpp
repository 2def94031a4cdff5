// Grid preprocessor: median-dual side structure and agglomerated multigrid
// levels (the data FLITE3D's Gen3d writes to <case>.sol_1).
//
// Host-side C++, O(N log N).  The declarations and the reference routines each
// part restates are listed in include/flite3d_prepro.h.
#include "../../include/flite3d_prepro.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

// optional progress report after each phase of a build (fp_set_progress)
void (*g_progress)(const char *) = nullptr;
void progress(const std::string &msg) {
  if (g_progress) g_progress(msg.c_str());
}

struct Level {
  int n_nodes = 0, n_bnodes = 0;
  std::vector<int> e1, e2;       // sides, e1 < e2
  std::vector<double> w;         // 3 per side, oriented e1 -> e2
  std::vector<int> b1, b2, bind; // boundary sides (b1 <= b2)
  std::vector<double> wb;        // 3 per boundary side (outward)
  std::vector<double> vol;       // dual volume
  std::vector<double> xyz;       // 3 per node (fine level only)
  std::vector<double> wdist;
  std::vector<int> wnode;
  std::vector<int> node_ind;     // indicator of boundary node i < n_bnodes
  std::vector<int> seed;         // fine seed node (coarse levels)
  std::vector<int> f2c;          // map from the next finer level
  std::vector<double> rweight;   // V_fine / V_coarse
};

}  // namespace

struct fp_mesh {
  std::vector<Level> lev;
  std::vector<int> old_index;  // fine new -> input index
  // kept for fp_metric_vjp only: positively oriented tets and, per boundary
  // triangle, its three boundary sides (fine numbering)
  bool keep_cells = false;
  std::vector<std::array<int, 4>> tets;
  std::vector<int> bface, bface_side;  // 3 vertices, 3 boundary sides per triangle
};

namespace {

// Boundary node indicator priority, gen3d_boundary.f90:1606-1627 (fine) and
// :1127-1149 (coarse): symmetry (2) and inviscid wall (9) lowest, farfield
// and engine faces above wall, engine faces above farfield.
int merge_indicator(int cur, int incoming) {
  if (incoming == cur) return cur;
  if (cur <= 0) return incoming;
  if (cur == 9) return incoming;
  if (cur == 2 && incoming != 9) return incoming;
  if (cur == 1 && incoming > 2) return incoming;
  if (cur <= 4 && incoming > 4) return incoming;
  return cur;
}

inline void cross(const double *a, const double *b, double *c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

// Sorted adjacency lookup: out-neighbours (j > i) of each node.
struct EdgeTable {
  std::vector<long long> start;  // n+1
  std::vector<int> nbr;          // j for edge ids start[i]..start[i+1]-1
  int find(int a, int b) const {  // a < b
    auto lo = nbr.begin() + start[a], hi = nbr.begin() + start[a + 1];
    auto it = std::lower_bound(lo, hi, b);
    if (it == hi || *it != b) return -1;
    return int(it - nbr.begin());
  }
};

// kd-tree nearest neighbour over a point set (wall nodes).
// Nearest point by a k-d tree over an implicit median layout: the subtree
// over idx[lo, hi) keeps its point at mid = (lo + hi) / 2, splits on the
// axis of its largest extent, and carries the bounding box of its points,
// so a query far from a nearly flat point set (a wall seen from the far
// field) prunes by the full 3-D distance to each box.  Exact: among points
// at the smallest distance the smallest index wins, whatever the visiting
// order.
struct KdTree {
  std::vector<int> idx;
  std::vector<unsigned char> axis;  // split axis per subtree (at its mid)
  std::vector<double> box;          // 6 per subtree (at its mid): min xyz, max xyz
  const double *p = nullptr;
  void build(const double *pts, std::vector<int> ids) {
    p = pts;
    idx = std::move(ids);
    axis.assign(idx.size(), 0);
    box.assign(6 * idx.size(), 0.0);
    build_rec(0, int(idx.size()));
  }
  void build_rec(int lo, int hi) {
    if (hi <= lo) return;
    int mid = (lo + hi) / 2;
    double b[6] = {1e300, 1e300, 1e300, -1e300, -1e300, -1e300};
    for (int k = lo; k < hi; ++k)
      for (int c = 0; c < 3; ++c) {
        double v = p[3 * size_t(idx[k]) + c];
        b[c] = std::min(b[c], v);
        b[3 + c] = std::max(b[3 + c], v);
      }
    int ax = 0;
    for (int c = 1; c < 3; ++c)
      if (b[3 + c] - b[c] > b[3 + ax] - b[ax]) ax = c;
    std::copy(b, b + 6, box.begin() + 6 * size_t(mid));
    axis[mid] = (unsigned char)ax;
    if (hi - lo <= 1) return;
    std::nth_element(idx.begin() + lo, idx.begin() + mid, idx.begin() + hi,
                     [&](int a, int b2) { return p[3 * size_t(a) + ax] < p[3 * size_t(b2) + ax]; });
    build_rec(lo, mid);
    build_rec(mid + 1, hi);
  }
  void query(const double *q, int lo, int hi, double &best, int &besti) const {
    if (hi <= lo) return;
    int mid = (lo + hi) / 2;
    const double *b = box.data() + 6 * size_t(mid);
    double lb = 0.0;
    for (int c = 0; c < 3; ++c) {
      double d = q[c] < b[c] ? b[c] - q[c] : (q[c] > b[3 + c] ? q[c] - b[3 + c] : 0.0);
      lb += d * d;
    }
    if (lb > best) return;
    const double *c = p + 3 * size_t(idx[mid]);
    double d2 = (c[0] - q[0]) * (c[0] - q[0]) + (c[1] - q[1]) * (c[1] - q[1]) +
                (c[2] - q[2]) * (c[2] - q[2]);
    if (d2 < best || (d2 == best && idx[mid] < besti)) {
      best = d2;
      besti = idx[mid];
    }
    int ax = axis[mid];
    double diff = q[ax] - c[ax];
    int nlo = diff < 0 ? lo : mid + 1, nhi = diff < 0 ? mid : hi;
    int flo = diff < 0 ? mid + 1 : lo, fhi = diff < 0 ? hi : mid;
    query(q, nlo, nhi, best, besti);
    query(q, flo, fhi, best, besti);
  }
};

void build_fine(fp_mesh &m, int n, const double *xyz_in, int nt,
                const int *tets_in, int nb, const int *btri_in,
                const int *bind_in) {
  for (int t = 0; t < 4 * nt; ++t)
    if (tets_in[t] < 0 || tets_in[t] >= n) throw std::runtime_error("tet node index out of range");
  for (int t = 0; t < 3 * nb; ++t)
    if (btri_in[t] < 0 || btri_in[t] >= n) throw std::runtime_error("boundary node index out of range");

  // boundary nodes first (input order), then interior (input order);
  // Gen3d renumbers the same way (gen3d_grid.f90:127-383).
  std::vector<char> isb(n, 0);
  for (int t = 0; t < 3 * nb; ++t) isb[btri_in[t]] = 1;
  std::vector<int> o2n(n), n2o(n);
  int k = 0;
  for (int i = 0; i < n; ++i)
    if (isb[i]) { o2n[i] = k; n2o[k++] = i; }
  int nbn = k;
  for (int i = 0; i < n; ++i)
    if (!isb[i]) { o2n[i] = k; n2o[k++] = i; }
  m.old_index = n2o;

  Level L;
  L.n_nodes = n;
  L.n_bnodes = nbn;
  L.xyz.resize(3 * size_t(n));
  for (int i = 0; i < n; ++i)
    for (int c = 0; c < 3; ++c) L.xyz[3 * size_t(i) + c] = xyz_in[3 * size_t(n2o[i]) + c];
  const double *X = L.xyz.data();

  std::vector<std::array<int, 4>> tets(nt);
  for (int t = 0; t < nt; ++t) {
    std::array<int, 4> q;
    for (int c = 0; c < 4; ++c) q[c] = o2n[tets_in[4 * size_t(t) + c]];
    double a[3], b[3], cc[3], r[3];
    for (int c = 0; c < 3; ++c) {
      a[c] = X[3 * q[1] + c] - X[3 * q[0] + c];
      b[c] = X[3 * q[2] + c] - X[3 * q[0] + c];
      cc[c] = X[3 * q[3] + c] - X[3 * q[0] + c];
    }
    cross(a, b, r);
    double det = r[0] * cc[0] + r[1] * cc[1] + r[2] * cc[2];
    if (det == 0.0) throw std::runtime_error("degenerate tetrahedron");
    if (det < 0) std::swap(q[2], q[3]);
    tets[t] = q;
  }

  progress("fine level: " + std::to_string(nt) + " tets oriented");
  // unique edges: out-neighbour lists from node->tet incidence
  std::vector<long long> tstart(n + 1, 0);
  for (auto &q : tets) for (int c = 0; c < 4; ++c) tstart[q[c] + 1]++;
  for (int i = 0; i < n; ++i) tstart[i + 1] += tstart[i];
  std::vector<int> tlist(tstart[n]);
  {
    std::vector<long long> pos(tstart.begin(), tstart.end() - 1);
    for (int t = 0; t < nt; ++t) for (int c = 0; c < 4; ++c) tlist[pos[tets[t][c]]++] = t;
  }
  EdgeTable et;
  et.start.assign(n + 1, 0);
  {
    std::vector<int> tmp;
    std::vector<std::vector<int>> outs;  // built per node then flattened
    std::vector<int> flat;
    flat.reserve(size_t(n) * 8);
    for (int i = 0; i < n; ++i) {
      tmp.clear();
      for (long long s = tstart[i]; s < tstart[i + 1]; ++s)
        for (int c = 0; c < 4; ++c) {
          int j = tets[tlist[s]][c];
          if (j > i) tmp.push_back(j);
        }
      std::sort(tmp.begin(), tmp.end());
      tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
      et.start[i + 1] = et.start[i] + (long long)tmp.size();
      flat.insert(flat.end(), tmp.begin(), tmp.end());
    }
    et.nbr = std::move(flat);
  }
  const long long ne = et.start[n];
  if (ne > 2000000000LL) throw std::runtime_error("too many edges for int32 indexing");
  L.e1.resize(ne);
  L.e2.resize(ne);
  for (int i = 0; i < n; ++i)
    for (long long s = et.start[i]; s < et.start[i + 1]; ++s) { L.e1[s] = i; L.e2[s] = et.nbr[s]; }
  L.w.assign(3 * size_t(ne), 0.0);
  L.vol.assign(n, 0.0);

  progress("fine level: " + std::to_string(ne) + " sides");
  // median-dual coefficients, gen3d_grid.f90:2092-2332 with the per-side
  // formula of registerSideGCL (gen3d_grid.f90:786-855): for edge (a,b) of
  // a positively oriented tet with (a,b,c,d) an even permutation,
  //   C = 1/2 (g-m) x (f_abd-m) + 1/2 (f_abc-m) x (g-m),
  // side coefficient += C/2 oriented from the lower to the higher node,
  // each end node gets (m - x_a).C / 3 of dual volume.
  static const int perm[6][4] = {{0, 1, 2, 3}, {0, 2, 3, 1}, {0, 3, 1, 2},
                                 {1, 2, 0, 3}, {1, 3, 2, 0}, {2, 3, 0, 1}};
  for (int t = 0; t < nt; ++t) {
    const auto &q = tets[t];
    double g[3];
    for (int c = 0; c < 3; ++c)
      g[c] = 0.25 * (X[3 * q[0] + c] + X[3 * q[1] + c] + X[3 * q[2] + c] + X[3 * q[3] + c]);
    for (int e = 0; e < 6; ++e) {
      int a = q[perm[e][0]], b = q[perm[e][1]], c3 = q[perm[e][2]], d = q[perm[e][3]];
      double mm[3], fabc[3], fabd[3], r1[3], r2[3], r3[3], r4[3], c1[3], c2[3], C[3], rv[3];
      for (int c = 0; c < 3; ++c) {
        mm[c] = 0.5 * (X[3 * a + c] + X[3 * b + c]);
        fabc[c] = (X[3 * a + c] + X[3 * b + c] + X[3 * c3 + c]) / 3.;
        fabd[c] = (X[3 * a + c] + X[3 * b + c] + X[3 * d + c]) / 3.;
        r1[c] = g[c] - mm[c];
        r2[c] = fabd[c] - mm[c];
        r3[c] = fabc[c] - mm[c];
        r4[c] = g[c] - mm[c];
        rv[c] = mm[c] - X[3 * a + c];
      }
      cross(r1, r2, c1);
      cross(r3, r4, c2);
      for (int c = 0; c < 3; ++c) C[c] = 0.5 * c1[c] + 0.5 * c2[c];
      double nv = (rv[0] * C[0] + rv[1] * C[1] + rv[2] * C[2]) / 3.;
      L.vol[a] += nv;
      L.vol[b] += nv;
      int lo = std::min(a, b), hi = std::max(a, b);
      int eid = et.find(lo, hi);
      double sg = (a == lo) ? 1.0 : -1.0;
      for (int c = 0; c < 3; ++c) L.w[3 * size_t(eid) + c] += sg * 0.5 * C[c];
    }
  }

  progress("fine level: median-dual coefficients and volumes");
  // boundary node indicators in triangle order (gen3d_boundary.f90:1602-1627)
  L.node_ind.assign(nbn, 0);
  for (int f = 0; f < nb; ++f)
    for (int c = 0; c < 3; ++c) {
      int i = o2n[btri_in[3 * f + c]];
      L.node_ind[i] = merge_indicator(L.node_ind[i], bind_in[f]);
    }

  // boundary sides, makeCalcDataForBoundary (gen3d_boundary.f90:511-620) and
  // registerBoundarySideGCL (:289-372): each triangle gives 1/4 of the area
  // vector of one (vertex, edge-midpoint, centroid) sub-triangle to each of
  // its edges; the side keeps the indicator of the first face creating it.
  {
    std::vector<std::vector<std::pair<int, int>>> bsearch(nbn);  // (hi, side id) per lo
    for (int f = 0; f < nb; ++f) {
      int p[3];
      for (int c = 0; c < 3; ++c) p[c] = o2n[btri_in[3 * f + c]];
      if (m.keep_cells) m.bface.insert(m.bface.end(), p, p + 3);
      const double *x1 = X + 3 * p[0], *x2 = X + 3 * p[1], *x3 = X + 3 * p[2];
      double xm[3], xb1[3], xb2[3], xb3[3];
      for (int c = 0; c < 3; ++c) {
        xm[c] = (x1[c] + x2[c] + x3[c]) / 3.;
        xb1[c] = 0.5 * (x1[c] + x2[c]);
        xb2[c] = 0.5 * (x2[c] + x3[c]);
        xb3[c] = 0.5 * (x3[c] + x1[c]);
      }
      struct { int a, b; double r1[3], r2[3]; } sides[3];
      for (int c = 0; c < 3; ++c) {
        sides[0].r1[c] = xm[c] - x1[c];  sides[0].r2[c] = xb1[c] - x1[c];
        sides[1].r1[c] = xb3[c] - x1[c]; sides[1].r2[c] = xm[c] - x1[c];
        sides[2].r1[c] = xm[c] - x2[c];  sides[2].r2[c] = xb2[c] - x2[c];
      }
      sides[0].a = p[0]; sides[0].b = p[1];
      sides[1].a = p[0]; sides[1].b = p[2];
      sides[2].a = p[1]; sides[2].b = p[2];
      for (auto &s : sides) {
        double cx[3];
        cross(s.r1, s.r2, cx);
        int lo = std::min(s.a, s.b), hi = std::max(s.a, s.b);
        int sid = -1;
        for (auto &pr : bsearch[lo]) if (pr.first == hi) { sid = pr.second; break; }
        if (sid < 0) {
          sid = int(L.b1.size());
          bsearch[lo].push_back({hi, sid});
          L.b1.push_back(lo);
          L.b2.push_back(hi);
          L.bind.push_back(bind_in[f]);
          L.wb.push_back(0.0); L.wb.push_back(0.0); L.wb.push_back(0.0);
        }
        for (int c = 0; c < 3; ++c) L.wb[3 * size_t(sid) + c] += 0.25 * (0.5 * cx[c]);
        if (m.keep_cells) m.bface_side.push_back(sid);
      }
    }
  }

  progress("fine level: " + std::to_string(L.b1.size()) + " boundary sides");
  // wall distance: distance to the nearest node of a wall (indicator 1) face
  {
    std::vector<char> iswall(n, 0);
    for (int f = 0; f < nb; ++f)
      if (bind_in[f] == 1)
        for (int c = 0; c < 3; ++c) iswall[o2n[btri_in[3 * f + c]]] = 1;
    std::vector<int> wids;
    for (int i = 0; i < n; ++i) if (iswall[i]) wids.push_back(i);
    L.wdist.assign(n, 0.0);
    L.wnode.assign(n, 0);
    if (!wids.empty()) {
      KdTree kd;
      kd.build(X, wids);
      for (int i = 0; i < n; ++i) {
        if (iswall[i]) { L.wdist[i] = 0.0; L.wnode[i] = i; continue; }
        double best = 1e300;
        int bi = -1;
        kd.query(X + 3 * size_t(i), 0, int(kd.idx.size()), best, bi);
        L.wdist[i] = std::sqrt(best);
        L.wnode[i] = bi;
      }
    } else {
      std::fill(L.wdist.begin(), L.wdist.end(), 1.0e10);
      for (int i = 0; i < n; ++i) L.wnode[i] = i;
    }
  }
  L.seed.resize(n);
  std::iota(L.seed.begin(), L.seed.end(), 0);
  if (m.keep_cells) m.tets = std::move(tets);
  m.lev.push_back(std::move(L));
}

// Transpose of the metric computation of build_fine: given gw = dF/dw per
// side and gwb = dF/dw_b per boundary side (column-major (E,3), (Eb,3),
// fine numbering), accumulate gx = dF/dx (3 per node, fine numbering).  The
// median-dual coefficients are sums of cross products of differences of
// vertex, edge-midpoint, face-centroid and tet-centroid positions; the
// transpose of G.(u x v) is (v x G) for u and (G x u) for v.
// transpose of the dual volumes: each vertex of a positively oriented tet
// holds a quarter of its volume v = e1.(e2 x e3)/6, e_k = x_k - x_0, whose
// derivatives are (e2 x e3)/6, (e3 x e1)/6, (e1 x e2)/6 and minus their sum
void volume_vjp(const fp_mesh &m, const double *gvol, double *gx) {
  const double *X = m.lev[0].xyz.data();
  for (const auto &q : m.tets) {
    double e1[3], e2[3], e3[3], d1[3], d2[3], d3[3];
    for (int c = 0; c < 3; ++c) {
      e1[c] = X[3 * q[1] + c] - X[3 * q[0] + c];
      e2[c] = X[3 * q[2] + c] - X[3 * q[0] + c];
      e3[c] = X[3 * q[3] + c] - X[3 * q[0] + c];
    }
    cross(e2, e3, d1);
    cross(e3, e1, d2);
    cross(e1, e2, d3);
    double gv = 0.25 * (gvol[q[0]] + gvol[q[1]] + gvol[q[2]] + gvol[q[3]]) / 6.0;
    for (int c = 0; c < 3; ++c) {
      gx[3 * size_t(q[1]) + c] += gv * d1[c];
      gx[3 * size_t(q[2]) + c] += gv * d2[c];
      gx[3 * size_t(q[3]) + c] += gv * d3[c];
      gx[3 * size_t(q[0]) + c] -= gv * (d1[c] + d2[c] + d3[c]);
    }
  }
}

void metric_vjp(const fp_mesh &m, const double *gw, const double *gwb, double *gx) {
  const Level &L = m.lev[0];
  const double *X = L.xyz.data();
  const size_t ne = L.e1.size(), nb = L.b1.size();
  static const int perm[6][4] = {{0, 1, 2, 3}, {0, 2, 3, 1}, {0, 3, 1, 2},
                                 {1, 2, 0, 3}, {1, 3, 2, 0}, {2, 3, 0, 1}};
  // sides are sorted by (e1, e2): the same lookup as build_fine's table
  std::vector<long long> start(size_t(L.n_nodes) + 1, 0);
  for (size_t e = 0; e < ne; ++e) start[L.e1[e] + 1]++;
  for (int i = 0; i < L.n_nodes; ++i) start[i + 1] += start[i];
  auto find = [&](int lo, int hi) {
    auto first = L.e2.begin() + start[lo], last = L.e2.begin() + start[lo + 1];
    return int(std::lower_bound(first, last, hi) - L.e2.begin());
  };
  for (const auto &q : m.tets) {
    double g[3];
    for (int c = 0; c < 3; ++c)
      g[c] = 0.25 * (X[3 * q[0] + c] + X[3 * q[1] + c] + X[3 * q[2] + c] + X[3 * q[3] + c]);
    double dg[3] = {0, 0, 0};
    for (int e = 0; e < 6; ++e) {
      int a = q[perm[e][0]], b = q[perm[e][1]], c3 = q[perm[e][2]], d = q[perm[e][3]];
      int lo = std::min(a, b), hi = std::max(a, b);
      int eid = find(lo, hi);
      double sg = (a == lo) ? 1.0 : -1.0;
      // w_e += sg C / 2 with C = (r1 x r2 + r3 x r4) / 2,
      // r1 = r4 = g - m, r2 = f_abd - m, r3 = f_abc - m
      double G[3], r1[3], r2[3], r3[3];
      for (int c = 0; c < 3; ++c) {
        G[c] = sg * 0.25 * gw[c * ne + eid];
        double mm = 0.5 * (X[3 * a + c] + X[3 * b + c]);
        r1[c] = g[c] - mm;
        r2[c] = (X[3 * a + c] + X[3 * b + c] + X[3 * d + c]) / 3. - mm;
        r3[c] = (X[3 * a + c] + X[3 * b + c] + X[3 * c3 + c]) / 3. - mm;
      }
      double d1[3], d2[3], d3[3], d4[3];
      cross(r2, G, d1);  // d/dr1
      cross(G, r1, d2);  // d/dr2
      cross(r1, G, d3);  // d/dr3 (r4 = r1)
      cross(G, r3, d4);  // d/dr4
      for (int c = 0; c < 3; ++c) {
        double dm = -(d1[c] + d2[c] + d3[c] + d4[c]);
        dg[c] += d1[c] + d4[c];
        gx[3 * size_t(a) + c] += 0.5 * dm + (d2[c] + d3[c]) / 3.;
        gx[3 * size_t(b) + c] += 0.5 * dm + (d2[c] + d3[c]) / 3.;
        gx[3 * size_t(c3) + c] += d3[c] / 3.;
        gx[3 * size_t(d) + c] += d2[c] / 3.;
      }
    }
    for (int k = 0; k < 4; ++k)
      for (int c = 0; c < 3; ++c) gx[3 * size_t(q[k]) + c] += 0.25 * dg[c];
  }
  // boundary triangles: w_b(side) += (r1 x r2) / 8 for the three
  // (vertex, edge midpoint, centroid) sub-triangles, as build_fine
  const size_t nf = m.bface.size() / 3;
  for (size_t f = 0; f < nf; ++f) {
    const int p[3] = {m.bface[3 * f], m.bface[3 * f + 1], m.bface[3 * f + 2]};
    const double *x1 = X + 3 * p[0], *x2 = X + 3 * p[1], *x3 = X + 3 * p[2];
    double xm[3], xb1[3], xb2[3], xb3[3];
    for (int c = 0; c < 3; ++c) {
      xm[c] = (x1[c] + x2[c] + x3[c]) / 3.;
      xb1[c] = 0.5 * (x1[c] + x2[c]);
      xb2[c] = 0.5 * (x2[c] + x3[c]);
      xb3[c] = 0.5 * (x3[c] + x1[c]);
    }
    // gradients of the sub-triangles' points: vertex x1, x2, x3, centroid, midpoints b1, b2, b3
    double dx1[3] = {0, 0, 0}, dx2[3] = {0, 0, 0}, dxm[3] = {0, 0, 0};
    double db1[3] = {0, 0, 0}, db2[3] = {0, 0, 0}, db3[3] = {0, 0, 0};
    for (int k = 0; k < 3; ++k) {
      int sid = m.bface_side[3 * f + k];
      double G[3], r1[3], r2[3], d1[3], d2[3];
      for (int c = 0; c < 3; ++c) G[c] = 0.125 * gwb[c * nb + sid];
      for (int c = 0; c < 3; ++c) {
        if (k == 0) { r1[c] = xm[c] - x1[c]; r2[c] = xb1[c] - x1[c]; }
        else if (k == 1) { r1[c] = xb3[c] - x1[c]; r2[c] = xm[c] - x1[c]; }
        else { r1[c] = xm[c] - x2[c]; r2[c] = xb2[c] - x2[c]; }
      }
      cross(r2, G, d1);
      cross(G, r1, d2);
      for (int c = 0; c < 3; ++c) {
        if (k == 0) { dxm[c] += d1[c]; db1[c] += d2[c]; dx1[c] -= d1[c] + d2[c]; }
        else if (k == 1) { db3[c] += d1[c]; dxm[c] += d2[c]; dx1[c] -= d1[c] + d2[c]; }
        else { dxm[c] += d1[c]; db2[c] += d2[c]; dx2[c] -= d1[c] + d2[c]; }
      }
    }
    for (int c = 0; c < 3; ++c) {
      gx[3 * size_t(p[0]) + c] += dx1[c] + dxm[c] / 3. + 0.5 * (db1[c] + db3[c]);
      gx[3 * size_t(p[1]) + c] += dx2[c] + dxm[c] / 3. + 0.5 * (db1[c] + db2[c]);
      gx[3 * size_t(p[2]) + c] += dxm[c] / 3. + 0.5 * (db2[c] + db3[c]);
    }
  }
}

// One agglomeration step, mergeNodes (gen3d_grid.f90:2358-2901) with
// directionalityParameter = 0 (gen3d_generator.f90:126): seeds are taken in
// order of decreasing available connectivity, boundary nodes first; a seed
// absorbs every neighbour not yet agglomerated; each absorbed node lowers
// the available connectivity of its free neighbours.
// The initial order is Gen3d's (gen3d_grid.f90:2407-2436).  The updates are
// not: Gen3d keeps the priority as a list in which a node whose available
// connectivity drops to c moves down only to just before the first later
// entry with connectivity <= c (the side's first node) or < c (its second
// node), and entries of agglomerated nodes count as connectivity 0, so they
// stop the move; with no such entry the node stays where it is
// (gen3d_grid.f90:2504-2570).  Here a decremented node goes to the back of
// a FIFO bucket of connectivity c, so a seed always has the largest
// available connectivity left, and seeds of equal connectivity can come in
// another order than Gen3d's.  The fine sides here are sorted by (lower,
// upper) node, where Gen3d numbers them in order of first appearance over
// the elements (gen3d_grid.f90:2110-2330), so a seed's neighbours come in
// another order too.  The boundary-first renumbering below moves whole
// agglomerates; Gen3d's swap (gen3d_grid.f90:2739-2796) exchanges the
// members but not the seeds, whose mapping values are negated at that point.
// Lonesome agglomerates join their one neighbour as in gen3d_grid.f90:
// 2611-2700, and the boundary count is that of the agglomerates holding a
// fine boundary node.
Level agglomerate(const Level &F) {
  const int n = F.n_nodes, nbf = F.n_bnodes;
  const long long ne = (long long)F.e1.size();
  // node -> incident side ids
  std::vector<long long> st(n + 1, 0);
  for (long long e = 0; e < ne; ++e) { st[F.e1[e] + 1]++; st[F.e2[e] + 1]++; }
  for (int i = 0; i < n; ++i) st[i + 1] += st[i];
  std::vector<int> inc(st[n]);
  {
    std::vector<long long> pos(st.begin(), st.end() - 1);
    for (long long e = 0; e < ne; ++e) { inc[pos[F.e1[e]]++] = int(e); inc[pos[F.e2[e]]++] = int(e); }
  }
  auto other = [&](int e, int i) { return F.e1[e] == i ? F.e2[e] : F.e1[e]; };

  std::vector<int> avail(n);
  int maxc = 0;
  for (int i = 0; i < n; ++i) {
    avail[i] = int(st[i + 1] - st[i]);
    if (avail[i] <= 0) throw std::runtime_error("node without sides in agglomeration");
    maxc = std::max(maxc, avail[i]);
  }
  // buckets per segment: intrusive doubly linked lists
  std::vector<int> nxt(n, -1), prv(n, -1);
  std::vector<int> head[2], tail[2];
  for (int s = 0; s < 2; ++s) { head[s].assign(maxc + 1, -1); tail[s].assign(maxc + 1, -1); }
  auto seg = [&](int i) { return i < nbf ? 0 : 1; };
  auto push_back = [&](int i) {
    int s = seg(i), c = avail[i];
    prv[i] = tail[s][c]; nxt[i] = -1;
    if (tail[s][c] >= 0) nxt[tail[s][c]] = i; else head[s][c] = i;
    tail[s][c] = i;
  };
  auto unlink = [&](int i) {
    int s = seg(i), c = avail[i];
    if (prv[i] >= 0) nxt[prv[i]] = nxt[i]; else head[s][c] = nxt[i];
    if (nxt[i] >= 0) prv[nxt[i]] = prv[i]; else tail[s][c] = prv[i];
    prv[i] = nxt[i] = -1;
  };
  // initial order: decreasing connectivity, ties by decreasing index
  // (the order gen3d_grid.f90:2407-2436 builds)
  for (int s = 0; s < 2; ++s) {
    int lo = s == 0 ? 0 : nbf, hi = s == 0 ? nbf : n;
    for (int i = hi - 1; i >= lo; --i) push_back(i);
  }
  std::vector<char> agg(n, 0);
  std::vector<int> map(n, -1);
  std::vector<int> seeds;
  std::vector<char> cbnd;  // coarse node contains a fine boundary node
  for (int s = 0; s < 2; ++s) {
    int c = maxc;
    while (true) {
      while (c >= 0 && head[s][c] < 0) --c;
      if (c < 0) break;
      int seed = head[s][c];
      unlink(seed);
      agg[seed] = 1;
      int cn = int(seeds.size());
      seeds.push_back(seed);
      map[seed] = cn;
      bool onb = seed < nbf;
      for (long long k = st[seed]; k < st[seed + 1]; ++k) {
        int nb = other(inc[k], seed);
        if (agg[nb]) continue;
        unlink(nb);
        agg[nb] = 1;
        map[nb] = cn;
        if (nb < nbf) onb = true;
        for (long long k2 = st[nb]; k2 < st[nb + 1]; ++k2) {
          int y = other(inc[k2], nb);
          if (agg[y]) continue;
          unlink(y);
          avail[y]--;
          if (avail[y] < 0) throw std::runtime_error("negative available connectivity");
          push_back(y);
        }
      }
      cbnd.push_back(onb ? 1 : 0);
      c = maxc;  // decremented nodes can only move down; restart scan at top
    }
  }
  int nc = int(seeds.size());
  // lonesome coarse nodes (exactly one distinct coarse neighbour) join that
  // neighbour (gen3d_grid.f90:2611-2700)
  {
    std::vector<long long> firstside(nc, -1);
    std::vector<char> multi(nc, 0);
    std::vector<int> onlynb(nc, -1);
    for (long long e = 0; e < ne; ++e) {
      int a = map[F.e1[e]], b = map[F.e2[e]];
      if (a == b) continue;
      for (int r = 0; r < 2; ++r) {
        int x = r == 0 ? a : b, y = r == 0 ? b : a;
        if (onlynb[x] < 0) onlynb[x] = y;
        else if (onlynb[x] != y) multi[x] = 1;
      }
    }
    std::vector<int> redirect(nc);
    std::iota(redirect.begin(), redirect.end(), 0);
    for (int c = 0; c < nc; ++c)
      if (!multi[c] && onlynb[c] >= 0) redirect[c] = onlynb[c];
    // a lonesome pair pointing at each other keeps the lower index
    for (int c = 0; c < nc; ++c)
      if (redirect[c] != c && redirect[redirect[c]] != redirect[c]) {
        int o = redirect[c];
        if (c < o) redirect[c] = c; else redirect[o] = o;
      }
    std::vector<int> newid(nc, -1);
    int k = 0;
    for (int c = 0; c < nc; ++c) if (redirect[c] == c) newid[c] = k++;
    std::vector<char> cb2(k, 0);
    std::vector<int> seeds2(k);
    for (int c = 0; c < nc; ++c) {
      int t = newid[redirect[c]];
      if (cbnd[c]) cb2[t] = 1;
      if (redirect[c] == c) seeds2[t] = seeds[c];
    }
    for (int i = 0; i < n; ++i) map[i] = newid[redirect[map[i]]];
    nc = k;
    cbnd.swap(cb2);
    seeds.swap(seeds2);
  }
  // boundary coarse nodes first (stable), gen3d_grid.f90:2733-2770
  std::vector<int> order;
  order.reserve(nc);
  for (int c = 0; c < nc; ++c) if (cbnd[c]) order.push_back(c);
  int nbc = int(order.size());
  for (int c = 0; c < nc; ++c) if (!cbnd[c]) order.push_back(c);
  std::vector<int> renum(nc);
  for (int k = 0; k < nc; ++k) renum[order[k]] = k;
  for (int i = 0; i < n; ++i) map[i] = renum[map[i]];

  Level C;
  C.n_nodes = nc;
  C.n_bnodes = nbc;
  C.f2c = map;
  C.seed.resize(nc);
  for (int k = 0; k < nc; ++k) C.seed[k] = F.seed[seeds[order[k]]];
  C.vol.assign(nc, 0.0);
  for (int i = 0; i < n; ++i) C.vol[map[i]] += F.vol[i];
  C.rweight.resize(n);
  for (int i = 0; i < n; ++i) C.rweight[i] = F.vol[i] / C.vol[map[i]];  // gen3d_grid.f90:3771-3772
  C.wdist.assign(nc, 0.0);
  C.wnode.assign(nc, -1);
  for (int i = 0; i < n; ++i) {  // gen3d_grid.f90:5838-5847
    int c = map[i];
    C.wdist[c] += F.vol[i] * F.wdist[i] / C.vol[c];
    if (C.wnode[c] < 0) C.wnode[c] = map[F.wnode[i]];
  }
  // coarse sides: sums of cut fine sides, createSides/registerSideData
  {
    std::vector<std::pair<long long, int>> keys;  // (lo*nc+hi, fine side)
    keys.reserve(ne);
    for (long long e = 0; e < ne; ++e) {
      int a = map[F.e1[e]], b = map[F.e2[e]];
      if (a == b) continue;
      long long key = (long long)std::min(a, b) * nc + std::max(a, b);
      keys.push_back({key, int(e)});
    }
    std::stable_sort(keys.begin(), keys.end(),
                     [](const std::pair<long long, int> &x, const std::pair<long long, int> &y) { return x.first < y.first; });
    for (size_t s = 0; s < keys.size();) {
      size_t t = s;
      double acc[3] = {0, 0, 0};
      int lo = int(keys[s].first / nc), hi = int(keys[s].first % nc);
      while (t < keys.size() && keys[t].first == keys[s].first) {
        int e = keys[t].second;
        double sg = (map[F.e1[e]] == lo) ? 1.0 : -1.0;
        for (int c = 0; c < 3; ++c) acc[c] += sg * F.w[3 * size_t(e) + c];
        ++t;
      }
      C.e1.push_back(lo);
      C.e2.push_back(hi);
      C.w.push_back(acc[0]); C.w.push_back(acc[1]); C.w.push_back(acc[2]);
      s = t;
    }
  }
  // coarse boundary sides (agglomerateBoundary, gen3d_boundary.f90:922-1000):
  // plain sums of fine coefficients; sides collapsing onto one coarse node
  // are kept (i1 == i2); first fine side sets the indicator.
  {
    std::vector<std::vector<std::pair<int, int>>> bs(nc);
    for (size_t s = 0; s < F.b1.size(); ++s) {
      int a = map[F.b1[s]], b = map[F.b2[s]];
      int lo = std::min(a, b), hi = std::max(a, b);
      int sid = -1;
      for (auto &pr : bs[lo]) if (pr.first == hi) { sid = pr.second; break; }
      if (sid < 0) {
        sid = int(C.b1.size());
        bs[lo].push_back({hi, sid});
        C.b1.push_back(lo); C.b2.push_back(hi); C.bind.push_back(F.bind[s]);
        C.wb.push_back(0); C.wb.push_back(0); C.wb.push_back(0);
      }
      for (int c = 0; c < 3; ++c) C.wb[3 * size_t(sid) + c] += F.wb[3 * s + c];
    }
  }
  // coarse boundary node indicators (gen3d_boundary.f90:1127-1149)
  C.node_ind.assign(nbc, 0);
  for (int i = 0; i < nbf; ++i) {
    int c = map[i];
    if (c >= nbc) throw std::runtime_error("fine boundary node mapped to interior coarse node");
    C.node_ind[c] = merge_indicator(C.node_ind[c], F.node_ind[i]);
  }
  return C;
}

void register_from_ind(const Level &L, int *reg, int *arr) {
  // nodeIndicatorRegister(-20..0): for indicator k (1..10) start at
  // index 2k-2 and end at 2k-1 of reg[] (gen3d_grid.f90:4541-4558)
  int cnt[11] = {0};
  for (int i = 0; i < L.n_bnodes; ++i) {
    int k = L.node_ind[i];
    if (k >= 1 && k <= 10) cnt[k]++;
  }
  int r[21] = {0};
  int end = 0;
  for (int k = 1; k <= 10; ++k) {
    r[2 * k - 2] = end + 1;
    end += cnt[k];
    r[2 * k - 1] = end;
  }
  if (reg) for (int i = 0; i < 21; ++i) reg[i] = r[i];
  if (arr) {
    int pos = 0;
    for (int k = 1; k <= 9; ++k)
      for (int i = 0; i < L.n_bnodes; ++i)
        if (L.node_ind[i] == k) arr[pos++] = i;
  }
}

}  // namespace

extern "C" {

const char *fp_last_error(void) { return g_err.c_str(); }

void fp_set_progress(void (*cb)(const char *msg)) { g_progress = cb; }

fp_mesh *fp_mesh_build(int n_nodes, const double *xyz, int n_tets,
                       const int *tets, int n_btri, const int *btri,
                       const int *btri_ind, int n_levels) {
  try {
    if (n_nodes <= 0 || n_tets <= 0 || n_btri <= 0 || n_levels < 1)
      throw std::runtime_error("bad sizes");
    auto *m = new fp_mesh;
    build_fine(*m, n_nodes, xyz, n_tets, tets, n_btri, btri, btri_ind);
    progress("fine level: wall distance");
    for (int l = 1; l < n_levels; ++l) {
      m->lev.push_back(agglomerate(m->lev.back()));
      progress("level " + std::to_string(l + 1) + ": " + std::to_string(m->lev.back().n_nodes) + " nodes");
    }
    return m;
  } catch (const std::exception &e) {
    g_err = e.what();
    return nullptr;
  }
}

void fp_mesh_free(fp_mesh *m) { delete m; }

int fp_mesh_nlevels(const fp_mesh *m) { return m ? int(m->lev.size()) : -1; }

int fp_level_sizes(const fp_mesh *m, int level, int *nn, int *ns, int *nbs, int *nbn) {
  if (!m || level < 0 || level >= int(m->lev.size())) { g_err = "bad level"; return -1; }
  const Level &L = m->lev[level];
  if (nn) *nn = L.n_nodes;
  if (ns) *ns = int(L.e1.size());
  if (nbs) *nbs = int(L.b1.size());
  if (nbn) *nbn = L.n_bnodes;
  return 0;
}

int fp_level_copy(const fp_mesh *m, int level, int *side_index, double *side_w,
                  int *bside_index, double *bside_w, double *volume,
                  double *coords, double *wall_dist, int *wall_node, int *reg,
                  int *node_ind_array, int *node_ind, int *seed, int *f2c,
                  double *rweight) {
  if (!m || level < 0 || level >= int(m->lev.size())) { g_err = "bad level"; return -1; }
  const Level &L = m->lev[level];
  const size_t n = L.n_nodes, ne = L.e1.size(), nb = L.b1.size();
  if (side_index) for (size_t e = 0; e < ne; ++e) { side_index[e] = L.e1[e]; side_index[ne + e] = L.e2[e]; }
  if (side_w) for (size_t e = 0; e < ne; ++e) for (int c = 0; c < 3; ++c) side_w[c * ne + e] = L.w[3 * e + c];
  if (bside_index) for (size_t s = 0; s < nb; ++s) { bside_index[s] = L.b1[s]; bside_index[nb + s] = L.b2[s]; bside_index[2 * nb + s] = L.bind[s]; }
  if (bside_w) for (size_t s = 0; s < nb; ++s) for (int c = 0; c < 3; ++c) bside_w[c * nb + s] = L.wb[3 * s + c];
  if (volume) std::copy(L.vol.begin(), L.vol.end(), volume);
  if (coords) {
    for (size_t i = 0; i < n; ++i)
      for (int c = 0; c < 3; ++c) coords[c * n + i] = L.xyz.empty() ? 0.0 : L.xyz[3 * i + c];
  }
  if (wall_dist) std::copy(L.wdist.begin(), L.wdist.end(), wall_dist);
  if (wall_node) std::copy(L.wnode.begin(), L.wnode.end(), wall_node);
  register_from_ind(L, reg, node_ind_array);
  if (node_ind) std::copy(L.node_ind.begin(), L.node_ind.end(), node_ind);
  if (seed) std::copy(L.seed.begin(), L.seed.end(), seed);
  if (f2c) { if (level == 0) { g_err = "fine level has no transfer"; return -2; } std::copy(L.f2c.begin(), L.f2c.end(), f2c); }
  if (rweight) { if (level == 0) { g_err = "fine level has no transfer"; return -2; } std::copy(L.rweight.begin(), L.rweight.end(), rweight); }
  return 0;
}

int fp_metric_vjp(int n_nodes, const double *xyz, int n_tets, const int *tets,
                  int n_btri, const int *btri, const int *btri_ind,
                  const double *gw, const double *gwb, double *gx) {
  try {
    if (n_nodes <= 0 || n_tets <= 0 || n_btri <= 0) throw std::runtime_error("bad sizes");
    if (!gw || !gwb || !gx) throw std::runtime_error("null array");
    fp_mesh m;
    m.keep_cells = true;
    build_fine(m, n_nodes, xyz, n_tets, tets, n_btri, btri, btri_ind);
    std::vector<double> g(3 * size_t(n_nodes), 0.0);
    metric_vjp(m, gw, gwb, g.data());
    for (int i = 0; i < n_nodes; ++i)
      for (int c = 0; c < 3; ++c) gx[3 * size_t(m.old_index[i]) + c] = g[3 * size_t(i) + c];
    return 0;
  } catch (const std::exception &e) {
    g_err = e.what();
    return -1;
  }
}

int fp_mesh_vjp(int n_nodes, const double *xyz, int n_tets, const int *tets,
                int n_btri, const int *btri, const int *btri_ind,
                const double *gw, const double *gwb, const double *gvol, const double *gxn, double *gx) {
  try {
    if (n_nodes <= 0 || n_tets <= 0 || n_btri <= 0) throw std::runtime_error("bad sizes");
    if (!gw || !gwb || !gx) throw std::runtime_error("null array");
    fp_mesh m;
    m.keep_cells = true;
    build_fine(m, n_nodes, xyz, n_tets, tets, n_btri, btri, btri_ind);
    std::vector<double> g(3 * size_t(n_nodes), 0.0);
    metric_vjp(m, gw, gwb, g.data());
    if (gvol) volume_vjp(m, gvol, g.data());
    if (gxn)
      for (size_t k = 0; k < g.size(); ++k) g[k] += gxn[k];
    for (int i = 0; i < n_nodes; ++i)
      for (int c = 0; c < 3; ++c) gx[3 * size_t(m.old_index[i]) + c] = g[3 * size_t(i) + c];
    return 0;
  } catch (const std::exception &e) {
    g_err = e.what();
    return -1;
  }
}

int fp_fine_old_index(const fp_mesh *m, int *old_index) {
  if (!m) return -1;
  std::copy(m->old_index.begin(), m->old_index.end(), old_index);
  return 0;
}

}  // extern "C"
