// smg_tri.cpp -- 2D triangulation of a job's particles: Voronoi-face edges from the
// GPU (k_vor2d), then the integer stage of Graphmeshcls::Reconstruct_tris
// (graph_mesh.cpp:37-407) and Tri_quality_statistics (graph_mesh.cpp:411-671) on the
// host.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "smg.h"
#include "smg_engine.hpp"
#include "smg_internal.hpp"
#include "smg_kernels.hpp"

namespace smg {

namespace {

using Tri = std::array<int32_t, 3>;   // i > j > k

struct Graph {
  std::vector<std::vector<int32_t>> adj;   // sorted, unique, no self loops
  explicit Graph(int64_t n) : adj(size_t(n)) {}
  void finish() {
    for (auto& a : adj) {
      std::sort(a.begin(), a.end());
      a.erase(std::unique(a.begin(), a.end()), a.end());
    }
  }
  void add(int32_t a, int32_t b) {
    if (a == b) return;
    adj[a].push_back(b);
    adj[b].push_back(a);
  }
  bool has(int32_t a, int32_t b) const { return std::binary_search(adj[a].begin(), adj[a].end(), b); }
  int64_t n_edges() const {
    int64_t s = 0;
    for (auto& a : adj) s += int64_t(a.size());
    return s / 2;
  }
};

// every 3-clique once, as (i > j > k), in lexicographic order of (i, j, k)
void cliques(const Graph& g, std::vector<Tri>& out) {
  out.clear();
  for (int32_t a = 0; a < int32_t(g.adj.size()); a++) {
    const auto& na = g.adj[a];
    for (size_t x = 0; x < na.size() && na[x] < a; x++)
      for (size_t y = 0; y < x; y++)
        if (g.has(na[x], na[y])) out.push_back({a, na[x], na[y]});
  }
  std::sort(out.begin(), out.end());
}

double dist2d(const double* p, const double* q) {
  const double dx = q[0] - p[0], dy = q[1] - p[1];
  return std::sqrt(dx * dx + dy * dy);
}

// Tri_quality_statistics: angles, edges, areas, aspect ratio and G of every triangle
void quality(const std::vector<Tri>& tris, const std::vector<double>& x, smg_tri_stats* st) {
  const double rad = 180. / std::acos(-1.);
  st->min_angle = 1.e15;
  st->max_angle = 0.;
  st->min_edge = 1.e15;
  st->max_edge = 0.;
  st->min_area = 1.e15;
  st->max_area = 0.;
  st->G_min = 1.e20;
  double g_sum = 0., minang_sum = 0.;
  for (const Tri& t : tris) {
    const double* v0 = &x[3 * size_t(t[0])];
    const double* v1 = &x[3 * size_t(t[1])];
    const double* v2 = &x[3 * size_t(t[2])];
    const double a1[2] = {v1[0] - v0[0], v1[1] - v0[1]};
    const double a2[2] = {v2[0] - v1[0], v2[1] - v1[1]};
    const double a3[2] = {v0[0] - v2[0], v0[1] - v2[1]};
    const double d1 = std::sqrt(a1[0] * a1[0] + a1[1] * a1[1]);
    const double d2 = std::sqrt(a2[0] * a2[0] + a2[1] * a2[1]);
    const double d3 = std::sqrt(a3[0] * a3[0] + a3[1] * a3[1]);
    const double dmax = std::max(d1, std::max(d2, d3)), dmin = std::min(d1, std::min(d2, d3));
    st->max_edge = std::max(st->max_edge, dmax);
    st->min_edge = std::min(st->min_edge, dmin);
    st->max_aspect = std::max(st->max_aspect, dmax / (dmin + 1.e-20));
    double al[3] = {rad * std::acos(-(a1[0] * a3[0] + a1[1] * a3[1]) / (d3 * d1 + 1.e-20)),
                    rad * std::acos(-(a1[0] * a2[0] + a1[1] * a2[1]) / (d2 * d1 + 1.e-20)),
                    rad * std::acos(-(a2[0] * a3[0] + a2[1] * a3[1]) / (d3 * d2 + 1.e-20))};
    double mn = 1.e300, mx = 0.;
    for (double& a : al) {
      a = std::max(0., std::min(a, 180.));
      mn = std::min(mn, a);
      mx = std::max(mx, a);
      st->angle_hist[std::min(44, int(a / 4.0))] += 1;   // 180 degrees goes to the last bin
    }
    st->max_angle = std::max(st->max_angle, mx);
    st->min_angle = std::min(st->min_angle, mn);
    const double lim[4] = {10., 20., 30., 40.};
    for (int q = 0; q < 4; q++) st->n_min_angle_below[q] += mn < lim[q];
    minang_sum += mn;
    const double area =
        0.5 * (d3 * d1) * std::sin(std::acos(-(a1[0] * a3[0] + a1[1] * a3[1]) / (d3 * d1 + 1.e-20)));
    st->total_area += area;
    st->max_area = std::max(st->max_area, area);
    st->min_area = std::min(st->min_area, area);
    const double G = 2. * std::sqrt(3.) * area / (dmax * (d1 + d2 + d3) / 2.);
    st->G_min = std::min(st->G_min, G);
    g_sum += G;
  }
  st->n_tri = int64_t(tris.size());
  if (!tris.empty()) {
    st->avg_min_angle = minang_sum / double(tris.size());
    st->G_avg = g_sum / double(tris.size());
  }
}

}  // namespace

int tri_cliques_host(int64_t nv, const int32_t* edges, int64_t ne, std::vector<std::array<int32_t, 3>>& out) {
  Graph g(nv);
  for (int64_t e = 0; e < ne; e++) {
    const int32_t a = edges[2 * e], b = edges[2 * e + 1];
    if (a < 0 || b < 0 || a >= nv || b >= nv) { set_error("smg_tri_cliques: vertex out of range"); return SMG_ERR_ARG; }
    g.add(a, b);
  }
  g.finish();
  cliques(g, out);
  return SMG_OK;
}

#define HIP_TRY_T(expr)                                                              \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess) {                                                          \
      set_error(std::string(#expr) + ": " + hipGetErrorString(e_));                  \
      rc = SMG_ERR_HIP;                                                              \
      goto done;                                                                     \
    }                                                                                \
  } while (0)

int Engine::triangulate(int job, std::vector<std::array<int32_t, 3>>& tris, smg_tri_stats* st) {
  std::memset(st, 0, sizeof(*st));
  tris.clear();
  if (D_ != 2) { set_error("triangulate: 2D only (the reference meshes 3D with TetGen)"); return SMG_ERR_ARG; }
  if (step_ == 0) {
    // after Pre_run and before any step (the post mode): lists on the placed positions
    if (!prerun_done_) { set_error("triangulate: call prerun (or seed and step) first"); return SMG_ERR_STATE; }
    const int rb = build_lists();
    if (rb != SMG_OK) return rb;
  }
  const JobParams& J = jp_[job];
  const int64_t n = J.n, cap = J.cap;
  const int KE = 32;
  int rc = SMG_OK;
  int32_t* d_nb = nullptr;
  int32_t* d_cnt = nullptr;
  double* d_pts = nullptr;
  double* d_phi = nullptr;
  std::vector<int32_t> nb(size_t(n) * KE), cnt(n), id(cap), ty(cap);
  std::vector<float> xh(cap), xl(cap);
  std::vector<double> x(3 * size_t(n), 0.);
  std::vector<int32_t> type(n);
  std::vector<Tri> t1;
  std::vector<std::pair<int32_t, int32_t>> del;
  Graph g(n);
  {
    HIP_TRY_T(hipStreamSynchronize(st_));
    HIP_TRY_T(hipMalloc(&d_nb, sizeof(int32_t) * n * KE));
    HIP_TRY_T(hipMalloc(&d_cnt, sizeof(int32_t) * n));
    Work Ws = W_;
    Ws.tile0 = J.p0 / kTile;
    launch_vor2d(st_, cap / kTile, Ws, A_, d_nb, d_cnt, KE);
    HIP_TRY_T(hipGetLastError());
    HIP_TRY_T(hipMemcpyAsync(nb.data(), d_nb, sizeof(int32_t) * n * KE, hipMemcpyDeviceToHost, st_));
    HIP_TRY_T(hipMemcpyAsync(cnt.data(), d_cnt, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st_));
    HIP_TRY_T(hipMemcpyAsync(id.data(), A_.id + J.p0, sizeof(int32_t) * cap, hipMemcpyDeviceToHost, st_));
    HIP_TRY_T(hipMemcpyAsync(ty.data(), A_.type + J.p0, sizeof(int32_t) * cap, hipMemcpyDeviceToHost, st_));
    HIP_TRY_T(hipStreamSynchronize(st_));
    for (int a = 0; a < 2; a++) {
      HIP_TRY_T(hipMemcpy(xh.data(), A_.xh[a] + J.p0, sizeof(float) * cap, hipMemcpyDeviceToHost));
      HIP_TRY_T(hipMemcpy(xl.data(), A_.xl[a] + J.p0, sizeof(float) * cap, hipMemcpyDeviceToHost));
      for (int64_t s = 0; s < n; s++) x[3 * size_t(id[s]) + a] = double(xh[s]) + double(xl[s]);
    }
    for (int64_t s = 0; s < n; s++) type[id[s]] = ty[s];
    // Voronoi-face edges (Add_edge keeps one undirected edge per pair)
    for (int64_t s = 0; s < n; s++) {
      if (cnt[s] < 0) {
        st->n_cell_overflow++;
        continue;
      }
      for (int c = 0; c < cnt[s]; c++) g.add(id[s], nb[size_t(s) * KE + c]);
    }
    g.finish();
    st->n_edge = g.n_edges();
    last_edges_.clear();
    for (int32_t a = 0; a < int32_t(n); a++)
      for (int32_t b : g.adj[a])
        if (b < a) last_edges_.push_back({a, b});
    cliques(g, t1);
    // centroid test of the triangles with three non-REAL particles: the particle phi
    // and the centroid phi sampled on the GPU like every other level-set sample
    std::vector<size_t> test;
    for (size_t q = 0; q < t1.size(); q++)
      if (type[t1[q][0]] != SMG_REAL && type[t1[q][1]] != SMG_REAL && type[t1[q][2]] != SMG_REAL) test.push_back(q);
    const int64_t np = n + int64_t(test.size());
    std::vector<double> pts(3 * size_t(np), 0.), ph(np);
    std::copy(x.begin(), x.end(), pts.begin());
    for (size_t u = 0; u < test.size(); u++) {
      const Tri& t = t1[test[u]];
      for (int a = 0; a < 2; a++)
        pts[3 * (n + u) + a] = 1. / 3. * (x[3 * size_t(t[0]) + a] + x[3 * size_t(t[1]) + a] + x[3 * size_t(t[2]) + a]);
    }
    HIP_TRY_T(hipMalloc(&d_pts, sizeof(double) * 3 * np));
    HIP_TRY_T(hipMalloc(&d_phi, sizeof(double) * np));
    HIP_TRY_T(hipMemcpy(d_pts, pts.data(), sizeof(double) * 3 * np, hipMemcpyHostToDevice));
    launch_sample_phi(st_, D_, W_, job, np, d_pts, d_phi);
    HIP_TRY_T(hipGetLastError());
    HIP_TRY_T(hipMemcpyAsync(ph.data(), d_phi, sizeof(double) * np, hipMemcpyDeviceToHost, st_));
    HIP_TRY_T(hipStreamSynchronize(st_));
    std::vector<uint8_t> keep(t1.size(), 1);
    for (size_t u = 0; u < test.size(); u++) {
      const Tri& t = t1[test[u]];
      const double phi_size = 1. / 3. * (ph[t[0]] + ph[t[1]] + ph[t[2]]);
      if (ph[n + u] > 0.0001 * phi_size) continue;
      keep[test[u]] = 0;
      // the longest edge goes (first of (i,j), (j,k), (i,k) on ties)
      const double* v0 = &x[3 * size_t(t[0])];
      const double* v1 = &x[3 * size_t(t[1])];
      const double* v2 = &x[3 * size_t(t[2])];
      double d = dist2d(v0, v1);
      std::pair<int32_t, int32_t> e = {t[0], t[1]};
      const double d2 = dist2d(v1, v2);
      if (d2 > d) { d = d2; e = {t[1], t[2]}; }
      const double d3 = dist2d(v2, v0);
      if (d3 > d) { d = d3; e = {t[0], t[2]}; }
      del.push_back(e);
    }
    std::sort(del.begin(), del.end());
    del.erase(std::unique(del.begin(), del.end()), del.end());
    st->n_deleted_edge = int64_t(del.size());
    // the graph again, from the kept triangles that have no deleted edge
    Graph g2(n);
    auto deleted = [&](int32_t a, int32_t b) { return std::binary_search(del.begin(), del.end(), std::make_pair(a, b)); };
    for (size_t q = 0; q < t1.size(); q++) {
      if (!keep[q]) continue;
      const Tri& t = t1[q];
      if (deleted(t[0], t[1]) || deleted(t[0], t[2]) || deleted(t[1], t[2])) continue;
      g2.add(t[0], t[1]);
      g2.add(t[1], t[2]);
      g2.add(t[0], t[2]);
    }
    g2.finish();
    cliques(g2, tris);
    quality(tris, x, st);
  }
done:
  if (d_nb) (void)hipFree(d_nb);
  if (d_cnt) (void)hipFree(d_cnt);
  if (d_pts) (void)hipFree(d_pts);
  if (d_phi) (void)hipFree(d_phi);
  return rc;
}

}  // namespace smg

extern "C" {

int smg_tri_edges(smg_job* job, int32_t* edges, int64_t capacity, int64_t* n_edges);   // smg_api.cpp

int smg_tri_cliques(int64_t nv, const int32_t* edges, int64_t ne, int32_t* tris, int64_t capacity, int64_t* n_tri) {
  if (nv < 0 || ne < 0 || (ne > 0 && !edges) || !n_tri) { smg::set_error("smg_tri_cliques: bad argument"); return SMG_ERR_ARG; }
  std::vector<std::array<int32_t, 3>> out;
  const int rc = smg::tri_cliques_host(nv, edges, ne, out);
  if (rc != SMG_OK) return rc;
  *n_tri = int64_t(out.size());
  if (!tris) return SMG_OK;
  if (int64_t(out.size()) > capacity) { smg::set_error("smg_tri_cliques: capacity too small"); return SMG_ERR_CAPACITY; }
  for (size_t q = 0; q < out.size(); q++)
    for (int a = 0; a < 3; a++) tris[3 * q + a] = out[q][a];
  return SMG_OK;
}

}  // extern "C"
