// smg_tet.cpp -- 3D tetrahedralisation of a job's particles: Delaunay tetrahedra from
// the GPU's Voronoi cells (k_vor3d), then the host stage of Graphmeshcls::Reconstruct_tets
// (graph_mesh.cpp:981-1963): the deletion rule of Check_whether_this_tet_should_be_delete
// (graph_mesh.cpp:675-979), the boundary triangles (graph_mesh.cpp:1153-1176) and
// Tet_quality_statistics (graph_mesh.cpp:1967-2460), in FP64.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "smg.h"
#include "smg_engine.hpp"
#include "smg_internal.hpp"
#include "smg_kernels.hpp"

namespace smg {

namespace {

using Tet = std::array<int32_t, 4>;
using Tri3 = std::array<int32_t, 3>;

// the reference's PI (glbparam.h:185); its dihedral angles are acos(.) / PI * 180
constexpr double kPiRef = 3.1415926;
constexpr double kDefaultDeleteThres = 10.0;   // tet_delete_thres, sph_mesh_generation.cpp:705

struct TetGeo {
  double vol;                 // |volume|
  double l2min, l2max;        // squared edge lengths
  double aspect;              // longest edge / smallest height
  double radius_ratio;        // 3 r_in / r_circ
  double dih[6];              // dihedral angles, degrees
};

inline void sub(const double* a, const double* b, double* r) {
  r[0] = a[0] - b[0]; r[1] = a[1] - b[1]; r[2] = a[2] - b[2];
}
inline void cross(const double* a, const double* b, double* r) {
  r[0] = a[1] * b[2] - a[2] * b[1];
  r[1] = a[2] * b[0] - a[0] * b[2];
  r[2] = a[0] * b[1] - a[1] * b[0];
}
inline double dot(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// Shape measures of one tetrahedron.  The reference takes them from an LU solve of the
// edge matrix (face normals = columns of its inverse); here from face cross products:
// the inverse height of a face is its area / (3 volume), the inradius 3 volume / total
// area, the circumcentre the closed-form solution of the same 3x3 system.
TetGeo tet_geometry(const double* p0, const double* p1, const double* p2, const double* p3) {
  TetGeo g;
  const double* P[4] = {p0, p1, p2, p3};
  double V[6][3];
  sub(p0, p3, V[0]); sub(p1, p3, V[1]); sub(p2, p3, V[2]);
  sub(p1, p0, V[3]); sub(p2, p1, V[4]); sub(p0, p2, V[5]);
  g.l2min = g.l2max = dot(V[0], V[0]);
  for (int e = 1; e < 6; e++) {
    const double l2 = dot(V[e], V[e]);
    g.l2min = std::min(g.l2min, l2);
    g.l2max = std::max(g.l2max, l2);
  }
  double c12[3], c20[3], c01[3];
  cross(V[1], V[2], c12); cross(V[2], V[0], c20); cross(V[0], V[1], c01);
  const double det = dot(V[0], c12);   // 6 x signed volume
  g.vol = std::fabs(det) / 6.;
  // outward unit normals and areas of the faces opposite each vertex
  double N[4][3], area[4], atot = 0.;
  for (int f = 0; f < 4; f++) {
    const double* a = P[(f + 1) % 4];
    const double* b = P[(f + 2) % 4];
    const double* c = P[(f + 3) % 4];
    double ab[3], ac[3], ao[3];
    sub(b, a, ab); sub(c, a, ac); sub(P[f], a, ao);
    cross(ab, ac, N[f]);
    if (dot(N[f], ao) > 0.)
      for (int k = 0; k < 3; k++) N[f][k] = -N[f][k];
    const double nn = std::sqrt(dot(N[f], N[f]));
    area[f] = 0.5 * nn;
    atot += area[f];
    for (int k = 0; k < 3; k++) N[f][k] = nn > 0. ? N[f][k] / nn : 0.;
  }
  double hmax = 0.;
  for (int f = 0; f < 4; f++) hmax = std::max(hmax, area[f] / (3. * g.vol));
  g.aspect = std::sqrt(g.l2max) * hmax;
  if (g.vol > 0.) {
    const double r_in = 3. * g.vol / atot;
    double cc[3];
    for (int k = 0; k < 3; k++)
      cc[k] = (dot(V[0], V[0]) * c12[k] + dot(V[1], V[1]) * c20[k] + dot(V[2], V[2]) * c01[k]) / (2. * det);
    g.radius_ratio = 3. * r_in / std::sqrt(dot(cc, cc));
  } else {
    g.aspect = std::numeric_limits<double>::infinity();
    g.radius_ratio = 0.;
  }
  int e = 0;
  for (int a = 0; a < 4; a++)
    for (int b = a + 1; b < 4; b++) {
      double c = -dot(N[a], N[b]);
      c = c < -1. ? -1. : (c > 1. ? 1. : c);
      g.dih[e++] = std::acos(c) / kPiRef * 180.;
    }
  return g;
}

void stats_init(smg_tet_stats* st) {
  st->min_edge = st->min_volume = st->min_aspect = st->min_radius_ratio = st->min_dihedral =
      std::numeric_limits<double>::infinity();
  st->max_edge = st->max_volume = st->max_aspect = st->max_radius_ratio = st->max_dihedral = 0.;
}

// Tet_quality_statistics (graph_mesh.cpp:1967-2460) over the kept tetrahedra
void quality(const std::vector<Tet>& tets, const double* x, smg_tet_stats* st) {
  static const double aspect_table[11] = {1.5, 2.0, 2.5, 3.0, 4.0, 6.0, 10.0, 15.0, 25.0, 50.0, 100.0};
  const double dratio = 0.02;
  const int nradius = int(1.1 / dratio) + 1;   // 56 bins, the last one open
  double l2min = std::numeric_limits<double>::infinity(), l2max = 0.;
  double asum = 0., rsum = 0., dsum = 0.;
  for (const Tet& t : tets) {
    const TetGeo g = tet_geometry(&x[3 * size_t(t[0])], &x[3 * size_t(t[1])], &x[3 * size_t(t[2])],
                                  &x[3 * size_t(t[3])]);
    st->total_volume += g.vol;
    st->min_volume = std::min(st->min_volume, g.vol);
    st->max_volume = std::max(st->max_volume, g.vol);
    l2min = std::min(l2min, g.l2min);
    l2max = std::max(l2max, g.l2max);
    double dmin = 180., dmax = 0.;
    for (double d : g.dih) {
      dmin = std::min(dmin, d);
      dmax = std::max(dmax, d);
      st->dihedral_hist[std::min(44, int(std::floor(d / 4.0)))] += 1;   // 180 degrees: the last bin
    }
    st->min_dihedral = std::min(st->min_dihedral, dmin);
    st->max_dihedral = std::max(st->max_dihedral, dmax);
    const double lim[4] = {10., 20., 30., 40.};
    for (int q = 0; q < 4; q++) st->n_min_dihedral_below[q] += dmin < lim[q];
    dsum += dmin;
    st->min_radius_ratio = std::min(st->min_radius_ratio, g.radius_ratio);
    st->max_radius_ratio = std::max(st->max_radius_ratio, g.radius_ratio);
    rsum += g.radius_ratio;
    const int ri = int(std::floor(g.radius_ratio / dratio));
    st->radius_hist[ri < nradius - 1 ? ri : nradius - 1] += 1;
    st->min_aspect = std::min(st->min_aspect, g.aspect);
    st->max_aspect = std::max(st->max_aspect, g.aspect);
    asum += g.aspect;
    int ai = 0;
    while (ai < 11 && g.aspect > aspect_table[ai]) ai++;
    st->aspect_hist[ai] += 1;
  }
  st->n_tet = int64_t(tets.size());
  if (!tets.empty()) {
    st->min_edge = std::sqrt(l2min);
    st->max_edge = std::sqrt(l2max);
    st->avg_aspect = asum / double(tets.size());
    st->avg_radius_ratio = rsum / double(tets.size());
    st->avg_min_dihedral = dsum / double(tets.size());
  } else {
    stats_init(st);
    st->min_edge = st->min_volume = st->min_aspect = st->min_radius_ratio = st->min_dihedral = 0.;
  }
}

}  // namespace

// deletion rule, boundary triangles and statistics of given tetrahedra (host only)
void tet_cell_quality(const double* x, const int32_t* tets, int64_t n, float* radius_ratio, float* min_dihedral) {
  for (int64_t q = 0; q < n; q++) {
    const int32_t* t = tets + 4 * q;
    const TetGeo g = tet_geometry(&x[3 * size_t(t[0])], &x[3 * size_t(t[1])], &x[3 * size_t(t[2])],
                                  &x[3 * size_t(t[3])]);
    double dmin = 180.;
    for (double d : g.dih) dmin = std::min(dmin, d);
    radius_ratio[q] = float(g.radius_ratio);
    min_dihedral[q] = float(dmin);
  }
}

int tet_finish_host(int64_t nv, const double* x, const int32_t* type, const std::vector<Tet>& in, double thres,
                    std::vector<Tet>& kept, std::vector<Tri3>& surf, smg_tet_stats* st) {
  if (!(thres > 0.)) thres = kDefaultDeleteThres;
  kept.clear();
  surf.clear();
  stats_init(st);
  for (const Tet& t : in)
    for (int a = 0; a < 4; a++)
      if (t[a] < 0 || t[a] >= nv) { set_error("tet_finish: vertex out of range"); return SMG_ERR_ARG; }
  for (const Tet& t : in) {
    int n_real = 0;
    for (int a = 0; a < 4; a++) n_real += type[t[a]] == SMG_REAL;
    if (n_real == 0) {
      const TetGeo g = tet_geometry(&x[3 * size_t(t[0])], &x[3 * size_t(t[1])], &x[3 * size_t(t[2])],
                                    &x[3 * size_t(t[3])]);
      if (g.aspect > thres) {
        st->n_deleted++;
        continue;
      }
    }
    kept.push_back(t);
    if (n_real == 1) {
      Tri3 f;
      int m = 0;
      for (int a = 0; a < 4; a++)
        if (type[t[a]] != SMG_REAL) f[m++] = t[a];
      std::sort(f.begin(), f.end());
      surf.push_back(f);
    }
  }
  std::sort(surf.begin(), surf.end());
  surf.erase(std::unique(surf.begin(), surf.end()), surf.end());
  st->n_surface_tri = int64_t(surf.size());
  quality(kept, x, st);
  return SMG_OK;
}

namespace {

// device buffer owned by this translation unit's calls
template <typename T>
struct DevBuf {
  T* p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  hipError_t alloc(size_t n) {
    if (p) (void)hipFree(p);
    p = nullptr;
    return hipMalloc(&p, sizeof(T) * (n ? n : 1));
  }
};

}  // namespace

#define HIP_TRY_TET(expr)                                              \
  do {                                                                 \
    hipError_t e_ = (expr);                                            \
    if (e_ != hipSuccess) {                                            \
      set_error(std::string(#expr) + ": " + hipGetErrorString(e_));    \
      return SMG_ERR_HIP;                                              \
    }                                                                  \
  } while (0)

int Engine::tetrahedralize(int job, double delete_thres, smg_tet_stats* st) {
  const auto t_start = std::chrono::steady_clock::now();
  std::memset(st, 0, sizeof(*st));
  tet_ = TetResult();
  if (D_ != 3) { set_error("tetrahedralize: 3D only (2D jobs use smg_triangulate)"); return SMG_ERR_ARG; }
  if (!prerun_done_) { set_error("tetrahedralize: seed the job first"); return SMG_ERR_STATE; }
  const JobParams& J = jp_[job];
  const int64_t n = J.n;
  if (n <= 0 || n > int64_t(std::numeric_limits<int32_t>::max() / 4)) {
    set_error("tetrahedralize: particle count out of range");
    return SMG_ERR_ARG;
  }
  std::vector<double> x(3 * size_t(n)), h(n);
  std::vector<int32_t> type(n);
  {
    const int rc = get_state(job, x.data(), nullptr, type.data(), h.data(), nullptr);
    if (rc != SMG_OK) return rc;
  }
  // uniform grid over the job's box, cell size = the mean spacing
  V3Grid G;
  double vol = 1.;
  for (int a = 0; a < 3; a++) vol *= J.box_r[a] - J.box_l[a];
  G.s = std::cbrt(vol / double(n));
  int64_t ncell = 1;
  for (int a = 0; a < 3; a++) {
    G.lo[a] = J.box_l[a];
    G.g[a] = std::max(1, int(std::ceil((J.box_r[a] - J.box_l[a]) / G.s)));
    ncell *= G.g[a];
  }
  if (ncell > 8 * n + 64) { set_error("tetrahedralize: degenerate job box"); return SMG_ERR_ARG; }
  // search radius: Delaunay circumradii stay below ~1.4 h, and a vertex at distance rho
  // needs every point within 2 rho clipped
  double hmax = 0.;
  for (int64_t i = 0; i < n; i++) hmax = std::max(hmax, h[i]);
  G.rmax = std::min(16, std::max(2, int(std::ceil(4. * hmax / G.s)) + 1));
  // Coincident particles: the reference's seeding clamps coordinates to the level-set
  // box shrunk by one package (sph_mesh_generation.cpp:1327-1342), so a few particles
  // start on the same corner point, and with no force between them they stay there.  A
  // tetrahedralisation keeps one vertex per point (TetGen marks the others duplicated
  // and leaves them out of every tetrahedron): here the smallest id; the others take
  // part in no cell.
  std::vector<uint8_t> dup(n, 0);
  {
    std::vector<int32_t> by(n);
    for (int64_t i = 0; i < n; i++) by[i] = int32_t(i);
    std::sort(by.begin(), by.end(), [&](int32_t a, int32_t b) {
      for (int k = 0; k < 3; k++)
        if (x[3 * size_t(a) + k] != x[3 * size_t(b) + k]) return x[3 * size_t(a) + k] < x[3 * size_t(b) + k];
      return a < b;
    });
    for (int64_t q = 1; q < n; q++) {
      const int32_t a = by[q - 1], b = by[q];
      if (x[3 * size_t(a)] == x[3 * size_t(b)] && x[3 * size_t(a) + 1] == x[3 * size_t(b) + 1] &&
          x[3 * size_t(a) + 2] == x[3 * size_t(b) + 2])
        dup[b] = 1;   // sorted by id within equal points: the first keeps the point
    }
  }
  std::vector<int32_t> cell_of(n), start(ncell + 1, 0), order;
  order.reserve(n);
  for (int64_t i = 0; i < n; i++) {
    if (dup[i]) continue;
    int64_t c = 0;
    for (int a = 2; a >= 0; a--) {
      int k = int(std::floor((x[3 * i + a] - G.lo[a]) / G.s));
      k = k < 0 ? 0 : (k >= G.g[a] ? G.g[a] - 1 : k);
      c = c * G.g[a] + k;
    }
    cell_of[i] = int32_t(c);
    start[c + 1]++;
  }
  for (int64_t c = 0; c < ncell; c++) start[c + 1] += start[c];
  const int64_t n_cells = start[ncell];   // particles with a cell of their own
  order.assign(size_t(n_cells), 0);
  {
    std::vector<int32_t> fill(start.begin(), start.end() - 1);
    for (int64_t i = 0; i < n; i++)
      if (!dup[i]) order[fill[cell_of[i]]++] = int32_t(i);
  }
  G.n = n_cells;
  DevBuf<double> d_x, d_rcert;
  DevBuf<int32_t> d_start, d_order, d_tets;
  DevBuf<int8_t> d_status;
  DevBuf<unsigned long long> d_count;
  HIP_TRY_TET(hipStreamSynchronize(st_));
  HIP_TRY_TET(d_x.alloc(3 * size_t(n)));
  HIP_TRY_TET(d_start.alloc(size_t(ncell) + 1));
  HIP_TRY_TET(d_order.alloc(size_t(n_cells)));
  HIP_TRY_TET(d_status.alloc(size_t(n)));
  HIP_TRY_TET(d_rcert.alloc(size_t(n)));
  HIP_TRY_TET(d_count.alloc(1));
  HIP_TRY_TET(hipMemcpy(d_x.p, x.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice));
  HIP_TRY_TET(hipMemcpy(d_start.p, start.data(), sizeof(int32_t) * (ncell + 1), hipMemcpyHostToDevice));
  HIP_TRY_TET(hipMemcpy(d_order.p, order.data(), sizeof(int32_t) * n_cells, hipMemcpyHostToDevice));
  G.x = d_x.p;
  G.start = d_start.p;
  G.order = d_order.p;
  // pass-2 inputs: blocks of 4^3 grid cells with the bounding box of their points
  V3Blocks K;
  K.s = 4. * G.s;
  int64_t nblk = 1;
  for (int a = 0; a < 3; a++) {
    K.lo[a] = G.lo[a];
    K.g[a] = std::max(1, int(std::ceil((J.box_r[a] - J.box_l[a]) / K.s)));
    nblk *= K.g[a];
  }
  std::vector<int32_t> bstart(nblk + 1, 0), border(static_cast<size_t>(n_cells));
  std::vector<double> bbox(6 * size_t(nblk));
  {
    std::vector<int32_t> blk_of(static_cast<size_t>(n_cells));
    for (int64_t q = 0; q < n_cells; q++) {
      const int32_t i = order[q];
      int64_t b = 0;
      for (int a = 2; a >= 0; a--) {
        int k = int(std::floor((x[3 * size_t(i) + a] - K.lo[a]) / K.s));
        k = k < 0 ? 0 : (k >= K.g[a] ? K.g[a] - 1 : k);
        b = b * K.g[a] + k;
      }
      blk_of[q] = int32_t(b);
      bstart[b + 1]++;
    }
    for (int64_t b = 0; b < nblk; b++) bstart[b + 1] += bstart[b];
    std::vector<int32_t> fill(bstart.begin(), bstart.end() - 1);
    for (int64_t b = 0; b < nblk; b++)
      for (int a = 0; a < 3; a++) {
        bbox[6 * b + a] = std::numeric_limits<double>::infinity();
        bbox[6 * b + 3 + a] = -std::numeric_limits<double>::infinity();
      }
    for (int64_t q = 0; q < n_cells; q++) {
      const int32_t i = order[q];
      const int64_t b = blk_of[q];
      border[fill[b]++] = i;
      for (int a = 0; a < 3; a++) {
        bbox[6 * b + a] = std::min(bbox[6 * b + a], x[3 * size_t(i) + a]);
        bbox[6 * b + 3 + a] = std::max(bbox[6 * b + 3 + a], x[3 * size_t(i) + a]);
      }
    }
  }
  // the second pass's starting box: far beyond the point set, so that a circumcentre
  // it cuts off would be 10^6 box lengths away
  double ext = 0.;
  for (int a = 0; a < 3; a++) ext = std::max(ext, J.box_r[a] - J.box_l[a]);
  const double Bfar = 1.e6 * ext;
  DevBuf<int32_t> d_bstart, d_border, d_open;
  DevBuf<double> d_bbox;
  HIP_TRY_TET(d_bstart.alloc(size_t(nblk) + 1));
  HIP_TRY_TET(d_border.alloc(size_t(n_cells)));
  HIP_TRY_TET(d_bbox.alloc(6 * size_t(nblk)));
  HIP_TRY_TET(d_open.alloc(size_t(n_cells)));
  HIP_TRY_TET(hipMemcpy(d_bstart.p, bstart.data(), sizeof(int32_t) * (nblk + 1), hipMemcpyHostToDevice));
  HIP_TRY_TET(hipMemcpy(d_border.p, border.data(), sizeof(int32_t) * n_cells, hipMemcpyHostToDevice));
  HIP_TRY_TET(hipMemcpy(d_bbox.p, bbox.data(), sizeof(double) * 6 * nblk, hipMemcpyHostToDevice));
  K.start = d_bstart.p;
  K.order = d_border.p;
  K.box = d_bbox.p;
  std::vector<int8_t> st1(n);
  std::vector<int32_t> open_ids;
  int64_t cap = 12 * n;
  unsigned long long count = 0ull;
  hipEvent_t ev[2] = {nullptr, nullptr};
  struct EvGuard {
    hipEvent_t* e;
    ~EvGuard() {
      for (int k = 0; k < 2; k++)
        if (e[k]) (void)hipEventDestroy(e[k]);
    }
  } ev_guard{ev};
  HIP_TRY_TET(hipEventCreate(&ev[0]));
  HIP_TRY_TET(hipEventCreate(&ev[1]));
  float cell_ms = 0.f;
  for (int attempt = 0; attempt < 2; attempt++) {
    HIP_TRY_TET(d_tets.alloc(4 * size_t(cap)));
    HIP_TRY_TET(hipMemsetAsync(d_count.p, 0, sizeof(unsigned long long), st_));
    V3Out O;
    O.tets = d_tets.p;
    O.count = d_count.p;
    O.cap = cap;
    O.status = d_status.p;
    O.rcert = d_rcert.p;
    HIP_TRY_TET(hipEventRecord(ev[0], st_));
    launch_vor3d(st_, G, O);
    HIP_TRY_TET(hipGetLastError());
    // the cells pass 1 left open (hull particles, cavity spanners) go to pass 2
    HIP_TRY_TET(hipMemcpyAsync(st1.data(), d_status.p, sizeof(int8_t) * n, hipMemcpyDeviceToHost, st_));
    HIP_TRY_TET(hipStreamSynchronize(st_));
    open_ids.clear();
    for (int64_t q = 0; q < n_cells; q++)
      if (st1[order[q]] != 0) open_ids.push_back(order[q]);
    if (!open_ids.empty()) {
      HIP_TRY_TET(hipMemcpyAsync(d_open.p, open_ids.data(), sizeof(int32_t) * open_ids.size(), hipMemcpyHostToDevice,
                                 st_));
      launch_vor3d_full(st_, G, K, d_open.p, int64_t(open_ids.size()), Bfar, O);
      HIP_TRY_TET(hipGetLastError());
    }
    HIP_TRY_TET(hipEventRecord(ev[1], st_));
    HIP_TRY_TET(hipMemcpyAsync(&count, d_count.p, sizeof(count), hipMemcpyDeviceToHost, st_));
    HIP_TRY_TET(hipStreamSynchronize(st_));
    HIP_TRY_TET(hipEventElapsedTime(&cell_ms, ev[0], ev[1]));
    if (int64_t(count) <= cap) break;
    cap = int64_t(count);
  }
  if (int64_t(count) > cap) { set_error("tetrahedralize: output capacity"); return SMG_ERR_CAPACITY; }
  std::vector<int32_t> raw(4 * size_t(count));
  tet_.status.resize(n);
  tet_.rcert.resize(n);
  HIP_TRY_TET(hipMemcpy(raw.data(), d_tets.p, sizeof(int32_t) * raw.size(), hipMemcpyDeviceToHost));
  HIP_TRY_TET(hipMemcpy(tet_.status.data(), d_status.p, sizeof(int8_t) * n, hipMemcpyDeviceToHost));
  HIP_TRY_TET(hipMemcpy(tet_.rcert.data(), d_rcert.p, sizeof(double) * n, hipMemcpyDeviceToHost));
  for (int64_t i = 0; i < n; i++)
    if (dup[i]) {
      tet_.status[i] = 6;
      tet_.rcert[i] = 0.;
    }
  tet_.delaunay.resize(count);
  for (size_t q = 0; q < count; q++)
    for (int a = 0; a < 4; a++) tet_.delaunay[q][a] = raw[4 * q + a];
  std::sort(tet_.delaunay.begin(), tet_.delaunay.end());
  smg_tet_stats s2;
  std::memset(&s2, 0, sizeof(s2));
  {
    const int rc = tet_finish_host(n, x.data(), type.data(), tet_.delaunay, delete_thres, tet_.kept, tet_.surface, &s2);
    if (rc != SMG_OK) return rc;
  }
  *st = s2;
  st->n_delaunay = int64_t(tet_.delaunay.size());
  for (int64_t i = 0; i < n; i++) {
    st->n_open += tet_.status[i] == 1;
    st->n_cell_overflow += tet_.status[i] >= 2 && tet_.status[i] <= 5;
    st->n_duplicate += tet_.status[i] == 6;
  }
  st->n_second_pass = int64_t(open_ids.size());
  st->cell_ms = double(cell_ms);
  st->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
  return SMG_OK;
}

}  // namespace smg

extern "C" {

int smg_tet_finish(int64_t n_vertices, const double* x, const int32_t* type, const int32_t* tets, int64_t n_tets,
                   double delete_thres, int32_t* kept, int64_t kept_capacity, int64_t* n_kept, int32_t* tris,
                   int64_t tri_capacity, int64_t* n_tri, smg_tet_stats* stats) {
  if (n_vertices < 0 || n_tets < 0 || (n_vertices > 0 && (!x || !type)) || (n_tets > 0 && !tets) || !n_kept ||
      !n_tri || !stats) {
    smg::set_error("smg_tet_finish: bad argument");
    return SMG_ERR_ARG;
  }
  std::vector<smg::Tet> in(static_cast<size_t>(n_tets));
  for (int64_t q = 0; q < n_tets; q++)
    for (int a = 0; a < 4; a++) in[q][a] = tets[4 * q + a];
  std::vector<smg::Tet> out;
  std::vector<smg::Tri3> surf;
  std::memset(stats, 0, sizeof(*stats));
  const int rc = smg::tet_finish_host(n_vertices, x, type, in, delete_thres, out, surf, stats);
  if (rc != SMG_OK) return rc;
  stats->n_delaunay = n_tets;
  *n_kept = int64_t(out.size());
  *n_tri = int64_t(surf.size());
  if (kept) {
    if (int64_t(out.size()) > kept_capacity) { smg::set_error("smg_tet_finish: capacity too small"); return SMG_ERR_CAPACITY; }
    for (size_t q = 0; q < out.size(); q++)
      for (int a = 0; a < 4; a++) kept[4 * q + a] = out[q][a];
  }
  if (tris) {
    if (int64_t(surf.size()) > tri_capacity) { smg::set_error("smg_tet_finish: capacity too small"); return SMG_ERR_CAPACITY; }
    for (size_t q = 0; q < surf.size(); q++)
      for (int a = 0; a < 3; a++) tris[3 * q + a] = surf[q][a];
  }
  return SMG_OK;
}

}  // extern "C"
