// smg_vor3d.hip -- 3D Delaunay tetrahedra of a job's particles from their Voronoi cells.
//
// The reference tetrahedralises the particle positions with TetGen ("JCnQ": the
// Delaunay tetrahedralisation of the point set; Graphmeshcls::Reconstruct_tets,
// graph_mesh.cpp:981-1034).  Here every particle's Voronoi cell is clipped on the GPU,
// one wavefront per cell, and each cell vertex made by three bisector planes is a
// Delaunay tetrahedron (the particle and the three neighbours whose bisectors meet
// there).
//
// Cell representation (the dual of the cell's boundary): planes n.p <= d relative to
// the particle, and triangles (a, b, c) of plane indices, one per cell vertex, oriented
// so that det(n_a, n_b, n_c) > 0.  Two triangles that share a cell edge hold it in
// opposite directions.  Clipping by a new plane removes the triangles whose vertex lies
// beyond it; the directed edges x -> y of removed triangles whose twin y -> x is not
// removed bound the removed patch, and each gets a new triangle (x, y, new).
//
// Pass 1 (k_vor3d): candidates come from a uniform grid of cell size s in rings of
// cells around the particle.  After ring r every point within r s has been clipped.  A
// cell vertex at distance rho is a true Voronoi vertex once every point within 2 rho has
// been clipped (its empty sphere lies in that ball), so the search stops when
// 2 max(rho) <= r s, or at ring rmax.  A cell that closes this way emits its vertices.
//
// Pass 2 (k_vor3d_full) takes the cells pass 1 left open: the convex-hull particles,
// whose cells are unbounded, and those whose cells reach across a cavity.  The cell
// starts from a box far larger than the point set, is clipped by the same rings, and
// then every vertex v (box vertices included) is verified: a point q lies strictly
// inside the sphere through the particle centred at v exactly when v is beyond q's
// bisector, so the cell is final once no point is inside any vertex's sphere.  The
// points are found through coarse blocks of the grid with the bounding box of their
// points; a block is scanned when that box meets the sphere.  Each point found removes
// the vertex it was found for, so the loop ends.  Vertices made by three bisectors are
// then exactly the particle's Delaunay tetrahedra, however large their circumspheres.
//
// Every tetrahedron is emitted once, by its smallest particle id.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "smg_kernels.hpp"

namespace smg {

namespace {

constexpr int kV3Waves = 4;      // cells per workgroup in pass 1
constexpr int kV3Box = 6;        // the six planes of the starting box
constexpr int kV3MaxVerify = 8192;   // pass 2: vertex verifications and clips per cell

// N planes and N triangles (cell vertices), N a multiple of 64 and at most 256 (plane
// indices are 8 bits in a triangle)
template <int N>
struct V3Cell {
  double pn[N][4];           // n.xyz, d
  int32_t pid[N];            // particle id of a bisector plane
  uint32_t tri[2][N];        // a | b << 8 | c << 16, ping-pong buffers
  double vx[2][N], vy[2][N], vz[2][N];
  uint32_t emask[N][N / 32]; // directed edges of the removed triangles, bit (x, y)
  uint32_t bnd[N];           // boundary edges x | y << 8; plane renumbering in a compaction
  uint32_t ref[N / 32];      // planes some vertex uses
  uint8_t ver[2][N];         // pass 2: the vertex's sphere holds no point
  int nb;
};

// LDS traffic of one wavefront: order it and make it visible to the other lanes
__device__ __forceinline__ void wsync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

__device__ __forceinline__ int wave_max_i(int v) {
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return v;
}

__device__ __forceinline__ uint32_t lanes_below(uint64_t m, int lane) {
  return uint32_t(__popcll(m & ((1ull << lane) - 1ull)));
}

// the point where planes a, b, c meet (Cramer's rule)
__device__ __forceinline__ void meet(const double* A, const double* B, const double* C, double& x, double& y,
                                     double& z, bool& ok) {
  const double bc[3] = {B[1] * C[2] - B[2] * C[1], B[2] * C[0] - B[0] * C[2], B[0] * C[1] - B[1] * C[0]};
  const double ca[3] = {C[1] * A[2] - C[2] * A[1], C[2] * A[0] - C[0] * A[2], C[0] * A[1] - C[1] * A[0]};
  const double ab[3] = {A[1] * B[2] - A[2] * B[1], A[2] * B[0] - A[0] * B[2], A[0] * B[1] - A[1] * B[0]};
  const double det = A[0] * bc[0] + A[1] * bc[1] + A[2] * bc[2];
  const double r = 1.0 / det;
  x = (A[3] * bc[0] + B[3] * ca[0] + C[3] * ab[0]) * r;
  y = (A[3] * bc[1] + B[3] * ca[1] + C[3] * ab[1]) * r;
  z = (A[3] * bc[2] + B[3] * ca[2] + C[3] * ab[2]) * r;
  ok = det != 0. && isfinite(x) && isfinite(y) && isfinite(z);
}

// Whether the vertex (vx, vy, vz) lies beyond the bisector of the point u (both relative
// to the particle): u is strictly inside the vertex's sphere.  Pass 2 finds points with
// this same expression, so a point found for a vertex always removes it.
__device__ __forceinline__ bool beyond(double vx, double vy, double vz, double px, double py, double pz, double c) {
  return vx * px + vy * py + vz * pz - c > 0.;
}

// The box |p|_inf <= B around the particle: 6 planes, 8 vertices.
template <int N>
__device__ void init_box(V3Cell<N>& C, double B, int lane) {
  if (lane < kV3Box) {
    const int ax = lane >> 1;
    const double sg = (lane & 1) ? -1. : 1.;
    C.pn[lane][0] = ax == 0 ? sg : 0.;
    C.pn[lane][1] = ax == 1 ? sg : 0.;
    C.pn[lane][2] = ax == 2 ? sg : 0.;
    C.pn[lane][3] = B;
    C.pid[lane] = -1;
  }
  for (int q = lane; q < N * (N / 32); q += 64) (&C.emask[0][0])[q] = 0u;
  if (lane < 8) {
    // corner (sx, sy, sz): planes +x/-x = 0/1, +y/-y = 2/3, +z/-z = 4/5; the order
    // (x, y, z) has det = sx sy sz, so a negative corner swaps the first two
    const int bx = lane & 1, by = (lane >> 1) & 1, bz = (lane >> 2) & 1;
    const uint32_t px = uint32_t(bx), py = uint32_t(2 + by), pz = uint32_t(4 + bz);
    const bool neg = (bx + by + bz) & 1;
    C.tri[0][lane] = neg ? (py | px << 8 | pz << 16) : (px | py << 8 | pz << 16);
    C.vx[0][lane] = bx ? -B : B;
    C.vy[0][lane] = by ? -B : B;
    C.vz[0][lane] = bz ? -B : B;
    C.ver[0][lane] = 0;
  }
  wsync();
}

// Drops the bisector planes that no vertex uses any more and renumbers the rest (the
// box planes keep 0..5); returns the new plane count.
template <int N>
__device__ int compact_planes(V3Cell<N>& C, int nT, int nP, int cb, int lane) {
  constexpr int H = N / 64;
  if (lane < N / 32) C.ref[lane] = 0u;
  wsync();
#pragma unroll
  for (int h = 0; h < H; h++) {
    const int t = lane + 64 * h;
    if (t >= nT) continue;
    const uint32_t tr = C.tri[cb][t];
    for (int e = 0; e < 3; e++) {
      const uint32_t a = (tr >> (8 * e)) & 255u;
      atomicOr(&C.ref[a >> 5], 1u << (a & 31u));
    }
  }
  wsync();
  bool k[H];
  int q[H];
  int base = 0;
#pragma unroll
  for (int h = 0; h < H; h++) {
    const int p = lane + 64 * h;
    k[h] = p < nP && (p < kV3Box || ((C.ref[p >> 5] >> (p & 31)) & 1u));
    const uint64_t m = __ballot(k[h]);
    q[h] = base + int(lanes_below(m, lane));
    base += __popcll(m);
  }
  double a[H][4];
  int32_t id[H];
#pragma unroll
  for (int h = 0; h < H; h++) {
    const int p = lane + 64 * h;
    for (int c = 0; c < 4; c++) a[h][c] = k[h] ? C.pn[p][c] : 0.;
    id[h] = k[h] ? C.pid[p] : 0;
    if (k[h]) C.bnd[p] = uint32_t(q[h]);
  }
  wsync();
#pragma unroll
  for (int h = 0; h < H; h++) {
    if (!k[h]) continue;
    for (int c = 0; c < 4; c++) C.pn[q[h]][c] = a[h][c];
    C.pid[q[h]] = id[h];
  }
#pragma unroll
  for (int h = 0; h < H; h++) {
    const int t = lane + 64 * h;
    if (t >= nT) continue;
    const uint32_t tr = C.tri[cb][t];
    C.tri[cb][t] = C.bnd[tr & 255u] | C.bnd[(tr >> 8) & 255u] << 8 | C.bnd[(tr >> 16) & 255u] << 16;
  }
  wsync();
  return base;
}

// Clips the cell by the bisector of the point (px, py, pz) (relative to the particle)
// of particle pq.  Returns 0 when the plane misses the cell, 1 when it clipped (R2 is
// then the squared distance of the farthest vertex), else the status of a cell that
// cannot be built: 2 capacity, 3 empty cut boundary, 4 non-convex cut, 5 singular vertex.
template <int N>
__device__ int clip(V3Cell<N>& C, int& nT, int& nP, int& cb, double px, double py, double pz, int32_t pq,
                    int lane, double& R2) {
  constexpr int H = N / 64;
  const double u2 = px * px + py * py + pz * pz;
  const double c = 0.5 * u2;
  // classify the vertices
  bool r[H];
  uint64_t m[H], any = 0ull;
#pragma unroll
  for (int h = 0; h < H; h++) {
    const int t = lane + 64 * h;
    r[h] = t < nT && beyond(C.vx[cb][t], C.vy[cb][t], C.vz[cb][t], px, py, pz, c);
    m[h] = __ballot(r[h]);
    any |= m[h];
  }
  if (any == 0ull) return 0;   // the plane misses the cell
  if (nP >= N) nP = compact_planes(C, nT, nP, cb, lane);
  if (nP >= N) return 2;
  const int p = nP;
  if (lane == 0) {
    C.pn[p][0] = px; C.pn[p][1] = py; C.pn[p][2] = pz; C.pn[p][3] = c;
    C.pid[p] = pq;
    C.nb = 0;
  }
  // mark the directed edges of the removed triangles
  uint32_t my[H];
#pragma unroll
  for (int h = 0; h < H; h++) my[h] = r[h] ? C.tri[cb][lane + 64 * h] : 0xffffffffu;
#pragma unroll
  for (int h = 0; h < H; h++) {
    if (my[h] == 0xffffffffu) continue;
    const uint32_t v[3] = {my[h] & 255u, (my[h] >> 8) & 255u, (my[h] >> 16) & 255u};
    for (int e = 0; e < 3; e++) {
      const uint32_t x = v[e], y = v[(e + 1) % 3];
      atomicOr(&C.emask[x][y >> 5], 1u << (y & 31u));
    }
  }
  wsync();
  // boundary: a removed edge x -> y whose twin y -> x is not removed
#pragma unroll
  for (int h = 0; h < H; h++) {
    if (my[h] == 0xffffffffu) continue;
    const uint32_t v[3] = {my[h] & 255u, (my[h] >> 8) & 255u, (my[h] >> 16) & 255u};
    for (int e = 0; e < 3; e++) {
      const uint32_t x = v[e], y = v[(e + 1) % 3];
      if (!((C.emask[y][x >> 5] >> (x & 31u)) & 1u)) {
        const int s = atomicAdd(&C.nb, 1);
        if (s < N) C.bnd[s] = x | y << 8;
      }
    }
  }
  wsync();
#pragma unroll
  for (int h = 0; h < H; h++) {
    if (my[h] == 0xffffffffu) continue;
    const uint32_t v[3] = {my[h] & 255u, (my[h] >> 8) & 255u, (my[h] >> 16) & 255u};
    for (int e = 0; e < 3; e++) C.emask[v[e]][v[(e + 1) % 3] >> 5] = 0u;
  }
  const int nb = C.nb;
  uint64_t keep[H];
  int nK = 0;
#pragma unroll
  for (int h = 0; h < H; h++) {
    const int lo = 64 * h;
    const uint64_t valid = nT >= lo + 64 ? ~0ull : (nT > lo ? ((1ull << (nT - lo)) - 1ull) : 0ull);
    keep[h] = valid & ~m[h];
    nK += __popcll(keep[h]);
  }
  if (nK + nb > N || nb > 64) return 2;
  if (nb < 3) return 3;
  // a plane starts at most one boundary edge of a convex cut
  bool dup = false;
  if (lane < nb) {
    const uint32_t x = C.bnd[lane] & 255u;
    for (int s = 0; s < nb; s++) dup = dup || (s != lane && (C.bnd[s] & 255u) == x);
  }
  if (__ballot(dup)) return 4;
  const int nx = cb ^ 1;
  int pos0 = 0;
#pragma unroll
  for (int h = 0; h < H; h++) {
    const int t = lane + 64 * h;
    if ((keep[h] >> lane) & 1ull) {
      const int pos = pos0 + int(lanes_below(keep[h], lane));
      C.tri[nx][pos] = C.tri[cb][t];
      C.vx[nx][pos] = C.vx[cb][t]; C.vy[nx][pos] = C.vy[cb][t]; C.vz[nx][pos] = C.vz[cb][t];
      C.ver[nx][pos] = C.ver[cb][t];
    }
    pos0 += __popcll(keep[h]);
  }
  bool okv = true;
  if (lane < nb) {
    const uint32_t x = C.bnd[lane] & 255u, y = (C.bnd[lane] >> 8) & 255u;
    double vx, vy, vz;
    meet(C.pn[x], C.pn[y], C.pn[p], vx, vy, vz, okv);
    C.tri[nx][nK + lane] = x | y << 8 | uint32_t(p) << 16;
    C.vx[nx][nK + lane] = vx; C.vy[nx][nK + lane] = vy; C.vz[nx][nK + lane] = vz;
    C.ver[nx][nK + lane] = 0;
  }
  if (__ballot(!okv)) return 5;
  wsync();
  nT = nK + nb;
  nP = p + 1;
  cb = nx;
  double d2 = 0.;
#pragma unroll
  for (int h = 0; h < H; h++) {
    const int t = lane + 64 * h;
    if (t < nT) d2 = fmax(d2, C.vx[cb][t] * C.vx[cb][t] + C.vy[cb][t] * C.vy[cb][t] + C.vz[cb][t] * C.vz[cb][t]);
  }
  R2 = wave_max(d2);
  return 1;
}

// Clips by the grid's points in rings 0..rmax of cells around the particle, skipping
// points beyond twice the farthest vertex (they cannot cut the cell).  In pass 1 the
// search stops early once 4 R2 <= (r s)^2; returns the last ring searched.
template <int N>
__device__ int clip_rings(const V3Grid& G, V3Cell<N>& C, int& nT, int& nP, int& cb, int32_t i, double xi, double yi,
                          double zi, const int* ci, bool stop_when_certified, int lane, double& R2, int& why) {
  int r_done = -1;
  for (int r = 0; r <= G.rmax && !why; r++) {
    const int side = 2 * r + 1;
    const int ncube = side * side * side;
    for (int base = 0; base < ncube && !why; base += 64) {
      const int idx = base + lane;
      int32_t st = 0, en = 0;
      if (idx < ncube) {
        const int dx = idx % side - r, dy = (idx / side) % side - r, dz = idx / (side * side) - r;
        const int ax = dx < 0 ? -dx : dx, ay = dy < 0 ? -dy : dy, az = dz < 0 ? -dz : dz;
        const int cheb = max(ax, max(ay, az));
        const int c0 = ci[0] + dx, c1 = ci[1] + dy, c2 = ci[2] + dz;
        if (cheb == r && c0 >= 0 && c0 < G.g[0] && c1 >= 0 && c1 < G.g[1] && c2 >= 0 && c2 < G.g[2]) {
          const int64_t cell = (int64_t(c2) * G.g[1] + c1) * G.g[0] + c0;
          st = G.start[cell];
          en = G.start[cell + 1];
        }
      }
      const int mc = wave_max_i(en - st);
      for (int t = 0; t < mc && !why; t++) {
        int32_t q = -1;
        double ux = 0., uy = 0., uz = 0.;
        bool cand = false;
        if (st + t < en) {
          q = G.order[st + t];
          ux = G.x[3 * int64_t(q)] - xi;
          uy = G.x[3 * int64_t(q) + 1] - yi;
          uz = G.x[3 * int64_t(q) + 2] - zi;
          cand = q != i && ux * ux + uy * uy + uz * uz < 4. * R2;
        }
        uint64_t m = __ballot(cand);
        while (m && !why) {
          const int L = __ffsll((unsigned long long)m) - 1;
          m &= m - 1ull;
          const double px = __shfl(ux, L, 64), py = __shfl(uy, L, 64), pz = __shfl(uz, L, 64);
          const int32_t pq = __shfl(q, L, 64);
          const double u2 = px * px + py * py + pz * pz;
          if (!(u2 < 4. * R2)) continue;   // the cell shrank since the lane's test
          const int rc = clip(C, nT, nP, cb, px, py, pz, pq, lane, R2);
          if (rc > 1) why = rc;
        }
      }
    }
    r_done = r;
    const double rs = double(r) * G.s;
    if (stop_when_certified && 4. * R2 <= rs * rs) break;   // every vertex is certified
  }
  return r_done;
}

// Emits the vertices made by three bisectors (emit_all, or within the certified radius
// rs: 4 rho^2 <= rs^2) whose particle ids all exceed i, as ascending-id tetrahedra.
template <int N>
__device__ void emit_tets(const V3Cell<N>& C, int nT, int cb, int32_t i, bool emit_all, double rs, const V3Out& O,
                          int lane) {
  constexpr int H = N / 64;
  int32_t tet[H][4];
  bool em[H];
  uint64_t e[H];
  int ne = 0;
#pragma unroll
  for (int h = 0; h < H; h++) {
    em[h] = false;
    const int t = lane + 64 * h;
    if (t < nT) {
      const uint32_t tr = C.tri[cb][t];
      const uint32_t a = tr & 255u, b = (tr >> 8) & 255u, c = (tr >> 16) & 255u;
      const double d2 = C.vx[cb][t] * C.vx[cb][t] + C.vy[cb][t] * C.vy[cb][t] + C.vz[cb][t] * C.vz[cb][t];
      if (a >= kV3Box && b >= kV3Box && c >= kV3Box && (emit_all || 4. * d2 <= rs * rs)) {
        int32_t v[4] = {i, C.pid[a], C.pid[b], C.pid[c]};
        if (i < v[1] && i < v[2] && i < v[3]) {
          // ascending ids
          for (int x = 1; x < 4; x++)
            for (int y = x; y > 0 && v[y - 1] > v[y]; y--) {
              const int32_t s = v[y - 1];
              v[y - 1] = v[y];
              v[y] = s;
            }
          for (int x = 0; x < 4; x++) tet[h][x] = v[x];
          em[h] = true;
        }
      }
    }
    e[h] = __ballot(em[h]);
    ne += __popcll(e[h]);
  }
  if (ne == 0) return;
  unsigned long long base = 0ull;
  if (lane == 0) base = atomicAdd(O.count, (unsigned long long)ne);
  base = __shfl(base, 0, 64);
#pragma unroll
  for (int h = 0; h < H; h++) {
    if (em[h]) {
      const unsigned long long pos = base + lanes_below(e[h], lane);
      if (int64_t(pos) < O.cap)
        for (int x = 0; x < 4; x++) O.tets[4 * int64_t(pos) + x] = tet[h][x];
    }
    base += uint32_t(__popcll(e[h]));
  }
}

__device__ __forceinline__ void cell_of(const V3Grid& G, double xi, double yi, double zi, int* ci) {
  const double p[3] = {xi, yi, zi};
  for (int a = 0; a < 3; a++) {
    const int c = int(floor((p[a] - G.lo[a]) / G.s));
    ci[a] = c < 0 ? 0 : (c >= G.g[a] ? G.g[a] - 1 : c);
  }
}

}  // namespace

__global__ __launch_bounds__(64 * kV3Waves) void k_vor3d(V3Grid G, V3Out O) {
  __shared__ V3Cell<128> cells[kV3Waves];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t k = int64_t(blockIdx.x) * kV3Waves + w;
  if (k >= G.n) return;   // wave-uniform
  V3Cell<128>& C = cells[w];
  const int32_t i = G.order[k];
  const double xi = G.x[3 * int64_t(i)], yi = G.x[3 * int64_t(i) + 1], zi = G.x[3 * int64_t(i) + 2];
  int ci[3];
  cell_of(G, xi, yi, zi, ci);
  // No vertex beyond rmax s / 2 can be certified, so the box need not reach further.
  const double B = 0.5 * double(G.rmax) * G.s;
  init_box(C, B, lane);
  int nT = 8, nP = kV3Box, cb = 0;
  int why = 0;   // status of a cell that could not be built
  double R2 = 3. * B * B;   // squared distance of the farthest vertex
  const int r_done = clip_rings(G, C, nT, nP, cb, i, xi, yi, zi, ci, true, lane, R2, why);
  const double rs = double(r_done) * G.s;
  const bool closed = !why && 4. * R2 <= rs * rs;
  if (lane == 0) {
    O.status[i] = why ? int8_t(why) : (closed ? int8_t(0) : int8_t(1));
    O.rcert[i] = why ? 0. : rs;
  }
  // an open cell emits nothing here: pass 2 builds it whole
  if (why || !closed) return;
  emit_tets(C, nT, cb, i, false, rs, O, lane);
}

// Pass 2: the cells of ids[0..nids) (left open by pass 1), one wavefront each.
__global__ __launch_bounds__(64) void k_vor3d_full(V3Grid G, V3Blocks K, const int32_t* ids, int64_t nids, double B,
                                                   V3Out O) {
  __shared__ V3Cell<256> C;
  const int lane = threadIdx.x;
  const int64_t k = blockIdx.x;
  if (k >= nids) return;
  const int32_t i = ids[k];
  const double xi = G.x[3 * int64_t(i)], yi = G.x[3 * int64_t(i) + 1], zi = G.x[3 * int64_t(i) + 2];
  int ci[3];
  cell_of(G, xi, yi, zi, ci);
  init_box(C, B, lane);
  int nT = 8, nP = kV3Box, cb = 0;
  int why = 0;
  double R2 = 3. * B * B;
  clip_rings(G, C, nT, nP, cb, i, xi, yi, zi, ci, false, lane, R2, why);
  // verify every vertex's sphere
  int iter = 0;
  while (!why) {
    int t = -1;
    for (int h = 0; h < 4; h++) {
      const int tt = lane + 64 * h;
      const uint64_t m = __ballot(tt < nT && !C.ver[cb][tt]);
      if (t < 0 && m) t = 64 * h + __ffsll((unsigned long long)m) - 1;
    }
    if (t < 0) break;   // every vertex verified: the cell is final
    if (++iter > kV3MaxVerify) { why = 2; break; }
    const double vx = C.vx[cb][t], vy = C.vy[cb][t], vz = C.vz[cb][t];
    const double rho2 = vx * vx + vy * vy + vz * vz;
    const double rho = sqrt(rho2);
    const double cc[3] = {xi + vx, yi + vy, zi + vz};   // the sphere's centre
    int b0[3], b1[3];
    int64_t nbk = 1;
    for (int a = 0; a < 3; a++) {
      const double lo = (cc[a] - rho - K.lo[a]) / K.s, hi = (cc[a] + rho - K.lo[a]) / K.s;
      b0[a] = lo < 0. ? 0 : (lo >= double(K.g[a]) ? K.g[a] : int(lo));
      b1[a] = hi < 0. ? -1 : (hi >= double(K.g[a]) ? K.g[a] - 1 : int(hi));
      nbk *= b1[a] >= b0[a] ? int64_t(b1[a] - b0[a] + 1) : 0;
    }
    const int nb0 = b1[0] - b0[0] + 1, nb1 = b1[1] - b0[1] + 1;
    bool clipped = false;
    for (int64_t base = 0; base < nbk && !clipped && !why; base += 64) {
      const int64_t idx = base + lane;
      bool hit = false;
      int32_t st = 0, en = 0;
      if (idx < nbk) {
        const int c0 = b0[0] + int(idx % nb0), c1 = b0[1] + int((idx / nb0) % nb1), c2 = b0[2] + int(idx / (int64_t(nb0) * nb1));
        const int64_t blk = (int64_t(c2) * K.g[1] + c1) * K.g[0] + c0;
        st = K.start[blk];
        en = K.start[blk + 1];
        if (en > st) {
          // squared distance from the centre to the block's point box (with a margin:
          // the box only culls)
          double d2 = 0.;
          for (int a = 0; a < 3; a++) {
            const double lo = K.box[6 * blk + a], hi = K.box[6 * blk + 3 + a];
            const double dd = cc[a] < lo ? lo - cc[a] : (cc[a] > hi ? cc[a] - hi : 0.);
            d2 += dd * dd;
          }
          hit = d2 <= rho2 * (1. + 1.e-9) + 1.e-24;
        }
      }
      uint64_t hm = __ballot(hit);
      while (hm && !clipped && !why) {
        const int L = __ffsll((unsigned long long)hm) - 1;
        hm &= hm - 1ull;
        const int32_t bst = __shfl(st, L, 64), ben = __shfl(en, L, 64);
        for (int32_t kk = bst; kk < ben && !why; kk += 64) {
          int32_t q = -1;
          double ux = 0., uy = 0., uz = 0.;
          bool inside = false;
          if (kk + lane < ben) {
            q = K.order[kk + lane];
            ux = G.x[3 * int64_t(q)] - xi;
            uy = G.x[3 * int64_t(q) + 1] - yi;
            uz = G.x[3 * int64_t(q) + 2] - zi;
            const double u2 = ux * ux + uy * uy + uz * uz;
            inside = q != i && beyond(vx, vy, vz, ux, uy, uz, 0.5 * u2);
          }
          uint64_t im = __ballot(inside);
          while (im && !why) {
            const int L2 = __ffsll((unsigned long long)im) - 1;
            im &= im - 1ull;
            const double px = __shfl(ux, L2, 64), py = __shfl(uy, L2, 64), pz = __shfl(uz, L2, 64);
            const int32_t pq = __shfl(q, L2, 64);
            // a point that already bounds the cell lies on its vertices' spheres
            bool isp = false;
            for (int h = 0; h < 4; h++) {
              const int pp = lane + 64 * h;
              isp = isp || (pp >= kV3Box && pp < nP && C.pid[pp] == pq);
            }
            if (__ballot(isp)) continue;
            const int rc = clip(C, nT, nP, cb, px, py, pz, pq, lane, R2);
            if (rc > 1) why = rc;
            else if (rc == 1) clipped = true;
          }
        }
      }
    }
    if (!clipped && !why) {
      if (lane == 0) C.ver[cb][t] = 1;
      wsync();
    }
  }
  if (lane == 0) {
    O.status[i] = int8_t(why);
    O.rcert[i] = why ? 0. : 1.0e300;
  }
  if (why) return;
  emit_tets(C, nT, cb, i, true, 0., O, lane);
}

void launch_vor3d(hipStream_t st, const V3Grid& G, const V3Out& O) {
  if (G.n <= 0) return;
  const int64_t blocks = (G.n + kV3Waves - 1) / kV3Waves;
  k_vor3d<<<blocks, 64 * kV3Waves, 0, st>>>(G, O);
}

void launch_vor3d_full(hipStream_t st, const V3Grid& G, const V3Blocks& K, const int32_t* ids, int64_t nids, double B,
                       const V3Out& O) {
  if (nids <= 0) return;
  k_vor3d_full<<<nids, 64, 0, st>>>(G, K, ids, nids, B, O);
}

}  // namespace smg
