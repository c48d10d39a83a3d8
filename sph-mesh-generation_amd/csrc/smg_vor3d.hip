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
// Candidates come from a uniform grid of cell size s in rings of cells around the
// particle.  After ring r every point within r s has been clipped.  A cell vertex at
// distance rho is a true Voronoi vertex once every point within 2 rho has been clipped
// (its empty sphere lies in that ball), so the search stops when 2 max(rho) <= r s, or
// at ring rmax.  Only vertices with 2 rho <= (searched radius) are emitted, so every
// emitted tetrahedron is Delaunay, and every Delaunay tetrahedron with a certified
// circumsphere is emitted exactly once, by its smallest particle id.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "smg_kernels.hpp"

namespace smg {

namespace {

constexpr int kV3Planes = 128;   // 6 box planes + the bisectors that cut the cell
constexpr int kV3Tris = 128;     // cell vertices
constexpr int kV3Waves = 4;      // cells per workgroup
constexpr int kV3Box = 6;

struct V3Cell {
  double pn[kV3Planes][4];        // n.xyz, d
  int32_t pid[kV3Planes];         // particle id of a bisector plane
  uint32_t tri[2][kV3Tris];       // a | b << 8 | c << 16, ping-pong buffers
  double vx[2][kV3Tris], vy[2][kV3Tris], vz[2][kV3Tris];
  uint32_t emask[kV3Planes][4];   // directed edges of the removed triangles, bit (x, y)
  uint32_t bnd[kV3Tris];          // boundary edges x | y << 8; plane renumbering in a compaction
  uint32_t ref[kV3Planes / 32];   // planes some vertex uses
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

// Drops the bisector planes that no vertex uses any more and renumbers the rest (the
// box planes keep 0..5); returns the new plane count.
__device__ int compact_planes(V3Cell& C, int nT, int nP, int cb, int lane) {
  if (lane < kV3Planes / 32) C.ref[lane] = 0u;
  wsync();
  for (int h = 0; h < 2; h++) {
    const int t = lane + 64 * h;
    if (t >= nT) continue;
    const uint32_t tr = C.tri[cb][t];
    for (int e = 0; e < 3; e++) {
      const uint32_t a = (tr >> (8 * e)) & 255u;
      atomicOr(&C.ref[a >> 5], 1u << (a & 31u));
    }
  }
  wsync();
  const int p0 = lane, p1 = lane + 64;
  const bool k0 = p0 < nP && (p0 < kV3Box || ((C.ref[p0 >> 5] >> (p0 & 31)) & 1u));
  const bool k1 = p1 < nP && ((C.ref[p1 >> 5] >> (p1 & 31)) & 1u);
  const uint64_t m0 = __ballot(k0), m1 = __ballot(k1);
  const int n0 = __popcll(m0);
  const int q0 = int(lanes_below(m0, lane)), q1 = n0 + int(lanes_below(m1, lane));
  double a0[4], a1[4];
  int32_t i0 = 0, i1 = 0;
  for (int c = 0; c < 4; c++) {
    a0[c] = k0 ? C.pn[p0][c] : 0.;
    a1[c] = k1 ? C.pn[p1][c] : 0.;
  }
  if (k0) i0 = C.pid[p0];
  if (k1) i1 = C.pid[p1];
  if (k0) C.bnd[p0] = uint32_t(q0);
  if (k1) C.bnd[p1] = uint32_t(q1);
  wsync();
  for (int c = 0; c < 4; c++) {
    if (k0) C.pn[q0][c] = a0[c];
    if (k1) C.pn[q1][c] = a1[c];
  }
  if (k0) C.pid[q0] = i0;
  if (k1) C.pid[q1] = i1;
  for (int h = 0; h < 2; h++) {
    const int t = lane + 64 * h;
    if (t >= nT) continue;
    const uint32_t tr = C.tri[cb][t];
    C.tri[cb][t] = C.bnd[tr & 255u] | C.bnd[(tr >> 8) & 255u] << 8 | C.bnd[(tr >> 16) & 255u] << 16;
  }
  wsync();
  return n0 + __popcll(m1);
}

}  // namespace

__global__ __launch_bounds__(64 * kV3Waves) void k_vor3d(V3Grid G, V3Out O) {
  __shared__ V3Cell cells[kV3Waves];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t k = int64_t(blockIdx.x) * kV3Waves + w;
  if (k >= G.n) return;   // wave-uniform
  V3Cell& C = cells[w];
  const int32_t i = G.order[k];
  const double xi = G.x[3 * int64_t(i)], yi = G.x[3 * int64_t(i) + 1], zi = G.x[3 * int64_t(i) + 2];
  int ci[3];
  {
    const double p[3] = {xi, yi, zi};
    for (int a = 0; a < 3; a++) {
      int c = int(floor((p[a] - G.lo[a]) / G.s));
      ci[a] = c < 0 ? 0 : (c >= G.g[a] ? G.g[a] - 1 : c);
    }
  }
  // the box |p|_inf <= B around the particle: 6 planes, 8 vertices.  No vertex beyond
  // rmax s / 2 can be certified, so the box need not reach further.
  const double B = 0.5 * double(G.rmax) * G.s;
  if (lane < kV3Box) {
    const int ax = lane >> 1;
    const double sg = (lane & 1) ? -1. : 1.;
    C.pn[lane][0] = ax == 0 ? sg : 0.;
    C.pn[lane][1] = ax == 1 ? sg : 0.;
    C.pn[lane][2] = ax == 2 ? sg : 0.;
    C.pn[lane][3] = B;
    C.pid[lane] = -1;
  }
  for (int q = lane; q < kV3Planes * 4; q += 64) (&C.emask[0][0])[q] = 0u;
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
  }
  wsync();
  int nT = 8, nP = kV3Box, cb = 0;
  bool bad = false;
  int why = 0;   // status of a cell that could not be built
  double R2 = 3. * B * B;   // squared distance of the farthest vertex
  int r_done = -1;
  for (int r = 0; r <= G.rmax && !bad; r++) {
    const int side = 2 * r + 1;
    const int ncube = side * side * side;
    for (int base = 0; base < ncube && !bad; base += 64) {
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
      for (int t = 0; t < mc && !bad; t++) {
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
        while (m && !bad) {
          const int L = __ffsll((unsigned long long)m) - 1;
          m &= m - 1ull;
          const double px = __shfl(ux, L, 64), py = __shfl(uy, L, 64), pz = __shfl(uz, L, 64);
          const int32_t pq = __shfl(q, L, 64);
          const double u2 = px * px + py * py + pz * pz;
          if (!(u2 < 4. * R2)) continue;   // the cell shrank since the lane's test
          const double c = 0.5 * u2;
          // classify the vertices
          const int t0 = lane, t1 = lane + 64;
          bool r0 = false, r1 = false;
          if (t0 < nT) r0 = C.vx[cb][t0] * px + C.vy[cb][t0] * py + C.vz[cb][t0] * pz - c > 0.;
          if (t1 < nT) r1 = C.vx[cb][t1] * px + C.vy[cb][t1] * py + C.vz[cb][t1] * pz - c > 0.;
          const uint64_t m0 = __ballot(r0), m1 = __ballot(r1);
          if ((m0 | m1) == 0ull) continue;   // the plane misses the cell
          if (nP >= kV3Planes) nP = compact_planes(C, nT, nP, cb, lane);
          if (nP >= kV3Planes) { bad = true; why = 2; break; }
          const int p = nP;
          if (lane == 0) {
            C.pn[p][0] = px; C.pn[p][1] = py; C.pn[p][2] = pz; C.pn[p][3] = c;
            C.pid[p] = pq;
            C.nb = 0;
          }
          // mark the directed edges of the removed triangles
          uint32_t my[2] = {r0 ? C.tri[cb][t0] : 0xffffffffu, r1 ? C.tri[cb][t1] : 0xffffffffu};
          for (int h = 0; h < 2; h++) {
            if (my[h] == 0xffffffffu) continue;
            const uint32_t v[3] = {my[h] & 255u, (my[h] >> 8) & 255u, (my[h] >> 16) & 255u};
            for (int e = 0; e < 3; e++) {
              const uint32_t x = v[e], y = v[(e + 1) % 3];
              atomicOr(&C.emask[x][y >> 5], 1u << (y & 31u));
            }
          }
          wsync();
          // boundary: a removed edge x -> y whose twin y -> x is not removed
          for (int h = 0; h < 2; h++) {
            if (my[h] == 0xffffffffu) continue;
            const uint32_t v[3] = {my[h] & 255u, (my[h] >> 8) & 255u, (my[h] >> 16) & 255u};
            for (int e = 0; e < 3; e++) {
              const uint32_t x = v[e], y = v[(e + 1) % 3];
              if (!((C.emask[y][x >> 5] >> (x & 31u)) & 1u)) {
                const int s = atomicAdd(&C.nb, 1);
                if (s < kV3Tris) C.bnd[s] = x | y << 8;
              }
            }
          }
          wsync();
          for (int h = 0; h < 2; h++) {
            if (my[h] == 0xffffffffu) continue;
            const uint32_t v[3] = {my[h] & 255u, (my[h] >> 8) & 255u, (my[h] >> 16) & 255u};
            for (int e = 0; e < 3; e++) C.emask[v[e]][v[(e + 1) % 3] >> 5] = 0u;
          }
          const int nb = C.nb;
          const uint64_t valid0 = nT >= 64 ? ~0ull : ((1ull << nT) - 1ull);
          const uint64_t valid1 = nT >= 128 ? ~0ull : (nT > 64 ? ((1ull << (nT - 64)) - 1ull) : 0ull);
          const uint64_t keep0 = valid0 & ~m0, keep1 = valid1 & ~m1;
          const int nK = __popcll(keep0) + __popcll(keep1);
          if (nK + nb > kV3Tris || nb > 64) { bad = true; why = 2; break; }
          if (nb < 3) { bad = true; why = 3; break; }
          // a plane starts at most one boundary edge of a convex cut
          bool dup = false;
          if (lane < nb) {
            const uint32_t x = C.bnd[lane] & 255u;
            for (int s = 0; s < nb; s++) dup = dup || (s != lane && (C.bnd[s] & 255u) == x);
          }
          if (__ballot(dup)) { bad = true; why = 4; break; }
          const int nx = cb ^ 1;
          if ((keep0 >> lane) & 1ull) {
            const int pos = int(lanes_below(keep0, lane));
            C.tri[nx][pos] = C.tri[cb][t0];
            C.vx[nx][pos] = C.vx[cb][t0]; C.vy[nx][pos] = C.vy[cb][t0]; C.vz[nx][pos] = C.vz[cb][t0];
          }
          if ((keep1 >> lane) & 1ull) {
            const int pos = __popcll(keep0) + int(lanes_below(keep1, lane));
            C.tri[nx][pos] = C.tri[cb][t1];
            C.vx[nx][pos] = C.vx[cb][t1]; C.vy[nx][pos] = C.vy[cb][t1]; C.vz[nx][pos] = C.vz[cb][t1];
          }
          bool okv = true;
          if (lane < nb) {
            const uint32_t x = C.bnd[lane] & 255u, y = (C.bnd[lane] >> 8) & 255u;
            double vx, vy, vz;
            meet(C.pn[x], C.pn[y], C.pn[p], vx, vy, vz, okv);
            C.tri[nx][nK + lane] = x | y << 8 | uint32_t(p) << 16;
            C.vx[nx][nK + lane] = vx; C.vy[nx][nK + lane] = vy; C.vz[nx][nK + lane] = vz;
          }
          if (__ballot(!okv)) { bad = true; why = 5; break; }
          wsync();
          nT = nK + nb;
          nP = p + 1;
          cb = nx;
          double d2 = 0.;
          if (lane < nT) d2 = C.vx[cb][lane] * C.vx[cb][lane] + C.vy[cb][lane] * C.vy[cb][lane] + C.vz[cb][lane] * C.vz[cb][lane];
          if (lane + 64 < nT) {
            const int t = lane + 64;
            d2 = fmax(d2, C.vx[cb][t] * C.vx[cb][t] + C.vy[cb][t] * C.vy[cb][t] + C.vz[cb][t] * C.vz[cb][t]);
          }
          R2 = wave_max(d2);
        }
      }
    }
    r_done = r;
    const double rs = double(r) * G.s;
    if (4. * R2 <= rs * rs) break;   // every vertex is certified
  }
  const double rs = double(r_done) * G.s;
  const bool closed = !bad && 4. * R2 <= rs * rs;
  if (lane == 0) {
    O.status[i] = bad ? int8_t(why) : (closed ? int8_t(0) : int8_t(1));
    O.rcert[i] = bad ? 0. : rs;
  }
  if (bad) return;
  // emit the certified vertices made by three bisectors, by the smallest id
  int32_t tet[2][4];
  bool em[2] = {false, false};
  for (int h = 0; h < 2; h++) {
    const int t = lane + 64 * h;
    if (t >= nT) continue;
    const uint32_t tr = C.tri[cb][t];
    const uint32_t a = tr & 255u, b = (tr >> 8) & 255u, c = (tr >> 16) & 255u;
    if (a < kV3Box || b < kV3Box || c < kV3Box) continue;
    const double d2 = C.vx[cb][t] * C.vx[cb][t] + C.vy[cb][t] * C.vy[cb][t] + C.vz[cb][t] * C.vz[cb][t];
    if (!(4. * d2 <= rs * rs)) continue;
    int32_t v[4] = {i, C.pid[a], C.pid[b], C.pid[c]};
    if (!(i < v[1] && i < v[2] && i < v[3])) continue;
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
  const uint64_t e0 = __ballot(em[0]), e1 = __ballot(em[1]);
  const int ne = __popcll(e0) + __popcll(e1);
  if (ne == 0) return;
  unsigned long long base = 0ull;
  if (lane == 0) base = atomicAdd(O.count, (unsigned long long)ne);
  base = __shfl(base, 0, 64);
  for (int h = 0; h < 2; h++) {
    if (!em[h]) continue;
    const unsigned long long pos =
        base + (h == 0 ? lanes_below(e0, lane) : uint32_t(__popcll(e0)) + lanes_below(e1, lane));
    if (int64_t(pos) < O.cap)
      for (int x = 0; x < 4; x++) O.tets[4 * int64_t(pos) + x] = tet[h][x];
  }
}

void launch_vor3d(hipStream_t st, const V3Grid& G, const V3Out& O) {
  if (G.n <= 0) return;
  const int64_t blocks = (G.n + kV3Waves - 1) / kV3Waves;
  k_vor3d<<<blocks, 64 * kV3Waves, 0, st>>>(G, O);
}

}  // namespace smg
