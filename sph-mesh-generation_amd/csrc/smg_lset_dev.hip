// smg_lset_dev.hip -- the level-set finishing of smg_lset_build on the GPU.
//
// Input: cell-centred phi on the grid with one full margin package per side.  Output:
// phi, normal and cut tag on the one-cell-halo grid and the particle budget, equal to
// the host builder's (csrc/smg_lset.cpp): every value is formed by the same sequence
// of IEEE operations (no contraction), so phi, normals, tags, volume, area and the
// uniform- and linear-profile masses are bit-identical; with the tanh profile the
// per-cell spacing goes through the device tanh (within 2 ulp of the host's).
//
//   k_lset_grid  one thread per halo-grid cell: normal (Levelset_package::Get_normal /
//                Get_BC_normal, lset_package.cpp:332-389), phi, cut tag (Is_cut_cell,
//                :664-761 with Get_phi_at_corners), maximum_phi (Get_global_scale,
//                :765-785), and the list of cut cells the particle count measures.
//   k_lset_cut   the cut cells' volume fraction and isosurface area (Calculate_volume_mass,
//                lset_package.cpp:1283-1396, with GeometryCalculator's adaptive
//                subdivision, geometry_calculator.cpp:12-566): 3D, one wavefront per cell,
//                a lane per second-level sub-cube (8 sub-cubes x 8 children) walking its
//                subtree depth first; 2D, 16 lanes per cell (4 sub-squares x 4).
//                Every term of the volume fraction is a multiple of 2^-23 (3D) / 2^-16
//                (2D) no larger than 1, so its sum is exact in any order.
//   k_lset_pkg   one wavefront per group of packages: the per-cell volume, area, mass
//                and surface mass, summed inside each package in the cell order of
//                Calculate_volume_mass (k fastest) as the host builder does.
// The package sums come back to the host and are added in package order
// (Get_num_particles, sph_mesh_generation.cpp:900-941).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "smg.h"
#include "smg_internal.hpp"
#include "smg_mc.hpp"

namespace smg {
namespace {

struct LGrid {
  const double* ext;   // phi, full-margin grid, x fastest
  int D;
  int nc[3], ext_n[3], halo[3];
  double dcell[3];
  double dl;           // smallest cell size
};

struct LProf {
  int profile;
  double min_dl, max_dl, tanh25, dx0, dx1, band, dl;
  double max_phi_cfg;  // > 0: the config's maximum_phi, else the measured one
};

struct LCount {
  unsigned long long max_phi_bits;   // max |phi| over interior phi > 0 (bit pattern)
  unsigned long long n_cut;          // interior cut cells
  unsigned long long n_list;         // cut cells the particle count measures
  unsigned long long n_queue;        // split lev-2 nodes of 3D cut cells
};

// the full-margin grid holds fewer than 2^31 cells (checked on the host): 32-bit indexing
__device__ __forceinline__ double at(const LGrid& g, int i, int j, int k) {
  const int a = i + 4, b = g.D > 1 ? j + 4 : 0, c = g.D > 2 ? k + 4 : 0;
  return g.ext[(c * g.ext_n[1] + b) * g.ext_n[0] + a];
}

__device__ __forceinline__ int sign3(double v) { return v > 0. ? 1 : (v < 0. ? -1 : 0); }

__device__ __forceinline__ void cswap(double& a, double& b) {
  const double lo = b < a ? b : a, hi = b < a ? a : b;
  a = lo;
  b = hi;
}

// mean of the 8 cells around a corner, summed in ascending order (sorting network)
__device__ double corner_value(const LGrid& g, int i, int j, int k, int sx, int sy, int sz) {
  const int z0 = k, z1 = g.D > 2 ? k + sz : k;
  const int y0 = g.D > 1 ? j + sy : j, y1 = j;
  const int x0 = i + sx, x1 = i;
  double v[8] = {at(g, x0, y0, z0), at(g, x1, y0, z0), at(g, x0, y1, z0), at(g, x1, y1, z0),
                 at(g, x0, y0, z1), at(g, x1, y0, z1), at(g, x0, y1, z1), at(g, x1, y1, z1)};
  // Batcher's odd-even merge sort for 8 inputs (19 compare-exchanges)
  cswap(v[0], v[1]); cswap(v[2], v[3]); cswap(v[4], v[5]); cswap(v[6], v[7]);
  cswap(v[0], v[2]); cswap(v[1], v[3]); cswap(v[4], v[6]); cswap(v[5], v[7]);
  cswap(v[1], v[2]); cswap(v[5], v[6]);
  cswap(v[0], v[4]); cswap(v[1], v[5]); cswap(v[2], v[6]); cswap(v[3], v[7]);
  cswap(v[2], v[4]); cswap(v[3], v[5]);
  cswap(v[1], v[2]); cswap(v[3], v[4]); cswap(v[5], v[6]);
  double s = 0.0;
#pragma unroll
  for (int q = 0; q < 8; q++) s += v[q];
  return s * 0.125;
}

__device__ bool cut_cell(const LGrid& g, int i, int j, int k) {
  const int s0 = sign3(at(g, i, j, k));
  if (s0 != 0) {   // all 3^D cells around of one sign: no corner can change sign
    const int zr = g.D > 2 ? 1 : 0, yr = g.D > 1 ? 1 : 0;
    // every load issued up front (no short circuit: a chain of dependent loads)
    bool same = true;
    for (int c = -zr; c <= zr; c++)
      for (int b = -yr; b <= yr; b++)
#pragma unroll
        for (int a = -1; a <= 1; a++) same &= sign3(at(g, i + a, j + b, k + c)) == s0;
    if (same) return false;
  }
  // the corner order of Get_phi_at_corners
  const int cs[8][3] = {{-1, -1, -1}, {-1, -1, 1}, {1, -1, 1}, {1, -1, -1},
                        {-1, 1, -1}, {-1, 1, 1}, {1, 1, 1}, {1, 1, -1}};
  const int c0 = sign3(corner_value(g, i, j, k, cs[0][0], cs[0][1], cs[0][2]));
  for (int u = 1; u < 8; u++)
    if (sign3(corner_value(g, i, j, k, cs[u][0], cs[u][1], cs[u][2])) != c0) return true;
  return s0 != c0;
}

__global__ __launch_bounds__(256) void k_lset_grid(LGrid g, int zero_margin, double* __restrict__ phi,
                                                   float* __restrict__ normal, int8_t* __restrict__ tag,
                                                   int32_t* __restrict__ list, LCount* __restrict__ cnt,
                                                   double* __restrict__ bmax) {
  // block (64, 4): 64 cells along x, 4 rows of y; blockIdx.z = z
  const int hx = int(blockIdx.x) * 64 + int(threadIdx.x & 63);
  const int hy = int(blockIdx.y) * 4 + int(threadIdx.x >> 6);
  const int hz = int(blockIdx.z);
  const bool live = hx < g.halo[0] && hy < g.halo[1];
  const int64_t o = (int64_t(hz) * g.halo[1] + hy) * g.halo[0] + hx;
  double pmax = 0.0;
  bool is_cut = false, listed = false;
  int32_t cell = 0;
  if (live) {
    const int id3[3] = {hx - 1, g.D > 1 ? hy - 1 : 0, g.D > 2 ? hz - 1 : 0};
    const int i = id3[0], j = id3[1], k = id3[2];
    const bool in = i >= 0 && i < g.nc[0] && (g.D < 2 || (j >= 0 && j < g.nc[1])) &&
                    (g.D < 3 || (k >= 0 && k < g.nc[2]));
    double grad[3] = {0.0, 0.0, 0.0};
    for (int a = 0; a < g.D; a++) {
      int lo3[3] = {i, j, k}, hi3[3] = {i, j, k};
      bool central = true;
      if (!in) {   // margin package: one-sided at its own faces
        const int c = id3[a];
        const int local = c & 3;   // place in the package, also for c < 0
        if (local == 0) { hi3[a] = c + 1; central = false; }
        else if (local == 3) { lo3[a] = c - 1; central = false; }
        else { lo3[a] = c - 1; hi3[a] = c + 1; }
      } else {
        lo3[a] = id3[a] - 1;
        hi3[a] = id3[a] + 1;
      }
      const double diff = at(g, hi3[0], hi3[1], hi3[2]) - at(g, lo3[0], lo3[1], lo3[2]);
      grad[a] = (central ? 0.5 * diff : diff) / g.dcell[a];
    }
    double len = 0.0;
    for (int a = 0; a < g.D; a++) len += grad[a] * grad[a];
    len = sqrt(len) + 1.e-12;
    const double inv = 1. / len;
    const bool blank = !in && zero_margin;
    for (int a = 0; a < 3; a++) normal[3 * o + a] = (blank || a >= g.D) ? 0.0f : float(grad[a] * inv);
    const double p = at(g, i, j, k);
    phi[o] = blank ? 0.0 : p;
    is_cut = in && cut_cell(g, i, j, k);
    tag[o] = is_cut ? 1 : 0;
    if (in) {
      if (p > 0.) pmax = fabs(p);
      listed = is_cut && (p > 0. || fabs(p) < g.dl);
      cell = int32_t((int64_t(k) * (g.D > 1 ? g.nc[1] : 1) + j) * g.nc[0] + i);
    }
  }
  // block-aggregated counters: one atomic per block with cut cells, and the block's
  // maximum phi to bmax[block] (a global maximum over all blocks by atomics on one
  // address would serialise)
  __shared__ unsigned int s_cut[4], s_list[4];
  __shared__ double s_max[4];
  __shared__ unsigned long long s_base;
  const unsigned long long cut_mask = __ballot(is_cut);
  const unsigned long long list_mask = __ballot(listed);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int off = 32; off > 0; off >>= 1) pmax = fmax(pmax, __shfl_xor(pmax, off));
  if (lane == 0) {
    s_cut[wave] = (unsigned int)__popcll(cut_mask);
    s_list[wave] = (unsigned int)__popcll(list_mask);
    s_max[wave] = pmax;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned int nc = s_cut[0] + s_cut[1] + s_cut[2] + s_cut[3];
    const unsigned int nl = s_list[0] + s_list[1] + s_list[2] + s_list[3];
    if (nc) atomicAdd(&cnt->n_cut, (unsigned long long)nc);
    s_base = nl ? atomicAdd(&cnt->n_list, (unsigned long long)nl) : 0ull;
    bmax[(int64_t(blockIdx.z) * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] =
        fmax(fmax(s_max[0], s_max[1]), fmax(s_max[2], s_max[3]));
  }
  __syncthreads();
  unsigned long long base = s_base;
  for (int w = 0; w < wave; w++) base += s_list[w];
  if (listed) list[base + __popcll(list_mask & ((1ull << lane) - 1ull))] = cell;
}

// maximum of the blocks' maxima (one block)
__global__ __launch_bounds__(1024) void k_lset_max(const double* __restrict__ bmax, int64_t nb,
                                                   LCount* __restrict__ cnt) {
  __shared__ double sm[16];
  double m = 0.0;
  for (int64_t b = threadIdx.x; b < nb; b += blockDim.x) m = fmax(m, bmax[b]);
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < int(blockDim.x >> 6); w++) m = fmax(m, sm[w]);
    cnt->max_phi_bits = (unsigned long long)__double_as_longlong(m);
  }
}

// corner / edge / face / centre samples of a cut cell from its 3^D neighbourhood
// (GeometryCalculator::Get_corner), in the evaluation order of the host builder
template <int D>
__device__ __forceinline__ void cell_samples(const LGrid& g, int i, int j, int k, double C[3][3][3]) {
  constexpr int dy = D > 1, dz = D > 2;
  double P[3][3][3];
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++)
#pragma unroll
      for (int e = 0; e < 3; e++) {
        P[a][b][e] = (b < (dy ? 3 : 1) && e < (dz ? 3 : 1)) ? at(g, i - 1 + a, dy ? j - 1 + b : 0, dz ? k - 1 + e : 0)
                                                           : 0.0;
        C[a][b][e] = 0.0;
      }
  C[1][dy][dz] = P[1][dy][dz];
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 1 + dy; b++)
      for (int e = 0; e < 1 + dz; e++)
        C[2 * a][2 * b][2 * e] = 0.125 * (P[a][b][e] + P[a + 1][b][e] + P[a][b + dy][e] + P[a + 1][b + dy][e] +
                                          P[a][b][e + dz] + P[a + 1][b][e + dz] + P[a][b + dy][e + dz] +
                                          P[a + 1][b + dy][e + dz]);
  if (dz)
    for (int a = 0; a < 2; a++)
      for (int b = 0; b < 2; b++)
        C[2 * a][2 * b][1] = 0.25 * (P[a][b][1] + P[a + 1][b][1] + P[a][b + 1][1] + P[a + 1][b + 1][1]);
  for (int b = 0; b < 1 + dy; b++)
    for (int e = 0; e < 1 + dz; e++)
      C[1][2 * b][2 * e] = 0.25 * (P[1][b][e] + P[1][b][e + dz] + P[1][b + dy][e] + P[1][b + dy][e + dz]);
  if (dy)
    for (int a = 0; a < 2; a++)
      for (int e = 0; e < 1 + dz; e++)
        C[2 * a][1][2 * e] = 0.25 * (P[a][1][e] + P[a + 1][1][e] + P[a][1][e + dz] + P[a + 1][1][e + dz]);
  if (dz)
    for (int a = 0; a < 2; a++) {
      C[2 * a][1][1] = 0.5 * (P[a][1][1] + P[a + 1][1][1]);
      C[1][2 * a][1] = 0.5 * (P[1][a][1] + P[1][a + 1][1]);
      C[1][1][2 * a] = 0.5 * (P[1][1][a] + P[1][1][a + 1]);
    }
}

// ---- 3D volume fraction: sub-cube corners c[q], q = 4 qx + 2 qy + qz ----
//
// Node weights are 2^-(3 + 3 lev): lev 0 (the cell's 8 half-size sub-cubes) 2^-3 down
// to the leaves at lev 6, 2^-21.

// the refined 3x3x3 samples of a cube, midpoints formed in the host builder's order
// (z edges, then y, then x)
__device__ __forceinline__ void refine3(const double* c, double r[3][3][3]) {
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int e = 0; e < 2; e++) r[2 * a][2 * b][2 * e] = c[4 * a + 2 * b + e];
#pragma unroll
  for (int a = 0; a < 3; a += 2)
#pragma unroll
    for (int b = 0; b < 3; b += 2) r[a][b][1] = 0.5 * (r[a][b][0] + r[a][b][2]);
#pragma unroll
  for (int e = 0; e < 3; e++)
#pragma unroll
    for (int a = 0; a < 3; a += 2) r[a][1][e] = 0.5 * (r[a][0][e] + r[a][2][e]);
#pragma unroll
  for (int b = 0; b < 3; b++)
#pragma unroll
    for (int e = 0; e < 3; e++) r[1][b][e] = 0.5 * (r[0][b][e] + r[2][b][e]);
}

// corners of child ch = 4 ca + 2 cb + ce of a refined 3x3x3 set, r[ca + p][cb + q][ce + t],
// by selection one axis at a time (ch may differ between lanes; no dynamic indexing)
__device__ __forceinline__ void select_child(const double r[3][3][3], int ch, double* out) {
  const bool ca = ch & 4, cb = ch & 2, ce = ch & 1;
  double tx[2][3][3], ty[2][2][3];
#pragma unroll
  for (int p = 0; p < 2; p++)
#pragma unroll
    for (int y = 0; y < 3; y++)
#pragma unroll
      for (int z = 0; z < 3; z++) tx[p][y][z] = ca ? r[1 + p][y][z] : r[p][y][z];
#pragma unroll
  for (int p = 0; p < 2; p++)
#pragma unroll
    for (int q = 0; q < 2; q++)
#pragma unroll
      for (int z = 0; z < 3; z++) ty[p][q][z] = cb ? tx[p][1 + q][z] : tx[p][q][z];
#pragma unroll
  for (int p = 0; p < 2; p++)
#pragma unroll
    for (int q = 0; q < 2; q++)
#pragma unroll
      for (int t = 0; t < 2; t++) out[4 * p + 2 * q + t] = ce ? ty[p][q][1 + t] : ty[p][q][t];
}

// corners of child ch of a cube (ch may differ between lanes): the same midpoints as
// refine3, formed only for the child's quadrant.  Per x-plane of the parent: the z
// midpoints of both rows, the y midpoints of both columns and the centre; the quadrant
// (cb, ce) picks four of the nine; the x midpoints of the two planes' picks complete it.
__device__ __forceinline__ void child3(const double* c, int ch, double* out) {
  const bool ca = ch & 4, cb = ch & 2, ce = ch & 1;
  double v[2][2][2];   // [x plane][q][t]
#pragma unroll
  for (int x = 0; x < 2; x++) {
    const double p00 = c[4 * x], p01 = c[4 * x + 1], p10 = c[4 * x + 2], p11 = c[4 * x + 3];
    const double zm0 = 0.5 * (p00 + p01), zm1 = 0.5 * (p10 + p11);
    const double ym0 = 0.5 * (p00 + p10), ym1 = 0.5 * (p01 + p11);
    const double yzm = 0.5 * (zm0 + zm1);
    v[x][0][0] = cb ? (ce ? yzm : ym0) : (ce ? zm0 : p00);
    v[x][0][1] = cb ? (ce ? ym1 : yzm) : (ce ? p01 : zm0);
    v[x][1][0] = cb ? (ce ? zm1 : p10) : (ce ? yzm : ym0);
    v[x][1][1] = cb ? (ce ? p11 : zm1) : (ce ? ym1 : yzm);
  }
#pragma unroll
  for (int q = 0; q < 2; q++)
#pragma unroll
    for (int t = 0; t < 2; t++) {
      const double xm = 0.5 * (v[0][q][t] + v[1][q][t]);
      out[2 * q + t] = ca ? xm : v[0][q][t];
      out[4 + 2 * q + t] = ca ? v[1][q][t] : xm;
    }
}

// value of a node with no further split: full, empty, or the leaf rule at depth > 5
// (corner sum in the host's order); returns false when the node must be split
__device__ __forceinline__ bool node3(const double* c, int lev, double w, double& v) {
  int f = 0;
#pragma unroll
  for (int q = 0; q < 8; q++) f += (c[q] > 0.0) - (c[q] < 0.0);
  if (f == 8) { v = w; return true; }
  if (f == -8) { v = 0.0; return true; }
  if (lev > 5) {
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < 8; q++) s += c[q];
    v = s > 0.0 ? 0.75 * w : (s == 0.0 ? 0.5 * w : 0.25 * w);
    return true;
  }
  return false;
}

// the 8 leaves (lev 6) of a split lev-5 node
__device__ __forceinline__ double leaves6(const double* c) {
  double r[3][3][3];
  refine3(c, r);
  double acc = 0.0;
#pragma unroll
  for (int ch = 0; ch < 8; ch++) {
    const int ca = ch >> 2, cb = (ch >> 1) & 1, ce = ch & 1;
    double s[8];
#pragma unroll
    for (int p = 0; p < 2; p++)
#pragma unroll
      for (int q = 0; q < 2; q++)
#pragma unroll
        for (int t = 0; t < 2; t++) s[4 * p + 2 * q + t] = r[ca + p][cb + q][ce + t];
    double v;
    node3(s, 6, 0x1p-21, v);
    acc += v;
  }
  return acc;
}

// the lev-2 node of a cut cell at path = 64 sub-cube + 8 lev-1 child + lev-2 child
__device__ __forceinline__ void node2_of(const double C[3][3][3], int path, double* c2) {
  double s0[8], c1[8];
  select_child(C, path >> 6, s0);
  child3(s0, (path >> 3) & 7, c1);
  child3(c1, path & 7, c2);
}

// one queued lev-2 node
struct Q2 {
  int32_t cell, path;
};

// the marching-cubes triangle table (smg_mc.hpp), for the cut-cell area
__constant__ uint64_t kMcTableD[256] = SMG_MC_TRI_TABLE;

// get_area_MQ of a cell with corner samples C over a group of lanes: lane t < 8 of the
// group evaluates sub-cube t (sub-cube indices i, j, k with k fastest); every lane of the
// group then adds all triangles in the reference's order, as mc_cell_area does
__device__ double mc_area_lanes(const double C[3][3][3], double dx, double dy, double dz, int t, int width) {
  double a[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  int n = 0;
  if (t < 8) n = mc_subcube_areas(C, t >> 2, (t >> 1) & 1, t & 1, dx, dy, dz, kMcTableD, a);
  double area = 0.0;
  for (int sb = 0; sb < 8; sb++) {
    const int ns = __shfl(n, sb, width);
#pragma unroll
    for (int q = 0; q < 5; q++) {
      const double v = __shfl(a[q], sb, width);
      if (q < ns) area += v;
    }
  }
  return area;
}

constexpr int kCutWaves = 4;


// 3D, first pass: one wavefront per listed cut cell, a lane per lev-1 node (8 sub-cubes
// x 8 children).  Levels 0-2 are evaluated here and split lev-2 nodes go to the queue;
// the block reserves its queue range with one atomic.  Writes the partial volume
// fraction and the isosurface area (plain stores, so the pass can be run again).
__global__ __launch_bounds__(64 * kCutWaves) void k_lset_cut3a(LGrid g, const int32_t* __restrict__ list,
                                                               int64_t nlist, double* __restrict__ hsum,
                                                               double* __restrict__ carea, Q2* __restrict__ queue,
                                                               int64_t qcap, unsigned long long* __restrict__ qn) {
  __shared__ unsigned int s_tot[kCutWaves];
  __shared__ unsigned long long s_base;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t w = int64_t(blockIdx.x) * kCutWaves + wave;
  const bool live = w < nlist;   // a wavefront past the list works on cell 0 and writes nothing
  const int32_t cell = live ? list[w] : 0;
  const int i = cell % g.nc[0], j = (cell / g.nc[0]) % g.nc[1], k = cell / (g.nc[0] * g.nc[1]);
  double C[3][3][3];
  cell_samples<3>(g, i, j, k, C);
  const double dx = g.dcell[0], dy = g.dcell[1], dz = g.dcell[2];
  int f = 0;
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++)
#pragma unroll
      for (int e = 0; e < 3; e++) f += (C[a][b][e] > 0.0) - (C[a][b][e] < 0.0);
  double H = f == 27 ? 1.0 : 0.0, area = 0.0;
  const int sc = lane >> 3, ch1 = lane & 7;
  double c1[8];
  int nopen = 0;
  if (live && f != 27 && f != -27) {   // uniform over the wavefront
    double s0[8], v;
    select_child(C, sc, s0);
    double part = 0.0;
    if (node3(s0, 0, 0x1p-3, v)) {
      part = ch1 == 0 ? v : 0.0;
    } else {
      child3(s0, ch1, c1);
      if (node3(c1, 1, 0x1p-6, v)) {
        part = v;
      } else {
#pragma unroll 1
        for (int ch2 = 0; ch2 < 8; ch2++) {
          double c2[8];
          child3(c1, ch2, c2);
          if (node3(c2, 2, 0x1p-9, v)) part += v;
          else nopen++;
        }
      }
    }
    for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off);
    H = part;
    area = mc_area_lanes(C, dx, dy, dz, lane, 64);
  }
  // queue range of the block: inclusive scan of nopen over the wavefront, then the
  // wavefronts' totals
  int inc = nopen;
  for (int off = 1; off < 64; off <<= 1) {
    const int y = __shfl_up(inc, off);
    if (lane >= off) inc += y;
  }
  if (lane == 63) s_tot[wave] = (unsigned int)inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long tot = 0;
    for (int q = 0; q < kCutWaves; q++) tot += s_tot[q];
    s_base = tot ? atomicAdd(qn, tot) : 0ull;
  }
  __syncthreads();
  if (nopen > 0) {
    unsigned long long slot = s_base + (unsigned long long)(inc - nopen);
    for (int q = 0; q < wave; q++) slot += s_tot[q];
#pragma unroll 1
    for (int ch2 = 0; ch2 < 8; ch2++) {
      double c2[8], v;
      child3(c1, ch2, c2);
      if (node3(c2, 2, 0x1p-9, v)) continue;
      // past the capacity the entry is dropped; the host sees the count, grows the
      // queue and runs this pass again
      if (slot < (unsigned long long)qcap) queue[slot] = Q2{cell, (sc << 6) | (ch1 << 3) | ch2};
      slot++;
    }
  }
  if (live && lane == 0) {
    hsum[cell] = H;
    carea[cell] = area;
  }
}

// 3D, second pass: one lane per queued lev-2 node.  Each lane walks levels 3 and 4 of its
// subtree depth first (its open lev-3 node in registers); split lev-4 nodes go to a
// per-wavefront LDS buffer, and whenever it holds 64 the whole wavefront expands 64 of
// them at once (the 8 children of each from one refinement); their split lev-5 children
// go to a second buffer, expanded 64 at a time into their 8 leaves.  Results are added
// to the owning lane's sum with LDS atomics: every term is an exact dyadic number, so the
// order of the additions does not matter.  Each cell gets one global atomic per
// wavefront.
constexpr int kBuf = 128;

struct NodeBuf {
  double node[kBuf][8];
  int32_t own[kBuf];
};

template <int CH>
__device__ __forceinline__ void pick(const double r[3][3][3], double* out) {
  constexpr int ca = CH >> 2, cb = (CH >> 1) & 1, ce = CH & 1;
#pragma unroll
  for (int p = 0; p < 2; p++)
#pragma unroll
    for (int q = 0; q < 2; q++)
#pragma unroll
      for (int t = 0; t < 2; t++) out[4 * p + 2 * q + t] = r[ca + p][cb + q][ce + t];
}

// child ch of a refined set, ch the same in every lane (a scalar branch, no selects)
__device__ __forceinline__ void pick_uniform(const double r[3][3][3], int ch, double* out) {
  switch (ch) {
    case 0: pick<0>(r, out); break;
    case 1: pick<1>(r, out); break;
    case 2: pick<2>(r, out); break;
    case 3: pick<3>(r, out); break;
    case 4: pick<4>(r, out); break;
    case 5: pick<5>(r, out); break;
    case 6: pick<6>(r, out); break;
    default: pick<7>(r, out); break;
  }
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// append the nodes of the lanes that want to the buffer (nb < 64 on entry)
__device__ __forceinline__ void push_nodes(NodeBuf& B, int& nb, bool want, const double* n, int owner, int lane) {
  const unsigned long long m = __ballot(want);
  if (want) {
    const int slot = nb + __popcll(m & ((1ull << lane) - 1ull));
#pragma unroll
    for (int q = 0; q < 8; q++) B.node[slot][q] = n[q];
    B.own[slot] = owner;
  }
  nb += __popcll(m);
  wave_sync();
}

// the leaves of the top min(64, n5) lev-5 nodes
__device__ __forceinline__ void drain5(NodeBuf& B5, int& n5, double (&accl)[64], int lane) {
  const int take = n5 >= 64 ? 64 : n5;
  const int base = n5 - take;
  if (lane < take) {
    double c[8];
#pragma unroll
    for (int q = 0; q < 8; q++) c[q] = B5.node[base + lane][q];
    atomicAdd(&accl[B5.own[base + lane]], leaves6(c));
  }
  n5 = base;
  wave_sync();
}

// the children of the top min(64, n4) lev-4 nodes; split ones go to B5
__device__ __forceinline__ void drain4(NodeBuf& B4, int& n4, NodeBuf& B5, int& n5, double (&accl)[64], int lane) {
  const int take = n4 >= 64 ? 64 : n4;
  const int base = n4 - take;
  const bool have = lane < take;
  double r[3][3][3];
  int own = 0;
  {
    double c[8] = {0., 0., 0., 0., 0., 0., 0., 0.};
    if (have) {
#pragma unroll
      for (int q = 0; q < 8; q++) c[q] = B4.node[base + lane][q];
      own = B4.own[base + lane];
    }
    refine3(c, r);
  }
  n4 = base;
  wave_sync();   // the taken slots are free again
  double acc = 0.0;
#pragma unroll 1
  for (int ch = 0; ch < 8; ch++) {
    double c5[8], v;
    pick_uniform(r, ch, c5);
    bool want = false;
    if (have) {
      if (node3(c5, 5, 0x1p-18, v)) acc += v;
      else want = true;
    }
    push_nodes(B5, n5, want, c5, own, lane);
    if (n5 >= 64) drain5(B5, n5, accl, lane);
  }
  if (have) atomicAdd(&accl[own], acc);
  wave_sync();
}

__global__ __launch_bounds__(256) void k_lset_cut3b(LGrid g, const Q2* __restrict__ queue, int64_t nq,
                                                    double* __restrict__ hsum) {
  __shared__ NodeBuf b4[4], b5[4];
  __shared__ double accl[4][64];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (int64_t(blockIdx.x) * blockDim.x + 64 * wave >= nq) return;   // whole wavefronts leave together
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const bool live = t < nq;
  const Q2 e = live ? queue[t] : Q2{0, 0};
  double c2[8], c3[8];
  {
    const int i = e.cell % g.nc[0], j = (e.cell / g.nc[0]) % g.nc[1], k = e.cell / (g.nc[0] * g.nc[1]);
    double C[3][3][3];
    cell_samples<3>(g, i, j, k, C);
    node2_of(C, e.path, c2);
  }
  accl[wave][lane] = 0.0;
  double acc = 0.0;
  int L = 3, ch3 = 0, ch4 = 0;
  bool active = live;
  int n4 = 0, n5 = 0;   // the same in every lane (formed from ballots)
  while (true) {
    bool want = false;
    double n[8];
    if (active) {
      double par[8];
#pragma unroll
      for (int q = 0; q < 8; q++) par[q] = L == 3 ? c2[q] : c3[q];
      child3(par, L == 3 ? ch3 : ch4, n);
      double v;
      bool next = true;
      if (node3(n, L, L == 3 ? 0x1p-12 : 0x1p-15, v)) {
        acc += v;
      } else if (L == 4) {
        want = true;
      } else {   // open the lev-3 node
#pragma unroll
        for (int q = 0; q < 8; q++) c3[q] = n[q];
        ch4 = 0;
        L = 4;
        next = false;
      }
      if (next) {
        if (L == 4 && ++ch4 < 8) {
        } else {
          L = 3;
          active = ++ch3 < 8;
        }
      }
    }
    push_nodes(b4[wave], n4, want, n, lane, lane);
    const bool done = __ballot(active) == 0;
    while (n4 >= 64 || (done && n4 > 0)) drain4(b4[wave], n4, b5[wave], n5, accl[wave], lane);
    if (done) {
      while (n5 > 0) drain5(b5[wave], n5, accl[wave], lane);
      break;
    }
  }
  acc += accl[wave][lane];
  // one atomic per cell present in the wavefront (a cell's entries are contiguous)
  bool pending = live;
  while (true) {
    const unsigned long long m = __ballot(pending);
    if (!m) break;
    const int lead = __ffsll((long long)m) - 1;
    const int32_t c = __shfl(e.cell, lead);
    const bool mine = pending && e.cell == c;
    double v = mine ? acc : 0.0;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if (lane == lead) atomicAdd(&hsum[c], v);
    pending = pending && !mine;
  }
}

// ---- 2D area fraction: sub-square corners c[2 a + b] ----

__device__ __forceinline__ void child2(const double* c, int ch, double* out) {
  double r[3][3];
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++) r[2 * a][2 * b] = c[2 * a + b];
  for (int a = 0; a < 3; a += 2) r[a][1] = 0.5 * (r[a][0] + r[a][2]);
  for (int b = 0; b < 3; b++) r[1][b] = 0.5 * (r[0][b] + r[2][b]);
  const int ca = (ch >> 1) & 1, cb = ch & 1;
  for (int p = 0; p < 2; p++)
    for (int q = 0; q < 2; q++) out[2 * p + q] = r[ca + p][cb + q];
}

__device__ __forceinline__ bool node2(const double* c, int lev, double w, double& v) {
  int f = 0;
  for (int q = 0; q < 4; q++) f += (c[q] > 0.0) - (c[q] < 0.0);
  if (f == 4) { v = w; return true; }
  if (f == -4) { v = 0.0; return true; }
  if (lev > 5) {
    double s = 0.0;
    for (int q = 0; q < 4; q++) s += c[q];
    v = s > 0.0 ? 0.75 * w : (s == 0.0 ? 0.5 * w : 0.25 * w);
    return true;
  }
  return false;
}

__device__ double subtree2(const double* c0, int lev0, double w0) {
  double stk[7][4];
  int nxt[7];
  double acc = 0.0, v;
  for (int q = 0; q < 4; q++) stk[lev0][q] = c0[q];
  int L = lev0;
  double w = w0;
  while (true) {
    if (node2(stk[L], L, w, v)) {
      acc += v;
      while (true) {
        if (L == lev0) return acc;
        L--;
        w *= 4.0;
        if (++nxt[L] < 4) break;
      }
      child2(stk[L], nxt[L], stk[L + 1]);
      L++;
      w *= 0.25;
    } else {
      nxt[L] = 0;
      child2(stk[L], 0, stk[L + 1]);
      L++;
      w *= 0.25;
    }
  }
}


// 2D: 16 lanes per listed cut cell, 4 cells per wavefront
__global__ __launch_bounds__(64 * kCutWaves) void k_lset_cut2(LGrid g, const int32_t* __restrict__ list, int64_t nlist,
                                                              double* __restrict__ cvol, double* __restrict__ carea) {
  const int64_t w = (int64_t(blockIdx.x) * kCutWaves + (threadIdx.x >> 6)) * 4 + ((threadIdx.x >> 4) & 3);
  const int lane = threadIdx.x & 15;
  const bool live = w < nlist;
  const int32_t cell = live ? list[w] : 0;
  const int i = cell % g.nc[0], j = cell / g.nc[0];
  double C[3][3][3];
  cell_samples<2>(g, i, j, 0, C);
  const double dx = g.dcell[0], dy = g.dcell[1];
  int f = 0;
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) f += (C[a][b][0] > 0.0) - (C[a][b][0] < 0.0);
  const bool split = f != 9 && f != -9;
  // every lane of the group takes part in the shuffles below
  const int sq = lane >> 2, ch = lane & 3;
  const int a = sq >> 1, b = sq & 1;
  double part = 0.0, v;
  if (split) {
    double s0[4] = {C[a][b][0], C[a][b + 1][0], C[a + 1][b][0], C[a + 1][b + 1][0]};
    if (node2(s0, 0, 0.25, v)) {
      part = ch == 0 ? v : 0.0;
    } else {
      double c1[4];
      child2(s0, ch, c1);
      part = subtree2(c1, 1, 0.25 * 0.25);
    }
  }
  for (int off = 8; off > 0; off >>= 1) part += __shfl_xor(part, off, 16);
  // the area as the reference measures a 2D cell: three equal layers of unit depth
  double E[3][3][3];
  for (int r = 0; r < 3; r++)
    for (int q = 0; q < 3; q++)
      for (int t = 0; t < 3; t++) E[r][q][t] = C[r][q][0];
  const double am = mc_area_lanes(E, dx, dy, 1.0, lane, 16);   // every lane of the group
  const double area = split ? am : 0.0;
  double H = f == 9 ? 1.0 : 0.0;
  if (split) H = part < 1.0 ? part : 1.0;
  if (live && lane == 0) {
    cvol[cell] = H;
    carea[cell] = split ? area : 0.0;
  }
}

__device__ __forceinline__ double prof_dx(const LProf& P, double max_phi, double phi) {
  if (P.profile == SMG_PROFILE_LINEAR) {
    if (phi > 0. || fabs(phi) < P.dl) {
      double t = phi > 0. ? phi / P.band : 0.;
      t = t < 1.0 ? t : 1.0;
      return P.dx0 + t * (P.dx1 - P.dx0);
    }
    return P.dx0;
  }
  if (phi > 0. || fabs(phi) < P.dl) return P.min_dl + tanh(2.5 * fabs(phi) / max_phi) / P.tanh25 * (P.max_dl - P.min_dl);
  return P.min_dl;
}

// one wavefront per 64 / 4^D packages; lane = package * 4^D + the cell's place in the
// host builder's loop (ci, then cj, then ck fastest)
__global__ __launch_bounds__(256) void k_lset_pkg(LGrid g, LProf P, const int8_t* __restrict__ tag,
                                                  const double* __restrict__ cvol, const double* __restrict__ carea,
                                                  const LCount* __restrict__ cnt, int64_t npkg, int npk1, int npk2,
                                                  double* __restrict__ psum) {
  __shared__ double terms[4][64][4];
  const int D = g.D;
  const int cpp = D == 3 ? 64 : 16;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int per_wave = 64 / cpp;
  const int64_t pk_lin = (int64_t(blockIdx.x) * 4 + wave) * per_wave + lane / cpp;
  const int cl = lane % cpp;
  double tv = 0.0, ta = 0.0, tm = 0.0, tms = 0.0;
  if (pk_lin < npkg) {
    // package (pi, pj, pk) with pk fastest
    const int pk = D > 2 ? int(pk_lin % npk2) : 0;
    const int pj = int((pk_lin / npk2) % npk1);
    const int pi = int(pk_lin / (int64_t(npk1) * npk2));
    const int ck = D > 2 ? cl % 4 : 0;
    const int cj = D > 2 ? (cl / 4) % 4 : cl % 4;
    const int ci = D > 2 ? cl / 16 : cl / 4;
    const int i = 4 * pi + ci, j = 4 * pj + cj, k = D > 2 ? 4 * pk + ck : 0;
    const double p = at(g, i, j, k);
    if (p > 0. || fabs(p) < g.dl) {
      const double max_phi = P.max_phi_cfg > 0.0 ? P.max_phi_cfg : __longlong_as_double((long long)cnt->max_phi_bits);
      const double sc = prof_dx(P, max_phi, p);
      const int64_t hidx = (int64_t(D > 2 ? k + 1 : 0) * g.halo[1] + (D > 1 ? j + 1 : 0)) * g.halo[0] + (i + 1);
      const int64_t cell = (int64_t(k) * (D > 1 ? g.nc[1] : 1) + j) * g.nc[0] + i;
      double vol, area = 0.0;
      if (tag[hidx] == 1) {
        // the volume fraction, clamped to 1 as GeometryCalculator does
        const double H = cvol[cell] < 1.0 ? cvol[cell] : 1.0;
        vol = H * g.dcell[0] * g.dcell[1] * (D > 2 ? g.dcell[2] : 1.0);
        area = carea[cell];
        double sp = 1.0;
        for (int q = 0; q < D - 1; q++) sp *= sc;
        ta = area;
        tms = area / sp;
      } else {
        vol = g.dcell[0] * (D > 1 ? g.dcell[1] : 1.0) * (D > 2 ? g.dcell[2] : 1.0);
      }
      double sd = 1.0;
      for (int q = 0; q < D; q++) sd *= sc;
      tm = vol / sd;
      tv = vol;
    }
  }
  double* T = terms[wave][lane];
  T[0] = tv;
  T[1] = ta;
  T[2] = tm;
  T[3] = tms;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  // lane 4 q + m sums term m of package q of this wavefront, in cell order
  if (lane < 4 * per_wave) {
    const int q = lane >> 2, m = lane & 3;
    const int64_t pq = (int64_t(blockIdx.x) * 4 + wave) * per_wave + q;
    if (pq < npkg) {
      double s = 0.0;
      for (int c = 0; c < cpp; c++) s += terms[wave][q * cpp + c][m];
      psum[4 * pq + m] = s;
    }
  }
}

#define HIP_TRY_LS(expr)                                               \
  do {                                                                 \
    hipError_t e_ = (expr);                                            \
    if (e_ != hipSuccess) {                                            \
      set_error(std::string(#expr) + ": " + hipGetErrorString(e_));    \
      rc = SMG_ERR_HIP;                                                \
      goto done;                                                       \
    }                                                                  \
  } while (0)

}  // namespace

int lset_build_device(const smg_config* cfg, const double* phi_ext, int zero_margin, double* phi, float* normal,
                      int8_t* tag, smg_lset_stats* st, double* kernel_ms) {
  LGrid g{};
  g.D = cfg->dim;
  if (g.D < 2 || g.D > 3) { set_error("smg_lset_build_gpu: dim must be 2 or 3"); return SMG_ERR_ARG; }
  double dl = 1.e20;
  int64_t n_ext = 1, n_halo = 1, n_int = 1, npkg = 1;
  for (int a = 0; a < 3; a++) {
    const bool on = a < g.D;
    if (on && cfg->num_pkg[a] < 1) { set_error("smg_lset_build_gpu: num_pkg < 1"); return SMG_ERR_ARG; }
    g.nc[a] = on ? 4 * cfg->num_pkg[a] : 1;
    g.ext_n[a] = on ? g.nc[a] + 8 : 1;
    g.halo[a] = on ? g.nc[a] + 2 : 1;
    const double dpkg = on ? cfg->domain[a] / double(cfg->num_pkg[a]) : 1.0;
    g.dcell[a] = on ? dpkg / 4.0 : 1.0;
    if (on) dl = std::min(dl, g.dcell[a]);
    n_ext *= g.ext_n[a];
    n_halo *= g.halo[a];
    n_int *= g.nc[a];
    npkg *= on ? cfg->num_pkg[a] : 1;
  }
  if (n_ext > (int64_t(1) << 31) - 1 || n_halo > (int64_t(1) << 31) - 1 || n_int > (int64_t(1) << 31) - 1) {
    set_error("smg_lset_build_gpu: grid larger than 2^31 cells");
    return SMG_ERR_ARG;
  }
  g.dl = dl;
  LProf P{};
  {
    const SpacingProfile hp(*cfg, dl, 1.0);
    P.profile = hp.profile;
    P.min_dl = hp.min_dl;
    P.max_dl = hp.max_dl;
    P.tanh25 = hp.tanh25;
    P.dx0 = hp.dx0;
    P.dx1 = hp.dx1;
    P.band = hp.band;
    P.dl = dl;
    P.max_phi_cfg = cfg->max_phi > 0.0 ? cfg->max_phi : 0.0;
  }
  int rc = SMG_OK;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  double *d_ext = nullptr, *d_phi = nullptr, *d_cvol = nullptr, *d_carea = nullptr, *d_psum = nullptr;
  float* d_nrm = nullptr;
  int8_t* d_tag = nullptr;
  int32_t* d_list = nullptr;
  double* d_bmax = nullptr;
  const dim3 grid_blocks(unsigned((g.halo[0] + 63) / 64), unsigned((g.halo[1] + 3) / 4), unsigned(g.halo[2]));
  const int64_t nblk = int64_t(grid_blocks.x) * grid_blocks.y * grid_blocks.z;
  Q2* d_queue = nullptr;
  LCount* d_cnt = nullptr;
  LCount hc{};
  std::vector<double> psum;
  int64_t nlist = 0, qcap = 0, nq = 0;
  const int npk1 = g.D > 1 ? cfg->num_pkg[1] : 1, npk2 = g.D > 2 ? cfg->num_pkg[2] : 1;
  const int cpp = g.D == 3 ? 64 : 16;
  const int64_t pkg_per_block = 4 * (64 / cpp);
  float ms = 0.f;
  HIP_TRY_LS(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  HIP_TRY_LS(hipEventCreate(&e0));
  HIP_TRY_LS(hipEventCreate(&e1));
  HIP_TRY_LS(hipMalloc(&d_ext, sizeof(double) * n_ext));
  HIP_TRY_LS(hipMalloc(&d_phi, sizeof(double) * n_halo));
  HIP_TRY_LS(hipMalloc(&d_nrm, sizeof(float) * 3 * n_halo));
  HIP_TRY_LS(hipMalloc(&d_tag, n_halo));
  HIP_TRY_LS(hipMalloc(&d_list, sizeof(int32_t) * n_int));
  HIP_TRY_LS(hipMalloc(&d_cvol, sizeof(double) * n_int));
  HIP_TRY_LS(hipMalloc(&d_carea, sizeof(double) * n_int));
  HIP_TRY_LS(hipMalloc(&d_psum, sizeof(double) * 4 * npkg));
  HIP_TRY_LS(hipMalloc(&d_cnt, sizeof(LCount)));
  HIP_TRY_LS(hipMalloc(&d_bmax, sizeof(double) * nblk));
  HIP_TRY_LS(hipMemcpyAsync(d_ext, phi_ext, sizeof(double) * n_ext, hipMemcpyHostToDevice, s));
  HIP_TRY_LS(hipMemsetAsync(d_cnt, 0, sizeof(LCount), s));
  g.ext = d_ext;
  HIP_TRY_LS(hipEventRecord(e0, s));
  k_lset_grid<<<grid_blocks, dim3(256), 0, s>>>(g, zero_margin, d_phi, d_nrm, d_tag, d_list, d_cnt, d_bmax);
  HIP_TRY_LS(hipGetLastError());
  k_lset_max<<<dim3(1), dim3(1024), 0, s>>>(d_bmax, nblk, d_cnt);
  HIP_TRY_LS(hipGetLastError());
  HIP_TRY_LS(hipMemcpyAsync(&hc, d_cnt, sizeof(LCount), hipMemcpyDeviceToHost, s));
  HIP_TRY_LS(hipStreamSynchronize(s));
  nlist = int64_t(hc.n_list);
  if (nlist > n_int) { set_error("smg_lset_build_gpu: cut-cell list overflow"); rc = SMG_ERR_HIP; goto done; }
  if (nlist > 0) {
    if (g.D == 3) {
      // about 50 split lev-2 nodes per cut cell on smooth surfaces; k_lset_cut3a drops
      // the entries past qcap but still counts them, and the host then regrows the
      // queue to the count and runs the pass again.
      // The environment variable SMG_LSET_QUEUE_PER_CELL (a positive integer) overrides
      // the 96 entries per cut cell, so that a test can force the regrow; any other
      // value is an error.
      int64_t per_cell = 96;
      if (const char* env = std::getenv("SMG_LSET_QUEUE_PER_CELL")) {
        char* end = nullptr;
        errno = 0;
        const long v = std::strtol(env, &end, 10);
        if (end == env || *end != '\0' || errno != 0 || v < 1 || v > (1L << 20)) {
          set_error("smg_lset_build_gpu: SMG_LSET_QUEUE_PER_CELL must be an integer in [1, 2^20]");
          rc = SMG_ERR_ARG;
          goto done;
        }
        per_cell = v;
      }
      qcap = per_cell == 96 ? std::max<int64_t>(96 * nlist, int64_t(1) << 16) : per_cell * nlist;
      for (int pass = 0; pass < 2; pass++) {   // the second pass only if the queue was short
        HIP_TRY_LS(hipMalloc(&d_queue, sizeof(Q2) * qcap));
        HIP_TRY_LS(hipMemsetAsync(&d_cnt->n_queue, 0, sizeof(unsigned long long), s));
        k_lset_cut3a<<<dim3(unsigned((nlist + kCutWaves - 1) / kCutWaves)), dim3(64 * kCutWaves), 0, s>>>(
            g, d_list, nlist, d_cvol, d_carea, d_queue, qcap, &d_cnt->n_queue);
        HIP_TRY_LS(hipGetLastError());
        HIP_TRY_LS(hipMemcpyAsync(&hc, d_cnt, sizeof(LCount), hipMemcpyDeviceToHost, s));
        HIP_TRY_LS(hipStreamSynchronize(s));
        if (int64_t(hc.n_queue) <= qcap) break;
        HIP_TRY_LS(hipFree(d_queue));
        d_queue = nullptr;
        qcap = int64_t(hc.n_queue);
      }
      nq = int64_t(hc.n_queue);
      if (nq > qcap) { set_error("smg_lset_build_gpu: queue sizing"); rc = SMG_ERR_HIP; goto done; }
      if (nq > 0)
        k_lset_cut3b<<<dim3(unsigned((nq + 255) / 256)), dim3(256), 0, s>>>(g, d_queue, nq, d_cvol);
    } else
      k_lset_cut2<<<dim3(unsigned((nlist + 4 * kCutWaves - 1) / (4 * kCutWaves))), dim3(64 * kCutWaves), 0, s>>>(
          g, d_list, nlist, d_cvol, d_carea);
    HIP_TRY_LS(hipGetLastError());
  }
  k_lset_pkg<<<dim3(unsigned((npkg + pkg_per_block - 1) / pkg_per_block)), dim3(256), 0, s>>>(
      g, P, d_tag, d_cvol, d_carea, d_cnt, npkg, npk1, npk2, d_psum);
  HIP_TRY_LS(hipGetLastError());
  HIP_TRY_LS(hipEventRecord(e1, s));
  psum.resize(4 * size_t(npkg));
  HIP_TRY_LS(hipMemcpyAsync(psum.data(), d_psum, sizeof(double) * 4 * npkg, hipMemcpyDeviceToHost, s));
  HIP_TRY_LS(hipMemcpyAsync(phi, d_phi, sizeof(double) * n_halo, hipMemcpyDeviceToHost, s));
  HIP_TRY_LS(hipMemcpyAsync(normal, d_nrm, sizeof(float) * 3 * n_halo, hipMemcpyDeviceToHost, s));
  HIP_TRY_LS(hipMemcpyAsync(tag, d_tag, n_halo, hipMemcpyDeviceToHost, s));
  HIP_TRY_LS(hipStreamSynchronize(s));
  HIP_TRY_LS(hipEventElapsedTime(&ms, e0, e1));
  if (kernel_ms) *kernel_ms = ms;
  if (st) {
    double t_vol = 0.0, t_area = 0.0, t_mass = 0.0, t_msurf = 0.0;
    for (size_t q = 0; q < psum.size(); q += 4) {
      t_vol += psum[q];
      t_area += psum[q + 1];
      t_mass += psum[q + 2];
      t_msurf += psum[q + 3];
    }
    double max_phi;
    std::memcpy(&max_phi, &hc.max_phi_bits, sizeof(double));
    const int64_t n_cut = int64_t(hc.n_cut);
    st->maximum_phi = max_phi;
    st->total_volume = t_vol;
    st->total_area = t_area;
    st->total_mass = t_mass;
    st->total_mass_surface = t_msurf;
    st->n_interface_cell = n_cut;
    int64_t n_surf = std::max<int64_t>(int64_t(t_msurf), 0);
    const int64_t n_mesh = std::max<int64_t>(int64_t(t_mass - 0.5 * t_msurf), 0);
    if (n_cut < n_surf) n_surf = n_cut;
    st->num_particle_mesh = n_mesh;
    st->num_particle_surface = n_surf;
    st->num_particle = n_mesh + n_surf;
  }
done:
  if (s) (void)hipStreamSynchronize(s);
  for (void* ptr : {(void*)d_ext, (void*)d_phi, (void*)d_nrm, (void*)d_tag, (void*)d_list, (void*)d_cvol,
                    (void*)d_carea, (void*)d_psum, (void*)d_cnt, (void*)d_queue, (void*)d_bmax})
    if (ptr) (void)hipFree(ptr);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (s) (void)hipStreamDestroy(s);
  return rc;
}

}  // namespace smg
