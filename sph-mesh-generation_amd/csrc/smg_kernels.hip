// smg_kernels.hip -- gfx950 kernels of one relaxation step.
//
// Pass structure (one step = SPH_run_mesh_generation, sph_mesh_generation.cpp:1371-1463):
//   k_reset    per-job reduction slots of this step
//   k_map      P1: x += v dt (Update_coord_full_step), level-set sample
//              (Calculate_particle_infor), surface projection / re-seeding
//              (Find_and_map_particle), target spacing (Set_particle_scale), cell key
//   k_scan_*   exclusive scan of the per-cell counts (cell starts)
//   k_scatter  counting-sort scatter by cell key
//   k_segsort  order inside each cell by particle id (makes the sort deterministic)
//   k_gather   move the particle fields into the new order
//   k_nb       neighbour lists with mirror ghosts, FP64 predicate
//              (Level_info::Refresh_neighbor_info, level_infor.cpp:469-506)
//   k_force1   P3: kernel summation + repulsion + stage-1 dt
//              (Get_kernel_summation, Accumulate_particle_force(1), Set_timestep(1))
//   k_force2   P4: viscous force + stage-2 dt + nu
//              (Accumulate_particle_force(2), Set_timestep(2))
// The two velocity half steps fold into the next k_map: v = a1 dt1/2 + a2 dt2/2.
//
// FP64 is used where the result must equal the reference's bit for bit (level-set
// sampling, projection, neighbour predicate) and for the dt / residual reductions;
// the pair sums run in FP32.  The file is compiled with -ffp-contract=off so that no
// FMA changes an FP64 rounding.
#include <mutex>
#include <hip/hip_runtime.h>
#include <cstdint>
#include <type_traits>

#include "smg_device.hpp"
#include "smg_kernels.hpp"

namespace smg {

// ---------------------------------------------------------------------------
// helpers

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ uint64_t rng_key(uint64_t seed, uint64_t stream, uint64_t id) {
  const uint64_t k = mix64(seed + 0x632BE59BD9B4E019ull * (stream + 1ull));
  return mix64(k ^ (id * 0xD1B54A32D192ED03ull + 0x8CB92BA72F3D8DD7ull));
}
__device__ __forceinline__ double rng_u(uint64_t key, uint64_t counter) {
  return double(mix64(key + counter * 0x9E3779B97F4A7C15ull) >> 11) * 0x1.0p-53;
}

__device__ __forceinline__ double amin(double a, double b) { return (a <= b) ? a : b; }  // AMIN1
__device__ __forceinline__ double amax(double a, double b) { return (a >= b) ? a : b; }  // AMAX1

__device__ __forceinline__ unsigned long long dbits(double v) { return (unsigned long long)__double_as_longlong(v); }
__device__ __forceinline__ double bitsd(unsigned long long v) { return __longlong_as_double((long long)v); }

template <typename T, typename Op>
__device__ __forceinline__ T wave_reduce(T v, Op op) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = op(v, __shfl_xor(v, o, 64));
  return v;
}

// reduce over a 256-thread workgroup; the result is valid in thread 0
template <typename T, typename Op>
__device__ __forceinline__ T block_reduce(T v, Op op, T* lds4) {
  v = wave_reduce(v, op);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) lds4[w] = v;
  __syncthreads();
  T r = lds4[0];
  if (threadIdx.x == 0) {
    for (int q = 1; q < (int)(blockDim.x >> 6); q++) r = op(r, lds4[q]);
  }
  __syncthreads();
  return r;
}

struct OpMin { __device__ double operator()(double a, double b) const { return a < b ? a : b; } };
struct OpMax { __device__ double operator()(double a, double b) const { return a > b ? a : b; } };
struct OpAdd { __device__ int operator()(int a, int b) const { return a + b; } };
struct OpMaxI { __device__ int operator()(int a, int b) const { return a > b ? a : b; } };
struct OpAddU {
  __device__ unsigned long long operator()(unsigned long long a, unsigned long long b) const { return a + b; }
};

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own L2.  The
// tile of workgroup b is chosen so that every XCD works through one contiguous run of
// tiles -- cell-sorted slots, hence a compact region of space -- and the neighbour
// records a tile reads are mostly in the L2 of the XCD that reads them.
constexpr int kXcds = 8;
__device__ __forceinline__ int64_t tile_id(const Work& W) {
  const int64_t b = blockIdx.x, G = gridDim.x;
  const int64_t x = b % kXcds, i = b / kXcds;
  const int64_t q = G / kXcds, r = G % kXcds;
  return W.tile0 + x * q + (x < r ? x : r) + i;
}

// Whether any job rebuilds its lists at this step (W.ctr[1]: set by k_decide, by the
// host before a single job's sort, cleared by k_reset).  The rebuild-only passes test
// it first, one scalar load for every workgroup, and return without touching their
// tiles on the steps where no job rebuilds.
__device__ __forceinline__ bool rebuild_any(const Work& W) { return W.ctr[1] != 0ull; }

// neighbour entry k of a slot (spill offset off is read only when the list spills)
__device__ __forceinline__ int64_t nbl_index(const Work& W, int64_t slot, int k) {
  return (int64_t(k >> 2) * W.cap_total + slot) * 4 + (k & 3);
}
__device__ __forceinline__ uint32_t nb_entry(const Work& W, int64_t slot, int k, int32_t off) {
  return k < W.K ? uint32_t(W.nbl[nbl_index(W, slot, k)]) : uint32_t(W.spill[int64_t(off) + (k - W.K)]);
}
// entries 4g .. 4g + 3 of a slot, g < K / 4, in one 16-byte load
__device__ __forceinline__ uint4 nb_group(const Work& W, int64_t slot, int g) {
  return reinterpret_cast<const uint4*>(W.nbl)[int64_t(g) * W.cap_total + slot];
}

// |v| dt2 / dx of one particle, its fixed-point form for the order-free sum, and |v| dt2
struct Disp {
  double r, md;
  unsigned long long rq;
};
__device__ __forceinline__ Disp disp_of(double U, double dt2, double h, const JobParams& J) {
  Disp d;
  d.md = U * dt2;
  d.r = d.md / (h * J.cut_off / J.cell_ratio);
  const double q = d.r * kResScale + 0.5;
  d.rq = (q < 9.0e15) ? (unsigned long long)q : 9000000000000000ull;   // clamp: 2^53 > 9e15
  return d;
}

__device__ __forceinline__ int64_t ls_index(const JobParams& J, int D, int gx, int gy, int gz) {
  const int a = gx + 1, b = (D > 1) ? gy + 1 : 0, c = (D > 2) ? gz + 1 : 0;
  return J.lset_off + ls_cell_offset(D, J.ng, a, b, c);
}

// clamped shift into the level-set box (particle_mesh_generation.cpp:182-185) and
// package index (get_cell_id with glbl_pkg_start = -1, glbl_pkg_end = num_pkg + 1)
template <int D>
__device__ __forceinline__ void ls_locate(const JobParams& J, const double* x, double* cs, int* pk) {
#pragma unroll
  for (int a = 0; a < 3; a++) {
    if (a < D) {
      double t = x[a] - J.box_l[a];
      t = (t >= 0.) ? t : 0.;
      const double hi = J.domain[a] - 1.e-8;
      cs[a] = (t <= hi) ? t : hi;
      int p = int(floor(cs[a] / J.dpkg[a]));
      const int phi_ = J.num_pkg[a];
      p = (p <= phi_) ? p : phi_;
      p = (-1 >= p) ? -1 : p;
      pk[a] = p;
    } else {
      cs[a] = 0.;
      pk[a] = 0;
    }
  }
}

// Calculate_particle_infor (particle_mesh_generation.cpp:175-265): bi/trilinear phi
// and raw normal; the normal is left unnormalised (quirk Q3)
template <int D, bool kNormal>
__device__ void ls_sample_base(const JobParams& J, const Work& W, const double* x, double& phi, double* n) {
  double cs[3];
  int pk[3];
  ls_locate<D>(J, x, cs, pk);
  int m[3] = {0, 0, 0};
  double f[3] = {0., 0., 0.};
#pragma unroll
  for (int a = 0; a < D; a++) {
    double rr = (cs[a] - J.dpkg[a] * pk[a]) / J.dcell[a] - 0.5;
    rr = (rr <= 3.5) ? rr : 3.5;
    rr = (-0.5 >= rr) ? -0.5 : rr;
    m[a] = int(floor(rr));
    f[a] = rr - double(m[a]);
  }
  phi = 0.;
  if (kNormal) n[0] = n[1] = n[2] = 0.;
#pragma unroll
  for (int r = 0; r < 2; r++)
#pragma unroll
    for (int s = 0; s < (D > 1 ? 2 : 1); s++)
#pragma unroll
      for (int t = 0; t < (D > 2 ? 2 : 1); t++) {
        const int64_t c = ls_index(J, D, pk[0] * 4 + m[0] + r, pk[1] * 4 + m[1] + s, pk[2] * 4 + m[2] + t);
        double coef = fabs(f[0] - (1.0 - r));
        if (D > 1) coef = coef * fabs(f[1] - (1.0 - s));
        if (D > 2) coef = coef * fabs(f[2] - (1.0 - t));
        phi += W.lsp[c] * coef;
        if (kNormal) {
          const float4 q = W.lsn[c];
          n[0] += double(q.x) * coef;
          n[1] += double(q.y) * coef;
          n[2] += double(q.z) * coef;
        }
      }
}

// Refined level l (0-based: refinement l + 1) of a multi-resolution level set.  The
// cell with global indices g (-1 .. 4 lnp per axis) in the level's bricks, or -1 where
// the level does not store its package.
template <int D>
__device__ __forceinline__ int64_t lvl_cell(const JobParams& J, const Work& W, int l, const int* g) {
  int p[3] = {0, 0, 0}, c[3] = {0, 0, 0};
#pragma unroll
  for (int a = 0; a < D; a++) {
    p[a] = (g[a] + 4) / 4 - 1;   // floor(g / 4) for g >= -4
    c[a] = g[a] - 4 * p[a];
  }
  int64_t e = p[0] + 1;
  if (D > 1) e += int64_t(p[1] + 1) * (J.lnp[l][0] + 2);
  if (D > 2) e += int64_t(p[2] + 1) * (J.lnp[l][0] + 2) * (J.lnp[l][1] + 2);
  const int32_t b = W.lpm[J.lmap_base + J.lmap_off[l] + e];
  if (b < 0) return -1;
  return J.lset_off + J.lcell_off[l] + int64_t(b) * (D > 2 ? 64 : 16) + ((c[2] * 4 + c[1]) * 4 + c[0]);
}

// Calculate_particle_infor's interpolation on refined level l: the same arithmetic as
// the base grid with the level's package and cell sizes.  False (nothing written)
// when a corner's package is not stored.
template <int D, bool kNormal>
__device__ bool ls_sample_level(const JobParams& J, const Work& W, int l, const double* x, double& phi, double* n) {
  int pk[3] = {0, 0, 0}, m[3] = {0, 0, 0};
  double f[3] = {0., 0., 0.};
#pragma unroll
  for (int a = 0; a < D; a++) {
    double t = x[a] - J.box_l[a];
    t = (t >= 0.) ? t : 0.;
    const double hi = J.domain[a] - 1.e-8;
    const double cs = (t <= hi) ? t : hi;
    int p = int(floor(cs / J.ldpkg[l][a]));
    p = (p <= J.lnp[l][a]) ? p : J.lnp[l][a];
    p = (-1 >= p) ? -1 : p;
    pk[a] = p;
    double rr = (cs - J.ldpkg[l][a] * p) / J.ldcell[l][a] - 0.5;
    rr = (rr <= 3.5) ? rr : 3.5;
    rr = (-0.5 >= rr) ? -0.5 : rr;
    m[a] = int(floor(rr));
    f[a] = rr - double(m[a]);
  }
  int64_t cell[8];
  bool ok = true;
#pragma unroll
  for (int r = 0; r < 2; r++)
#pragma unroll
    for (int s = 0; s < (D > 1 ? 2 : 1); s++)
#pragma unroll
      for (int t = 0; t < (D > 2 ? 2 : 1); t++) {
        const int g[3] = {pk[0] * 4 + m[0] + r, pk[1] * 4 + m[1] + s, pk[2] * 4 + m[2] + t};
        const int64_t c = lvl_cell<D>(J, W, l, g);
        cell[(r * 2 + s) * 2 + t] = c;
        ok = ok && c >= 0;
      }
  if (!ok) return false;
  phi = 0.;
  if (kNormal) n[0] = n[1] = n[2] = 0.;
#pragma unroll
  for (int r = 0; r < 2; r++)
#pragma unroll
    for (int s = 0; s < (D > 1 ? 2 : 1); s++)
#pragma unroll
      for (int t = 0; t < (D > 2 ? 2 : 1); t++) {
        const int64_t c = cell[(r * 2 + s) * 2 + t];
        double coef = fabs(f[0] - (1.0 - r));
        if (D > 1) coef = coef * fabs(f[1] - (1.0 - s));
        if (D > 2) coef = coef * fabs(f[2] - (1.0 - t));
        phi += W.lsp[c] * coef;
        if (kNormal) {
          const float4 q = W.lsn[c];
          n[0] += double(q.x) * coef;
          n[1] += double(q.y) * coef;
          n[2] += double(q.z) * coef;
        }
      }
  return true;
}

// the finest level that stores the sample's corners, else the base grid
template <int D, bool kNormal>
__device__ void ls_sample(const JobParams& J, const Work& W, const double* x, double& phi, double* n) {
  for (int l = J.ls_nl - 1; l >= 0; l--)
    if (ls_sample_level<D, kNormal>(J, W, l, x, phi, n)) return;
  ls_sample_base<D, kNormal>(J, W, x, phi, n);
}

template <int D>
__device__ int ls_tag_base(const JobParams& J, const Work& W, const double* x) {
  double cs[3];
  int pk[3];
  ls_locate<D>(J, x, cs, pk);
  int g[3] = {0, 0, 0};
#pragma unroll
  for (int a = 0; a < D; a++) {
    int c = int(floor((cs[a] - J.dpkg[a] * pk[a]) / J.dcell[a]));
    c = (c <= 3) ? c : 3;
    c = (0 >= c) ? 0 : c;
    g[a] = pk[a] * 4 + c;
  }
  return W.lst[ls_index(J, D, g[0], g[1], g[2])];
}

// Levelset_levelinfo::Get_id_pkg_cell (lset_level_infor.cpp:350-366) -> tag_interface,
// from the finest level that stores the cell
template <int D>
__device__ int ls_tag(const JobParams& J, const Work& W, const double* x) {
  for (int l = J.ls_nl - 1; l >= 0; l--) {
    int g[3] = {0, 0, 0};
#pragma unroll
    for (int a = 0; a < D; a++) {
      double t = x[a] - J.box_l[a];
      t = (t >= 0.) ? t : 0.;
      const double hi = J.domain[a] - 1.e-8;
      const double cs = (t <= hi) ? t : hi;
      int p = int(floor(cs / J.ldpkg[l][a]));
      p = (p <= J.lnp[l][a]) ? p : J.lnp[l][a];
      p = (-1 >= p) ? -1 : p;
      int c = int(floor((cs - J.ldpkg[l][a] * p) / J.ldcell[l][a]));
      c = (c <= 3) ? c : 3;
      c = (0 >= c) ? 0 : c;
      g[a] = p * 4 + c;
    }
    const int64_t c = lvl_cell<D>(J, W, l, g);
    if (c >= 0) return W.lst[c];
  }
  return ls_tag_base<D>(J, W, x);
}

// Levelset::Get_scale (level_set.cpp:311-351) and the linear-profile extension
__device__ __forceinline__ double get_scale(const JobParams& J, double phi) {
  if (J.profile == SMG_PROFILE_LINEAR) {
    if (phi > 0. || fabs(phi) < J.dl) {
      double t = phi > 0. ? phi / J.lin_band : 0.;
      t = t < 1.0 ? t : 1.0;
      return J.lin_dx0 + t * (J.lin_dx1 - J.lin_dx0);
    }
    return J.lin_dx0;
  }
  if (phi > 0. || fabs(phi) < J.dl) return J.min_dl + tanh(2.5 * fabs(phi) / J.max_phi) / J.tanh25 * (J.max_dl - J.min_dl);
  return J.min_dl;
}
// Set_particle_info (particle_mesh_generation.cpp:361-381): h only; rho and vol
// follow from dx = h * cut_off / cell_ratio where they are used
__device__ __forceinline__ double h_of_phi(const JobParams& J, double phi) {
  return get_scale(J, phi) * J.cell_ratio / J.cut_off;
}

// cell-list level of a particle: the finest level whose cells hold its list radius
__device__ __forceinline__ int level_of(const JobParams& J, double h) {
  const double r = h * J.cut_v + 2.e-10;
  int l = 0;
  while (l + 1 < J.nlev && J.clen[l + 1] >= r) l++;
  return l;
}

__device__ __forceinline__ uint32_t spread2(uint32_t v) {
  v &= 0x0000FFFFu;
  v = (v | (v << 8)) & 0x00FF00FFu;
  v = (v | (v << 4)) & 0x0F0F0F0Fu;
  v = (v | (v << 2)) & 0x33333333u;
  v = (v | (v << 1)) & 0x55555555u;
  return v;
}
__device__ __forceinline__ uint32_t spread3(uint32_t v) {
  v &= 0x000003FFu;
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}
template <int D>
__device__ __forceinline__ uint32_t morton(const int* c) {
  if (D == 2) return spread2(uint32_t(c[0])) | (spread2(uint32_t(c[1])) << 1);
  return spread3(uint32_t(c[0])) | (spread3(uint32_t(c[1])) << 1) | (spread3(uint32_t(c[2])) << 2);
}

template <int D>
__device__ __forceinline__ void cell_at(const JobParams& J, int L, const double* x, int* c) {
#pragma unroll
  for (int a = 0; a < D; a++) {
    // clamp in FP64 before the conversion so that a non-finite position still lands
    // in a valid cell of the job's table
    double t = floor((x[a] - J.cbox_l[a]) * J.cinv[L]);
    const double hi = double(J.lev_n[L][a] - 1);
    t = (t >= 0.) ? t : 0.;
    t = (t <= hi) ? t : hi;
    c[a] = int(t);
  }
}

__device__ __forceinline__ void split_store(double x, float* hi, float* lo, int64_t s) {
  const float h = float(x);
  hi[s] = h;
  lo[s] = float(x - double(h));
}

// ---------------------------------------------------------------------------
// hyperbolic kernel in FP32 (particle.cpp:31-401, 405-807 with PI = 3.1415926)
template <int DK>
__device__ __forceinline__ float kC() {
  return DK == 1 ? 1.0f / 7.0f : (DK == 2 ? float(1.0 / (3.0 * 3.1415926)) : float(15.0 / (62.0 * 3.1415926)));
}
template <int DK>
__device__ __forceinline__ float ipow(float h) {
  return DK == 1 ? h : (DK == 2 ? h * h : (DK == 3 ? h * h * h : h * h * h * h));
}
template <int DK>
__device__ __forceinline__ float Wk(float r, float h) {
  const float C = kC<DK>() / ipow<DK>(h);
  const float s = r / h;
  if (s < 1.f) return C * (6.0f - 6.0f * s + s * s * s);
  if (s < 2.f) {
    const float t = 2.f - s;
    return C * t * t * t;
  }
  return 0.f;
}
template <int DK>
__device__ __forceinline__ float dWk(float r, float h) {   // (dW/dr) / (r + 1e-20)
  const float C = kC<DK>() / ipow<DK + 1>(h);
  const float s = r / h;
  float v = 0.f;
  if (s < 1.f) v = C * (-6.0f + 3.f * s * s);
  else if (s < 2.f) {
    const float t = 2.f - s;
    v = C * (-3.0f) * t * t;
  }
  return v / (r + 1.e-20f);
}

// Pair-loop forms with the hardware reciprocal and square root (v_rcp_f32,
// v_sqrt_f32, 1 ulp) in place of the IEEE division sequences: W from s = r / h and
// C = kC / h^DK of the particle's own h, both computed once per particle; dW from the
// pair's mean h with one reciprocal.
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float frsq(float x) { return __builtin_amdgcn_rsqf(x); }
// Both branches are evaluated and selected (no divergent control flow); t = 0 for
// s >= 2 makes the outer branch the zero of the kernel's support.
template <int DK>
__device__ __forceinline__ float Wk_s(float s, float C) {
  const float t = fmaxf(2.f - s, 0.f);
  const float a = C * (6.0f - 6.0f * s + s * s * s);
  const float b = C * t * t * t;
  return s < 1.f ? a : b;
}
template <int DK>
__device__ __forceinline__ float dWk_f(float r, float h) {
  const float ih = frcp(h);
  const float C = kC<DK>() * ipow<DK + 1>(ih);
  const float s = r * ih;
  const float t = fmaxf(2.f - s, 0.f);
  const float a = C * (-6.0f + 3.f * s * s);
  const float b = C * (-3.0f) * t * t;
  return (s < 1.f ? a : b) * frcp(r + 1.e-20f);
}

// relative position x_i - image(x_q) in FP32 from the hi/lo pairs; the image of a
// mirror ghost is 2w - x_q, so its difference is (x_i - w) + (x_q - w)
template <int D>
__device__ __forceinline__ void rel_pos(const JobParams& J, const float4& qh, const float4& ql, uint32_t code,
                                        const float* xhi, const float* xli, float* dr) {
  const float hq[3] = {qh.x, qh.y, qh.z};
  const float lq[3] = {ql.x, ql.y, ql.z};
#pragma unroll
  for (int a = 0; a < D; a++) {
    const uint32_t m = (code >> (2 * a)) & 3u;
    if (m == 0) dr[a] = (xhi[a] - hq[a]) + (xli[a] - lq[a]);
    else if (m == 1) dr[a] = ((xhi[a] - J.wl_h[a]) + (hq[a] - J.wl_h[a])) + ((xli[a] + lq[a]) - 2.f * J.wl_l[a]);
    else dr[a] = ((xhi[a] - J.wh_h[a]) + (hq[a] - J.wh_h[a])) + ((xli[a] + lq[a]) - 2.f * J.wh_l[a]);
  }
}

// ---------------------------------------------------------------------------
// reduction slots of this step

// the accumulator copies of job j, and the copy workgroup b adds into
__device__ __forceinline__ JobAcc* acc_of(const Work& W, int j) { return W.acc + int64_t(j) * kAccCopies; }
__device__ __forceinline__ JobAcc& acc_mine(const Work& W, int j) { return W.acc[int64_t(j) * kAccCopies + blockIdx.x % kAccCopies]; }
// dt1 of this step folded from the copies (wave-uniform scalar loads)
__device__ __forceinline__ double acc_dt(const Work& W, int j, int par, bool second) {
  const JobAcc* A = acc_of(W, j);
  unsigned long long m = second ? A[0].dt2[par] : A[0].dt1[par];
#pragma unroll
  for (int c = 1; c < kAccCopies; c++) {
    const unsigned long long v = second ? A[c].dt2[par] : A[c].dt1[par];
    m = v < m ? v : m;   // non-negative doubles order as their bit patterns
  }
  return bitsd(m);
}
template <typename Op>
__device__ __forceinline__ unsigned long long wave_fold(unsigned long long v, Op op) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = op(v, (unsigned long long)__shfl_xor((long long)v, o, 64));
  return v;
}
struct OpMinU { __device__ unsigned long long operator()(unsigned long long a, unsigned long long b) const { return a < b ? a : b; } };
struct OpMaxU { __device__ unsigned long long operator()(unsigned long long a, unsigned long long b) const { return a > b ? a : b; } };

// Start of a step, one wavefront per job: the previous step's dt1, dt2 and max v
// folded from the copies into JobRed (the integration in k_map reads them), then this
// step's accumulators cleared.
__global__ __launch_bounds__(64) void k_reset(Work W, int njobs, int par) {
  const int j = blockIdx.x, l = threadIdx.x;
  if (j == 0 && l == 0) {   // spill queue and the any-job-rebuilds flag of this step
    W.ctr[0] = 0ull;
    W.ctr[1] = 0ull;
  }
  if (j >= njobs) return;
  JobRed& R = W.red[j];
  JobAcc* A = acc_of(W, j);
  const bool c = l < kAccCopies;
  const unsigned long long d1 = wave_fold(c ? A[l].dt1[par ^ 1] : ~0ull, OpMinU());
  const unsigned long long d2 = wave_fold(c ? A[l].dt2[par ^ 1] : ~0ull, OpMinU());
  const unsigned long long mv = wave_fold(c ? A[l].maxv[par ^ 1] : 0ull, OpMaxU());
  if (c) {
    A[l].dt1[par] = 0x7FF0000000000000ull;
    A[l].dt2[par] = 0x7FF0000000000000ull;
    A[l].maxv[par] = 0ull;
    A[l].res[par ^ 1] = 0ull;
    A[l].maxdisp[par ^ 1] = 0ull;
    A[l].rsum[par ^ 1] = 0ull;
    A[l].drift = 0ull;
  }
  if (l == 0) {
    R.step += 1;   // re-seeding RNG stream of this step (graph replays advance it too)
    R.rebuild = 0;
    R.dt1[par ^ 1] = d1;
    R.dt2[par ^ 1] = d2;
    R.maxv[par ^ 1] = mv;
    R.nremap[par] = 0;
  }
}

// ---------------------------------------------------------------------------
// seeding: Random_particle_distribution (sph_mesh_generation.cpp:1302-1367)

template <int D>
__global__ __launch_bounds__(kTile) void k_seed(Work W, PState S) {
  const int j = W.tile_job[tile_id(W)];
  if (j < 0) return;   // tile not owned by a job
  const JobParams& J = W.jp[j];
  const int64_t slot = tile_id(W) * kTile + threadIdx.x;
  const int64_t loc = slot - J.p0;
  int bad = 0;
  if (loc < J.n) {
    const uint64_t key = rng_key(J.seed, 0xFFFFFFFFull, uint64_t(loc));
    double x[3] = {0., 0., 0.}, phi = 0., n[3] = {0., 0., 0.}, h = 0.;
    const double pad = J.seed_pad;   // one package (of the finest level-set level)
    bool ok = false;
    for (int att = 0; att < J.remap_max && !ok; att++) {
#pragma unroll
      for (int a = 0; a < D; a++) {
        const double box = J.pbox_r[a] - J.pbox_l[a];
        const double u0 = rng_u(key, uint64_t(att) * 6 + 2 * a);
        const double u1 = rng_u(key, uint64_t(att) * 6 + 2 * a + 1);
        double v = u0 * box + u1 * 1.e-3 + J.pbox_l[a];
        const double lo = J.pbox_l[a] + 1.e-10, hi = J.pbox_r[a] - 1.e-10;
        v = (v <= hi) ? v : hi;
        v = (lo >= v) ? lo : v;
        const double llo = J.box_l[a] + pad, lhi = J.box_r[a] - pad;
        v = (lhi <= v) ? lhi : v;
        v = (llo >= v) ? llo : v;
        x[a] = v;
      }
      ls_sample<D, false>(J, W, x, phi, n);
      h = h_of_phi(J, phi);
      ok = phi > h;
    }
    bad = !ok;
#pragma unroll
    for (int a = 0; a < 3; a++) {
      if (a < D) split_store(x[a], S.xh[a], S.xl[a], slot);
      else { S.xh[a][slot] = 0.f; S.xl[a][slot] = 0.f; }
      S.nrm[a][slot] = a < D ? float(n[a]) : 0.f;
      W.a1[a][slot] = 0.f;
      W.a2[a][slot] = 0.f;
    }
    S.h[slot] = h;
    S.type[slot] = SMG_REAL;
    S.id[slot] = int32_t(loc);
    W.nu_b[slot] = 0.f;
  } else {
    S.id[slot] = -1;
    S.type[slot] = -1;
    S.h[slot] = 0.;
#pragma unroll
    for (int a = 0; a < 3; a++) {
      S.xh[a][slot] = 0.f; S.xl[a][slot] = 0.f; S.nrm[a][slot] = 0.f;
      W.a1[a][slot] = 0.f; W.a2[a][slot] = 0.f;
    }
    W.nu_b[slot] = 0.f;
  }
  if (bad) atomicOr(&W.red[j].fail, 1);
}

// ---------------------------------------------------------------------------
// P1: integrate, sample, map, scale, key

// mode (kModeInit, kModePrerun, kModeStep) is a template parameter: each launch
// compiles only its own path (the step pass 114 VGPRs instead of 124)
template <int D, int mode>
__global__ __launch_bounds__(kTile) void k_map(Work W, PState S, int par) {
  __shared__ double lds_d[4];
  __shared__ int lds_i[4];
  __shared__ unsigned long long lds_u[4];
  const int j = W.tile_job[tile_id(W)];
  if (j < 0) return;
  const JobParams& J = W.jp[j];
  const int64_t slot = tile_id(W) * kTile + threadIdx.x;
  const int64_t loc = slot - J.p0;
  double res = 0., md = 0., drift = 0.;
  unsigned long long rq = 0ull;
  int nre = 0, bad = 0;
  if (loc < J.n) {
    double x[3] = {0., 0., 0.};
#pragma unroll
    for (int a = 0; a < D; a++) x[a] = double(S.xh[a][slot]) + double(S.xl[a][slot]);
    double h = S.h[slot];
    int type = S.type[slot];
    // the build position for the drift test, loaded with the particle's state so that
    // its latency overlaps the level-set gathers
    const float4 b = mode == kModeStep ? W.recb[slot] : make_float4(0.f, 0.f, 0.f, 0.f);
    const bool built = mode == kModeStep && W.red[j].built;
    if (mode == kModeStep) {
      // the two Update_velocity_half_step of the previous step, then Update_coord_full_step
      const double dt1 = bitsd(W.red[j].dt1[par ^ 1]), dt2 = bitsd(W.red[j].dt2[par ^ 1]);
      double v[3] = {0., 0., 0.}, U = 0.;
#pragma unroll
      for (int a = 0; a < D; a++) {
        v[a] = double(W.a1[a][slot]) * (0.5 * dt1);
        v[a] = v[a] + double(W.a2[a][slot]) * (0.5 * dt2);
        U += v[a] * v[a];
      }
      U = sqrt(U);
      const Disp dd = disp_of(U, dt2, h, J);
      res = dd.r;
      md = dd.md;
      rq = dd.rq;
      if (type != SMG_SINGULARITY) {
#pragma unroll
        for (int a = 0; a < D; a++) x[a] = x[a] + v[a] * dt2;
      }
    }
    // phi only: the normal matters to a particle in a cut cell (projection) and to a
    // surface particle, and is sampled there; an interior particle's stays 0, unread
    double phi, n[3] = {0., 0., 0.};
    ls_sample<D, false>(J, W, x, phi, n);
    if (mode == kModeInit) {
      h = h_of_phi(J, phi);
    } else {
      if (mode == kModePrerun) h = h_of_phi(J, phi);
      // Find_and_map_particle (particle_mesh_generation.cpp:408-467)
      const int tag = ls_tag<D>(J, W, x);
      if (tag == 1 && (type == SMG_REAL || type == SMG_SURFACE)) {
        ls_sample<D, true>(J, W, x, phi, n);   // the same phi, and the normal
        type = SMG_SURFACE;
        const double dist = fabs(phi);
        double nn = 0.;
#pragma unroll
        for (int a = 0; a < D; a++) nn += n[a] * n[a];
        const double div = 1.0 / (sqrt(nn) + 1.0e-15);
#pragma unroll
        for (int a = 0; a < D; a++) {
          const double nd = phi < 0.0 ? n[a] * div : -(n[a] * div);
          x[a] = x[a] + nd * dist;
        }
      } else {
        type = SMG_REAL;
      }
      if (type == SMG_SURFACE) {   // Get_dispersed_position(1e-5 h) with the pre-projection phi, n (Q12)
        const double dx = 0.00001 * h;
        double nn = 0.;
#pragma unroll
        for (int a = 0; a < D; a++) nn += n[a] * n[a];
        const double div = 1.0 / (sqrt(nn) + 1.0e-15);
#pragma unroll
        for (int a = 0; a < D; a++) {
          const double nd = phi < 0.0 ? n[a] * div : -(n[a] * div);
          x[a] = x[a] + nd * dx;
        }
      }
      // sample again where the particle moved (a REAL particle keeps x, phi and n)
      if (type == SMG_SURFACE) ls_sample<D, true>(J, W, x, phi, n);
      if (phi < 0. && type == SMG_REAL) {
        // Remap_this_particle_to_positive_phase (sph_mesh_generation.cpp:1885-1928)
        nre = 1;
        const uint64_t stream = mode == kModeStep ? uint64_t(W.red[j].step) : 0ull;
        const uint64_t key_r = rng_key(J.seed, stream, uint64_t(S.id[slot]));
        bool ok = false;
        for (int att = 0; att < J.remap_max && !ok; att++) {
#pragma unroll
          for (int a = 0; a < D; a++) {
            const double box = J.pbox_r[a] - J.pbox_l[a];
            const double u0 = rng_u(key_r, uint64_t(att) * 6 + 2 * a);
            const double u1 = rng_u(key_r, uint64_t(att) * 6 + 2 * a + 1);
            double v = u0 * box + u1 * 1.e-3 + J.pbox_l[a];
            const double lo = J.pbox_l[a] + 1.e-10, hi = J.pbox_r[a] - 1.e-10;
            v = (v <= hi) ? v : hi;
            v = (lo >= v) ? lo : v;
            x[a] = v;
          }
          ls_sample<D, false>(J, W, x, phi, n);
          ok = phi > h;
        }
        if (ok) {
          type = SMG_REAL;
          h = h_of_phi(J, phi);
        } else {
          bad = 1;
        }
      }
      if (mode == kModeStep) h = h_of_phi(J, phi);   // Set_particle_scale
    }
#pragma unroll
    for (int a = 0; a < D; a++) {
      split_store(x[a], S.xh[a], S.xl[a], slot);
      S.nrm[a][slot] = float(n[a]);
    }
    S.h[slot] = h;
    S.type[slot] = type;
    float hp[3] = {0.f, 0.f, 0.f}, lp[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int a = 0; a < D; a++) {
      hp[a] = float(x[a]);
      lp[a] = float(x[a] - double(hp[a]));
    }
    const float4 r0 = make_float4(hp[0], hp[1], hp[2], float(h));
    bool jump = false;
    if (mode == kModeStep) {
      // drift since the last list build, as a fraction of h_b.  A particle whose drift
      // reaches the limit becomes a jumper (its pairs are found without the lists until
      // the next build); the job rebuilds when the other particles' drift reaches it or
      // when there are more than kJumpMax jumpers.
      if (b.w < 0.f) {
        jump = true;
      } else {
        const double dx0 = double(r0.x) - double(b.x), dx1 = double(r0.y) - double(b.y);
        const double dx2 = D > 2 ? double(r0.z) - double(b.z) : 0.;
        const double e = sqrt(dx0 * dx0 + dx1 * dx1 + dx2 * dx2) + 4.0 * double(J.pf_tol) +
                         J.cut_off * fabs(h - double(b.w));
        double rho = e / double(b.w) * (1.0 + 1.e-5) + 1.e-9;
        rho = (rho <= 1.e300) ? rho : 1.e300;   // also for b.w = 0 or NaN
        if (rho >= double(J.skin_lim) && J.skin_lim >= 0.f && built) {
          const int ji = atomicAdd(&W.red[j].njump, 1);
          if (ji < kJumpMax) {
            W.jset[int64_t(j) * kJumpMax + ji] = int32_t(slot);
            W.recb[slot].w = -b.w;
            jump = true;
          } else {
            atomicOr(&W.red[j].jover, 1);
          }
        } else {
          drift = rho;
        }
      }
    }
    const float4 r1 = make_float4(lp[0], lp[1], lp[2], __int_as_float(type | (jump ? kJumpBit : 0)));
    W.rec[4 * int64_t(slot) + 0] = r0;
    W.rec[4 * int64_t(slot) + 1] = r1;
    W.rec[4 * int64_t(slot) + 3] = make_float4(float(n[0]), D > 1 ? float(n[1]) : 0.f, D > 2 ? float(n[2]) : 0.f, 0.f);
  }
  drift = block_reduce(drift, OpMax(), lds_d);
  res = block_reduce(res, OpMax(), lds_d);
  md = block_reduce(md, OpMax(), lds_d);
  rq = block_reduce(rq, OpAddU(), lds_u);
  nre = block_reduce(nre, OpAdd(), lds_i);
  bad = block_reduce(bad, OpMaxI(), lds_i);
  if (threadIdx.x == 0) {
    JobRed& R = W.red[j];
    if (mode == kModeStep) {   // residual of the previous step, whose motion this pass applied
      JobAcc& A = acc_mine(W, j);
      atomicMax(&A.drift, dbits(drift));
      atomicMax(&A.res[par ^ 1], dbits(res));
      atomicMax(&A.maxdisp[par ^ 1], dbits(md));
      if (rq) {
        atomicAdd(&A.rsum[par ^ 1], rq);
        atomicAdd(&A.rcum, rq);
      }
    }
    if (nre) atomicAdd(&R.nremap[par], nre);
    if (bad) atomicOr(&R.fail, 1);
  }
}

// ---------------------------------------------------------------------------
// list rebuild decision per job: no lists yet, forced, or some particle's drift
// reached skin * cut_off / 2 of its h at the last build

// One wavefront per job: the rebuild decision, and otherwise the job's jumpers put in
// slot order (a rank sort in registers: the slots are distinct), so that every pass
// visits them in the same order.
static_assert(kJumpMax == 64, "one lane per jumper");
__global__ __launch_bounds__(64) void k_decide(Work W, int njobs, int force) {
  const int j = blockIdx.x, l = threadIdx.x;
  if (j >= njobs) return;
  JobRed& R = W.red[j];
  if (W.jp[j].n == 0) {   // an empty job index of a pool
    if (l == 0) R.rebuild = 0;
    return;
  }
  const JobAcc* A = acc_of(W, j);
  const unsigned long long drift = wave_fold(l < kAccCopies ? A[l].drift : 0ull, OpMaxU());
  const int rb = force || !R.built || R.jover || bitsd(drift) >= double(W.jp[j].skin_lim);
  const int njump = R.njump;
  __syncthreads();   // every lane has read R before lane 0 writes it
  if (rb) {
    if (l == 0) {
      R.rebuild = 1;
      R.spill_used = 0ull;
      R.items_used = 0ull;
      R.njump = 0;
      R.jover = 0;
      if (!force) R.nrebuild += 1;
      W.ctr[1] = 1ull;   // the table scan runs this step
    }
    return;
  }
  if (l == 0) R.rebuild = 0;
  int32_t* js = W.jset + int64_t(j) * kJumpMax;
  const int nj = njump < kJumpMax ? njump : kJumpMax;
  const int32_t v = l < nj ? js[l] : 0x7fffffff;
  int rank = 0;
  for (int k = 0; k < nj; k++) rank += __shfl(v, k, 64) < v;
  if (l < nj) js[rank] = v;
}

// cell key and count of every slot of a rebuilding job
template <int D>
__global__ __launch_bounds__(kTile) void k_key(Work W, PState S) {
  if (!rebuild_any(W)) return;
  const int j = W.tile_job[tile_id(W)];
  if (j < 0 || !W.red[j].rebuild) return;
  const JobParams& J = W.jp[j];
  const int64_t slot = tile_id(W) * kTile + threadIdx.x;
  uint32_t key;
  if (slot - J.p0 < J.n) {
    const double h = S.h[slot];
    double x[3] = {0., 0., 0.};
#pragma unroll
    for (int a = 0; a < D; a++) x[a] = double(S.xh[a][slot]) + double(S.xl[a][slot]);
    const int L = level_of(J, h);
    int c[3] = {0, 0, 0};
    cell_at<D>(J, L, x, c);
    key = uint32_t(J.tbl_base + J.lev_off[L]) + morton<D>(c);
  } else {
    key = uint32_t(J.tbl_base + J.tbl_size - 1);
  }
  W.key[slot] = key;
  atomicAdd(&W.counts[key], 1);
}

// ---------------------------------------------------------------------------
// cell starts: exclusive scan of counts over all job tables

// gate (may be NULL): skip the scan on a step where no job rebuilds.  blk_job (may be
// NULL): the job owning each 4096-entry block (job table segments are block aligned);
// the blocks of jobs that do not rebuild count as empty and keep their starts, which
// stay valid because a job's starts are only ever read relative to its first entry.
__device__ __forceinline__ bool scan_skip(const int32_t* blk_job, const JobRed* red) {
  if (!blk_job) return false;
  const int j = blk_job[blockIdx.x];
  return j < 0 || !red[j].rebuild;
}
__global__ __launch_bounds__(kTile) void k_scan_reduce(const int32_t* __restrict__ counts, int64_t n, int32_t* bsum,
                                                       const unsigned long long* gate, const int32_t* blk_job,
                                                       const JobRed* red) {
  if (gate && !*gate) return;
  if (scan_skip(blk_job, red)) {
    if (threadIdx.x == 0) bsum[blockIdx.x] = 0;
    return;
  }
  __shared__ int lds[4];
  const int64_t base = int64_t(blockIdx.x) * kScanItems + int64_t(threadIdx.x) * 16;
  int s = 0;
#pragma unroll
  for (int q = 0; q < 16; q++) s += (base + q < n) ? counts[base + q] : 0;
  s = block_reduce(s, OpAdd(), lds);
  if (threadIdx.x == 0) bsum[blockIdx.x] = s;
}

__global__ __launch_bounds__(1024) void k_scan_top(int32_t* bsum, int64_t nb, const unsigned long long* gate) {
  if (gate && !*gate) return;
  __shared__ int lds[1024];
  int carry = 0;
  for (int64_t c0 = 0; c0 < nb; c0 += 1024) {
    const int64_t i = c0 + threadIdx.x;
    const int v = i < nb ? bsum[i] : 0;
    lds[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
      const int t = threadIdx.x >= (unsigned)o ? lds[threadIdx.x - o] : 0;
      __syncthreads();
      lds[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < nb) bsum[i] = carry + lds[threadIdx.x] - v;
    const int tot = lds[1023];
    __syncthreads();
    carry += tot;
  }
  if (threadIdx.x == 0) bsum[nb] = carry;
}

__global__ __launch_bounds__(kTile) void k_scan_down(const int32_t* __restrict__ counts, int64_t n,
                                                     const int32_t* __restrict__ bsum, int64_t nb, int32_t* starts,
                                                     const unsigned long long* gate, const int32_t* blk_job,
                                                     const JobRed* red) {
  if (gate && !*gate) return;
  if (scan_skip(blk_job, red)) return;
  __shared__ int lds[kTile];
  const int64_t base = int64_t(blockIdx.x) * kScanItems + int64_t(threadIdx.x) * 16;
  int v[16];
  int s = 0;
#pragma unroll
  for (int q = 0; q < 16; q++) {
    v[q] = (base + q < n) ? counts[base + q] : 0;
    s += v[q];
  }
  lds[threadIdx.x] = s;
  __syncthreads();
  for (int o = 1; o < kTile; o <<= 1) {
    const int t = threadIdx.x >= (unsigned)o ? lds[threadIdx.x - o] : 0;
    __syncthreads();
    lds[threadIdx.x] += t;
    __syncthreads();
  }
  int run = bsum[blockIdx.x] + lds[threadIdx.x] - s;
#pragma unroll
  for (int q = 0; q < 16; q++) {
    if (base + q < n) starts[base + q] = run;
    run += v[q];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) starts[n] = bsum[nb];   // end sentinel
}

// ---------------------------------------------------------------------------
// counting sort by cell key

// A job's cell starts are taken relative to the start of its table segment, so the
// slot ranges of the jobs may sit anywhere (gaps included) and still sort in place.
__global__ __launch_bounds__(kTile) void k_scatter(Work W, PState S) {
  if (!rebuild_any(W)) return;
  const int j = W.tile_job[tile_id(W)];
  if (j < 0 || !W.red[j].rebuild) return;
  const JobParams& J = W.jp[j];
  const int64_t p = tile_id(W) * kTile + threadIdx.x;
  const uint32_t k = W.key[p];
  const int r = atomicSub(&W.counts[k], 1) - 1;
  const int64_t d = J.p0 + int64_t(W.starts[k] - W.starts[J.tbl_base]) + r;
  W.perm[d] = int32_t(p);
  W.skey[d] = k;
  W.sid[d] = S.id[p];
}
// Order inside each cell by particle id (deterministic whatever order the scatter's
// atomics took), a rank sort: slot d of cell [c0, c1) goes to c0 + #{a : sid[a] <
// sid[d]}; the ids of a job are distinct.  Dead slots (the pad cell) keep their place.
__global__ __launch_bounds__(kTile) void k_segsort(Work W) {
  if (!rebuild_any(W)) return;
  const int j = W.tile_job[tile_id(W)];
  if (j < 0 || !W.red[j].rebuild) return;
  const JobParams& J = W.jp[j];
  const int64_t d = tile_id(W) * kTile + threadIdx.x;
  const uint32_t k = W.skey[d];
  if (k == uint32_t(J.tbl_base + J.tbl_size - 1)) {
    W.sperm[d] = W.perm[d];
    return;
  }
  const int64_t base = J.p0 - int64_t(W.starts[J.tbl_base]);
  const int64_t c0 = base + W.starts[k], c1 = base + W.starts[k + 1];
  const int32_t me = W.sid[d];
  int rank = 0;
  for (int64_t a = c0; a < c1; a++) rank += W.sid[a] < me ? 1 : 0;
  W.sperm[c0 + rank] = W.perm[d];
}

// new order into B (the scratch set); k_copyback then moves it into A, so A is the
// current state at every step whichever jobs re-sorted
template <int D>
__global__ __launch_bounds__(kTile) void k_gather(Work W, PState A, PState B) {
  if (!rebuild_any(W)) return;
  const int j = W.tile_job[tile_id(W)];
  if (j < 0 || !W.red[j].rebuild) return;
  const int64_t d = tile_id(W) * kTile + threadIdx.x;
  const int64_t p = W.sperm[d];
#pragma unroll
  for (int a = 0; a < D; a++) {
    B.xh[a][d] = A.xh[a][p];
    B.xl[a][d] = A.xl[a][p];
    B.nrm[a][d] = A.nrm[a][p];
  }
  B.h[d] = A.h[p];
  B.type[d] = A.type[p];
  B.id[d] = A.id[p];
  W.nu_a[d] = W.nu_b[p];
}

template <int D>
__global__ __launch_bounds__(kTile) void k_copyback(Work W, PState A, PState B) {
  if (!rebuild_any(W)) return;
  const int j = W.tile_job[tile_id(W)];
  if (j < 0 || !W.red[j].rebuild) return;
  const int64_t d = tile_id(W) * kTile + threadIdx.x;
  float xh[3] = {0.f, 0.f, 0.f}, xl[3] = {0.f, 0.f, 0.f}, nn[3] = {0.f, 0.f, 0.f};
#pragma unroll
  for (int a = 0; a < D; a++) {
    xh[a] = B.xh[a][d];
    xl[a] = B.xl[a][d];
    nn[a] = B.nrm[a][d];
    A.xh[a][d] = xh[a];
    A.xl[a][d] = xl[a];
    A.nrm[a][d] = nn[a];
  }
  const double h = B.h[d];
  const int32_t ty = B.type[d];
  A.h[d] = h;
  A.type[d] = ty;
  A.id[d] = B.id[d];
  W.nu_b[d] = W.nu_a[d];
  W.rec[4 * int64_t(d) + 0] = make_float4(xh[0], xh[1], xh[2], float(h));
  W.rec[4 * int64_t(d) + 1] = make_float4(xl[0], xl[1], xl[2], __int_as_float(ty));
  W.recb[d] = make_float4(xh[0], xh[1], xh[2], float(h));   // build position and h
  W.rec[4 * int64_t(d) + 3] = make_float4(nn[0], nn[1], nn[2], 0.f);
}

// ---------------------------------------------------------------------------
// neighbour lists (Level_info::Refresh_neighbor_info with mirror ghosts)

template <int D>
__device__ __forceinline__ void image_of(const JobParams& J, uint32_t code, const double* x, double* g) {
#pragma unroll
  for (int a = 0; a < D; a++) {
    const uint32_t m = (code >> (2 * a)) & 3u;
    g[a] = (m == 1) ? 2. * J.box_l[a] - x[a] : ((m == 2) ? 2. * J.box_r[a] - x[a] : x[a]);
  }
}

// The FP32 relative position to the image (code) of a slot q: dr_a = A_a + sg_a xh_q,
// from particle i's hi parts (code 0: A = xh_i, sg = -1, so dr = xh_i - xh_q exactly)
template <int D>
__device__ __forceinline__ void image_affine(const JobParams& J, uint32_t code, const float* xh, float* A, float* sg) {
#pragma unroll
  for (int a = 0; a < D; a++) {
    const uint32_t m = (code >> (2 * a)) & 3u;
    A[a] = m == 0 ? xh[a] : (m == 1 ? xh[a] - 2.f * J.wl_h[a] : xh[a] - 2.f * J.wh_h[a]);
    sg[a] = m == 0 ? -1.f : 1.f;
  }
}

// Enumerates the list candidates of one particle in a fixed order (mirror images,
// then levels, then cells, then slots): every slot (or mirror image) within
// cut_v * max(h_i, h_q) by the FP32 distance of the hi parts, widened by that
// distance's error bound -- a superset, so no FP64 decision is needed here.  Hands
// each entry (k, slot | code << 26, near) to emit and returns the count; near: within
// the cutoff radius itself by that distance (the pair passes take such entries first).
template <int D, class Emit>
__device__ __forceinline__ int nb_enumerate(const Work& W, const PState& S, const JobParams& J, int64_t slot,
                                            Emit emit) {
  int cnt = 0;
  const float4 r0 = W.rec[4 * int64_t(slot) + 0], r1 = W.rec[4 * int64_t(slot) + 1];
  const float xh[3] = {r0.x, r0.y, r0.z}, xl[3] = {r1.x, r1.y, r1.z};
  double xi[3] = {0., 0., 0.};
#pragma unroll
  for (int a = 0; a < D; a++) xi[a] = double(xh[a]) + double(xl[a]);
  const double hi = S.h[slot];
  const float rfi = float(hi * J.cut_v);
  const int li = level_of(J, hi);
  const int32_t soff = int32_t(J.p0) - W.starts[J.tbl_base];   // table-relative starts -> slots
  // mirror options per axis: none, low wall, high wall
  int opt[3][3], nopt[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    nopt[a] = 1;
    opt[a][0] = 0;
    if (a < D && J.sym[a]) {
      if (xi[a] - J.box_l[a] < J.c0) opt[a][nopt[a]++] = 1;
      if (J.box_r[a] - xi[a] < J.c0) opt[a][nopt[a]++] = 2;
    }
  }
  const float cutf = float(J.cut_v);
  // the job's constants in registers: JobParams sits in global memory, and the list
  // stores of emit would otherwise make the loops below read them again each time
  const float pf_tol = J.pf_tol;
  const float nfac = float(J.cut_off / J.cut_v);   // the cutoff radius over the list radius
  const int64_t p0 = J.p0;
  const int nlev = J.nlev;
  const int64_t tbl = J.tbl_base;
  // one image: every level and stencil cell; the plain image (code 0, the first for
  // every lane) tests xh_i - xh_q directly, a mirror image A + sg xh_q
  auto image = [&](uint32_t code, auto plain) {
    double y[3] = {0., 0., 0.};
    image_of<D>(J, code, xi, y);
    // FP32 relative position: dr_a = A_a + s_a * xh_q
    float A[3] = {0.f, 0.f, 0.f}, sg[3] = {0.f, 0.f, 0.f};
    image_affine<D>(J, code, xh, A, sg);
    // one candidate: the FP32 distance of the hi parts against the widened radius (bit
    // 0), and whether it is within the cutoff radius itself (bit 1: near)
    auto test = [&](const float4& qr) -> int {
      const float qv[3] = {qr.x, qr.y, qr.z};
      float d2 = 0.f;
#pragma unroll
      for (int a = 0; a < D; a++) {
        const float t = decltype(plain)::value ? xh[a] - qv[a] : A[a] + sg[a] * qv[a];
        d2 += t * t;
      }
      const float r = fmaxf(rfi, qr.w * cutf);
      const float rhi = r + (pf_tol + 1.e-6f * r);
      const float rn = r * nfac;
      return (d2 <= rhi * rhi ? 1 : 0) | (d2 <= rn * rn ? 2 : 0);
    };
    for (int L = 0; L < nlev; L++) {
      const int s = L < li ? L : li;
      const int shift = D * (L - s);
      int cy[3] = {0, 0, 0};
      cell_at<D>(J, s, y, cy);
      const int ln[3] = {J.lev_n[s][0], J.lev_n[s][1], D > 2 ? J.lev_n[s][2] : 1};
      const int32_t* st = W.starts + (tbl + J.lev_off[L]);
      // Stencil cells that no candidate can reach are skipped.  A slot of the level-L
      // table has h cut_v + 2e-10 <= clen[L] (level_of), so a candidate passes only
      // within R = max(rfi, clen[L]) (+ the FP32 slack: pf_tol of the radius, pf_tol of
      // the FP32 distance, relative roundings) of y.  The slots of stencil cell
      // cy + o lie beyond the faces y's own cell shares with it (a position clamped
      // into a border cell lies beyond its inner face too), so their distance from y
      // is at least the gap gq[a][o_a + 1] to those faces on each axis.
      // A stencil cell of a finer table (L > s) is a range of level-L cells; its 2^D
      // halves at level s + 1 are consecutive sub-ranges, each tested the same way on
      // the one face of it that faces y: gh[a][2 (o_a + 1) + b_a] for half b_a of
      // offset o_a (the halves' ranges, taken in order, keep the candidates' order).
      float gq[3][3], gh[3][6];
      {
        const double cl = J.clen[s];
#pragma unroll
        for (int a = 0; a < 3; a++) {
          gq[a][0] = gq[a][1] = gq[a][2] = 0.f;
#pragma unroll
          for (int b = 0; b < 6; b++) gh[a][b] = 0.f;
          if (a < D) {
            const double f = J.cbox_l[a] + double(cy[a]) * cl;
            const float gm = float(fmax(0., y[a] - f)), gp = float(fmax(0., f + cl - y[a]));
            gq[a][0] = gm * gm;
            gq[a][2] = gp * gp;
            const float g[6] = {float(fmax(0., y[a] - (f - 0.5 * cl))), gm, float(fmax(0., y[a] - (f + 0.5 * cl))),
                                float(fmax(0., (f + 0.5 * cl) - y[a])), gp, float(fmax(0., (f + 1.5 * cl) - y[a]))};
#pragma unroll
            for (int b = 0; b < 6; b++) gh[a][b] = g[b] * g[b];
          }
        }
      }
      const bool halves = shift >= D;   // the stencil cells of a finer table
      const float Rl = fmaxf(rfi, float(J.clen[L])) * (1.f + 1.e-5f) + 2.f * pf_tol + 1.e-9f;
      const float R2 = Rl * Rl;
      for (int dz = (D > 2 ? -1 : 0); dz <= (D > 2 ? 1 : 0); dz++)
        for (int dy = -1; dy <= 1; dy++) {
          // the slot ranges of the row's three cells, their bounds loaded together
          int32_t lo[3], hi2[3];
          const float g2 = gq[1][dy + 1] + gq[2][dz + 1];
          // the row's half gaps on axes 1 and 2 (selects: no indexing by dy, dz)
          const float g1l = dy < 0 ? gh[1][0] : (dy == 0 ? gh[1][2] : gh[1][4]);
          const float g1h = dy < 0 ? gh[1][1] : (dy == 0 ? gh[1][3] : gh[1][5]);
          const float g2l = dz < 0 ? gh[2][0] : (dz == 0 ? gh[2][2] : gh[2][4]);
          const float g2h = dz < 0 ? gh[2][1] : (dz == 0 ? gh[2][3] : gh[2][5]);
#pragma unroll
          for (int u = 0; u < 3; u++) {
            int c[3] = {cy[0] + u - 1, cy[1] + dy, cy[2] + dz};
            bool ok = c[0] >= 0 && c[0] < ln[0] && c[1] >= 0 && c[1] < ln[1];
            if (D > 2) ok = ok && c[2] >= 0 && c[2] < ln[2];
            ok = ok && g2 + gq[0][u] <= R2;
            lo[u] = 0;
            hi2[u] = 0;
            if (ok) {
              const uint64_t m = morton<D>(c);
              lo[u] = st[int64_t(m << shift)] + soff;
              hi2[u] = st[int64_t((m + 1) << shift)] + soff;
            }
          }
          // candidates four at a time: the four records are loaded before any is
          // tested, from one address (recb has kRecbPad records past its last slot, so
          // the loads past the range stay inside it; their tests are masked)
          auto scan = [&](int32_t qlo, int32_t qhi) __attribute__((always_inline)) {
            for (int32_t q = qlo; q < qhi; q += 4) {
              float4 qr[4];
              const float4* rq = W.recb + uint32_t(q);   // == rec[0] at a build
#pragma unroll
              for (int v = 0; v < 4; v++) qr[v] = rq[v];
#pragma unroll
              for (int v = 0; v < 4; v++) {
                const int tv = q + v < qhi ? test(qr[v]) : 0;
                if (tv & 1) {
                  emit(cnt, int32_t(nb_pack(q + v, p0, code)), (tv & 2) != 0);
                  cnt++;
                }
              }
            }
          };
          // cell u of the row (u a compile-time constant: the row's arrays stay in registers)
          auto cell = [&](auto u_) __attribute__((always_inline)) {
            constexpr int u = decltype(u_)::value;
            if (!halves) {
              scan(lo[u], hi2[u]);
            } else if (lo[u] < hi2[u]) {
              const int c[3] = {cy[0] + u - 1, cy[1] + dy, cy[2] + dz};
              const uint64_t mc = uint64_t(morton<D>(c)) << D;
              const int sh = shift - D;
              const float g0l = gh[0][2 * u], g0h = gh[0][2 * u + 1];
              int32_t hl = lo[u];
              for (int b = 0; b < (1 << D); b++) {
                const int32_t hh2 = b + 1 < (1 << D) ? st[int64_t((mc + uint64_t(b + 1)) << sh)] + soff : hi2[u];
                float g = ((b & 1) ? g0h : g0l) + ((b & 2) ? g1h : g1l);
                if (D > 2) g += (b & 4) ? g2h : g2l;
                if (g <= R2) scan(hl, hh2);
                hl = hh2;
              }
            }
          };
          cell(std::integral_constant<int, 0>{});
          cell(std::integral_constant<int, 1>{});
          cell(std::integral_constant<int, 2>{});
        }
    }
  };
  image(0u, std::true_type{});
  for (int o0 = 0; o0 < nopt[0]; o0++)
    for (int o1 = 0; o1 < nopt[1]; o1++)
      for (int o2 = 0; o2 < nopt[2]; o2++) {
        if (o0 + o1 + o2 == 0) continue;   // the plain image, done above
        image(uint32_t(opt[0][o0] | (opt[1][o1] << 2) | (opt[2][o2] << 4)), std::false_type{});
      }
  return cnt;
}

// candidate lists of the rebuilding jobs: the first K entries inline, a slot with
// more joins the spill queue (k_copyback left the build positions and h in recb)
template <int D>
__global__ __launch_bounds__(kTile) void k_nb(Work W, PState S) {
  if (!rebuild_any(W)) return;
  const int j = W.tile_job[tile_id(W)];
  if (j < 0 || !W.red[j].rebuild) return;
  const JobParams& J = W.jp[j];
  const int64_t slot = tile_id(W) * kTile + threadIdx.x;
  const int64_t loc = slot - J.p0;
  int cnt = 0;
  if (loc < J.n) {
    const int64_t cap = W.cap_total;
    const int K = W.K;
    // each group of four inline entries is collected in registers and written with
    // one 16-byte store; a last partial group is written as it stands
    // the near bits of the inline entries, 32 to a word, likewise
    uint4* nb4 = reinterpret_cast<uint4*>(W.nbl);
    uint4 grp = make_uint4(0u, 0u, 0u, 0u);
    uint32_t nw = 0u;
    cnt = nb_enumerate<D>(W, S, J, slot, [&](int k, int32_t e, bool near) {
      if (k >= K) return;
      const int u = k & 3;
      grp.x = u == 0 ? uint32_t(e) : grp.x;
      grp.y = u == 1 ? uint32_t(e) : grp.y;
      grp.z = u == 2 ? uint32_t(e) : grp.z;
      grp.w = u == 3 ? uint32_t(e) : grp.w;
      if (u == 3) nb4[int64_t(k >> 2) * cap + slot] = grp;
      nw |= near ? 1u << (k & 31) : 0u;
      if ((k & 31) == 31) {
        W.nbnear[int64_t(k >> 5) * cap + slot] = nw;
        nw = 0u;
      }
    });
    if ((cnt & 3) && cnt < K) nb4[int64_t(cnt >> 2) * cap + slot] = grp;
    // K is a multiple of four, not of 32: a spilling list can end its inline part
    // inside a word too
    const int ni = cnt < K ? cnt : K;
    if (ni & 31) W.nbnear[int64_t(ni >> 5) * cap + slot] = nw;
    if (cnt > K) {
      const unsigned long long qi = atomicAdd(&W.ctr[0], 1ull);
      W.ovq[qi] = int32_t(slot);   // at most one entry per slot: qi < cap_total
    }
  }
  W.nbc[slot] = cnt;
  if (threadIdx.x == 0) W.red[j].built = 1;
}

// entries K.. of the queued slots, written to the job's part of the spill pool; the
// enumeration repeats the list pass exactly, so the entries keep their order.  A
// slot that finds its job's part full keeps its first K entries and flags the job.
template <int D>
__global__ __launch_bounds__(kTile) void k_nb_spill(Work W, PState S) {
  const unsigned long long nq = W.ctr[0];
  const int K = W.K;
  for (unsigned long long i = uint64_t(blockIdx.x) * kTile + threadIdx.x; i < nq; i += uint64_t(gridDim.x) * kTile) {
    const int64_t slot = W.ovq[i];
    const int j = W.tile_job[slot / kTile];
    const JobParams& J = W.jp[j];
    const int cnt = W.nbc[slot];
    const unsigned long long need = (unsigned long long)(cnt - K);
    const unsigned long long off = atomicAdd(&W.red[j].spill_used, need);
    if (off + need > (unsigned long long)J.spill_cap) {
      W.nbc[slot] = K;
      atomicOr(&W.red[j].overflow, kOverSpill);
      continue;
    }
    const int64_t at = J.spill_base + int64_t(off);
    W.nbo[slot] = int32_t(at);
    int32_t* dst = W.spill + at;
    nb_enumerate<D>(W, S, J, slot, [&](int k, int32_t e, bool) {
      if (k >= K) dst[k - K] = e;
    });
  }
}

// The reference's neighbour predicate (Level_info::Refresh_neighbor_info with
// CUT_OFF * h of either particle, level_infor.cpp:469-506) on one list entry: the
// FP32 distance d2 of the compensated relative position decides outside its error
// band; inside it the FP64 distance of the stored positions decides.
template <int D>
__device__ __forceinline__ bool is_neighbor(const JobParams& J, const Work& W, const PState& S, int64_t q,
                                            uint32_t code, float d2, float rfi, float hq, const float* xhi,
                                            const float* xli, double hid) {
  const float r = fmaxf(rfi, hq * float(J.cut_off));
  const float band = J.pf_tol + 1.e-6f * r;
  const float rlo = r - band, rhi = r + band;
  if (d2 < rlo * rlo) return true;
  if (d2 > rhi * rhi) return false;
  const float4 qr = W.rec[4 * int64_t(q) + 0], ql = W.rec[4 * int64_t(q) + 1];
  const double xq[3] = {double(qr.x) + double(ql.x), double(qr.y) + double(ql.y), double(qr.z) + double(ql.z)};
  double g[3];
  image_of<D>(J, code, xq, g);
  double dist = 0.0;
#pragma unroll
  for (int a = 0; a < D; a++) {
    const double dr = (double(xhi[a]) + double(xli[a])) - g[a];
    dist += dr * dr;
  }
  dist = sqrt(dist);
  return dist <= hid * J.cut_off + 1.e-10 || dist <= S.h[q] * J.cut_off + 1.e-10;
}

// Candidate lists of this step's jumpers: the cell search of the last build's tables
// around each jumper's current position (the other particles are within their drift
// bound of their build positions, so the lists cover the jumper's neighbours), other
// jumpers left out (every particle pairs with the jumpers separately).  One wavefront
// per jumper: lane l takes the (mirror image, level, cell) triples l, l + 64, ...;
// counts, a wave prefix sum and a second pass write the entries in triple order.  A
// list longer than kJumpList makes the job rebuild.
template <int D>
__global__ __launch_bounds__(64) void k_jump_lists(Work W, PState S) {
  const int j = blockIdx.y, t = blockIdx.x;
  const JobParams& J = W.jp[j];
  if (J.n == 0 || W.red[j].jover) return;
  const int nj0 = W.red[j].njump;
  if (t >= (nj0 < kJumpMax ? nj0 : kJumpMax)) return;
  const int lane = threadIdx.x;
  const int64_t slot = W.jset[int64_t(j) * kJumpMax + t];
  int32_t* lst = W.jl + (int64_t(j) * kJumpMax + t) * kJumpList;
  const float4 r0 = W.rec[4 * slot + 0], r1 = W.rec[4 * slot + 1];
  const float xh[3] = {r0.x, r0.y, r0.z}, xl[3] = {r1.x, r1.y, r1.z};
  double xi[3] = {0., 0., 0.};
#pragma unroll
  for (int a = 0; a < D; a++) xi[a] = double(xh[a]) + double(xl[a]);
  const double hi = S.h[slot];
  const float rfi = float(hi * J.cut_v);
  const int li = level_of(J, hi);
  const int32_t soff = int32_t(J.p0) - W.starts[J.tbl_base];
  int opt[3][3], nopt[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    nopt[a] = 1;
    opt[a][0] = 0;
    if (a < D && J.sym[a]) {
      if (xi[a] - J.box_l[a] < J.c0) opt[a][nopt[a]++] = 1;
      if (J.box_r[a] - xi[a] < J.c0) opt[a][nopt[a]++] = 2;
    }
  }
  const int ncell = D > 2 ? 27 : 9;
  const int T = nopt[0] * nopt[1] * nopt[2] * J.nlev * ncell;
  const float cutf = float(J.cut_v);
  int total = 0;
  for (int base = 0; base < T; base += 64) {
    const int idx = base + lane;
    int32_t lo = 0, hi2 = 0;
    uint32_t code = 0u;
    float A[3] = {0.f, 0.f, 0.f}, sg[3] = {0.f, 0.f, 0.f};
    if (idx < T) {
      const int cell = idx % ncell;
      const int L = (idx / ncell) % J.nlev;
      const int ci = idx / (ncell * J.nlev);
      const int o0 = ci % nopt[0], o1 = (ci / nopt[0]) % nopt[1], o2 = ci / (nopt[0] * nopt[1]);
      code = uint32_t(opt[0][o0] | (opt[1][o1] << 2) | (opt[2][o2] << 4));
      double y[3] = {0., 0., 0.};
      image_of<D>(J, code, xi, y);
#pragma unroll
      for (int a = 0; a < D; a++) {
        const uint32_t m = (code >> (2 * a)) & 3u;
        A[a] = m == 0 ? xh[a] : (m == 1 ? xh[a] - 2.f * J.wl_h[a] : xh[a] - 2.f * J.wh_h[a]);
        sg[a] = m == 0 ? -1.f : 1.f;
      }
      const int sL = L < li ? L : li;
      const int shift = D * (L - sL);
      int cy[3] = {0, 0, 0};
      cell_at<D>(J, sL, y, cy);
      int c[3] = {cy[0] + cell % 3 - 1, cy[1] + (cell / 3) % 3 - 1, cy[2] + (D > 2 ? cell / 9 - 1 : 0)};
      bool ok = c[0] >= 0 && c[0] < J.lev_n[sL][0] && c[1] >= 0 && c[1] < J.lev_n[sL][1];
      if (D > 2) ok = ok && c[2] >= 0 && c[2] < J.lev_n[sL][2];
      if (ok) {
        const int64_t tb = J.tbl_base + J.lev_off[L];
        const uint64_t m = morton<D>(c);
        lo = W.starts[tb + int64_t(m << shift)] + soff;
        hi2 = W.starts[tb + int64_t((m + 1) << shift)] + soff;
      }
    }
    auto cand = [&](int32_t q) {
      if (__float_as_int(W.rec[4 * int64_t(q) + 1].w) & kJumpBit) return false;
      const float4 qr = W.recb[q];
      float d2 = 0.f;
#pragma unroll
      for (int a = 0; a < D; a++) {
        const float tt = A[a] + sg[a] * (a == 0 ? qr.x : (a == 1 ? qr.y : qr.z));
        d2 += tt * tt;
      }
      const float r = fmaxf(rfi, qr.w * cutf);
      const float rr = r + (J.pf_tol + 1.e-6f * r);
      return d2 <= rr * rr;
    };
    int cnt = 0;
    for (int32_t q = lo; q < hi2; q++) cnt += cand(q);
    // exclusive prefix over the lanes (lane order = triple order)
    int pre = cnt;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int v = __shfl_up(pre, o, 64);
      if (lane >= o) pre += v;
    }
    const int sum = __shfl(pre, 63, 64);
    pre -= cnt;
    int w = total + pre;
    for (int32_t q = lo; q < hi2; q++) {
      if (!cand(q)) continue;
      if (w < kJumpList) lst[w] = int32_t(nb_pack(q, J.p0, code));
      w++;
    }
    total += sum;
  }
  if (lane == 0) {
    if (total > kJumpList) atomicOr(&W.red[j].jover, 1);
    W.jcnt[int64_t(j) * kJumpMax + t] = total < kJumpList ? total : kJumpList;
    W.nbo[slot] = t;
  }
}

// Jumpers that may pair with a tile: the tile's bounding box (hi parts of its live
// particles) against every mirror image of each jumper, with the largest pair radius
// c0.  Called by every thread of the workgroup; the survivors, in jset order, go to
// LDS.
struct JumpTile {
  int n;
  int32_t slot[kJumpMax];
  float box[6];
};

template <int D>
__device__ void jumper_prepare(const Work& W, const JobParams& J, int j, bool live, const float* xh, JumpTile& jt) {
  const int nj0 = W.red[j].njump;
  const int nj = nj0 < kJumpMax ? nj0 : kJumpMax;
  if (nj == 0) {   // uniform over the workgroup
    if (threadIdx.x == 0) jt.n = 0;
    __syncthreads();
    return;
  }
  __shared__ float red_lds[6][4];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
  for (int a = 0; a < D; a++) {
    float lo = live ? xh[a] : 3.0e38f, hi = live ? xh[a] : -3.0e38f;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      lo = fminf(lo, __shfl_xor(lo, o, 64));
      hi = fmaxf(hi, __shfl_xor(hi, o, 64));
    }
    if (l == 0) {
      red_lds[2 * a][w] = lo;
      red_lds[2 * a + 1][w] = hi;
    }
  }
  __syncthreads();
  if (threadIdx.x < 64) {
    float box[6];
#pragma unroll
    for (int a = 0; a < D; a++) {
      box[2 * a] = fminf(fminf(red_lds[2 * a][0], red_lds[2 * a][1]), fminf(red_lds[2 * a][2], red_lds[2 * a][3]));
      box[2 * a + 1] = fmaxf(fmaxf(red_lds[2 * a + 1][0], red_lds[2 * a + 1][1]),
                             fmaxf(red_lds[2 * a + 1][2], red_lds[2 * a + 1][3]));
    }
    bool keep = false;
    int32_t s = -1;
    if (threadIdx.x < nj) {
      s = W.jset[int64_t(j) * kJumpMax + threadIdx.x];
      const float4 q = W.rec[4 * int64_t(s)];
      const float qv[3] = {q.x, q.y, q.z};
      const float R = float(J.c0) + 8.f * J.pf_tol;
      float d2 = 0.f;
#pragma unroll
      for (int a = 0; a < D; a++) {
        // distance from the box to the nearest image of the jumper along this axis
        float g = fmaxf(0.f, fmaxf(box[2 * a] - qv[a], qv[a] - box[2 * a + 1]));
        if (J.sym[a]) {
          const float lo = 2.f * J.wl_h[a] - qv[a], hi = 2.f * J.wh_h[a] - qv[a];
          g = fminf(g, fmaxf(0.f, fmaxf(box[2 * a] - lo, lo - box[2 * a + 1])));
          g = fminf(g, fmaxf(0.f, fmaxf(box[2 * a] - hi, hi - box[2 * a + 1])));
        }
        d2 += g * g;
      }
      keep = d2 <= R * R;
    }
    const uint64_t m = __ballot(keep);
    if (keep) jt.slot[__popcll(m & ((1ull << threadIdx.x) - 1ull))] = s;
    if (threadIdx.x == 0) jt.n = __popcll(m);
  }
  __syncthreads();
}

// Every neighbour of one particle, in a fixed order, handed to fn(k, q, code, qh, ql,
// dr, d2), with k the list entry or -1: first the list entries (a non-jumper) or a
// cell search of the last build's tables (a jumper), skipping jumpers; then the
// jumpers near the tile, each with the particle's mirror options.  The neighbour
// predicate is the reference's (is_neighbor).
template <int D, class Fn>
__device__ __forceinline__ void for_each_neighbor(const Work& W, const PState& S, const JobParams& J, int64_t slot,
                                                  const JumpTile& jt, Fn fn, bool jumpers_only = false) {
  const float4 r0 = W.rec[4 * slot + 0], r1 = W.rec[4 * slot + 1];
  const float xhi[3] = {r0.x, r0.y, r0.z}, xli[3] = {r1.x, r1.y, r1.z};
  const double hd = S.h[slot];
  const float rfi = float(hd * J.cut_off);
  const bool ijump = __float_as_int(r1.w) & kJumpBit;
  auto visit_rec = [&](int k, int64_t q, uint32_t code, const float4& qh, const float4& ql) {
    float dr[3] = {0.f, 0.f, 0.f};
    rel_pos<D>(J, qh, ql, code, xhi, xli, dr);
    float d2 = 0.f;
#pragma unroll
    for (int a = 0; a < D; a++) d2 += dr[a] * dr[a];
    if (is_neighbor<D>(J, W, S, q, code, d2, rfi, qh.w, xhi, xli, hd)) fn(k, q, code, qh, ql, dr, d2);
  };
  auto visit = [&](int k, int64_t q, uint32_t code) { visit_rec(k, q, code, W.rec[4 * q + 0], W.rec[4 * q + 1]); };
  if (jumpers_only && !ijump) {
    // the caller walked the list itself
  } else if (!ijump) {
    const int cnt = W.nbc[slot];
    const int ni = cnt < W.K ? cnt : W.K;
    // inline entries in groups of four: one 16-byte load brings a group, its four
    // records are gathered together, then the next group's load is issued while this
    // group is used.  Jumpers come from the jumper scan.
    uint4 gn = make_uint4(0u, 0u, 0u, 0u);
    if (ni > 0) gn = nb_group(W, slot, 0);
    for (int g = 0; 4 * g < ni; g++) {
      const uint32_t ev[4] = {gn.x, gn.y, gn.z, gn.w};
      float4 qh[4], ql[4];
      // past the end of the list a lane loads its own record (always valid) and skips it
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int64_t q = 4 * g + u < ni ? nb_slot(ev[u], J.p0) : slot;
        qh[u] = W.rec[4 * q + 0];
        ql[u] = W.rec[4 * q + 1];
      }
      if (4 * g + 4 < ni) gn = nb_group(W, slot, g + 1);
#pragma unroll
      for (int u = 0; u < 4; u++) {
        if (4 * g + u >= ni || (__float_as_int(ql[u].w) & kJumpBit)) continue;
        visit_rec(4 * g + u, nb_slot(ev[u], J.p0), ev[u] >> kJBits, qh[u], ql[u]);
      }
    }
    for (int k = ni; k < cnt; k++) {   // the spilled tail
      const uint32_t e = uint32_t(W.spill[int64_t(W.nbo[slot]) + (k - W.K)]);
      const int64_t q = nb_slot(e, J.p0);
      const float4 qh = W.rec[4 * q + 0], ql = W.rec[4 * q + 1];
      if (__float_as_int(ql.w) & kJumpBit) continue;
      visit_rec(k, q, e >> kJBits, qh, ql);
    }
  } else {
    // a jumper walks the list k_jump_lists built for it this step (nbo holds its index)
    const int64_t t = int64_t(W.tile_job[slot / kTile]) * kJumpMax + W.nbo[slot];
    const int cnt = W.jcnt[t];
    const int32_t* lst = W.jl + t * kJumpList;
    for (int k = 0; k < cnt; k++) {
      const uint32_t e = uint32_t(lst[k]);
      visit(-1, nb_slot(e, J.p0), e >> kJBits);
    }
  }
  if (jt.n == 0) return;
  double xi[3] = {0., 0., 0.};
#pragma unroll
  for (int a = 0; a < D; a++) xi[a] = double(xhi[a]) + double(xli[a]);
  int opt[3][3], nopt[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    nopt[a] = 1;
    opt[a][0] = 0;
    if (a < D && J.sym[a]) {
      if (xi[a] - J.box_l[a] < J.c0) opt[a][nopt[a]++] = 1;
      if (J.box_r[a] - xi[a] < J.c0) opt[a][nopt[a]++] = 2;
    }
  }
  for (int t = 0; t < jt.n; t++) {
    const int64_t q = jt.slot[t];
    for (int o0 = 0; o0 < nopt[0]; o0++)
      for (int o1 = 0; o1 < nopt[1]; o1++)
        for (int o2 = 0; o2 < nopt[2]; o2++)
          visit(-1, q, uint32_t(opt[0][o0] | (opt[1][o1] << 2) | (opt[2][o2] << 4)));
  }
}

// neighbour sets for download: counts (pass 0) or nb_pack entries at offs
template <int D>
__global__ __launch_bounds__(kTile) void k_nb_out(Work W, PState S, int pass, const int64_t* offs, int32_t* cnt,
                                                  int32_t* out) {
  __shared__ JumpTile jt;
  const int j = W.tile_job[tile_id(W)];
  if (j < 0) return;
  const JobParams& J = W.jp[j];
  const int64_t slot = tile_id(W) * kTile + threadIdx.x;
  const bool live = slot - J.p0 < J.n;
  const float4 r0 = W.rec[4 * slot];
  const float xh[3] = {r0.x, r0.y, r0.z};
  jumper_prepare<D>(W, J, j, live, xh, jt);
  if (!live) return;
  int c = 0;
  const int64_t o = pass ? offs[slot - J.p0] : 0;
  for_each_neighbor<D>(W, S, J, slot, jt, [&](int, int64_t q, uint32_t code, const float4&, const float4&,
                                               const float*, float) {
    if (pass) out[o + c] = int32_t(nb_pack(q, J.p0, code));
    c++;
  });
  if (!pass) cnt[slot - J.p0] = c;
}

// ---------------------------------------------------------------------------
// Neighbour tiles of the pair passes: the halo and item build (list-build steps)

// the hash of a halo key into a table of HN slots (HN a power of two)
template <int HN>
__device__ __forceinline__ uint32_t halo_hash(uint32_t e) {
  static_assert(HN >= 1024 && (HN & (HN - 1)) == 0, "halo_hash yields log2(HN) bits");
  return (e * 0x9E3779B1u) >> (32 - __builtin_ctz(unsigned(HN)));
}
// the key subset of pass `pass` when a tile's halo takes P passes (P a power of two)
__device__ __forceinline__ uint32_t halo_part(uint32_t e, int P) { return ((e * 0x85EBCA6Bu) >> 7) & uint32_t(P - 1); }

// exclusive prefix sum over the 256 threads of a workgroup; total = the sum
__device__ __forceinline__ int block_scan_excl(int v, int* lds4, int& total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    x += lane >= o ? y : 0;
  }
  if (lane == 63) lds4[w] = x;
  __syncthreads();
  int base = 0, tot = 0;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int s = lds4[q];
    base += q < w ? s : 0;
    tot += s;
  }
  __syncthreads();
  total = tot;
  return base + x - v;
}

// the candidate entries of a slot in list order: fn(k, entry)
template <class Fn>
__device__ __forceinline__ void for_entries(const Work& W, int64_t slot, int cnt, int32_t off, Fn fn) {
  const int ni = cnt < W.K ? cnt : W.K;
  // the next group's load is issued before this group is used
  uint4 vn = ni > 0 ? nb_group(W, slot, 0) : make_uint4(0u, 0u, 0u, 0u);
  for (int g = 0; 4 * g < ni; g++) {
    const uint4 v = vn;
    if (4 * g + 4 < ni) vn = nb_group(W, slot, g + 1);
    const uint32_t ev[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int u = 0; u < 4; u++)
      if (4 * g + u < ni) fn(4 * g + u, ev[u]);
  }
  for (int k = ni; k < cnt; k++) fn(k, uint32_t(W.spill[int64_t(off) + (k - W.K)]));
}

// the same walk with each entry's near bit: k_nb's nbnear words for the inline entries,
// spill_near(entry) for the spilled ones; fn(k, entry, near)
template <class Fn, class Near>
__device__ __forceinline__ void for_entries_near(const Work& W, int64_t slot, int cnt, int32_t off, Fn fn,
                                                 Near spill_near) {
  const int ni = cnt < W.K ? cnt : W.K;
  uint4 vn = ni > 0 ? nb_group(W, slot, 0) : make_uint4(0u, 0u, 0u, 0u);
  uint32_t nw = 0u;
  for (int g = 0; 4 * g < ni; g++) {
    const uint4 v = vn;
    if (4 * g + 4 < ni) vn = nb_group(W, slot, g + 1);
    if ((g & 7) == 0) nw = W.nbnear[int64_t(g >> 3) * W.cap_total + slot];
    const uint32_t ev[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int u = 0; u < 4; u++)
      if (4 * g + u < ni) fn(4 * g + u, ev[u], ((nw >> ((4 * g + u) & 31)) & 1u) != 0u);
  }
  for (int k = ni; k < cnt; k++) {
    const uint32_t e = uint32_t(W.spill[int64_t(off) + (k - W.K)]);
    fn(k, e, spill_near(e));
  }
}

// One workgroup per tile of a rebuilding job.  The tile's halo: its own slots (halo
// index = slot - tile start), then every other (slot, mirror code) its candidate lists
// name, each once, deduplicated in an LDS hash table; the candidate lists become items
// (halo indices): the entries within the cutoff at the build positions first, in list
// order, then the rest in reverse list order.  A halo too large for one table is built
// in P passes, each taking the keys of one hash class.  The halo order is that of the
// table (it may differ between runs); a particle's item order depends only on its list
// and the build positions, and a halo index only says where the record sits, so every
// pair sum is formed in the same order each run.
template <int D, int HN>
__global__ __launch_bounds__(kTile) void k_halo(Work W, PState S) {
  // The long-list variant (the larger table) orders each pass's halo entries by how many
  // of the tile's list entries name them, most-named first, so that the records the
  // pair passes stage in LDS (the first ones) serve most item reads of a halo too large
  // to stage whole; hv holds each key's count while the pass inserts.
  constexpr bool kOrd = HN > kHashN;
  using HV = typename std::conditional<kOrd, uint32_t, uint16_t>::type;
  __shared__ uint32_t hk[HN];
  __shared__ HV hv[HN];
  __shared__ int s_bin[kOrd ? kTile : 1];
  __shared__ int s_scan[4];
  __shared__ int s_len[kTile];
  __shared__ int s_cnt, s_over;
  __shared__ unsigned long long s_ibase;
  if (!rebuild_any(W)) return;
  const int64_t tile = tile_id(W);
  const int j = W.tile_job[tile];
  if (j < 0 || !W.red[j].rebuild) return;
  const JobParams& J = W.jp[j];
  // the job's first slot in a register: JobParams sits in global memory, and the
  // stores below would otherwise make every entry read it again
  const int64_t p0 = J.p0;
  const int64_t ts = tile * kTile, slot = ts + threadIdx.x;
  const bool live = slot - p0 < J.n;
  const int cnt = live ? W.nbc[slot] : 0;
  const int32_t off = cnt > W.K ? W.nbo[slot] : 0;
  // The pair passes' thread order: rank by (list length, slot), so that a wavefront's
  // threads walk lists of about the same length.  Items are stored by wavefront, eight
  // to a 16-byte group, lane-interleaved: group g of the particle at wave w, lane l at
  // woff[w] + (64 g + l) 8, where wave w's part holds as many groups as its longest
  // list needs; a wavefront's loads of one group are 1 KB contiguous.
  s_len[threadIdx.x] = cnt;
  __syncthreads();
  int rank = 0;
  for (int q = 0; q < kTile; q++) {
    const int cq = s_len[q];
    rank += (cq < cnt || (cq == cnt && q < int(threadIdx.x))) ? 1 : 0;
  }
  __syncthreads();
  // rank r goes to wavefront r / 64, lane r % 64: the lists of a wavefront are a
  // quarter of the sorted lengths, so its padding to its longest list stays small
  s_len[rank] = cnt;   // sorted: wave w's longest list has rank 64 w + 63
  const int rw = rank >> 6, rl = rank & 63;
  W.lane[ts + rank] = uint8_t(threadIdx.x);
  __syncthreads();
  int M = 0, woff = 0;
  for (int w = 0; w < kTile / 64; w++) {
    const int g = (s_len[64 * w + 63] + 7) >> 3;
    if (w < rw) woff += g * 512;
    M += g * 512;
  }
  const int excl = woff + rl * 8;
  W.ioff[slot] = excl;
  // A tile that does not fit is marked invalid (hcnt = 0: the pair passes read none of
  // its items) and its job flagged.
  auto fail = [&](int bit) {
    if (threadIdx.x == 0) {
      atomicOr(&W.red[j].overflow, bit);
      W.hcnt[tile] = 0;
    }
  };
  // the tile's M items from its job's part of the pool (where a tile lands depends on
  // the order of the atomics; what it holds does not)
  if (threadIdx.x == 0) s_ibase = atomicAdd(&W.red[j].items_used, (unsigned long long)M);
  __syncthreads();
  const unsigned long long ib = s_ibase;   // the same in every thread, like M
  if (ib + (unsigned long long)M > (unsigned long long)(J.cap * W.ips)) {
    fail(kOverItems);
    return;
  }
  const int64_t tb = p0 * W.ips + int64_t(ib);
  if (threadIdx.x == 0) W.tbase[tile] = tb;
  uint16_t* it0 = W.items + tb + excl;
  auto item_at = [&](int k) -> uint16_t& { return it0[(k >> 3) * 512 + (k & 7)]; };
  uint32_t* hal = W.halo + tile * W.ht;
  const int cap = int(W.ht < int64_t(kMaxHalo) ? W.ht : int64_t(kMaxHalo)) - kTile;
  auto own = [&](uint32_t e) {
    const int64_t q = nb_slot(e, p0);
    return (e >> kJBits) == 0u && q >= ts && q < ts + kTile;
  };
  // k_nb's near test for the spilled entries (its near bits cover the inline ones):
  // the same FP32 operations on the same records, so a list gets the same near flags
  // whatever part of it is spilled
  float xb[3] = {0.f, 0.f, 0.f}, rfi = 0.f;
  {
    const float4 rb = W.rec[4 * slot + 0];
    xb[0] = rb.x; xb[1] = rb.y; xb[2] = rb.z;
    if (live) rfi = float(S.h[slot] * J.cut_v);
  }
  const float cutv = float(J.cut_v), nfac = float(J.cut_off / J.cut_v);
  auto spill_near = [&](uint32_t e) {
    float A[3] = {0.f, 0.f, 0.f}, sg[3] = {0.f, 0.f, 0.f};
    image_affine<D>(J, e >> kJBits, xb, A, sg);
    const float4 qr = W.recb[nb_slot(e, p0)];
    const float qv[3] = {qr.x, qr.y, qr.z};
    float d2 = 0.f;
#pragma unroll
    for (int a = 0; a < D; a++) {
      const float t = A[a] + sg[a] * qv[a];
      d2 += t * t;
    }
    const float rn = fmaxf(rfi, qr.w * cutv) * nfac;
    return d2 <= rn * rn;
  };
  int P = 1, base = 0;
  for (int pass = 0; pass < P;) {
    for (int s = threadIdx.x; s < HN; s += kTile) {
      hk[s] = 0xFFFFFFFFu;
      if constexpr (kOrd) hv[s] = 0u;
    }
    if constexpr (kOrd) s_bin[threadIdx.x] = 0;
    if (threadIdx.x == 0) {
      s_cnt = 0;
      s_over = 0;
    }
    __syncthreads();
    // Insert this pass's keys.  Past 3/4 of the table a thread stops after its next new
    // key, so at most 256 more keys arrive: the table never fills and every probe ends.
    bool full = false;
    for_entries(W, slot, cnt, off, [&](int, uint32_t e) {
      if (full || own(e) || halo_part(e, P) != uint32_t(pass)) return;
      uint32_t s = halo_hash<HN>(e);
      while (true) {
        const uint32_t old = atomicCAS(&hk[s], 0xFFFFFFFFu, e);
        if (old == 0xFFFFFFFFu) {
          full = atomicAdd(&s_cnt, 1) >= HN / 4 * 3;
          break;
        }
        if (old == e) break;
        s = (s + 1u) & uint32_t(HN - 1);
      }
      if constexpr (kOrd) atomicAdd(&hv[s], 1u);
    });
    if (full) s_over = 1;
    __syncthreads();
    if (s_over) {   // the same in every thread: start again with twice the passes
      __syncthreads();
      if (P >= 64) {
        fail(kOverHalo);
        return;
      }
      P *= 2;
      pass = 0;
      base = 0;
      continue;
    }
    // halo indices for the occupied table slots (thread t takes slots t, t + 256, ...)
    constexpr int kPer = HN / kTile;
    int U = 0;
    if constexpr (!kOrd) {   // in table order
      int c = 0;
#pragma unroll 8
      for (int u = 0; u < kPer; u++) c += hk[u * kTile + threadIdx.x] != 0xFFFFFFFFu ? 1 : 0;
      int idx = base + block_scan_excl(c, s_scan, U);
      if (base + U > cap) {
        fail(kOverHalo);
        return;
      }
      for (int u = 0; u < kPer; u++) {
        const int s = u * kTile + threadIdx.x;
        const uint32_t key = hk[s];
        if (key != 0xFFFFFFFFu) {
          hv[s] = HV(kTile + idx);
          hal[idx] = key;
          idx++;
        }
      }
    } else {   // by count, most-named first: a counting sort over the bins 256 - count
      auto bin_of = [&](uint32_t n) { return kTile - int(n < uint32_t(kTile) ? n : uint32_t(kTile)); };
      for (int u = 0; u < kPer; u++) {
        const int s = u * kTile + threadIdx.x;
        if (hk[s] != 0xFFFFFFFFu) atomicAdd(&s_bin[bin_of(hv[s])], 1);
      }
      __syncthreads();
      const int b0 = block_scan_excl(s_bin[threadIdx.x], s_scan, U);   // thread t: bin t
      if (base + U > cap) {
        fail(kOverHalo);
        return;
      }
      s_bin[threadIdx.x] = base + b0;   // the bin's next index
      __syncthreads();
      for (int u = 0; u < kPer; u++) {
        const int s = u * kTile + threadIdx.x;
        const uint32_t key = hk[s];
        if (key != 0xFFFFFFFFu) {
          const int idx = atomicAdd(&s_bin[bin_of(hv[s])], 1);
          hv[s] = HV(kTile + idx);
          hal[idx] = key;
        }
      }
    }
    __syncthreads();
    // Items near-first: the entries k_nb found within the cutoff radius at the build
    // positions from the front, the others from the back.  Every pass places every
    // entry, so each entry gets the same position whatever pass writes it.
    int kn = 0, kf = 0;
    for_entries_near(W, slot, cnt, off, [&](int, uint32_t e, bool near) {
      const int k = near ? kn++ : cnt - 1 - kf++;
      if (own(e)) {
        if (pass == 0) item_at(k) = uint16_t(nb_slot(e, p0) - ts);
        return;
      }
      if (halo_part(e, P) != uint32_t(pass)) return;
      uint32_t s = halo_hash<HN>(e);
      while (hk[s] != e) s = (s + 1u) & uint32_t(HN - 1);
      item_at(k) = hv[s];
    }, spill_near);
    base += U;
    pass++;
    __syncthreads();
  }
  if (threadIdx.x == 0) W.hcnt[tile] = kTile + base;
}

// slot and mirror code of halo entry e of a tile
__device__ __forceinline__ void halo_entry(const Work& W, int64_t tile, int64_t ts, int64_t p0, uint32_t e,
                                           int64_t& q, uint32_t& code) {
  if (e < uint32_t(kTile)) {
    q = ts + e;
    code = 0u;
    return;
  }
  const uint32_t v = W.halo[tile * W.ht + (e - kTile)];
  q = nb_slot(v, p0);
  code = v >> kJBits;
}

// A halo record as staged in LDS: the position of slot q's image as an unevaluated
// pair of floats, (hi, h) and (lo, type word) -- the slot's own record for a plain
// entry, the image 2w - x formed in FP64 and split again for a mirror image.  The type
// word carries the type bits, the jumper bit and kHMir for an image.  Records are in
// the box's frame, so two particles at the same point get the same relative positions
// to every neighbour whatever tile they sit in.
template <int D>
__device__ __forceinline__ void image_rec(const JobParams& J, uint32_t code, float4& rh, float4& rl) {
  const double x[3] = {double(rh.x) + double(rl.x), double(rh.y) + double(rl.y), double(rh.z) + double(rl.z)};
  double g[3] = {x[0], x[1], x[2]};
  image_of<D>(J, code, x, g);
  const float gh[3] = {float(g[0]), float(g[1]), float(g[2])};
  rh = make_float4(gh[0], gh[1], gh[2], rh.w);
  rl = make_float4(float(g[0] - double(gh[0])), float(g[1] - double(gh[1])), float(g[2] - double(gh[2])),
                   __int_as_float(__float_as_int(rl.w) | int(kHMir)));
}

template <int D>
__device__ __forceinline__ void halo_rec(const Work& W, const JobParams& J, int64_t q, uint32_t code, float4& rh,
                                         float4& rl) {
  rh = W.rec[4 * q + 0];
  rl = W.rec[4 * q + 1];
  if (code) image_rec<D>(J, code, rh, rl);
}

// The pair passes' compact form of a record's second half: the lo parts rounded to
// bfloat16 (8 significant bits, round to nearest even) and a 16-bit type word (type,
// kQJump, kQMir), 8 bytes.  A lo part is at most half an ulp of its hi part, so the
// rounding moves a position by at most 2^-9 of that: 2^-33 |x| (the pair passes widen
// their cutoff band by it).  A particle's own lo parts are rounded the same way, so
// two particles at one point still sit at one point.
constexpr uint32_t kQJump = 1u << 14, kQMir = 1u << 15, kQType = kQJump - 1u;
__device__ __forceinline__ uint32_t bf16_rn(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;
}
__device__ __forceinline__ uint2 pack_lo(const float4& rl) {
  const int tw = __float_as_int(rl.w);
  const uint32_t t16 = (uint32_t(tw) & kQType) | ((tw & kJumpBit) ? kQJump : 0u) | ((tw & int(kHMir)) ? kQMir : 0u);
  return make_uint2(bf16_rn(rl.x) | (bf16_rn(rl.y) << 16), bf16_rn(rl.z) | (t16 << 16));
}
__device__ __forceinline__ void unpack_lo(const uint2& v, float* lo, int& tw) {
  lo[0] = __uint_as_float(v.x << 16);
  lo[1] = __uint_as_float(v.x & 0xFFFF0000u);
  lo[2] = __uint_as_float(v.y << 16);
  tw = int(v.y >> 16);
}

// Stages the records of a tile's halo entries [0, HS) in LDS (hh: hi parts and h; hq:
// the compact lo parts and type word; hv: the third record word, when V),
// HS <= NU kTile.  Each thread takes entries threadIdx.x + u kTile: all its halo words
// are loaded first, then all its records, so the loads of one thread are in flight
// together.
template <int D, int NU, bool V>
__device__ __forceinline__ void stage_halo(const Work& W, const JobParams& J, int64_t tile, int64_t ts, int HS,
                                           float4* hh, uint2* hq, float4* hv) {
  int64_t q[NU];
  uint32_t code[NU];
#pragma unroll
  for (int u = 0; u < NU; u++) {
    const int e = threadIdx.x + u * kTile;
    q[u] = ts;   // a valid slot for the entries past HS (their records are never stored)
    code[u] = 0u;
    if (e < HS) halo_entry(W, tile, ts, J.p0, uint32_t(e), q[u], code[u]);
  }
  float4 rh[NU], rl[NU];
#pragma unroll
  for (int u = 0; u < NU; u++) {
    const int e = threadIdx.x + u * kTile;
    if (e < HS) {
      rh[u] = W.rec[4 * q[u] + 0];
      rl[u] = W.rec[4 * q[u] + 1];
      if (V) hv[e] = W.rec[4 * q[u] + 2];
    }
  }
#pragma unroll
  for (int u = 0; u < NU; u++) {
    const int e = threadIdx.x + u * kTile;
    if (e < HS) {
      if (code[u]) image_rec<D>(J, code[u], rh[u], rl[u]);
      hh[e] = rh[u];
      hq[e] = pack_lo(rl[u]);
    }
  }
}

// The reference's predicate in FP64 on the stored positions (for pairs whose FP32
// distance falls in the error band around the cutoff)
template <int D>
__device__ __forceinline__ bool exact_pred(const Work& W, const JobParams& J, const PState& S, int64_t si, int64_t q,
                                           uint32_t code) {
  const float4 a0 = W.rec[4 * si + 0], a1 = W.rec[4 * si + 1];
  const float4 b0 = W.rec[4 * q + 0], b1 = W.rec[4 * q + 1];
  const double xi[3] = {double(a0.x) + double(a1.x), double(a0.y) + double(a1.y), double(a0.z) + double(a1.z)};
  const double xq[3] = {double(b0.x) + double(b1.x), double(b0.y) + double(b1.y), double(b0.z) + double(b1.z)};
  double g[3];
  image_of<D>(J, code, xq, g);
  double dist = 0.0;
#pragma unroll
  for (int a = 0; a < D; a++) {
    const double dr = xi[a] - g[a];
    dist += dr * dr;
  }
  dist = sqrt(dist);
  return dist <= S.h[si] * J.cut_off + 1.e-10 || dist <= S.h[q] * J.cut_off + 1.e-10;
}

// x_i - x_q from the hi / lo pairs (the image already applied to q's record)
template <int D>
__device__ __forceinline__ void rel_hl(const float* xhi, const float* xli, const float4& qh, const float* lo,
                                       float* dr) {
  dr[0] = (xhi[0] - qh.x) + (xli[0] - lo[0]);
  if (D > 1) dr[1] = (xhi[1] - qh.y) + (xli[1] - lo[1]);
  if (D > 2) dr[2] = (xhi[2] - qh.z) + (xli[2] - lo[2]);
}

// ---------------------------------------------------------------------------
// P3: kernel summation + stage-1 force + stage-1 dt

// the particle a pair pass is summing for
struct Own {
  float xh[3], xl[3];
  float h, ih, Ck, kd3, vol2i, vol, eta, rf;
  float kdx, cth;   // the job's cut_off / cell_ratio and cos(angle_thres), in registers
  float bq;         // the cutoff band's absolute part (FP32 rounding and the bf16 lo parts)
  float n[3], v[3];
  bool surf, jump;
};
struct Acc1 {
  float sum, g[3];
  int nc, ntr;
};

template <int D>
__device__ __forceinline__ void own_common(const JobParams& J, const float4& rh, const float4& rl, const float* n, Own& o) {
  const int tw = __float_as_int(rl.w);
  const int type = tw & ~(kJumpBit | int(kHMir));
  o.xh[0] = rh.x; o.xh[1] = rh.y; o.xh[2] = rh.z;
  // the lo parts as the compact records hold them (bf16)
  o.xl[0] = __uint_as_float(bf16_rn(rl.x) << 16);
  o.xl[1] = __uint_as_float(bf16_rn(rl.y) << 16);
  o.xl[2] = __uint_as_float(bf16_rn(rl.z) << 16);
  o.h = rh.w;
  o.surf = type == SMG_SURFACE;
  o.jump = (tw & kJumpBit) != 0;
#pragma unroll
  for (int a = 0; a < 3; a++) o.n[a] = n[a];
  const float dxi = o.h * J.kdx;
  o.vol = ipow<D>(dxi);
  o.vol2i = o.vol * o.vol;
  o.ih = 1.f / o.h;
  const float Cw = kC<D>() * ipow<D>(o.ih), Cs = kC<D - 1>() * ipow<D - 1>(o.ih);
  o.Ck = o.surf ? Cs : Cw;
  o.kd3 = 3.f * (o.surf ? kC<D - 1>() : kC<D>());
  o.rf = o.h * float(J.cut_off);
  o.kdx = J.kdx;
  o.cth = J.cos_thres;
  // every coordinate of a record (mirror images included) lies within c0 of the box,
  // and rounding a lo part moves it by at most 2^-33 |x|: 2^-30 max |x| bounds what
  // that adds to a distance, with room to spare
  double xm = 0.;
#pragma unroll
  for (int a = 0; a < D; a++) xm = fmax(xm, fmax(fabs(J.box_l[a] - J.c0), fabs(J.box_r[a] + J.c0)));
  o.bq = 2.e-10f + 9.32e-10f * float(xm);
}

// One stage-1 pair (Get_kernel_summation and the stage-1 Accumulate_particle_force_real
// / _surf, particle_mesh_generation.cpp:690-765, 883-1118) of a verified neighbour:
// dr0 = x_i - x_q, d2 = |dr0|^2.  A surface particle uses the (D-1)-dimensional kernel
// on the tangential part of dr (raw normal, Q3) and only surface neighbours with
// aligned normals.  TS: the particle's tile may hold surface particles (the general
// form); without TS no particle of the tile is a surface one.  In the general form a
// REAL particle has n = 0, so the projection leaves dr and d2 unchanged bit for bit and
// both forms give the same sums.  With s = r / h_i and s2 = r / h_m (h_m the pair's
// mean h), W = Ck w(s) with w = 6 - 6s + s^3 (s < 1), (2 - s)^3 (s < 2), and
// (dW/dr) / r = 3 kC h_m^-(DK+1) dq(s2) / r with dq = s2^2 - 2 (s2 < 1), -(2 - s2)^2:
// the cubic spline and its derivative, factored.  1 / r is the hardware reciprocal
// square root of r^2 (of 1e-30 at r = 0, where dr = 0 makes the term vanish) and
// r = r^2 / r.  Returns whether the pair enters the stage-2 pass (a non-self
// neighbour the particle uses).
template <int D, bool TS>
__device__ __forceinline__ bool pair1(const Work& W, const JobParams& J, const Own& o, const float* dr0, float d2,
                                      float hq, int tq, bool self, int64_t q, Acc1& acc, bool valid = true) {
  // valid = false: the pair is not a neighbour (the sums are left as they are)
  const bool osurf = TS && o.surf;
  acc.ntr += valid ? 1 : 0;
  bool use = valid && !(osurf && tq == SMG_REAL);
  if (TS && use && osurf && tq == SMG_SURFACE) {
    const float4 qn = W.rec[4 * q + 3];
    const float nq[3] = {qn.x, qn.y, qn.z};
    float ca = 0.f;
#pragma unroll
    for (int a = 0; a < D; a++) ca += o.n[a] * nq[a];
    use = fabsf(ca) > o.cth;
  }
  float dr[3] = {dr0[0], dr0[1], dr0[2]};
  float rt2 = d2;
  if (TS) {
    float dn = 0.f;
#pragma unroll
    for (int a = 0; a < D; a++) dn += dr0[a] * o.n[a];
    rt2 = 0.f;
#pragma unroll
    for (int a = 0; a < D; a++) {
      dr[a] = dr0[a] - o.n[a] * dn;
      rt2 += dr[a] * dr[a];
    }
  }
  const float irt = frsq(fmaxf(rt2, 1.e-30f));
  const float rt = rt2 * irt;
  const bool counts = use && (!osurf || tq == SMG_SURFACE);   // the kernel summation
  const float s = rt * o.ih;
  const float t = fmaxf(2.f - s, 0.f);
  const float w = s < 1.f ? (6.0f - 6.0f * s) + s * s * s : t * t * t;
  acc.sum = counts ? acc.sum + o.Ck * w : acc.sum;
  acc.nc += counts ? 1 : 0;
  const float dxq = hq * o.kdx;
  const float vol2q = ipow<D>(dxq) * ipow<D>(dxq);
  const float ihm = frcp(0.5f * (o.h + hq));
  const float C3 = osurf ? o.kd3 * ipow<D>(ihm) : o.kd3 * ipow<D + 1>(ihm);
  const float s2 = rt * ihm;
  const float t2 = fmaxf(2.f - s2, 0.f);
  const float dq = s2 < 1.f ? s2 * s2 - 2.f : -(t2 * t2);
  const float aw = C3 * dq * irt * (o.vol2i + vol2q);
  const bool f = use && !self;
#pragma unroll
  for (int a = 0; a < D; a++) acc.g[a] = f ? acc.g[a] + dr[a] * aw : acc.g[a];
  return f;
}

// One workgroup per tile; thread t sums for the tile's particle of rank t in the order
// of list lengths (so each wavefront walks lists of about the same length), over its
// items in item order, then the pairs with this step's jumpers.  Records come from the
// tile's halo, staged in LDS (the first kHL1 entries; the rest read from the records).
template <int D>
__global__ __launch_bounds__(kTile) __attribute__((amdgpu_waves_per_eu(4))) void k_force1(Work W, PState S, int par) {
  __shared__ float4 hh[kHL1];
  __shared__ uint2 hq[kHL1];
  __shared__ double lds_d[4];
  __shared__ int lds_i[4];
  __shared__ int s_surf;
  __shared__ JumpTile jt;
  const int64_t tile = tile_id(W);
  const int j = W.tile_job[tile];
  if (j < 0) return;
  const JobParams& J = W.jp[j];
  const int64_t ts = tile * kTile;
  // the particle of this thread
  const int p = W.lane[ts + threadIdx.x];
  const int64_t slot = ts + p;
  const bool live = slot - J.p0 < J.n;
  const float4 r0 = W.rec[4 * slot + 0], r1 = W.rec[4 * slot + 1];
  const float xhi[3] = {r0.x, r0.y, r0.z};
  if (threadIdx.x == 0) s_surf = 0;
  jumper_prepare<D>(W, J, j, live, xhi, jt);   // ends with a barrier
  const int H = W.hcnt[tile];
  const int HS = H < kHL1 ? H : kHL1;
  stage_halo<D, kHL1 / kTile, false>(W, J, tile, ts, HS, hh, hq, nullptr);
  if (live && (__float_as_int(r1.w) & ~kJumpBit) == SMG_SURFACE) s_surf = 1;
  // this lane's item count (hcnt = 0: the tile's build failed; a jumper's list is
  // stale) and the wavefront's largest
  const int cnt = (live && !(__float_as_int(r1.w) & kJumpBit) && H > 0) ? W.nbc[slot] : 0;
  const int kmax = __builtin_amdgcn_readfirstlane(wave_reduce(cnt, OpMaxI()));
  __syncthreads();
  const bool tsurf = s_surf != 0;
  double dt = 1.0e300;
  int ntr = 0;
  if (live) {
    float ni[3] = {0.f, 0.f, 0.f};
    if ((__float_as_int(r1.w) & ~kJumpBit) == SMG_SURFACE) {
#pragma unroll
      for (int c = 0; c < D; c++) ni[c] = S.nrm[c][slot];
    }
    Own o;
    own_common<D>(J, r0, r1, ni, o);
    const float nub = W.nu_b[slot];   // stored with a1 below
    Acc1 acc = {};
    const float cutf = float(J.cut_off);
    if (!o.jump) {   // a jumper's list is stale: it walks its own list below
      int c2 = 0;
      if (kmax > 0) {
        const int64_t i0 = W.tbase[tile] + W.ioff[slot];
        // items eight to a 16-byte group, groups 64 apart (the wavefront's lanes between)
        const uint4* it = reinterpret_cast<const uint4*>(W.items + i0);
        uint16_t* st2 = W.items2 + i0;   // the same group layout
        // One item of the walk: the record (qh, ql) of halo entry e; q, code: its slot
        // and mirror code if already known (q < 0: looked up when a rare path needs
        // them).  An item past the lane's list (valid = false) is its own record and
        // adds nothing.  A kept pair (not self, not a mirror image) goes to the lane's
        // stage-2 items.
        auto body = [&](auto ts_, uint32_t e, const float4& qh, const uint2& qc, int64_t q, uint32_t code,
                        bool valid) __attribute__((always_inline)) {
          constexpr bool TS = decltype(ts_)::value;
          float lo[3];
          int tw;
          unpack_lo(qc, lo, tw);
          float dr[3] = {0.f, 0.f, 0.f};
          rel_hl<D>(o.xh, o.xl, qh, lo, dr);
          float d2 = 0.f;
#pragma unroll
          for (int c = 0; c < D; c++) d2 += dr[c] * dr[c];
          // The FP32 distance of the compensated relative position is within 3 2^-24 d
          // (+ 1e-13, + the bf16 lo parts' share) of the FP64 one and r within 2^-23 r;
          // the FP64 predicate decides inside about 1e-6 r + bq of the cutoff (its own
          // +1e-10 included).
          const float r = fmaxf(o.rf, qh.w * cutf);
          const float rlo = r * (1.f - 1.e-6f) - o.bq;
          bool nb = d2 < rlo * rlo;
          if (valid && !nb) {
            const float rhi = r * (1.f + 1.e-6f) + o.bq;
            if (d2 <= rhi * rhi) {
              if (q < 0) halo_entry(W, tile, ts, J.p0, e, q, code);
              nb = exact_pred<D>(W, J, S, slot, q, code);
            }
          }
          nb = valid && nb && !(tw & int(kQJump));   // a jumper: paired through the jumper scan
          const int tq = tw & int(kQType);
          if (TS && nb && o.surf && tq == SMG_SURFACE && q < 0) halo_entry(W, tile, ts, J.p0, e, q, code);
          // An item that is no lane's neighbour adds nothing (pair1 with valid = false
          // leaves the sums as they are): the wavefront skips the pair arithmetic.  Items
          // are near-first, so past the near parts of its lists most items are skipped.
          if (__ballot(nb) != 0ull) {
            const bool f = pair1<D, TS>(W, J, o, dr, d2, qh.w, tq, e == uint32_t(p), q, acc, nb);
            if (f && !(tw & int(kQMir))) {
              st2[(c2 >> 3) * 512 + (c2 & 7)] = uint16_t(e);
              c2++;
            }
          }
        };
        // The walk runs to the wavefront's longest list, so its loop and group loads are
        // the same in every lane; each group of eight items is unrolled.
        auto walk = [&](auto ts_, auto fits_) __attribute__((always_inline)) {
          constexpr bool kFits = decltype(fits_)::value;
          const int ngw = (kmax + 7) >> 3;
          uint4 gn = it[0];
          for (int g = 0; g < ngw; g++) {
            uint4 gc = gn;
            // the group is taken into registers here, so the wait for its load stays
            // out of the items
            asm volatile("" : "+v"(gc.x), "+v"(gc.y), "+v"(gc.z), "+v"(gc.w));
            gn = it[(g + 1 < ngw ? g + 1 : g) * 64];
            uint32_t w0 = gc.x, w1 = gc.y, w2 = gc.z, w3 = gc.w;
#pragma unroll 1
            for (int v = 0; v < 4; v++) {   // two items from one word
#pragma unroll
              for (int u = 0; u < 2; u++) {
                const bool valid = 8 * g + 2 * v + u < cnt;
                const uint32_t e = valid ? (u ? w0 >> 16 : w0 & 0xFFFFu) : uint32_t(p);
                float4 qh;
                uint2 qc;
                int64_t q = -1;
                uint32_t code = 0u;
                if (kFits || e < uint32_t(HS)) {
                  qh = hh[e];
                  qc = hq[e];
                } else {   // past the staged records (a tile whose halo exceeds the LDS)
                  halo_entry(W, tile, ts, J.p0, e, q, code);
                  float4 rl;
                  halo_rec<D>(W, J, q, code, qh, rl);
                  qc = pack_lo(rl);
                }
                body(ts_, e, qh, qc, q, code, valid);
              }
              w0 = w1;
              w1 = w2;
              w2 = w3;
            }
          }
        };
        if (H <= kHL1) {
          if (tsurf) walk(std::true_type{}, std::true_type{});
          else walk(std::false_type{}, std::true_type{});
        } else {
          walk(std::true_type{}, std::false_type{});
        }
      }
      W.nbc2[slot] = c2;
    } else {
      W.nbc2[slot] = 0;
    }
    // pairs with this step's jumpers (and, for a jumper, its own list), after the items
    for_each_neighbor<D>(W, S, J, slot, jt, [&](int, int64_t q, uint32_t code, const float4& qh, const float4& ql,
                                                 const float* dr0, float d2) {
      pair1<D, true>(W, J, o, dr0, d2, qh.w, __float_as_int(ql.w) & ~kJumpBit, q == slot && code == 0u, q, acc);
    }, true);
    float sum = acc.sum;
    if (acc.nc > 1) sum = 1.f / (sum + 1.e-15f);
    W.sum[slot] = sum;
    W.ntrue[slot] = acc.ntr;
    ntr = acc.ntr;
    float F[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int c = 0; c < D; c++) F[c] = -sum * acc.g[c];
    if (o.surf) {   // Map_surface_particle_force with the raw normal
      float nf = 0.f;
#pragma unroll
      for (int c = 0; c < D; c++) nf += F[c] * ni[c];
#pragma unroll
      for (int c = 0; c < D; c++) F[c] = F[c] - ni[c] * nf;
    }
    double A = 0.;
#pragma unroll
    for (int c = 0; c < D; c++) {
      W.a1[c][slot] = F[c];
      A += double(F[c]) * double(F[c]);
    }
    W.rec[4 * slot + 2] = make_float4(F[0], F[1], D > 2 ? F[2] : 0.f, nub);
    A = sqrt(A);
    // Set_timestep(1): v was zeroed by Reset_particle_force(1), so U = 0
    const double hd = S.h[slot];
    const double t1 = 0.25 * sqrt(hd / (A + 1.e-15));
    const double t2 = hd / 40. / (0. + 1.e-15);
    dt = amin(t1, t2);
  }
  dt = block_reduce(dt, OpMin(), lds_d);
  const int mx = block_reduce(ntr, OpMaxI(), lds_i);
  if (threadIdx.x == 0) {
    JobAcc& A = acc_mine(W, j);
    if (dt < 1.0e300) atomicMin(&A.dt1[par], dbits(dt));
    if (mx > 0) atomicMax(&A.maxnb, mx);
  }
}

// ---------------------------------------------------------------------------
// P4: stage-2 (viscous) force, stage-2 dt and nu

// One stage-2 pair (Accumulate_particle_force_real / _surf, stage 2,
// particle_mesh_generation.cpp:995-1118): eta_ij the harmonic mean of nu rho, a
// non-REAL neighbour of a REAL particle moving at -v_i (Q8); dr0 = x_i - x_q, rr = |dr0|^2.
template <int D, bool TS>
__device__ __forceinline__ void pair2(const JobParams& J, const Own& o, float hdt, const float* dr0, float rr, float hq,
                                      int tq, const float4& qv, bool use, float* F) {
  const bool osurf = TS && o.surf;
  const float dxq = hq * o.kdx;
  const float volq = ipow<D>(dxq);
  const float rhoq = frcp(volq);
  const float etaq = qv.w * rhoq;
  // eta_ij = 0; and a pair at zero distance (particles seeded onto the same clamped
  // point) adds nothing: its term is (v_i - v_j) dW(0) / 1e-20 with v_i = v_j, both
  // sums of the same terms in the same order
  use = use && etaq != 0.f && rr > 0.f;
  float rt2 = rr;
  if (TS) {
    float dn = 0.f;
#pragma unroll
    for (int a = 0; a < D; a++) dn += dr0[a] * o.n[a];
    rt2 = 0.f;
#pragma unroll
    for (int a = 0; a < D; a++) {
      const float t = dr0[a] - o.n[a] * dn;
      rt2 += t * t;
    }
  }
  // (dW/dr) / r as in the stage-1 pair
  const float irt = frsq(fmaxf(rt2, 1.e-30f));
  const float rt = rt2 * irt;
  const float ihm = frcp(0.5f * (o.h + hq));
  const float C3 = osurf ? o.kd3 * ipow<D>(ihm) : o.kd3 * ipow<D + 1>(ihm);
  const float s2 = rt * ihm;
  const float t2 = fmaxf(2.f - s2, 0.f);
  const float dq = s2 < 1.f ? s2 * s2 - 2.f : -(t2 * t2);
  const float aw = C3 * dq * irt;
  const bool mirror_v = !osurf && tq != SMG_REAL;   // Q8: v_j := -v_i
  const float etaij = 2.f * etaq * o.eta * frcp(o.eta + etaq + 1.e-15f);
  const float temp0 = aw * etaij * (o.vol * o.vol + volq * volq);
#pragma unroll
  for (int a = 0; a < D; a++) {
    const float qa = a == 0 ? qv.x : (a == 1 ? qv.y : qv.z);
    const float vq = mirror_v ? -o.v[a] : qa * hdt;
    F[a] = use ? F[a] + (o.v[a] - vq) * temp0 : F[a];
  }
}

// The stage-2 pairs the stage-1 pass kept (verified, unmirrored, not self), with
// staged records carrying (a1, nu of the previous step) for the neighbour's velocity
// and viscosity; the same thread-to-particle order as the stage-1 pass.
template <int D>
__global__ __launch_bounds__(kTile) __attribute__((amdgpu_waves_per_eu(4))) void k_force2(Work W, PState S, int par) {
  __shared__ float4 hh[kHL2];
  __shared__ uint2 hq[kHL2];
  __shared__ float4 hv[kHL2];
  __shared__ double lds_d[4];
  __shared__ int s_surf;
  __shared__ JumpTile jt;
  const int64_t tile = tile_id(W);
  const int j = W.tile_job[tile];
  if (j < 0) return;
  const JobParams& J = W.jp[j];
  const int64_t ts = tile * kTile;
  const int p = W.lane[ts + threadIdx.x];
  const int64_t slot = ts + p;
  const bool live = slot - J.p0 < J.n;
  const float4 r0 = W.rec[4 * slot + 0], r1 = W.rec[4 * slot + 1], r2 = W.rec[4 * slot + 2];
  {
    const float xb[3] = {r0.x, r0.y, r0.z};
    if (threadIdx.x == 0) s_surf = 0;
    jumper_prepare<D>(W, J, j, live, xb, jt);   // ends with a barrier
  }
  const double dt1 = acc_dt(W, j, par, false);
  const float hdt = float(0.5 * dt1);
  const int H = W.hcnt[tile];
  const int HS = H < kHL2 ? H : kHL2;
  stage_halo<D, kHL2 / kTile, true>(W, J, tile, ts, HS, hh, hq, hv);
  if (live && (__float_as_int(r1.w) & ~kJumpBit) == SMG_SURFACE) s_surf = 1;
  // this lane's stage-2 item count (0 when its nu is 0) and the wavefront's largest
  const int cnt = (live && r2.w != 0.f) ? W.nbc2[slot] : 0;
  const int kmax = __builtin_amdgcn_readfirstlane(wave_reduce(cnt, OpMaxI()));
  __syncthreads();
  const bool tsurf = s_surf != 0;
  double dt = 1.0e300, U = 0.;
  if (live) {
    float ni[3] = {0.f, 0.f, 0.f};
    if ((__float_as_int(r1.w) & ~kJumpBit) == SMG_SURFACE) {
#pragma unroll
      for (int c = 0; c < D; c++) ni[c] = S.nrm[c][slot];
    }
    Own o;
    own_common<D>(J, r0, r1, ni, o);
    o.eta = r2.w * (1.f / o.vol);
    const float a1i[3] = {r2.x, r2.y, r2.z};
    double vd[3] = {0., 0., 0.};
#pragma unroll
    for (int c = 0; c < 3; c++) o.v[c] = 0.f;
#pragma unroll
    for (int c = 0; c < D; c++) {
      o.v[c] = a1i[c] * hdt;
      vd[c] = double(a1i[c]) * (0.5 * dt1);
    }
    float F[3] = {0.f, 0.f, 0.f};
    if (o.eta != 0.f) {   // eta_ij = 0 for every pair when this particle's nu is 0
      if (kmax > 0) {
        const uint4* it = reinterpret_cast<const uint4*>(W.items2 + W.tbase[tile] + W.ioff[slot]);
        // the items of the stage-1 walk, to the wavefront's longest list; an item past
        // the lane's list is its own record (rr = 0: it adds nothing)
        auto walk = [&](auto ts_, auto fits_) __attribute__((always_inline)) {
          constexpr bool TS = decltype(ts_)::value, kFits = decltype(fits_)::value;
          const int ngw = (kmax + 7) >> 3;
          uint4 gn = it[0];
          for (int g = 0; g < ngw; g++) {
            uint4 gc = gn;
            asm volatile("" : "+v"(gc.x), "+v"(gc.y), "+v"(gc.z), "+v"(gc.w));
            gn = it[(g + 1 < ngw ? g + 1 : g) * 64];
            uint32_t w0 = gc.x, w1 = gc.y, w2 = gc.z, w3 = gc.w;
#pragma unroll 1
            for (int v = 0; v < 8; v++) {   // an item: the low half of w0, then the high half
              const bool valid = 8 * g + v < cnt;
              const uint32_t e = valid ? w0 & 0xFFFFu : uint32_t(p);
              w0 = (v & 1) ? w1 : w0 >> 16;
              if (v & 1) {
                w1 = w2;
                w2 = w3;
              }
              float4 qh, qv;
              uint2 qc;
              if (kFits || e < uint32_t(HS)) {
                qh = hh[e];
                qc = hq[e];
                qv = hv[e];
              } else {
                int64_t q;
                uint32_t code;
                halo_entry(W, tile, ts, J.p0, e, q, code);
                float4 rl;
                halo_rec<D>(W, J, q, code, qh, rl);
                qc = pack_lo(rl);
                qv = W.rec[4 * q + 2];
              }
              float lo[3];
              int tw;
              unpack_lo(qc, lo, tw);
              float dr[3] = {0.f, 0.f, 0.f};
              rel_hl<D>(o.xh, o.xl, qh, lo, dr);
              float rr = 0.f;
#pragma unroll
              for (int c = 0; c < D; c++) rr += dr[c] * dr[c];
              pair2<D, TS>(J, o, hdt, dr, rr, qh.w, tw & int(kQType), qv, valid, F);
            }
          }
        };
        if (H <= kHL2) {
          if (tsurf) walk(std::true_type{}, std::true_type{});
          else walk(std::false_type{}, std::true_type{});
        } else {
          walk(std::true_type{}, std::false_type{});
        }
      }
      // pairs with this step's jumpers (and, for a jumper, its own list), after the items
      const float xhi[3] = {r0.x, r0.y, r0.z}, xli[3] = {r1.x, r1.y, r1.z};
      for_each_neighbor<D>(W, S, J, slot, jt, [&](int, int64_t q, uint32_t code, const float4& qh, const float4& ql,
                                                   const float*, float) {
        const int tq = __float_as_int(ql.w) & ~kJumpBit;
        const bool use = code == 0u && q != slot && !(o.surf && tq == SMG_REAL);
        bool aligned = true;
        if (use && o.surf && tq == SMG_SURFACE) {
          const float4 qn = W.rec[4 * q + 3];
          const float nq[3] = {qn.x, qn.y, qn.z};
          float ca = 0.f;
#pragma unroll
          for (int c = 0; c < D; c++) ca += ni[c] * nq[c];
          aligned = fabsf(ca) > J.cos_thres;
        }
        float dr[3] = {0.f, 0.f, 0.f};
        rel_pos<D>(J, qh, ql, 0u, xhi, xli, dr);
        float rr = 0.f;
#pragma unroll
        for (int c = 0; c < D; c++) rr += dr[c] * dr[c];
        pair2<D, true>(J, o, hdt, dr, rr, qh.w, tq, W.rec[4 * q + 2], use && aligned, F);
      }, true);
    }
    if (o.surf) {
      float nf = 0.f;
#pragma unroll
      for (int c = 0; c < D; c++) nf += F[c] * ni[c];
#pragma unroll
      for (int c = 0; c < D; c++) F[c] = F[c] - ni[c] * nf;
    }
    double A = 0.;
#pragma unroll
    for (int c = 0; c < D; c++) {
      W.a2[c][slot] = F[c];
      A += double(F[c]) * double(F[c]);
      U += vd[c] * vd[c];
    }
    A = sqrt(A);
    U = sqrt(U);
    const double hd = S.h[slot];
    double t = 0.25 * sqrt(hd / (A + 1.e-15));
    t = amin(t, hd / 40. / (U + 1.e-15));
    const double nu = J.nu_coeff * hd * U;
    t = amin(t, 0.125 * hd * hd / (nu + 1.e-15));
    W.nu_b[slot] = float(nu);
    dt = t;
  }
  dt = block_reduce(dt, OpMin(), lds_d);
  U = block_reduce(U, OpMax(), lds_d);
  if (threadIdx.x == 0 && dt < 1.0e300) {
    JobAcc& A = acc_mine(W, j);
    atomicMin(&A.dt2[par], dbits(dt));
    atomicMax(&A.maxv[par], dbits(U));
  }
}

// ---------------------------------------------------------------------------
// 2D triangulation, geometric stage: the Voronoi cell of each particle among its
// neighbours (Voronoi_construction::Local_VD_generation, voronoi_construction.cpp:6-127)
// inside the box x_i +- 1.5 cut_off h_i that the reference gives voro++, built by
// clipping the box with the bisector half-planes in FP64.  Each cell edge that a
// neighbour's bisector made is a Voronoi face: the particle id of that neighbour (a
// mirror image counts as its source particle, as the reference's ghosts carry their
// source's local_id) goes to out_nb[loc * KE + c].  out_cnt[loc] = number of faces, or
// -1 when the cell needs more than kVorVerts vertices or KE faces.

constexpr int kVorVerts = 48;

template <int D>
__global__ __launch_bounds__(kTile) void k_vor2d(Work W, PState S, int32_t* out_nb, int32_t* out_cnt, int KE) {
  __shared__ JumpTile jt;
  const int j = W.tile_job[tile_id(W)];
  if (j < 0) return;
  const JobParams& J = W.jp[j];
  const int64_t slot = tile_id(W) * kTile + threadIdx.x;
  const int64_t loc = slot - J.p0;
  const bool live = loc < J.n;
  const float4 r0 = W.rec[4 * slot + 0], r1 = W.rec[4 * slot + 1];
  const float xhi[3] = {r0.x, r0.y, r0.z};
  jumper_prepare<D>(W, J, j, live, xhi, jt);
  if (!live) return;
  const double xi = double(xhi[0]) + double(r1.x), yi = double(xhi[1]) + double(r1.y);
  const double sc = S.h[slot] * J.cut_off * 1.5;
  double px[kVorVerts], py[kVorVerts];
  int32_t lab[kVorVerts];   // label of the edge from vertex v to v+1: -1 box wall, else the neighbour entry
  int nv = 4;
  px[0] = xi - sc; py[0] = yi - sc; lab[0] = -1;
  px[1] = xi + sc; py[1] = yi - sc; lab[1] = -1;
  px[2] = xi + sc; py[2] = yi + sc; lab[2] = -1;
  px[3] = xi - sc; py[3] = yi + sc; lab[3] = -1;
  bool bad = false;
  for_each_neighbor<D>(W, S, J, slot, jt, [&](int, int64_t q, uint32_t code, const float4& qh, const float4& ql,
                                               const float*, float) {
    if (bad || nv == 0 || (q == slot && code == 0)) return;
    const double xq[3] = {double(qh.x) + double(ql.x), double(qh.y) + double(ql.y), 0.};
    double g[3];
    image_of<D>(J, code, xq, g);
    // keep the side closer to x_i: f(p) = (p - x_i).(g - x_i) - |g - x_i|^2 / 2 <= 0
    const double ux = g[0] - xi, uy = g[1] - yi;
    const double c = 0.5 * (ux * ux + uy * uy);
    double f[kVorVerts];
    bool any_out = false, any_in = false;
    for (int v = 0; v < nv; v++) {
      f[v] = (px[v] - xi) * ux + (py[v] - yi) * uy - c;
      if (f[v] > 0.) any_out = true;
      else any_in = true;
    }
    if (!any_out) return;   // the bisector misses the cell
    if (!any_in) {          // only for a point coincident with x_i
      nv = 0;
      return;
    }
    double nx[kVorVerts], ny[kVorVerts];
    int32_t nl[kVorVerts];
    int m = 0;
    const int32_t label = int32_t(nb_pack(q, J.p0, code));
    for (int v = 0; v < nv; v++) {
      const int w = v + 1 < nv ? v + 1 : 0;
      const bool vin = f[v] <= 0., win = f[w] <= 0.;
      if (vin) {
        if (m >= kVorVerts) { bad = true; return; }
        nx[m] = px[v]; ny[m] = py[v];
        nl[m] = lab[v];   // the edge from v (to w or to the cut point) keeps its label
        m++;
      }
      if (vin != win) {   // the edge crosses the bisector: the cut point starts a new edge
        const double t = f[v] / (f[v] - f[w]);
        if (m >= kVorVerts) { bad = true; return; }
        nx[m] = px[v] + t * (px[w] - px[v]);
        ny[m] = py[v] + t * (py[w] - py[v]);
        // leaving the kept side: the new edge runs along the bisector; entering it: the
        // edge continues the old edge v -> w
        nl[m] = vin ? label : lab[v];
        m++;
      }
    }
    for (int v = 0; v < m; v++) {
      px[v] = nx[v]; py[v] = ny[v]; lab[v] = nl[v];
    }
    nv = m;
  });
  int nf = 0;
  if (!bad) {
    for (int v = 0; v < nv; v++) {
      if (lab[v] < 0) continue;
      const int32_t src = S.id[nb_slot(uint32_t(lab[v]), J.p0)];
      if (src == S.id[slot]) continue;   // the cell face with the particle's own mirror image
      bool seen = false;
      for (int u = 0; u < nf; u++) seen = seen || out_nb[loc * KE + u] == src;
      if (seen) continue;
      if (nf >= KE) { bad = true; break; }
      out_nb[loc * KE + nf] = src;
      nf++;
    }
  }
  out_cnt[loc] = bad ? -1 : nf;
}

// FP64 level-set phi at arbitrary points of one job (Calculate_particle_infor)
template <int D>
__global__ void k_sample_phi(Work W, int job, int64_t n, const double* pts, double* phi, double* nrm) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const JobParams& J = W.jp[job];
  const double x[3] = {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
  double p, nn[3] = {0., 0., 0.};
  if (nrm) {   // and the interpolated normal (Get_normal), not renormalised
    ls_sample<D, true>(J, W, x, p, nn);
    for (int a = 0; a < 3; a++) nrm[3 * i + a] = nn[a];
  } else {
    ls_sample<D, false>(J, W, x, p, nn);
  }
  phi[i] = p;
}

// residual of the last step without moving anything (for statistics); the same
// per-particle arithmetic as the residual that the next k_map accumulates
template <int D>
__global__ __launch_bounds__(kTile) void k_residual(Work W, PState S, int par) {
  __shared__ double lds_d[4];
  __shared__ unsigned long long lds_u[4];
  const int j = W.tile_job[tile_id(W)];
  if (j < 0) return;
  const JobParams& J = W.jp[j];
  const int64_t slot = tile_id(W) * kTile + threadIdx.x;
  double res = 0., md = 0.;
  unsigned long long rq = 0ull;
  if (slot - J.p0 < J.n) {
    const double dt1 = acc_dt(W, j, par, false), dt2 = acc_dt(W, j, par, true);
    double U = 0.;
#pragma unroll
    for (int a = 0; a < D; a++) {
      double v = double(W.a1[a][slot]) * (0.5 * dt1);
      v = v + double(W.a2[a][slot]) * (0.5 * dt2);
      U += v * v;
    }
    U = sqrt(U);
    const Disp dd = disp_of(U, dt2, S.h[slot], J);
    res = dd.r;
    md = dd.md;
    rq = dd.rq;
  }
  res = block_reduce(res, OpMax(), lds_d);
  md = block_reduce(md, OpMax(), lds_d);
  rq = block_reduce(rq, OpAddU(), lds_u);
  if (threadIdx.x == 0) {
    JobAcc& A = acc_mine(W, j);
    atomicMax(&A.res[par], dbits(res));
    atomicMax(&A.maxdisp[par], dbits(md));
    if (rq) atomicAdd(&A.rsum[par], rq);
  }
}

__global__ void k_clear_res(Work W, int njobs, int par) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;   // job * kAccCopies + copy
  if (i < int64_t(njobs) * kAccCopies) {
    W.acc[i].res[par] = 0ull;
    W.acc[i].maxdisp[par] = 0ull;
    W.acc[i].rsum[par] = 0ull;
  }
}

// FP64 conformance probe: sqrt, division, floor and tanh of given inputs
__global__ void k_fp64_probe(int64_t n, const double* a, const double* b, double* out) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[4 * i + 0] = sqrt(a[i]);
  out[4 * i + 1] = a[i] / b[i];
  out[4 * i + 2] = floor(a[i] / b[i]);
  out[4 * i + 3] = tanh(a[i]);
}

// ---------------------------------------------------------------------------
// host-side launchers

#define SMG_DISPATCH(D, KERNEL, ...)                                      \
  do {                                                                    \
    if ((D) == 2) KERNEL<2><<<__VA_ARGS__>>>;                             \
    else KERNEL<3><<<__VA_ARGS__>>>;                                      \
  } while (0)

void launch_decide(hipStream_t st, const Work& W, int njobs, int force) {
  if (njobs > 0) k_decide<<<njobs, 64, 0, st>>>(W, njobs, force);
}
void launch_key(hipStream_t st, int D, int64_t ntiles, const Work& W, const PState& S) {
  if (D == 2) k_key<2><<<ntiles, kTile, 0, st>>>(W, S);
  else k_key<3><<<ntiles, kTile, 0, st>>>(W, S);
}
void launch_copyback(hipStream_t st, int D, int64_t ntiles, const Work& W, const PState& A, const PState& B) {
  if (D == 2) k_copyback<2><<<ntiles, kTile, 0, st>>>(W, A, B);
  else k_copyback<3><<<ntiles, kTile, 0, st>>>(W, A, B);
}
void launch_nb_out(hipStream_t st, int D, int64_t ntiles, const Work& W, const PState& S, int pass, const int64_t* offs,
                   int32_t* cnt, int32_t* out) {
  if (D == 2) k_nb_out<2><<<ntiles, kTile, 0, st>>>(W, S, pass, offs, cnt, out);
  else k_nb_out<3><<<ntiles, kTile, 0, st>>>(W, S, pass, offs, cnt, out);
}
void launch_jump_lists(hipStream_t st, int D, const Work& W, const PState& S, int njobs) {
  const dim3 g(kJumpMax, njobs);
  if (D == 2) k_jump_lists<2><<<g, 64, 0, st>>>(W, S);
  else k_jump_lists<3><<<g, 64, 0, st>>>(W, S);
}
void launch_reset(hipStream_t st, const Work& W, int njobs, int par) {
  if (njobs > 0) k_reset<<<njobs, 64, 0, st>>>(W, njobs, par);
}
void launch_seed(hipStream_t st, int D, int64_t ntiles, const Work& W, const PState& S) {
  if (D == 2) k_seed<2><<<ntiles, kTile, 0, st>>>(W, S);
  else k_seed<3><<<ntiles, kTile, 0, st>>>(W, S);
}
void launch_map(hipStream_t st, int D, int64_t ntiles, const Work& W, const PState& S, int mode, int par) {
  if (mode == kModeStep) { if (D == 2) k_map<2, kModeStep><<<ntiles, kTile, 0, st>>>(W, S, par); else k_map<3, kModeStep><<<ntiles, kTile, 0, st>>>(W, S, par); }
  else if (mode == kModePrerun) { if (D == 2) k_map<2, kModePrerun><<<ntiles, kTile, 0, st>>>(W, S, par); else k_map<3, kModePrerun><<<ntiles, kTile, 0, st>>>(W, S, par); }
  else { if (D == 2) k_map<2, kModeInit><<<ntiles, kTile, 0, st>>>(W, S, par); else k_map<3, kModeInit><<<ntiles, kTile, 0, st>>>(W, S, par); }
}
void launch_scan(hipStream_t st, const Work& W, bool gated, bool per_job) {
  const int64_t nb = (W.n_tbl + kScanItems - 1) / kScanItems;
  if (nb == 0) return;
  const unsigned long long* gate = gated ? W.ctr + 1 : nullptr;
  const int32_t* bj = per_job ? W.sblk_job : nullptr;
  k_scan_reduce<<<nb, kTile, 0, st>>>(W.counts, W.n_tbl, W.bsum, gate, bj, W.red);
  k_scan_top<<<1, 1024, 0, st>>>(W.bsum, nb, gate);
  k_scan_down<<<nb, kTile, 0, st>>>(W.counts, W.n_tbl, W.bsum, nb, W.starts, gate, bj, W.red);
}
void launch_scatter(hipStream_t st, int64_t ntiles, const Work& W, const PState& S) {
  k_scatter<<<ntiles, kTile, 0, st>>>(W, S);
}
void launch_segsort(hipStream_t st, int64_t ntiles, const Work& W, const PState& S) {
  (void)S;
  k_segsort<<<ntiles, kTile, 0, st>>>(W);
}
void launch_gather(hipStream_t st, int D, int64_t ntiles, const Work& W, const PState& A, const PState& B) {
  if (D == 2) k_gather<2><<<ntiles, kTile, 0, st>>>(W, A, B);
  else k_gather<3><<<ntiles, kTile, 0, st>>>(W, A, B);
}
void launch_nb(hipStream_t st, int D, int64_t ntiles, const Work& W, const PState& S) {
  if (D == 2) k_nb<2><<<ntiles, kTile, 0, st>>>(W, S);
  else k_nb<3><<<ntiles, kTile, 0, st>>>(W, S);
}
void launch_nb_spill(hipStream_t st, int D, int64_t ntiles, const Work& W, const PState& S) {
  const int64_t g = ntiles < 1024 ? ntiles : 1024;   // grid-stride over the (usually empty) queue
  if (D == 2) k_nb_spill<2><<<g, kTile, 0, st>>>(W, S);
  else k_nb_spill<3><<<g, kTile, 0, st>>>(W, S);
}
void launch_halo(hipStream_t st, int D, int64_t ntiles, const Work& W, const PState& S) {
  // Lists of up to 256 items per slot (inline + spill) keep a tile's halo within one
  // pass of a 4,096-slot table (24 KB, 5 workgroups per CU); longer lists take the
  // 8,192-slot table (48 KB, 3 per CU), so that their halos take half the passes.
  if (W.ips <= kHashSmallIps) {
    if (D == 2) k_halo<2, kHashN><<<ntiles, kTile, 0, st>>>(W, S);
    else k_halo<3, kHashN><<<ntiles, kTile, 0, st>>>(W, S);
  } else {
    if (D == 2) k_halo<2, 2 * kHashN><<<ntiles, kTile, 0, st>>>(W, S);
    else k_halo<3, 2 * kHashN><<<ntiles, kTile, 0, st>>>(W, S);
  }
}
void launch_force1(hipStream_t st, int D, int64_t ntiles, const Work& W, const PState& S, int par) {
  if (D == 2) k_force1<2><<<ntiles, kTile, 0, st>>>(W, S, par);
  else k_force1<3><<<ntiles, kTile, 0, st>>>(W, S, par);
}
void launch_force2(hipStream_t st, int D, int64_t ntiles, const Work& W, const PState& S, int par) {
  if (D == 2) k_force2<2><<<ntiles, kTile, 0, st>>>(W, S, par);
  else k_force2<3><<<ntiles, kTile, 0, st>>>(W, S, par);
}
void launch_residual(hipStream_t st, int D, int64_t ntiles, const Work& W, const PState& S, int njobs, int par) {
  k_clear_res<<<(njobs * kAccCopies + 255) / 256, 256, 0, st>>>(W, njobs, par);
  if (D == 2) k_residual<2><<<ntiles, kTile, 0, st>>>(W, S, par);
  else k_residual<3><<<ntiles, kTile, 0, st>>>(W, S, par);
}
void launch_vor2d(hipStream_t st, int64_t ntiles, const Work& W, const PState& S, int32_t* out_nb, int32_t* out_cnt,
                  int KE) {
  k_vor2d<2><<<ntiles, kTile, 0, st>>>(W, S, out_nb, out_cnt, KE);
}
void launch_sample_phi(hipStream_t st, int D, const Work& W, int job, int64_t n, const double* pts, double* phi,
                       double* nrm) {
  if (n <= 0) return;
  const int64_t g = (n + 255) / 256;
  if (D == 2) k_sample_phi<2><<<g, 256, 0, st>>>(W, job, n, pts, phi, nrm);
  else k_sample_phi<3><<<g, 256, 0, st>>>(W, job, n, pts, phi, nrm);
}
void launch_fp64_probe(hipStream_t st, int64_t n, const double* a, const double* b, double* out) {
  k_fp64_probe<<<(n + 255) / 256, 256, 0, st>>>(n, a, b, out);
}

// Staging area of smg_sample_phi_static: points (3 doubles each) then phi (1 double
// each), in a file-scope device array so that the entry point allocates nothing.
// Every GPU holds its own copy, at its own address, until the process ends; in host
// code the name `scratch` is only a stand-in for the symbol, so the address each
// device's copy has is looked up once per device and kept.
static __device__ __align__(256) double scratch[kStaticSampleDoubles];

int static_sample_scratch(double** ptr) {
  constexpr int kMaxDevices = 64;
  static std::mutex mu;
  static double* addr[kMaxDevices] = {};
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return -1;
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return -1;
    addr[dev] = static_cast<double*>(p);
  }
  *ptr = addr[dev];
  return 0;
}

}  // namespace smg
