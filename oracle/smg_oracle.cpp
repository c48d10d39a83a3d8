// smg_oracle.cpp -- FP64 CPU restatement of SPH-mesh-generation's particle-relaxation
// step.  TEST INFRASTRUCTURE ONLY: this is the checker for the HIP engine.  Only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
// product library never links or calls it.
//
// Every function cites the reference file:line it restates (paths relative to the
// reference's src/).  Arithmetic is written in the reference's evaluation order,
// in double, and the file is compiled with -ffp-contract=off so no FMA changes a
// rounding.  rand() is replaced by the counter-based generator or_uniform() that the
// HIP engine implements identically.
//
// Pinning (see tests/test_oracle_*.py): kernel normalisation and continuity, lattice
// symmetry (zero force on a perfect lattice), hand two-particle cases, brute-force
// O(N^2) neighbour sets, the reference's own tree search restated in its default and
// _SCLL_ builds (neighbors_tree), closed-form level-set sampling, mirror geometry, and an
// independent long-double evaluation of a step.  Functions with no such check are
// listed as "parity unpinned" in DESIGN.md.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <cstdio>
#include <vector>
#include <algorithm>
#include <omp.h>

#include "../include/smg.h"

namespace {

const double OR_PI = 3.1415926;   // glbparam.h:185 (quirk Q4: truncated pi)

// When set, positions are rounded to the engine's storage format (an unevaluated
// pair of floats, hi + lo) at the points where the engine stores them between
// passes: after seeding and after the mapping pass.  Bit-exact comparisons use it so
// that both sides see the same stored positions; the arithmetic is unchanged.
int g_round_positions = 0;
inline double store_round(double x) {
  if (!g_round_positions) return x;
  const float hi = float(x);
  const float lo = float(x - double(hi));
  return double(hi) + double(lo);
}

// ---------------------------------------------------------------------------
// counter-based RNG (replaces rand(), sph_mesh_generation.cpp:1327-1342, 1909-1921)
inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
inline uint64_t rng_key(uint64_t seed, uint64_t stream, uint64_t id) {
  uint64_t k = mix64(seed + 0x632BE59BD9B4E019ull * (stream + 1ull));
  return mix64(k ^ (id * 0xD1B54A32D192ED03ull + 0x8CB92BA72F3D8DD7ull));
}
inline double rng_u(uint64_t key, uint64_t counter) {
  return double(mix64(key + counter * 0x9E3779B97F4A7C15ull) >> 11) * 0x1.0p-53;
}

// ---------------------------------------------------------------------------
// hyperbolic kernel, particle.cpp:288-310 (W) and :686-709 (dW/dr / r)
inline double powern(double a, int n) {   // glbfunc.h:42-54
  double r = 1.0;
  for (int i = 0; i < n; i++) r *= a;
  return r;
}
inline double kernel_C(int dim) {
  if (dim == 1) return 1.0 / 7.0;
  if (dim == 2) return 1.0 / (3.0 * OR_PI);
  return 15.0 / (62.0 * OR_PI);
}
inline double W(int dim, double dist, double h) {
  double C = kernel_C(dim) / powern(h, dim);
  double s = dist / h, val;
  if (s < 1) val = C * (6.0 - 6.0 * s + s * s * s);
  else if (s < 2) val = C * (2 - s) * (2 - s) * (2 - s);
  else val = 0.0;
  return val;
}
inline double dW(int dim, double dist, double h) {
  double C = kernel_C(dim) / powern(h, dim + 1);
  double s = dist / h, val;
  if (s < 1) val = C * (-6.0 + 3. * s * s);
  else if (s < 2) val = C * (-3.0) * (2 - s) * (2 - s);
  else val = 0.0;
  val /= (dist + 1.e-20);
  return val;
}

struct V3 { double x[3]; };

struct Ctx {
  const smg_config* cfg;
  int D;
  int nc[3];          // interior level-set cells per axis
  int ng[3];          // grid cells per axis incl. one-cell halo
  double dpkg[3], dcell[3], dl;
  double box_r[3];
  const double* phi;  // one-cell-halo grid
  const float* nrm;
  const int8_t* tag;
  // refined level-set levels (the multi-resolution extension; none for the reference's
  // single-level set): level l's grid and which of its packages are stored
  int nlev = 0;
  const Ctx* lv[3] = {nullptr, nullptr, nullptr};
  const uint8_t* act[3] = {nullptr, nullptr, nullptr};
};

// the refined levels or_set_levels registered (applied to every context built after)
struct LevelReg {
  smg_config cfg;
  const double* phi;
  const float* nrm;
  const int8_t* tag;
  const uint8_t* active;
};
LevelReg g_lreg[3];
int g_nlreg = 0;
Ctx g_lctx[3];

inline int64_t gidx(const Ctx& c, int gi, int gj, int gk) {  // global cell (may be -1 .. nc)
  int64_t i = gi + 1, j = (c.D >= 2) ? gj + 1 : 0, k = (c.D >= 3) ? gk + 1 : 0;
  return (k * c.ng[1] + j) * c.ng[0] + i;
}

void ctx_init(Ctx& c, const smg_config* cfg, const double* phi, const float* nrm, const int8_t* tag) {
  c.cfg = cfg; c.D = cfg->dim;
  c.dl = 1.e20;
  for (int a = 0; a < 3; a++) {
    bool on = a < c.D;
    c.nc[a] = on ? 4 * cfg->num_pkg[a] : 1;
    c.ng[a] = on ? c.nc[a] + 2 : 1;
    // lset_level_infor.cpp:33-36: dpkg = domain/num_pkg, dcell = dpkg/ICP
    c.dpkg[a] = on ? cfg->domain[a] / double(cfg->num_pkg[a]) : 1.0;
    c.dcell[a] = on ? c.dpkg[a] / 4.0 : 1.0;
    if (on) c.dl = std::min(c.dl, c.dcell[a]);
    c.box_r[a] = cfg->box_l[a] + cfg->domain[a];
  }
  c.phi = phi; c.nrm = nrm; c.tag = tag;
  c.nlev = 0;
  if (phi && cfg != nullptr) {
    for (int l = 0; l < g_nlreg; l++) {
      Ctx& L = g_lctx[l];
      L.cfg = &g_lreg[l].cfg; L.D = c.D; L.dl = 1.e20; L.nlev = 0;
      for (int a = 0; a < 3; a++) {
        bool on = a < c.D;
        L.nc[a] = on ? 4 * g_lreg[l].cfg.num_pkg[a] : 1;
        L.ng[a] = on ? L.nc[a] + 2 : 1;
        L.dpkg[a] = on ? g_lreg[l].cfg.domain[a] / double(g_lreg[l].cfg.num_pkg[a]) : 1.0;
        L.dcell[a] = on ? L.dpkg[a] / 4.0 : 1.0;
        if (on) { L.dl = std::min(L.dl, L.dcell[a]); c.dl = std::min(c.dl, L.dcell[a]); }
        L.box_r[a] = c.box_r[a];
      }
      L.phi = g_lreg[l].phi; L.nrm = g_lreg[l].nrm; L.tag = g_lreg[l].tag;
      c.lv[l] = &L;
      c.act[l] = g_lreg[l].active;
      c.nlev = l + 1;
    }
  }
}

// clamped shift into the level-set box, particle_mesh_generation.cpp:182-185
inline void coord_shift(const Ctx& c, const double* x, double* cs) {
  for (int a = 0; a < 3; a++) {
    if (a < c.D) {
      double t = x[a] - c.cfg->box_l[a];
      t = (t >= 0.) ? t : 0.;                                   // AMAX1(., 0.)
      double hi = c.cfg->domain[a] - 1.e-8;
      cs[a] = (t <= hi) ? t : hi;                               // AMIN1(., domain-1e-8)
    } else cs[a] = 0.;
  }
}
// get_cell_id (glbfunc.h:196-208) for the package grid with glbl start -1 / end num_pkg+1
inline int pkg_of(const Ctx& c, int a, double cs) {
  if (a >= c.D) return 0;
  int p = int(std::floor(cs / c.dpkg[a]));
  int lo = -1, hi = c.cfg->num_pkg[a] + 1 - 1;
  p = (p <= hi) ? p : hi;
  p = (lo >= p) ? lo : p;
  return p;
}

// Calculate_particle_infor(SOLVER*), particle_mesh_generation.cpp:175-265
void sample_grid(const Ctx& c, const double* x, double& phi, double* n) {
  double cs[3]; coord_shift(c, x, cs);
  int pk[3]; double rr[3] = {0., 0., 0.};
  for (int a = 0; a < 3; a++) pk[a] = pkg_of(c, a, cs[a]);
  for (int a = 0; a < 3; a++) {
    if (a < c.D) {
      double csp = cs[a] - c.dpkg[a] * pk[a];
      rr[a] = csp / c.dcell[a] - 0.5;
    } else rr[a] = 0.;
    double hi = (a < c.D) ? 4.0 - 0.5 : 1.0 - 0.5;
    rr[a] = (rr[a] <= hi) ? rr[a] : hi;
    rr[a] = (-0.5 >= rr[a]) ? -0.5 : rr[a];
  }
  int m[3] = {0, 0, 0}; double f[3] = {0., 0., 0.};
  for (int a = 0; a < c.D; a++) { m[a] = int(std::floor(rr[a])); f[a] = rr[a] - double(m[a]); }
  int ex = 1, ey = (c.D >= 2) ? 1 : 0, ez = (c.D >= 3) ? 1 : 0;
  phi = 0.; n[0] = n[1] = n[2] = 0.;
  for (int r = 0; r <= ex; r++)
    for (int s = 0; s <= ey; s++)
      for (int t = 0; t <= ez; t++) {
        int64_t g = gidx(c, pk[0] * 4 + m[0] + r, pk[1] * 4 + m[1] + s, pk[2] * 4 + m[2] + t);
        double coef = (std::fabs(f[0] - (1.0 - r))) * (ey ? std::fabs(f[1] - (1.0 - s)) : 1.0) *
                      (ez ? std::fabs(f[2] - (1.0 - t)) : 1.0);
        phi += c.phi[g] * coef;
        n[0] += double(c.nrm[3 * g + 0]) * coef;
        n[1] += double(c.nrm[3 * g + 1]) * coef;
        n[2] += double(c.nrm[3 * g + 2]) * coef;
      }
  // particle_mesh_generation.cpp:262-263: the renormalised copy is discarded (Q3)
}

// A refined level stores package p (clamped into the level: a halo package is stored
// where the interior package next to it is)
bool level_stores(const Ctx& c, int l, const int* p) {
  const Ctx& L = *c.lv[l];
  int q[3] = {0, 0, 0};
  for (int a = 0; a < c.D; a++) {
    const int np = L.cfg->num_pkg[a];
    q[a] = p[a] < 0 ? 0 : (p[a] > np - 1 ? np - 1 : p[a]);
  }
  const int np0 = L.cfg->num_pkg[0], np1 = c.D >= 2 ? L.cfg->num_pkg[1] : 1;
  return c.act[l][(int64_t(q[2]) * np1 + q[1]) * np0 + q[0]] != 0;
}

// The sample on the finest refined level that stores the packages of all 2^D corner
// cells of the interpolation, else on the base grid (the multi-resolution extension)
void sample(const Ctx& c, const double* x, double& phi, double* n) {
  for (int l = c.nlev - 1; l >= 0; l--) {
    const Ctx& L = *c.lv[l];
    double cs[3]; coord_shift(L, x, cs);
    int pk[3] = {0, 0, 0}, m[3] = {0, 0, 0};
    for (int a = 0; a < c.D; a++) {
      pk[a] = pkg_of(L, a, cs[a]);
      double rr = (cs[a] - L.dpkg[a] * pk[a]) / L.dcell[a] - 0.5;
      rr = (rr <= 3.5) ? rr : 3.5;
      rr = (-0.5 >= rr) ? -0.5 : rr;
      m[a] = int(std::floor(rr));
    }
    bool ok = true;
    for (int corner = 0; corner < (1 << c.D) && ok; corner++) {
      int p[3] = {0, 0, 0};
      for (int a = 0; a < c.D; a++) {
        const int g = pk[a] * 4 + m[a] + ((corner >> a) & 1);
        p[a] = (g + 4) / 4 - 1;   // floor(g / 4), g >= -4
      }
      ok = level_stores(c, l, p);
    }
    if (ok) { sample_grid(L, x, phi, n); return; }
  }
  sample_grid(c, x, phi, n);
}

int tag_grid(const Ctx& c, const double* x);

// the tag of the cell holding x on the finest refined level that stores that cell's
// package, else on the base grid
int tag_at(const Ctx& c, const double* x) {
  for (int l = c.nlev - 1; l >= 0; l--) {
    const Ctx& L = *c.lv[l];
    double cs[3]; coord_shift(L, x, cs);
    int p[3] = {0, 0, 0};
    for (int a = 0; a < c.D; a++) p[a] = pkg_of(L, a, cs[a]);
    if (level_stores(c, l, p)) return tag_grid(L, x);
  }
  return tag_grid(c, x);
}

// Get_level_set_char_cell_tag_and_idx -> Levelset_levelinfo::Get_id_pkg_cell
// (particle_mesh_generation.cpp:552-573, lset_level_infor.cpp:350-366)
int tag_grid(const Ctx& c, const double* x) {
  double cs[3]; coord_shift(c, x, cs);
  int g[3] = {0, 0, 0};
  for (int a = 0; a < c.D; a++) {
    int p = pkg_of(c, a, cs[a]);
    double ps = cs[a] - c.dpkg[a] * p;
    int ci = int(std::floor(ps / c.dcell[a]));
    ci = (ci <= 3) ? ci : 3;
    ci = (0 >= ci) ? 0 : ci;
    g[a] = p * 4 + ci;
  }
  return c.tag[gidx(c, g[0], g[1], g[2])];
}

// Levelset::Get_scale (level_set.cpp:311-351) plus the linear-profile extension
double get_scale(const Ctx& c, double phi) {
  const smg_config* f = c.cfg;
  if (f->profile == SMG_PROFILE_LINEAR) {
    if (phi > 0. || std::fabs(phi) < c.dl) {
      double t = phi > 0. ? phi / f->lin_band : 0.;
      t = t < 1.0 ? t : 1.0;
      return f->lin_dx0 + t * (f->lin_dx1 - f->lin_dx0);
    }
    return f->lin_dx0;
  }
  if (phi > 0. || std::fabs(phi) < c.dl) {
    double factor = 2.5;
    return f->min_dl + std::tanh(factor * std::fabs(phi) / f->max_phi) / std::tanh(factor) * (f->max_dl - f->min_dl);
  }
  return f->min_dl;
}

struct P {            // one particle (the fields of particle.h:15-32 and particle_mesh_generation.h:59-73 the path uses)
  double x[3], v[3], a[3], F1[3], F2[3], n[3];
  double h, rho, vol, sum, nu, phi;
  int type;
  int id;
};

// Set_particle_info, particle_mesh_generation.cpp:361-381
void set_particle_info(const Ctx& c, P& p) {
  double h_tmp = get_scale(c, p.phi);
  p.h = h_tmp * c.cfg->cell_ratio / c.cfg->cut_off;
  p.vol = std::pow(h_tmp, c.D);
  p.rho = 1.0 / p.vol;   // mass = 1
}

// Remap_this_particle_to_positive_phase (sph_mesh_generation.cpp:1885-1928), rand() -> or_uniform
bool remap(const Ctx& c, P& p, uint64_t stream) {
  const smg_config* f = c.cfg;
  uint64_t key = rng_key(f->seed, stream, uint64_t(p.id));
  for (int att = 0; att < f->remap_max_attempts; att++) {
    for (int a = 0; a < c.D; a++) {
      double box = f->particle_box_r[a] - f->particle_box_l[a];
      double u0 = rng_u(key, uint64_t(att) * 6 + 2 * a);
      double u1 = rng_u(key, uint64_t(att) * 6 + 2 * a + 1);
      double x = u0 * box + u1 * 1.e-3 + f->particle_box_l[a];
      double lo = f->particle_box_l[a] + 1.e-10, hi = f->particle_box_r[a] - 1.e-10;
      x = (x <= hi) ? x : hi;
      x = (lo >= x) ? lo : x;
      p.x[a] = x;
    }
    sample(c, p.x, p.phi, p.n);
    if (p.phi > p.h) {
      p.type = SMG_REAL;
      set_particle_info(c, p);
      return true;
    }
  }
  return false;
}

// Find_and_map_particle FILL branch for one particle (sph_mesh_generation.cpp:1515-1543 ->
// particle_mesh_generation.cpp:408-467, Get_dispersed_position :385-404)
bool map_particle(const Ctx& c, P& p, uint64_t stream, int& remapped) {
  int tag = tag_at(c, p.x);
  if (tag == 1 && (p.type == SMG_REAL || p.type == SMG_SURFACE)) {
    p.type = SMG_SURFACE;
    double dist = std::fabs(p.phi);
    double nn = 0.; for (int a = 0; a < c.D; a++) nn += p.n[a] * p.n[a];
    double div = 1.0 / (std::sqrt(nn) + 1.0e-15);
    for (int a = 0; a < c.D; a++) {
      double nd = p.phi < 0.0 ? p.n[a] * div : -(p.n[a] * div);
      p.x[a] = p.x[a] + nd * dist;
    }
  } else {
    p.type = SMG_REAL;
  }
  if (p.type == SMG_SURFACE) {
    double dx = 0.00001 * p.h;
    double nn = 0.; for (int a = 0; a < c.D; a++) nn += p.n[a] * p.n[a];
    double div = 1.0 / (std::sqrt(nn) + 1.0e-15);
    for (int a = 0; a < c.D; a++) {   // phi and n from before the projection (Q12)
      double nd = p.phi < 0.0 ? p.n[a] * div : -(p.n[a] * div);
      p.x[a] = p.x[a] + nd * dx;
    }
  }
  sample(c, p.x, p.phi, p.n);
  if (p.phi < 0. && p.type == SMG_REAL) {
    remapped = 1;
    return remap(c, p, stream);
  }
  return true;
}

// ---------------------------------------------------------------------------
// neighbour sets with mirror ghosts (level_infor.cpp:469-506, 786-856; sph.cpp:749-758)
struct Img { double x[3]; int src; int code; };
struct NbEntry { int j; int code; };

inline double dist_of(int D, const double* xi, const double* xj) {   // get_distance, glbfunc.h:224-232
  double d = 0.0;
  for (int a = 0; a < D; a++) { double dr = xi[a] - xj[a]; d += dr * dr; }
  return std::sqrt(d);
}

inline bool nb_pred(double dist, double hi, double hj, double cut) {
  return (dist <= (hi * cut + 1.e-10)) || (dist <= (hj * cut + 1.e-10));
}

// image of x under mirror code (2 bits per axis: 1 low wall, 2 high wall)
inline void image(const Ctx& c, int code, const double* x, double* g) {
  for (int a = 0; a < 3; a++) {
    int m = (code >> (2 * a)) & 3;
    if (m == 1) g[a] = 2. * c.cfg->box_l[a] - x[a];
    else if (m == 2) g[a] = 2. * c.box_r[a] - x[a];
    else g[a] = x[a];
  }
}

void build_images(const Ctx& c, const std::vector<P>& ps, double band, std::vector<Img>& out) {
  out.clear();
  int n = int(ps.size());
  for (int j = 0; j < n; j++) {
    const P& p = ps[j];
    int opt[3][3]; int nopt[3];
    for (int a = 0; a < 3; a++) {
      nopt[a] = 1; opt[a][0] = 0;
      if (a < c.D && c.cfg->sym[a]) {
        if (p.x[a] - c.cfg->box_l[a] < band) opt[a][nopt[a]++] = 1;
        if (c.box_r[a] - p.x[a] < band) opt[a][nopt[a]++] = 2;
      }
    }
    for (int i0 = 0; i0 < nopt[0]; i0++)
      for (int i1 = 0; i1 < nopt[1]; i1++)
        for (int i2 = 0; i2 < nopt[2]; i2++) {
          int code = opt[0][i0] | (opt[1][i1] << 2) | (opt[2][i2] << 4);
          Img im; im.src = j; im.code = code;
          image(c, code, p.x, im.x);
          out.push_back(im);
        }
  }
}

// grid-accelerated search; the predicate is the only decision, the grid just
// enumerates candidates (cell >= the largest cutoff, so the 3^D scan is complete)
void neighbors(const Ctx& c, const std::vector<P>& ps, std::vector<std::vector<NbEntry>>& nb) {
  int n = int(ps.size()), D = c.D;
  double cut = c.cfg->cut_off;
  double hmax = 0.;
  for (const P& p : ps) hmax = std::max(hmax, p.h);
  double R = hmax * cut + 1.e-9;
  std::vector<Img> im;
  build_images(c, ps, R, im);
  double lo[3], inv = 1.0 / R; int dims[3];
  for (int a = 0; a < 3; a++) {
    if (a < D) { lo[a] = c.cfg->box_l[a] - 2 * R; dims[a] = int(std::ceil((c.cfg->domain[a] + 4 * R) * inv)) + 1; }
    else { lo[a] = 0.; dims[a] = 1; }
  }
  auto cell_of = [&](const double* x, int* ci) {
    for (int a = 0; a < 3; a++) {
      if (a < D) { int k = int(std::floor((x[a] - lo[a]) * inv)); ci[a] = std::min(std::max(k, 0), dims[a] - 1); }
      else ci[a] = 0;
    }
  };
  int64_t ncell = int64_t(dims[0]) * dims[1] * dims[2];
  std::vector<int> cnt(ncell + 1, 0), order(im.size());
  std::vector<int64_t> key(im.size());
  for (size_t q = 0; q < im.size(); q++) {
    int ci[3]; cell_of(im[q].x, ci);
    key[q] = (int64_t(ci[2]) * dims[1] + ci[1]) * dims[0] + ci[0];
    cnt[key[q] + 1]++;
  }
  for (int64_t k = 0; k < ncell; k++) cnt[k + 1] += cnt[k];
  std::vector<int> fill(cnt.begin(), cnt.end() - 1);
  for (size_t q = 0; q < im.size(); q++) order[fill[key[q]]++] = int(q);
  nb.assign(n, {});
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < n; i++) {
    const P& p = ps[i];
    int ci[3]; cell_of(p.x, ci);
    int rz = D >= 3 ? 1 : 0, ry = D >= 2 ? 1 : 0;
    std::vector<NbEntry>& L = nb[i];
    for (int dz = -rz; dz <= rz; dz++)
      for (int dy = -ry; dy <= ry; dy++)
        for (int dx = -1; dx <= 1; dx++) {
          int cx = ci[0] + dx, cy = ci[1] + dy, cz = ci[2] + dz;
          if (cx < 0 || cy < 0 || cz < 0 || cx >= dims[0] || cy >= dims[1] || cz >= dims[2]) continue;
          int64_t k = (int64_t(cz) * dims[1] + cy) * dims[0] + cx;
          for (int q = cnt[k]; q < cnt[k + 1]; q++) {
            const Img& g = im[order[q]];
            double d = dist_of(D, p.x, g.x);
            if (nb_pred(d, p.h, ps[g.src].h, cut)) L.push_back({g.src, g.code});
          }
        }
    std::sort(L.begin(), L.end(), [](const NbEntry& a, const NbEntry& b) {
      return a.j != b.j ? a.j < b.j : a.code < b.code; });
  }
}

// O(N^2) reference enumeration of the same set (pinning)
void neighbors_brute(const Ctx& c, const std::vector<P>& ps, std::vector<std::vector<NbEntry>>& nb) {
  int n = int(ps.size()), D = c.D;
  double cut = c.cfg->cut_off;
  nb.assign(n, {});
  int ncode = 1; for (int a = 0; a < D; a++) ncode *= 3;
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < n; i++) {
    for (int j = 0; j < n; j++)
      for (int cc = 0; cc < ncode; cc++) {
        int code = 0, t = cc, ok = 1;
        for (int a = 0; a < D; a++) { int m = t % 3; t /= 3; if (m && !c.cfg->sym[a]) ok = 0; code |= m << (2 * a); }
        if (!ok) continue;
        double g[3]; image(c, code, ps[j].x, g);
        double d = dist_of(D, ps[i].x, g);
        if (nb_pred(d, ps[i].h, ps[j].h, cut)) nb[i].push_back({j, code});
      }
    std::sort(nb[i].begin(), nb[i].end(), [](const NbEntry& a, const NbEntry& b) {
      return a.j != b.j ? a.j < b.j : a.code < b.code; });
  }
}

// The reference's own neighbour search, restated with its data structures, to pin the
// predicate form of neighbors() above; scll selects the _SCLL_ build.
// * The tree: max_scale = maximum_dl CELL_RATIO (= the largest h CUT_OFF),
//   min_scale likewise (sph_mesh_generation.cpp:781-783); Define_level_infor_tight
//   (sph.cpp:144-180) fits max_scale to the domain and counts the levels; level L has
//   ini_num_cell 2^L cells per axis of dcell = domain / num_cell and a margin of 2^L
//   cells on each mirrored side (Level_info::Initialize, level_infor.cpp:35-102).
// * A particle's level: Set_particle_scale from max_scale (particle.cpp:957-975,
//   particle_mesh_generation.cpp:57), matched to the level scales
//   (Map_the_particle_to_tree, sph.cpp:247-276); its cell: get_cell_id (glbfunc.h).
// * Mirror ghosts per level, copied from the mirrored cells before finer particles are
//   merged (Construct_symmetric_BC_particles, level_infor.cpp:786-856, called before
//   Update_every_level_info at sph_mesh_generation.cpp:2061-2068), then every level's
//   cells take the finer levels' particles and ghosts (Update_every_level_info,
//   level_infor.cpp:407-464).
// * Every particle and every ghost scans (SPH::Refresh_neighbor_info, sph.cpp:345-360):
//   the default build the 3^D cells around it, adding the neighbour to itself and
//   itself to a neighbour of another level (level_infor.cpp:469-506); the _SCLL_ build
//   the sub-cells (cells split in 3 per axis, lists sorted by sub-cell key with start
//   and end per key: cell_list.cpp:103-137, keys as in Update_cell_list,
//   level_infor.cpp:325-344) within int((h CUT_OFF + dsub - 1e-10) / dsub) of its own,
//   adding both ways when h_current > h_neighbour and one way when equal
//   (level_infor.cpp:508-567).  What lands in a ghost's list is dropped: ghosts are
//   rebuilt every step and their lists are never read.
// Returns the real particles' lists (entries: source id and mirror code), sorted.
void neighbors_tree(const Ctx& c, const std::vector<P>& ps, bool scll, std::vector<std::vector<NbEntry>>& nb) {
  const int n = int(ps.size()), D = c.D;
  const double cut = c.cfg->cut_off;
  double hmax = 0., hmin = 1.e300;
  for (const P& p : ps) { hmax = std::max(hmax, p.h); hmin = std::min(hmin, p.h); }
  // Define_level_infor_tight
  double dmin = 1.e300;
  for (int a = 0; a < D; a++) dmin = std::min(dmin, c.cfg->domain[a]);
  const double max0 = hmax * cut, min_scale0 = hmin * cut;
  const double ini = dmin / std::floor(dmin / max0);
  double st = ini;
  int nlev = 1;
  while ((st + 1.e-6) > min_scale0) { st /= 2.; nlev++; }
  nlev -= 1;
  int ini_nc[3];
  for (int a = 0; a < 3; a++) ini_nc[a] = a < D ? int(c.cfg->domain[a] / ini) : 1;
  struct Lev { int nc[3], cs[3], ce[3]; double dc[3], scale; };
  std::vector<Lev> lv(nlev);
  for (int L = 0; L < nlev; L++) {
    Lev& v = lv[L];
    v.scale = 1.e300;
    for (int a = 0; a < 3; a++) {
      const bool on = a < D;
      v.nc[a] = on ? ini_nc[a] << L : 1;
      v.dc[a] = on ? c.cfg->domain[a] / double(v.nc[a]) : 1.;
      if (on) v.scale = std::min(v.scale, v.dc[a]);
      const int dN = (on && c.cfg->sym[a]) ? (1 << L) : 0;
      v.cs[a] = -dN; v.ce[a] = v.nc[a] + dN;
    }
  }
  // an entry of a cell list: a real particle (code 0) or a ghost, at a level
  struct Ent { double x[3]; int src, code, level; double h; bool ghost; };
  std::vector<Ent> ents;
  auto level_of = [&](double h) {
    double scale = ini;
    for (int guard = 0; guard < 64; guard++) {
      if (h * cut > scale + 1.e-15) scale *= 2.;
      else if (h * cut < scale / 2. - 1.e-15) scale /= 2.;
      else break;
    }
    int lev = 0;
    for (int L = nlev - 1; L >= 0; L--)
      if (std::fabs(scale - lv[L].scale) <= 1.e-10) { lev = L; break; }
    return lev;
  };
  auto cell_id = [&](const Lev& v, const double* x, int* ci) {   // get_cell_id, clamped
    for (int a = 0; a < 3; a++) {
      if (a >= D) { ci[a] = 0; continue; }
      int k = int(std::floor((x[a] - c.cfg->box_l[a]) / v.dc[a]));
      ci[a] = std::min(std::max(k, v.cs[a]), v.ce[a] - 1);
    }
  };
  std::vector<int> plev(n);
  for (int i = 0; i < n; i++) {
    plev[i] = level_of(ps[i].h);
    Ent e; for (int a = 0; a < 3; a++) e.x[a] = ps[i].x[a];
    e.src = i; e.code = 0; e.level = plev[i]; e.h = ps[i].h; e.ghost = false;
    ents.push_back(e);
  }
  // ghosts of each level's own particles, from the cells the margin mirrors
  for (int i = 0; i < n; i++) {
    const Lev& v = lv[plev[i]];
    int ci[3]; cell_id(v, ps[i].x, ci);
    int opt[3][2], nopt[3];
    for (int a = 0; a < 3; a++) {
      nopt[a] = 1; opt[a][0] = 0;
      if (a < D && c.cfg->sym[a]) {
        const int dN = -v.cs[a];
        if (ci[a] < dN) opt[a][nopt[a]++] = 1;                    // mirrored into the low margin
        if (ci[a] >= v.nc[a] - dN) opt[a][nopt[a]++] = 2;         // ... the high margin
      }
    }
    for (int o0 = 0; o0 < nopt[0]; o0++)
      for (int o1 = 0; o1 < nopt[1]; o1++)
        for (int o2 = 0; o2 < nopt[2]; o2++) {
          const int code = opt[0][o0] | (opt[1][o1] << 2) | (opt[2][o2] << 4);
          if (!code) continue;
          Ent e; image(c, code, ps[i].x, e.x);
          e.src = i; e.code = code; e.level = plev[i]; e.h = ps[i].h; e.ghost = true;
          ents.push_back(e);
        }
  }
  // the cell lists of every level: entries of that level and finer (a finer entry sits
  // in the coarse cell that holds its finer cell)
  std::vector<std::vector<std::vector<int>>> cells(nlev);
  auto cidx = [&](const Lev& v, const int* ci) -> int64_t {
    return (int64_t(ci[2] - v.cs[2]) * (v.ce[1] - v.cs[1]) + (ci[1] - v.cs[1])) * (v.ce[0] - v.cs[0]) + (ci[0] - v.cs[0]);
  };
  for (int L = 0; L < nlev; L++) {
    const Lev& v = lv[L];
    cells[L].assign(size_t(v.ce[0] - v.cs[0]) * (v.ce[1] - v.cs[1]) * (v.ce[2] - v.cs[2]), {});
    for (size_t q = 0; q < ents.size(); q++) {
      if (ents[q].level < L) continue;
      int ci[3]; cell_id(v, ents[q].x, ci);
      // a ghost lies in the margin; an entry outside the grid's margin is not in a cell
      bool inside = true;
      for (int a = 0; a < D; a++) {
        const int k = int(std::floor((ents[q].x[a] - c.cfg->box_l[a]) / v.dc[a]));
        inside = inside && k >= v.cs[a] && k < v.ce[a];
      }
      if (!inside) continue;
      cells[L][cidx(v, ci)].push_back(int(q));
    }
  }
  // _SCLL_: each cell's list sorted by sub-cell key (a stable sort, as the LSD radix
  // sort of parallelfunc.h:39-96), start and end per key
  const int NS = 3;
  auto subkey = [&](const Lev& v, const double* x, const int* ci) {
    int ps3[3] = {0, 0, 0};
    for (int a = 0; a < D; a++) {
      const double dsub = v.dc[a] / NS;
      const double cs = (x[a] - c.cfg->box_l[a]) - v.dc[a] * ci[a];
      int k = int(std::floor(cs / dsub));
      ps3[a] = std::max(0, std::min(k, NS - 1));
    }
    return (ps3[2] * NS + ps3[1]) * NS + ps3[0];
  };
  if (scll)
    for (int L = 0; L < nlev; L++)
      for (size_t k = 0; k < cells[L].size(); k++) {
        auto& lst = cells[L][k];
        if (lst.empty()) continue;
        const Lev& v = lv[L];
        int ci[3];
        int64_t r = int64_t(k);
        const int wx = v.ce[0] - v.cs[0], wy = v.ce[1] - v.cs[1];
        ci[0] = int(r % wx) + v.cs[0]; r /= wx;
        ci[1] = int(r % wy) + v.cs[1]; r /= wy;
        ci[2] = int(r) + v.cs[2];
        std::stable_sort(lst.begin(), lst.end(), [&](int a, int b) {
          return subkey(v, ents[a].x, ci) < subkey(v, ents[b].x, ci); });
      }
  std::vector<std::vector<NbEntry>> all(ents.size());
  auto add = [&](int to, int from) {   // "to" gains "from" (ghost lists are dropped)
    if (ents[to].ghost) return;
    all[to].push_back({ents[from].src, ents[from].code});
  };
  for (size_t q = 0; q < ents.size(); q++) {
    const Ent& cur = ents[q];
    const int L = cur.level;
    const Lev& v = lv[L];
    int ci[3]; cell_id(v, cur.x, ci);
    if (!scll) {
      const int r[3] = {1, D >= 2 ? 1 : 0, D >= 3 ? 1 : 0};
      for (int t = std::max(ci[0] - r[0], v.cs[0]); t <= std::min(ci[0] + r[0], v.ce[0] - 1); t++)
        for (int u = std::max(ci[1] - r[1], v.cs[1]); u <= std::min(ci[1] + r[1], v.ce[1] - 1); u++)
          for (int w = std::max(ci[2] - r[2], v.cs[2]); w <= std::min(ci[2] + r[2], v.ce[2] - 1); w++) {
            const int cc[3] = {t, u, w};
            for (int e : cells[L][cidx(v, cc)]) {
              const Ent& nbr = ents[e];
              const double d = dist_of(D, cur.x, nbr.x);
              if (nb_pred(d, cur.h, nbr.h, cut)) {
                add(int(q), e);
                if (nbr.level != L) add(e, int(q));
              }
            }
          }
    } else {
      // sub-cell of cur and its search range (level_infor.cpp:510-523)
      int sc[3], rg[3], s0[3], s1[3];
      for (int a = 0; a < 3; a++) {
        if (a >= D) { sc[a] = 0; rg[a] = 0; s0[a] = 0; s1[a] = 1; continue; }
        const double dsub = v.dc[a] / NS;
        const double cs = (cur.x[a] - c.cfg->box_l[a]) - v.dc[a] * ci[a];
        sc[a] = ci[a] * NS + std::max(0, std::min(int(std::floor(cs / dsub)), NS - 1));
        rg[a] = int((cur.h * cut + dsub - 1.e-10) / dsub);
        s0[a] = std::max(sc[a] - rg[a], v.cs[a] * NS);
        s1[a] = std::min(sc[a] + rg[a] + 1, v.ce[a] * NS);
      }
      for (int sx = s0[0]; sx < s1[0]; sx++)
        for (int sy = s0[1]; sy < s1[1]; sy++)
          for (int sz = s0[2]; sz < s1[2]; sz++) {
            const int sj[3] = {sx, sy, sz};
            int tc[3], sub[3] = {0, 0, 0};
            for (int a = 0; a < 3; a++) {
              tc[a] = a < D ? int(std::floor(double(sj[a]) / NS)) : 0;
              if (a < D) sub[a] = sj[a] - NS * tc[a];
            }
            const int key = (sub[2] * NS + sub[1]) * NS + sub[0];
            for (int e : cells[L][cidx(v, tc)]) {
              if (subkey(v, ents[e].x, tc) != key) continue;   // this key's run of the sorted list
              const Ent& nbr = ents[e];
              const double d = dist_of(D, cur.x, nbr.x);
              if (nb_pred(d, cur.h, nbr.h, cut)) {
                if (cur.h > nbr.h) { add(int(q), e); add(e, int(q)); }
                else if (cur.h == nbr.h) add(int(q), e);
              }
            }
          }
    }
  }
  nb.assign(n, {});
  for (int i = 0; i < n; i++) {
    nb[i] = all[i];
    std::sort(nb[i].begin(), nb[i].end(), [](const NbEntry& a, const NbEntry& b) {
      return a.j != b.j ? a.j < b.j : a.code < b.code; });
  }
}

inline double dot(int D, const double* a, const double* b) {   // get_dot_product, glbfunc.h:263-271
  double v = 0.0;
  for (int k = 0; k < D; k++) v += a[k] * b[k];
  return v;
}

// Get_kernel_summation, particle_mesh_generation.cpp:690-765
void kernel_summation(const Ctx& c, std::vector<P>& ps, const std::vector<std::vector<NbEntry>>& nb) {
  int D = c.D, n = int(ps.size());
  double cthr = std::cos(c.cfg->angle_thres);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++) {
    P& p = ps[i];
    double sum = 0.; int ncount = 0;
    for (const NbEntry& e : nb[i]) {
      const P& q = ps[e.j];
      double g[3]; image(c, e.code, q.x, g);
      if (p.type == SMG_REAL) {
        double dr[3]; for (int a = 0; a < 3; a++) dr[a] = p.x[a] - g[a];
        double dist = 0.; for (int a = 0; a < D; a++) dist += dr[a] * dr[a];
        dist = std::sqrt(dist);
        sum += W(D, dist, p.h);
        ncount++;
      } else if (p.type == SMG_SURFACE) {
        double vec[3]; for (int a = 0; a < 3; a++) vec[a] = p.x[a] - g[a];
        double dn = dot(D, vec, p.n);
        double dr[3]; for (int a = 0; a < 3; a++) dr[a] = vec[a] - p.n[a] * dn;
        double dist = 0.; for (int a = 0; a < D; a++) dist += dr[a] * dr[a];
        dist = std::sqrt(dist);
        if (q.type == SMG_SURFACE) {
          double ca = dot(D, p.n, q.n);      // mirror ghosts keep the source normal
          if (std::fabs(ca) > cthr) { sum += W(D - 1, dist, p.h); ncount++; }
        }
      }
    }
    if (ncount > 1) sum = 1. / (sum + 1.e-15);
    p.sum = sum;
  }
}

// Accumulate_particle_force_real (:995-1118) / _surf (:883-991) + Map_surface_particle_force
// (:1122-1126); sigma = 1 (one rank), mass = P = 1, damping coefficient 0.
void accumulate_force(const Ctx& c, std::vector<P>& ps, const std::vector<std::vector<NbEntry>>& nb, int stage) {
  int D = c.D, n = int(ps.size());
  double cthr = std::cos(c.cfg->angle_thres);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++) {
    P& p = ps[i];
    double F1[3] = {0, 0, 0}, F2[3] = {0, 0, 0};
    for (const NbEntry& e : nb[i]) {
      if (e.j == i && e.code == 0) continue;             // id != neighbor->id
      const P& q = ps[e.j];
      bool ghost = e.code != 0;
      double g[3]; image(c, e.code, q.x, g);
      double nu_j = ghost ? 0. : q.nu;                    // ghosts carry nu = 0 (Q6)
      double vj[3];
      for (int a = 0; a < 3; a++) vj[a] = q.v[a];
      if (ghost) for (int a = 0; a < 3; a++) if ((e.code >> (2 * a)) & 3) vj[a] = -q.v[a];
      double rho_j = q.rho, h_j = q.h;
      double vol_i = 1.0 / powern(p.rho, 2), vol_j = 1.0 / powern(rho_j, 2);   // P/rho^2 with P = 1
      double aw, dr[3];
      if (p.type == SMG_REAL) {
        if (q.type != SMG_REAL) for (int a = 0; a < 3; a++) vj[a] = p.v[a] * -1.0;   // Q8
        for (int a = 0; a < 3; a++) dr[a] = p.x[a] - g[a];
        double dist = 0.; for (int a = 0; a < D; a++) dist += dr[a] * dr[a];
        dist = std::sqrt(dist);
        aw = dW(D, dist, 0.5 * (p.h + h_j));
      } else if (p.type == SMG_SURFACE) {
        if (q.type == SMG_REAL) continue;
        double vec[3]; for (int a = 0; a < 3; a++) vec[a] = p.x[a] - g[a];
        double dn = dot(D, vec, p.n);
        for (int a = 0; a < 3; a++) dr[a] = vec[a] - p.n[a] * dn;
        double dist = 0.; for (int a = 0; a < D; a++) dist += dr[a] * dr[a];
        dist = std::sqrt(dist);
        if (q.type == SMG_SURFACE) {
          double ca = dot(D, p.n, q.n);
          if (std::fabs(ca) <= cthr) continue;
        }
        aw = dW(D - 1, dist, 0.5 * (p.h + h_j));
      } else continue;
      if (stage == 1) {
        double Pij = 1. * 1. * p.sum * (vol_i + vol_j);
        for (int a = 0; a < D; a++) F1[a] += dr[a] * aw * -Pij;
      } else {
        double eta_i = p.nu * p.rho, eta_j = nu_j * rho_j;
        double eta_ij = 2. * eta_j * eta_i / (eta_i + eta_j + 1.e-15);
        double temp0 = aw * eta_ij * 1. * (1. / p.rho / p.rho + 1. / rho_j / rho_j);
        for (int a = 0; a < D; a++) F2[a] += (p.v[a] - vj[a]) * temp0;
      }
    }
    for (int a = 0; a < 3; a++) p.a[a] = (a < D) ? F1[a] + F2[a] : 0.;
    if (p.type == SMG_SURFACE) {
      double nf = dot(D, p.a, p.n);
      for (int a = 0; a < D; a++) p.a[a] = p.a[a] - p.n[a] * nf;
    }
  }
}

// Particle_mesh_generation::Set_timestep(stage) (:127-145) + ParallelGetTimestep
// (parallelfunc.cpp:7-58): returns global min dt and max |v|.
void set_timestep(const Ctx& c, std::vector<P>& ps, int stage, double& dt, double& maxv) {
  int D = c.D, n = int(ps.size());
  std::vector<double> ts(n), us(n);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++) {
    P& p = ps[i];
    double U = 0.; for (int a = 0; a < D; a++) U += p.v[a] * p.v[a]; U = std::sqrt(U);
    double A = 0.; for (int a = 0; a < D; a++) A += p.a[a] * p.a[a]; A = std::sqrt(A);
    double t = 0.25 * std::sqrt(p.h / (A + 1.e-15));
    double t2 = p.h / 40. / (U + 1.e-15);
    t = (t <= t2) ? t : t2;
    if (stage == 2) {
      p.nu = c.cfg->nu_coeff * p.h * U;
      double t3 = 0.125 * p.h * p.h / (p.nu + 1.e-15);
      t = (t <= t3) ? t : t3;
    }
    ts[i] = t; us[i] = U;
  }
  dt = 9.223372036854775807e18;   // LONG_MAX
  maxv = 0.0;
  for (int i = 0; i < n; i++) { dt = std::min(dt, ts[i]); maxv = std::max(maxv, us[i]); }
}

struct StepOut {
  double dt1, dt2, maxv, residual, max_disp;
  int64_t n_remapped;
  int ok;
};

void do_map_all(const Ctx& c, std::vector<P>& ps, uint64_t stream, bool scale_after, StepOut& o) {
  int n = int(ps.size());
  int64_t nre = 0; int bad = 0;
#pragma omp parallel for schedule(static) reduction(+ : nre, bad)
  for (int i = 0; i < n; i++) {
    int rm = 0;
    if (!map_particle(c, ps[i], stream, rm)) bad++;
    nre += rm;
    if (scale_after) set_particle_info(c, ps[i]);
    for (int a = 0; a < c.D; a++) ps[i].x[a] = store_round(ps[i].x[a]);
  }
  o.n_remapped = nre;
  if (bad) o.ok = 0;
}

}  // namespace

// ===========================================================================
// C API for the tests (ctypes).  Particle arrays are by id, x/v as n*3 doubles.
extern "C" {

// Registers n refined level-set levels (finest last) for every later call, or clears
// them (n = 0).  The arrays must stay alive while they are registered.
int or_set_levels(int n, const smg_config* cfgs, const double* const* phi, const float* const* nrm,
                  const int8_t* const* tag, const uint8_t* const* active) {
  if (n < 0 || n > 3) return SMG_ERR_ARG;
  for (int l = 0; l < n; l++) g_lreg[l] = LevelReg{cfgs[l], phi[l], nrm[l], tag[l], active[l]};
  g_nlreg = n;
  return SMG_OK;
}

struct or_state {
  int64_t n;
  double* x; double* v; double* h; double* nu; int32_t* type;
  double dt;                 // glbl_timestep after the last Set_timestep(2)
};

double or_kernel_W(int dim, double r, double h) { return W(dim, r, h); }
double or_kernel_dW(int dim, double r, double h) { return dW(dim, r, h); }
void or_rng(uint64_t seed, uint64_t stream, uint64_t id, uint64_t counter, int n, double* out) {
  uint64_t key = rng_key(seed, stream, id);
  for (int i = 0; i < n; i++) out[i] = rng_u(key, counter + uint64_t(i));
}

static void load_particles(const Ctx& c, const or_state* s, std::vector<P>& ps) {
  ps.assign(s->n, P());
  for (int64_t i = 0; i < s->n; i++) {
    P& p = ps[i]; std::memset(&p, 0, sizeof(P));
    for (int a = 0; a < 3; a++) { p.x[a] = a < c.D ? s->x[3 * i + a] : 0.; p.v[a] = a < c.D ? s->v[3 * i + a] : 0.; }
    p.h = s->h[i]; p.nu = s->nu[i]; p.type = s->type[i]; p.id = int(i);
    double dx = p.h * c.cfg->cut_off / c.cfg->cell_ratio;
    (void)dx;
  }
}
static void store_particles(const Ctx& c, const std::vector<P>& ps, or_state* s) {
  for (int64_t i = 0; i < s->n; i++) {
    const P& p = ps[i];
    for (int a = 0; a < 3; a++) { s->x[3 * i + a] = a < c.D ? p.x[a] : 0.; s->v[3 * i + a] = a < c.D ? p.v[a] : 0.; }
    s->h[i] = p.h; s->nu[i] = p.nu; s->type[i] = p.type;
  }
}

// sample the level set at m points
void or_sample(const smg_config* cfg, const double* lphi, const float* lnrm, const int8_t* ltag,
               int64_t m, const double* x, double* phi, double* nrm, int32_t* tag) {
  Ctx c; ctx_init(c, cfg, lphi, lnrm, ltag);
  for (int64_t i = 0; i < m; i++) {
    double xx[3] = {x[3 * i], x[3 * i + 1], x[3 * i + 2]}, n[3];
    sample(c, xx, phi[i], n);
    for (int a = 0; a < 3; a++) nrm[3 * i + a] = n[a];
    tag[i] = tag_at(c, xx);
  }
}
double or_get_scale(const smg_config* cfg, const double* lphi, const float* lnrm, const int8_t* ltag, double phi) {
  Ctx c; ctx_init(c, cfg, lphi, lnrm, ltag);
  return get_scale(c, phi);
}

// Random_particle_distribution (sph_mesh_generation.cpp:1302-1367): positions, h, type
int or_seed(const smg_config* cfg, const double* lphi, const float* lnrm, const int8_t* ltag, or_state* s) {
  Ctx c; ctx_init(c, cfg, lphi, lnrm, ltag);
  int bad = 0;
#pragma omp parallel for schedule(static) reduction(+ : bad)
  for (int64_t i = 0; i < s->n; i++) {
    P p; std::memset(&p, 0, sizeof(P)); p.id = int(i); p.type = SMG_REAL;
    uint64_t key = rng_key(cfg->seed, 0xFFFFFFFFull, uint64_t(i));
    double pad = c.dpkg[0];   // lset_level_info[0]->scale
    bool okk = false;
    for (int att = 0; att < cfg->remap_max_attempts && !okk; att++) {
      for (int a = 0; a < c.D; a++) {
        double box = cfg->particle_box_r[a] - cfg->particle_box_l[a];
        double u0 = rng_u(key, uint64_t(att) * 6 + 2 * a);
        double u1 = rng_u(key, uint64_t(att) * 6 + 2 * a + 1);
        double x = u0 * box + u1 * 1.e-3 + cfg->particle_box_l[a];
        double lo = cfg->particle_box_l[a] + 1.e-10, hi = cfg->particle_box_r[a] - 1.e-10;
        x = (x <= hi) ? x : hi;
        x = (lo >= x) ? lo : x;
        double llo = cfg->box_l[a] + pad, lhi = c.box_r[a] - pad;
        x = (lhi <= x) ? lhi : x;
        x = (llo >= x) ? llo : x;
        p.x[a] = x;
      }
      sample(c, p.x, p.phi, p.n);
      set_particle_info(c, p);
      if (p.phi > p.h) okk = true;
    }
    if (!okk) bad++;
    for (int a = 0; a < 3; a++) { s->x[3 * i + a] = a < c.D ? store_round(p.x[a]) : 0.; s->v[3 * i + a] = 0.; }
    s->h[i] = p.h; s->nu[i] = 0.; s->type[i] = SMG_REAL;
  }
  s->dt = 0.;
  if (g_round_positions) {   // the engine re-samples h at the stored position
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < s->n; i++) {
      P p; std::memset(&p, 0, sizeof(P));
      for (int a = 0; a < 3; a++) p.x[a] = s->x[3 * i + a];
      sample(c, p.x, p.phi, p.n);
      set_particle_info(c, p);
      s->h[i] = p.h;
    }
  }
  return bad ? SMG_ERR_REMAP : SMG_OK;
}

// Pre_run (sph_mesh_generation.cpp:1010-1052): the force passes see sum = 0 and
// nu = 0 and leave a = 0; what remains is the mapping, v = 0, nu = 0 and the dt.
int or_prerun(const smg_config* cfg, const double* lphi, const float* lnrm, const int8_t* ltag, or_state* s) {
  Ctx c; ctx_init(c, cfg, lphi, lnrm, ltag);
  std::vector<P> ps; load_particles(c, s, ps);
  for (P& p : ps) { sample(c, p.x, p.phi, p.n); set_particle_info(c, p); }
  StepOut o{}; o.ok = 1;
  do_map_all(c, ps, 0, false, o);
  for (P& p : ps) { for (int a = 0; a < 3; a++) { p.v[a] = 0.; p.a[a] = 0.; } p.nu = 0.; p.sum = 0.; }
  double dt, mv;
  set_timestep(c, ps, 2, dt, mv);
  s->dt = dt;
  store_particles(c, ps, s);
  return o.ok ? SMG_OK : SMG_ERR_REMAP;
}

// One SPH_run_mesh_generation (sph_mesh_generation.cpp:1371-1463).  step_index is the
// re-seeding RNG stream (1 for the first step after Pre_run).  Optional outputs
// (may be NULL): sum[n], a1[n*3], a2[n*3], diag[6] = {dt1, dt2, max_v, residual, max_disp, res_max},
// nb_offs[n+1] + nb_ids/nb_codes (capacity nb_cap) for the neighbour sets.
int or_step(const smg_config* cfg, const double* lphi, const float* lnrm, const int8_t* ltag,
            or_state* s, int64_t step_index, double* sum_out, double* a1_out, double* a2_out,
            double* diag, int64_t* nb_offs, int32_t* nb_ids, int32_t* nb_codes, int64_t nb_cap,
            int64_t* n_remapped) {
  Ctx c; ctx_init(c, cfg, lphi, lnrm, ltag);
  int D = c.D;
  std::vector<P> ps; load_particles(c, s, ps);
  StepOut o{}; o.ok = 1;
  // Update_coord_full_step
  for (P& p : ps) for (int a = 0; a < D; a++) p.x[a] = p.x[a] + p.v[a] * s->dt;
  // Get_particle_info
  for (P& p : ps) sample(c, p.x, p.phi, p.n);
  // Find_and_map_particle then Set_particle_scale
  do_map_all(c, ps, uint64_t(step_index), true, o);
  // Prepare_for_the_current_step
  std::vector<std::vector<NbEntry>> nb;
  neighbors(c, ps, nb);
  // Reset_particle_force(1): v = 0 every step (v_reini = 1, quirk Q2)
  for (P& p : ps) { for (int a = 0; a < 3; a++) p.v[a] = 0.; p.sum = 0.; }
  kernel_summation(c, ps, nb);
  accumulate_force(c, ps, nb, 1);
  double dt1, mv1; set_timestep(c, ps, 1, dt1, mv1);
  if (a1_out) for (int64_t i = 0; i < s->n; i++) for (int a = 0; a < 3; a++) a1_out[3 * i + a] = ps[i].a[a];
  if (sum_out) for (int64_t i = 0; i < s->n; i++) sum_out[i] = ps[i].sum;
  for (P& p : ps) for (int a = 0; a < D; a++) p.v[a] = p.v[a] + p.a[a] * (0.5 * dt1);
  accumulate_force(c, ps, nb, 2);   // nu from the previous step (Q5)
  double dt2, mv2; set_timestep(c, ps, 2, dt2, mv2);
  if (a2_out) for (int64_t i = 0; i < s->n; i++) for (int a = 0; a < 3; a++) a2_out[3 * i + a] = ps[i].a[a];
  for (P& p : ps) for (int a = 0; a < D; a++) p.v[a] = p.v[a] + p.a[a] * (0.5 * dt2);
  s->dt = dt2;
  // convergence residual: mean over particles of |v| dt2 / dx_i (the reference has
  // none; its closest quantity is the global max |v| of Set_timestep,
  // sph_mesh_generation.cpp:983-998); rmax = the max of the same ratio
  double rsum = 0., rmax = 0., md = 0.;
  for (P& p : ps) {
    double U = 0.; for (int a = 0; a < D; a++) U += p.v[a] * p.v[a]; U = std::sqrt(U);
    double dx = p.h * cfg->cut_off / cfg->cell_ratio;
    const double r = U * dt2 / dx;
    rsum += r; rmax = std::max(rmax, r); md = std::max(md, U * dt2);
  }
  if (diag) {
    diag[0] = dt1; diag[1] = dt2; diag[2] = mv2; diag[3] = ps.empty() ? 0. : rsum / double(ps.size());
    diag[4] = md; diag[5] = rmax;
  }
  if (n_remapped) *n_remapped = o.n_remapped;
  if (nb_offs) {
    int64_t k = 0; nb_offs[0] = 0;
    for (int64_t i = 0; i < s->n; i++) {
      for (const NbEntry& e : nb[i]) {
        if (nb_ids && k < nb_cap) { nb_ids[k] = e.j; nb_codes[k] = e.code; }
        k++;
      }
      nb_offs[i + 1] = k;
    }
  }
  store_particles(c, ps, s);
  return o.ok ? SMG_OK : SMG_ERR_REMAP;
}

// Neighbour sets of given (x, h) with mirror ghosts; brute = 1 for the O(N^2) enumeration,
// 2 / 3 for the restatement of the reference's tree search (default / _SCLL_ build).
int or_neighbors(const smg_config* cfg, int64_t n, const double* x, const double* h, int brute,
                 int64_t* offs, int32_t* ids, int32_t* codes, int64_t cap) {
  Ctx c; ctx_init(c, cfg, nullptr, nullptr, nullptr);
  std::vector<P> ps(n);
  for (int64_t i = 0; i < n; i++) {
    std::memset(&ps[i], 0, sizeof(P));
    for (int a = 0; a < 3; a++) ps[i].x[a] = a < c.D ? x[3 * i + a] : 0.;
    ps[i].h = h[i]; ps[i].id = int(i);
  }
  std::vector<std::vector<NbEntry>> nb;
  if (brute == 1) neighbors_brute(c, ps, nb);
  else if (brute == 2 || brute == 3) neighbors_tree(c, ps, brute == 3, nb);
  else neighbors(c, ps, nb);
  int64_t k = 0; offs[0] = 0;
  for (int64_t i = 0; i < n; i++) {
    for (const NbEntry& e : nb[i]) { if (ids && k < cap) { ids[k] = e.j; codes[k] = e.code; } k++; }
    offs[i + 1] = k;
  }
  return k <= cap || !ids ? SMG_OK : SMG_ERR_CAPACITY;
}

// Pair terms for a given state (x, v, h, nu, type, n) with no mapping: kernel summation,
// stage-1 and stage-2 accelerations; dt1 is the half-step used for v before stage 2.
int or_pair_terms(const smg_config* cfg, int64_t n, const double* x, const double* nrm, const double* h,
                  const double* nu, const int32_t* type, double* sum, double* a1, double* a2, double* dts) {
  Ctx c; ctx_init(c, cfg, nullptr, nullptr, nullptr);
  int D = c.D;
  std::vector<P> ps(n);
  for (int64_t i = 0; i < n; i++) {
    P& p = ps[i]; std::memset(&p, 0, sizeof(P));
    for (int a = 0; a < 3; a++) { p.x[a] = a < D ? x[3 * i + a] : 0.; p.n[a] = a < D ? nrm[3 * i + a] : 0.; }
    p.h = h[i]; p.nu = nu[i]; p.type = type[i]; p.id = int(i);
    double dx = p.h * cfg->cut_off / cfg->cell_ratio;
    p.vol = std::pow(dx, D); p.rho = 1.0 / p.vol;
  }
  std::vector<std::vector<NbEntry>> nb;
  neighbors(c, ps, nb);
  kernel_summation(c, ps, nb);
  accumulate_force(c, ps, nb, 1);
  double dt1, mv; set_timestep(c, ps, 1, dt1, mv);
  for (int64_t i = 0; i < n; i++) { sum[i] = ps[i].sum; for (int a = 0; a < 3; a++) a1[3 * i + a] = ps[i].a[a]; }
  for (P& p : ps) for (int a = 0; a < D; a++) p.v[a] = p.v[a] + p.a[a] * (0.5 * dt1);
  accumulate_force(c, ps, nb, 2);
  double dt2, mv2; set_timestep(c, ps, 2, dt2, mv2);
  for (int64_t i = 0; i < n; i++) for (int a = 0; a < 3; a++) a2[3 * i + a] = ps[i].a[a];
  if (dts) { dts[0] = dt1; dts[1] = dt2; dts[2] = mv2; }
  return SMG_OK;
}

// Stage 2 alone for a given state (x, v, h, nu, type, n), no mapping: the viscous
// accelerations (Accumulate_particle_force(2)), then Set_timestep(2), which also sets
// nu = nu_coeff h |v|.  out_a2[n*3], out_nu[n] (may be NULL), dts = {dt2, max |v|}.
int or_stage2(const smg_config* cfg, int64_t n, const double* x, const double* nrm, const double* h,
              const double* nu, const double* v, const int32_t* type, double* out_a2, double* out_nu,
              double* dts) {
  Ctx c; ctx_init(c, cfg, nullptr, nullptr, nullptr);
  int D = c.D;
  std::vector<P> ps(n);
  for (int64_t i = 0; i < n; i++) {
    P& p = ps[i]; std::memset(&p, 0, sizeof(P));
    for (int a = 0; a < 3; a++) {
      p.x[a] = a < D ? x[3 * i + a] : 0.;
      p.n[a] = a < D ? nrm[3 * i + a] : 0.;
      p.v[a] = a < D ? v[3 * i + a] : 0.;
    }
    p.h = h[i]; p.nu = nu[i]; p.type = type[i]; p.id = int(i);
    double dx = p.h * cfg->cut_off / cfg->cell_ratio;
    p.vol = std::pow(dx, D); p.rho = 1.0 / p.vol;
  }
  std::vector<std::vector<NbEntry>> nb;
  neighbors(c, ps, nb);
  accumulate_force(c, ps, nb, 2);
  double dt2, mv2; set_timestep(c, ps, 2, dt2, mv2);
  for (int64_t i = 0; i < n; i++) {
    for (int a = 0; a < 3; a++) out_a2[3 * i + a] = ps[i].a[a];
    if (out_nu) out_nu[i] = ps[i].nu;
  }
  if (dts) { dts[0] = dt2; dts[1] = mv2; }
  return SMG_OK;
}

int or_num_threads(void) { return omp_get_max_threads(); }
void or_set_position_rounding(int on) { g_round_positions = on ? 1 : 0; }
void or_set_threads(int n) { omp_set_num_threads(n); }

}  // extern "C"
