// smg_lset.cpp -- host-side level-set finishing for the relaxation engine.
//
// Input is the cell-centred signed distance on the package grid including one full
// margin package per side (the layout Levelset_levelinfo::Initialize allocates,
// lset_level_infor.cpp:17-117).  Output is what the device kernels sample: phi,
// the normal and the cut-cell tag on the grid with a one-cell halo, plus the
// scalars the relaxation needs (maximum_phi, particle count).
//
//  normal   Levelset_package::Get_normal / Get_BC_normal   lset_package.cpp:332-389
//  tag      Get_interface_tag / Is_cut_cell / corners       lset_package.cpp:450-469, 664-761
//  max phi  Levelset_package::Get_global_scale              lset_package.cpp:765-785
//  volume   Calculate_volume_mass + GeometryCalculator       lset_package.cpp:1283-1396,
//                                                            geometry_calculator.cpp:12-566
//  count    SPH_mesh_generation::Get_num_particles           sph_mesh_generation.cpp:900-941
//
// The margin packages are either kept (zero_margin = 0, first pass of the reference)
// or zeroed (zero_margin = 1): after Load_lset_restart_file only interior packages
// are read back, so margin phi and normals are zero in the relaxation run (quirk Q9);
// the tags always come from the first pass, as in the restart file.
// Cut-cell surface area: the reference's marching cubes over the cell's 2x2x2
// sub-cubes (Get_surface_area / get_area_MQ, geometry_calculator.cpp:73-157, :284-345),
// a 2D cell as three equal layers of unit depth; smg_mc.hpp holds the form shared with
// the GPU builder.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <thread>
#include <vector>

#include "smg.h"
#include "smg_internal.hpp"
#include "smg_mc.hpp"

namespace smg {
namespace {

struct Grid {
  int D;
  int nc[3], ext[3], halo[3];
  double dcell[3];
  const double* phi_ext;
  inline double at(int i, int j, int k) const {
    const int64_t a = i + 4, b = (D > 1) ? j + 4 : 0, c = (D > 2) ? k + 4 : 0;
    return phi_ext[(c * ext[1] + b) * ext[0] + a];
  }
  inline int64_t hidx(int i, int j, int k) const {
    const int64_t a = i + 1, b = (D > 1) ? j + 1 : 0, c = (D > 2) ? k + 1 : 0;
    return (c * halo[1] + b) * halo[0] + a;
  }
  inline bool inside(int i, int j, int k) const {
    return i >= 0 && i < nc[0] && (D < 2 || (j >= 0 && j < nc[1])) && (D < 3 || (k >= 0 && k < nc[2]));
  }
};

inline int sign3(double v) { return v > 0. ? 1 : (v < 0. ? -1 : 0); }

// corner value = mean of the 8 (2D: 4, each twice) cells around the corner, summed in
// ascending order as Get_phi_at_corners does
double corner_value(const Grid& g, int i, int j, int k, int sx, int sy, int sz) {
  double v[8];
  int n = 0;
  const int zs[2] = {k, (g.D > 2) ? k + sz : k};
  const int ys[2] = {(g.D > 1) ? j + sy : j, j};
  const int xs[2] = {i + sx, i};
  // order within the 8-vector does not matter: it is sorted before summing
  for (int c = 0; c < 2; c++)
    for (int b = 0; b < 2; b++)
      for (int a = 0; a < 2; a++) v[n++] = g.at(xs[a], ys[b], zs[c]);
  std::sort(v, v + 8);
  double s = 0.0;
  for (int q = 0; q < 8; q++) s += v[q];
  return s * 0.125;
}

// all 3^D cells around (i, j, k) strictly of one sign -> every corner average has that
// sign too, so the cell cannot be cut
bool same_sign_block(const Grid& g, int i, int j, int k) {
  const int s0 = sign3(g.at(i, j, k));
  if (s0 == 0) return false;
  const int zr = g.D > 2 ? 1 : 0, yr = g.D > 1 ? 1 : 0;
  for (int c = -zr; c <= zr; c++)
    for (int b = -yr; b <= yr; b++)
      for (int a = -1; a <= 1; a++)
        if (sign3(g.at(i + a, j + b, k + c)) != s0) return false;
  return true;
}

bool cut_cell(const Grid& g, int i, int j, int k) {
  if (same_sign_block(g, i, j, k)) return false;
  // the corner order of Get_phi_at_corners: (sx, sy, sz) for u = 0..7
  static const int cs[8][3] = {{-1, -1, -1}, {-1, -1, 1}, {1, -1, 1}, {1, -1, -1},
                               {-1, 1, -1}, {-1, 1, 1}, {1, 1, 1}, {1, 1, -1}};
  const int s0 = sign3(corner_value(g, i, j, k, cs[0][0], cs[0][1], cs[0][2]));
  for (int u = 1; u < 8; u++)
    if (sign3(corner_value(g, i, j, k, cs[u][0], cs[u][1], cs[u][2])) != s0) return true;
  return sign3(g.at(i, j, k)) != s0;
}

int face_flag4(const double c[2][2]) {
  int f = 0;
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++) f += (c[a][b] > 0.0) - (c[a][b] < 0.0);
  return f;
}

// adaptive sub-face area fraction, GeometryCalculator::get_face
double face_fraction(int lev, double w, const double c[2][2]) {
  const int f = face_flag4(c);
  if (f == 4) return w;
  if (f == -4) return 0.0;
  if (lev > 5) {
    double s = 0.0;
    for (int a = 0; a < 2; a++)
      for (int b = 0; b < 2; b++) s += c[a][b];
    return s > 0.0 ? 0.75 * w : (s == 0.0 ? 0.5 * w : 0.25 * w);
  }
  double r[3][3];
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++) r[2 * a][2 * b] = c[a][b];
  for (int a = 0; a < 3; a += 2) r[a][1] = 0.5 * (r[a][0] + r[a][2]);
  for (int b = 0; b < 3; b++) r[1][b] = 0.5 * (r[0][b] + r[2][b]);
  double acc = 0.0;
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++) {
      const double s[2][2] = {{r[a][b], r[a][b + 1]}, {r[a + 1][b], r[a + 1][b + 1]}};
      acc += face_fraction(lev + 1, w * 0.25, s);
    }
  return acc;
}

// adaptive sub-cell volume fraction, GeometryCalculator::get_subvolume
double cube_fraction(int lev, double w, const double c[2][2][2]) {
  int f = 0;
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++)
      for (int e = 0; e < 2; e++) f += (c[a][b][e] > 0.0) - (c[a][b][e] < 0.0);
  if (f == 8) return w;
  if (f == -8) return 0.0;
  if (lev > 5) {
    double s = 0.0;
    for (int a = 0; a < 2; a++)
      for (int b = 0; b < 2; b++)
        for (int e = 0; e < 2; e++) s += c[a][b][e];
    return s > 0.0 ? 0.75 * w : (s == 0.0 ? 0.5 * w : 0.25 * w);
  }
  double r[3][3][3];
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++)
      for (int e = 0; e < 2; e++) r[2 * a][2 * b][2 * e] = c[a][b][e];
  for (int a = 0; a < 3; a += 2)
    for (int b = 0; b < 3; b += 2) r[a][b][1] = 0.5 * (r[a][b][0] + r[a][b][2]);
  for (int e = 0; e < 3; e++)
    for (int a = 0; a < 3; a += 2) r[a][1][e] = 0.5 * (r[a][0][e] + r[a][2][e]);
  for (int b = 0; b < 3; b++)
    for (int e = 0; e < 3; e++) r[1][b][e] = 0.5 * (r[0][b][e] + r[2][b][e]);
  double acc = 0.0;
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++)
      for (int e = 0; e < 2; e++) {
        double s[2][2][2];
        for (int p = 0; p < 2; p++)
          for (int q = 0; q < 2; q++)
            for (int t = 0; t < 2; t++) s[p][q][t] = r[a + p][b + q][e + t];
        acc += cube_fraction(lev + 1, w * 0.125, s);
      }
  return acc;
}

constexpr uint64_t kMcTable[256] = SMG_MC_TRI_TABLE;

// corner / edge / face / centre samples of a cut cell from its 3^D neighbourhood,
// GeometryCalculator::Get_corner
void cell_samples(const Grid& g, int i, int j, int k, double C[3][3][3]) {
  const int D = g.D, dy = D > 1, dz = D > 2;
  double P[3][3][3] = {};
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < (dy ? 3 : 1); b++)
      for (int e = 0; e < (dz ? 3 : 1); e++) P[a][b][e] = g.at(i - 1 + a, dy ? j - 1 + b : 0, dz ? k - 1 + e : 0);
  std::memset(C, 0, sizeof(double) * 27);
  C[1][dy][dz] = P[1][dy][dz];
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 1 + dy; b++)
      for (int e = 0; e < 1 + dz; e++)
        C[2 * a][2 * b][2 * e] = 0.125 * (P[a][b][e] + P[a + 1][b][e] + P[a][b + dy][e] + P[a + 1][b + dy][e] +
                                          P[a][b][e + dz] + P[a + 1][b][e + dz] + P[a][b + dy][e + dz] +
                                          P[a + 1][b + dy][e + dz]);
  if (dz)
    for (int a = 0; a < 2; a++)
      for (int b = 0; b < 2; b++) C[2 * a][2 * b][1] = 0.25 * (P[a][b][1] + P[a + 1][b][1] + P[a][b + 1][1] + P[a + 1][b + 1][1]);
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

void cut_cell_measure(const Grid& g, int i, int j, int k, double& vol, double& area) {
  double C[3][3][3];
  cell_samples(g, i, j, k, C);
  const double dx = g.dcell[0], dy = g.D > 1 ? g.dcell[1] : 1.0, dz = g.D > 2 ? g.dcell[2] : 1.0;
  area = 0.0;
  if (g.D == 2) {
    int f = 0;
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) f += (C[a][b][0] > 0.0) - (C[a][b][0] < 0.0);
    double H = f == 9 ? 1.0 : 0.0;
    if (f != 9 && f != -9) {
      H = 0.0;
      for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) {
          const double s[2][2] = {{C[a][b][0], C[a][b + 1][0]}, {C[a + 1][b][0], C[a + 1][b + 1][0]}};
          H += face_fraction(0, 0.25, s);
        }
      H = std::min(H, 1.0);
      double E[3][3][3];
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++)
          for (int e = 0; e < 3; e++) E[a][b][e] = C[a][b][0];
      area = mc_cell_area(E, dx, dy, 1.0, kMcTable);
    }
    vol = H * dx * dy;
    return;
  }
  int f = 0;
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++)
      for (int e = 0; e < 3; e++) f += (C[a][b][e] > 0.0) - (C[a][b][e] < 0.0);
  double H = f == 27 ? 1.0 : 0.0;
  if (f != 27 && f != -27) {
    H = 0.0;
    for (int a = 0; a < 2; a++)
      for (int b = 0; b < 2; b++)
        for (int e = 0; e < 2; e++) {
          double s[2][2][2];
          for (int q = 0; q < 8; q++) {
            const int qx = q & 1, qy = (q >> 1) & 1, qz = (q >> 2) & 1;
            s[qx][qy][qz] = C[a + qx][b + qy][e + qz];
          }
          H += cube_fraction(0, 0.125, s);
        }
    H = std::min(H, 1.0);
    area = mc_cell_area(C, dx, dy, dz, kMcTable);
  }
  vol = H * dx * dy * dz;
}

template <typename F>
void parallel_range(int lo, int hi, F fn) {
  const int n = hi - lo;
  if (n <= 0) return;
  unsigned nt = std::thread::hardware_concurrency();
  if (nt == 0) nt = 1;
  nt = std::min<unsigned>(nt, 32u);
  nt = std::min<unsigned>(nt, unsigned(n));
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; t++)
    th.emplace_back([=, &fn]() {
      for (int q = lo + int(t); q < hi; q += int(nt)) fn(q);
    });
  for (auto& x : th) x.join();
}

}  // namespace

int lset_build(const smg_config* cfg, const double* phi_ext, int zero_margin, double* phi, float* normal,
               int8_t* tag, smg_lset_stats* st) {
  Grid g;
  g.D = cfg->dim;
  g.phi_ext = phi_ext;
  double dl = 1.e20;
  for (int a = 0; a < 3; a++) {
    const bool on = a < g.D;
    g.nc[a] = on ? 4 * cfg->num_pkg[a] : 1;
    g.ext[a] = on ? g.nc[a] + 8 : 1;
    g.halo[a] = on ? g.nc[a] + 2 : 1;
    const double dpkg = on ? cfg->domain[a] / double(cfg->num_pkg[a]) : 1.0;
    g.dcell[a] = on ? dpkg / 4.0 : 1.0;
    if (on) dl = std::min(dl, g.dcell[a]);
  }
  const int D = g.D;
  const int klo = D > 2 ? -1 : 0, khi = D > 2 ? g.nc[2] + 1 : 1;
  const int jlo = D > 1 ? -1 : 0, jhi = D > 1 ? g.nc[1] + 1 : 1;
  // slices are independent: each cell's outputs depend only on the input grid
  parallel_range(D > 2 ? klo : jlo, D > 2 ? khi : jhi, [&](int outer) {
  for (int k = (D > 2 ? outer : 0); k < (D > 2 ? outer + 1 : 1); k++)
    for (int j = (D > 2 ? jlo : outer); j < (D > 2 ? jhi : outer + 1); j++)
      for (int i = -1; i < g.nc[0] + 1; i++) {
        const int64_t o = g.hidx(i, j, k);
        const bool in = g.inside(i, j, k);
        const int id3[3] = {i, j, k};
        double grad[3] = {0.0, 0.0, 0.0};
        for (int a = 0; a < D; a++) {
          int lo3[3] = {i, j, k}, hi3[3] = {i, j, k};
          double w = 0.5;
          if (!in) {  // margin package: one-sided at its own faces
            const int c = id3[a];
            const int local = ((c % 4) + 4) % 4;
            if (local == 0) { hi3[a] = c + 1; w = 1.0; }
            else if (local == 3) { lo3[a] = c - 1; w = 1.0; }
            else { lo3[a] = c - 1; hi3[a] = c + 1; }
          } else {
            lo3[a] = id3[a] - 1;
            hi3[a] = id3[a] + 1;
          }
          const double diff = g.at(hi3[0], hi3[1], hi3[2]) - g.at(lo3[0], lo3[1], lo3[2]);
          grad[a] = (w == 0.5 ? 0.5 * diff : diff) / g.dcell[a];
        }
        double len = 0.0;
        for (int a = 0; a < D; a++) len += grad[a] * grad[a];
        len = std::sqrt(len) + 1.e-12;
        const double inv = 1. / len;
        const bool blank = !in && zero_margin;
        for (int a = 0; a < 3; a++) normal[3 * o + a] = (blank || a >= D) ? 0.0f : float(grad[a] * inv);
        phi[o] = blank ? 0.0 : g.at(i, j, k);
        tag[o] = (in && cut_cell(g, i, j, k)) ? 1 : 0;
      }
  });

  double max_phi = 0.0;
  int64_t n_cut = 0;
  for (int k = 0; k < (D > 2 ? g.nc[2] : 1); k++)
    for (int j = 0; j < (D > 1 ? g.nc[1] : 1); j++)
      for (int i = 0; i < g.nc[0]; i++) {
        const double p = g.at(i, j, k);
        if (p > 0.) max_phi = std::max(max_phi, std::fabs(p));
        n_cut += tag[g.hidx(i, j, k)] == 1;
      }

  SpacingProfile prof(*cfg, dl, cfg->max_phi > 0.0 ? cfg->max_phi : max_phi);
  double t_vol = 0.0, t_area = 0.0, t_mass = 0.0, t_msurf = 0.0;
  const int npk[3] = {cfg->num_pkg[0], D > 1 ? cfg->num_pkg[1] : 1, D > 2 ? cfg->num_pkg[2] : 1};
  // packages in the order Initial_pkg_list_and_reset_tags pushes them (k fastest),
  // cells inside a package with k fastest, as Calculate_volume_mass loops; package
  // sums are formed in parallel and added up in that order
  std::vector<double> psum(size_t(npk[0]) * npk[1] * npk[2] * 4, 0.0);
  parallel_range(0, npk[0], [&](int pi) {
  for (int pj = 0; pj < npk[1]; pj++)
      for (int pk = 0; pk < npk[2]; pk++) {
        double pv = 0.0, pa = 0.0, pm = 0.0, pms = 0.0;
        for (int ci = 0; ci < 4; ci++)
          for (int cj = 0; cj < (D > 1 ? 4 : 1); cj++)
            for (int ck = 0; ck < (D > 2 ? 4 : 1); ck++) {
              const int i = 4 * pi + ci, j = D > 1 ? 4 * pj + cj : 0, k = D > 2 ? 4 * pk + ck : 0;
              const double p = g.at(i, j, k);
              if (!(p > 0. || std::fabs(p) < dl)) continue;
              const double sc = prof.dx(p);
              double vol, area = 0.0;
              if (tag[g.hidx(i, j, k)] == 1) {
                cut_cell_measure(g, i, j, k, vol, area);
                pa += area;
                double sp = 1.0;
                for (int q = 0; q < D - 1; q++) sp *= sc;
                pms += area / sp;
              } else {
                vol = g.dcell[0] * (D > 1 ? g.dcell[1] : 1.0) * (D > 2 ? g.dcell[2] : 1.0);
              }
              double sd = 1.0;
              for (int q = 0; q < D; q++) sd *= sc;
              pm += vol / sd;
              pv += vol;
            }
        double* ps = &psum[((size_t(pi) * npk[1] + pj) * npk[2] + pk) * 4];
        ps[0] = pv;
        ps[1] = pa;
        ps[2] = pm;
        ps[3] = pms;
      }
  });
  for (size_t q = 0; q < psum.size(); q += 4) {
    t_vol += psum[q];
    t_area += psum[q + 1];
    t_mass += psum[q + 2];
    t_msurf += psum[q + 3];
  }
  if (st) {
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
  return SMG_OK;
}

}  // namespace smg
