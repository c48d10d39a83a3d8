// lset_oracle.cpp -- FP64 CPU restatement of the level-set finishing steps that feed
// the relaxation path: normals, cut-cell tags, maximum_phi, cut-cell volume/area and
// the particle count.  TEST INFRASTRUCTURE ONLY (checker for smg_lset_build).
//
// Restates: Levelset_package::Get_normal / Get_BC_normal (lset_package.cpp:332-389),
// Get_interface_tag / Is_cut_cell / Get_phi_at_corners (:450-469, :664-761),
// Get_global_scale (:765-785), Calculate_volume_mass (:1283-1396),
// GeometryCalculator::Get_corner / Get_cutcell_volume / get_face / get_subvolume
// (geometry_calculator.cpp:12-68, :166-241, :447-566) and
// SPH_mesh_generation::Get_num_particles (sph_mesh_generation.cpp:900-941).
// The cut-cell surface area is restated with marching squares (2D) and a
// six-tetrahedron split (3D) instead of the marching-cube tables of polygonise.h;
// both converge to the same isosurface area, and the particle count built on it is
// pinned by closed forms (circle perimeter, sphere area), not bit for bit.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#include <numeric>
#include <array>

#include "../include/smg.h"

namespace {

struct G {
  int D, nc[3], E[3], O[3];
  double dcell[3], dpkg[3], dl;
  const double* ext;
  double ph(int i, int j, int k) const {   // global cell index, -4 .. nc+3
    int64_t a = i + 4, b = D >= 2 ? j + 4 : 0, c = D >= 3 ? k + 4 : 0;
    return ext[(c * E[1] + b) * E[0] + a];
  }
  int64_t out(int i, int j, int k) const {
    int64_t a = i + 1, b = D >= 2 ? j + 1 : 0, c = D >= 3 ? k + 1 : 0;
    return (c * O[1] + b) * O[0] + a;
  }
};

inline int sgn(double x) { return (0. < x) - (x < 0.); }

// Get_phi_at_corners + Is_cut_cell
bool is_cut(const G& g, int i, int j, int k) {
  int ip = i + 1, im = i - 1;
  int jp = g.D >= 2 ? j + 1 : j, jm = g.D >= 2 ? j - 1 : j;
  int kp = g.D >= 3 ? k + 1 : k, km = g.D >= 3 ? k - 1 : k;
  auto P = [&](int a, int b, int c) { return g.ph(a, b, c); };
  double v[8][8] = {
      {P(im, jm, km), P(i, jm, km), P(im, j, km), P(i, j, km), P(im, jm, k), P(i, jm, k), P(im, j, k), P(i, j, k)},
      {P(im, jm, kp), P(i, jm, kp), P(im, j, kp), P(i, j, kp), P(im, jm, k), P(i, jm, k), P(im, j, k), P(i, j, k)},
      {P(i, jm, kp), P(ip, jm, kp), P(i, j, kp), P(ip, j, kp), P(i, jm, k), P(ip, jm, k), P(i, j, k), P(ip, j, k)},
      {P(i, jm, km), P(ip, jm, km), P(i, j, km), P(ip, j, km), P(i, jm, k), P(ip, jm, k), P(i, j, k), P(ip, j, k)},
      {P(im, jp, km), P(i, jp, km), P(im, j, km), P(i, j, km), P(im, jp, k), P(i, jp, k), P(im, j, k), P(i, j, k)},
      {P(im, jp, kp), P(i, jp, kp), P(im, j, kp), P(i, j, kp), P(im, jp, k), P(i, jp, k), P(im, j, k), P(i, j, k)},
      {P(i, jp, kp), P(ip, jp, kp), P(i, j, kp), P(ip, j, kp), P(i, jp, k), P(ip, jp, k), P(i, j, k), P(ip, j, k)},
      {P(i, jp, km), P(ip, jp, km), P(i, j, km), P(ip, j, km), P(i, jp, k), P(ip, jp, k), P(i, j, k), P(ip, j, k)}};
  double cv[8];
  for (int u = 0; u < 8; u++) {
    std::vector<double> w(v[u], v[u] + 8);
    std::sort(w.begin(), w.end());
    cv[u] = std::accumulate(w.begin(), w.end(), 0.0);
    cv[u] *= 0.125;
  }
  for (int c = 0; c < 8; c++) if (sgn(cv[c]) != sgn(cv[0])) return true;
  if (sgn(P(i, j, k)) != sgn(cv[0])) return true;
  return false;
}

// geometry_calculator.cpp:376-395, 447-499
int check_face(const double c[2][2]) {
  int f = 0;
  for (int i = 0; i < 2; i++) for (int j = 0; j < 2; j++) { if (c[i][j] > 0.0) f++; else if (c[i][j] < 0.0) f--; }
  return f == 4 ? 1 : (f == -4 ? 0 : -1);
}
double get_face(int lev, double multi, const double cor[2][2]) {
  int flag = check_face(cor);
  if (flag == 1) return multi;
  if (flag == 0) return 0.0;
  double vof = 0.0;
  if (lev <= 5) {
    multi *= 0.25; lev += 1;
    double rf[3][3];
    for (int i = 0; i < 2; i++) for (int j = 0; j < 2; j++) rf[2 * i][2 * j] = cor[i][j];
    for (int i = 0; i < 3; i += 2) rf[i][1] = 0.5 * (rf[i][0] + rf[i][2]);
    for (int j = 0; j < 3; j++) rf[1][j] = 0.5 * (rf[0][j] + rf[2][j]);
    for (int i = 0; i < 2; i++) for (int j = 0; j < 2; j++) {
      double sub[2][2];
      for (int r = 0; r < 2; r++) for (int s = 0; s < 2; s++) sub[r][s] = rf[i + r][j + s];
      vof += get_face(lev, multi, sub);
    }
  } else {
    double cvl = 0.0;
    for (int i = 0; i < 2; i++) for (int j = 0; j < 2; j++) cvl += cor[i][j];
    vof = cvl > 0.0 ? 0.75 * multi : (cvl == 0.0 ? 0.5 * multi : 0.25 * multi);
  }
  return vof;
}
int check_volume(const double c[2][2][2]) {
  int f = 0;
  for (int i = 0; i < 2; i++) for (int j = 0; j < 2; j++) for (int k = 0; k < 2; k++) {
    if (c[i][j][k] > 0.0) f++; else if (c[i][j][k] < 0.0) f--; }
  return f == 8 ? 1 : (f == -8 ? 0 : -1);
}
double get_subvolume(int lev, double multi, const double cor[2][2][2]) {
  int flag = check_volume(cor);
  if (flag == 1) return multi;
  if (flag == 0) return 0.0;
  double vof = 0.0;
  if (lev <= 5) {
    multi *= 0.125; lev += 1;
    double rf[3][3][3];
    for (int i = 0; i < 2; i++) for (int j = 0; j < 2; j++) for (int k = 0; k < 2; k++) rf[2 * i][2 * j][2 * k] = cor[i][j][k];
    for (int i = 0; i < 3; i += 2) for (int j = 0; j < 3; j += 2) rf[i][j][1] = 0.5 * (rf[i][j][0] + rf[i][j][2]);
    for (int k = 0; k < 3; k++) for (int i = 0; i < 3; i += 2) rf[i][1][k] = 0.5 * (rf[i][0][k] + rf[i][2][k]);
    for (int j = 0; j < 3; j++) for (int k = 0; k < 3; k++) rf[1][j][k] = 0.5 * (rf[0][j][k] + rf[2][j][k]);
    for (int i = 0; i < 2; i++) for (int j = 0; j < 2; j++) for (int k = 0; k < 2; k++) {
      double sub[2][2][2];
      for (int r = 0; r < 2; r++) for (int s = 0; s < 2; s++) for (int t = 0; t < 2; t++) sub[r][s][t] = rf[i + r][j + s][k + t];
      vof += get_subvolume(lev, multi, sub);
    }
  } else {
    double cvl = 0.0;
    for (int i = 0; i < 2; i++) for (int j = 0; j < 2; j++) for (int k = 0; k < 2; k++) cvl += cor[i][j][k];
    vof = cvl > 0.0 ? 0.75 * multi : (cvl == 0.0 ? 0.5 * multi : 0.25 * multi);
  }
  return vof;
}

// isoline length in one square (marching squares, linear edge interpolation)
double square_len(const double v[4], const double px[4], const double py[4]) {
  // corners in cyclic order 0-1-2-3; edges (0,1),(1,2),(2,3),(3,0)
  double ex[4], ey[4]; int ne = 0;
  for (int e = 0; e < 4; e++) {
    int a = e, b = (e + 1) & 3;
    bool ia = v[a] < 0.0, ib = v[b] < 0.0;
    if (ia != ib) {
      double t = (0.0 - v[a]) / (v[b] - v[a]);
      ex[ne] = px[a] + t * (px[b] - px[a]); ey[ne] = py[a] + t * (py[b] - py[a]); ne++;
    }
  }
  // lengths as sqrt of the sum of squares (geometry_calculator.cpp:340-342)
  auto len = [](double dx, double dy) { return std::sqrt(dx * dx + dy * dy); };
  double L = 0.;
  if (ne == 2) L = len(ex[1] - ex[0], ey[1] - ey[0]);
  else if (ne == 4) L = len(ex[1] - ex[0], ey[1] - ey[0]) + len(ex[3] - ex[2], ey[3] - ey[2]);
  return L;
}

// isosurface area in one tetrahedron
double tet_area(const double v[4], const double p[4][3]) {
  int in[4], out[4], ni = 0, no = 0;
  for (int a = 0; a < 4; a++) { if (v[a] < 0.0) in[ni++] = a; else out[no++] = a; }
  if (ni == 0 || no == 0) return 0.;
  auto cut = [&](int a, int b, double* q) {
    double t = (0.0 - v[a]) / (v[b] - v[a]);
    for (int k = 0; k < 3; k++) q[k] = p[a][k] + t * (p[b][k] - p[a][k]);
  };
  auto tri = [](const double* A, const double* B, const double* C) {
    double u[3], w[3];
    for (int k = 0; k < 3; k++) { u[k] = B[k] - A[k]; w[k] = C[k] - A[k]; }
    double cx = u[1] * w[2] - u[2] * w[1], cy = u[2] * w[0] - u[0] * w[2], cz = u[0] * w[1] - u[1] * w[0];
    return 0.5 * std::sqrt(cx * cx + cy * cy + cz * cz);
  };
  if (ni == 1 || no == 1) {
    int s = ni == 1 ? in[0] : out[0];
    const int* o = ni == 1 ? out : in;
    double q[3][3];
    for (int k = 0; k < 3; k++) cut(s, o[k], q[k]);
    return tri(q[0], q[1], q[2]);
  }
  double q[4][3];
  cut(in[0], out[0], q[0]); cut(in[0], out[1], q[1]); cut(in[1], out[1], q[2]); cut(in[1], out[0], q[3]);
  return tri(q[0], q[1], q[2]) + tri(q[0], q[2], q[3]);
}

}  // namespace

extern "C" {

// Same contract as smg_lset_build (include/smg.h).
int or_lset_build(const smg_config* cfg, const double* phi_ext, int zero_margin,
                  double* phi, float* normal, int8_t* tag, smg_lset_stats* st) {
  G g; g.D = cfg->dim; g.ext = phi_ext; g.dl = 1.e20;
  for (int a = 0; a < 3; a++) {
    bool on = a < g.D;
    g.nc[a] = on ? 4 * cfg->num_pkg[a] : 1;
    g.E[a] = on ? g.nc[a] + 8 : 1;
    g.O[a] = on ? g.nc[a] + 2 : 1;
    g.dpkg[a] = on ? cfg->domain[a] / double(cfg->num_pkg[a]) : 1.0;
    g.dcell[a] = on ? g.dpkg[a] / 4.0 : 1.0;
    if (on) g.dl = std::min(g.dl, g.dcell[a]);
  }
  int D = g.D;
  int lo[3], hi[3];
  for (int a = 0; a < 3; a++) { lo[a] = a < D ? -1 : 0; hi[a] = a < D ? g.nc[a] + 1 : 1; }
  auto interior = [&](int i, int j, int k) {
    return i >= 0 && i < g.nc[0] && (D < 2 || (j >= 0 && j < g.nc[1])) && (D < 3 || (k >= 0 && k < g.nc[2]));
  };
  for (int k = lo[2]; k < hi[2]; k++)
    for (int j = lo[1]; j < hi[1]; j++)
      for (int i = lo[0]; i < hi[0]; i++) {
        int64_t o = g.out(i, j, k);
        double d[3] = {0., 0., 0.};
        bool in = interior(i, j, k);
        if (in) {
          d[0] = 0.5 * (g.ph(i + 1, j, k) - g.ph(i - 1, j, k)) / g.dcell[0];
          if (D >= 2) d[1] = 0.5 * (g.ph(i, j + 1, k) - g.ph(i, j - 1, k)) / g.dcell[1];
          if (D >= 3) d[2] = 0.5 * (g.ph(i, j, k + 1) - g.ph(i, j, k - 1)) / g.dcell[2];
        } else {
          // Get_BC_normal: one-sided at the faces of the margin package
          int idx[3] = {i, j, k};
          for (int a = 0; a < D; a++) {
            int c = idx[a];
            int loc = c - 4 * int(std::floor(c / 4.0));
            int e[3] = {i, j, k}, f[3] = {i, j, k};
            if (loc == 0) { e[a] = c + 1; d[a] = (g.ph(e[0], e[1], e[2]) - g.ph(i, j, k)) / g.dcell[a]; }
            else if (loc == 3) { f[a] = c - 1; d[a] = (g.ph(i, j, k) - g.ph(f[0], f[1], f[2])) / g.dcell[a]; }
            else { e[a] = c + 1; f[a] = c - 1; d[a] = 0.5 * (g.ph(e[0], e[1], e[2]) - g.ph(f[0], f[1], f[2])) / g.dcell[a]; }
          }
        }
        double dist = 0.; for (int a = 0; a < D; a++) dist += d[a] * d[a];
        dist = std::sqrt(dist) + 1.e-12;
        double inv = 1. / dist;
        bool zero = (!in && zero_margin);
        for (int a = 0; a < 3; a++) normal[3 * o + a] = zero || a >= D ? 0.f : float(d[a] * inv);
        phi[o] = zero ? 0. : g.ph(i, j, k);
        tag[o] = (in && is_cut(g, i, j, k)) ? 1 : 0;
      }
  // maximum_phi over interior cells (Get_global_scale)
  double mphi = 0.; int64_t ncut = 0;
  for (int k = 0; k < (D >= 3 ? g.nc[2] : 1); k++)
    for (int j = 0; j < (D >= 2 ? g.nc[1] : 1); j++)
      for (int i = 0; i < g.nc[0]; i++) {
        double p = g.ph(i, j, k);
        if (p > 0.) mphi = std::max(mphi, std::fabs(p));
        if (tag[g.out(i, j, k)] == 1) ncut++;
      }
  // Calculate_volume_mass with Get_scale of the config (maximum_phi from above if unset)
  smg_config c2 = *cfg;
  if (c2.max_phi <= 0.) c2.max_phi = mphi;
  auto scale_of = [&](double p) {
    if (c2.profile == SMG_PROFILE_LINEAR) {
      if (p > 0. || std::fabs(p) < g.dl) {
        double t = p > 0. ? p / c2.lin_band : 0.; t = t < 1.0 ? t : 1.0;
        return c2.lin_dx0 + t * (c2.lin_dx1 - c2.lin_dx0);
      }
      return c2.lin_dx0;
    }
    if (p > 0. || std::fabs(p) < g.dl)
      return c2.min_dl + std::tanh(2.5 * std::fabs(p) / c2.max_phi) / std::tanh(2.5) * (c2.max_dl - c2.min_dl);
    return c2.min_dl;
  };
  double tv = 0., ta = 0., tm = 0., tms = 0.;
  int np[3] = {cfg->num_pkg[0], D >= 2 ? cfg->num_pkg[1] : 1, D >= 3 ? cfg->num_pkg[2] : 1};
  for (int pi = 0; pi < np[0]; pi++)
    for (int pj = 0; pj < np[1]; pj++)
      for (int pk = 0; pk < np[2]; pk++) {
        double pv = 0., pa = 0., pm = 0., pms = 0.;
        for (int ci = 0; ci < 4; ci++)
          for (int cj = 0; cj < (D >= 2 ? 4 : 1); cj++)
            for (int ck = 0; ck < (D >= 3 ? 4 : 1); ck++) {
              int i = 4 * pi + ci, j = D >= 2 ? 4 * pj + cj : 0, k = D >= 3 ? 4 * pk + ck : 0;
              double p = g.ph(i, j, k);
              if (!(p > 0. || std::fabs(p) < g.dl)) continue;
              double sc = scale_of(p);
              double vol = 0., area = 0.;
              if (tag[g.out(i, j, k)] == 1) {
                int xw = 3, yw = D >= 2 ? 3 : 1, zw = D >= 3 ? 3 : 1;
                double P0[3][3][3] = {}, C[3][3][3] = {};
                for (int r = 0; r < xw; r++) for (int s = 0; s < yw; s++) for (int t = 0; t < zw; t++)
                  P0[r][s][t] = g.ph(i - 1 + r, D >= 2 ? j - 1 + s : 0, D >= 3 ? k - 1 + t : 0);
                int DX = 1, DY = D >= 2 ? 1 : 0, DZ = D >= 3 ? 1 : 0;
                // Get_corner
                C[DX][DY][DZ] = P0[DX][DY][DZ];
                for (int r = 0; r < 1 + DX; r++) for (int s = 0; s < 1 + DY; s++) for (int t = 0; t < 1 + DZ; t++)
                  C[2 * r][2 * s][2 * t] = 0.125 * (P0[r][s][t] + P0[r + DX][s][t] + P0[r][s + DY][t] + P0[r + DX][s + DY][t] +
                                                    P0[r][s][t + DZ] + P0[r + DX][s][t + DZ] + P0[r][s + DY][t + DZ] + P0[r + DX][s + DY][t + DZ]);
                if (DZ == 1)
                  for (int r = 0; r < 1 + DX; r++) for (int s = 0; s < 1 + DY; s++)
                    C[2 * r][2 * s][1] = 0.25 * (P0[r][s][1] + P0[r + DX][s][1] + P0[r][s + DY][1] + P0[r + DX][s + DY][1]);
                for (int s = 0; s < 1 + DY; s++) for (int t = 0; t < 1 + DZ; t++)
                  C[1][2 * s][2 * t] = 0.25 * (P0[1][s][t] + P0[1][s][t + DZ] + P0[1][s + DY][t] + P0[1][s + DY][t + DZ]);
                if (DY == 1)
                  for (int r = 0; r < 1 + DX; r++) for (int t = 0; t < 1 + DZ; t++)
                    C[2 * r][1][2 * t] = 0.25 * (P0[r][1][t] + P0[r + DX][1][t] + P0[r][1][t + DZ] + P0[r + DX][1][t + DZ]);
                if (D == 3)
                  for (int r = 0; r < 2; r++) {
                    C[2 * r][1][1] = 0.5 * (P0[r][1][1] + P0[r + 1][1][1]);
                    C[1][2 * r][1] = 0.5 * (P0[1][r][1] + P0[1][r + 1][1]);
                    C[1][1][2 * r] = 0.5 * (P0[1][1][r] + P0[1][1][r + 1]);
                  }
                double dx = g.dcell[0], dy = D >= 2 ? g.dcell[1] : 1., dz = D >= 3 ? g.dcell[2] : 1.;
                if (D == 2) {
                  double face[3][3];
                  int flag = 0;
                  for (int r = 0; r < 3; r++) for (int s = 0; s < 3; s++) {
                    face[r][s] = C[r][s][0];
                    if (face[r][s] > 0.0) flag++; else if (face[r][s] < 0.0) flag--;
                  }
                  double H;
                  if (flag == 9) H = 1.0;
                  else if (flag == -9) H = 0.0;
                  else {
                    H = 0.;
                    for (int a = 0; a < 2; a++) for (int b = 0; b < 2; b++) {
                      double sub[2][2];
                      for (int r = 0; r < 2; r++) for (int s = 0; s < 2; s++) sub[r][s] = face[a + r][b + s];
                      H += get_face(0, 0.25, sub);
                    }
                    H = std::min(H, 1.0);
                    for (int a = 0; a < 2; a++) for (int b = 0; b < 2; b++) {
                      double v4[4] = {face[a][b], face[a + 1][b], face[a + 1][b + 1], face[a][b + 1]};
                      double px[4] = {a * 0.5 * dx, (a + 1) * 0.5 * dx, (a + 1) * 0.5 * dx, a * 0.5 * dx};
                      double py[4] = {b * 0.5 * dy, b * 0.5 * dy, (b + 1) * 0.5 * dy, (b + 1) * 0.5 * dy};
                      area += square_len(v4, px, py);
                    }
                  }
                  vol = H * dx * dy;
                } else {
                  int flag = 0;
                  for (int r = 0; r < 3; r++) for (int s = 0; s < 3; s++) for (int t = 0; t < 3; t++) {
                    if (C[r][s][t] > 0.0) flag++; else if (C[r][s][t] < 0.0) flag--; }
                  double H;
                  if (flag == 27) H = 1.0;
                  else if (flag == -27) H = 0.0;
                  else {
                    H = 0.;
                    for (int a = 0; a < 2; a++) for (int b = 0; b < 2; b++) for (int e = 0; e < 2; e++) {
                      double sub[2][2][2];
                      for (int r = 0; r < 2; r++) for (int s = 0; s < 2; s++) for (int t = 0; t < 2; t++) sub[r][s][t] = C[a + r][b + s][e + t];
                      H += get_subvolume(0, 0.125, sub);
                    }
                    H = std::min(H, 1.0);
                    static const int tets[6][4] = {{0, 1, 3, 7}, {0, 3, 2, 7}, {0, 2, 6, 7}, {0, 6, 4, 7}, {0, 4, 5, 7}, {0, 5, 1, 7}};
                    for (int a = 0; a < 2; a++) for (int b = 0; b < 2; b++) for (int e = 0; e < 2; e++) {
                      double cv[8], cp[8][3];
                      for (int q = 0; q < 8; q++) {
                        int qx = q & 1, qy = (q >> 1) & 1, qz = (q >> 2) & 1;
                        cv[q] = C[a + qx][b + qy][e + qz];
                        cp[q][0] = (a + qx) * 0.5 * dx; cp[q][1] = (b + qy) * 0.5 * dy; cp[q][2] = (e + qz) * 0.5 * dz;
                      }
                      for (int t = 0; t < 6; t++) {
                        double tv4[4], tp[4][3];
                        for (int q = 0; q < 4; q++) { tv4[q] = cv[tets[t][q]]; for (int m = 0; m < 3; m++) tp[q][m] = cp[tets[t][q]][m]; }
                        area += tet_area(tv4, tp);
                      }
                    }
                  }
                  vol = H * dx * dy * dz;
                }
                pa += area;
                double sp = 1.0; for (int q = 0; q < D - 1; q++) sp *= sc;
                pms += area / sp;
              } else {
                vol = g.dcell[0] * (D >= 2 ? g.dcell[1] : 1.) * (D >= 3 ? g.dcell[2] : 1.);
              }
              double sd = 1.0; for (int q = 0; q < D; q++) sd *= sc;
              pm += vol / sd; pv += vol;
            }
        tv += pv; ta += pa; tm += pm; tms += pms;
      }
  if (st) {
    st->maximum_phi = mphi; st->total_volume = tv; st->total_area = ta;
    st->total_mass = tm; st->total_mass_surface = tms; st->n_interface_cell = ncut;
    int64_t nsurf = std::max<int64_t>(int64_t(tms * 1. - 0.5 * 0.), 0);
    int64_t nmesh = std::max<int64_t>(int64_t(tm * 1. - 0.5 * tms * 1.), 0);
    if (ncut < nsurf) nsurf = ncut;
    st->num_particle_mesh = nmesh; st->num_particle_surface = nsurf; st->num_particle = nmesh + nsurf;
  }
  return SMG_OK;
}

}  // extern "C"
