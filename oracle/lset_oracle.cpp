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
// The cut-cell surface area restates Get_surface_area / get_area_MQ / get_area_tri
// (geometry_calculator.cpp:73-157, :284-345) and Polygonise_MQ::Polygonise /
// VertexInterp (polygonise.cpp:10-95): marching cubes over the 2x2x2 sub-cubes of the
// cell's corner samples (a 2D cell as three equal layers of unit depth), triangle areas
// as 0.5 |v3| |v1| sin(acos(c)) with the host's libm, summed in the reference's order.
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


// ---- marching-cube area (get_area_MQ) ----
// The published marching-cubes triangle table (the one polygonise.h carries), each
// case's triangles as the hex digits of their edge indices, three per triangle.
const char* const kMcTri[256] = {
    "", "083", "019", "183981", "12a", "08312a", "92a029", "2832a8a98",
    "3b2", "0b28b0", "19023b", "1b219b98b", "3a1ba3", "0a108a8ba", "3903b9ba9", "98aa8b",
    "478", "430734", "019847", "419471731", "12a847", "34730412a", "92a902847", "2a9297273794",
    "8473b2", "b47b24204", "90184723b", "47b94b9b2921", "3a13ba784", "1ba14b1047b4", "47890b9bab03", "47b4b99ba",
    "954", "954083", "054150", "854835315", "12a954", "30812a495", "52a542402", "2a5325354348",
    "95423b", "0b208b495", "05401523b", "21525828b485", "a3ba13954", "4950818a18ba", "54050b5bab03", "54858aa8b",
    "978579", "930953573", "078017157", "153357", "978957a12", "a12950530573", "802825857a52", "2a5253357",
    "7957893b2", "95797292027b", "23b018178157", "b21b17715", "958857a13a3b", "5705097b010aba0", "ba0b03a50807570", "ba57b5",
    "a65", "0835a6", "9015a6", "1831985a6", "165261", "165126308", "965906026", "598582526328",
    "23ba65", "b08b20a65", "01923b5a6", "5a61929b298b", "63b653513", "08b0b50515b6", "3b6036065059", "65969bb98",
    "5a6478", "43047365a", "1905a6847", "a65197173794", "612651478", "125526304347", "847905065026", "739794329596269",
    "3b2784a65", "5a647242027b", "01947823b5a6", "9219b294b7b45a6", "8473b53515b6", "51b5b610b7b404b", "059065036b63847", "65969b4797b9",
    "a4964a", "4a649a083", "a01a60640", "83181686461a", "149124264", "308129249264", "024426", "832824426",
    "a49a64b23", "08228b49a4a6", "3b201606461a", "64161a48121b8b1", "964936913b63", "8b1810b61914641", "3b6360064", "648b68",
    "7a678a89a", "0730a709a67a", "a671a7178180", "a67a71173", "126168189867", "269291679093739", "780706602", "732672",
    "23ba68a89867", "20727b09767a9a7", "1801781a767a23b", "b21b17a61671", "896867916b63136", "091b67", "7807063b0b60", "7b6",
    "76b", "308b76", "019b76", "819831b76", "a126b7", "12a3086b7", "2902a96b7", "6b72a3a83a98",
    "723627", "708760620", "276237019", "162186198876", "a76a17137", "a7617a187108", "03707a0a96a7", "76a7a88a9",
    "684b86", "36b306046", "86b846901", "946963931b36", "6846b82a1", "12a30b06b046", "4b846b0292a9", "a93a32943b36463",
    "823842462", "042462", "190234246438", "194142246", "8138618466a1", "a10a06604", "4634386a3039a93", "a946a4",
    "49576b", "083495b76", "50154076b", "b76834354315", "954a1276b", "6b712a083495", "76b54a42a402", "348354325a52b76",
    "723762549", "954086062687", "362376150540", "628687218485158", "954a16176137", "16a176107870954", "40a4a503a6a737a", "76a7a854a48a",
    "6956b9b89", "36b063056095", "0b805b01556b", "6b3635531", "12a95b9b8b56", "0b306b09656912a", "b85b56805a52025", "6b36352a3a53",
    "589528562382", "956960062", "158180568382628", "156216", "13616a386569896", "a10a06950560", "03856a", "a56",
    "b5a75b", "b5ab75830", "5b75ab190", "a75ab7981831", "b12b71751", "08312717572b", "9759279022b7", "75272b592328982",
    "25a235375", "820852875a25", "9015a35373a2", "982921872a25752", "135375", "087071175", "903935537", "987597",
    "5845a8ab8", "5045b05abb30", "01984a8aba45", "ab4a45b34941314", "2512852b8458", "04b0b345b2b151b", "0250592b5458b85", "9452b3",
    "25a352345384", "5a2524420", "3a235a385458019", "5a2524192942", "845853351", "045105", "845853905035", "945",
    "4b749b9ab", "0834979b79ab", "1ab1b414074b", "3143481a474bab4", "4b79b492b912", "9749b791b2b1083", "b74b42240", "b74b42834324",
    "29a279237749", "9a7974a27870207", "37a3a274a1a040a", "1a2874", "491417713", "491417081871", "403743", "487",
    "9a8ab8", "30939bb9a", "01a0a88ab", "31ab3a", "12b1b99b8", "30939b1292b9", "02b80b", "32b",
    "23828aa89", "9a2092", "23828a0181a8", "1a2", "138918", "091", "038", "",
};
// sub-cube vertex v at corner offset kMcV[v] (the reference's numbering: 1 is +y, 3 is
// +x), and the two vertices of edge e
const int kMcV[8][3] = {{0, 0, 0}, {0, 1, 0}, {1, 1, 0}, {1, 0, 0}, {0, 0, 1}, {0, 1, 1}, {1, 1, 1}, {1, 0, 1}};
const int kMcE[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6}, {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};

void mq_interp(const double* p1, const double* p2, double v1, double v2, double* out) {
  const double iso = 0.0;
  if (std::fabs(iso - v1) < 1.0e-15) { for (int c = 0; c < 3; c++) out[c] = p1[c]; return; }
  if (std::fabs(iso - v2) < 1.0e-15) { for (int c = 0; c < 3; c++) out[c] = p2[c]; return; }
  if (std::fabs(v1 - v2) < 1.0e-15) { for (int c = 0; c < 3; c++) out[c] = p1[c]; return; }
  const double mu = (iso - v1) / (v2 - v1);
  for (int c = 0; c < 3; c++) out[c] = p1[c] + mu * (p2[c] - p1[c]);
}

double mq_tri(const double* a, const double* b, const double* c) {
  const double v1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
  const double v3[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
  const double d1 = std::sqrt(v1[0] * v1[0] + v1[1] * v1[1] + v1[2] * v1[2]);
  const double d3 = std::sqrt(v3[0] * v3[0] + v3[1] * v3[1] + v3[2] * v3[2]);
  double area = 0.5 * (d3 * d1) * std::sin(std::acos((v1[0] * v3[0] + v1[1] * v3[1] + v1[2] * v3[2]) / (d3 * d1 + 1.e-20)));
  if (std::isnan(area)) area = 0.;
  return area;
}

double area_mq(const double C[3][3][3], double dx, double dy, double dz) {
  double area = 0.0;
  for (int i = 0; i < 2; i++) for (int j = 0; j < 2; j++) for (int k = 0; k < 2; k++) {
    double p[8][3], val[8];
    int idx = 0;
    for (int v = 0; v < 8; v++) {
      const int vx = i + kMcV[v][0], vy = j + kMcV[v][1], vz = k + kMcV[v][2];
      p[v][0] = vx * 0.5 * dx; p[v][1] = vy * 0.5 * dy; p[v][2] = vz * 0.5 * dz;
      val[v] = C[vx][vy][vz];
      if (val[v] < 0.0) idx |= 1 << v;
    }
    const char* t = kMcTri[idx];
    for (int q = 0; t[q]; q += 3) {
      double tp[3][3];
      for (int s = 0; s < 3; s++) {
        const char ch = t[q + s];
        const int e = ch <= '9' ? ch - '0' : ch - 'a' + 10;
        mq_interp(p[kMcE[e][0]], p[kMcE[e][1]], val[kMcE[e][0]], val[kMcE[e][1]], tp[s]);
      }
      area += mq_tri(tp[0], tp[1], tp[2]);
    }
  }
  return area;
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
                    double ext[3][3][3];
                    for (int r = 0; r < 3; r++) for (int s = 0; s < 3; s++) for (int t = 0; t < 3; t++) ext[r][s][t] = face[r][s];
                    area = area_mq(ext, dx, dy, 1.0);
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
                    area = area_mq(C, dx, dy, dz);
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
