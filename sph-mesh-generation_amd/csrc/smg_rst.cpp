// smg_rst.cpp -- the reference's level-set restart file (./rstfile/levelset.rst): the
// input of its relaxation run.  Layout, all native-endian binary (Real = double):
//
//  case header      SPH_mesh_generation::Load_rst_case   sph_mesh_generation.cpp:341-373
//    int DIM; Real domain[3], box_l[3], box_r[3], particle_box[3], particle_box_l[3],
//    particle_box_r[3]; int resolution[3], Lmin, Lmax, ini_num_cell[3], num_level
//  level-set header Levelset::Load_lset_restart_file    level_set.cpp:1098-1142
//    Real convergence_error; int num_singularity; Real xyz[3] x num_singularity;
//    int num_segment; Real xyz[6] x num_segment; Real fac_maximum_dl, fac_middel_dl,
//    fac_minimum_dl, maximum_curv_artificial, total_mass, total_mass_surface,
//    total_mass_segment, total_volume, total_area, total_length, maximum_phi,
//    maximum_psi, maximum_dl, middel_dl, minimum_dl, maximum_curv, minimum_curv;
//    int glbl_num_interface_cell
//  level info       Levelset_levelinfo::Load_level_set_rstfile  lset_level_infor.cpp:932-981
//    int glbl_total_num_pkg; Real scale, dl; int num_pkg[3], glbl_num_pkg[3],
//    glbl_pkg_start[3], glbl_pkg_end[3], pkg_start[3], pkg_end[3], num_pkg_margin[3];
//    Real dpkg[3], dcell[3]
//  cell fields      lset_level_infor.cpp:983-1146, each over the interior cells, x fastest
//    Real phi, n_x, n_y, n_z, curv, psi; int tag_interface, tag_characteristic,
//    idx_characteristic; Real scale
//
// After a restart the reference holds only the interior cells; the margin packages
// stay zero (quirk Q9).  The reader fills the engine's one-cell-halo arrays the same
// way and returns phi also on the full-margin grid of smg_lset_build (margins zero),
// so that the particle count of a restart run can be recomputed as the reference's
// restart mode does.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "smg.h"
#include "smg_internal.hpp"

namespace smg {
namespace {

constexpr int kCutCell = 1;     // CUT_CELL, glbparam.h:49
constexpr int kNormalCell = 0;  // NORMAL_CELL, glbparam.h:48

struct Reader {
  FILE* f = nullptr;
  bool ok = true;
  template <typename T>
  T get() {
    T v{};
    if (ok && std::fread(&v, sizeof(T), 1, f) != 1) ok = false;
    return v;
  }
  void skip(size_t bytes) {
    if (ok && std::fseek(f, long(bytes), SEEK_CUR) != 0) ok = false;
  }
};

struct Writer {
  FILE* f = nullptr;
  bool ok = true;
  template <typename T>
  void put(T v) {
    if (ok && std::fwrite(&v, sizeof(T), 1, f) != 1) ok = false;
  }
};

struct FileGuard {
  FILE* f;
  ~FileGuard() {
    if (f) std::fclose(f);
  }
};

}  // namespace
}  // namespace smg

extern "C" {

int smg_lset_read_rst(const char* path, smg_config* cfg, smg_rst_info* info, double* phi, float* normal,
                      int8_t* tag, double* phi_ext) {
  using namespace smg;
  if (!path || !cfg || !info) { set_error("smg_lset_read_rst: NULL argument"); return SMG_ERR_ARG; }
  FILE* fp = std::fopen(path, "rb");
  if (!fp) { set_error(std::string("smg_lset_read_rst: cannot open ") + path); return SMG_ERR_ARG; }
  FileGuard guard{fp};
  Reader r;
  r.f = fp;
  std::memset(info, 0, sizeof(*info));
  const int D = r.get<int>();
  if (!r.ok || (D != 2 && D != 3)) { set_error("smg_lset_read_rst: bad dimension in the case header"); return SMG_ERR_ARG; }
  {
    const int rc = smg_default_config(D, cfg);
    if (rc != SMG_OK) return rc;
  }
  double domain[3], box_l[3], box_r[3], pbox[3], pbox_l[3], pbox_r[3];
  for (double& v : domain) v = r.get<double>();
  for (double& v : box_l) v = r.get<double>();
  for (double& v : box_r) v = r.get<double>();
  for (double& v : pbox) v = r.get<double>();
  for (double& v : pbox_l) v = r.get<double>();
  for (double& v : pbox_r) v = r.get<double>();
  for (int a = 0; a < 3; a++) info->resolution[a] = r.get<int>();
  info->Lmin = r.get<int>();
  info->Lmax = r.get<int>();
  for (int a = 0; a < 3; a++) info->ini_num_cell[a] = r.get<int>();
  info->num_level = r.get<int>();
  // level-set header
  info->convergence_error = r.get<double>();
  info->num_singularity = r.get<int>();
  if (!r.ok || info->num_singularity < 0) { set_error("smg_lset_read_rst: truncated level-set header"); return SMG_ERR_ARG; }
  r.skip(size_t(info->num_singularity) * 3 * sizeof(double));
  info->num_segment = r.get<int>();
  if (!r.ok || info->num_segment < 0) { set_error("smg_lset_read_rst: truncated level-set header"); return SMG_ERR_ARG; }
  r.skip(size_t(info->num_segment) * 6 * sizeof(double));
  double* hv[] = {&info->fac_maximum_dl, &info->fac_middel_dl, &info->fac_minimum_dl,
                  &info->maximum_curv_artificial, &info->total_mass, &info->total_mass_surface,
                  &info->total_mass_segment, &info->total_volume, &info->total_area, &info->total_length,
                  &info->maximum_phi, &info->maximum_psi, &info->maximum_dl, &info->middel_dl, &info->minimum_dl,
                  &info->maximum_curv, &info->minimum_curv};
  for (double* p : hv) *p = r.get<double>();
  info->glbl_num_interface_cell = r.get<int>();
  // level info
  r.get<int>();   // glbl_total_num_pkg
  info->scale = r.get<double>();
  info->dl = r.get<double>();
  int num_pkg[3], skip6[18];
  for (int& v : num_pkg) v = r.get<int>();
  for (int& v : skip6) v = r.get<int>();   // glbl_num_pkg .. num_pkg_margin
  double dpkg[3], dcell[3];
  for (double& v : dpkg) v = r.get<double>();
  for (double& v : dcell) v = r.get<double>();
  if (!r.ok) { set_error("smg_lset_read_rst: truncated file header"); return SMG_ERR_ARG; }
  for (int a = 0; a < 3; a++) {
    const bool on = a < D;
    if (on && (num_pkg[a] <= 0 || num_pkg[a] > 4096)) { set_error("smg_lset_read_rst: bad package count"); return SMG_ERR_ARG; }
    cfg->box_l[a] = box_l[a];
    cfg->domain[a] = domain[a];
    cfg->particle_box_l[a] = pbox_l[a];
    cfg->particle_box_r[a] = pbox_r[a];
    cfg->num_pkg[a] = on ? num_pkg[a] : 1;
  }
  // the spacing profile the file carries (Levelset::Get_scale, level_set.cpp:311-351)
  cfg->profile = SMG_PROFILE_TANH;
  cfg->min_dl = info->minimum_dl;
  cfg->max_dl = info->maximum_dl;
  cfg->max_phi = info->maximum_phi;
  const int nc[3] = {4 * cfg->num_pkg[0], D > 1 ? 4 * cfg->num_pkg[1] : 1, D > 2 ? 4 * cfg->num_pkg[2] : 1};
  const int halo[3] = {nc[0] + 2, D > 1 ? nc[1] + 2 : 1, D > 2 ? nc[2] + 2 : 1};
  const int ext[3] = {nc[0] + 8, D > 1 ? nc[1] + 8 : 1, D > 2 ? nc[2] + 8 : 1};
  info->n_halo_cells = int64_t(halo[0]) * halo[1] * halo[2];
  info->n_ext_cells = int64_t(ext[0]) * ext[1] * ext[2];
  info->n_cells = int64_t(nc[0]) * nc[1] * nc[2];
  {
    // the cell fields fill the rest of the file exactly: 7 Real and 3 int per cell
    const long here = std::ftell(fp);
    if (here < 0 || std::fseek(fp, 0, SEEK_END) != 0) { set_error("smg_lset_read_rst: cannot seek"); return SMG_ERR_ARG; }
    const long end = std::ftell(fp);
    if (std::fseek(fp, here, SEEK_SET) != 0) { set_error("smg_lset_read_rst: cannot seek"); return SMG_ERR_ARG; }
    const int64_t want = info->n_cells * int64_t(7 * sizeof(double) + 3 * sizeof(int32_t));
    if (int64_t(end - here) != want) {
      set_error("smg_lset_read_rst: the cell fields do not match the header (truncated or foreign file)");
      return SMG_ERR_ARG;
    }
  }
  if (!phi && !normal && !tag && !phi_ext) return SMG_OK;   // sizes only
  const int64_t nin = info->n_cells;
  std::vector<double> fld(static_cast<size_t>(nin) * 6);
  std::vector<int32_t> itag(static_cast<size_t>(nin));
  for (int q = 0; q < 6; q++)   // phi, n_x, n_y, n_z, curv, psi
    if (r.ok && std::fread(&fld[size_t(q) * nin], sizeof(double), size_t(nin), fp) != size_t(nin)) r.ok = false;
  if (r.ok && std::fread(itag.data(), sizeof(int32_t), size_t(nin), fp) != size_t(nin)) r.ok = false;
  if (!r.ok) { set_error("smg_lset_read_rst: truncated cell fields"); return SMG_ERR_ARG; }
  // tag_characteristic, idx_characteristic and scale follow; a characteristic cell
  // (singularity / segment) is outside what the engine models
  {
    std::vector<int32_t> ch(static_cast<size_t>(nin));
    if (std::fread(ch.data(), sizeof(int32_t), size_t(nin), fp) != size_t(nin)) {
      set_error("smg_lset_read_rst: truncated cell fields");
      return SMG_ERR_ARG;
    }
    for (int32_t v : ch)
      if (v != kNormalCell) { set_error("smg_lset_read_rst: singularity / segment cells are not supported"); return SMG_ERR_ARG; }
  }
  auto hidx = [&](int i, int j, int k) { return (int64_t(k + (D > 2)) * halo[1] + (j + (D > 1))) * halo[0] + (i + 1); };
  auto eidx = [&](int i, int j, int k) { return (int64_t(k + (D > 2 ? 4 : 0)) * ext[1] + (j + (D > 1 ? 4 : 0))) * ext[0] + (i + 4); };
  if (phi) std::memset(phi, 0, sizeof(double) * size_t(info->n_halo_cells));
  if (normal) std::memset(normal, 0, sizeof(float) * 3 * size_t(info->n_halo_cells));
  if (tag) std::memset(tag, 0, size_t(info->n_halo_cells));
  if (phi_ext) std::memset(phi_ext, 0, sizeof(double) * size_t(info->n_ext_cells));
  int64_t c = 0;
  for (int k = 0; k < nc[2]; k++)
    for (int j = 0; j < nc[1]; j++)
      for (int i = 0; i < nc[0]; i++, c++) {
        const int64_t o = hidx(i, j, k);
        if (phi) phi[o] = fld[size_t(c)];
        if (normal)
          for (int a = 0; a < 3; a++) normal[3 * o + a] = a < D ? float(fld[size_t(1 + a) * nin + c]) : 0.f;
        if (tag) tag[o] = itag[size_t(c)] == kCutCell ? 1 : 0;
        if (phi_ext) phi_ext[eidx(i, j, k)] = fld[size_t(c)];
      }
  return SMG_OK;
}

int smg_lset_read_sdf(const char* path, const smg_config* cfg, double* phi_ext) {
  using namespace smg;
  if (!path || !cfg || !phi_ext) { set_error("smg_lset_read_sdf: NULL argument"); return SMG_ERR_ARG; }
  if (cfg->dim != 3) { set_error("smg_lset_read_sdf: SDF files are 3D (Load_SDF_file rejects 2D)"); return SMG_ERR_ARG; }
  FILE* fp = std::fopen(path, "r");
  if (!fp) { set_error(std::string("smg_lset_read_sdf: cannot open ") + path); return SMG_ERR_ARG; }
  FileGuard guard{fp};
  long long g[3];
  double origin[3], dx;
  if (std::fscanf(fp, "%lld %lld %lld", &g[0], &g[1], &g[2]) != 3 ||
      std::fscanf(fp, "%lf %lf %lf", &origin[0], &origin[1], &origin[2]) != 3 || std::fscanf(fp, "%lf", &dx) != 1) {
    set_error("smg_lset_read_sdf: bad header (ni nj nk / origin / dx)");
    return SMG_ERR_ARG;
  }
  if (g[0] <= 0 || g[1] <= 0 || g[2] <= 0) { set_error("smg_lset_read_sdf: bad grid size"); return SMG_ERR_ARG; }
  int nc[3], ext[3];
  double dl = 1.e20;
  for (int a = 0; a < 3; a++) {
    nc[a] = 4 * cfg->num_pkg[a];
    ext[a] = nc[a] + 8;
    dl = std::fmin(dl, cfg->domain[a] / double(cfg->num_pkg[a]) / 4.0);
  }
  if (std::fabs(dx - dl) > 1.e-15) {   // the reference aborts on this (lset_level_infor.cpp:1196-1199)
    set_error("smg_lset_read_sdf: the file's dx differs from the level-set cell size");
    return SMG_ERR_ARG;
  }
  // the margin packages start at zero, as Levelset_cell::Initialize leaves them
  std::memset(phi_ext, 0, sizeof(double) * size_t(ext[0]) * ext[1] * ext[2]);
  for (long long k = 0; k < g[2]; k++)
    for (long long j = 0; j < g[1]; j++)
      for (long long i = 0; i < g[0]; i++) {
        double v;
        if (std::fscanf(fp, "%lf", &v) != 1) { set_error("smg_lset_read_sdf: truncated values"); return SMG_ERR_ARG; }
        if (i < nc[0] && j < nc[1] && k < nc[2])
          phi_ext[((k + 4) * ext[1] + (j + 4)) * int64_t(ext[0]) + (i + 4)] = v;
      }
  return SMG_OK;
}

int smg_lset_write_rst(const char* path, const smg_config* cfg, const double* phi, const float* normal,
                       const int8_t* tag, const smg_lset_stats* st) {
  using namespace smg;
  if (!path || !cfg || !phi || !normal || !tag || !st) { set_error("smg_lset_write_rst: NULL argument"); return SMG_ERR_ARG; }
  const int D = cfg->dim;
  if (D != 2 && D != 3) { set_error("smg_lset_write_rst: dim must be 2 or 3"); return SMG_ERR_ARG; }
  FILE* fp = std::fopen(path, "wb");
  if (!fp) { set_error(std::string("smg_lset_write_rst: cannot create ") + path); return SMG_ERR_ARG; }
  FileGuard guard{fp};
  Writer w;
  w.f = fp;
  int np[3];
  double dpkg[3], dcell[3], box_r[3], pbox[3];
  double dl = 1.e20;
  for (int a = 0; a < 3; a++) {
    const bool on = a < D;
    np[a] = on ? cfg->num_pkg[a] : 1;
    dpkg[a] = cfg->domain[a] / double(np[a]);
    dcell[a] = on ? dpkg[a] / 4.0 : dpkg[a];
    if (on) dl = std::fmin(dl, dcell[a]);
    box_r[a] = cfg->box_l[a] + cfg->domain[a];
    pbox[a] = cfg->particle_box_r[a] - cfg->particle_box_l[a];
  }
  // case header
  w.put<int>(D);
  for (int a = 0; a < 3; a++) w.put<double>(cfg->domain[a]);
  for (int a = 0; a < 3; a++) w.put<double>(cfg->box_l[a]);
  for (int a = 0; a < 3; a++) w.put<double>(box_r[a]);
  for (int a = 0; a < 3; a++) w.put<double>(pbox[a]);
  for (int a = 0; a < 3; a++) w.put<double>(cfg->particle_box_l[a]);
  for (int a = 0; a < 3; a++) w.put<double>(cfg->particle_box_r[a]);
  for (int a = 0; a < 3; a++) w.put<int>(4 * np[a]);   // resolution
  w.put<int>(0);                                        // Lmin
  w.put<int>(0);                                        // Lmax
  for (int a = 0; a < 3; a++) w.put<int>(np[a]);       // ini_num_cell
  w.put<int>(1);                                        // num_level
  // level-set header
  const double max_phi = st->maximum_phi;
  w.put<double>(0.0);   // convergence_error
  w.put<int>(0);        // num_singularity
  w.put<int>(0);        // num_segment
  const double fmax = cfg->max_dl / dl, fmin = cfg->min_dl / dl;
  const double hv[] = {fmax, std::sqrt(fmax * fmin), fmin, 0.0, st->total_mass, st->total_mass_surface, 0.0,
                       st->total_volume, st->total_area, 0.0, max_phi, 0.0, cfg->max_dl,
                       std::sqrt(cfg->max_dl * cfg->min_dl), cfg->min_dl, 0.0, 0.0};
  for (double v : hv) w.put<double>(v);
  w.put<int>(int(st->n_interface_cell));
  // level info: one rank, one margin package per side
  w.put<int>((np[0] + 2) * (D > 1 ? np[1] + 2 : 1) * (D > 2 ? np[2] + 2 : 1));   // glbl_total_num_pkg
  w.put<double>(dpkg[0]);   // scale
  w.put<double>(dl);
  for (int a = 0; a < 3; a++) w.put<int>(np[a]);                       // num_pkg
  for (int a = 0; a < 3; a++) w.put<int>(a < D ? np[a] + 2 : 1);       // glbl_num_pkg
  for (int a = 0; a < 3; a++) w.put<int>(a < D ? -1 : 0);              // glbl_pkg_start
  for (int a = 0; a < 3; a++) w.put<int>(a < D ? np[a] + 1 : 1);       // glbl_pkg_end
  for (int a = 0; a < 3; a++) w.put<int>(a < D ? -1 : 0);              // pkg_start
  for (int a = 0; a < 3; a++) w.put<int>(a < D ? np[a] + 1 : 1);       // pkg_end
  for (int a = 0; a < 3; a++) w.put<int>(a < D ? 1 : 0);               // num_pkg_margin
  for (int a = 0; a < 3; a++) w.put<double>(dpkg[a]);
  for (int a = 0; a < 3; a++) w.put<double>(dcell[a]);
  // cell fields over the interior cells, x fastest
  const int nc[3] = {4 * np[0], D > 1 ? 4 * np[1] : 1, D > 2 ? 4 * np[2] : 1};
  const int halo[3] = {nc[0] + 2, D > 1 ? nc[1] + 2 : 1, D > 2 ? nc[2] + 2 : 1};
  auto hidx = [&](int i, int j, int k) { return (int64_t(k + (D > 2)) * halo[1] + (j + (D > 1))) * halo[0] + (i + 1); };
  SpacingProfile prof(*cfg, dl, cfg->max_phi > 0. ? cfg->max_phi : max_phi);
  for (int q = 0; q < 10; q++)
    for (int k = 0; k < nc[2]; k++)
      for (int j = 0; j < nc[1]; j++)
        for (int i = 0; i < nc[0]; i++) {
          const int64_t o = hidx(i, j, k);
          switch (q) {
            case 0: w.put<double>(phi[o]); break;
            case 1: case 2: case 3: w.put<double>(q - 1 < D ? double(normal[3 * o + q - 1]) : 0.0); break;
            case 4: case 5: w.put<double>(0.0); break;   // curv, psi
            case 6: w.put<int>(tag[o] ? kCutCell : kNormalCell); break;
            case 7: w.put<int>(kNormalCell); break;       // tag_characteristic
            case 8: w.put<int>(-1); break;                // idx_characteristic
            default: w.put<double>(prof.dx(phi[o])); break;   // scale
          }
        }
  if (!w.ok) { set_error(std::string("smg_lset_write_rst: write failed on ") + path); return SMG_ERR_ARG; }
  return SMG_OK;
}

}  // extern "C"
