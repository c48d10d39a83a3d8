// smg_internal.hpp -- declarations shared by the engine's translation units.
#pragma once
#include <cmath>
#include <cstdint>
#include <string>

#include "smg.h"

namespace smg {

// Levelset::Get_scale (level_set.cpp:311-351) for host code; the device copy of the
// same arithmetic lives in smg_engine.hip.
struct SpacingProfile {
  int profile;
  double min_dl, max_dl, max_phi, dl, tanh25;
  double dx0, dx1, band;
  SpacingProfile(const smg_config& c, double dl_, double max_phi_)
      : profile(c.profile), min_dl(c.min_dl), max_dl(c.max_dl), max_phi(max_phi_), dl(dl_),
        tanh25(std::tanh(2.5)), dx0(c.lin_dx0), dx1(c.lin_dx1), band(c.lin_band) {}
  double dx(double phi) const {
    if (profile == SMG_PROFILE_LINEAR) {
      if (phi > 0. || std::fabs(phi) < dl) {
        double t = phi > 0. ? phi / band : 0.;
        t = t < 1.0 ? t : 1.0;
        return dx0 + t * (dx1 - dx0);
      }
      return dx0;
    }
    if (phi > 0. || std::fabs(phi) < dl) return min_dl + std::tanh(2.5 * std::fabs(phi) / max_phi) / tanh25 * (max_dl - min_dl);
    return min_dl;
  }
  double dx_min() const { return profile == SMG_PROFILE_LINEAR ? std::min(dx0, dx1) : std::min(min_dl, max_dl); }
  double dx_max() const { return profile == SMG_PROFILE_LINEAR ? std::max(dx0, dx1) : std::max(min_dl, max_dl); }
};

int lset_build(const smg_config* cfg, const double* phi_ext, int zero_margin, double* phi, float* normal,
               int8_t* tag, smg_lset_stats* st);
// the same outputs computed on the current GPU (csrc/smg_lset_dev.hip); kernel_ms (may
// be null) gets the device time from the first kernel to the last
int lset_build_device(const smg_config* cfg, const double* phi_ext, int zero_margin, double* phi, float* normal,
                      int8_t* tag, smg_lset_stats* st, double* kernel_ms);

// per-tetrahedron radius ratio (3 r_in / r_circ) and smallest dihedral angle in the
// reference's degrees (csrc/smg_tet.cpp); tets: 4 vertex indices into x (3 per vertex)
void tet_cell_quality(const double* x, const int32_t* tets, int64_t n, float* radius_ratio, float* min_dihedral);

void set_error(const std::string& msg);

// cells of a one-cell-halo level-set grid with num_pkg << shift packages per axis
int64_t halo_cells(const smg_config& c, int shift = 0);
// maximum_phi over the interior cells of such a grid, as Parallel_get_global_scale
// (parallelfunc_mesh.cpp:46-54)
double interior_max_phi(const smg_config& c, const double* phi, int shift);

}  // namespace smg
