// smg_kernels.hpp -- host-side launchers of the gfx950 kernels (smg_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "smg.h"
#include "smg_device.hpp"

namespace smg {

enum { kModeInit = 0, kModePrerun = 1, kModeStep = 2 };

void launch_reset(hipStream_t st, const Work& W, int njobs, int par);
void launch_decide(hipStream_t st, const Work& W, int njobs, int force);
void launch_jump_lists(hipStream_t st, int D, const Work& W, const PState& S, int njobs);
void launch_key(hipStream_t st, int D, int64_t ntiles, const Work& W, const PState& S);
void launch_copyback(hipStream_t st, int D, int64_t ntiles, const Work& W, const PState& A, const PState& B);
void launch_nb_out(hipStream_t st, int D, int64_t ntiles, const Work& W, const PState& S, int pass, const int64_t* offs,
                   int32_t* cnt, int32_t* out);
void launch_seed(hipStream_t st, int D, int64_t ntiles, const Work& W, const PState& S);
void launch_map(hipStream_t st, int D, int64_t ntiles, const Work& W, const PState& S, int mode, int par);
void launch_scan(hipStream_t st, const Work& W, bool gated, bool per_job);
void launch_scatter(hipStream_t st, int64_t ntiles, const Work& W, const PState& S);
void launch_segsort(hipStream_t st, int64_t ntiles, const Work& W, const PState& S);
void launch_gather(hipStream_t st, int D, int64_t ntiles, const Work& W, const PState& A, const PState& B);
void launch_nb(hipStream_t st, int D, int64_t ntiles, const Work& W, const PState& S);
void launch_nb_spill(hipStream_t st, int D, int64_t ntiles, const Work& W, const PState& S);
void launch_halo(hipStream_t st, int D, int64_t ntiles, const Work& W, const PState& S);
void launch_force1(hipStream_t st, int D, int64_t ntiles, const Work& W, const PState& S, int par);
void launch_force2(hipStream_t st, int D, int64_t ntiles, const Work& W, const PState& S, int par);
void launch_residual(hipStream_t st, int D, int64_t ntiles, const Work& W, const PState& S, int njobs, int par);
void launch_vor2d(hipStream_t st, int64_t ntiles, const Work& W, const PState& S, int32_t* out_nb, int32_t* out_cnt,
                  int KE);
// phi (and, when nrm is not null, the interpolated normal, 3 per point) at points
void launch_sample_phi(hipStream_t st, int D, const Work& W, int job, int64_t n, const double* pts, double* phi,
                       double* nrm = nullptr);

// 3D Voronoi cells -> Delaunay tetrahedra (smg_vor3d.hip)
struct V3Grid {
  const double* x;        // 3 per particle, by particle id
  const int32_t* start;   // cell starts, g0 g1 g2 + 1 entries (x fastest)
  const int32_t* order;   // particle ids in cell order
  int64_t n;
  int32_t g[3];
  double lo[3];           // grid origin
  double s;               // cell size
  int32_t rmax;           // outermost ring of cells searched
};
struct V3Out {
  int32_t* tets;                // 4 ascending particle ids per tetrahedron
  unsigned long long* count;    // tetrahedra found (may exceed cap: then none past cap is written)
  int64_t cap;
  int8_t* status;               // by id: 0 cell closed, 1 open at rmax; not built: 2 capacity,
                                //   3 empty cut boundary, 4 non-convex cut, 5 singular vertex
  double* rcert;                // by id: every point within this radius was clipped (1e300:
                                //   closed by the second pass, every vertex verified)
};
// coarse blocks of the grid for the second pass: the points of each block and the
// bounding box of those points
struct V3Blocks {
  const int32_t* start;   // block starts, g0 g1 g2 + 1 entries (x fastest)
  const int32_t* order;   // particle ids in block order
  const double* box;      // 6 per block: min xyz, max xyz of its points
  int32_t g[3];
  double lo[3];           // origin (the grid's)
  double s;               // block size
};
void launch_vor3d(hipStream_t st, const V3Grid& G, const V3Out& O);
// the cells of ids (those pass 1 left open), started from a box of half-size B and
// verified vertex by vertex; status 0 when closed
void launch_vor3d_full(hipStream_t st, const V3Grid& G, const V3Blocks& K, const int32_t* ids, int64_t nids, double B,
                       const V3Out& O);
void launch_fp64_probe(hipStream_t st, int64_t n, const double* a, const double* b, double* out);

// the no-allocation sampling entry point's staging array (doubles): points then phi,
// 4 doubles per point
constexpr int64_t kStaticSampleDoubles = int64_t(1) << 20;   // 8 MiB
constexpr int64_t kStaticSampleMax = kStaticSampleDoubles / 4;
// this device's address of the staging array (looked up on the first call on each
// device, with that device current); 0 on success
int static_sample_scratch(double** ptr);

}  // namespace smg
