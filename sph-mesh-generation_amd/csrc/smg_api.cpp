// smg_api.cpp -- the extern "C" boundary declared in include/smg.h.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <array>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "smg.h"
#include "smg_engine.hpp"
#include "smg_internal.hpp"
#include "smg_kernels.hpp"

namespace smg {
static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }
}  // namespace smg

using smg::set_error;

struct smg_job {
  smg_config cfg;
  std::vector<double> phi;
  std::vector<float> normal;
  std::vector<int8_t> tag;
  double max_phi = 0.;
  bool has_lset = false;
  struct Level {
    std::vector<double> phi;
    std::vector<float> normal;
    std::vector<int8_t> tag;
    std::vector<uint8_t> active;
  };
  std::vector<Level> levels;   // refined levels 1, 2, ... (smg_set_levelset_levels)
  std::unique_ptr<smg::Engine> eng;
  bool profile = false;
};

namespace smg {

int64_t halo_cells(const smg_config& c, int shift) {
  int64_t m = 1;
  for (int a = 0; a < c.dim; a++) m *= 4 * (int64_t(c.num_pkg[a]) << shift) + 2;
  return m;
}

// maximum_phi over the interior cells of a one-cell-halo grid with num_pkg << shift
// packages, as Parallel_get_global_scale (parallelfunc_mesh.cpp:46-54)
double interior_max_phi(const smg_config& c, const double* phi, int shift) {
  const int nx = 4 * (c.num_pkg[0] << shift), ny = c.dim > 1 ? 4 * (c.num_pkg[1] << shift) : 1;
  const int nz = c.dim > 2 ? 4 * (c.num_pkg[2] << shift) : 1;
  const int gx = nx + 2, gy = c.dim > 1 ? ny + 2 : 1;
  double mp = 0.;
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        const int64_t o = ((int64_t(c.dim > 2 ? k + 1 : 0) * gy) + (c.dim > 1 ? j + 1 : 0)) * gx + i + 1;
        if (phi[o] > 0.) mp = std::max(mp, std::fabs(phi[o]));
      }
  return mp;
}

}  // namespace smg

using smg::halo_cells;
using smg::interior_max_phi;

namespace {

int check_config(const smg_config* c) {
  if (!c) { set_error("config is NULL"); return SMG_ERR_ARG; }
  if (c->dim != 2 && c->dim != 3) { set_error("dim must be 2 or 3"); return SMG_ERR_ARG; }
  for (int a = 0; a < c->dim; a++) {
    if (c->num_pkg[a] <= 0) { set_error("num_pkg must be positive"); return SMG_ERR_ARG; }
    if (!(c->domain[a] > 0.)) { set_error("domain must be positive"); return SMG_ERR_ARG; }
    if (!(c->particle_box_r[a] > c->particle_box_l[a])) { set_error("particle box is empty"); return SMG_ERR_ARG; }
  }
  if (!(c->cell_ratio > 0.) || !(c->cut_off > 0.)) { set_error("cell_ratio/cut_off must be positive"); return SMG_ERR_ARG; }
  if (c->profile != SMG_PROFILE_TANH && c->profile != SMG_PROFILE_LINEAR) { set_error("unknown spacing profile"); return SMG_ERR_ARG; }
  if (c->profile == SMG_PROFILE_TANH && !(c->min_dl > 0. && c->max_dl >= c->min_dl)) {
    set_error("tanh profile needs 0 < min_dl <= max_dl");
    return SMG_ERR_ARG;
  }
  if (c->profile == SMG_PROFILE_LINEAR && !(c->lin_dx0 > 0. && c->lin_dx1 > 0. && c->lin_band > 0.)) {
    set_error("linear profile needs positive lin_dx0, lin_dx1, lin_band");
    return SMG_ERR_ARG;
  }
  return SMG_OK;
}

int build_engine(smg_job* job, int64_t n) {
  if (!job->has_lset) { set_error("set the level set first (smg_set_levelset)"); return SMG_ERR_STATE; }
  smg::JobInput in;
  in.cfg = job->cfg;
  in.n = n;
  in.max_phi = job->max_phi;
  in.phi = job->phi.data();
  in.normal = job->normal.data();
  in.tag = job->tag.data();
  for (auto& L : job->levels) {
    smg::LevelInput li;
    li.phi = L.phi.data();
    li.normal = L.normal.data();
    li.tag = L.tag.data();
    li.active = L.active.data();
    in.levels.push_back(li);
  }
  job->eng.reset(new smg::Engine());
  job->eng->profile_enable(job->profile);
  const int rc = job->eng->setup({in}, job->cfg.max_neighbors, job->cfg.spill_per_slot, nullptr);
  if (rc != SMG_OK) job->eng.reset();
  return rc;
}

}  // namespace

extern "C" {

int smg_abi_version(void) { return SMG_ABI_VERSION; }
const char* smg_last_error(void) { return smg::g_err.c_str(); }

int smg_default_config(int dim, smg_config* c) {
  if (!c || (dim != 2 && dim != 3)) { set_error("smg_default_config: bad argument"); return SMG_ERR_ARG; }
  std::memset(c, 0, sizeof(*c));
  c->dim = dim;
  // SYMMETRIC = DIM2 with free slip (Makefile.in:183-196); all axes in 3D
  c->sym[0] = 1; c->sym[1] = 1; c->sym[2] = dim == 3 ? 1 : 0;
  for (int a = 0; a < 3; a++) {
    c->box_l[a] = 0.;
    c->domain[a] = 1.;
    c->particle_box_l[a] = 0.;
    c->particle_box_r[a] = 1.;
    c->num_pkg[a] = a < dim ? 32 : 1;
  }
  // glbparam.h:171-178
  c->cell_ratio = dim == 3 ? 1.45 : 2.0;
  c->cut_off = 2.0;
  c->angle_thres = 3.1415926 * 0.4;   // sph_mesh_generation.cpp:689 with PI of glbparam.h:185
  c->nu_coeff = 0.1;                  // sph_mesh_generation.cpp:691-692
  c->profile = SMG_PROFILE_TANH;
  // restart-mode spacing factors 2 / 16 times the level-set cell (sph_mesh_generation.cpp:659-661)
  const double dl = 1.0 / (4.0 * 32.0);
  c->min_dl = 2.0 * dl;
  c->max_dl = 16.0 * dl;
  c->max_phi = 0.;
  c->lin_dx0 = c->lin_dx1 = c->lin_band = 0.;
  c->seed = 20261203ull;
  c->remap_max_attempts = 100000;
  c->max_neighbors = 0;
  c->spill_per_slot = 0;
  return SMG_OK;
}

int smg_lset_build(const smg_config* cfg, const double* phi_ext, int zero_margin, double* phi, float* normal,
                   int8_t* tag, smg_lset_stats* stats) {
  if (check_config(cfg) != SMG_OK) return SMG_ERR_ARG;
  if (!phi_ext || !phi || !normal || !tag) { set_error("smg_lset_build: NULL buffer"); return SMG_ERR_ARG; }
  return smg::lset_build(cfg, phi_ext, zero_margin, phi, normal, tag, stats);
}

int smg_lset_build_gpu(const smg_config* cfg, const double* phi_ext, int zero_margin, double* phi, float* normal,
                       int8_t* tag, smg_lset_stats* stats, double* kernel_ms) {
  if (!cfg) { smg::set_error("smg_lset_build_gpu: NULL config"); return SMG_ERR_ARG; }
  if (!phi_ext || !phi || !normal || !tag) { smg::set_error("smg_lset_build_gpu: NULL buffer"); return SMG_ERR_ARG; }
  return smg::lset_build_device(cfg, phi_ext, zero_margin, phi, normal, tag, stats, kernel_ms);
}

int smg_create(const smg_config* cfg, smg_job** job) {
  if (!job) { set_error("smg_create: job is NULL"); return SMG_ERR_ARG; }
  *job = nullptr;
  const int rc = check_config(cfg);
  if (rc != SMG_OK) return rc;
  smg_job* j = new smg_job();
  j->cfg = *cfg;
  *job = j;
  return SMG_OK;
}

int smg_destroy(smg_job* job) {
  delete job;
  return SMG_OK;
}

int smg_set_levelset(smg_job* job, const double* phi, const float* normal, const int8_t* tag) {
  if (!job || !phi || !normal || !tag) { set_error("smg_set_levelset: NULL argument"); return SMG_ERR_ARG; }
  const int64_t m = halo_cells(job->cfg);
  job->phi.assign(phi, phi + m);
  job->normal.assign(normal, normal + 3 * m);
  job->tag.assign(tag, tag + m);
  job->levels.clear();
  job->max_phi = interior_max_phi(job->cfg, job->phi.data(), 0);
  job->has_lset = true;
  return SMG_OK;
}

int smg_set_levelset_levels(smg_job* job, int32_t nlev, const smg_lset_level* levels) {
  if (!job || nlev < 0 || nlev > smg::kMaxLsLevels || (nlev > 0 && !levels)) {
    set_error("smg_set_levelset_levels: bad argument (at most 3 levels)");
    return SMG_ERR_ARG;
  }
  if (!job->has_lset) { set_error("smg_set_levelset_levels: set the base level set first"); return SMG_ERR_STATE; }
  if (job->eng) { set_error("smg_set_levelset_levels: call before seeding"); return SMG_ERR_STATE; }
  std::vector<smg_job::Level> lv(nlev);
  for (int l = 0; l < nlev; l++) {
    const smg_lset_level& L = levels[l];
    if (!L.phi || !L.normal || !L.tag || !L.active) {
      set_error("smg_set_levelset_levels: NULL array in level " + std::to_string(l + 1));
      return SMG_ERR_ARG;
    }
    const int64_t m = halo_cells(job->cfg, l + 1);
    int64_t np = 1;
    for (int a = 0; a < job->cfg.dim; a++) np *= int64_t(job->cfg.num_pkg[a]) << (l + 1);
    lv[l].phi.assign(L.phi, L.phi + m);
    lv[l].normal.assign(L.normal, L.normal + 3 * m);
    lv[l].tag.assign(L.tag, L.tag + m);
    lv[l].active.assign(L.active, L.active + np);
  }
  job->levels = std::move(lv);
  job->max_phi = nlev > 0 ? interior_max_phi(job->cfg, job->levels.back().phi.data(), nlev)
                          : interior_max_phi(job->cfg, job->phi.data(), 0);
  return SMG_OK;
}

int smg_init_particles(smg_job* job, int64_t n, const double* x, const int32_t* type) {
  if (!job || !x || n <= 0) { set_error("smg_init_particles: bad argument"); return SMG_ERR_ARG; }
  int rc = build_engine(job, n);
  if (rc != SMG_OK) return rc;
  rc = job->eng->init_positions(0, x, type);
  if (rc != SMG_OK) return rc;
  return job->eng->init_finish();
}

int smg_seed_particles(smg_job* job, int64_t n) {
  if (!job || n <= 0) { set_error("smg_seed_particles: bad argument"); return SMG_ERR_ARG; }
  int rc = build_engine(job, n);
  if (rc != SMG_OK) return rc;
  rc = job->eng->seed();
  if (rc != SMG_OK) return rc;
  return job->eng->prerun();
}

int smg_prerun(smg_job* job) {
  if (!job || !job->eng) { set_error("smg_prerun: no particles"); return SMG_ERR_STATE; }
  return job->eng->prerun();
}

int smg_step(smg_job* job, int64_t nsteps) {
  if (!job || !job->eng) { set_error("smg_step: no particles"); return SMG_ERR_STATE; }
  if (nsteps < 0) { set_error("smg_step: negative step count"); return SMG_ERR_ARG; }
  int rc = job->eng->step(nsteps);
  if (rc != SMG_OK) return rc;
  rc = job->eng->sync();
  if (rc != SMG_OK) return rc;
  return job->eng->check_health();
}

int smg_get_timestep(smg_job* job, double* dt) {
  if (!job || !job->eng || !dt) { set_error("smg_get_timestep: bad argument"); return SMG_ERR_STATE; }
  return job->eng->timestep(0, dt);
}

int smg_get_stats(smg_job* job, smg_stats* st) {
  if (!job || !job->eng || !st) { set_error("smg_get_stats: bad argument"); return SMG_ERR_STATE; }
  return job->eng->get_stats(0, st);
}

int smg_get_state(smg_job* job, double* x, double* v, int32_t* type, double* h, float* nu) {
  if (!job || !job->eng) { set_error("smg_get_state: no particles"); return SMG_ERR_STATE; }
  return job->eng->get_state(0, x, v, type, h, nu);
}

int smg_get_pair_terms(smg_job* job, float* sum, float* a1, float* a2) {
  if (!job || !job->eng) { set_error("smg_get_pair_terms: no particles"); return SMG_ERR_STATE; }
  return job->eng->get_pair_terms(0, sum, a1, a2);
}

int smg_get_neighbors(smg_job* job, int64_t* offs, int32_t* ids, int32_t* codes, int64_t capacity) {
  if (!job || !job->eng || !offs) { set_error("smg_get_neighbors: bad argument"); return SMG_ERR_STATE; }
  if (ids && !codes) { set_error("smg_get_neighbors: codes is NULL"); return SMG_ERR_ARG; }
  return job->eng->get_neighbors(0, offs, ids, codes, capacity);
}

int smg_get_normals(smg_job* job, float* normal) {
  if (!job || !normal) { set_error("smg_get_normals: bad argument"); return SMG_ERR_ARG; }
  if (!job->eng) { set_error("smg_get_normals: seed or initialise the job first"); return SMG_ERR_STATE; }
  return job->eng->get_normals(0, normal);
}

int smg_sample_phi(smg_job* job, int64_t n, const double* pts, double* phi) {
  if (!job || n < 0 || (n > 0 && (!pts || !phi))) { set_error("smg_sample_phi: bad argument"); return SMG_ERR_ARG; }
  if (!job->eng) { set_error("smg_sample_phi: seed or initialise the job first"); return SMG_ERR_STATE; }
  return job->eng->sample_phi(0, n, pts, phi);
}

int smg_sample_phi_static(smg_job* job, int64_t n, const double* pts, double* phi) {
  if (!job || n < 0 || (n > 0 && (!pts || !phi))) { set_error("smg_sample_phi_static: bad argument"); return SMG_ERR_ARG; }
  if (!job->eng) { set_error("smg_sample_phi_static: seed or initialise the job first"); return SMG_ERR_STATE; }
  return job->eng->sample_phi_static(0, n, pts, phi);
}

int smg_profile_enable(smg_job* job, int enable) {
  if (!job) { set_error("smg_profile_enable: job is NULL"); return SMG_ERR_ARG; }
  job->profile = enable != 0;
  if (job->eng) job->eng->profile_enable(job->profile);
  return SMG_OK;
}

int smg_profile_get(smg_job* job, const char** names, double* avg_ms, int64_t* counts, int max_kernels, int* n_kernels) {
  if (!job || !job->eng || !n_kernels) { set_error("smg_profile_get: bad argument"); return SMG_ERR_STATE; }
  static thread_local std::vector<std::string> keep;
  std::vector<double> ms;
  std::vector<int64_t> cn;
  const int rc = job->eng->profile_get(keep, ms, cn);
  if (rc != SMG_OK) return rc;
  *n_kernels = int(keep.size());
  for (int q = 0; q < int(keep.size()) && q < max_kernels; q++) {
    if (names) names[q] = keep[q].c_str();
    if (avg_ms) avg_ms[q] = ms[q];
    if (counts) counts[q] = cn[q];
  }
  return SMG_OK;
}

int smg_step_async(smg_job* job, int64_t nsteps) {
  if (!job || !job->eng || nsteps < 0) { set_error("smg_step_async: bad argument"); return SMG_ERR_STATE; }
  if (job->eng->has_graph()) {
    const int64_t g = job->eng->graph_steps();
    if (nsteps % g) { set_error("smg_step_async: step count is not a multiple of the captured graph"); return SMG_ERR_ARG; }
    return job->eng->run_graph(nsteps / g);
  }
  return job->eng->step(nsteps);
}

int smg_query(smg_job* job, int32_t* busy) {
  if (!job || !job->eng || !busy) { set_error("smg_query: bad argument"); return SMG_ERR_STATE; }
  bool b = false;
  const int rc = job->eng->query(&b);
  *busy = b ? 1 : 0;
  return rc;
}

int smg_sync(smg_job* job) {
  if (!job || !job->eng) { set_error("smg_sync: no particles"); return SMG_ERR_STATE; }
  const int rc = job->eng->sync();
  if (rc != SMG_OK) return rc;
  return job->eng->check_health();
}

int smg_capture(smg_job* job, int64_t steps_per_graph) {
  if (!job || !job->eng) { set_error("smg_capture: no particles"); return SMG_ERR_STATE; }
  return job->eng->capture_graph(steps_per_graph);
}

int smg_footprint(smg_job* job, int64_t* bytes) {
  if (!job || !bytes) { set_error("smg_footprint: bad argument"); return SMG_ERR_ARG; }
  *bytes = job->eng ? job->eng->footprint() : 0;
  return SMG_OK;
}

// ---------------------------------------------------------------------------
// batched jobs

}  // extern "C"

struct smg_batch {
  std::unique_ptr<smg::Engine> eng[2];   // [0] = 2D jobs, [1] = 3D jobs
  std::vector<int> dim_of, local_of;
  std::vector<std::vector<double>> phi;
  std::vector<std::vector<float>> normal;
  std::vector<std::vector<int8_t>> tag;
};

extern "C" {

int smg_batch_create(int32_t njobs, const smg_config* cfgs, const int64_t* n_particles, const double* const* phi,
                     const float* const* normal, const int8_t* const* tag, smg_batch** batch) {
  if (!batch || njobs <= 0 || !cfgs || !n_particles || !phi || !normal || !tag) {
    set_error("smg_batch_create: bad argument");
    return SMG_ERR_ARG;
  }
  *batch = nullptr;
  std::unique_ptr<smg_batch> b(new smg_batch());
  std::vector<smg::JobInput> in[2];
  b->dim_of.resize(njobs);
  b->local_of.resize(njobs);
  b->phi.resize(njobs);
  b->normal.resize(njobs);
  b->tag.resize(njobs);
  for (int j = 0; j < njobs; j++) {
    int rc = check_config(&cfgs[j]);
    if (rc != SMG_OK) return rc;
    if (n_particles[j] <= 0 || !phi[j] || !normal[j] || !tag[j]) {
      set_error("smg_batch_create: job " + std::to_string(j) + " has no particles or no level set");
      return SMG_ERR_ARG;
    }
    const int64_t m = halo_cells(cfgs[j]);
    b->phi[j].assign(phi[j], phi[j] + m);
    b->normal[j].assign(normal[j], normal[j] + 3 * m);
    b->tag[j].assign(tag[j], tag[j] + m);
    smg::JobInput ji;
    ji.cfg = cfgs[j];
    ji.n = n_particles[j];
    double mp = 0.;
    const smg_config& c = cfgs[j];
    const int nx = 4 * c.num_pkg[0], ny = c.dim > 1 ? 4 * c.num_pkg[1] : 1, nz = c.dim > 2 ? 4 * c.num_pkg[2] : 1;
    const int gx = nx + 2, gy = c.dim > 1 ? ny + 2 : 1;
    for (int k = 0; k < nz; k++)
      for (int jj = 0; jj < ny; jj++)
        for (int i = 0; i < nx; i++) {
          const int64_t o = ((int64_t(c.dim > 2 ? k + 1 : 0) * gy) + (c.dim > 1 ? jj + 1 : 0)) * gx + i + 1;
          if (b->phi[j][o] > 0.) mp = std::max(mp, std::fabs(b->phi[j][o]));
        }
    ji.max_phi = mp;
    ji.phi = b->phi[j].data();
    ji.normal = b->normal[j].data();
    ji.tag = b->tag[j].data();
    const int e = cfgs[j].dim == 3 ? 1 : 0;
    b->dim_of[j] = e;
    b->local_of[j] = int(in[e].size());
    in[e].push_back(ji);
  }
  int maxnb = 0, sps = 0;
  for (int j = 0; j < njobs; j++) {
    maxnb = std::max(maxnb, int(cfgs[j].max_neighbors));
    sps = std::max(sps, int(cfgs[j].spill_per_slot));
  }
  for (int e = 0; e < 2; e++) {
    if (in[e].empty()) continue;
    b->eng[e].reset(new smg::Engine());
    const int rc = b->eng[e]->setup(in[e], maxnb, sps, nullptr);
    if (rc != SMG_OK) return rc;
  }
  *batch = b.release();
  return SMG_OK;
}

int smg_batch_destroy(smg_batch* batch) {
  delete batch;
  return SMG_OK;
}

#define SMG_FOR_ENGINES(b, call)                 \
  for (int e_ = 0; e_ < 2; e_++) {               \
    if (!(b)->eng[e_]) continue;                 \
    const int rc_ = (b)->eng[e_]->call;          \
    if (rc_ != SMG_OK) return rc_;               \
  }

int smg_batch_seed(smg_batch* batch) {
  if (!batch) { set_error("smg_batch_seed: batch is NULL"); return SMG_ERR_ARG; }
  SMG_FOR_ENGINES(batch, seed());
  SMG_FOR_ENGINES(batch, prerun());
  return SMG_OK;
}

int smg_batch_step(smg_batch* batch, int64_t nsteps) {
  if (!batch || nsteps < 0) { set_error("smg_batch_step: bad argument"); return SMG_ERR_ARG; }
  SMG_FOR_ENGINES(batch, step(nsteps));   // both dimensions enqueue before either waits
  SMG_FOR_ENGINES(batch, sync());
  SMG_FOR_ENGINES(batch, check_health());
  return SMG_OK;
}

int smg_batch_capture(smg_batch* batch, int64_t steps_per_graph) {
  if (!batch) { set_error("smg_batch_capture: batch is NULL"); return SMG_ERR_ARG; }
  SMG_FOR_ENGINES(batch, capture_graph(steps_per_graph));
  return SMG_OK;
}

int smg_batch_run(smg_batch* batch, int64_t launches) {
  if (!batch || launches < 0) { set_error("smg_batch_run: bad argument"); return SMG_ERR_ARG; }
  SMG_FOR_ENGINES(batch, run_graph(launches));
  return SMG_OK;
}

int smg_batch_sync(smg_batch* batch) {
  if (!batch) { set_error("smg_batch_sync: batch is NULL"); return SMG_ERR_ARG; }
  SMG_FOR_ENGINES(batch, sync());
  SMG_FOR_ENGINES(batch, check_health());
  return SMG_OK;
}

int smg_batch_get_stats(smg_batch* batch, int32_t job, smg_stats* st) {
  if (!batch || !st || job < 0 || job >= int32_t(batch->dim_of.size())) { set_error("smg_batch_get_stats: bad argument"); return SMG_ERR_ARG; }
  return batch->eng[batch->dim_of[job]]->get_stats(batch->local_of[job], st);
}

int smg_batch_get_state(smg_batch* batch, int32_t job, double* x, double* v, int32_t* type, double* h, float* nu) {
  if (!batch || job < 0 || job >= int32_t(batch->dim_of.size())) { set_error("smg_batch_get_state: bad argument"); return SMG_ERR_ARG; }
  return batch->eng[batch->dim_of[job]]->get_state(batch->local_of[job], x, v, type, h, nu);
}

int smg_triangulate(smg_job* job, int32_t* tris, int64_t capacity, int64_t* n_tri, smg_tri_stats* stats) {
  if (!job || !job->eng || !n_tri) { set_error("smg_triangulate: bad argument"); return SMG_ERR_STATE; }
  std::vector<std::array<int32_t, 3>> t;
  smg_tri_stats st;
  const int rc = job->eng->triangulate(0, t, &st);
  if (rc != SMG_OK) return rc;
  *n_tri = int64_t(t.size());
  if (stats) *stats = st;
  if (!tris) return SMG_OK;
  if (int64_t(t.size()) > capacity) { set_error("smg_triangulate: capacity too small"); return SMG_ERR_CAPACITY; }
  for (size_t q = 0; q < t.size(); q++)
    for (int a = 0; a < 3; a++) tris[3 * q + a] = t[q][a];
  return SMG_OK;
}

int smg_tri_edges(smg_job* job, int32_t* edges, int64_t capacity, int64_t* n_edges) {
  if (!job || !job->eng || !n_edges) { set_error("smg_tri_edges: bad argument"); return SMG_ERR_STATE; }
  const auto& e = job->eng->last_edges();
  *n_edges = int64_t(e.size());
  if (!edges) return SMG_OK;
  if (int64_t(e.size()) > capacity) { set_error("smg_tri_edges: capacity too small"); return SMG_ERR_CAPACITY; }
  for (size_t q = 0; q < e.size(); q++) {
    edges[2 * q] = e[q][0];
    edges[2 * q + 1] = e[q][1];
  }
  return SMG_OK;
}

int smg_tetrahedralize(smg_job* job, double delete_thres, int32_t* tets, int64_t capacity, int64_t* n_tet,
                       smg_tet_stats* stats) {
  if (!job || !job->eng || !n_tet) { set_error("smg_tetrahedralize: bad argument"); return SMG_ERR_STATE; }
  smg_tet_stats st;
  const int rc = job->eng->tetrahedralize(0, delete_thres, &st);
  if (rc != SMG_OK) return rc;
  const auto& t = job->eng->last_tets().kept;
  *n_tet = int64_t(t.size());
  if (stats) *stats = st;
  if (!tets) return SMG_OK;
  if (int64_t(t.size()) > capacity) { set_error("smg_tetrahedralize: capacity too small"); return SMG_ERR_CAPACITY; }
  for (size_t q = 0; q < t.size(); q++)
    for (int a = 0; a < 4; a++) tets[4 * q + a] = t[q][a];
  return SMG_OK;
}

int smg_tet_surface(smg_job* job, int32_t* tris, int64_t capacity, int64_t* n_tri) {
  if (!job || !job->eng || !n_tri) { set_error("smg_tet_surface: bad argument"); return SMG_ERR_STATE; }
  const auto& t = job->eng->last_tets().surface;
  *n_tri = int64_t(t.size());
  if (!tris) return SMG_OK;
  if (int64_t(t.size()) > capacity) { set_error("smg_tet_surface: capacity too small"); return SMG_ERR_CAPACITY; }
  for (size_t q = 0; q < t.size(); q++)
    for (int a = 0; a < 3; a++) tris[3 * q + a] = t[q][a];
  return SMG_OK;
}

int smg_tet_delaunay(smg_job* job, int32_t* tets, int64_t capacity, int64_t* n_tet, int8_t* status, double* rcert) {
  if (!job || !job->eng || !n_tet) { set_error("smg_tet_delaunay: bad argument"); return SMG_ERR_STATE; }
  const auto& R = job->eng->last_tets();
  *n_tet = int64_t(R.delaunay.size());
  if (status) std::copy(R.status.begin(), R.status.end(), status);
  if (rcert) std::copy(R.rcert.begin(), R.rcert.end(), rcert);
  if (!tets) return SMG_OK;
  if (int64_t(R.delaunay.size()) > capacity) { set_error("smg_tet_delaunay: capacity too small"); return SMG_ERR_CAPACITY; }
  for (size_t q = 0; q < R.delaunay.size(); q++)
    for (int a = 0; a < 4; a++) tets[4 * q + a] = R.delaunay[q][a];
  return SMG_OK;
}

int smg_autotune(const smg_config* cfg, const double* phi, const float* normal, const int8_t* tag, int64_t n,
                 int32_t warmup, int32_t steps, smg_tune_result* res) {
  if (check_config(cfg) != SMG_OK) return SMG_ERR_ARG;
  if (!phi || !normal || !tag || !res || n <= 0 || warmup < 0 || steps <= 0) {
    set_error("smg_autotune: bad argument");
    return SMG_ERR_ARG;
  }
  std::memset(res, 0, sizeof(*res));
  const int64_t m = halo_cells(*cfg);
  double mp = 0.;
  for (int64_t q = 0; q < m; q++) mp = std::max(mp, phi[q]);
  const double skins[] = {0.05, 0.1, 0.15, 0.25, 0.4};
  const int32_t depths[] = {64, 96, 128};   // 15 candidates (the result holds 16)
  const int64_t gsteps = 10;   // steps per captured graph (even)
  const int64_t timed = std::max<int64_t>(gsteps, (steps + gsteps - 1) / gsteps * gsteps);
  res->best_ms_per_step = 1.e300;
  for (double sk : skins)
    for (int32_t K : depths) {
      smg::JobInput in;
      in.cfg = *cfg;
      in.cfg.skin = sk;
      in.cfg.max_neighbors = K;
      in.n = n;
      in.max_phi = mp;
      in.phi = phi;
      in.normal = normal;
      in.tag = tag;
      smg::Engine eng;
      int rc = eng.setup({in}, K, cfg->spill_per_slot, nullptr);
      if (rc == SMG_OK) rc = eng.seed();
      if (rc == SMG_OK) rc = eng.prerun();
      if (rc == SMG_OK) rc = eng.step(warmup);
      if (rc == SMG_OK) rc = eng.sync();
      if (rc == SMG_OK) rc = eng.capture_graph(gsteps);
      if (rc != SMG_OK) return rc;
      const auto t0 = std::chrono::steady_clock::now();
      rc = eng.run_graph(timed / gsteps);
      if (rc == SMG_OK) rc = eng.sync();
      if (rc == SMG_OK) rc = eng.check_health();
      if (rc != SMG_OK) return rc;
      const double ms =
          std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / double(timed);
      const int q = res->n_tried++;
      res->skin[q] = sk;
      res->max_neighbors[q] = K;
      res->ms_per_step[q] = ms;
      if (ms < res->best_ms_per_step) {
        res->best_ms_per_step = ms;
        res->best_skin = sk;
        res->best_max_neighbors = K;
      }
    }
  return SMG_OK;
}

int smg_fp64_probe(int64_t n, const double* a, const double* b, double* out) {
  if (n <= 0 || !a || !b || !out) { set_error("smg_fp64_probe: bad argument"); return SMG_ERR_ARG; }
  double *da = nullptr, *db = nullptr, *dout = nullptr;
  if (hipMalloc(&da, sizeof(double) * n) != hipSuccess || hipMalloc(&db, sizeof(double) * n) != hipSuccess ||
      hipMalloc(&dout, sizeof(double) * 4 * n) != hipSuccess) {
    (void)hipFree(da); (void)hipFree(db); (void)hipFree(dout);
    set_error("smg_fp64_probe: hipMalloc failed");
    return SMG_ERR_HIP;
  }
  (void)hipMemcpy(da, a, sizeof(double) * n, hipMemcpyHostToDevice);
  (void)hipMemcpy(db, b, sizeof(double) * n, hipMemcpyHostToDevice);
  smg::launch_fp64_probe(nullptr, n, da, db, dout);
  hipError_t e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(out, dout, sizeof(double) * 4 * n, hipMemcpyDeviceToHost);
  (void)hipFree(da); (void)hipFree(db); (void)hipFree(dout);
  if (e != hipSuccess) { set_error(std::string("smg_fp64_probe: ") + hipGetErrorString(e)); return SMG_ERR_HIP; }
  return SMG_OK;
}

}  // extern "C"
