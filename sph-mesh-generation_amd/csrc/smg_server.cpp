// smg_server.cpp -- continuous batching: a pool engine that admits relaxation jobs
// between chunks of steps, steps all of them in the same launches, and retires the
// ones that converged.  No counterpart in the reference (one geometry per MPI job).
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <vector>

#include "smg.h"
#include "smg_device.hpp"
#include "smg_engine.hpp"
#include "smg_internal.hpp"

using smg::set_error;

namespace {

struct Result {
  smg_server_job info{};
  std::vector<double> x, h;
  std::vector<int32_t> type;
  smg_stats st{};
};

struct Active {
  int32_t ticket = -1;
  int64_t steps = 0;
  unsigned long long rcum_prev = 0ull;
};

// device bytes per slot / table entry / level-set cell / job, as Engine allocates them
struct Sizes {
  int64_t per_slot, per_table, per_cell, per_job;
};
Sizes sizes_for(int K, int sps) {
  Sizes z;
  const int64_t state = 9 * 4 + 8 + 4 + 4;   // xh, xl, nrm (3 each), h, type, id
  // state (A, B), nu_a / nu_b, a1 / a2, sum, rec, key / skey / perm / sperm / sid,
  // inline lists and their near bits, nbc / nbo / ovq / ntrue, recb, spill pool, items and stage-2 items
  // (2 B each), halo (ht per 256 slots), ioff, nbc2, perm, and the per-tile words
  const int64_t ips = int64_t(K) + sps;
  const int64_t ht = std::min<int64_t>(smg::kMaxHalo, int64_t(smg::kTile) * std::max<int64_t>(16, ips / 8));
  z.per_slot = 2 * state + 2 * 4 + 6 * 4 + 4 + 4 * 16 + 5 * 4 + int64_t(K) * 4 + int64_t((K + 31) / 32) * 4 + 4 * 4 +
               16 + int64_t(sps) * 4 +
               ips * 4 + ht / smg::kTile * 4 + 4 + 4 + 1 + 1 + 1;
  z.per_table = 2 * 4 + 1;
  z.per_cell = smg::kLsCellBytes;
  z.per_job = int64_t(sizeof(smg::JobParams) + sizeof(smg::JobRed) + smg::kAccCopies * sizeof(smg::JobAcc));
  return z;
}

}  // namespace

struct smg_server {
  smg_server_config cfg{};
  smg::Engine eng;
  std::map<int32_t, Active> by_engine;   // engine job index -> active job
  std::map<int32_t, Result> results;     // finished, not yet collected
  std::deque<smg_server_job> finished;   // not yet reported by smg_server_finished
  int32_t next_ticket = 0;
  int64_t active_particles = 0;
};

extern "C" {

int smg_server_footprint_of(const smg_server_config* c, int64_t* bytes) {
  if (!c || !bytes) { set_error("smg_server_footprint_of: NULL argument"); return SMG_ERR_ARG; }
  const Sizes z = sizes_for(c->max_neighbors > 0 ? (c->max_neighbors + 3) / 4 * 4 : 128, c->spill_per_slot > 0 ? c->spill_per_slot : 64);
  *bytes = c->max_slots * z.per_slot + c->max_table * z.per_table + c->max_lset_cells * z.per_cell +
           int64_t(c->max_jobs) * z.per_job + std::max<int64_t>(c->max_lset_maps, 1) * int64_t(sizeof(int32_t));
  return SMG_OK;
}

int smg_server_job_size(const smg_config* cfg, int64_t n, int64_t* slots, int64_t* table, int64_t* lset_cells) {
  if (!cfg || n <= 0) { set_error("smg_server_job_size: bad argument"); return SMG_ERR_ARG; }
  // parameters only: the level-set arrays are not read
  static const double dphi = 0.;
  static const float dn[3] = {0.f, 0.f, 0.f};
  static const int8_t dt = 0;
  smg::JobInput in;
  in.cfg = *cfg;
  in.n = n;
  in.phi = &dphi;
  in.normal = dn;
  in.tag = &dt;
  in.max_phi = cfg->max_phi > 0. ? cfg->max_phi : 1.0;
  smg::JobParams J;
  const int rc = smg::make_job_params(in, cfg->dim, J);
  if (rc != SMG_OK) return rc;
  if (slots) *slots = J.cap;
  if (table) *table = smg::table_blocks(J) * smg::kScanItems;
  if (lset_cells) *lset_cells = smg::job_lset_cells(J);
  return SMG_OK;
}

int smg_server_create(const smg_server_config* c, smg_server** server) {
  if (!c || !server) { set_error("smg_server_create: NULL argument"); return SMG_ERR_ARG; }
  if (c->chunk_steps <= 0 || !(c->tol >= 0.) || c->min_steps < 0 || c->max_steps < 0 || c->max_lset_maps < 0) {
    set_error("smg_server_create: chunk_steps > 0, tol >= 0, min_steps >= 0, max_steps >= 0, max_lset_maps >= 0");
    return SMG_ERR_ARG;
  }
  std::unique_ptr<smg_server> s(new smg_server());
  s->cfg = *c;
  const int rc = s->eng.setup_pool(c->dim, c->max_slots, c->max_table, c->max_lset_cells, c->max_jobs,
                                   c->max_neighbors, c->spill_per_slot, nullptr, c->max_lset_maps);
  if (rc != SMG_OK) return rc;
  *server = s.release();
  return SMG_OK;
}

int smg_server_destroy(smg_server* server) {
  delete server;
  return SMG_OK;
}

int smg_server_admit(smg_server* s, const smg_config* cfg, const double* phi, const float* normal,
                     const int8_t* tag, int64_t n, int32_t* ticket) {
  return smg_server_admit_levels(s, cfg, phi, normal, tag, 0, nullptr, n, ticket);
}

int smg_server_admit_levels(smg_server* s, const smg_config* cfg, const double* phi, const float* normal,
                            const int8_t* tag, int32_t nlev, const smg_lset_level* levels, int64_t n,
                            int32_t* ticket) {
  if (!s || !cfg || !phi || !normal || !tag || !ticket || n <= 0 || nlev < 0 || nlev > smg::kMaxLsLevels ||
      (nlev > 0 && !levels)) {
    set_error("smg_server_admit: bad argument");
    return SMG_ERR_ARG;
  }
  if (cfg->dim != s->cfg.dim) { set_error("smg_server_admit: job dim differs from the server's"); return SMG_ERR_ARG; }
  if (s->cfg.max_active_particles > 0 && !s->by_engine.empty() &&
      s->active_particles + n > s->cfg.max_active_particles) {
    set_error("smg_server_admit: the pool's particle limit is reached");
    return SMG_ERR_CAPACITY;
  }
  smg::JobInput in;
  in.cfg = *cfg;
  in.n = n;
  for (int l = 0; l < nlev; l++) {
    const smg_lset_level& L = levels[l];
    if (!L.phi || !L.normal || !L.tag || !L.active) {
      set_error("smg_server_admit: NULL array in refined level " + std::to_string(l + 1));
      return SMG_ERR_ARG;
    }
    smg::LevelInput li;
    li.phi = L.phi;
    li.normal = L.normal;
    li.tag = L.tag;
    li.active = L.active;
    in.levels.push_back(li);
  }
  // maximum_phi of the finest level (Levelset::Get_max_phi) for the tanh profile, as
  // smg_set_levelset / smg_set_levelset_levels take it
  in.max_phi = nlev > 0 ? smg::interior_max_phi(*cfg, levels[nlev - 1].phi, nlev) : smg::interior_max_phi(*cfg, phi, 0);
  in.phi = phi;
  in.normal = normal;
  in.tag = tag;
  int j = -1;
  const int rc = s->eng.admit(in, &j);
  if (rc != SMG_OK) return rc;
  Active a;
  a.ticket = s->next_ticket++;
  s->by_engine[j] = a;
  s->active_particles += n;
  *ticket = a.ticket;
  return SMG_OK;
}

int smg_server_advance(smg_server* s, int32_t* n_active) {
  if (!s) { set_error("smg_server_advance: NULL server"); return SMG_ERR_ARG; }
  const int64_t chunk = s->cfg.chunk_steps;
  if (!s->by_engine.empty()) {
    // replayed from a captured graph of G steps: the largest even divisor of the chunk
    // up to 10 (an odd chunk steps eagerly)
    int64_t G = 0;
    for (int64_t g = 10; g >= 2; g -= 2)
      if (chunk % g == 0) { G = g; break; }
    int rc = G > 0 ? s->eng.step_graphed(chunk, G) : s->eng.step(chunk);
    if (rc != SMG_OK) return rc;
    rc = s->eng.download_reds();
    if (rc != SMG_OK) return rc;
    std::vector<int32_t> done;
    for (auto& kv : s->by_engine) {
      const int j = kv.first;
      Active& a = kv.second;
      const smg::JobRed& R = s->eng.red_host(j);
      const smg::JobParams& J = s->eng.params(j);
      a.steps += chunk;
      // the k_map of step s adds the residual of step s - 1: the window covers the
      // chunk's steps shifted by one, the same for every job and every mix
      const double win = double(R.rcum - a.rcum_prev) / smg::kResScale / double(chunk) / double(J.n);
      a.rcum_prev = R.rcum;
      int status = 2;   // keep going
      if (R.fail) status = SMG_ERR_REMAP;
      else if (R.overflow) status = SMG_ERR_CAPACITY;
      else if (a.steps >= s->cfg.min_steps && win < s->cfg.tol) status = SMG_JOB_CONVERGED;
      else if (s->cfg.max_steps > 0 && a.steps >= s->cfg.max_steps) status = SMG_JOB_MAX_STEPS;
      if (status == 2) continue;
      Result r;
      r.info.ticket = a.ticket;
      r.info.status = status;
      r.info.steps = a.steps;
      r.info.n_particle = J.n;
      r.info.window_residual = win;
      r.x.resize(size_t(3 * J.n));
      r.h.resize(size_t(J.n));
      r.type.resize(size_t(J.n));
      rc = s->eng.get_state(j, r.x.data(), nullptr, r.type.data(), r.h.data(), nullptr);
      if (rc != SMG_OK) return rc;
      rc = s->eng.get_stats(j, &r.st);
      if (rc != SMG_OK) return rc;
      r.st.step = a.steps;
      s->finished.push_back(r.info);
      s->results[a.ticket] = std::move(r);
      done.push_back(j);
    }
    for (int j : done) {
      s->active_particles -= s->eng.params(j).n;
      rc = s->eng.retire(j);
      if (rc != SMG_OK) return rc;
      s->by_engine.erase(j);
    }
  }
  if (n_active) *n_active = int32_t(s->by_engine.size());
  return SMG_OK;
}

int smg_server_finished(smg_server* s, smg_server_job* out, int32_t cap, int32_t* n) {
  if (!s || !n || (cap > 0 && !out)) { set_error("smg_server_finished: bad argument"); return SMG_ERR_ARG; }
  int32_t k = 0;
  while (k < cap && !s->finished.empty()) {
    out[k++] = s->finished.front();
    s->finished.pop_front();
  }
  *n = k;
  return SMG_OK;
}

int smg_server_result(smg_server* s, int32_t ticket, double* x, int32_t* type, double* h, smg_stats* st) {
  if (!s) { set_error("smg_server_result: NULL server"); return SMG_ERR_ARG; }
  auto it = s->results.find(ticket);
  if (it == s->results.end()) { set_error("smg_server_result: no finished job with this ticket"); return SMG_ERR_STATE; }
  const Result& r = it->second;
  if (x) std::memcpy(x, r.x.data(), sizeof(double) * r.x.size());
  if (type) std::memcpy(type, r.type.data(), sizeof(int32_t) * r.type.size());
  if (h) std::memcpy(h, r.h.data(), sizeof(double) * r.h.size());
  if (st) *st = r.st;
  s->results.erase(it);
  return SMG_OK;
}

int smg_server_usage(smg_server* s, int64_t* used_slots, int64_t* total_slots, int32_t* n_active) {
  if (!s) { set_error("smg_server_usage: NULL server"); return SMG_ERR_ARG; }
  if (used_slots) *used_slots = s->eng.used_tiles() * smg::kTile;
  if (total_slots) *total_slots = s->cfg.max_slots / smg::kTile * smg::kTile;
  if (n_active) *n_active = int32_t(s->by_engine.size());
  return SMG_OK;
}

}  // extern "C"
