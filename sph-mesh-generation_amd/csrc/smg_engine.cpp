// smg_engine.cpp -- host orchestration: job layout, device buffers, the step
// sequence of SPH_run_mesh_generation and result download.
#include "smg_engine.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

#include "smg_internal.hpp"
#include "smg_kernels.hpp"

namespace smg {

#define HIP_TRY(expr)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess) {                                                             \
      set_error(std::string(#expr) + ": " + hipGetErrorString(e_));                     \
      return SMG_ERR_HIP;                                                               \
    }                                                                                   \
  } while (0)

#define SMG_TRY(expr)        \
  do {                       \
    int r_ = (expr);         \
    if (r_ != SMG_OK) return r_; \
  } while (0)

Engine::~Engine() {
  if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
  if (graph_) (void)hipGraphDestroy(graph_);
  for (auto& r : recs_) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  for (auto e : ev_pool_) (void)hipEventDestroy(e);
  for (void* p : allocs_) (void)hipFree(p);
  if (own_stream_ && st_) (void)hipStreamDestroy(st_);
}

template <typename T>
T* Engine::alloc_(int64_t count) {
  void* p = nullptr;
  if (count <= 0) count = 1;
  if (hipMalloc(&p, size_t(count) * sizeof(T)) != hipSuccess) {
    alloc_failed_ = true;
    return nullptr;
  }
  allocs_.push_back(p);
  bytes_ += int64_t(count) * int64_t(sizeof(T));
  return static_cast<T*>(p);
}

int64_t Engine::total_particles() const {
  int64_t s = 0;
  for (auto& j : jp_) s += j.n;
  return s;
}

static int ceil_log2(int64_t v) {
  int b = 0;
  while ((int64_t(1) << b) < v) b++;
  return b;
}

// One job's parameters: everything except its slot, table and level-set ranges.
// package (p0, p1, p2) of a refined level's map (index 0 is the halo package before
// the first): stored where the interior package it is, or borders, is active
static bool level_pkg_active(const JobParams& J, int l, const LevelInput& L, int p0, int p1, int p2) {
  int i[3] = {p0 - 1, p1 - 1, p2 - 1};
  for (int a = 0; a < 3; a++) i[a] = a < J.D ? std::min(std::max(i[a], 0), J.lnp[l][a] - 1) : 0;
  return L.active[(int64_t(i[2]) * J.lnp[l][1] + i[1]) * J.lnp[l][0] + i[0]] != 0;
}

int make_job_params(const JobInput& in, int D, JobParams& J) {
  const smg_config& c = in.cfg;
  std::memset(&J, 0, sizeof(J));
  if (c.dim != D) { set_error("setup: all jobs of one engine need the same dim"); return SMG_ERR_ARG; }
  if (in.n <= 0) { set_error("setup: job without particles"); return SMG_ERR_ARG; }
  if (!in.phi || !in.normal || !in.tag) { set_error("setup: level set missing"); return SMG_ERR_STATE; }
  J.D = D;
  J.profile = c.profile;
  J.remap_max = c.remap_max_attempts > 0 ? c.remap_max_attempts : 100000;
  double dl = 1.e20, dmin = 1.e300;
  for (int a = 0; a < 3; a++) {
    const bool on = a < D;
    J.sym[a] = on ? c.sym[a] : 0;
    J.num_pkg[a] = on ? c.num_pkg[a] : 1;
    if (on && c.num_pkg[a] <= 0) { set_error("setup: num_pkg must be positive"); return SMG_ERR_ARG; }
    J.ng[a] = on ? 4 * c.num_pkg[a] + 2 : 1;
    J.box_l[a] = on ? c.box_l[a] : 0.;
    J.domain[a] = on ? c.domain[a] : 1.;
    J.box_r[a] = J.box_l[a] + J.domain[a];
    J.pbox_l[a] = on ? c.particle_box_l[a] : 0.;
    J.pbox_r[a] = on ? c.particle_box_r[a] : 1.;
    J.dpkg[a] = on ? c.domain[a] / double(c.num_pkg[a]) : 1.;
    J.dcell[a] = on ? J.dpkg[a] / 4.0 : 1.;
    if (on) { dl = std::min(dl, J.dcell[a]); dmin = std::min(dmin, J.domain[a]); }
    const float lh = float(J.box_l[a]), rh = float(J.box_r[a]);
    J.wl_h[a] = lh; J.wl_l[a] = float(J.box_l[a] - double(lh));
    J.wh_h[a] = rh; J.wh_l[a] = float(J.box_r[a] - double(rh));
  }
  // refined level-set levels: geometry, stored packages (a halo package is stored
  // where the interior package next to it is) and the ranges they take
  const int nl = int(in.levels.size());
  if (nl > kMaxLsLevels) { set_error("setup: at most 3 refined level-set levels"); return SMG_ERR_ARG; }
  J.ls_nl = nl;
  int64_t cells = ls_cells_padded(D, J.ng), maps = 0;
  for (int l = 0; l < nl; l++) {
    const LevelInput& L = in.levels[l];
    if (!L.phi || !L.normal || !L.tag || !L.active) { set_error("setup: refined level incomplete"); return SMG_ERR_ARG; }
    int np[3] = {1, 1, 1};
    for (int a = 0; a < 3; a++) {
      const bool on = a < D;
      J.lnp[l][a] = on ? (c.num_pkg[a] << (l + 1)) : 1;
      J.ldpkg[l][a] = on ? c.domain[a] / double(J.lnp[l][a]) : 1.;
      J.ldcell[l][a] = on ? J.ldpkg[l][a] / 4.0 : 1.;
      if (on) dl = std::min(dl, J.ldcell[l][a]);
      np[a] = on ? J.lnp[l][a] + 2 : 1;
    }
    int64_t nb = 0;
    for (int p2 = 0; p2 < np[2]; p2++)
      for (int p1 = 0; p1 < np[1]; p1++)
        for (int p0 = 0; p0 < np[0]; p0++) nb += level_pkg_active(J, l, L, p0, p1, p2) ? 1 : 0;
    J.lbricks[l] = nb;
    J.lcell_off[l] = cells;
    J.lmap_off[l] = maps;
    cells += nb * (D > 2 ? 64 : 16);
    maps += int64_t(np[0]) * np[1] * np[2];
  }
  J.dl = dl;   // the finest level-set cell (Get_scale's |phi| < dl)
  J.seed_pad = nl > 0 ? J.ldpkg[nl - 1][0] : J.dpkg[0];
  J.cell_ratio = c.cell_ratio;
  J.cut_off = c.cut_off;
  J.nu_coeff = c.nu_coeff;
  J.min_dl = c.min_dl; J.max_dl = c.max_dl;
  J.max_phi = c.max_phi > 0. ? c.max_phi : in.max_phi;
  J.tanh25 = std::tanh(2.5);
  J.lin_dx0 = c.lin_dx0; J.lin_dx1 = c.lin_dx1; J.lin_band = c.lin_band;
  J.cos_thres = float(std::cos(c.angle_thres));
  J.kdx = float(c.cut_off / c.cell_ratio);
  // candidate lists cover (1 + skin) cut_off h; skin < 0 rebuilds them every step
  const double skin = c.skin > 0. ? c.skin : (c.skin < 0. ? 0. : 0.1);
  J.cut_v = (1.0 + skin) * c.cut_off;
  J.skin_lim = c.skin < 0. ? -1.f : float(0.5 * skin * c.cut_off);
  {
    // bound on the FP32 error of a relative position formed from the hi parts
    double X = 0.;
    for (int a = 0; a < D; a++) X = std::max(X, std::max(std::fabs(J.box_l[a]), std::fabs(J.box_r[a])));
    J.pf_tol = float(32.0 * std::ldexp(1.0, -23) * (X + 1.0));
  }
  J.seed = c.seed;
  if (c.profile == SMG_PROFILE_TANH && !(J.max_phi > 0.) && c.max_dl != c.min_dl) {
    set_error("setup: tanh profile needs max_phi > 0");
    return SMG_ERR_ARG;
  }
  // cell-list levels: c0 >= 2 h_max (+ margin), c_L = c0 / 2^L >= 2 h_min (+ margin)
  SpacingProfile prof(c, dl, J.max_phi > 0. ? J.max_phi : 1.0);
  const double hmax = prof.dx_max() * c.cell_ratio / c.cut_off;
  const double hmin = prof.dx_min() * c.cell_ratio / c.cut_off;
  if (!(hmin > 0.)) { set_error("setup: spacing must be positive"); return SMG_ERR_ARG; }
  const double R0 = hmax * J.cut_v * (1. + 1.e-9) + 1.e-9;
  // The cell grid covers the particle box (seeding and re-seeding draw from it, and
  // the fluid lies in it) grown by one list radius, inside the level-set box: a small
  // geometry in a large box gets tables of its own size.
  dmin = 1.e300;
  for (int a = 0; a < 3; a++) {
    J.cbox_l[a] = J.box_l[a];
    J.cdom[a] = J.domain[a];
    if (a >= D) continue;
    const double lo = std::max(J.box_l[a], J.pbox_l[a] - R0), hi = std::min(J.box_r[a], J.pbox_r[a] + R0);
    if (hi > lo) {
      J.cbox_l[a] = lo;
      J.cdom[a] = hi - lo;
    }
    dmin = std::min(dmin, J.cdom[a]);
  }
  int64_t nbox = int64_t(std::floor(dmin / R0));
  if (nbox < 1) nbox = 1;
  J.c0 = dmin / double(nbox);
  const int max_bits = D == 2 ? 15 : 10;
  int nlev = 1;
  while (nlev < kMaxLevels && std::ldexp(J.c0, -nlev) >= hmin * J.cut_v * (1. + 1.e-9) + 2.e-10) nlev++;
  int bits0 = 0;
  for (int a = 0; a < D; a++) bits0 = std::max(bits0, ceil_log2(int64_t(std::ceil(J.cdom[a] / J.c0))));
  while (nlev > 1 && bits0 + nlev - 1 > max_bits) nlev--;
  if (bits0 > max_bits) { set_error("setup: domain has too many cells per axis"); return SMG_ERR_CAPACITY; }
  J.nlev = nlev;
  int64_t off = 0;
  for (int L = 0; L < nlev; L++) {
    J.clen[L] = std::ldexp(J.c0, -L);
    J.cinv[L] = 1.0 / J.clen[L];
    for (int a = 0; a < 3; a++) J.lev_n[L][a] = a < D ? int32_t(std::ceil(J.cdom[a] / J.clen[L])) : 1;
    J.lev_bits[L] = bits0 + L;
    J.lev_off[L] = off;
    off += int64_t(1) << (D * J.lev_bits[L]);
  }
  J.tbl_size = off + 1;   // + pad cell for dead slots
  J.n = in.n;
  J.cap = (J.n + kTile - 1) / kTile * kTile;
  if (J.cap > (int64_t(1) << kJBits)) { set_error("setup: a job holds at most 2^26 particles"); return SMG_ERR_CAPACITY; }
  return SMG_OK;
}

int64_t job_lset_cells(const JobParams& J) {
  int64_t m = ls_cells_padded(J.D, J.ng);
  for (int l = 0; l < J.ls_nl; l++) m += J.lbricks[l] * (J.D > 2 ? 64 : 16);
  return m;
}
int64_t job_lmap_entries(const JobParams& J) {
  int64_t m = 0;
  for (int l = 0; l < J.ls_nl; l++) {
    int64_t e = 1;
    for (int a = 0; a < J.D; a++) e *= J.lnp[l][a] + 2;
    m += e;
  }
  return m;
}

// 4096-entry scan blocks of a job's table segment (+ one entry for a sub-scan's end
// sentinel); segments start on block boundaries so that no scan block spans two jobs
int64_t table_blocks(const JobParams& J) { return (J.tbl_size + 1 + kScanItems - 1) / kScanItems; }

int Engine::set_table_owner_(const JobParams& J, int32_t owner) {
  const int64_t b0 = J.tbl_base / kScanItems, nb = table_blocks(J);
  for (int64_t b = b0; b < b0 + nb; b++) sblk_job_h_[b] = owner;
  HIP_TRY(hipMemcpyAsync(d_sblk_job_ + b0, sblk_job_h_.data() + b0, sizeof(int32_t) * nb, hipMemcpyHostToDevice, st_));
  HIP_TRY(hipStreamSynchronize(st_));
  return SMG_OK;
}

int Engine::alloc_buffers_(int64_t slots, int64_t tbl, int64_t ls, int njobs, int spill_per_slot, int64_t maps) {
  if (tbl >= (int64_t(1) << 31) - 2) { set_error("setup: cell tables exceed 2^31 entries"); return SMG_ERR_CAPACITY; }
  if (slots >= (int64_t(1) << 31) - kTile) { set_error("setup: more than 2^31 particle slots"); return SMG_ERR_CAPACITY; }
  cap_total_ = slots;
  n_tbl_ = tbl;
  n_ls_ = ls;
  d_jp_ = alloc_<JobParams>(njobs);
  d_tile_job_ = alloc_<int32_t>(slots / kTile);
  d_sblk_job_ = alloc_<int32_t>((tbl + kScanItems - 1) / kScanItems + 1);
  d_red_ = alloc_<JobRed>(njobs);
  d_acc_ = alloc_<JobAcc>(int64_t(njobs) * kAccCopies);
  d_lsp_ = alloc_<double>(n_ls_);
  d_lsn_ = alloc_<float4>(n_ls_);
  d_lst_ = alloc_<int8_t>(n_ls_);
  d_lpm_ = alloc_<int32_t>(std::max<int64_t>(maps, 1));
  PState* S2[2] = {&A_, &B_};
  for (PState* S : S2) {
    for (int a = 0; a < 3; a++) {
      S->xh[a] = alloc_<float>(cap_total_);
      S->xl[a] = alloc_<float>(cap_total_);
      S->nrm[a] = alloc_<float>(cap_total_);
    }
    S->h = alloc_<double>(cap_total_);
    S->type = alloc_<int32_t>(cap_total_);
    S->id = alloc_<int32_t>(cap_total_);
  }
  W_.nu_a = alloc_<float>(cap_total_);
  W_.nu_b = alloc_<float>(cap_total_);
  for (int a = 0; a < 3; a++) {
    W_.a1[a] = alloc_<float>(cap_total_);
    W_.a2[a] = alloc_<float>(cap_total_);
  }
  W_.sum = alloc_<float>(cap_total_);
  W_.rec = alloc_<float4>(4 * cap_total_);
  W_.key = alloc_<uint32_t>(cap_total_);
  W_.skey = alloc_<uint32_t>(cap_total_);
  W_.perm = alloc_<int32_t>(cap_total_);
  W_.sperm = alloc_<int32_t>(cap_total_);
  W_.sid = alloc_<int32_t>(cap_total_);
  W_.counts = alloc_<int32_t>(n_tbl_);
  W_.starts = alloc_<int32_t>(n_tbl_ + 1);
  const int64_t nb = (n_tbl_ + kScanItems - 1) / kScanItems;
  W_.bsum = alloc_<int32_t>(nb + 1);
  W_.nbl = alloc_<int32_t>(int64_t(K_) * cap_total_);
  W_.nbnear = alloc_<uint32_t>(int64_t((K_ + 31) / 32) * cap_total_);
  W_.nbc = alloc_<int32_t>(cap_total_);
  W_.nbo = alloc_<int32_t>(cap_total_);
  W_.ovq = alloc_<int32_t>(cap_total_);
  W_.ntrue = alloc_<int32_t>(cap_total_);
  W_.recb = alloc_<float4>(cap_total_ + kRecbPad);   // the list build reads up to 3 records past a range
  W_.jset = alloc_<int32_t>(int64_t(njobs) * kJumpMax);
  W_.jl = alloc_<int32_t>(int64_t(njobs) * kJumpMax * kJumpList);
  W_.jcnt = alloc_<int32_t>(int64_t(njobs) * kJumpMax);
  // spill pool: by default 64 entries per slot beyond the inline K (the unrelaxed
  // random start and strong h gradients are where lists run long); a job owns the
  // part [p0 * sps, (p0 + cap) * sps) of it
  sps_ = spill_per_slot > 0 ? spill_per_slot : 64;
  sps_ = int(std::min<int64_t>(sps_, ((int64_t(1) << 31) - 1) / std::max<int64_t>(cap_total_, 1)));
  if (sps_ < 1) { set_error("setup: too many slots for the spill pool"); return SMG_ERR_CAPACITY; }
  W_.spill_cap = int64_t(sps_) * cap_total_;
  W_.spill = alloc_<int32_t>(W_.spill_cap);
  // neighbour tiles: a job's tiles take their items from a pool of (K + spill_per_slot)
  // items per slot (the lists' average capacity); a tile holds (K + spill_per_slot) 32
  // halo entries (at least 4096, at most 65536)
  W_.ips = int64_t(K_) + sps_;
  W_.ht = std::min<int64_t>(kMaxHalo, int64_t(kTile) * std::max<int64_t>(16, W_.ips / 8));
  const int64_t ntl = cap_total_ / kTile;
  W_.items = alloc_<uint16_t>(W_.ips * cap_total_);
  W_.items2 = alloc_<uint16_t>(W_.ips * cap_total_);
  W_.halo = alloc_<uint32_t>(std::max<int64_t>(ntl, 1) * W_.ht);
  W_.hcnt = alloc_<int32_t>(std::max<int64_t>(ntl, 1));
  W_.tbase = alloc_<int64_t>(std::max<int64_t>(ntl, 1));
  W_.ioff = alloc_<int32_t>(cap_total_);
  W_.nbc2 = alloc_<int32_t>(cap_total_);
  W_.lane = alloc_<uint8_t>(cap_total_);
  W_.ctr = alloc_<unsigned long long>(2);
  if (alloc_failed_) { set_error("setup: device allocation failed (hipMalloc)"); return SMG_ERR_HIP; }
  W_.jp = d_jp_;
  W_.tile_job = d_tile_job_;
  W_.sblk_job = d_sblk_job_;
  W_.red = d_red_;
  W_.acc = d_acc_;
  W_.lsp = d_lsp_;
  W_.lsn = d_lsn_;
  W_.lst = d_lst_;
  W_.lpm = d_lpm_;
  W_.cap_total = cap_total_;
  W_.n_tbl = n_tbl_;
  W_.K = K_;
  W_.tile0 = 0;
  SMG_TRY(clear_reds_(0, njobs));
  HIP_TRY(hipMemsetAsync(d_jp_, 0, sizeof(JobParams) * njobs, st_));   // n = 0: no job
  HIP_TRY(hipMemsetAsync(d_tile_job_, 0xFF, sizeof(int32_t) * (slots / kTile), st_));   // -1: no job
  sblk_job_h_.assign((tbl + kScanItems - 1) / kScanItems + 1, -1);
  HIP_TRY(hipMemsetAsync(d_sblk_job_, 0xFF, sizeof(int32_t) * sblk_job_h_.size(), st_));
  HIP_TRY(hipMemsetAsync(W_.counts, 0, sizeof(int32_t) * n_tbl_, st_));
  HIP_TRY(hipMemsetAsync(W_.starts, 0, sizeof(int32_t) * (n_tbl_ + 1), st_));
  HIP_TRY(hipMemsetAsync(W_.nbc, 0, sizeof(int32_t) * cap_total_, st_));
  HIP_TRY(hipMemsetAsync(W_.ctr, 0, 2 * sizeof(unsigned long long), st_));
  HIP_TRY(hipMemsetAsync(W_.nbo, 0, sizeof(int32_t) * cap_total_, st_));
  HIP_TRY(hipMemsetAsync(W_.ntrue, 0, sizeof(int32_t) * cap_total_, st_));
  HIP_TRY(hipMemsetAsync(W_.recb, 0, sizeof(float4) * (cap_total_ + kRecbPad), st_));
  HIP_TRY(hipMemsetAsync(W_.hcnt, 0, sizeof(int32_t) * std::max<int64_t>(ntl, 1), st_));
  HIP_TRY(hipMemsetAsync(W_.ioff, 0, sizeof(int32_t) * cap_total_, st_));
  HIP_TRY(hipMemsetAsync(W_.nbc2, 0, sizeof(int32_t) * cap_total_, st_));
  HIP_TRY(hipMemsetAsync(W_.lane, 0, cap_total_, st_));
  HIP_TRY(hipMemsetAsync(W_.nu_a, 0, sizeof(float) * cap_total_, st_));
  HIP_TRY(hipMemsetAsync(W_.nu_b, 0, sizeof(float) * cap_total_, st_));
  HIP_TRY(hipMemsetAsync(W_.sum, 0, sizeof(float) * cap_total_, st_));
  HIP_TRY(hipMemsetAsync(W_.rec, 0, 4 * sizeof(float4) * cap_total_, st_));
  for (int a = 0; a < 3; a++) {
    HIP_TRY(hipMemsetAsync(W_.a1[a], 0, sizeof(float) * cap_total_, st_));
    HIP_TRY(hipMemsetAsync(W_.a2[a], 0, sizeof(float) * cap_total_, st_));
  }
  red_h_.assign(njobs, JobRed{});
  acc_h_.assign(size_t(njobs) * kAccCopies, JobAcc{});
  return SMG_OK;
}

// reductions and their accumulator copies of jobs [j0, j0 + nj) to zero (dt = 0: the
// first step after Pre_run moves nothing)
int Engine::clear_reds_(int j0, int nj) {
  HIP_TRY(hipMemsetAsync(d_red_ + j0, 0, sizeof(JobRed) * nj, st_));
  HIP_TRY(hipMemsetAsync(d_acc_ + int64_t(j0) * kAccCopies, 0, sizeof(JobAcc) * nj * kAccCopies, st_));
  return SMG_OK;
}

// the host copy of job j's reductions with the accumulator copies folded in
void Engine::fold_acc_(int j) {
  JobRed& R = red_h_[j];
  const JobAcc* A = acc_h_.data() + size_t(j) * kAccCopies;
  for (int p = 0; p < 2; p++) {
    R.dt1[p] = R.dt2[p] = ~0ull;
    R.maxv[p] = R.res[p] = R.maxdisp[p] = R.rsum[p] = 0ull;
  }
  R.rcum = R.drift = 0ull;
  R.maxnb = 0;
  for (int c = 0; c < kAccCopies; c++) {
    for (int p = 0; p < 2; p++) {   // non-negative doubles order as their bit patterns
      R.dt1[p] = std::min(R.dt1[p], A[c].dt1[p]);
      R.dt2[p] = std::min(R.dt2[p], A[c].dt2[p]);
      R.maxv[p] = std::max(R.maxv[p], A[c].maxv[p]);
      R.res[p] = std::max(R.res[p], A[c].res[p]);
      R.maxdisp[p] = std::max(R.maxdisp[p], A[c].maxdisp[p]);
      R.rsum[p] += A[c].rsum[p];
    }
    R.rcum += A[c].rcum;
    R.drift = std::max(R.drift, A[c].drift);
    R.maxnb = std::max(R.maxnb, A[c].maxnb);
  }
}

// level-set cells, parameters and tile ownership of job j (ranges already set in jp_[j])
int Engine::upload_job_(int j, const JobInput& in) {
  jp_[j].spill_base = jp_[j].p0 * sps_;
  jp_[j].spill_cap = jp_[j].cap * sps_;
  const JobParams& J = jp_[j];
  const int64_t m = job_lset_cells(J);
  std::vector<double> hp(m, 0.);   // the padding stays zero
  std::vector<float4> hn(m, make_float4(0.f, 0.f, 0.f, 0.f));
  int64_t q = 0;   // the caller's grid: x fastest
  for (int c = 0; c < J.ng[2]; c++)
    for (int b = 0; b < J.ng[1]; b++)
      for (int a = 0; a < J.ng[0]; a++, q++) {
        const int64_t o = ls_cell_offset(J.D, J.ng, a, b, c);
        hp[o] = in.phi[q];
        int32_t tag = in.tag[q];
        float tf;
        std::memcpy(&tf, &tag, sizeof(tf));
        hn[o] = make_float4(in.normal[3 * q + 0], in.normal[3 * q + 1], in.normal[3 * q + 2], tf);
      }
  // refined levels: package maps, and the bricks of the stored packages in map order
  // (cells outside the level's one-cell-halo grid stay zero; they are never sampled)
  std::vector<int32_t> hm(job_lmap_entries(J), -1);
  for (int l = 0; l < J.ls_nl; l++) {
    const LevelInput& L = in.levels[l];
    int np[3] = {1, 1, 1}, ngl[3] = {1, 1, 1};
    for (int a = 0; a < J.D; a++) {
      np[a] = J.lnp[l][a] + 2;
      ngl[a] = 4 * J.lnp[l][a] + 2;
    }
    const int cpb = J.D > 2 ? 64 : 16;
    int64_t b = 0, e = J.lmap_off[l];
    for (int p2 = 0; p2 < np[2]; p2++)
      for (int p1 = 0; p1 < np[1]; p1++)
        for (int p0 = 0; p0 < np[0]; p0++, e++) {
          if (!level_pkg_active(J, l, L, p0, p1, p2)) continue;
          hm[e] = int32_t(b);
          const int64_t base = J.lcell_off[l] + b * cpb;
          for (int c2 = 0; c2 < (J.D > 2 ? 4 : 1); c2++)
            for (int c1 = 0; c1 < 4; c1++)
              for (int c0 = 0; c0 < 4; c0++) {
                // global cell 4 (p - 1) + c, at index + 1 on the one-cell-halo grid
                const int g[3] = {4 * (p0 - 1) + c0 + 1, 4 * (p1 - 1) + c1 + 1, J.D > 2 ? 4 * (p2 - 1) + c2 + 1 : 0};
                bool in_grid = true;
                for (int a = 0; a < J.D; a++) in_grid = in_grid && g[a] >= 0 && g[a] < ngl[a];
                if (!in_grid) continue;
                const int64_t qq = (int64_t(g[2]) * ngl[1] + g[1]) * ngl[0] + g[0];
                const int64_t o = base + (c2 * 4 + c1) * 4 + c0;
                hp[o] = L.phi[qq];
                int32_t tag = L.tag[qq];
                float tf;
                std::memcpy(&tf, &tag, sizeof(tf));
                hn[o] = make_float4(L.normal[3 * qq + 0], L.normal[3 * qq + 1], L.normal[3 * qq + 2], tf);
              }
          b++;
        }
  }
  HIP_TRY(hipMemcpyAsync(d_lsp_ + J.lset_off, hp.data(), sizeof(double) * m, hipMemcpyHostToDevice, st_));
  HIP_TRY(hipMemcpyAsync(d_lsn_ + J.lset_off, hn.data(), sizeof(float4) * m, hipMemcpyHostToDevice, st_));
  std::vector<int8_t> ht(m);
  for (int64_t o = 0; o < m; o++) {
    int32_t tag;
    std::memcpy(&tag, &hn[o].w, sizeof(tag));
    ht[o] = int8_t(tag);
  }
  HIP_TRY(hipMemcpyAsync(d_lst_ + J.lset_off, ht.data(), size_t(m), hipMemcpyHostToDevice, st_));
  if (!hm.empty())
    HIP_TRY(hipMemcpyAsync(d_lpm_ + J.lmap_base, hm.data(), sizeof(int32_t) * hm.size(), hipMemcpyHostToDevice, st_));
  HIP_TRY(hipStreamSynchronize(st_));   // the host buffers go out of scope
  HIP_TRY(hipMemcpyAsync(d_jp_ + j, &J, sizeof(JobParams), hipMemcpyHostToDevice, st_));
  const int64_t t0 = J.p0 / kTile, nt = J.cap / kTile;
  for (int64_t t = t0; t < t0 + nt; t++) tile_job_h_[t] = int32_t(j);
  HIP_TRY(hipMemcpyAsync(d_tile_job_ + t0, tile_job_h_.data() + t0, sizeof(int32_t) * nt, hipMemcpyHostToDevice, st_));
  SMG_TRY(set_table_owner_(J, int32_t(j)));
  // the host staging buffer must outlive the copy
  HIP_TRY(hipStreamSynchronize(st_));
  return SMG_OK;
}

int Engine::begin_setup_(int D, int max_neighbors, hipStream_t stream) {
  D_ = D;
  if (D_ != 2 && D_ != 3) { set_error("setup: dim must be 2 or 3"); return SMG_ERR_ARG; }
  K_ = max_neighbors > 0 ? (max_neighbors + 3) / 4 * 4 : 128;   // whole groups of four entries
  if (stream) st_ = stream;
  else {
    HIP_TRY(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
    own_stream_ = true;
  }
  return SMG_OK;
}

int Engine::setup(const std::vector<JobInput>& jobs, int max_neighbors, int spill_per_slot, hipStream_t stream) {
  if (jobs.empty()) { set_error("setup: no jobs"); return SMG_ERR_ARG; }
  SMG_TRY(begin_setup_(jobs[0].cfg.dim, max_neighbors, stream));
  jp_.assign(jobs.size(), JobParams{});
  active_.assign(jobs.size(), 1);
  int64_t p0 = 0, tbl = 0, ls = 0, maps = 0;
  for (size_t j = 0; j < jobs.size(); j++) {
    JobParams& J = jp_[j];
    SMG_TRY(make_job_params(jobs[j], D_, J));
    J.tbl_base = tbl;
    tbl += table_blocks(J) * kScanItems;
    J.lset_off = ls;
    ls += job_lset_cells(J);
    J.lmap_base = maps;
    maps += job_lmap_entries(J);
    J.p0 = p0;
    p0 += J.cap;
  }
  tile_job_h_.assign(p0 / kTile, -1);
  SMG_TRY(alloc_buffers_(p0, tbl, ls, int(jobs.size()), spill_per_slot, maps));
  for (size_t j = 0; j < jobs.size(); j++) SMG_TRY(upload_job_(int(j), jobs[j]));
  ntiles_ = p0 / kTile;
  HIP_TRY(hipStreamSynchronize(st_));
  return SMG_OK;
}

// ---------------------------------------------------------------------------
// pool mode: capacity allocated once, jobs admitted into free ranges and retired

int64_t RangeSet::take(int64_t size) {
  for (auto it = free_.begin(); it != free_.end(); ++it) {
    if (it->second >= size) {
      const int64_t at = it->first, len = it->second;
      free_.erase(it);
      if (len > size) free_[at + size] = len - size;
      return at;
    }
  }
  return -1;
}
void RangeSet::give(int64_t at, int64_t size) {
  auto nx = free_.lower_bound(at);
  if (nx != free_.end() && at + size == nx->first) {   // merge with the next range
    size += nx->second;
    nx = free_.erase(nx);
  }
  if (nx != free_.begin()) {
    auto pv = std::prev(nx);
    if (pv->first + pv->second == at) {   // merge with the previous range
      pv->second += size;
      return;
    }
  }
  free_[at] = size;
}
int64_t RangeSet::high_water() const {
  // end of the last range in use = start of a trailing free range that reaches the end
  if (!free_.empty()) {
    auto last = std::prev(free_.end());
    if (last->first + last->second == total_) return last->first;
  }
  return total_;
}

int Engine::setup_pool(int D, int64_t max_slots, int64_t max_table, int64_t max_lset_cells, int max_jobs,
                       int max_neighbors, int spill_per_slot, hipStream_t stream, int64_t max_lset_maps) {
  if (max_jobs <= 0 || max_slots < kTile || max_table <= 0 || max_lset_cells <= 0 || max_lset_maps < 0) {
    set_error("setup_pool: capacities must be positive");
    return SMG_ERR_ARG;
  }
  SMG_TRY(begin_setup_(D, max_neighbors, stream));
  const int64_t slots = max_slots / kTile * kTile;
  jp_.assign(max_jobs, JobParams{});
  active_.assign(max_jobs, 0);
  tile_job_h_.assign(slots / kTile, -1);
  SMG_TRY(alloc_buffers_(slots, max_table, max_lset_cells, max_jobs, spill_per_slot, max_lset_maps));
  is_pool_ = true;
  tiles_free_.reset(slots / kTile);
  tbl_free_.reset(max_table / kScanItems);   // in 4096-entry scan blocks
  ls_free_.reset(max_lset_cells);
  maps_free_.reset(max_lset_maps);
  ntiles_ = 0;
  W_.n_tbl = 0;
  prerun_done_ = true;
  HIP_TRY(hipStreamSynchronize(st_));
  return SMG_OK;
}

void Engine::update_extent_() {
  ntiles_ = tiles_free_.high_water();
  W_.n_tbl = tbl_free_.high_water() * kScanItems;
}

// Seeds the job into its ranges, then Set_particle_info, sort, Pre_run and sort, all
// launched over the job's own tiles; the other jobs' state is untouched.
int Engine::admit(const JobInput& in, int* job) {
  if (!is_pool_) { set_error("admit: engine is not a pool"); return SMG_ERR_STATE; }
  int j = -1;
  for (int q = 0; q < int(active_.size()); q++)
    if (!active_[q]) { j = q; break; }
  if (j < 0) { set_error("admit: all job slots in use"); return SMG_ERR_CAPACITY; }
  JobParams J;
  SMG_TRY(make_job_params(in, D_, J));
  const int64_t nmap = job_lmap_entries(J);
  if (nmap > 0 && maps_free_.total() == 0) {
    set_error("admit: refined level sets need a pool with level-set map entries (max_lset_maps)");
    return SMG_ERR_ARG;
  }
  const int64_t nt = J.cap / kTile, nls = job_lset_cells(J);
  const int64_t t0 = tiles_free_.take(nt);
  if (t0 < 0) { set_error("admit: no free slot range"); return SMG_ERR_CAPACITY; }
  const int64_t tblocks = table_blocks(J);
  const int64_t tb = tbl_free_.take(tblocks);
  if (tb < 0) { tiles_free_.give(t0, nt); set_error("admit: no free cell-table range"); return SMG_ERR_CAPACITY; }
  const int64_t lo = ls_free_.take(nls);
  if (lo < 0) {
    tiles_free_.give(t0, nt);
    tbl_free_.give(tb, tblocks);
    set_error("admit: no free level-set range");
    return SMG_ERR_CAPACITY;
  }
  const int64_t mo = nmap > 0 ? maps_free_.take(nmap) : 0;
  if (mo < 0) {
    tiles_free_.give(t0, nt);
    tbl_free_.give(tb, tblocks);
    ls_free_.give(lo, nls);
    set_error("admit: no free level-set map range");
    return SMG_ERR_CAPACITY;
  }
  J.p0 = t0 * kTile;
  J.tbl_base = tb * kScanItems;
  J.lset_off = lo;
  J.lmap_base = mo;
  jp_[j] = J;
  active_[j] = 1;
  SMG_TRY(upload_job_(j, in));
  update_extent_();
  const JobRed zero{};
  JobRed sorting{};
  sorting.rebuild = 1;
  HIP_TRY(hipMemcpyAsync(d_red_ + j, &sorting, sizeof(JobRed), hipMemcpyHostToDevice, st_));
  HIP_TRY(hipMemsetAsync(W_.counts + J.tbl_base, 0, sizeof(int32_t) * tblocks * kScanItems, st_));
  Work Ws = W_;
  Ws.tile0 = t0;
  launch_seed(st_, D_, nt, Ws, A_);
  launch_map(st_, D_, nt, Ws, A_, kModeInit, 0);
  SMG_TRY(sort_job_(j));
  launch_map(st_, D_, nt, Ws, A_, kModePrerun, 0);
  SMG_TRY(sort_job_(j));
  const size_t fb = sizeof(float) * J.cap;
  HIP_TRY(hipMemsetAsync(W_.recb + J.p0, 0, sizeof(float4) * J.cap, st_));
  HIP_TRY(hipMemsetAsync(W_.nu_a + J.p0, 0, fb, st_));
  HIP_TRY(hipMemsetAsync(W_.nu_b + J.p0, 0, fb, st_));
  HIP_TRY(hipMemsetAsync(W_.sum + J.p0, 0, fb, st_));
  for (int a = 0; a < 3; a++) {
    HIP_TRY(hipMemsetAsync(W_.a1[a] + J.p0, 0, fb, st_));
    HIP_TRY(hipMemsetAsync(W_.a2[a] + J.p0, 0, fb, st_));
  }
  HIP_TRY(hipGetLastError());
  JobRed r;
  SMG_TRY(red_of(j, &r));
  if (r.fail) {
    SMG_TRY(retire(j));
    set_error("admit: no positive-phase position found within remap_max_attempts");
    return SMG_ERR_REMAP;
  }
  // the job starts from clean reductions, as after Pre_run of a job of its own
  HIP_TRY(hipMemcpyAsync(d_red_ + j, &zero, sizeof(JobRed), hipMemcpyHostToDevice, st_));
  HIP_TRY(hipMemsetAsync(d_acc_ + int64_t(j) * kAccCopies, 0, sizeof(JobAcc) * kAccCopies, st_));
  HIP_TRY(hipStreamSynchronize(st_));
  *job = j;
  return SMG_OK;
}

// counting sort of one job's slots inside its own table segment (its reductions
// carry rebuild = 1)
int Engine::sort_job_(int j) {
  const JobParams& J = jp_[j];
  const int64_t t0 = J.p0 / kTile, nt = J.cap / kTile;
  // the rebuild-only passes' any-job gate (the next step's k_reset clears it)
  static const unsigned long long kOne = 1ull;
  HIP_TRY(hipMemcpyAsync(W_.ctr + 1, &kOne, sizeof(kOne), hipMemcpyHostToDevice, st_));
  Work Ws = W_;
  Ws.tile0 = t0;
  launch_key(st_, D_, nt, Ws, A_);
  Work Wt = W_;
  Wt.counts = W_.counts + J.tbl_base;
  Wt.starts = W_.starts + J.tbl_base;
  Wt.n_tbl = J.tbl_size;
  launch_scan(st_, Wt, false, false);
  launch_scatter(st_, nt, Ws, A_);
  launch_segsort(st_, nt, Ws, A_);
  launch_gather(st_, D_, nt, Ws, A_, B_);
  launch_copyback(st_, D_, nt, Ws, A_, B_);
  HIP_TRY(hipGetLastError());
  return SMG_OK;
}

int Engine::retire(int j) {
  if (!is_pool_ || j < 0 || j >= int(active_.size()) || !active_[j]) { set_error("retire: no such job"); return SMG_ERR_ARG; }
  const JobParams& J = jp_[j];
  const int64_t t0 = J.p0 / kTile, nt = J.cap / kTile;
  for (int64_t t = t0; t < t0 + nt; t++) tile_job_h_[t] = -1;
  HIP_TRY(hipMemcpyAsync(d_tile_job_ + t0, tile_job_h_.data() + t0, sizeof(int32_t) * nt, hipMemcpyHostToDevice, st_));
  HIP_TRY(hipStreamSynchronize(st_));
  tiles_free_.give(t0, nt);
  tbl_free_.give(J.tbl_base / kScanItems, table_blocks(J));
  ls_free_.give(J.lset_off, job_lset_cells(J));
  if (job_lmap_entries(J) > 0) maps_free_.give(J.lmap_base, job_lmap_entries(J));
  SMG_TRY(set_table_owner_(J, -1));
  active_[j] = 0;
  jp_[j].n = 0;
  HIP_TRY(hipMemcpyAsync(d_jp_ + j, &jp_[j], sizeof(JobParams), hipMemcpyHostToDevice, st_));
  HIP_TRY(hipStreamSynchronize(st_));
  update_extent_();
  return SMG_OK;
}

int Engine::red_of(int j, JobRed* r) {
  HIP_TRY(hipMemcpyAsync(&red_h_[j], d_red_ + j, sizeof(JobRed), hipMemcpyDeviceToHost, st_));
  HIP_TRY(hipMemcpyAsync(acc_h_.data() + size_t(j) * kAccCopies, d_acc_ + int64_t(j) * kAccCopies,
                         sizeof(JobAcc) * kAccCopies, hipMemcpyDeviceToHost, st_));
  HIP_TRY(hipStreamSynchronize(st_));
  fold_acc_(j);
  *r = red_h_[j];
  return SMG_OK;
}

int Engine::download_reds() { return download_red_(); }

int Engine::sample_phi(int job, int64_t n, const double* pts, double* phi) {
  if (n == 0) return SMG_OK;
  double *d_pts = nullptr, *d_phi = nullptr;
  hipError_t e = hipMalloc(&d_pts, sizeof(double) * 3 * n);
  if (e == hipSuccess) e = hipMalloc(&d_phi, sizeof(double) * n);
  if (e == hipSuccess) e = hipMemcpyAsync(d_pts, pts, sizeof(double) * 3 * n, hipMemcpyHostToDevice, st_);
  if (e == hipSuccess) {
    launch_sample_phi(st_, D_, W_, job, n, d_pts, d_phi);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(phi, d_phi, sizeof(double) * n, hipMemcpyDeviceToHost, st_);
  if (e == hipSuccess) e = hipStreamSynchronize(st_);
  (void)hipFree(d_pts);
  (void)hipFree(d_phi);
  if (e != hipSuccess) { set_error(std::string("sample_phi: ") + hipGetErrorString(e)); return SMG_ERR_HIP; }
  return SMG_OK;
}

// sample_phi without device allocations: the points and the results pass through the
// file-scope staging array (calls that overlap on one GPU share it and race)
int Engine::sample_phi_static(int job, int64_t n, const double* pts, double* phi) {
  if (n == 0) return SMG_OK;
  if (n > kStaticSampleMax) {
    set_error("smg_sample_phi_static: " + std::to_string(n) + " points exceed the staging array's " +
              std::to_string(kStaticSampleMax) + " (use smg_sample_phi, or sample in parts)");
    return SMG_ERR_CAPACITY;
  }
  double* buf = nullptr;
  if (static_sample_scratch(&buf) != 0) { set_error("smg_sample_phi_static: no staging array on this device"); return SMG_ERR_HIP; }
  double* d_pts = buf;
  double* d_phi = buf + 3 * n;
  hipError_t e = hipMemcpyAsync(d_pts, pts, sizeof(double) * 3 * n, hipMemcpyHostToDevice, st_);
  if (e == hipSuccess) {
    launch_sample_phi(st_, D_, W_, job, n, d_pts, d_phi);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(phi, d_phi, sizeof(double) * n, hipMemcpyDeviceToHost, st_);
  if (e == hipSuccess) e = hipStreamSynchronize(st_);
  if (e != hipSuccess) { set_error(std::string("smg_sample_phi_static: ") + hipGetErrorString(e)); return SMG_ERR_HIP; }
  return SMG_OK;
}

void Engine::begin_(const char* name) {
  if (!profiling_) return;
  if (ev_pool_.size() < 2) {
    hipEvent_t e;
    for (int q = 0; q < 32; q++) {
      if (hipEventCreate(&e) == hipSuccess) ev_pool_.push_back(e);
    }
  }
  if (ev_pool_.size() < 2) return;
  cur_a_ = ev_pool_.back();
  ev_pool_.pop_back();
  cur_name_ = name;
  (void)hipEventRecord(cur_a_, st_);
}
void Engine::end_() {
  if (!profiling_ || !cur_a_) return;
  hipEvent_t b = ev_pool_.back();
  ev_pool_.pop_back();
  (void)hipEventRecord(b, st_);
  recs_.push_back({cur_name_, cur_a_, b});
  cur_a_ = nullptr;
}

// re-sort every job (set-up passes): forced rebuild flags, keys, scan, counting
// sort into B and back into A
int Engine::sort_particles() {
  launch_decide(st_, W_, int(jp_.size()), 1);
  launch_key(st_, D_, ntiles_, W_, A_);
  launch_scan(st_, W_, false, true);
  launch_scatter(st_, ntiles_, W_, A_);
  launch_segsort(st_, ntiles_, W_, A_);
  launch_gather(st_, D_, ntiles_, W_, A_, B_);
  launch_copyback(st_, D_, ntiles_, W_, A_, B_);
  HIP_TRY(hipGetLastError());
  return SMG_OK;
}

int Engine::seed() {
  SMG_TRY(clear_reds_(0, int(jp_.size())));
  launch_seed(st_, D_, ntiles_, W_, A_);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(st_));
  SMG_TRY(download_red_());
  for (auto& r : red_h_)
    if (r.fail) { set_error("seed: no positive-phase position found within remap_max_attempts"); return SMG_ERR_REMAP; }
  // Set_particle_info at the seeded positions, cell keys, first sort
  launch_map(st_, D_, ntiles_, W_, A_, kModeInit, 0);
  HIP_TRY(hipGetLastError());
  SMG_TRY(sort_particles());
  HIP_TRY(hipStreamSynchronize(st_));
  step_ = 0;
  prerun_done_ = false;
  return SMG_OK;
}

int Engine::init_positions(int job, const double* x, const int32_t* type) {
  const JobParams& J = jp_[job];
  const int64_t n = J.n, cap = J.cap;
  std::vector<float> hi(cap), lo(cap);
  for (int a = 0; a < 3; a++) {
    for (int64_t i = 0; i < cap; i++) {
      const double v = (i < n && a < D_) ? x[3 * i + a] : 0.;
      hi[i] = float(v);
      lo[i] = float(v - double(hi[i]));
    }
    HIP_TRY(hipMemcpyAsync(A_.xh[a] + J.p0, hi.data(), sizeof(float) * cap, hipMemcpyHostToDevice, st_));
    HIP_TRY(hipMemcpyAsync(A_.xl[a] + J.p0, lo.data(), sizeof(float) * cap, hipMemcpyHostToDevice, st_));
    HIP_TRY(hipStreamSynchronize(st_));
  }
  std::vector<int32_t> ty(cap), id(cap);
  for (int64_t i = 0; i < cap; i++) {
    ty[i] = i < n ? (type ? type[i] : SMG_REAL) : -1;
    id[i] = i < n ? int32_t(i) : -1;
  }
  HIP_TRY(hipMemcpyAsync(A_.type + J.p0, ty.data(), sizeof(int32_t) * cap, hipMemcpyHostToDevice, st_));
  HIP_TRY(hipMemcpyAsync(A_.id + J.p0, id.data(), sizeof(int32_t) * cap, hipMemcpyHostToDevice, st_));
  HIP_TRY(hipStreamSynchronize(st_));
  return SMG_OK;
}

int Engine::init_finish() {
  // Set_particle_info at the given positions, cell keys, first sort
  SMG_TRY(clear_reds_(0, int(jp_.size())));
  launch_map(st_, D_, ntiles_, W_, A_, kModeInit, 0);
  HIP_TRY(hipGetLastError());
  SMG_TRY(sort_particles());
  HIP_TRY(hipStreamSynchronize(st_));
  step_ = 0;
  prerun_done_ = false;
  return SMG_OK;
}

int Engine::prerun() {
  // Pre_run: Get_particle_info, Set_particle_scale, Find_and_map_particle; the force
  // passes that follow see sum = 0 and nu = 0 and leave a = 0, v = 0, nu = 0.
  SMG_TRY(clear_reds_(0, int(jp_.size())));
  begin_("k_map");
  launch_map(st_, D_, ntiles_, W_, A_, kModePrerun, 0);
  end_();
  HIP_TRY(hipGetLastError());
  SMG_TRY(sort_particles());
  HIP_TRY(hipMemsetAsync(W_.nu_a, 0, sizeof(float) * cap_total_, st_));
  HIP_TRY(hipMemsetAsync(W_.nu_b, 0, sizeof(float) * cap_total_, st_));
  for (int a = 0; a < 3; a++) {
    HIP_TRY(hipMemsetAsync(W_.a1[a], 0, sizeof(float) * cap_total_, st_));
    HIP_TRY(hipMemsetAsync(W_.a2[a], 0, sizeof(float) * cap_total_, st_));
  }
  HIP_TRY(hipStreamSynchronize(st_));
  SMG_TRY(download_red_());
  // the dt slots stay 0 so the first step moves nothing (v = 0 after Pre_run)
  for (auto& r : red_h_)
    if (r.fail) { set_error("prerun: re-seeding found no positive-phase position"); return SMG_ERR_REMAP; }
  step_ = 0;
  prerun_done_ = true;
  return SMG_OK;
}

// The candidate lists on the positions Pre_run left, without a step: its sort left the
// cell tables, the build records and every job's rebuild flag (k_decide with force), so
// the list passes alone run.  The 2D mesh stage reads these lists; a step would move
// the particles (its mapping pass) before they are meshed.
int Engine::build_lists() {
  if (!prerun_done_ || step_ != 0) { set_error("build_lists: only between prerun and the first step"); return SMG_ERR_STATE; }
  if (ntiles_ == 0) return SMG_OK;
  HIP_TRY(hipMemsetAsync(W_.ctr, 0, sizeof(unsigned long long), st_));   // the spill queue
  launch_nb(st_, D_, ntiles_, W_, A_);
  launch_nb_spill(st_, D_, ntiles_, W_, A_);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(st_));
  return check_health();
}

int Engine::step(int64_t nsteps) {
  if (!prerun_done_) { set_error("step: call prerun (or seed) first"); return SMG_ERR_STATE; }
  if (ntiles_ == 0) return SMG_OK;   // an empty pool
  for (int64_t s = 0; s < nsteps; s++) {
    const int64_t idx = step_ + 1;
    const int par = int(idx & 1);
    launch_reset(st_, W_, int(jp_.size()), par);
    begin_("k_map");
    launch_map(st_, D_, ntiles_, W_, A_, kModeStep, par);
    end_();
    // jobs whose drift reached the skin re-sort and rebuild their lists; the other
    // jobs' tiles return at once from the passes up to k_nb_spill
    begin_("k_jump_lists");
    launch_jump_lists(st_, D_, W_, A_, int(jp_.size()));
    end_();
    begin_("k_decide");
    launch_decide(st_, W_, int(jp_.size()), 0);
    end_();
    begin_("k_key");
    launch_key(st_, D_, ntiles_, W_, A_);
    end_();
    begin_("k_scan");
    launch_scan(st_, W_, true, true);
    end_();
    begin_("k_scatter");
    launch_scatter(st_, ntiles_, W_, A_);
    end_();
    begin_("k_segsort");
    launch_segsort(st_, ntiles_, W_, A_);
    end_();
    begin_("k_gather");
    launch_gather(st_, D_, ntiles_, W_, A_, B_);
    end_();
    begin_("k_copyback");
    launch_copyback(st_, D_, ntiles_, W_, A_, B_);
    end_();
    begin_("k_nb");
    launch_nb(st_, D_, ntiles_, W_, A_);
    end_();
    begin_("k_nb_spill");
    launch_nb_spill(st_, D_, ntiles_, W_, A_);
    end_();
    begin_("k_halo");
    launch_halo(st_, D_, ntiles_, W_, A_);
    end_();
    begin_("k_force1");
    launch_force1(st_, D_, ntiles_, W_, A_, par);
    end_();
    begin_("k_force2");
    launch_force2(st_, D_, ntiles_, W_, A_, par);
    end_();
    HIP_TRY(hipGetLastError());
    step_ = idx;
  }
  return SMG_OK;
}

int Engine::capture_graph(int64_t steps_per_graph) {
  if (steps_per_graph <= 0 || (steps_per_graph & 1)) { set_error("capture_graph: need an even step count"); return SMG_ERR_ARG; }
  if (graph_exec_) { (void)hipGraphExecDestroy(graph_exec_); graph_exec_ = nullptr; }
  if (graph_) { (void)hipGraphDestroy(graph_); graph_ = nullptr; }
  const bool prof = profiling_;
  profiling_ = false;
  const int64_t save = step_;
  HIP_TRY(hipStreamBeginCapture(st_, hipStreamCaptureModeThreadLocal));
  int rc = step(steps_per_graph);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(st_, &g);
  profiling_ = prof;
  step_ = save;   // capture records only; nothing ran
  if (rc != SMG_OK) return rc;
  if (e != hipSuccess) { set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(e)); return SMG_ERR_HIP; }
  graph_ = g;
  HIP_TRY(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0));
  graph_steps_ = steps_per_graph;
  graph_par_ = int(save & 1);
  graph_tiles_ = ntiles_;
  graph_tbl_ = W_.n_tbl;
  return SMG_OK;
}

int Engine::step_graphed(int64_t nsteps, int64_t G) {
  if (G <= 0 || (G & 1) || nsteps % G != 0 || (step_ & 1) || ntiles_ == 0) return step(nsteps);
  if (!graph_exec_ || graph_steps_ != G || graph_par_ != 0 || graph_tiles_ != ntiles_ || graph_tbl_ != W_.n_tbl)
    SMG_TRY(capture_graph(G));
  return run_graph(nsteps / G);
}

// Each replay runs graph_steps_ steps; the parity and the A/B buffer roles repeat
// every two steps, and the re-seeding RNG stream comes from the device step counter,
// so a replay is the same computation as the eager steps it stands for, provided the
// step index has the parity it had at capture (the parity picks the dt / max v slots
// the captured kernels read and write).
int Engine::run_graph(int64_t launches) {
  if (!graph_exec_) { set_error("run_graph: no graph"); return SMG_ERR_STATE; }
  if (int(step_ & 1) != graph_par_) {
    set_error("run_graph: the step index parity differs from the parity at capture (take one eager step or capture again)");
    return SMG_ERR_STATE;
  }
  for (int64_t q = 0; q < launches; q++) {
    HIP_TRY(hipGraphLaunch(graph_exec_, st_));
    step_ += graph_steps_;
  }
  return SMG_OK;
}

int Engine::sync() {
  HIP_TRY(hipStreamSynchronize(st_));
  return SMG_OK;
}

int Engine::query(bool* busy) {
  const hipError_t e = hipStreamQuery(st_);
  if (e == hipSuccess) { *busy = false; return SMG_OK; }
  if (e == hipErrorNotReady) { *busy = true; return SMG_OK; }
  set_error(std::string("hipStreamQuery: ") + hipGetErrorString(e));
  return SMG_ERR_HIP;
}

int Engine::download_red_() {
  HIP_TRY(hipMemcpyAsync(red_h_.data(), d_red_, sizeof(JobRed) * jp_.size(), hipMemcpyDeviceToHost, st_));
  HIP_TRY(hipMemcpyAsync(acc_h_.data(), d_acc_, sizeof(JobAcc) * acc_h_.size(), hipMemcpyDeviceToHost, st_));
  HIP_TRY(hipStreamSynchronize(st_));
  for (size_t j = 0; j < red_h_.size(); j++) fold_acc_(int(j));
  return SMG_OK;
}

int Engine::check_health() {
  SMG_TRY(download_red_());
  for (size_t j = 0; j < red_h_.size(); j++) {
    if (!active_[j]) continue;
    if (red_h_[j].fail) { set_error("job " + std::to_string(j) + ": re-seeding failed"); return SMG_ERR_REMAP; }
    const int ov = red_h_[j].overflow;
    if (ov & kOverSpill) {
      set_error("job " + std::to_string(j) + ": neighbour spill pool exhausted (longest list " +
                std::to_string(red_h_[j].maxnb) + ", inline " + std::to_string(K_) + ", pool " +
                std::to_string(W_.spill_cap) + " entries); raise max_neighbors or spill_per_slot");
      return SMG_ERR_CAPACITY;
    }
    if (ov & kOverItems) {
      set_error("job " + std::to_string(j) + ": its lists, padded to each wavefront's longest, exceed its " +
                std::to_string(W_.ips) +
                " pair items per particle (max_neighbors + spill_per_slot); raise max_neighbors or spill_per_slot");
      return SMG_ERR_CAPACITY;
    }
    if (ov) {
      set_error("job " + std::to_string(j) + ": a tile's neighbours exceed its " + std::to_string(W_.ht) +
                " halo entries; raise max_neighbors or spill_per_slot");
      return SMG_ERR_CAPACITY;
    }
  }
  return SMG_OK;
}

static double bits_to_double(unsigned long long b) {
  double d;
  std::memcpy(&d, &b, sizeof(d));
  return d;
}

int Engine::get_normals(int job, float* nrm) {
  // the level-set normal interpolated at every particle's current position (the
  // particle's norm of the reference's Get_particle_info; the relaxation itself samples
  // it only in cut cells)
  const int64_t n = jp_[job].n;
  if (n == 0) return SMG_OK;
  std::vector<double> x(3 * size_t(n)), nd(3 * size_t(n));
  SMG_TRY(get_state(job, x.data(), nullptr, nullptr, nullptr, nullptr));
  double *d_pts = nullptr, *d_phi = nullptr, *d_n = nullptr;
  hipError_t e = hipMalloc(&d_pts, sizeof(double) * 3 * n);
  if (e == hipSuccess) e = hipMalloc(&d_phi, sizeof(double) * n);
  if (e == hipSuccess) e = hipMalloc(&d_n, sizeof(double) * 3 * n);
  if (e == hipSuccess) e = hipMemcpyAsync(d_pts, x.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice, st_);
  if (e == hipSuccess) {
    launch_sample_phi(st_, D_, W_, job, n, d_pts, d_phi, d_n);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(nd.data(), d_n, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, st_);
  if (e == hipSuccess) e = hipStreamSynchronize(st_);
  (void)hipFree(d_pts);
  (void)hipFree(d_phi);
  (void)hipFree(d_n);
  if (e != hipSuccess) { set_error(std::string("get_normals: ") + hipGetErrorString(e)); return SMG_ERR_HIP; }
  for (size_t q = 0; q < nd.size(); q++) nrm[q] = float(nd[q]);
  return SMG_OK;
}

int Engine::get_state(int job, double* x, double* v, int32_t* type, double* h, float* nu) {
  const JobParams& J = jp_[job];
  const int64_t cap = J.cap, n = J.n;
  std::vector<int32_t> id(cap), ty(cap);
  std::vector<double> hh(cap);
  std::vector<float> xh(cap), xl(cap), a1(cap), a2(cap), nn(cap);
  HIP_TRY(hipStreamSynchronize(st_));
  HIP_TRY(hipMemcpy(id.data(), A_.id + J.p0, sizeof(int32_t) * cap, hipMemcpyDeviceToHost));
  if (type) {
    HIP_TRY(hipMemcpy(ty.data(), A_.type + J.p0, sizeof(int32_t) * cap, hipMemcpyDeviceToHost));
    for (int64_t s = 0; s < n; s++) type[id[s]] = ty[s];
  }
  if (h) {
    HIP_TRY(hipMemcpy(hh.data(), A_.h + J.p0, sizeof(double) * cap, hipMemcpyDeviceToHost));
    for (int64_t s = 0; s < n; s++) h[id[s]] = hh[s];
  }
  if (nu) {
    // after a step the current nu is the one the stage-2 pass just wrote
    HIP_TRY(hipMemcpy(nn.data(), (step_ > 0 ? W_.nu_b : W_.nu_a) + J.p0, sizeof(float) * cap, hipMemcpyDeviceToHost));
    for (int64_t s = 0; s < n; s++) nu[id[s]] = nn[s];
  }
  if (x) {
    for (int a = 0; a < 3; a++) {
      if (a >= D_) {
        for (int64_t i = 0; i < n; i++) x[3 * i + a] = 0.;
        continue;
      }
      HIP_TRY(hipMemcpy(xh.data(), A_.xh[a] + J.p0, sizeof(float) * cap, hipMemcpyDeviceToHost));
      HIP_TRY(hipMemcpy(xl.data(), A_.xl[a] + J.p0, sizeof(float) * cap, hipMemcpyDeviceToHost));
      for (int64_t s = 0; s < n; s++) x[3 * id[s] + a] = double(xh[s]) + double(xl[s]);
    }
  }
  if (v) {
    SMG_TRY(download_red_());
    const int par = int(step_ & 1);
    const double dt1 = step_ > 0 ? bits_to_double(red_h_[job].dt1[par]) : 0.;
    const double dt2 = step_ > 0 ? bits_to_double(red_h_[job].dt2[par]) : 0.;
    for (int a = 0; a < 3; a++) {
      if (a >= D_ || step_ == 0) {
        for (int64_t i = 0; i < n; i++) v[3 * i + a] = 0.;
        continue;
      }
      HIP_TRY(hipMemcpy(a1.data(), W_.a1[a] + J.p0, sizeof(float) * cap, hipMemcpyDeviceToHost));
      HIP_TRY(hipMemcpy(a2.data(), W_.a2[a] + J.p0, sizeof(float) * cap, hipMemcpyDeviceToHost));
      for (int64_t s = 0; s < n; s++) {
        double vv = double(a1[s]) * (0.5 * dt1);
        vv = vv + double(a2[s]) * (0.5 * dt2);
        v[3 * id[s] + a] = vv;
      }
    }
  }
  return SMG_OK;
}

int Engine::get_pair_terms(int job, float* sum, float* a1, float* a2) {
  const JobParams& J = jp_[job];
  const int64_t cap = J.cap, n = J.n;
  std::vector<int32_t> id(cap);
  std::vector<float> buf(cap);
  HIP_TRY(hipStreamSynchronize(st_));
  HIP_TRY(hipMemcpy(id.data(), A_.id + J.p0, sizeof(int32_t) * cap, hipMemcpyDeviceToHost));
  if (sum) {
    HIP_TRY(hipMemcpy(buf.data(), W_.sum + J.p0, sizeof(float) * cap, hipMemcpyDeviceToHost));
    for (int64_t s = 0; s < n; s++) sum[id[s]] = buf[s];
  }
  float* outs[2] = {a1, a2};
  for (int w = 0; w < 2; w++) {
    if (!outs[w]) continue;
    for (int a = 0; a < 3; a++) {
      if (a >= D_) {
        for (int64_t i = 0; i < n; i++) outs[w][3 * i + a] = 0.f;
        continue;
      }
      HIP_TRY(hipMemcpy(buf.data(), (w == 0 ? W_.a1[a] : W_.a2[a]) + J.p0, sizeof(float) * cap, hipMemcpyDeviceToHost));
      for (int64_t s = 0; s < n; s++) outs[w][3 * id[s] + a] = buf[s];
    }
  }
  return SMG_OK;
}

// The neighbour sets of the last step, as the force passes visit them (k_nb_out runs
// the same neighbour iterator), mapped to particle ids.
int Engine::get_neighbors(int job, int64_t* offs, int32_t* ids, int32_t* codes, int64_t capacity) {
  if (step_ == 0) { set_error("get_neighbors: no step taken yet"); return SMG_ERR_STATE; }
  const JobParams& J = jp_[job];
  const int64_t cap = J.cap, n = J.n;
  std::vector<int32_t> id(cap_total_), cnt(std::max<int64_t>(n, 1));
  std::vector<int64_t> co(n + 1, 0);
  std::vector<int32_t> ent;
  int32_t* d_cnt = nullptr;
  int32_t* d_out = nullptr;
  int64_t* d_off = nullptr;
  hipError_t e = hipStreamSynchronize(st_);
  Work Ws = W_;
  Ws.tile0 = J.p0 / kTile;
  if (e == hipSuccess) e = hipMemcpy(id.data(), A_.id, sizeof(int32_t) * cap_total_, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMalloc(&d_cnt, sizeof(int32_t) * std::max<int64_t>(n, 1));
  if (e == hipSuccess) e = hipMalloc(&d_off, sizeof(int64_t) * (n + 1));
  if (e == hipSuccess) {
    launch_nb_out(st_, D_, cap / kTile, Ws, A_, 0, d_off, d_cnt, nullptr);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(st_);
  if (e == hipSuccess) e = hipMemcpy(cnt.data(), d_cnt, sizeof(int32_t) * n, hipMemcpyDeviceToHost);
  if (e == hipSuccess) {
    for (int64_t s = 0; s < n; s++) co[s + 1] = co[s] + cnt[s];
    ent.resize(std::max<int64_t>(co[n], 1));
    e = hipMemcpy(d_off, co.data(), sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice);
  }
  if (e == hipSuccess) e = hipMalloc(&d_out, sizeof(int32_t) * std::max<int64_t>(co[n], 1));
  if (e == hipSuccess) {
    launch_nb_out(st_, D_, cap / kTile, Ws, A_, 1, d_off, d_cnt, d_out);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(st_);
  if (e == hipSuccess && co[n] > 0) e = hipMemcpy(ent.data(), d_out, sizeof(int32_t) * co[n], hipMemcpyDeviceToHost);
  if (d_cnt) (void)hipFree(d_cnt);
  if (d_out) (void)hipFree(d_out);
  if (d_off) (void)hipFree(d_off);
  if (e != hipSuccess) { set_error(std::string("get_neighbors: ") + hipGetErrorString(e)); return SMG_ERR_HIP; }
  std::vector<int64_t> slot_of(n, -1);
  for (int64_t s = 0; s < n; s++) slot_of[id[J.p0 + s]] = s;
  offs[0] = 0;
  for (int64_t i = 0; i < n; i++) offs[i + 1] = offs[i] + cnt[slot_of[i]];
  if (!ids) return SMG_OK;
  if (offs[n] > capacity) { set_error("get_neighbors: capacity too small"); return SMG_ERR_CAPACITY; }
  std::vector<std::pair<int32_t, int32_t>> tmp;
  for (int64_t i = 0; i < n; i++) {
    const int64_t s = slot_of[i];
    tmp.clear();
    for (int64_t k = co[s]; k < co[s + 1]; k++) {
      const uint32_t en = uint32_t(ent[k]);
      tmp.push_back({id[nb_slot(en, J.p0)], int32_t(en >> kJBits)});
    }
    std::sort(tmp.begin(), tmp.end());
    for (size_t k = 0; k < tmp.size(); k++) {
      ids[offs[i] + k] = tmp[k].first;
      codes[offs[i] + k] = tmp[k].second;
    }
  }
  return SMG_OK;
}

// glbl_timestep after the last step's Set_timestep(2) (0 before the first step)
int Engine::timestep(int job, double* dt) {
  *dt = 0.;
  if (step_ == 0) return SMG_OK;
  JobRed r;
  SMG_TRY(red_of(job, &r));
  *dt = bits_to_double(r.dt2[step_ & 1]);
  return SMG_OK;
}

int Engine::get_stats(int job, smg_stats* st) {
  std::memset(st, 0, sizeof(*st));
  const int par = int(step_ & 1);
  if (step_ > 0) {
    launch_residual(st_, D_, ntiles_, W_, A_, int(jp_.size()), par);
    HIP_TRY(hipGetLastError());
  }
  SMG_TRY(download_red_());
  const JobParams& J = jp_[job];
  const JobRed& R = red_h_[job];
  st->step = step_;
  st->n_particle = J.n;
  std::vector<int32_t> ty(J.cap);
  HIP_TRY(hipMemcpy(ty.data(), A_.type + J.p0, sizeof(int32_t) * J.cap, hipMemcpyDeviceToHost));
  for (int64_t s = 0; s < J.n; s++) {
    st->n_real += ty[s] == SMG_REAL;
    st->n_surface += ty[s] == SMG_SURFACE;
  }
  st->n_remapped = R.nremap[par];
  st->max_neighbors_seen = R.maxnb;
  st->num_level = J.nlev;
  if (step_ > 0) {
    st->dt1 = bits_to_double(R.dt1[par]);
    st->dt2 = bits_to_double(R.dt2[par]);
    st->max_v = bits_to_double(R.maxv[par]);
    st->residual = double(R.rsum[par]) / kResScale / double(J.n);
    st->res_max = bits_to_double(R.res[par]);
    st->max_disp = bits_to_double(R.maxdisp[par]);
    std::vector<int32_t> cnt(J.cap);
    HIP_TRY(hipMemcpy(cnt.data(), W_.ntrue + J.p0, sizeof(int32_t) * J.cap, hipMemcpyDeviceToHost));
    for (int64_t s = 0; s < J.n; s++) st->n_neighbor_entries += cnt[s];
    st->n_rebuild = R.nrebuild;
    st->drift = bits_to_double(R.drift);
    st->residual_sum = double(R.rcum) / kResScale / double(J.n);
  }
  return SMG_OK;
}

int Engine::profile_get(std::vector<std::string>& names, std::vector<double>& avg_ms, std::vector<int64_t>& counts) {
  HIP_TRY(hipStreamSynchronize(st_));
  for (auto& r : recs_) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
      auto& p = prof_[r.name];
      p.first += ms;
      p.second += 1;
      last_ms_[r.name] = ms;
    }
    ev_pool_.push_back(r.a);
    ev_pool_.push_back(r.b);
  }
  recs_.clear();
  names.clear();
  avg_ms.clear();
  counts.clear();
  for (auto& kv : prof_) {
    names.push_back(kv.first);
    avg_ms.push_back(kv.second.second ? kv.second.first / double(kv.second.second) : 0.);
    counts.push_back(kv.second.second);
  }
  return SMG_OK;
}

void Engine::profile_enable(bool on) {
  // (re-)enabling starts a new measurement window
  if (on) {
    (void)hipStreamSynchronize(st_);
    for (auto& r : recs_) {
      ev_pool_.push_back(r.a);
      ev_pool_.push_back(r.b);
    }
    recs_.clear();
    prof_.clear();
  }
  profiling_ = on;
}

double Engine::last_kernel_ms(const std::string& name) const {
  auto it = last_ms_.find(name);
  return it == last_ms_.end() ? 0. : it->second;
}

}  // namespace smg
