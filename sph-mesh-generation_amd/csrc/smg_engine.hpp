// smg_engine.hpp -- host orchestration of the relaxation kernels.
//
// An Engine holds the device state of one or more jobs of the same dimension in one
// slot space, on one HIP stream.  A single relaxation job (smg_create) is an Engine
// with one job; a batch is an Engine with many, launched together.
#pragma once
#include <hip/hip_runtime_api.h>
#include <cstdint>
#include <array>
#include <map>
#include <string>
#include <vector>

#include "smg.h"
#include "smg_device.hpp"

namespace smg {

// a refined level-set level: its one-cell-halo grid (4 (num_pkg << l) + 2 cells per
// axis) and one byte per package, nonzero where the level is stored
struct LevelInput {
  const double* phi = nullptr;
  const float* normal = nullptr;
  const int8_t* tag = nullptr;
  const uint8_t* active = nullptr;
};

struct JobInput {
  smg_config cfg;
  int64_t n = 0;
  double max_phi = 0.;            // level-set maximum_phi (tanh profile)
  const double* phi = nullptr;    // one-cell-halo level-set grid
  const float* normal = nullptr;
  const int8_t* tag = nullptr;
  std::vector<LevelInput> levels;   // refined levels 1, 2, ... (may be empty)
};

// first-fit allocator of [0, total) ranges with coalescing on release
class RangeSet {
 public:
  void reset(int64_t total) {
    total_ = total;
    free_.clear();
    if (total > 0) free_[0] = total;
  }
  int64_t take(int64_t size);            // start of a free range of this size, or -1
  void give(int64_t at, int64_t size);
  int64_t high_water() const;            // end of the last range in use
  int64_t total() const { return total_; }
  int64_t free_total() const {
    int64_t s = 0;
    for (auto& kv : free_) s += kv.second;
    return s;
  }

 private:
  int64_t total_ = 0;
  std::map<int64_t, int64_t> free_;   // start -> length
};

int make_job_params(const JobInput& in, int D, JobParams& J);
int64_t job_lset_cells(const JobParams& J);
int64_t job_lmap_entries(const JobParams& J);
int64_t table_blocks(const JobParams& J);

class Engine {
 public:
  Engine() = default;
  ~Engine();
  Engine(const Engine&) = delete;
  Engine& operator=(const Engine&) = delete;

  // Allocates and uploads everything for the given jobs (all of dimension D).
  int setup(const std::vector<JobInput>& jobs, int max_neighbors, int spill_per_slot = 0, hipStream_t stream = nullptr);

  // pool mode: capacities allocated once; jobs enter and leave between steps
  int setup_pool(int D, int64_t max_slots, int64_t max_table, int64_t max_lset_cells, int max_jobs,
                 int max_neighbors, int spill_per_slot, hipStream_t stream = nullptr, int64_t max_lset_maps = 0);
  int admit(const JobInput& in, int* job);   // seed + Pre_run of the new job only
  int retire(int job);
  bool active(int job) const { return job >= 0 && job < int(active_.size()) && active_[job]; }
  int red_of(int job, JobRed* r);
  int download_reds();
  int sample_phi(int job, int64_t n, const double* pts, double* phi);
  // the same with the file-scope staging array in place of device allocations
  int sample_phi_static(int job, int64_t n, const double* pts, double* phi);
  const JobRed& red_host(int job) const { return red_h_[job]; }
  int64_t used_tiles() const { return is_pool_ ? tiles_free_.total() - tiles_free_.free_total() : ntiles_; }

  int seed();                                                   // Random_particle_distribution
  int init_positions(int job, const double* x, const int32_t* type);  // upload one job's positions
  int init_finish();                                            // Set_particle_info + first sort
  int prerun();                                                 // Pre_run
  int build_lists();   // candidate lists on the current positions of a job that has not stepped
  int step(int64_t nsteps);                                     // SPH_run_mesh_generation x n
  // the same steps replayed from a captured graph of G steps (G even, nsteps a multiple
  // of G, even step index), captured again when G or the launch extent changed; any
  // other case steps eagerly
  int step_graphed(int64_t nsteps, int64_t G);
  int sync();
  int query(bool* busy);
  int64_t footprint() const { return bytes_; }
  int64_t graph_steps() const { return graph_steps_; }

  // results (host, by particle id of the job)
  int get_state(int job, double* x, double* v, int32_t* type, double* h, float* nu);
  int get_pair_terms(int job, float* sum, float* a1, float* a2);
  int get_normals(int job, float* nrm);   // 3 per particle, by id
  int get_neighbors(int job, int64_t* offs, int32_t* ids, int32_t* codes, int64_t cap);
  int get_stats(int job, smg_stats* st);
  int timestep(int job, double* dt);
  int triangulate(int job, std::vector<std::array<int32_t, 3>>& tris, smg_tri_stats* st);   // smg_tri.cpp
  // Voronoi-stage edges (i > j) of the last triangulate call
  const std::vector<std::array<int32_t, 2>>& last_edges() const { return last_edges_; }
  // 3D tetrahedralisation (smg_tet.cpp) and the results of its last call
  struct TetResult {
    std::vector<std::array<int32_t, 4>> delaunay, kept;   // ascending ids, sorted
    std::vector<std::array<int32_t, 3>> surface;
    std::vector<int8_t> status;                           // by particle id
    std::vector<double> rcert;
  };
  int tetrahedralize(int job, double delete_thres, smg_tet_stats* st);
  const TetResult& last_tets() const { return tet_; }
  int check_health();   // remap failures, neighbour overflow

  void profile_enable(bool on);
  int profile_get(std::vector<std::string>& names, std::vector<double>& avg_ms, std::vector<int64_t>& counts);

  int dim() const { return D_; }
  int64_t step_index() const { return step_; }
  int num_jobs() const { return int(jp_.size()); }
  int64_t total_particles() const;
  const JobParams& params(int j) const { return jp_[j]; }
  hipStream_t stream() const { return st_; }
  double last_kernel_ms(const std::string& name) const;

  // graph-capturable step loop body (two steps so the buffer swap returns to A)
  int capture_graph(int64_t steps_per_graph);
  int run_graph(int64_t launches);
  bool has_graph() const { return graph_exec_ != nullptr; }

 private:
  int sort_particles();
  int sort_job_(int job);
  int begin_setup_(int D, int max_neighbors, hipStream_t stream);
  int alloc_buffers_(int64_t slots, int64_t tbl, int64_t ls, int njobs, int spill_per_slot, int64_t maps = 0);
  int upload_job_(int job, const JobInput& in);
  void update_extent_();
  int set_table_owner_(const JobParams& J, int32_t owner);
  void begin_(const char* name);
  void end_();
  int download_red_();
  int clear_reds_(int j0, int nj);
  void fold_acc_(int j);

  int D_ = 0;
  hipStream_t st_ = nullptr;
  bool own_stream_ = false;
  std::vector<JobParams> jp_;
  std::vector<std::array<int32_t, 2>> last_edges_;
  TetResult tet_;
  std::vector<uint8_t> active_;
  bool is_pool_ = false;
  bool alloc_failed_ = false;
  RangeSet tiles_free_, tbl_free_, ls_free_, maps_free_;
  std::vector<int32_t> tile_job_h_;
  int64_t ntiles_ = 0, cap_total_ = 0, n_tbl_ = 0, n_ls_ = 0;
  int K_ = 64;
  int sps_ = 32;   // spill entries per slot
  int64_t step_ = 0;   // last completed step index (0 = pre-run)
  bool prerun_done_ = false;

  JobParams* d_jp_ = nullptr;
  int32_t* d_tile_job_ = nullptr;
  int32_t* d_sblk_job_ = nullptr;
  std::vector<int32_t> sblk_job_h_;
  JobRed* d_red_ = nullptr;
  JobAcc* d_acc_ = nullptr;
  double* d_lsp_ = nullptr;
  float4* d_lsn_ = nullptr;
  int8_t* d_lst_ = nullptr;
  int32_t* d_lpm_ = nullptr;
  PState A_{}, B_{};
  Work W_{};
  std::vector<void*> allocs_;
  int64_t bytes_ = 0;
  std::vector<JobRed> red_h_;
  std::vector<JobAcc> acc_h_;

  // profiling
  bool profiling_ = false;
  struct Rec { std::string name; hipEvent_t a, b; };
  std::vector<Rec> recs_;
  std::vector<hipEvent_t> ev_pool_;
  std::map<std::string, std::pair<double, int64_t>> prof_;
  std::map<std::string, double> last_ms_;
  hipEvent_t cur_a_ = nullptr;
  std::string cur_name_;

  hipGraph_t graph_ = nullptr;
  hipGraphExec_t graph_exec_ = nullptr;
  int64_t graph_steps_ = 0;
  int graph_par_ = 0;   // step index parity at capture; replays need the same parity
  int64_t graph_tiles_ = 0, graph_tbl_ = 0;   // launch extent at capture

  template <typename T>
  T* alloc_(int64_t count);
};

}  // namespace smg
