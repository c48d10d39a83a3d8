// smg_solver.hpp -- C++ host mirror of the reference's solver interface over the C ABI
// of include/smg.h, single rank (the reference's communicator argument is dropped).
//
//   Initialization        system.h:9-20, system.cpp:12-48
//   SPH_mesh_generation   the SOLVER of glbparam.h:23-25: Initialize_case, Set_current_time,
//                         SPH_run_mesh_generation, Get_timestep, Post_processing,
//                         Output_simulation_infor (sph_mesh_generation.h)
//   Solver                solver.h:28-45: Load_solver (solver.cpp:35-78) and
//                         SPH_Mesh_Generation_Solver (solver.cpp:257-321)
//
// Error behaviour follows the reference: a failure prints a message and ends the
// process with a non-zero status (the reference calls world.abort / exit).
#pragma once
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "smg.h"

namespace smg_host {

typedef double Real;

inline void check(int rc, const char* where) {
  if (rc != SMG_OK) {
    std::fprintf(stderr, "ERROR in %s (%d): %s\n", where, rc, smg_last_error());
    std::exit(1);
  }
}

// system parameters for global control (system.h:9-20)
class Initialization {
 public:
  Real start_time = 0.0, end_time = 20000.0;
  Real output_dt = 100.0;
  int num_output = 0;
  Initialization(Real start = 0.0, Real end = 20000.0, Real dt = 100.0)
      : start_time(start), end_time(end), output_dt(dt) {
    num_output = int(std::ceil((end_time - start_time) / output_dt));
  }
  Real Get_start_time() const { return start_time; }
  Real Get_dt() const { return output_dt; }
  int Get_num_timestep() const { return num_output; }
};

// The level set the case runs on: cell-centred phi with a full margin package per
// side, from which smg_lset_build_gpu derives normals, cut-cell tags and the particle
// count.
struct LevelSetCase {
  smg_config cfg{};
  std::vector<double> phi_ext;
  int zero_margin = 1;
};

// SPH_mesh_generation as the Solver drives it (one relaxation job on the GPU)
class SPH_mesh_generation {
 public:
  int i_iter = 0;
  Real run_time = 0.0;
  Real error_global = 0.0;   // the relaxation residual (mean |v| dt2 / dx)
  int64_t total_num_particle = 0;

  ~SPH_mesh_generation() {
    if (job_) smg_destroy(job_);
  }

  // Initialize_case (sph_mesh_generation.cpp:594): level set, particle budget
  // (Get_num_particles), Random_particle_distribution and Pre_run
  void Initialize_case(const LevelSetCase& c) {
    cfg_ = c.cfg;
    int64_t m = 1;
    for (int a = 0; a < cfg_.dim; a++) m *= 4 * int64_t(cfg_.num_pkg[a]) + 2;
    phi_.assign(m, 0.);
    normal_.assign(3 * m, 0.f);
    tag_.assign(m, 0);
    // the level-set finishing on the GPU, bit-identical to smg_lset_build for the
    // uniform (tanh with minimum_dl == maximum_dl) and linear spacing profiles; a tanh
    // profile with minimum_dl != maximum_dl has masses that depend on the device tanh
    // (within 2 ulp, the particle count within 1), so it keeps the host builder and
    // its count
    if (cfg_.profile == SMG_PROFILE_TANH && cfg_.min_dl != cfg_.max_dl)
      check(smg_lset_build(&cfg_, c.phi_ext.data(), c.zero_margin, phi_.data(), normal_.data(), tag_.data(), &lset_),
            "smg_lset_build");
    else
      check(smg_lset_build_gpu(&cfg_, c.phi_ext.data(), c.zero_margin, phi_.data(), normal_.data(), tag_.data(),
                               &lset_, nullptr),
            "smg_lset_build_gpu");
    total_num_particle = lset_.num_particle;
    check(smg_create(&cfg_, &job_), "smg_create");
    check(smg_set_levelset(job_, phi_.data(), normal_.data(), tag_.data()), "smg_set_levelset");
    check(smg_seed_particles(job_, total_num_particle), "smg_seed_particles");
  }
  void Set_current_time(Real t) { run_time = t; }
  // one SPH_run_mesh_generation (sph_mesh_generation.cpp:1371-1463)
  void SPH_run_mesh_generation() {
    check(smg_step(job_, 1), "smg_step");
  }
  Real Get_timestep() {
    Real dt = 0.;
    check(smg_get_timestep(job_, &dt), "smg_get_timestep");
    return dt;
  }
  // Post_processing: the final particle state, by particle id
  void Post_processing(int n, std::vector<double>& x, std::vector<int32_t>& type) {
    (void)n;
    x.assign(3 * total_num_particle, 0.);
    type.assign(total_num_particle, 0);
    check(smg_get_state(job_, x.data(), nullptr, type.data(), nullptr, nullptr), "smg_get_state");
  }
  // The mesh stage of Post_processing (sph_mesh_generation.cpp:2409-2436) on the
  // current particles.  2D: the triangles of Local_voronoi_diagram +
  // Reconstruct_tris (3 ids per triangle, i > j > k).  3D: the tetrahedra of
  // Reconstruct_tets after the deletion rule (4 ascending ids) and their boundary
  // triangles.  delete_thres <= 0 takes the reference's tet_delete_thres (10).
  struct Mesh {
    int dim = 0;
    std::vector<int32_t> cells;     // 3 (2D) or 4 (3D) particle ids per cell
    std::vector<int32_t> surface;   // 3D boundary triangles, 3 ids each
  };
  Mesh Mesh_generation(double delete_thres = 0.) {
    Mesh m;
    m.dim = cfg_.dim;
    int64_t nc = 0;
    if (cfg_.dim == 2) {
      smg_tri_stats ts{};
      check(smg_triangulate(job_, nullptr, 0, &nc, &ts), "smg_triangulate");
      m.cells.assign(3 * static_cast<size_t>(nc), 0);
      check(smg_triangulate(job_, m.cells.data(), nc, &nc, &ts), "smg_triangulate");
      return m;
    }
    smg_tet_stats ts{};
    // a Delaunay tetrahedralisation has about 6.5 tetrahedra per point; a short
    // buffer comes back as SMG_ERR_CAPACITY with the count, and the copy is redone
    int64_t cap = 8 * total_num_particle + 16;
    m.cells.assign(4 * static_cast<size_t>(cap), 0);
    int rc = smg_tetrahedralize(job_, delete_thres, m.cells.data(), cap, &nc, &ts);
    if (rc == SMG_ERR_CAPACITY && nc > cap) {
      cap = nc;
      m.cells.assign(4 * static_cast<size_t>(cap), 0);
      rc = smg_tetrahedralize(job_, delete_thres, m.cells.data(), cap, &nc, &ts);
    }
    check(rc, "smg_tetrahedralize");
    m.cells.resize(4 * static_cast<size_t>(nc));
    int64_t ns = 0;
    check(smg_tet_surface(job_, nullptr, 0, &ns), "smg_tet_surface");
    m.surface.assign(3 * static_cast<size_t>(ns), 0);
    check(smg_tet_surface(job_, m.surface.data(), ns, &ns), "smg_tet_surface");
    return m;
  }
  // Particle_binary_writer::Output_particle: particle_rank_0_step_<n>.vtu, .pvtu and
  // particle.pvd in dir (positions, type, h, neighbour counts, normals)
  void Output_particle(const char* dir, int n) {
    std::vector<double> x(3 * total_num_particle), h(total_num_particle);
    std::vector<int32_t> type(total_num_particle), nn(total_num_particle);
    std::vector<int64_t> offs(total_num_particle + 1);
    std::vector<float> nrm(3 * total_num_particle);
    check(smg_get_state(job_, x.data(), nullptr, type.data(), h.data(), nullptr), "smg_get_state");
    check(smg_get_neighbors(job_, offs.data(), nullptr, nullptr, 0), "smg_get_neighbors");
    for (int64_t i = 0; i < total_num_particle; i++) nn[i] = int32_t(offs[i + 1] - offs[i]);
    check(smg_get_normals(job_, nrm.data()), "smg_get_normals");
    const smg_vtu_points p{total_num_particle, x.data(), type.data(), h.data(), nn.data(), nullptr, nrm.data()};
    check(smg_write_particles_vtu(dir, n, run_time, &p), "smg_write_particles_vtu");
  }
  // Save_particle_post_file (sph_mesh_generation.cpp:2624-2667): step_<run_time>_rank_0.pst
  // and a list.csv line in dir
  void Save_particle_post_file(const char* dir, int n) {
    std::vector<double> x(3 * total_num_particle);
    check(smg_get_state(job_, x.data(), nullptr, nullptr, nullptr, nullptr), "smg_get_state");
    check(smg_write_post_file(dir, int64_t(run_time), total_num_particle, x.data(), n == 0), "smg_write_post_file");
  }
  // Mesh_binary_writer::Output_mesh of a Mesh_generation result
  void Output_mesh(const char* dir, int n, const Mesh& m) {
    std::vector<double> x(3 * total_num_particle), h(total_num_particle);
    std::vector<int32_t> type(total_num_particle);
    check(smg_get_state(job_, x.data(), nullptr, type.data(), h.data(), nullptr), "smg_get_state");
    const smg_vtu_points p{total_num_particle, x.data(), type.data(), h.data(), nullptr, nullptr, nullptr};
    const int per = m.dim == 2 ? 3 : 4;
    check(smg_write_mesh_vtu(dir, n, run_time, m.dim, &p, m.cells.data(), int64_t(m.cells.size() / per)),
          "smg_write_mesh_vtu");
  }
  // Output_simulation_infor: step count, dt, residual, particle counts
  smg_stats Output_simulation_infor() {
    smg_stats st{};
    check(smg_get_stats(job_, &st), "smg_get_stats");
    error_global = st.residual;
    return st;
  }
  const smg_lset_stats& lset_stats() const { return lset_; }
  smg_job* handle() { return job_; }

 private:
  smg_config cfg_{};
  smg_lset_stats lset_{};
  std::vector<double> phi_;
  std::vector<float> normal_;
  std::vector<int8_t> tag_;
  smg_job* job_ = nullptr;
};

class Solver {
 public:
  int total_iterate_num = 0;
  Real output_dt = 0.0, timestep = 0.0, run_time = 0.0, integral_time = 0.0;
  SPH_mesh_generation current_solver;

  explicit Solver(const Initialization& Ini) : total_iterate_num(Ini.num_output), output_dt(Ini.output_dt) {}

  void Set_current_time(Real t) { run_time = t; }

  // Load_solver (solver.cpp:35-78): Initialize_case, then num_output rounds of
  // SPH_Mesh_Generation_Solver; on_output is called after each round
  template <class OnOutput>
  void Load_solver(const Initialization& Ini, const LevelSetCase& c, OnOutput on_output) {
    current_solver.Initialize_case(c);
    Real Time = Ini.Get_start_time();
    for (int n = 0; n < total_iterate_num; n++) {
      Set_current_time(Time);
      SPH_Mesh_Generation_Solver(Ini, n);
      on_output(n, current_solver);
      Time += Ini.Get_dt();
    }
  }

  // SPH_Mesh_Generation_Solver (solver.cpp:257-321): output_dt relaxation steps, one
  // per unit of integral time
  void SPH_Mesh_Generation_Solver(const Initialization& Ini, int n) {
    (void)Ini;
    integral_time = 0.0;
    current_solver.i_iter = n;
    while (integral_time < output_dt) {
      integral_time += 1.;
      current_solver.Set_current_time(run_time + integral_time);
      current_solver.SPH_run_mesh_generation();
      timestep = current_solver.Get_timestep();
    }
  }
};

}  // namespace smg_host
