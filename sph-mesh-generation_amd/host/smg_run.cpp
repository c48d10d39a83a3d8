// smg_run.cpp -- command-line driver in the shape of the reference's main.cpp
// (Initialization, Solver, Load_solver), single rank, on an analytic level set:
// the unit square (2D) or cube (3D) minus a circle / sphere at the centre.
//
//   smg_run [--dim 2|3] [--outputs N] [--output-dt STEPS] [--dx DX] [--num-pkg P]
//           [--radius R] [--csv FILE] [--mesh FILE] [--outdir DIR]
//
// Prints one line per output round (steps, dt, residual, particle counts); with
// --csv, writes the final positions and types by particle id; with --mesh, runs the
// mesh stage of Post_processing and writes its cells (triangles in 2D, tetrahedra in
// 3D, one line of particle ids each) and, in 3D, the boundary triangles after a line
// "surface N"; with --outdir, writes the reference's output files there after every
// output round (particle VTU/PVTU/PVD, .pst post file and list.csv) and, with --mesh,
// the mesh VTU of the final state.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "smg_solver.hpp"

using smg_host::Initialization;
using smg_host::LevelSetCase;
using smg_host::Solver;
using smg_host::SPH_mesh_generation;

namespace {

// cell-centred phi on the grid with one margin package (4 cells) on each side,
// x fastest; phi > 0 outside the hole (the fluid side)
std::vector<double> hole_phi(const smg_config& c, double radius) {
  int ne[3] = {1, 1, 1};
  double dpkg[3] = {1., 1., 1.}, dcell[3] = {1., 1., 1.};
  for (int a = 0; a < c.dim; a++) {
    ne[a] = 4 * c.num_pkg[a] + 8;
    dpkg[a] = c.domain[a] / double(c.num_pkg[a]);
    dcell[a] = dpkg[a] / 4.0;
  }
  std::vector<double> phi(size_t(ne[0]) * ne[1] * ne[2]);
  for (int k = 0; k < ne[2]; k++)
    for (int j = 0; j < ne[1]; j++)
      for (int i = 0; i < ne[0]; i++) {
        const int g[3] = {i - 4, j - 4, k - 4};
        double r2 = 0.;
        for (int a = 0; a < c.dim; a++) {
          // centre of global cell g = 4 * package + cell
          const int pkg = g[a] >= 0 ? g[a] / 4 : -((-g[a] + 3) / 4);
          const int cell = g[a] - 4 * pkg;
          const double x = pkg * dpkg[a] + (cell + 0.5) * dcell[a] + c.box_l[a];
          const double d = x - (c.box_l[a] + 0.5 * c.domain[a]);
          r2 += d * d;
        }
        phi[(size_t(k) * ne[1] + j) * ne[0] + i] = std::sqrt(r2) - radius;
      }
  return phi;
}

}  // namespace

int main(int argc, char** argv) {
  int dim = 2, outputs = 5, num_pkg = 32;
  double output_dt = 100., dx = 0.01, radius = 0.2;
  const char* csv = nullptr;
  const char* mesh = nullptr;
  const char* outdir = nullptr;
  for (int i = 1; i < argc; i++) {
    auto next = [&](const char* name) -> const char* {
      if (i + 1 >= argc) {
        std::fprintf(stderr, "missing value for %s\n", name);
        std::exit(2);
      }
      return argv[++i];
    };
    if (!std::strcmp(argv[i], "--dim")) dim = std::atoi(next("--dim"));
    else if (!std::strcmp(argv[i], "--outputs")) outputs = std::atoi(next("--outputs"));
    else if (!std::strcmp(argv[i], "--output-dt")) output_dt = std::atof(next("--output-dt"));
    else if (!std::strcmp(argv[i], "--dx")) dx = std::atof(next("--dx"));
    else if (!std::strcmp(argv[i], "--num-pkg")) num_pkg = std::atoi(next("--num-pkg"));
    else if (!std::strcmp(argv[i], "--radius")) radius = std::atof(next("--radius"));
    else if (!std::strcmp(argv[i], "--csv")) csv = next("--csv");
    else if (!std::strcmp(argv[i], "--mesh")) mesh = next("--mesh");
    else if (!std::strcmp(argv[i], "--outdir")) outdir = next("--outdir");
    else {
      std::fprintf(stderr, "unknown argument %s\n", argv[i]);
      return 2;
    }
  }
  LevelSetCase c;
  smg_host::check(smg_default_config(dim, &c.cfg), "smg_default_config");
  for (int a = 0; a < dim; a++) c.cfg.num_pkg[a] = num_pkg;
  c.cfg.min_dl = c.cfg.max_dl = dx;   // uniform target spacing
  c.phi_ext = hole_phi(c.cfg, radius);
  c.zero_margin = 1;   // restart-mode margins, as Load_rst_case leaves them

  Initialization Ini(0.0, output_dt * outputs, output_dt);
  std::printf("start time %g, end time %g, output dt %g, outputs %d\n", Ini.start_time, Ini.end_time, Ini.output_dt,
              Ini.num_output);
  Solver solver(Ini);
  solver.Load_solver(Ini, c, [outdir](int n, SPH_mesh_generation& s) {
    if (outdir) {
      s.Output_particle(outdir, n);
      s.Save_particle_post_file(outdir, n);
    }
    const smg_stats st = s.Output_simulation_infor();
    std::printf("output %d: steps %lld dt %.6e residual %.3e particles %lld real %lld surface %lld rebuilds %lld\n", n,
                (long long)st.step, st.dt2, st.residual, (long long)st.n_particle, (long long)st.n_real,
                (long long)st.n_surface, (long long)st.n_rebuild);
    std::fflush(stdout);
  });
  if (csv) {
    std::vector<double> x;
    std::vector<int32_t> type;
    solver.current_solver.Post_processing(Ini.num_output, x, type);
    FILE* f = std::fopen(csv, "w");
    if (!f) {
      std::fprintf(stderr, "cannot write %s\n", csv);
      return 1;
    }
    std::fprintf(f, "id,x,y,z,type\n");
    for (size_t i = 0; i < type.size(); i++)
      std::fprintf(f, "%zu,%.17g,%.17g,%.17g,%d\n", i, x[3 * i], x[3 * i + 1], x[3 * i + 2], type[i]);
    std::fclose(f);
  }
  if (mesh) {
    const SPH_mesh_generation::Mesh m = solver.current_solver.Mesh_generation();
    const size_t per = m.dim == 2 ? 3 : 4;
    FILE* f = std::fopen(mesh, "w");
    if (!f) {
      std::fprintf(stderr, "cannot write %s\n", mesh);
      return 1;
    }
    std::fprintf(f, "cells %zu\n", m.cells.size() / per);
    for (size_t q = 0; q < m.cells.size(); q += per) {
      for (size_t u = 0; u < per; u++) std::fprintf(f, u ? " %d" : "%d", m.cells[q + u]);
      std::fprintf(f, "\n");
    }
    if (m.dim == 3) {
      std::fprintf(f, "surface %zu\n", m.surface.size() / 3);
      for (size_t q = 0; q < m.surface.size(); q += 3)
        std::fprintf(f, "%d %d %d\n", m.surface[q], m.surface[q + 1], m.surface[q + 2]);
    }
    std::fclose(f);
    std::printf("mesh: %zu cells, %zu boundary triangles\n", m.cells.size() / per, m.surface.size() / 3);
    if (outdir) solver.current_solver.Output_mesh(outdir, 0, m);
  }
  return 0;
}
