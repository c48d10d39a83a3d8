/*
 * smg.h -- C ABI of the MI355X (gfx950) particle-relaxation engine.
 *
 * This is the drop-in boundary for the adaptive particle-relaxation hot path
 * of SPH-mesh-generation.  In the reference the path is the compile-time
 * SOLVER class driven by `Solver` (src/solver.h:28-45, src/solver.cpp:257-321);
 * there is no FFI.  Each entry point below names the reference member it
 * replaces.  Conventions:
 *   - plain C types, caller-owned host buffers, library-owned device state;
 *   - every call returns an int status (SMG_OK or a negative SMG_ERR_*); no
 *     exceptions cross the ABI, smg_last_error() gives the message;
 *   - calls on one handle are single-threaded; each handle owns a HIP stream;
 *   - per-particle host arrays are indexed by particle id (0..n-1), the order
 *     the particles were created in, whatever order the device keeps them in.
 */
#ifndef SMG_H
#define SMG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SMG_ABI_VERSION 5

enum {
  SMG_OK = 0,
  SMG_ERR_ARG = -1,       /* bad argument / config                          */
  SMG_ERR_HIP = -2,       /* HIP runtime error (no GPU, launch failure, ...) */
  SMG_ERR_CAPACITY = -3,  /* neighbour capacity or table size exceeded       */
  SMG_ERR_STATE = -4,     /* call out of order (e.g. step before levelset)  */
  SMG_ERR_REMAP = -5,     /* re-seeding found no positive-phase position    */
  SMG_ERR_NONFINITE = -6  /* a NaN/Inf appeared in the particle state        */
};

/* particle types, src/glbparam.h:55-62 */
enum { SMG_REAL = 0, SMG_SURFACE = 1, SMG_SINGULARITY = 2, SMG_SEGMENT = 3 };

/* target-spacing profiles: tanh is Levelset::Get_scale
 * (src/level_set_solver/level_set.cpp:311-351); linear is the build's
 * extension dx = dx0 + min(max(phi,0)/band,1)*(dx1-dx0). */
enum { SMG_PROFILE_TANH = 0, SMG_PROFILE_LINEAR = 1 };

typedef struct smg_config {
  int32_t dim;               /* 2 or 3                                        */
  int32_t sym[3];            /* mirror-ghost walls per axis (SYM_DIM_X/Y/Z)   */
  double box_l[3];           /* SPH and level-set box corner (SPH::box_l)     */
  double domain[3];          /* box extent (SPH::domain)                      */
  double particle_box_l[3];  /* seeding / re-seeding box (particle_box_l/r)   */
  double particle_box_r[3];
  int32_t num_pkg[3];        /* level-set packages per axis, 4 cells each     */
  double cell_ratio;         /* CELL_RATIO, h = dx*cell_ratio/cut_off         */
  double cut_off;            /* CUT_OFF                                       */
  double angle_thres;        /* SPH_mesh_generation::angle_thres (0.4*PI)     */
  double nu_coeff;           /* nu_coeff_start == nu_coeff_end (0.1)          */
  int32_t profile;           /* SMG_PROFILE_*                                 */
  double min_dl, max_dl;     /* tanh profile: minimum_dl / maximum_dl         */
  double max_phi;            /* tanh profile: Levelset::maximum_phi           */
  double lin_dx0, lin_dx1;   /* linear profile end spacings                   */
  double lin_band;           /* linear profile band width in phi              */
  uint64_t seed;             /* counter-based RNG key (replaces rand())       */
  int32_t remap_max_attempts;/* bound on the re-seeding rejection loop        */
  int32_t max_neighbors;     /* inline list entries per particle, 0 = 128,    
                                rounded up to a multiple of 4; longer lists  
                                spill to a shared pool                       */
  int32_t spill_per_slot;    /* spill pool entries per particle slot, 0 = 64  */
  int32_t reserved0;
  double skin;               /* neighbour lists cover (1 + skin) * cut_off * h
                                and are rebuilt (with the cell sort) only when
                                some particle's drift since the last build,
                                |dx| + cut_off |dh|, reaches skin*cut_off*h/2;
                                0 = 0.1, < 0 = rebuild every step             */
} smg_config;

/* Level-set statistics from smg_lset_build (Levelset::Calculate_total_volume_mass,
 * level_set.cpp:930-972, and Parallel_get_global_scale, parallelfunc_mesh.cpp:46-54). */
typedef struct smg_lset_stats {
  double maximum_phi;
  double total_volume;
  double total_area;
  double total_mass;
  double total_mass_surface;
  int64_t n_interface_cell;
  int64_t num_particle_mesh;     /* SPH_mesh_generation::Get_num_particles */
  int64_t num_particle_surface;
  int64_t num_particle;
} smg_lset_stats;

/* Per-step diagnostics (the values SPH_mesh_generation keeps as members). */
typedef struct smg_stats {
  int64_t step;          /* relaxation steps taken (run_time)                   */
  int64_t n_particle;
  int64_t n_real, n_surface;
  int64_t n_remapped;    /* re-seeded in the last mapping pass                  */
  int64_t n_neighbor_entries;
  int32_t max_neighbors_seen;
  int32_t num_level;     /* cell-list levels in use                             */
  double dt1, dt2;       /* glbl_timestep after Set_timestep(1) / (2)           */
  double max_v;          /* glbl_max_v after Set_timestep(2)                    */
  double residual;       /* mean_i |v_i| dt2 / dx_i of the last step (the      */
                         /* convergence measure; summed in 2^-40 fixed point) */
  double max_disp;       /* max_i |v_i| dt2 of the last step                    */
  double res_max;        /* max_i |v_i| dt2 / dx_i of the last step             */
  int64_t n_rebuild;     /* neighbour-list (and cell-sort) builds so far        */
  double drift;          /* max_i (|x_i - x_b| + cut_off |h_i - h_b|) / h_b of   */
                         /* the last step (x_b, h_b: at the last list build)     */
  double residual_sum;   /* residual summed over the steps so far: the mapping  */
                         /* pass of step s adds step s-1's, so after s steps it  */
                         /* covers steps 1..s-1; differences of it give windowed */
                         /* means (the server's convergence test)                */
} smg_stats;

typedef struct smg_job smg_job;

int smg_abi_version(void);
const char* smg_last_error(void);

/* Reference constants for `dim` (glbparam.h:171-185, sph_mesh_generation.cpp:8-152,680-743). */
int smg_default_config(int dim, smg_config* cfg);

/* Host-only level-set finishing: normals (Levelset_package::Get_normal / Get_BC_normal,
 * lset_package.cpp:332-389), cut-cell tags (Get_interface_tag / Is_cut_cell,
 * :450-469, :664-761), maximum_phi and particle count
 * (SPH_mesh_generation::Get_num_particles, sph_mesh_generation.cpp:900-941).
 * phi_ext: cell-centred phi on the grid with a full margin package on each side,
 *          (4*num_pkg+8) cells per active axis, x fastest.
 * zero_margin: 1 = restart-mode margins (phi = n = 0 outside the box), 0 = analytic.
 * Outputs on the grid with a one-cell halo, (4*num_pkg+2) cells per active axis:
 * phi (double), normal (float, 3 per cell), tag (int8, 1 = CUT_CELL). */
int smg_lset_build(const smg_config* cfg, const double* phi_ext, int zero_margin,
                   double* phi, float* normal, int8_t* tag, smg_lset_stats* stats);
/* smg_lset_build on the GPU (2D and 3D): the same arguments and outputs, computed by
 * the same operation sequences, so phi, normal, tag, maximum_phi, volume, area and, for
 * the uniform and linear spacing profiles, masses and particle counts are bit-identical
 * to the host builder's; with the tanh profile the per-cell spacing uses the device tanh
 * (within 2 ulp).  Replaces the same reference routines as smg_lset_build.  kernel_ms
 * (may be NULL): device time of the build, transfers excluded.  No GPU: SMG_ERR_HIP.
 * The 3D cut-cell pass queues split subdivision nodes in a buffer sized at 96 entries
 * per cut cell and regrown when short; the environment variable SMG_LSET_QUEUE_PER_CELL
 * (an integer in [1, 2^20]) overrides the 96 for tests of the regrow, and any other
 * value gives SMG_ERR_ARG. */
int smg_lset_build_gpu(const smg_config* cfg, const double* phi_ext, int zero_margin,
                       double* phi, float* normal, int8_t* tag, smg_lset_stats* stats, double* kernel_ms);

/* The reference's level-set restart file (./rstfile/levelset.rst), the input of its
 * relaxation run: case header (SPH_mesh_generation::Load_rst_case,
 * sph_mesh_generation.cpp:323-373), level-set header (Levelset::Load_lset_restart_file,
 * level_set.cpp:1098-1142), level info and per-cell fields over the interior cells
 * (Levelset_levelinfo::Load_level_set_rstfile, lset_level_infor.cpp:932-1146). */
typedef struct smg_rst_info {
  int32_t resolution[3], Lmin, Lmax, ini_num_cell[3], num_level;
  int32_t num_singularity, num_segment, glbl_num_interface_cell, reserved;
  double convergence_error;
  double fac_maximum_dl, fac_middel_dl, fac_minimum_dl, maximum_curv_artificial;
  double total_mass, total_mass_surface, total_mass_segment, total_volume, total_area, total_length;
  double maximum_phi, maximum_psi, maximum_dl, middel_dl, minimum_dl, maximum_curv, minimum_curv;
  double scale, dl;
  int64_t n_cells;       /* interior cells (4 num_pkg per active axis)                  */
  int64_t n_halo_cells;  /* cells of the one-cell-halo grid (phi, normal, tag outputs)   */
  int64_t n_ext_cells;   /* cells of the full-margin grid (phi_ext output)               */
} smg_rst_info;

/* Reads a restart file into a config (box, packages, particle box, the file's tanh
 * spacing profile: minimum_dl, maximum_dl, maximum_phi) and the engine's level-set
 * arrays on the one-cell-halo grid of smg_lset_build: phi, normal (3 per cell) and
 * tag (1 = CUT_CELL) from the file, halo cells zero as after the reference's restart
 * (the margin packages are not in the file).  phi_ext (may be NULL) gets phi on the
 * full-margin grid with zero margins, the input smg_lset_build(zero_margin = 1) needs
 * to recompute the particle count as the reference's restart mode does.  With every
 * array NULL only cfg and info (with the array sizes) are filled.  Files with
 * singularity or segment cells are rejected (SMG_ERR_ARG). */
int smg_lset_read_rst(const char* path, smg_config* cfg, smg_rst_info* info, double* phi, float* normal,
                      int8_t* tag, double* phi_ext);
/* An SDFGen signed-distance file (text: ni nj nk / origin / dx / ni*nj*nk values, i
 * fastest), as Levelset_levelinfo::Load_SDF_file (lset_level_infor.cpp:1168-1240)
 * reads it: 3D only, dx must equal the level-set cell size (within 1e-15), values of
 * cells beyond the configured grid are skipped and cells the file does not cover stay
 * zero, as do the margin packages.  Output: phi on the full-margin grid of
 * smg_lset_build ((4 num_pkg + 8)^3 cells, x fastest), ready for smg_lset_build. */
int smg_lset_read_sdf(const char* path, const smg_config* cfg, double* phi_ext);
/* Writes a level set (one-cell-halo arrays of smg_lset_build and its statistics) in the
 * same layout, as Levelset::Save_lset_restart_file (level_set.cpp:1022-1094) and
 * Levelset_levelinfo::Write_level_set_rstfile (lset_level_infor.cpp:695-927) lay it
 * out for one rank: curv and psi 0, no characteristic cells, scale = the spacing
 * profile's dx(phi). */
int smg_lset_write_rst(const char* path, const smg_config* cfg, const double* phi, const float* normal,
                       const int8_t* tag, const smg_lset_stats* stats);

/* SPH_mesh_generation constructor + Initialize_case (sph_mesh_generation.cpp:8, :594). */
int smg_create(const smg_config* cfg, smg_job** job);
int smg_destroy(smg_job* job);

/* Level set on the one-cell-halo grid of smg_lset_build (copied to the device);
 * replaces Load_rst_case / Levelset::Load_lset_restart_file (:323-422). */
int smg_set_levelset(smg_job* job, const double* phi, const float* normal, const int8_t* tag);

/* Multi-resolution extension; no counterpart in the reference, whose level set is
 * single-level (level_set.cpp:53; Lmin = Lmax = 0, sph_mesh_generation.cpp:308-310).
 * Refined levels over the base level set of smg_set_levelset: level l = 1..nlev has
 * num_pkg << l packages per axis.  Its cells are given on its own one-cell-halo grid,
 * as smg_lset_build produces them for a config with num_pkg << l, and active (one byte
 * per package of the level, x fastest, nonzero = kept) selects the packages stored on
 * the device, typically a band near the surface.  A sample (Calculate_particle_infor)
 * uses the finest level that stores all 2^D corner cells, else the next coarser, down
 * to the base grid; the cut-cell tag the finest level that stores the cell.  With
 * every package active the result equals a single-level run on the finest grid.  The
 * spacing profile's level-set cell and maximum_phi come from the finest level.  At
 * most 3 levels; call before seeding.  A server pool takes such jobs through
 * smg_server_admit_levels. */
typedef struct smg_lset_level {
  const double* phi;
  const float* normal;
  const int8_t* tag;
  const uint8_t* active;
} smg_lset_level;
int smg_set_levelset_levels(smg_job* job, int32_t nlev, const smg_lset_level* levels);

/* The job's level-set phi at n points (n*3 doubles) as the relaxation samples it
 * (Calculate_particle_infor's interpolation, particle_mesh_generation.cpp:175-265, on
 * the finest stored level); the job must be seeded or initialised. */
int smg_sample_phi(smg_job* job, int64_t n, const double* pts, double* phi);
/* No counterpart in the reference: smg_sample_phi for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same arguments and results, bit for bit; the points and results
 * are staged in a fixed device array of 8 MiB that the library holds on each GPU until
 * the process ends, so n is at most 262144 (SMG_ERR_CAPACITY beyond).  Every call on a
 * GPU shares that GPU's array: calls that overlap on one GPU (from several threads or
 * jobs) race and must be serialised by the caller. */
int smg_sample_phi_static(smg_job* job, int64_t n, const double* pts, double* phi);

/* Particles at given positions (n*3 doubles, unused axes ignored); type may be NULL
 * (all REAL).  h is set from the level set (Set_particle_info). */
int smg_init_particles(smg_job* job, int64_t n, const double* x, const int32_t* type);
/* Random_particle_distribution (sph_mesh_generation.cpp:1302-1367) with the
 * counter-based RNG keyed by cfg->seed. */
int smg_seed_particles(smg_job* job, int64_t n);
/* Pre_run (sph_mesh_generation.cpp:1010-1052): map/project once, zero v and nu. */
int smg_prerun(smg_job* job);
/* nsteps x SPH_run_mesh_generation (sph_mesh_generation.cpp:1371-1463). */
int smg_step(smg_job* job, int64_t nsteps);
/* Solver::Get_timestep (solver.cpp:269). */
int smg_get_timestep(smg_job* job, double* dt);
int smg_get_stats(smg_job* job, smg_stats* st);

/* State by particle id; any pointer may be NULL.  x, v: n*3 doubles; h: dx*cell_ratio/cut_off. */
int smg_get_state(smg_job* job, double* x, double* v, int32_t* type, double* h, float* nu);
/* Last step's per-particle pair results: sum (kernel summation), a1 (stage-1
 * acceleration), a2 (stage-2 acceleration), n*3 floats for a1/a2. */
int smg_get_pair_terms(smg_job* job, float* sum, float* a1, float* a2);
/* Neighbour sets of the last step (Level_info::Refresh_neighbor_info,
 * level_infor.cpp:469-506): for particle id i, entries offs[i]..offs[i+1]-1 hold
 * (src id, mirror code), sorted.  code = 0 for a real particle, else 2 bits per axis
 * (1 = image across the low wall, 2 = across the high wall) for mirror ghosts
 * (Level_info::Construct_symmetric_BC_particles, level_infor.cpp:786-856).
 * With ids == NULL only offs is filled (to size the buffers). */
int smg_get_neighbors(smg_job* job, int64_t* offs, int32_t* ids, int32_t* codes, int64_t capacity);
/* The level-set normal interpolated at every particle's current position (3 per
 * particle, by id; not renormalised): the particle's norm that the reference's writers
 * print (particle_binary_writer.cpp, "norm"). */
int smg_get_normals(smg_job* job, float* normal);

/* Kernel timing: average device time in ms of each kernel over the launches since
 * profiling was last enabled (enabling starts a new window), measured with HIP events
 * recorded on the job's stream around each launch. */
int smg_profile_enable(smg_job* job, int enable);
int smg_profile_get(smg_job* job, const char** names, double* avg_ms, int64_t* counts, int max_kernels, int* n_kernels);

/* Asynchronous stepping for request loops (no counterpart in the reference): enqueue
 * nsteps on the job's stream and return; smg_query sets *busy to 1 while the stream
 * still has work; smg_sync waits.  smg_capture records steps_per_graph (even) steps as
 * a hipGraph, after which smg_step_async replays it (nsteps must then be a multiple). */
int smg_step_async(smg_job* job, int64_t nsteps);
int smg_query(smg_job* job, int32_t* busy);
int smg_sync(smg_job* job);
int smg_capture(smg_job* job, int64_t steps_per_graph);
/* device bytes held by the job (for admission by memory footprint) */
int smg_footprint(smg_job* job, int64_t* bytes);

/* ---- autotuner (no counterpart in the reference) ----
 * Times candidate list shapes on a scratch copy of the job -- the cell-size multiple
 * (1 + skin; cells hold (1 + skin) cut_off h) and the inline list depth max_neighbors
 * -- each seeded, warmed for `warmup` steps and timed over `steps` graph-replayed
 * steps, and returns the fastest.  The caller puts best_skin / best_max_neighbors
 * into its smg_config before smg_create.  Results are bit-identical for every K;
 * skin changes only the list and sort schedule, not the neighbour sets. */
typedef struct smg_tune_result {
  double best_skin;
  int32_t best_max_neighbors;
  int32_t n_tried;
  double best_ms_per_step;
  double skin[16];
  int32_t max_neighbors[16];
  double ms_per_step[16];   /* per candidate, in the order tried */
} smg_tune_result;
int smg_autotune(const smg_config* cfg, const double* phi, const float* normal, const int8_t* tag, int64_t n_particles,
                 int32_t warmup, int32_t steps, smg_tune_result* result);

/* ---- batched relaxation (many independent jobs per kernel launch) ----
 * No counterpart in the reference, which relaxes one geometry per MPI job.  Jobs of
 * the same dimension share one slot space: every kernel launch covers all of them,
 * with per-job parameter and offset tables (tile -> job); the 2D and the 3D jobs run
 * on two HIP streams and overlap on the device.  Arrays are indexed by job;
 * phi/normal/tag are one-cell-halo level sets as for smg_set_levelset. */
typedef struct smg_batch smg_batch;
int smg_batch_create(int32_t njobs, const smg_config* cfgs, const int64_t* n_particles,
                     const double* const* phi, const float* const* normal, const int8_t* const* tag,
                     smg_batch** batch);
int smg_batch_destroy(smg_batch* batch);
/* seed every job (Random_particle_distribution) and Pre_run */
int smg_batch_seed(smg_batch* batch);
/* nsteps steps of every job, one launch per pass */
int smg_batch_step(smg_batch* batch, int64_t nsteps);
/* capture steps_per_graph (even) steps of every job as one hipGraph per dimension;
 * smg_batch_run replays it `launches` times */
int smg_batch_capture(smg_batch* batch, int64_t steps_per_graph);
int smg_batch_run(smg_batch* batch, int64_t launches);
int smg_batch_sync(smg_batch* batch);
int smg_batch_get_stats(smg_batch* batch, int32_t job, smg_stats* st);
int smg_batch_get_state(smg_batch* batch, int32_t job, double* x, double* v, int32_t* type, double* h, float* nu);

/* ---- 2D triangulation (Graphmeshcls::Reconstruct_tris, graph_mesh.cpp:37-407) ----
 * Edges: the Voronoi faces of each particle's cell among its neighbours, inside the box
 * x +- 1.5 cut_off h (Voronoi_construction::Local_VD_generation,
 * voronoi_construction.cpp:6-127), on the GPU in FP64; a mirror image stands for its
 * source particle.  Then the reference's integer stage: all 3-cliques (i > j > k);
 * a triangle of three non-REAL particles whose centroid phi is <= 1e-4 of their mean
 * phi loses its longest edge; the graph is rebuilt from the surviving triangles and
 * its 3-cliques are the mesh.  Quality statistics follow Tri_quality_statistics
 * (graph_mesh.cpp:411-671). */
typedef struct smg_tri_stats {
  int64_t n_tri, n_edge;           /* final triangles, edges of the Voronoi stage        */
  int64_t n_deleted_edge;          /* edges removed by the centroid test                */
  int64_t n_cell_overflow;         /* cells that needed more vertices / faces than kept */
  double min_angle, max_angle;     /* degrees                                            */
  double avg_min_angle;
  double min_edge, max_edge;
  double min_area, max_area, total_area;
  double max_aspect;               /* longest / shortest edge                            */
  double G_min, G_avg;             /* 2 sqrt(3) area / (longest edge * half perimeter)   */
  int64_t n_min_angle_below[4];    /* triangles with min angle < 10, 20, 30, 40 degrees  */
  int32_t angle_hist[45];          /* angles in 4-degree bins                            */
  int32_t reserved;
} smg_tri_stats;

/* Triangulates a 2D job's current particles, after smg_prerun or any step (between
 * smg_prerun and the first step the candidate lists are built on the placed positions
 * and nothing moves: the post mode).  tris: 3 particle ids per triangle (i > j > k),
 * may be NULL to query the count; capacity in triangles. */
int smg_triangulate(smg_job* job, int32_t* tris, int64_t capacity, int64_t* n_tri, smg_tri_stats* stats);
/* The Voronoi-stage edges (particle ids, i > j, sorted) of the last smg_triangulate. */
int smg_tri_edges(smg_job* job, int32_t* edges, int64_t capacity, int64_t* n_edges);
/* The integer stage alone (host only): all 3-cliques (i > j > k) of an undirected
 * graph given as n_edges pairs (self loops and repeats ignored), sorted. */
int smg_tri_cliques(int64_t n_vertices, const int32_t* edges, int64_t n_edges, int32_t* tris, int64_t capacity,
                    int64_t* n_tri);

/* ---- 3D tetrahedralisation (Graphmeshcls::Reconstruct_tets, graph_mesh.cpp:981-1963) ----
 * Delaunay tetrahedra of the particle positions (the reference calls TetGen "JCnQ",
 * graph_mesh.cpp:1021) from each particle's Voronoi cell, clipped on the GPU in FP64.
 * A tetrahedron is produced when its circumsphere is certified empty by the cell
 * search (every particle within twice its circumradius of a vertex was clipped); the
 * convex-hull slivers and cavity-spanning tetrahedra of a global tetrahedralisation,
 * whose circumspheres are far larger than the spacing, are not produced (n_open counts
 * the particles whose cells stayed open).  Then the reference's deletion rule
 * (Check_whether_this_tet_should_be_delete, graph_mesh.cpp:675-979): a tetrahedron of
 * four non-REAL particles whose aspect ratio (longest edge / smallest height) exceeds
 * delete_thres goes (the phi tests there are computed but do not decide).  Boundary
 * triangles: the face of three non-REAL particles of a kept tetrahedron with exactly one
 * REAL particle (graph_mesh.cpp:1153-1176).  Quality statistics follow
 * Tet_quality_statistics (graph_mesh.cpp:1967-2460).  TetGen's flip / optimisation pass
 * ("JrO/1o/100Q", graph_mesh.cpp:1401) is TetGen-internal and not reproduced. */
typedef struct smg_tet_stats {
  int64_t n_tet;                   /* tetrahedra after the deletion rule                  */
  int64_t n_delaunay;              /* Delaunay tetrahedra of the GPU stage                 */
  int64_t n_deleted;               /* four non-REAL particles and aspect > delete_thres    */
  int64_t n_surface_tri;           /* boundary triangles                                   */
  int64_t n_open;                  /* particles whose cell neither pass closed             */
  int64_t n_cell_overflow;         /* particles whose cell could not be built             */
  int64_t n_duplicate;             /* particles on the point of a smaller id (in no tet)   */
  double min_edge, max_edge;
  double min_volume, max_volume, total_volume;
  double min_aspect, max_aspect, avg_aspect;                     /* longest edge / min height */
  double min_radius_ratio, max_radius_ratio, avg_radius_ratio;   /* 3 r_in / r_circ           */
  double min_dihedral, max_dihedral, avg_min_dihedral;           /* degrees                   */
  int64_t n_min_dihedral_below[4]; /* tetrahedra with min dihedral < 10, 20, 30, 40 degrees */
  int64_t dihedral_hist[45];       /* dihedral angles in 4-degree bins                     */
  int64_t aspect_hist[12];         /* bins bounded by 1.5 2 2.5 3 4 6 10 15 25 50 100      */
  int64_t radius_hist[56];         /* radius ratio in 0.02 bins, the last one open         */
  double cell_ms;                  /* GPU time of the Voronoi-cell passes                  */
  double total_ms;                 /* the whole call, downloads and host stage included    */
  int64_t n_second_pass;           /* cells the ring search left open (hull particles,     */
                                   /* cavity spanners), built by the verifying pass        */
} smg_tet_stats;

/* Tetrahedralises a 3D job's current particles.  tets: 4 particle ids per tetrahedron
 * (ascending), sorted; may be NULL to query the count; capacity in tetrahedra.
 * delete_thres <= 0 takes the reference's 10 (sph_mesh_generation.cpp:705). */
int smg_tetrahedralize(smg_job* job, double delete_thres, int32_t* tets, int64_t capacity, int64_t* n_tet,
                       smg_tet_stats* stats);
/* Boundary triangles (3 ascending particle ids, sorted) of the last smg_tetrahedralize. */
int smg_tet_surface(smg_job* job, int32_t* tris, int64_t capacity, int64_t* n_tri);
/* The GPU stage of the last smg_tetrahedralize: the Delaunay tetrahedra of all
 * particles (before deletion, ascending ids, sorted) and, by particle id, the cell
 * status (0 closed; 2 capacity, 3-5 degenerate cut: not built; 6 on the point of a
 * smaller id) and the certified radius (every point within it was clipped; 1e300 for
 * a cell closed by the verifying pass).  status and rcert may be NULL; they hold
 * n_particles entries. */
int smg_tet_delaunay(smg_job* job, int32_t* tets, int64_t capacity, int64_t* n_tet, int8_t* status, double* rcert);
/* The host stage alone (no GPU): deletion rule, boundary triangles and statistics of
 * given tetrahedra over positions x (3 per vertex) and particle types. */
int smg_tet_finish(int64_t n_vertices, const double* x, const int32_t* type, const int32_t* tets, int64_t n_tets,
                   double delete_thres, int32_t* kept, int64_t kept_capacity, int64_t* n_kept, int32_t* tris,
                   int64_t tri_capacity, int64_t* n_tri, smg_tet_stats* stats);

/* ---- serving: continuous batching over a device-resident job pool ----
 * No counterpart in the reference.  A server owns one pool of particle slots, cell
 * tables and level-set cells on the device (its HBM footprint, fixed at creation).
 * Jobs are admitted into free ranges of the pool between chunks of steps; every step
 * launch covers all admitted jobs; after each chunk the jobs whose mean residual over
 * the chunk fell below tol (or that reached max_steps) are retired, their results
 * copied to the host and their ranges freed for the next admission.  Admission and
 * retirement happen only at chunk boundaries, so a job's step count at every check is
 * a multiple of chunk_steps whatever else is in the pool, and its result is the same
 * bits alone or in any mix.  All calls on one server come from one host thread. */
typedef struct smg_server_config {
  int32_t dim;
  int32_t max_jobs;          /* jobs in the pool at once                              */
  int64_t max_slots;         /* particle slots (a job takes n rounded up to 256)      */
  int64_t max_table;         /* cell-table entries (a job takes its cell-list tables) */
  int64_t max_lset_cells;    /* level-set cells (a job takes (4 num_pkg + 2)^dim)     */
  int32_t max_neighbors;     /* inline list entries per slot, 0 = 128 (multiple of 4) */
  int32_t spill_per_slot;    /* spill pool entries per slot, 0 = 64                   */
  int32_t chunk_steps;       /* steps between admissions / convergence checks (> 0)  */
  int32_t min_steps;         /* no retirement by convergence before this many steps  */
  int32_t max_steps;         /* retire unconverged at this many steps (0 = no limit) */
  int32_t reserved;
  double tol;                /* converged: mean over the chunk of smg_stats.residual < tol */
  int64_t max_active_particles; /* admit only while the pool's particles stay within
                                   this (0 = no limit; an empty pool admits any job):
                                   bounds the work per chunk, hence latency          */
  int64_t max_lset_maps;     /* package-map entries of refined level-set levels (a job
                                with refined levels takes (num_pkg << l + 2)^dim per
                                level l); 0 = jobs with refined levels are refused */
} smg_server_config;

/* finished-job status */
enum { SMG_JOB_CONVERGED = 0, SMG_JOB_MAX_STEPS = 1 };   /* < 0: an SMG_ERR_* code */

typedef struct smg_server_job {
  int32_t ticket;            /* id returned by smg_server_admit                       */
  int32_t status;            /* SMG_JOB_* or a negative SMG_ERR_* code                */
  int64_t steps;             /* steps taken                                           */
  int64_t n_particle;
  double window_residual;    /* mean residual over the last chunk                      */
} smg_server_job;

typedef struct smg_server smg_server;
/* device bytes a server of this configuration allocates */
int smg_server_footprint_of(const smg_server_config* cfg, int64_t* bytes);
/* pool ranges one job takes: slots, table entries, level-set cells */
int smg_server_job_size(const smg_config* cfg, int64_t n_particles, int64_t* slots, int64_t* table,
                        int64_t* lset_cells);
int smg_server_create(const smg_server_config* cfg, smg_server** server);
int smg_server_destroy(smg_server* server);
/* Seeds the job (Random_particle_distribution + Pre_run) into free ranges of the pool.
 * Returns SMG_ERR_CAPACITY, with the pool unchanged, when the job does not fit now. */
int smg_server_admit(smg_server* server, const smg_config* cfg, const double* phi, const float* normal,
                     const int8_t* tag, int64_t n_particles, int32_t* ticket);
/* smg_server_admit for a job with refined level-set levels (the arrays and rules of
 * smg_set_levelset_levels): the levels' stored cells take pool level-set cells and
 * their package maps take max_lset_maps entries.  nlev = 0 is smg_server_admit. */
int smg_server_admit_levels(smg_server* server, const smg_config* cfg, const double* phi, const float* normal,
                            const int8_t* tag, int32_t nlev, const smg_lset_level* levels, int64_t n_particles,
                            int32_t* ticket);
/* One chunk of steps for every job in the pool, then convergence checks and
 * retirement.  n_active: jobs still in the pool afterwards (may be NULL).  With an even
 * chunk_steps the chunk replays a captured hipGraph of the step loop (captured again
 * only when the pool's extent changes); the results are those of eager steps. */
int smg_server_advance(smg_server* server, int32_t* n_active);
/* Jobs finished since the last call (at most cap entries). */
int smg_server_finished(smg_server* server, smg_server_job* out, int32_t cap, int32_t* n);
/* Final state of a finished job by particle id (any pointer may be NULL), then
 * releases it.  x: n*3 doubles. */
int smg_server_result(smg_server* server, int32_t ticket, double* x, int32_t* type, double* h, smg_stats* st);
/* pool occupancy */
int smg_server_usage(smg_server* server, int64_t* used_slots, int64_t* total_slots, int32_t* n_active);

/* Diagnostic with no counterpart in the reference: device FP64 sqrt(a), a/b,
 * floor(a/b) and tanh(a) for n inputs (4 outputs per input), to check that the FP64
 * operations the bit-exact paths use round as the host's do. */
int smg_fp64_probe(int64_t n, const double* a, const double* b, double* out);

/* ---- Output files (host only), single rank --------------------------------------- */

/* Points of an output file: positions and the point data the reference writes. */
typedef struct smg_vtu_points {
  int64_t n;
  const double* x;        /* 3 per point (required when n > 0)        */
  const int32_t* type;    /* per point; NULL writes zeros              */
  const double* h;        /* NULL writes zeros                         */
  const int32_t* nn;      /* neighbour counts; NULL writes zeros       */
  const int32_t* color;   /* rank colour; NULL writes zeros            */
  const float* normal;    /* 3 per point; NULL writes zeros            */
} smg_vtu_points;

/* Particle_binary_writer::Output_particle (particle_binary_writer.cpp:36-401): in dir,
 * particle_rank_0_step_<n_out>.vtu (VTK XML, raw appended data; point data id, local_id,
 * type, h, NN, color, norm; one vertex cell per point), particle_step_<n_out>.pvtu, and
 * particle.pvd, created at n_out = 0 and extended by one DataSet (timestep = time) at
 * every later n_out. */
int smg_write_particles_vtu(const char* dir, int n_out, double time, const smg_vtu_points* points);
/* Mesh_binary_writer::Output_mesh (mesh_binary_writer.cpp:36-650): 2D triangles
 * (3 ids each) to mesh_tris_rank_0_step_<n_out>.vtu, or 3D tetrahedra (4 ids) to
 * mesh_tets_... with cell data radius_ratio and mindihedangle; point data type, h,
 * color; plus the .pvtu and the mesh_tris.pvd / mesh_tets.pvd collection. */
int smg_write_mesh_vtu(const char* dir, int n_out, double time, int dim, const smg_vtu_points* points,
                       const int32_t* cells, int64_t n_cells);
/* SPH_mesh_generation::Save_particle_post_file (sph_mesh_generation.cpp:2624-2667):
 * dir/step_<step>_rank_0.pst (int32 count, 3 doubles per particle) and a line
 * "step count" in dir/list.csv, which first != 0 starts with the number of ranks (1). */
int smg_write_post_file(const char* dir, int64_t step, int64_t n, const double* x, int first);
/* Load_data_for_post (sph_mesh_generation.cpp:426-590), its list: the number of ranks
 * and the saved steps with their particle counts (steps/counts may be NULL to query). */
int smg_read_post_list(const char* dir, int32_t* n_ranks, int64_t* steps, int64_t* counts, int64_t capacity,
                       int64_t* n_steps);
/* ... and one step's particles from all rank files, in rank order; color (may be NULL)
 * gets each particle's rank.  x may be NULL to query the count. */
int smg_read_post_file(const char* dir, int64_t step, int32_t n_ranks, double* x, int32_t* color, int64_t capacity,
                       int64_t* n);

#ifdef __cplusplus
}
#endif
#endif
