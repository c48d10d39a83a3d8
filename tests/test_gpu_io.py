"""Output files of a relaxed job: the particle VTU carries the job's positions, types,
h, neighbour counts and interpolated normals; the mesh VTU its triangles (2D) or
tetrahedra with their quality (3D)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_job_output_2d(smg_mod, c1, tmp_path):
    cfg, lset = c1
    job = smg_mod.Job(cfg, lset)
    job.seed(int(lset["stats"].num_particle))
    job.step(20)
    tris, _ = smg_mod.triangulate(job)
    smg_mod.write_job_output(job, tmp_path, 0, time=20.0, mesh=(2, tris))
    st = job.state()
    offs, _, _ = job.neighbors()
    nrm = job.normals()
    v = smg_mod.read_vtu(tmp_path / "particle_rank_0_step_0.vtu")
    assert v["_points"] == job.n
    assert np.array_equal(v["Position"], st["x"].astype(np.float32))
    assert np.array_equal(v["type"], st["type"]) and np.array_equal(v["h"], st["h"].astype(np.float32))
    assert np.array_equal(v["NN"], np.diff(offs)) and np.array_equal(v["norm"], nrm)
    # particles near the hole (radius 0.2): the interpolated normal points away from its
    # centre
    d = np.linalg.norm(st["x"][:, :2] - 0.5, axis=1)
    near = np.abs(d - 0.2) < 0.02
    assert near.sum() > 50
    r = (st["x"][near, :2] - 0.5) / d[near, None]
    assert np.all(np.sum(nrm[near, :2] * r, axis=1) > 0.9)
    m = smg_mod.read_vtu(tmp_path / "mesh_tris_rank_0_step_0.vtu")
    assert m["_cells"] == len(tris) and np.array_equal(m["connectivity"].reshape(-1, 3), tris)
    job.close()


def test_job_mesh_output_3d(smg_mod, tmp_path):
    cfg, sdf = smg_mod.config_c2(scale=2.0, num_pkg=16)
    lset = smg_mod.make_levelset(cfg, sdf)
    job = smg_mod.Job(cfg, lset)
    job.seed(int(lset["stats"].num_particle))
    job.step(20)
    tets, st = smg_mod.tetrahedralize(job)
    smg_mod.write_job_output(job, tmp_path, 0, mesh=(3, tets))
    m = smg_mod.read_vtu(tmp_path / "mesh_tets_rank_0_step_0.vtu")
    assert m["_cells"] == len(tets) == st["n_tet"]
    assert np.all((m["radius_ratio"] > 0) & (m["radius_ratio"] <= 1.0 + 1e-6))
    assert np.all((m["mindihedangle"] > 0) & (m["mindihedangle"] < 180))
    assert float(m["mindihedangle"].min()) == pytest.approx(st["min_dihedral"], rel=1e-5)
    job.close()


def test_post_processing_mode(smg_mod, c1, tmp_path):
    """Load_data_for_post: particles saved with the post files are placed on the level
    set again and meshed; with no particle in a cut cell the placement moves nothing, so
    the triangles equal the relaxed job's own."""
    cfg, lset = c1
    job = smg_mod.Job(cfg, lset)
    job.seed(int(lset["stats"].num_particle))
    job.step(20)
    st = job.state()
    tris, _ = smg_mod.triangulate(job)
    smg_mod.write_post_file(tmp_path, 20, st["x"][:, :2], first=True)
    job.close()
    out = tmp_path / "outdata"
    out.mkdir()
    res = smg_mod.post_process(tmp_path, cfg, lset, outdir=out)
    assert [r[0] for r in res] == [20]
    assert np.count_nonzero(st["type"] == 1) == 0
    assert np.array_equal(res[0][1], tris)
    m = smg_mod.read_vtu(out / "mesh_tris_rank_0_step_0.vtu")
    assert np.array_equal(m["connectivity"].reshape(-1, 3), tris)
    assert 'timestep="20"' in (out / "mesh_tris.pvd").read_text()


@pytest.mark.parametrize("case", ["2d", "3d"])
def test_post_processing_places_cut_cell_particles(smg_mod, orc, c1, case, tmp_path):
    """Load_data_for_post on positions that need placing: the 2D case's seeds plus 200
    points just outside the circle (in its cut cells), the 3D case's seeds as drawn.
    The post mode places them (Get_particle_info, Set_particle_scale,
    Find_and_map_particle) exactly as the oracle's Pre_run places the same positions:
    the mesh VTU's types equal, its positions and h equal the oracle's rounded to its
    Float32 arrays, and some particles were projected onto the surface."""
    if case == "2d":
        cfg, lset = c1
    else:
        cfg, sdf = smg_mod.config_c2(scale=3.0, num_pkg=8)
        lset = smg_mod.make_levelset(cfg, sdf)
    D = cfg.dim
    n = int(lset["stats"].num_particle)
    x0 = orc.seed(cfg, lset, n).x[:, :D].copy()
    if case == "2d":
        rng = np.random.default_rng(5)
        th = rng.uniform(0.0, 2.0 * np.pi, 200)
        r = 0.2 + rng.uniform(0.0005, 0.004, 200)
        x0 = np.concatenate([x0, np.stack([0.5 + r * np.cos(th), 0.5 + r * np.sin(th)], 1)])
    n = len(x0)
    smg_mod.write_post_file(tmp_path, 0, x0, first=True)
    out = tmp_path / "outdata"
    out.mkdir()
    res = smg_mod.post_process(tmp_path, cfg, lset, outdir=out)
    assert [r_[0] for r_ in res] == [0] and len(res[0][1]) > 0
    o = orc.State.from_arrays(x0, np.zeros_like(x0), np.zeros(n), np.zeros(n), np.zeros(n, np.int32), 0.0)
    orc.prerun(cfg, lset, o)
    moved = np.any(o.x[:, :D] != x0, axis=1)
    assert np.count_nonzero(o.type == 1) > 0 and np.count_nonzero(moved) > 0
    m = smg_mod.read_vtu(out / f"mesh_{'tris' if D == 2 else 'tets'}_rank_0_step_0.vtu")
    assert m["_points"] == n
    assert np.array_equal(m["type"], o.type)
    assert np.array_equal(m["Position"][:, :D], o.x[:, :D].astype(np.float32))
    assert np.array_equal(m["h"], o.h.astype(np.float32))
