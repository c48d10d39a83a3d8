"""The C++ host mirror (Initialization / Solver / SPH_mesh_generation over the C ABI,
sph-mesh-generation_amd/host): it builds with the library, fails loudly without a
GPU, and on the GPU gives the same particles and the same mesh (Post_processing's
mesh stage: triangles in 2D, tetrahedra and boundary triangles in 3D) as the Python
binding of the same case."""
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "sph-mesh-generation_amd", "lib", "smg_run")


def _driver():
    if not os.path.exists(DRIVER):
        sys.path.insert(0, os.path.join(ROOT, "sph-mesh-generation_amd"))
        import build_lib
        build_lib.build_driver()
    return DRIVER


def test_driver_fails_loudly_without_gpu():
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1", ROCR_VISIBLE_DEVICES="-1")
    out = subprocess.run([_driver(), "--outputs", "1", "--output-dt", "2"], env=env, capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 1
    assert "ERROR in smg_lset_build_gpu (-2)" in out.stderr


@pytest.mark.gpu
def _read_mesh(path, per):
    lines = open(path).read().split("\n")
    assert lines[0].startswith("cells ")
    nc = int(lines[0].split()[1])
    cells = np.array([[int(v) for v in ln.split()] for ln in lines[1:1 + nc]], np.int32).reshape(-1, per)
    surf = None
    if per == 4:
        assert lines[1 + nc].startswith("surface ")
        ns = int(lines[1 + nc].split()[1])
        surf = np.array([[int(v) for v in ln.split()] for ln in lines[2 + nc:2 + nc + ns]], np.int32).reshape(-1, 3)
    return cells, surf


@pytest.mark.gpu
def test_driver_equals_python_binding(smg_mod, tmp_path):
    csv = tmp_path / "x.csv"
    mesh = tmp_path / "mesh.txt"
    outdir = tmp_path / "outdata"
    outdir.mkdir()
    out = subprocess.run([_driver(), "--outputs", "2", "--output-dt", "10", "--csv", str(csv), "--mesh", str(mesh),
                          "--outdir", str(outdir)], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    assert sum(line.startswith("output ") for line in out.stdout.splitlines()) == 2
    data = np.loadtxt(csv, delimiter=",", skiprows=1)
    cfg, sdf = smg_mod.config_c1()
    lset = smg_mod.make_levelset(cfg, sdf)
    job = smg_mod.Job(cfg, lset)
    job.seed(int(lset["stats"].num_particle))
    job.step(20)
    g = job.state()
    assert data.shape[0] == len(g["h"])
    assert np.array_equal(data[:, 1:3], g["x"][:, :2])
    assert np.array_equal(data[:, 4].astype(np.int32), g["type"])
    tris, _ = smg_mod.triangulate(job)
    cells, _ = _read_mesh(mesh, 3)
    assert len(cells) > 0
    assert np.array_equal(cells, tris)
    # the reference's output files of the two rounds and of the final mesh
    n = len(g["h"])
    assert (outdir / "list.csv").read_text() == f"1\n10 {n}\n20 {n}\n"
    ranks, steps = smg_mod.read_post_list(outdir)
    x, _ = smg_mod.read_post_file(outdir, steps[-1][0], ranks)
    assert np.array_equal(x[:, :2], g["x"][:, :2])
    assert (outdir / "particle.pvd").read_text().count("<DataSet") == 2
    v = smg_mod.read_vtu(outdir / "particle_rank_0_step_1.vtu")
    assert np.array_equal(v["Position"], g["x"].astype(np.float32)) and np.array_equal(v["type"], g["type"])
    m = smg_mod.read_vtu(outdir / "mesh_tris_rank_0_step_0.vtu")
    assert np.array_equal(m["connectivity"].reshape(-1, 3), tris)


@pytest.mark.gpu
def test_driver_mesh_3d_equals_python_binding(smg_mod, tmp_path):
    """3D: the unit cube minus a sphere r = 0.2, uniform dx = 0.04 on 32^3 level-set
    cells, 20 steps, then the tetrahedra (reference deletion rule, threshold 10) and
    their boundary triangles."""
    mesh = tmp_path / "mesh.txt"
    out = subprocess.run([_driver(), "--dim", "3", "--num-pkg", "8", "--dx", "0.04", "--outputs", "1",
                          "--output-dt", "20", "--mesh", str(mesh)], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    cfg = smg_mod.default_config(3)
    for a in range(3):
        cfg.num_pkg[a] = 8
    cfg.min_dl = cfg.max_dl = 0.04
    lset = smg_mod.make_levelset(cfg, smg_mod.sdf_circle_hole((0.5, 0.5, 0.5), 0.2))
    job = smg_mod.Job(cfg, lset)
    job.seed(int(lset["stats"].num_particle))
    job.step(20)
    tets, st = smg_mod.tetrahedralize(job)
    surf = smg_mod.tet_surface(job)
    cells, csurf = _read_mesh(mesh, 4)
    assert st["n_tet"] > 4 * job.n
    assert np.array_equal(cells, tets)
    assert np.array_equal(csurf, surf)
