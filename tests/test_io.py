"""The reference's output files (host only): VTU / PVTU / PVD for particles and meshes
(particle_binary_writer.cpp, mesh_binary_writer.cpp) and the .pst post files with
list.csv (sph_mesh_generation.cpp:2624-2667, read back as Load_data_for_post does)."""
import math
import os
import struct

import numpy as np
import pytest

PI_REF = 3.1415926   # glbparam.h:185, the reference's degrees


def _points(n, seed=0):
    rng = np.random.default_rng(seed)
    return (rng.random((n, 3)), rng.integers(0, 2, n).astype(np.int32), rng.random(n) * 0.01,
            rng.integers(5, 30, n).astype(np.int32), rng.standard_normal((n, 3)).astype(np.float32))


def test_particles_vtu_arrays(smg_mod, tmp_path):
    x, ty, h, nn, nrm = _points(257)
    smg_mod.write_particles_vtu(tmp_path, 0, 100.0, x, ty, h, nn, None, nrm)
    v = smg_mod.read_vtu(tmp_path / "particle_rank_0_step_0.vtu")
    assert v["_points"] == v["_cells"] == 257
    assert np.array_equal(v["id"], np.arange(257)) and np.array_equal(v["local_id"], np.arange(257))
    assert np.array_equal(v["type"], ty)
    assert np.array_equal(v["h"], h.astype(np.float32))
    assert np.array_equal(v["NN"], nn)
    assert np.array_equal(v["color"], np.zeros(257, np.int32))
    assert np.array_equal(v["norm"], nrm)
    assert np.array_equal(v["Position"], x.astype(np.float32))
    # one vertex cell per particle (VTK_VERTEX = 1)
    assert np.array_equal(v["connectivity"], np.arange(257))
    assert np.array_equal(v["offsets"], np.arange(1, 258))
    assert np.all(v["types"] == 1)
    text = (tmp_path / "particle_step_0.pvtu").read_text()
    assert 'Source="./particle_rank_0_step_0.vtu"' in text and 'Name="norm" NumberOfComponents="3"' in text


def test_pvd_collects_every_output(smg_mod, tmp_path):
    x, ty, h, nn, nrm = _points(10)
    for k, t in enumerate((0.0, 100.0, 200.0)):
        smg_mod.write_particles_vtu(tmp_path, k, t, x, ty, h, nn, None, nrm)
    text = (tmp_path / "particle.pvd").read_text().splitlines()
    assert text[0] == '<VTKFile type="Collection" version="0.1">' and text[-1] == "</VTKFile>"
    ds = [ln for ln in text if "<DataSet" in ln]
    assert [d.split('timestep="')[1].split('"')[0] for d in ds] == ["0", "100", "200"]
    assert [d.split('file="')[1].split('"')[0] for d in ds] == [f"./particle_step_{k}.pvtu" for k in range(3)]
    # n_out > 0 without the collection: a state error, nothing written
    with pytest.raises(smg_mod.SmgError):
        smg_mod.write_particles_vtu(tmp_path / "nothere", 1, 1.0, x)


def test_mesh_vtu_triangles(smg_mod, tmp_path):
    x = np.array([[0, 0], [1, 0], [0, 1], [1, 1]], float)
    tris = np.array([[2, 1, 0], [3, 2, 1]], np.int32)
    smg_mod.write_mesh_vtu(tmp_path, 0, 5.0, 2, x, tris, type=np.array([0, 1, 0, 1], np.int32))
    v = smg_mod.read_vtu(tmp_path / "mesh_tris_rank_0_step_0.vtu")
    assert v["_points"] == 4 and v["_cells"] == 2
    assert np.array_equal(v["connectivity"], tris.ravel())
    assert np.array_equal(v["offsets"], [3, 6])
    assert np.all(v["types"] == 5)   # VTK_TRIANGLE
    assert np.array_equal(v["Position"][:, 2], np.zeros(4, np.float32))
    assert np.array_equal(v["type"], [0, 1, 0, 1])
    assert os.path.exists(tmp_path / "mesh_tris.pvd") and os.path.exists(tmp_path / "mesh_tris_step_0.pvtu")


def test_mesh_vtu_tetrahedra_quality(smg_mod, tmp_path):
    """Regular tetrahedron: radius ratio 1, every dihedral acos(1/3); a right-corner
    tetrahedron: radius ratio 3 r_in / r_circ in closed form, smallest dihedral
    acos(1/sqrt 3) (degrees with the reference's PI)."""
    reg = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], float)
    cor = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float) + 5.0
    x = np.vstack([reg, cor])
    tets = np.array([[0, 1, 2, 3], [4, 5, 6, 7]], np.int32)
    smg_mod.write_mesh_vtu(tmp_path, 0, 0.0, 3, x, tets)
    v = smg_mod.read_vtu(tmp_path / "mesh_tets_rank_0_step_0.vtu")
    assert np.all(v["types"] == 10) and np.array_equal(v["offsets"], [4, 8])
    r_in = 1.0 / (3.0 + math.sqrt(3.0))           # volume 1/6, area (3 + sqrt 3) / 2
    r_circ = math.sqrt(3.0) / 2.0
    assert v["radius_ratio"][0] == pytest.approx(1.0, abs=1e-6)
    assert v["radius_ratio"][1] == pytest.approx(3 * r_in / r_circ, rel=1e-6)
    assert v["mindihedangle"][0] == pytest.approx(math.acos(1 / 3) / PI_REF * 180, rel=1e-6)
    assert v["mindihedangle"][1] == pytest.approx(math.acos(1 / math.sqrt(3)) / PI_REF * 180, rel=1e-6)


def test_mesh_rejects_bad_vertex(smg_mod, tmp_path):
    x = np.zeros((3, 3))
    with pytest.raises(smg_mod.SmgError):
        smg_mod.write_mesh_vtu(tmp_path, 0, 0.0, 2, x, np.array([[0, 1, 3]], np.int32))


def test_post_files_round_trip(smg_mod, tmp_path):
    rng = np.random.default_rng(3)
    xa, xb = rng.random((50, 3)), rng.random((70, 2))
    smg_mod.write_post_file(tmp_path, 0, xa, first=True)
    smg_mod.write_post_file(tmp_path, 1000, xb, first=False)
    assert (tmp_path / "list.csv").read_text() == "1\n0 50\n1000 70\n"
    raw = (tmp_path / "step_1000_rank_0.pst").read_bytes()
    assert struct.unpack("<i", raw[:4])[0] == 70 and len(raw) == 4 + 70 * 24
    ranks, steps = smg_mod.read_post_list(tmp_path)
    assert ranks == 1 and steps == [(0, 50), (1000, 70)]
    x, color = smg_mod.read_post_file(tmp_path, 1000, 1)
    assert np.array_equal(x[:, :2], xb) and np.all(x[:, 2] == 0.0) and np.all(color == 0)


def test_post_file_of_several_ranks(smg_mod, tmp_path):
    """Load_data_for_post reads every rank's file of a step in rank order and colours
    the particles by rank."""
    parts = [np.arange(6, dtype=float).reshape(2, 3), np.arange(9, dtype=float).reshape(3, 3) + 100]
    (tmp_path / "list.csv").write_text("2\n7 5\n")
    for r, p in enumerate(parts):
        (tmp_path / f"step_7_rank_{r}.pst").write_bytes(struct.pack("<i", len(p)) + p.astype("<f8").tobytes())
    ranks, steps = smg_mod.read_post_list(tmp_path)
    assert ranks == 2 and steps == [(7, 5)]
    x, color = smg_mod.read_post_file(tmp_path, 7, ranks)
    assert np.array_equal(x, np.vstack(parts)) and np.array_equal(color, [0, 0, 1, 1, 1])
    # a truncated rank file is an error
    (tmp_path / "step_7_rank_1.pst").write_bytes(struct.pack("<i", 3) + parts[1][:2].tobytes())
    with pytest.raises(smg_mod.SmgError):
        smg_mod.read_post_file(tmp_path, 7, ranks)
