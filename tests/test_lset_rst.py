"""The reference's level-set restart file (./rstfile/levelset.rst): the library writes
and reads it.  The layout is checked by a parser written here from the reference's
read order (Load_rst_case, sph_mesh_generation.cpp:341-373; Load_lset_restart_file,
level_set.cpp:1098-1142; Load_level_set_rstfile, lset_level_infor.cpp:952-1146), a
round trip keeps every value, and the particle count recomputed from a restart equals
the oracle's level-set builder with zero margins (the reference's restart mode)."""
import struct

import numpy as np
import pytest


def _parse(path):
    """A plain restatement of the reference's reader: returns the headers and the cell
    fields (interior cells, x fastest)."""
    b = open(path, "rb").read()
    o = 0

    def take(fmt):
        nonlocal o
        v = struct.unpack_from("<" + fmt, b, o)
        o += struct.calcsize("<" + fmt)
        return v

    h = {}
    (h["DIM"],) = take("i")
    for k in ("domain", "box_l", "box_r", "particle_box", "particle_box_l", "particle_box_r"):
        h[k] = take("3d")
    h["resolution"] = take("3i")
    h["Lmin"], h["Lmax"] = take("2i")
    h["ini_num_cell"] = take("3i")
    (h["num_level"],) = take("i")
    (h["convergence_error"],) = take("d")
    (ns,) = take("i")
    take("%dd" % (3 * ns))
    (nseg,) = take("i")
    take("%dd" % (6 * nseg))
    names = ["fac_maximum_dl", "fac_middel_dl", "fac_minimum_dl", "maximum_curv_artificial", "total_mass",
             "total_mass_surface", "total_mass_segment", "total_volume", "total_area", "total_length",
             "maximum_phi", "maximum_psi", "maximum_dl", "middel_dl", "minimum_dl", "maximum_curv", "minimum_curv"]
    for k, v in zip(names, take("17d")):
        h[k] = v
    (h["glbl_num_interface_cell"],) = take("i")
    (h["glbl_total_num_pkg"],) = take("i")
    h["scale"], h["dl"] = take("2d")
    for k in ("num_pkg", "glbl_num_pkg", "glbl_pkg_start", "glbl_pkg_end", "pkg_start", "pkg_end", "num_pkg_margin"):
        h[k] = take("3i")
    h["dpkg"] = take("3d")
    h["dcell"] = take("3d")
    # the loops run over k in [max(pkg_start,0), min(glbl_num_pkg,num_pkg)) * ICP, etc.
    D = h["DIM"]
    icp = [4, 4, 4 if D > 2 else 1]
    n = 1
    for a in range(3):
        n *= (min(h["glbl_num_pkg"][a], h["num_pkg"][a]) - max(h["pkg_start"][a], 0)) * icp[a]
    f = {}
    for k in ("phi", "n_x", "n_y", "n_z", "curv", "psi"):
        f[k] = np.array(take("%dd" % n))
    for k in ("tag_interface", "tag_characteristic", "idx_characteristic"):
        f[k] = np.array(take("%di" % n))
    f["scale"] = np.array(take("%dd" % n))
    assert o == len(b), "trailing bytes"
    return h, f


def _interior(cfg, arr, comps=1):
    D = cfg.dim
    nc = [4 * cfg.num_pkg[a] if a < D else 1 for a in range(3)]
    halo = [nc[a] + 2 if a < D else 1 for a in range(3)]
    a = np.asarray(arr).reshape(halo[2], halo[1], halo[0], comps) if comps > 1 else \
        np.asarray(arr).reshape(halo[2], halo[1], halo[0])
    sl = (slice(1, -1) if D > 2 else slice(None), slice(1, -1), slice(1, -1))
    return a[sl]


@pytest.fixture(scope="module", params=["2d", "3d"])
def case(request, smg_mod):
    if request.param == "2d":
        cfg, sdf = smg_mod.config_c1(num_pkg=16, dx=0.02)
    else:
        # the reference's tanh spacing profile (the only one a restart file carries)
        cfg, sdf = smg_mod.config_c2(scale=3.0, num_pkg=8)
        cfg.profile = smg_mod.PROFILE_TANH
        cfg.min_dl, cfg.max_dl = 0.024, 0.072
    return cfg, smg_mod.make_levelset(cfg, sdf)


def test_layout_follows_the_reference_reader(smg_mod, case, tmp_path):
    cfg, ls = case
    path = tmp_path / "levelset.rst"
    smg_mod.write_levelset_rst(str(path), cfg, ls)
    h, f = _parse(path)
    D = cfg.dim
    assert h["DIM"] == D
    assert h["domain"] == tuple(cfg.domain) and h["box_l"] == tuple(cfg.box_l)
    assert h["particle_box_l"] == tuple(cfg.particle_box_l) and h["particle_box_r"] == tuple(cfg.particle_box_r)
    assert h["num_pkg"][:D] == tuple(cfg.num_pkg)[:D]
    assert h["dcell"][0] == pytest.approx(cfg.domain[0] / cfg.num_pkg[0] / 4.0, rel=1e-15)
    st = ls["stats"]
    assert h["maximum_phi"] == st.maximum_phi and h["total_mass"] == st.total_mass
    assert h["glbl_num_interface_cell"] == st.n_interface_cell
    assert h["maximum_dl"] == cfg.max_dl and h["minimum_dl"] == cfg.min_dl
    # cell fields: interior cells, x fastest
    assert np.array_equal(f["phi"], _interior(cfg, ls["phi"]).ravel())
    nrm = _interior(cfg, ls["normal"], 3).reshape(-1, 3)
    for a, k in enumerate(("n_x", "n_y", "n_z")):
        assert np.array_equal(f[k], nrm[:, a].astype(np.float64))
    assert np.array_equal(f["tag_interface"], _interior(cfg, ls["tag"]).ravel().astype(np.int64))
    assert f["tag_interface"].sum() == st.n_interface_cell
    assert np.all(f["tag_characteristic"] == 0)
    assert np.all(f["scale"] > 0)


def test_round_trip(smg_mod, case, tmp_path):
    cfg, ls = case
    path = tmp_path / "levelset.rst"
    smg_mod.write_levelset_rst(str(path), cfg, ls)
    c2, l2, info, phi_ext = smg_mod.read_levelset_rst(str(path))
    D = cfg.dim
    assert c2.dim == D and list(c2.num_pkg)[:D] == list(cfg.num_pkg)[:D]
    assert list(c2.box_l) == list(cfg.box_l) and list(c2.domain) == list(cfg.domain)
    assert c2.min_dl == cfg.min_dl and c2.max_dl == cfg.max_dl
    # interior cells as written, the halo zero (the margin packages are not in the file)
    for k, comps in (("phi", 1), ("normal", 3), ("tag", 1)):
        assert np.array_equal(_interior(cfg, l2[k], comps), _interior(cfg, ls[k], comps)), k
    halo_mask = np.ones_like(l2["phi"], bool)
    halo_mask.reshape(_interior_shape(cfg))[_interior_slices(cfg)] = False
    assert np.all(l2["phi"][halo_mask] == 0) and np.all(l2["tag"][halo_mask] == 0)
    assert info.maximum_phi == ls["stats"].maximum_phi and info.n_cells == _interior(cfg, ls["phi"]).size


def _interior_shape(cfg):
    D = cfg.dim
    return tuple(4 * cfg.num_pkg[a] + 2 if a < D else 1 for a in (2, 1, 0))


def _interior_slices(cfg):
    return (slice(1, -1) if cfg.dim > 2 else slice(None), slice(1, -1), slice(1, -1))


def test_restart_particle_count_equals_oracle(smg_mod, orc, case, tmp_path):
    """Restart mode: the particle count from phi with zero margins (quirk Q9), by the
    library's builder and by the oracle's independent one."""
    cfg, ls = case
    path = tmp_path / "levelset.rst"
    smg_mod.write_levelset_rst(str(path), cfg, ls)
    c2, l2, info, phi_ext = smg_mod.read_levelset_rst(str(path))
    ref = orc.lset_build(c2, phi_ext, zero_margin=True)
    assert l2["stats"].num_particle == ref["stats"].num_particle
    assert l2["stats"].total_volume == pytest.approx(ref["stats"].total_volume, rel=1e-12)
    # the interior is the same field: the counts of restart and first pass agree closely
    assert abs(l2["stats"].num_particle - ls["stats"].num_particle) <= 0.02 * ls["stats"].num_particle


def test_bad_files_are_rejected(smg_mod, tmp_path):
    bad = tmp_path / "bad.rst"
    bad.write_bytes(struct.pack("<i", 7) + b"\0" * 64)
    with pytest.raises(smg_mod.SmgError):
        smg_mod.read_levelset_rst(str(bad))
    trunc = tmp_path / "trunc.rst"
    cfg, sdf = smg_mod.config_c1(num_pkg=8, dx=0.04)
    smg_mod.write_levelset_rst(str(trunc), cfg, smg_mod.make_levelset(cfg, sdf))
    trunc.write_bytes(trunc.read_bytes()[:-100])
    with pytest.raises(smg_mod.SmgError):
        smg_mod.read_levelset_rst(str(trunc))
    with pytest.raises(smg_mod.SmgError):
        smg_mod.read_levelset_rst(str(tmp_path / "missing.rst"))


def _write_sdf(path, phi, origin, dx):
    """SDFGen's text layout: ni nj nk / origin / dx / values, i fastest."""
    nk, nj, ni = phi.shape
    with open(path, "w") as f:
        f.write(f"{ni} {nj} {nk}\n{origin[0]} {origin[1]} {origin[2]}\n{dx!r}\n")
        for v in phi.ravel():
            f.write(f"{float(v)!r}\n")


def test_sdf_file_fills_the_interior(smg_mod, tmp_path):
    cfg, sdf = smg_mod.config_c2(scale=3.0, num_pkg=8)
    nc = 4 * cfg.num_pkg[0]
    dl = cfg.domain[0] / cfg.num_pkg[0] / 4.0
    X, Y, Z = smg_mod.cell_centres_ext(cfg)
    full = sdf(X, Y, Z)
    interior = full[4:-4, 4:-4, 4:-4]
    path = tmp_path / "case.sdf"
    _write_sdf(path, interior, (0.0, 0.0, 0.0), dl)
    phi_ext = smg_mod.read_levelset_sdf(str(path), cfg).reshape(full.shape)
    assert np.array_equal(phi_ext[4:-4, 4:-4, 4:-4], interior)
    margin = np.ones(full.shape, bool)
    margin[4:-4, 4:-4, 4:-4] = False
    assert np.all(phi_ext[margin] == 0.0)
    # a file larger than the grid: the extra cells are skipped; a smaller one leaves zeros
    big = np.pad(interior, ((0, 2), (0, 3), (0, 1)), constant_values=9.0)
    _write_sdf(path, big, (0.0, 0.0, 0.0), dl)
    assert np.array_equal(smg_mod.read_levelset_sdf(str(path), cfg).reshape(full.shape)[4:-4, 4:-4, 4:-4], interior)
    _write_sdf(path, interior[:, :, : nc - 4], (0.0, 0.0, 0.0), dl)
    part = smg_mod.read_levelset_sdf(str(path), cfg).reshape(full.shape)[4:-4, 4:-4, 4:-4]
    assert np.array_equal(part[:, :, : nc - 4], interior[:, :, : nc - 4]) and np.all(part[:, :, nc - 4:] == 0)
    # the level set built from it: the sphere hole's volume to the grid's accuracy
    ls = smg_mod.lset_build(cfg, phi_ext, zero_margin=True)
    assert ls["stats"].total_volume == pytest.approx(1.0 - 4.0 / 3.0 * np.pi * 0.3 ** 3, rel=0.05)


def test_sdf_file_errors(smg_mod, tmp_path):
    cfg, sdf = smg_mod.config_c2(scale=3.0, num_pkg=8)
    path = tmp_path / "case.sdf"
    _write_sdf(path, np.zeros((4, 4, 4)), (0.0, 0.0, 0.0), 0.5)   # dx differs from the cell size
    with pytest.raises(smg_mod.SmgError):
        smg_mod.read_levelset_sdf(str(path), cfg)
    cfg2, _ = smg_mod.config_c1()
    with pytest.raises(smg_mod.SmgError):
        smg_mod.read_levelset_sdf(str(path), cfg2)   # 2D
    dl = cfg.domain[0] / cfg.num_pkg[0] / 4.0
    path.write_text(f"4 4 4\n0 0 0\n{dl!r}\n1.0\n2.0\n")   # truncated values
    with pytest.raises(smg_mod.SmgError):
        smg_mod.read_levelset_sdf(str(path), cfg)
