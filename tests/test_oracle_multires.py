"""The oracle's multi-resolution sampling (oracle.set_levels), restated from the rule
alone: a sample is the single-level interpolation on the finest refined level that
stores the packages of all 2^D corner cells of the interpolation (a halo package
counts as its interior neighbour), else on the base grid; a cell tag comes from the
finest level that stores the cell's package, else the base grid.  Single-level
samples are pinned by test_oracle_lset.py."""
import numpy as np
import pytest


def _levels(smg_mod, dim):
    if dim == 2:
        cfg, sdf = smg_mod.config_c1(num_pkg=8, dx=0.02)   # base 32^2, finest 128^2
    else:
        cfg, sdf = smg_mod.config_c3(num_pkg=4, dx0=0.012, ratio=2.0, band=0.2)
    base = smg_mod.make_levelset(cfg, sdf)
    lv = smg_mod.make_levelset_levels(cfg, sdf, 2, band=2.0)
    assert all(0 < L["active"].sum() < L["active"].size for L in lv)
    return cfg, base, lv


def _geometry(cl, pts, D):
    """Package and interpolation base cell of each point on level config cl."""
    np_ = [cl.num_pkg[a] for a in range(D)]
    pk = np.zeros((len(pts), D), int)
    m = np.zeros((len(pts), D), int)
    ci = np.zeros((len(pts), D), int)
    for a in range(D):
        dpkg = cl.domain[a] / float(cl.num_pkg[a])
        dcell = dpkg / 4.0
        cs = np.minimum(np.maximum(pts[:, a] - cl.box_l[a], 0.), cl.domain[a] - 1.e-8)
        p = np.clip(np.floor(cs / dpkg).astype(int), -1, np_[a])
        rr = np.clip((cs - dpkg * p) / dcell - 0.5, -0.5, 3.5)
        pk[:, a], m[:, a] = p, np.floor(rr).astype(int)
        ci[:, a] = np.clip(np.floor((cs - dpkg * p) / dcell).astype(int), 0, 3)
    return np_, pk, m, ci


def _stored(L, np_, idx, D):
    act = L["active"].reshape([np_[a] for a in reversed(range(D))])
    cl = [np.clip(idx[a], 0, np_[a] - 1) for a in range(D)]   # a halo package: its interior neighbour
    return act[tuple(reversed(cl))] != 0


@pytest.mark.parametrize("dim", [2, 3])
def test_multilevel_samples_and_tags(smg_mod, orc, dim):
    cfg, base, lv = _levels(smg_mod, dim)
    rng = np.random.default_rng(3 + dim)
    pts = rng.uniform(-0.02, 1.02, (20000, dim))
    phi_w, n_w, tag_w = orc.sample(cfg, base, pts)
    lev_s = np.zeros(len(pts), int)
    lev_t = np.zeros(len(pts), int)
    for l, L in enumerate(lv):
        cl = L["cfg"]
        np_, pk, m, ci = _geometry(cl, pts, dim)
        ok = np.ones(len(pts), bool)
        for corner in range(1 << dim):
            g = [pk[:, a] * 4 + m[:, a] + ((corner >> a) & 1) for a in range(dim)]
            ok &= _stored(L, np_, [np.floor_divide(x, 4) for x in g], dim)
        ph, nn, tg = orc.sample(cl, L, pts)
        phi_w = np.where(ok, ph, phi_w)
        n_w = np.where(ok[:, None], nn, n_w)
        lev_s = np.where(ok, l + 1, lev_s)
        okt = _stored(L, np_, [pk[:, a] for a in range(dim)], dim)
        tag_w = np.where(okt, tg, tag_w)
        lev_t = np.where(okt, l + 1, lev_t)
    assert np.bincount(lev_s, minlength=3).min() > 0 and np.bincount(lev_t, minlength=3).min() > 0
    orc.set_levels(lv)
    try:
        phi, n, tag = orc.sample(cfg, base, pts)
    finally:
        orc.set_levels(None)
    assert np.array_equal(phi, phi_w)
    assert np.array_equal(n, n_w)
    assert np.array_equal(tag, tag_w)
