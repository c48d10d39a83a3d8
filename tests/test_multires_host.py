"""Host side of the multi-resolution level set: the seeded ellipsoid-union signed
distance of the multi-resolution test case and the band selection of the refined
levels (no GPU)."""
import numpy as np


def test_ellipsoid_distance_reduces_to_a_sphere(smg_mod):
    # equal semi-axes: the exact branch must give |p - c| - r for any rotation
    f = smg_mod.sdf_ellipsoid_union(n=1, seed=7, amin=0.15, amax=0.15, exact_band=0.05)
    rng = np.random.default_rng(0)
    c = np.random.default_rng(7).uniform(0.2, 0.8, (1, 3))[0]
    d = rng.normal(size=(2000, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    rad = 0.15 + rng.uniform(-0.04, 0.04, 2000)
    P = c + d * rad[:, None]
    phi = f(P[:, 0], P[:, 1], P[:, 2])
    assert np.allclose(phi, 0.15 - rad, atol=1e-12, rtol=0)


def test_ellipsoid_distance_closest_point_is_normal(smg_mod):
    # an axis-aligned ellipsoid with distinct axes: the distance is the length of a
    # segment normal to the surface (checked against a dense surface sampling)
    f = smg_mod.sdf_ellipsoid_union(n=1, seed=11, exact_band=0.05)
    rng = np.random.default_rng(11)
    c = rng.uniform(0.2, 0.8, (1, 3))[0]
    a = rng.uniform(0.05, 0.2, (1, 3))[0]
    u = rng.uniform(0., 1., (1, 3))[0]
    q = np.array([np.sqrt(1 - u[0]) * np.sin(2 * np.pi * u[1]), np.sqrt(1 - u[0]) * np.cos(2 * np.pi * u[1]),
                  np.sqrt(u[0]) * np.sin(2 * np.pi * u[2]), np.sqrt(u[0]) * np.cos(2 * np.pi * u[2])])
    x, y, z, w = q
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    th, ph = np.meshgrid(np.linspace(0, np.pi, 700), np.linspace(0, 2 * np.pi, 1400), indexing="ij")
    S = np.stack([a[0] * np.sin(th) * np.cos(ph), a[1] * np.sin(th) * np.sin(ph), a[2] * np.cos(th)], -1).reshape(-1, 3)
    W = S @ R.T + c   # surface points in world coordinates
    P = W[rng.integers(0, len(W), 60)] + rng.normal(size=(60, 3)) * 0.02   # around the surface
    phi = f(P[:, 0], P[:, 1], P[:, 2])
    near = (np.abs(phi) > 0.005) & (np.abs(phi) < 0.05)
    assert near.sum() > 5
    for p, v in zip(P[near], phi[near]):
        dmin = np.sqrt(((W - p) ** 2).sum(1)).min()
        assert abs(abs(v) - dmin) < 2e-4   # the surface sampling's resolution


def test_bands_select_packages_near_the_surface(smg_mod):
    cfg, sdf = smg_mod.config_c1(num_pkg=4, dx=0.02)
    full = smg_mod.make_levelset_levels(cfg, sdf, 2, band=None)
    assert all(L["active"].all() for L in full)
    part = smg_mod.make_levelset_levels(cfg, sdf, 2, band=1.0)
    for l, L in enumerate(part, start=1):
        npk = cfg.num_pkg[0] << l
        assert L["active"].size == npk * npk
        act = L["active"].reshape(npk, npk)
        # package centres: active ones lie within band + half a package diagonal
        dpkg = 1.0 / npk
        cx = (np.arange(npk) + 0.5) * dpkg
        X, Y = np.meshgrid(cx, cx)
        dist = np.abs(np.sqrt((X - 0.5) ** 2 + (Y - 0.5) ** 2) - 0.2)
        assert (dist[act == 1] <= 1.0 * dpkg + np.sqrt(2) * dpkg / 2 + 1e-12).all()
        assert (dist[act == 0] >= 1.0 * dpkg - np.sqrt(2) * dpkg / 2 - 1e-12).all()
        assert 0 < act.sum() < act.size
