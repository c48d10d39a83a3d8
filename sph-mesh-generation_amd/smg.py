"""Python binding of the MI355X relaxation engine (ctypes over libsmg.so).

The C ABI (include/smg.h) is the product boundary; this module is a thin host
mirror of it plus the level-set and configuration builders for the benchmark
geometries.  It loads lib/libsmg.so from this directory and raises if that file is
missing -- there is no CPU fallback for the relaxation step.
"""
import ctypes as C
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libsmg.so")

SMG_OK = 0
SMG_ERR_ARG, SMG_ERR_HIP, SMG_ERR_CAPACITY, SMG_ERR_STATE, SMG_ERR_REMAP, SMG_ERR_NONFINITE = -1, -2, -3, -4, -5, -6
SMG_REAL, SMG_SURFACE = 0, 1
PROFILE_TANH, PROFILE_LINEAR = 0, 1
PI_REF = 3.1415926   # glbparam.h:185


class SmgConfig(C.Structure):
    _fields_ = [
        ("dim", C.c_int32), ("sym", C.c_int32 * 3),
        ("box_l", C.c_double * 3), ("domain", C.c_double * 3),
        ("particle_box_l", C.c_double * 3), ("particle_box_r", C.c_double * 3),
        ("num_pkg", C.c_int32 * 3),
        ("cell_ratio", C.c_double), ("cut_off", C.c_double),
        ("angle_thres", C.c_double), ("nu_coeff", C.c_double),
        ("profile", C.c_int32),
        ("min_dl", C.c_double), ("max_dl", C.c_double), ("max_phi", C.c_double),
        ("lin_dx0", C.c_double), ("lin_dx1", C.c_double), ("lin_band", C.c_double),
        ("seed", C.c_uint64), ("remap_max_attempts", C.c_int32), ("max_neighbors", C.c_int32),
        ("spill_per_slot", C.c_int32), ("reserved0", C.c_int32), ("skin", C.c_double),
    ]


class SmgLsetStats(C.Structure):
    _fields_ = [
        ("maximum_phi", C.c_double), ("total_volume", C.c_double), ("total_area", C.c_double),
        ("total_mass", C.c_double), ("total_mass_surface", C.c_double),
        ("n_interface_cell", C.c_int64), ("num_particle_mesh", C.c_int64),
        ("num_particle_surface", C.c_int64), ("num_particle", C.c_int64),
    ]


class SmgLsetLevel(C.Structure):
    _fields_ = [("phi", C.c_void_p), ("normal", C.c_void_p), ("tag", C.c_void_p), ("active", C.c_void_p)]


class SmgStats(C.Structure):
    _fields_ = [
        ("step", C.c_int64), ("n_particle", C.c_int64), ("n_real", C.c_int64), ("n_surface", C.c_int64),
        ("n_remapped", C.c_int64), ("n_neighbor_entries", C.c_int64),
        ("max_neighbors_seen", C.c_int32), ("num_level", C.c_int32),
        ("dt1", C.c_double), ("dt2", C.c_double), ("max_v", C.c_double),
        ("residual", C.c_double), ("max_disp", C.c_double), ("res_max", C.c_double),
        ("n_rebuild", C.c_int64), ("drift", C.c_double), ("residual_sum", C.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def lib():
    """The loaded libsmg.so; raises FileNotFoundError when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} is missing: build it with build_lib.py")
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        L.smg_last_error.restype = C.c_char_p
        L.smg_default_config.argtypes = [C.c_int, P(SmgConfig)]
        L.smg_lset_build.argtypes = [P(SmgConfig), C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, P(SmgLsetStats)]
        L.smg_create.argtypes = [P(SmgConfig), P(C.c_void_p)]
        L.smg_destroy.argtypes = [C.c_void_p]
        L.smg_set_levelset.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.smg_set_levelset_levels.argtypes = [C.c_void_p, C.c_int32, P(SmgLsetLevel)]
        L.smg_sample_phi.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]
        L.smg_sample_phi_static.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]
        L.smg_init_particles.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]
        L.smg_seed_particles.argtypes = [C.c_void_p, C.c_int64]
        L.smg_prerun.argtypes = [C.c_void_p]
        L.smg_step.argtypes = [C.c_void_p, C.c_int64]
        L.smg_get_timestep.argtypes = [C.c_void_p, P(C.c_double)]
        L.smg_get_stats.argtypes = [C.c_void_p, P(SmgStats)]
        L.smg_get_state.argtypes = [C.c_void_p] + [C.c_void_p] * 5
        L.smg_get_pair_terms.argtypes = [C.c_void_p] + [C.c_void_p] * 3
        L.smg_get_neighbors.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64]
        L.smg_profile_enable.argtypes = [C.c_void_p, C.c_int]
        L.smg_profile_get.argtypes = [C.c_void_p, P(C.c_char_p), P(C.c_double), P(C.c_int64), C.c_int, P(C.c_int)]
        L.smg_fp64_probe.argtypes = [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p]
        L.smg_step_async.argtypes = [C.c_void_p, C.c_int64]
        L.smg_query.argtypes = [C.c_void_p, P(C.c_int32)]
        L.smg_sync.argtypes = [C.c_void_p]
        L.smg_capture.argtypes = [C.c_void_p, C.c_int64]
        L.smg_footprint.argtypes = [C.c_void_p, P(C.c_int64)]
        _lib = L
    return _lib


class SmgError(RuntimeError):
    def __init__(self, rc, where):
        super().__init__(f"{where} failed ({rc}): {lib().smg_last_error().decode()}")
        self.rc = rc


def _check(rc, where):
    if rc != SMG_OK:
        raise SmgError(rc, where)


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


# ---------------------------------------------------------------------------
# configuration and level sets

def default_config(dim):
    cfg = SmgConfig()
    _check(lib().smg_default_config(dim, C.byref(cfg)), "smg_default_config")
    return cfg


def config_copy(cfg):
    c2 = SmgConfig()
    C.memmove(C.byref(c2), C.byref(cfg), C.sizeof(SmgConfig))
    return c2


def cell_centres_ext(cfg):
    """Cell centres of the grid with a full margin package (4 cells) on each side."""
    D = cfg.dim
    axes = []
    for a in range(3):
        if a < D:
            dpkg = cfg.domain[a] / float(cfg.num_pkg[a])
            dcell = dpkg / 4.0
            g = np.arange(-4, 4 * cfg.num_pkg[a] + 4)
            # pkg*dpkg + (cell+0.5)*dcell + box_l with global cell g = 4*pkg + cell
            pkg = np.floor_divide(g, 4)
            cell = g - 4 * pkg
            axes.append(pkg * dpkg + (cell + 0.5) * dcell + cfg.box_l[a])
        else:
            axes.append(np.zeros(1))
    Z, Y, X = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
    return X, Y, Z


def lset_build(cfg, phi_ext, zero_margin=True):
    """Normals, cut tags and statistics of a level set given on the extended grid."""
    D = cfg.dim
    halo = [4 * cfg.num_pkg[a] + 2 if a < D else 1 for a in range(3)]
    m = halo[0] * halo[1] * halo[2]
    phi = np.zeros(m, np.float64)
    normal = np.zeros(3 * m, np.float32)
    tag = np.zeros(m, np.int8)
    st = SmgLsetStats()
    ext = np.ascontiguousarray(phi_ext, np.float64).ravel()
    _check(lib().smg_lset_build(C.byref(cfg), _ptr(ext), int(zero_margin), _ptr(phi), _ptr(normal), _ptr(tag),
                                C.byref(st)), "smg_lset_build")
    return {"phi": phi, "normal": normal, "tag": tag, "stats": st}


def lset_build_gpu(cfg, phi_ext, zero_margin=True):
    """lset_build on the GPU (smg_lset_build_gpu): the same dict, plus "kernel_ms",
    the device time of the build (transfers excluded)."""
    D = cfg.dim
    halo = [4 * cfg.num_pkg[a] + 2 if a < D else 1 for a in range(3)]
    ext_n = [4 * cfg.num_pkg[a] + 8 if a < D else 1 for a in range(3)]
    m = halo[0] * halo[1] * halo[2]
    ext = np.ascontiguousarray(phi_ext, np.float64).ravel()
    if ext.size != ext_n[0] * ext_n[1] * ext_n[2]:
        raise ValueError(f"phi_ext has {ext.size} values, the grid {ext_n} needs {ext_n[0] * ext_n[1] * ext_n[2]}")
    phi = np.zeros(m, np.float64)
    normal = np.zeros(3 * m, np.float32)
    tag = np.zeros(m, np.int8)
    st = SmgLsetStats()
    ms = C.c_double()
    _check(lib().smg_lset_build_gpu(C.byref(cfg), _ptr(ext), int(zero_margin), _ptr(phi), _ptr(normal), _ptr(tag),
                                    C.byref(st), C.byref(ms)), "smg_lset_build_gpu")
    return {"phi": phi, "normal": normal, "tag": tag, "stats": st, "kernel_ms": ms.value}


def sdf_circle_hole(center, radius):
    """phi > 0 outside the circle / sphere (the fluid), SPH-mesh-generation's sign convention."""
    c = np.asarray(center, float)

    def f(X, Y, Z):
        r2 = (X - c[0]) ** 2 + (Y - c[1]) ** 2
        if len(c) > 2:
            r2 = r2 + (Z - c[2]) ** 2
        return np.sqrt(r2) - radius
    return f


def sdf_sphere_union(centers, radii, dim):
    """phi > 0 inside the union of circles / spheres (the fluid is the union)."""
    cs = np.asarray(centers, float)
    rs = np.asarray(radii, float)

    def f(X, Y, Z):
        out = None
        for c, r in zip(cs, rs):
            r2 = (X - c[0]) ** 2 + (Y - c[1]) ** 2
            if dim > 2:
                r2 = r2 + (Z - c[2]) ** 2
            v = r - np.sqrt(r2)
            out = v if out is None else np.maximum(out, v)
        return out
    return f


class SmgRstInfo(C.Structure):
    _fields_ = [
        ("resolution", C.c_int32 * 3), ("Lmin", C.c_int32), ("Lmax", C.c_int32), ("ini_num_cell", C.c_int32 * 3),
        ("num_level", C.c_int32), ("num_singularity", C.c_int32), ("num_segment", C.c_int32),
        ("glbl_num_interface_cell", C.c_int32), ("reserved", C.c_int32),
        ("convergence_error", C.c_double),
        ("fac_maximum_dl", C.c_double), ("fac_middel_dl", C.c_double), ("fac_minimum_dl", C.c_double),
        ("maximum_curv_artificial", C.c_double),
        ("total_mass", C.c_double), ("total_mass_surface", C.c_double), ("total_mass_segment", C.c_double),
        ("total_volume", C.c_double), ("total_area", C.c_double), ("total_length", C.c_double),
        ("maximum_phi", C.c_double), ("maximum_psi", C.c_double), ("maximum_dl", C.c_double),
        ("middel_dl", C.c_double), ("minimum_dl", C.c_double), ("maximum_curv", C.c_double),
        ("minimum_curv", C.c_double), ("scale", C.c_double), ("dl", C.c_double),
        ("n_cells", C.c_int64), ("n_halo_cells", C.c_int64), ("n_ext_cells", C.c_int64),
    ]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}
        for k in ("resolution", "ini_num_cell"):
            d[k] = list(d[k])
        return d


def write_levelset_rst(path, cfg, lset):
    """The reference's levelset.rst layout (smg_lset_write_rst) for a level set of
    make_levelset / lset_build."""
    L = lib()
    L.smg_lset_write_rst.argtypes = [C.c_char_p, C.POINTER(SmgConfig), C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.POINTER(SmgLsetStats)]
    _check(L.smg_lset_write_rst(os.fsencode(path), C.byref(cfg), _ptr(lset["phi"]), _ptr(lset["normal"]),
                                _ptr(lset["tag"]), C.byref(lset["stats"])), "smg_lset_write_rst")


def read_levelset_sdf(path, cfg):
    """phi on the full-margin grid from an SDFGen .sdf file (smg_lset_read_sdf)."""
    L = lib()
    L.smg_lset_read_sdf.argtypes = [C.c_char_p, C.POINTER(SmgConfig), C.c_void_p]
    ext = [4 * cfg.num_pkg[a] + 8 if a < cfg.dim else 1 for a in range(3)]
    phi_ext = np.zeros(ext[0] * ext[1] * ext[2])
    _check(L.smg_lset_read_sdf(os.fsencode(path), C.byref(cfg), _ptr(phi_ext)), "smg_lset_read_sdf")
    return phi_ext


def read_levelset_rst(path):
    """(cfg, lset, info, phi_ext) from a levelset.rst (smg_lset_read_rst): lset holds
    the file's phi / normal / tag on the one-cell-halo grid and, as "stats", the
    statistics smg_lset_build recomputes from phi with zero margins (the reference's
    restart mode); info is the file's header."""
    L = lib()
    L.smg_lset_read_rst.argtypes = [C.c_char_p, C.POINTER(SmgConfig), C.POINTER(SmgRstInfo), C.c_void_p,
                                    C.c_void_p, C.c_void_p, C.c_void_p]
    cfg, info = SmgConfig(), SmgRstInfo()
    _check(L.smg_lset_read_rst(os.fsencode(path), C.byref(cfg), C.byref(info), None, None, None, None),
           "smg_lset_read_rst")
    nh, ne = int(info.n_halo_cells), int(info.n_ext_cells)
    phi = np.zeros(nh); normal = np.zeros(3 * nh, np.float32); tag = np.zeros(nh, np.int8)
    phi_ext = np.zeros(ne)
    _check(L.smg_lset_read_rst(os.fsencode(path), C.byref(cfg), C.byref(info), _ptr(phi), _ptr(normal), _ptr(tag),
                               _ptr(phi_ext)), "smg_lset_read_rst")
    stats = lset_build(cfg, phi_ext, zero_margin=True)["stats"]
    return cfg, {"phi": phi, "normal": normal, "tag": tag, "stats": stats}, info, phi_ext


def make_levelset(cfg, sdf, zero_margin=True):
    X, Y, Z = cell_centres_ext(cfg)
    return lset_build(cfg, sdf(X, Y, Z), zero_margin)


def config_refined(cfg, level):
    """The config of refined level-set level `level` (num_pkg << level packages)."""
    c = config_copy(cfg)
    for a in range(cfg.dim):
        c.num_pkg[a] = cfg.num_pkg[a] << level
    return c


def make_levelset_levels(cfg, sdf, nlev, band=4.0, zero_margin=True):
    """Refined levels 1..nlev over the base level set of `cfg` (the multi-resolution
    extension, smg_set_levelset_levels): level l is built like a single-level set with
    num_pkg << l packages, and keeps the packages that hold a cell with |phi| < band
    times the level's package size; band = None keeps every package."""
    D = cfg.dim
    out = []
    for l in range(1, nlev + 1):
        cl = config_refined(cfg, l)
        ls = make_levelset(cl, sdf, zero_margin)
        np_ = [cl.num_pkg[a] if a < D else 1 for a in range(3)]
        halo = [4 * np_[a] + 2 if a < D else 1 for a in range(3)]
        phi = ls["phi"].reshape(halo[2], halo[1], halo[0])
        inner = phi[1:-1, 1:-1, 1:-1] if D > 2 else phi[:, 1:-1, 1:-1]
        if band is None:
            act = np.ones((np_[2], np_[1], np_[0]), np.uint8)
        else:
            dpkg = cl.domain[0] / float(cl.num_pkg[0])
            near = np.abs(inner) < band * dpkg
            if D > 2:
                act = near.reshape(np_[2], 4, np_[1], 4, np_[0], 4).any(axis=(1, 3, 5))
            else:
                act = near.reshape(1, np_[1], 4, np_[0], 4).any(axis=(2, 4))
            act = act.astype(np.uint8)
        ls["active"] = np.ascontiguousarray(act.ravel())
        ls["cfg"] = cl
        out.append(ls)
    return out


def sdf_ellipsoid_union(n=12, seed=20261203, lo=0.2, hi=0.8, amin=0.05, amax=0.2, exact_band=0.04):
    """phi > 0 inside the union of n seeded ellipsoids (centres uniform in [lo, hi]^3,
    semi-axes uniform in [amin, amax], rotations from uniform unit quaternions): the
    signed distance to each ellipsoid, exact (the closest-point equation solved by
    bisection on its Lagrange multiplier) within exact_band of the surface and the
    algebraic estimate (|p/a| - 1) min(a) further out, combined by max."""
    rng = np.random.default_rng(seed)
    cs = rng.uniform(lo, hi, (n, 3))
    ax = rng.uniform(amin, amax, (n, 3))
    # Shoemake: uniform unit quaternions
    u = rng.uniform(0., 1., (n, 3))
    q = np.stack([np.sqrt(1 - u[:, 0]) * np.sin(2 * np.pi * u[:, 1]), np.sqrt(1 - u[:, 0]) * np.cos(2 * np.pi * u[:, 1]),
                  np.sqrt(u[:, 0]) * np.sin(2 * np.pi * u[:, 2]), np.sqrt(u[:, 0]) * np.cos(2 * np.pi * u[:, 2])], 1)
    Rs = []
    for x, y, z, w in q:
        Rs.append(np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                            [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                            [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]]))

    def dist(P, a):   # signed distance to the axis-aligned ellipsoid (positive outside)
        P = np.abs(P)
        g = np.sqrt(((P / a) ** 2).sum(1))
        out = (g - 1.0) * a.min()
        sel = np.abs(out) < exact_band
        if sel.any():
            p = np.maximum(P[sel], 1e-300)
            a2 = a * a
            # F(t) = sum (a_i p_i / (t + a_i^2))^2 - 1 decreases in t > -min a^2; its root
            # gives the closest point x_i = a_i^2 p_i / (t + a_i^2)
            t_lo = np.full(len(p), -a2.min() + 1e-14)
            t_hi = np.full(len(p), np.linalg.norm(p, axis=1) * a.max())
            for _ in range(48):   # the bracket starts below |p| max(a) <= 0.2: 2^-48 of it
                t = 0.5 * (t_lo + t_hi)
                F = ((a * p / (t[:, None] + a2)) ** 2).sum(1) - 1.0
                t_lo = np.where(F > 0, t, t_lo)
                t_hi = np.where(F > 0, t_hi, t)
            t = 0.5 * (t_lo + t_hi)
            xc = a2 * p / (t[:, None] + a2)
            d = np.linalg.norm(p - xc, axis=1)
            out[sel] = np.where(g[sel] >= 1.0, d, -d)
        return out

    def f(X, Y, Z):
        P = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
        best = None
        for c, a, R in zip(cs, ax, Rs):
            v = -dist((P - c) @ R, a)
            best = v if best is None else np.maximum(best, v)
        return best.reshape(X.shape)
    return f


def config_c3(num_pkg=16, dx0=0.0015, ratio=8.0, band=0.1, seed=20261203):
    """3D multi-resolution case: inside the union of 12 seeded ellipsoids in the unit
    cube; dx linear in phi from dx0 at the surface to ratio * dx0 at phi >= band; the
    base level set (4 num_pkg)^3 cells, refined twice (make_levelset_levels)."""
    cfg = default_config(3)
    for a in range(3):
        cfg.num_pkg[a] = num_pkg
    cfg.profile = PROFILE_LINEAR
    cfg.lin_dx0 = dx0
    cfg.lin_dx1 = dx0 * ratio
    cfg.lin_band = band
    cfg.seed = seed
    return cfg, sdf_ellipsoid_union(seed=seed)


def config_c1(h_over_dx=1.0, num_pkg=32, dx=0.01, seed=20261203):
    """2D: [0,1]^2 minus a circle r = 0.2 at the centre, uniform dx, lset 128^2 cells."""
    cfg = default_config(2)
    cfg.num_pkg[0] = cfg.num_pkg[1] = num_pkg
    cfg.cell_ratio = 2.0 * h_over_dx
    cfg.cut_off = 2.0
    cfg.min_dl = cfg.max_dl = dx
    cfg.seed = seed
    sdf = sdf_circle_hole((0.5, 0.5), 0.2)
    return cfg, sdf


def config_c2(scale=0.564, num_pkg=64, seed=20261203):
    """3D: unit cube minus a sphere r = 0.3 at the centre; dx linear in phi from
    0.008*scale at the surface to 0.024*scale at phi >= 0.2; lset (4*num_pkg)^3."""
    cfg = default_config(3)
    for a in range(3):
        cfg.num_pkg[a] = num_pkg
    cfg.profile = PROFILE_LINEAR
    cfg.lin_dx0 = 0.008 * scale
    cfg.lin_dx1 = 0.024 * scale
    cfg.lin_band = 0.2
    cfg.seed = seed
    sdf = sdf_circle_hole((0.5, 0.5, 0.5), 0.3)
    return cfg, sdf


# ---------------------------------------------------------------------------
# jobs

class Job:
    """One relaxation job on the GPU (SPH_mesh_generation driven through the C ABI)."""

    def __init__(self, cfg, lset, profile=False, levels=None):
        self.cfg = config_copy(cfg)
        self.D = cfg.dim
        h = C.c_void_p()
        _check(lib().smg_create(C.byref(self.cfg), C.byref(h)), "smg_create")
        self.h = h
        self.lset = lset
        _check(lib().smg_set_levelset(self.h, _ptr(lset["phi"]), _ptr(lset["normal"]), _ptr(lset["tag"])),
               "smg_set_levelset")
        self.levels = levels or []
        if self.levels:
            arr = (SmgLsetLevel * len(self.levels))()
            for i, L in enumerate(self.levels):
                arr[i] = SmgLsetLevel(_ptr(L["phi"]), _ptr(L["normal"]), _ptr(L["tag"]), _ptr(L["active"]))
            _check(lib().smg_set_levelset_levels(self.h, len(self.levels), arr), "smg_set_levelset_levels")
        if profile:
            _check(lib().smg_profile_enable(self.h, 1), "smg_profile_enable")
        self.n = 0

    def close(self):
        if self.h:
            lib().smg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def seed(self, n):
        self.n = int(n)
        _check(lib().smg_seed_particles(self.h, self.n), "smg_seed_particles")

    def init_particles(self, x, types=None):
        x3 = np.zeros((len(x), 3), np.float64)
        x3[:, :x.shape[1]] = x
        self.n = len(x)
        tp = None if types is None else np.ascontiguousarray(types, np.int32)
        _check(lib().smg_init_particles(self.h, self.n, _ptr(x3), None if tp is None else _ptr(tp)),
               "smg_init_particles")

    def prerun(self):
        _check(lib().smg_prerun(self.h), "smg_prerun")

    def step(self, n=1):
        _check(lib().smg_step(self.h, int(n)), "smg_step")

    def step_async(self, n):
        _check(lib().smg_step_async(self.h, int(n)), "smg_step_async")

    def busy(self):
        b = C.c_int32()
        _check(lib().smg_query(self.h, C.byref(b)), "smg_query")
        return bool(b.value)

    def sync(self):
        _check(lib().smg_sync(self.h), "smg_sync")

    def capture(self, steps_per_graph):
        _check(lib().smg_capture(self.h, int(steps_per_graph)), "smg_capture")

    def footprint(self):
        b = C.c_int64()
        _check(lib().smg_footprint(self.h, C.byref(b)), "smg_footprint")
        return int(b.value)

    def state(self):
        n = self.n
        x = np.zeros((n, 3)); v = np.zeros((n, 3))
        t = np.zeros(n, np.int32); hh = np.zeros(n); nu = np.zeros(n, np.float32)
        _check(lib().smg_get_state(self.h, _ptr(x), _ptr(v), _ptr(t), _ptr(hh), _ptr(nu)), "smg_get_state")
        return {"x": x, "v": v, "type": t, "h": hh, "nu": nu}

    def pair_terms(self):
        n = self.n
        s = np.zeros(n, np.float32); a1 = np.zeros((n, 3), np.float32); a2 = np.zeros((n, 3), np.float32)
        _check(lib().smg_get_pair_terms(self.h, _ptr(s), _ptr(a1), _ptr(a2)), "smg_get_pair_terms")
        return {"sum": s, "a1": a1, "a2": a2}

    def neighbors(self):
        offs = np.zeros(self.n + 1, np.int64)
        _check(lib().smg_get_neighbors(self.h, _ptr(offs), None, None, 0), "smg_get_neighbors")
        m = int(offs[-1])
        ids = np.zeros(max(m, 1), np.int32); codes = np.zeros(max(m, 1), np.int32)
        _check(lib().smg_get_neighbors(self.h, _ptr(offs), _ptr(ids), _ptr(codes), m), "smg_get_neighbors")
        return offs, ids[:m], codes[:m]

    def normals(self):
        """The particles' interpolated level-set normals (n x 3, float32, by id)."""
        out = np.zeros((self.n, 3), np.float32)
        _check(lib().smg_get_normals(self.h, _ptr(out)), "smg_get_normals")
        return out

    def sample_phi(self, pts):
        """Level-set phi at points (n x 3) as the relaxation samples it."""
        p = np.zeros((len(pts), 3))
        p[:, :np.asarray(pts).shape[1]] = pts
        out = np.zeros(len(p))
        _check(lib().smg_sample_phi(self.h, len(p), _ptr(p), _ptr(out)), "smg_sample_phi")
        return out

    def stats(self):
        st = SmgStats()
        _check(lib().smg_get_stats(self.h, C.byref(st)), "smg_get_stats")
        return st.as_dict()

    def profile(self):
        names = (C.c_char_p * 64)(); ms = (C.c_double * 64)(); cnt = (C.c_int64 * 64)(); nk = C.c_int()
        _check(lib().smg_profile_get(self.h, names, ms, cnt, 64, C.byref(nk)), "smg_profile_get")
        return {names[i].decode(): (ms[i], cnt[i]) for i in range(nk.value)}


def fp64_probe(a, b):
    a = np.ascontiguousarray(a, np.float64); b = np.ascontiguousarray(b, np.float64)
    out = np.zeros((len(a), 4))
    _check(lib().smg_fp64_probe(len(a), _ptr(a), _ptr(b), _ptr(out)), "smg_fp64_probe")
    return out


# ---------------------------------------------------------------------------
# batched jobs

def _batch_sigs(L):
    P = C.POINTER
    L.smg_batch_create.argtypes = [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                   P(C.c_void_p)]
    L.smg_batch_destroy.argtypes = [C.c_void_p]
    for f in ("smg_batch_seed", "smg_batch_sync"):
        getattr(L, f).argtypes = [C.c_void_p]
    for f in ("smg_batch_step", "smg_batch_capture", "smg_batch_run"):
        getattr(L, f).argtypes = [C.c_void_p, C.c_int64]
    L.smg_batch_get_stats.argtypes = [C.c_void_p, C.c_int32, P(SmgStats)]
    L.smg_batch_get_state.argtypes = [C.c_void_p, C.c_int32] + [C.c_void_p] * 5


class Batch:
    """Many independent jobs relaxed together (smg_batch_* of include/smg.h)."""

    def __init__(self, jobs):
        """jobs: list of (cfg, lset, n)."""
        L = lib()
        _batch_sigs(L)
        nj = len(jobs)
        cfgs = (SmgConfig * nj)()
        for i, (c, _, _) in enumerate(jobs):
            C.memmove(C.byref(cfgs[i]), C.byref(c), C.sizeof(SmgConfig))
        ns = np.array([int(n) for _, _, n in jobs], np.int64)
        self._keep = [ls for _, ls, _ in jobs]
        phis = (C.c_void_p * nj)(*[_ptr(ls["phi"]).value for ls in self._keep])
        nrms = (C.c_void_p * nj)(*[_ptr(ls["normal"]).value for ls in self._keep])
        tags = (C.c_void_p * nj)(*[_ptr(ls["tag"]).value for ls in self._keep])
        h = C.c_void_p()
        _check(L.smg_batch_create(nj, C.cast(cfgs, C.c_void_p), _ptr(ns), C.cast(phis, C.c_void_p),
                                  C.cast(nrms, C.c_void_p), C.cast(tags, C.c_void_p), C.byref(h)), "smg_batch_create")
        self.h = h
        self.ns = ns
        self.dims = [c.dim for c, _, _ in jobs]

    def close(self):
        if self.h:
            lib().smg_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def seed(self):
        _check(lib().smg_batch_seed(self.h), "smg_batch_seed")

    def step(self, n):
        _check(lib().smg_batch_step(self.h, int(n)), "smg_batch_step")

    def capture(self, steps_per_graph):
        _check(lib().smg_batch_capture(self.h, int(steps_per_graph)), "smg_batch_capture")

    def run(self, launches):
        _check(lib().smg_batch_run(self.h, int(launches)), "smg_batch_run")

    def sync(self):
        _check(lib().smg_batch_sync(self.h), "smg_batch_sync")

    def stats(self, j):
        st = SmgStats()
        _check(lib().smg_batch_get_stats(self.h, int(j), C.byref(st)), "smg_batch_get_stats")
        return st.as_dict()

    def state(self, j):
        n = int(self.ns[j])
        x = np.zeros((n, 3)); v = np.zeros((n, 3))
        t = np.zeros(n, np.int32); hh = np.zeros(n); nu = np.zeros(n, np.float32)
        _check(lib().smg_batch_get_state(self.h, int(j), _ptr(x), _ptr(v), _ptr(t), _ptr(hh), _ptr(nu)),
               "smg_batch_get_state")
        return {"x": x, "v": v, "type": t, "h": hh, "nu": nu}


# ---------------------------------------------------------------------------
# serving: continuous batching over a device-resident pool (smg_server_*)

class SmgTriStats(C.Structure):
    _fields_ = [
        ("n_tri", C.c_int64), ("n_edge", C.c_int64), ("n_deleted_edge", C.c_int64), ("n_cell_overflow", C.c_int64),
        ("min_angle", C.c_double), ("max_angle", C.c_double), ("avg_min_angle", C.c_double),
        ("min_edge", C.c_double), ("max_edge", C.c_double),
        ("min_area", C.c_double), ("max_area", C.c_double), ("total_area", C.c_double),
        ("max_aspect", C.c_double), ("G_min", C.c_double), ("G_avg", C.c_double),
        ("n_min_angle_below", C.c_int64 * 4), ("angle_hist", C.c_int32 * 45), ("reserved", C.c_int32),
    ]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}
        d["n_min_angle_below"] = list(self.n_min_angle_below)
        d["angle_hist"] = list(self.angle_hist)
        return d


def _tri_sigs(L):
    P = C.POINTER
    L.smg_triangulate.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, P(C.c_int64), P(SmgTriStats)]
    L.smg_tri_cliques.argtypes = [C.c_int64, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, P(C.c_int64)]
    L.smg_tri_edges.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, P(C.c_int64)]


def tri_edges(job):
    """Voronoi-stage edges (i > j) of the job's last triangulate() call."""
    L = lib(); _tri_sigs(L)
    m = C.c_int64()
    _check(L.smg_tri_edges(job.h, None, 0, C.byref(m)), "smg_tri_edges")
    e = np.zeros((max(m.value, 1), 2), np.int32)
    _check(L.smg_tri_edges(job.h, _ptr(e), m.value, C.byref(m)), "smg_tri_edges")
    return e[:m.value]


def triangulate(job):
    """(tris (m x 3 particle ids, i > j > k), stats) of a 2D job's current particles."""
    L = lib(); _tri_sigs(L)
    m = C.c_int64()
    st = SmgTriStats()
    _check(L.smg_triangulate(job.h, None, 0, C.byref(m), C.byref(st)), "smg_triangulate")
    tris = np.zeros((max(m.value, 1), 3), np.int32)
    _check(L.smg_triangulate(job.h, _ptr(tris), m.value, C.byref(m), C.byref(st)), "smg_triangulate")
    return tris[:m.value], st.as_dict()


def tri_cliques(n_vertices, edges):
    """All 3-cliques (i > j > k) of an undirected graph (host-only integer stage)."""
    L = lib(); _tri_sigs(L)
    e = np.ascontiguousarray(edges, np.int32).reshape(-1, 2)
    m = C.c_int64()
    _check(L.smg_tri_cliques(int(n_vertices), _ptr(e), len(e), None, 0, C.byref(m)), "smg_tri_cliques")
    out = np.zeros((max(m.value, 1), 3), np.int32)
    _check(L.smg_tri_cliques(int(n_vertices), _ptr(e), len(e), _ptr(out), m.value, C.byref(m)), "smg_tri_cliques")
    return out[:m.value]


class SmgTetStats(C.Structure):
    _fields_ = [
        ("n_tet", C.c_int64), ("n_delaunay", C.c_int64), ("n_deleted", C.c_int64), ("n_surface_tri", C.c_int64),
        ("n_open", C.c_int64), ("n_cell_overflow", C.c_int64), ("n_duplicate", C.c_int64),
        ("min_edge", C.c_double), ("max_edge", C.c_double),
        ("min_volume", C.c_double), ("max_volume", C.c_double), ("total_volume", C.c_double),
        ("min_aspect", C.c_double), ("max_aspect", C.c_double), ("avg_aspect", C.c_double),
        ("min_radius_ratio", C.c_double), ("max_radius_ratio", C.c_double), ("avg_radius_ratio", C.c_double),
        ("min_dihedral", C.c_double), ("max_dihedral", C.c_double), ("avg_min_dihedral", C.c_double),
        ("n_min_dihedral_below", C.c_int64 * 4), ("dihedral_hist", C.c_int64 * 45),
        ("aspect_hist", C.c_int64 * 12), ("radius_hist", C.c_int64 * 56),
        ("cell_ms", C.c_double), ("total_ms", C.c_double), ("n_second_pass", C.c_int64),
    ]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        for k in ("n_min_dihedral_below", "dihedral_hist", "aspect_hist", "radius_hist"):
            d[k] = list(d[k])
        return d


def _tet_sigs(L):
    P = C.POINTER
    L.smg_tetrahedralize.argtypes = [C.c_void_p, C.c_double, C.c_void_p, C.c_int64, P(C.c_int64), P(SmgTetStats)]
    L.smg_tet_surface.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, P(C.c_int64)]
    L.smg_tet_delaunay.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, P(C.c_int64), C.c_void_p, C.c_void_p]
    L.smg_tet_finish.argtypes = [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_double,
                                 C.c_void_p, C.c_int64, P(C.c_int64), C.c_void_p, C.c_int64, P(C.c_int64),
                                 P(SmgTetStats)]


def tetrahedralize(job, delete_thres=0.0):
    """(tets (m x 4 ascending particle ids, sorted), stats) of a 3D job's current particles."""
    L = lib(); _tet_sigs(L)
    m = C.c_int64()
    st = SmgTetStats()
    # a Delaunay tetrahedralisation has about 6.5 tetrahedra per point
    tets = np.zeros((max(8 * job.n, 16), 4), np.int32)
    rc = L.smg_tetrahedralize(job.h, float(delete_thres), _ptr(tets), len(tets), C.byref(m), C.byref(st))
    if rc == -3 and m.value > len(tets):   # SMG_ERR_CAPACITY: computed, not copied
        tets = np.zeros((m.value, 4), np.int32)
        rc = L.smg_tetrahedralize(job.h, float(delete_thres), _ptr(tets), len(tets), C.byref(m), C.byref(st))
    _check(rc, "smg_tetrahedralize")
    return tets[:m.value], st.as_dict()


def tet_surface(job):
    """Boundary triangles (ascending ids, sorted) of the job's last tetrahedralize()."""
    L = lib(); _tet_sigs(L)
    m = C.c_int64()
    _check(L.smg_tet_surface(job.h, None, 0, C.byref(m)), "smg_tet_surface")
    t = np.zeros((max(m.value, 1), 3), np.int32)
    _check(L.smg_tet_surface(job.h, _ptr(t), m.value, C.byref(m)), "smg_tet_surface")
    return t[:m.value]


def tet_delaunay(job, n):
    """(certified Delaunay tets, per-particle cell status, certified radius) of the job's
    last tetrahedralize(); n = the job's particle count."""
    L = lib(); _tet_sigs(L)
    m = C.c_int64()
    _check(L.smg_tet_delaunay(job.h, None, 0, C.byref(m), None, None), "smg_tet_delaunay")
    t = np.zeros((max(m.value, 1), 4), np.int32)
    status = np.zeros(n, np.int8)
    rcert = np.zeros(n)
    _check(L.smg_tet_delaunay(job.h, _ptr(t), m.value, C.byref(m), _ptr(status), _ptr(rcert)), "smg_tet_delaunay")
    return t[:m.value], status, rcert


def tet_finish(x, types, tets, delete_thres=0.0):
    """Host stage alone: (kept tets, boundary triangles, stats) of given tetrahedra."""
    L = lib(); _tet_sigs(L)
    x = np.ascontiguousarray(x, np.float64).reshape(-1, 3)
    ty = np.ascontiguousarray(types, np.int32)
    t = np.ascontiguousarray(tets, np.int32).reshape(-1, 4)
    kept = np.zeros((max(len(t), 1), 4), np.int32)
    tris = np.zeros((max(4 * len(t), 1), 3), np.int32)
    nk, nt = C.c_int64(), C.c_int64()
    st = SmgTetStats()
    _check(L.smg_tet_finish(len(x), _ptr(x), _ptr(ty), _ptr(t), len(t), float(delete_thres), _ptr(kept), len(kept),
                            C.byref(nk), _ptr(tris), len(tris), C.byref(nt), C.byref(st)), "smg_tet_finish")
    return kept[:nk.value], tris[:nt.value], st.as_dict()


class SmgTuneResult(C.Structure):
    _fields_ = [("best_skin", C.c_double), ("best_max_neighbors", C.c_int32), ("n_tried", C.c_int32),
                ("best_ms_per_step", C.c_double), ("skin", C.c_double * 16), ("max_neighbors", C.c_int32 * 16),
                ("ms_per_step", C.c_double * 16)]


def autotune(cfg, lset, n, warmup=50, steps=50):
    """Fastest (skin, max_neighbors) for this job, timed on scratch copies (smg_autotune)."""
    L = lib()
    L.smg_autotune.argtypes = [C.POINTER(SmgConfig), C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_int32,
                               C.c_int32, C.POINTER(SmgTuneResult)]
    r = SmgTuneResult()
    _check(L.smg_autotune(C.byref(cfg), _ptr(lset["phi"]), _ptr(lset["normal"]), _ptr(lset["tag"]), int(n),
                          int(warmup), int(steps), C.byref(r)), "smg_autotune")
    k = r.n_tried
    return {"skin": r.best_skin, "max_neighbors": r.best_max_neighbors, "ms_per_step": r.best_ms_per_step,
            "tried": [(r.skin[i], r.max_neighbors[i], r.ms_per_step[i]) for i in range(k)]}


class SmgServerConfig(C.Structure):
    _fields_ = [
        ("dim", C.c_int32), ("max_jobs", C.c_int32), ("max_slots", C.c_int64), ("max_table", C.c_int64),
        ("max_lset_cells", C.c_int64), ("max_neighbors", C.c_int32), ("spill_per_slot", C.c_int32),
        ("chunk_steps", C.c_int32), ("min_steps", C.c_int32), ("max_steps", C.c_int32), ("reserved", C.c_int32),
        ("tol", C.c_double), ("max_active_particles", C.c_int64), ("max_lset_maps", C.c_int64),
    ]


class SmgServerJob(C.Structure):
    _fields_ = [("ticket", C.c_int32), ("status", C.c_int32), ("steps", C.c_int64), ("n_particle", C.c_int64),
                ("window_residual", C.c_double)]


JOB_CONVERGED, JOB_MAX_STEPS = 0, 1


def _server_sigs(L):
    P = C.POINTER
    L.smg_server_footprint_of.argtypes = [P(SmgServerConfig), P(C.c_int64)]
    L.smg_server_job_size.argtypes = [P(SmgConfig), C.c_int64, P(C.c_int64), P(C.c_int64), P(C.c_int64)]
    L.smg_server_create.argtypes = [P(SmgServerConfig), P(C.c_void_p)]
    L.smg_server_destroy.argtypes = [C.c_void_p]
    L.smg_server_admit.argtypes = [C.c_void_p, P(SmgConfig), C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64,
                                   P(C.c_int32)]
    L.smg_server_admit_levels.argtypes = [C.c_void_p, P(SmgConfig), C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32,
                                          P(SmgLsetLevel), C.c_int64, P(C.c_int32)]
    L.smg_server_advance.argtypes = [C.c_void_p, P(C.c_int32)]
    L.smg_server_finished.argtypes = [C.c_void_p, P(SmgServerJob), C.c_int32, P(C.c_int32)]
    L.smg_server_result.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, P(SmgStats)]
    L.smg_server_usage.argtypes = [C.c_void_p, P(C.c_int64), P(C.c_int64), P(C.c_int32)]


def server_config(dim, max_jobs, max_slots, max_table, max_lset_cells, chunk_steps=50, min_steps=100,
                  max_steps=3000, tol=1e-3, max_neighbors=0, spill_per_slot=0, max_active_particles=0,
                  max_lset_maps=0):
    return SmgServerConfig(dim, max_jobs, int(max_slots), int(max_table), int(max_lset_cells), max_neighbors,
                           spill_per_slot, chunk_steps, min_steps, max_steps, 0, tol, int(max_active_particles),
                           int(max_lset_maps))


def server_footprint(scfg):
    L = lib(); _server_sigs(L)
    b = C.c_int64()
    _check(L.smg_server_footprint_of(C.byref(scfg), C.byref(b)), "smg_server_footprint_of")
    return b.value


def job_size(cfg, n):
    """(slots, table entries, level-set cells) one job takes in a server pool."""
    L = lib(); _server_sigs(L)
    a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
    _check(L.smg_server_job_size(C.byref(cfg), int(n), C.byref(a), C.byref(b), C.byref(c)), "smg_server_job_size")
    return a.value, b.value, c.value


class Server:
    """Continuous batching: admit jobs into a device pool between chunks of steps,
    step all of them together, retire the converged ones (smg_server_* of include/smg.h)."""

    def __init__(self, scfg):
        L = lib(); _server_sigs(L)
        h = C.c_void_p()
        _check(L.smg_server_create(C.byref(scfg), C.byref(h)), "smg_server_create")
        self.h = h
        self.cfg = scfg
        self.n_of = {}

    def close(self):
        if self.h:
            lib().smg_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def admit(self, cfg, lset, n, levels=None):
        """Ticket of the admitted job, or None when it does not fit the pool now.
        levels: refined level-set levels as for Job (make_levelset_levels)."""
        t = C.c_int32()
        if levels:
            arr = (SmgLsetLevel * len(levels))()
            for i, L in enumerate(levels):
                arr[i] = SmgLsetLevel(_ptr(L["phi"]), _ptr(L["normal"]), _ptr(L["tag"]), _ptr(L["active"]))
            rc = lib().smg_server_admit_levels(self.h, C.byref(cfg), _ptr(lset["phi"]), _ptr(lset["normal"]),
                                               _ptr(lset["tag"]), len(levels), arr, int(n), C.byref(t))
        else:
            rc = lib().smg_server_admit(self.h, C.byref(cfg), _ptr(lset["phi"]), _ptr(lset["normal"]),
                                        _ptr(lset["tag"]), int(n), C.byref(t))
        if rc == SMG_ERR_CAPACITY:
            return None
        _check(rc, "smg_server_admit")
        self.n_of[t.value] = int(n)
        return t.value

    def advance(self):
        na = C.c_int32()
        _check(lib().smg_server_advance(self.h, C.byref(na)), "smg_server_advance")
        return na.value

    def finished(self):
        buf = (SmgServerJob * 256)()
        out = []
        while True:
            k = C.c_int32()
            _check(lib().smg_server_finished(self.h, buf, 256, C.byref(k)), "smg_server_finished")
            out += [{f: getattr(buf[i], f) for f, _ in SmgServerJob._fields_} for i in range(k.value)]
            if k.value < 256:
                return out

    def result(self, ticket):
        n = self.n_of.pop(ticket)
        x = np.zeros((n, 3)); t = np.zeros(n, np.int32); hh = np.zeros(n)
        st = SmgStats()
        _check(lib().smg_server_result(self.h, int(ticket), _ptr(x), _ptr(t), _ptr(hh), C.byref(st)),
               "smg_server_result")
        return {"x": x, "type": t, "h": hh, "stats": st.as_dict()}

    def usage(self):
        u, tot, na = C.c_int64(), C.c_int64(), C.c_int32()
        _check(lib().smg_server_usage(self.h, C.byref(u), C.byref(tot), C.byref(na)), "smg_server_usage")
        return {"used_slots": u.value, "total_slots": tot.value, "n_active": na.value}


def union_job(dim, n_target, rng, seed, max_pkg=None):
    """One batched-config job: a union of 1-5 circles / spheres (radii U[0.05, 0.25],
    centres U[r, 1-r]^dim) in the unit box, the fluid inside the union, uniform dx
    chosen so that the particle count (Get_num_particles) is about n_target; the level
    set covers the unit box and one package around it, the particle box the union."""
    k = int(rng.integers(1, 6))
    radii = rng.uniform(0.05, 0.25, size=k)
    centres = np.stack([rng.uniform(r, 1 - r, size=dim) for r in radii])
    sdf = sdf_sphere_union(centres, radii, dim)
    cfg = default_config(dim)
    cfg.seed = seed
    # volume and boundary measure of the union by sampling, then N(dx) = V/dx^D + A/(2 dx^(D-1))
    pts = rng.uniform(0, 1, size=(200000, 3))
    phi = sdf(pts[:, 0], pts[:, 1], pts[:, 2])
    vol = max(float(np.mean(phi > 0)), 1e-4)
    eps = 0.01
    area = float(np.mean(np.abs(phi) < eps)) / (2 * eps)
    lo, hi = 1e-5, 1.0
    for _ in range(80):
        mid = 0.5 * (lo + hi)
        if vol / mid ** dim + 0.5 * area / mid ** (dim - 1) > n_target:
            lo = mid
        else:
            hi = mid
    dx = 0.5 * (lo + hi)
    if max_pkg is None:
        max_pkg = 64 if dim == 2 else 16
    npk = int(np.clip(np.ceil(1.0 / (6.0 * dx)), 6, max_pkg))
    # The level-set box reaches one package beyond the unit box on every side: the
    # seeding clamps to the level-set box shrunk by one package
    # (sph_mesh_generation.cpp:1327-1342), which is then the unit box that holds the
    # union, so no seeds pile up on the clamp box's faces, edges and corners.
    dpkg = 1.0 / npk
    for a in range(dim):
        cfg.num_pkg[a] = npk + 2
        cfg.box_l[a] = -dpkg
        cfg.domain[a] = 1.0 + 2.0 * dpkg
        # seeds are drawn from the union's bounding box (one dx wider)
        cfg.particle_box_l[a] = max(0.0, float(np.min(centres[:, a] - radii)) - dx)
        cfg.particle_box_r[a] = min(1.0, float(np.max(centres[:, a] + radii)) + dx)
    cfg.min_dl = cfg.max_dl = dx
    ls = make_levelset(cfg, sdf)
    return cfg, ls, int(ls["stats"].num_particle)


def batched_jobs(njobs, n_min, n_max, seed=20261203, dims=(2, 3), max_pkg=None):
    rng = np.random.default_rng(seed)
    jobs = []
    for j in range(njobs):
        dim = dims[j % len(dims)]
        n_t = float(np.exp(rng.uniform(np.log(n_min), np.log(n_max))))
        jobs.append(union_job(dim, n_t, rng, seed + 7919 * (j + 1), max_pkg=max_pkg))
    return jobs


# ---- output files (host only) --------------------------------------------------------

class SmgVtuPoints(C.Structure):
    _fields_ = [("n", C.c_int64), ("x", C.c_void_p), ("type", C.c_void_p), ("h", C.c_void_p), ("nn", C.c_void_p),
                ("color", C.c_void_p), ("normal", C.c_void_p)]


def _vtu_points(x, type=None, h=None, nn=None, color=None, normal=None):
    keep = []

    def arr(a, dt, per):
        if a is None:
            return None
        a = np.ascontiguousarray(a, dt).reshape(-1)
        if a.size != per * len(x):
            raise ValueError("array length does not match the points")
        keep.append(a)
        return _ptr(a)
    x = np.ascontiguousarray(x, np.float64)
    if x.ndim != 2 or x.shape[1] not in (2, 3):
        raise ValueError("x must be n x 2 or n x 3")
    if x.shape[1] == 2:
        x = np.hstack([x, np.zeros((len(x), 1))])
    keep.append(x)
    p = SmgVtuPoints(len(x), _ptr(x), arr(type, np.int32, 1), arr(h, np.float64, 1), arr(nn, np.int32, 1),
                     arr(color, np.int32, 1), arr(normal, np.float32, 3))
    return p, keep


def write_particles_vtu(dir, n_out, time, x, type=None, h=None, nn=None, color=None, normal=None):
    """Particle_binary_writer::Output_particle: particle_rank_0_step_<n_out>.vtu, .pvtu, particle.pvd."""
    L = lib()
    L.smg_write_particles_vtu.argtypes = [C.c_char_p, C.c_int, C.c_double, C.POINTER(SmgVtuPoints)]
    p, keep = _vtu_points(x, type, h, nn, color, normal)
    _check(L.smg_write_particles_vtu(str(dir).encode(), int(n_out), float(time), C.byref(p)), "smg_write_particles_vtu")


def write_mesh_vtu(dir, n_out, time, dim, x, cells, type=None, h=None, color=None):
    """Mesh_binary_writer::Output_mesh: mesh_tris_* (dim 2) or mesh_tets_* (dim 3) .vtu, .pvtu, .pvd."""
    L = lib()
    L.smg_write_mesh_vtu.argtypes = [C.c_char_p, C.c_int, C.c_double, C.c_int, C.POINTER(SmgVtuPoints), C.c_void_p,
                                     C.c_int64]
    p, keep = _vtu_points(x, type, h, None, color, None)
    c = np.ascontiguousarray(cells, np.int32).reshape(-1, 3 if dim == 2 else 4)
    _check(L.smg_write_mesh_vtu(str(dir).encode(), int(n_out), float(time), int(dim), C.byref(p), _ptr(c), len(c)),
           "smg_write_mesh_vtu")


def write_job_output(job, dir, n_out, time=None, mesh=None):
    """The reference's per-output files for a job: particles (positions, type, h, neighbour
    counts, normals) and, when mesh = (dim, cells) is given, the mesh."""
    st = job.state()
    offs, _, _ = job.neighbors()
    t = float(job.stats()["step"]) if time is None else time
    write_particles_vtu(dir, n_out, t, st["x"], st["type"], st["h"], np.diff(offs).astype(np.int32), None,
                        job.normals())
    if mesh is not None:
        write_mesh_vtu(dir, n_out, t, mesh[0], st["x"], mesh[1], st["type"], st["h"])


def write_post_file(dir, step, x, first):
    """Save_particle_post_file: step_<step>_rank_0.pst and a list.csv line."""
    L = lib()
    L.smg_write_post_file.argtypes = [C.c_char_p, C.c_int64, C.c_int64, C.c_void_p, C.c_int]
    x3 = np.zeros((len(x), 3))
    x3[:, :np.asarray(x).shape[1]] = x
    _check(L.smg_write_post_file(str(dir).encode(), int(step), len(x3), _ptr(x3), int(bool(first))),
           "smg_write_post_file")


def read_post_list(dir):
    """(number of ranks, [(step, count)]) from list.csv."""
    L = lib()
    L.smg_read_post_list.argtypes = [C.c_char_p, C.POINTER(C.c_int32), C.c_void_p, C.c_void_p, C.c_int64,
                                     C.POINTER(C.c_int64)]
    r, k = C.c_int32(), C.c_int64()
    _check(L.smg_read_post_list(str(dir).encode(), C.byref(r), None, None, 0, C.byref(k)), "smg_read_post_list")
    steps = np.zeros(max(k.value, 1), np.int64)
    counts = np.zeros(max(k.value, 1), np.int64)
    _check(L.smg_read_post_list(str(dir).encode(), C.byref(r), _ptr(steps), _ptr(counts), len(steps), C.byref(k)),
           "smg_read_post_list")
    return r.value, list(zip(steps[:k.value].tolist(), counts[:k.value].tolist()))


def read_post_file(dir, step, n_ranks):
    """(positions n x 3, rank colour per particle) of one saved step, all ranks."""
    L = lib()
    L.smg_read_post_file.argtypes = [C.c_char_p, C.c_int64, C.c_int32, C.c_void_p, C.c_void_p, C.c_int64,
                                     C.POINTER(C.c_int64)]
    n = C.c_int64()
    _check(L.smg_read_post_file(str(dir).encode(), int(step), int(n_ranks), None, None, 0, C.byref(n)),
           "smg_read_post_file")
    x = np.zeros((max(n.value, 1), 3))
    color = np.zeros(max(n.value, 1), np.int32)
    _check(L.smg_read_post_file(str(dir).encode(), int(step), int(n_ranks), _ptr(x), _ptr(color), n.value,
                                C.byref(n)), "smg_read_post_file")
    return x[:n.value], color[:n.value]


def read_vtu(path):
    """A VTK XML file with raw appended data as written here: {array name: values}
    (shaped n x components) plus "_points" / "_cells" counts."""
    import re
    raw = open(path, "rb").read()
    head, _, rest = raw.partition(b"<AppendedData encoding=\"raw\">")
    data = rest[rest.index(b"_") + 1:]
    text = head.decode()
    m = re.search(r'NumberOfPoints="(\d+)" NumberOfCells="(\d+)"', text)
    out = {"_points": int(m.group(1)), "_cells": int(m.group(2))}
    types = {"Int32": np.int32, "Float32": np.float32, "UInt8": np.uint8}
    for a in re.finditer(r'<DataArray type="(\w+)" Name="(\w+)"(?: NumberOfComponents="(\d+)")? '
                         r'format="appended" offset="(\d+)"/>', text):
        dt, name, comps, off = types[a.group(1)], a.group(2), int(a.group(3) or 1), int(a.group(4))
        nbytes = int(np.frombuffer(data[off:off + 4], np.uint32)[0])
        v = np.frombuffer(data[off + 4:off + 4 + nbytes], dt)
        out[name] = v.reshape(-1, comps) if comps > 1 else v
    return out


def post_process(dir, cfg, lset, outdir=None):
    """SPH_mesh_generation::Load_data_for_post (sph_mesh_generation.cpp:426-590): for every
    step in dir/list.csv, the particles of all its rank files are placed on the level set
    (Get_particle_info, Set_particle_scale and Find_and_map_particle, here smg_prerun)
    and meshed at the placed positions: triangles in 2D (where the reference's post mode
    leaves the mesh empty; the lists are built on the placed positions, no step is
    taken), tetrahedra after the deletion rule in 3D; with outdir, each step's mesh goes
    to a mesh VTU (output n_post, time = the step).  Returns [(step, cells)]."""
    ranks, steps = read_post_list(dir)
    out = []
    for n_post, (step, count) in enumerate(steps):
        x, color = read_post_file(dir, step, ranks)
        if len(x) != count:
            raise ValueError(f"step {step}: list.csv says {count} particles, the files hold {len(x)}")
        job = Job(cfg, lset)
        try:
            job.init_particles(x[:, :cfg.dim], np.zeros(len(x), np.int32))
            job.prerun()
            if cfg.dim == 2:
                cells = triangulate(job)[0]
            else:
                cells = tetrahedralize(job)[0]
            if outdir is not None:
                st = job.state()
                write_mesh_vtu(outdir, n_post, float(step), cfg.dim, st["x"], cells, st["type"], st["h"], color)
        finally:
            job.close()
        out.append((step, cells))
    return out
