"""Test infrastructure only: ctypes bindings to the oracle (liboracle.so).

The oracle is a CPU restatement of the DVMpp reference time-step path
(oracle/dvm_oracle.cpp).  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module, and only as the checker.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

_dp = ctypes.POINTER(ctypes.c_double)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_i64 = ctypes.c_int64


class Params(ctypes.Structure):
    """orc_params (oracle/dvm_oracle.h); XML options of src/XmlHandler.cpp:162-202."""
    _fields_ = [
        ("rho", ctypes.c_double), ("nu", ctypes.c_double), ("dt", ctypes.c_double),
        ("ux", ctypes.c_double), ("uz", ctypes.c_double), ("cutoff_exp", ctypes.c_double),
        ("max_npv", ctypes.c_int), ("scheme", ctypes.c_int), ("crossing", ctypes.c_int),
        ("seed", ctypes.c_int), ("remesh_frequency", ctypes.c_int), ("remesh_kernel", ctypes.c_int),
        ("remesh_xmin", ctypes.c_double), ("remesh_xmax", ctypes.c_double),
        ("remesh_zmin", ctypes.c_double), ("remesh_zmax", ctypes.c_double),
        ("remesh_dx", ctypes.c_double), ("remesh_dz", ctypes.c_double),
        ("remesh_sigma", ctypes.c_double), ("remesh_circ_threshold", ctypes.c_double),
        ("merge_threshold", ctypes.c_double), ("merge_frequency", ctypes.c_int),
        ("rng_mode", ctypes.c_int),
    ]


def make_params(**kw):
    p = Params()
    defaults = dict(rho=1.0, nu=1e-3, dt=0.005, ux=1.0, uz=0.0, cutoff_exp=0.675, max_npv=3,
                    scheme=0, crossing=2, seed=0, remesh_frequency=0, remesh_kernel=0,
                    remesh_xmin=-5.0, remesh_xmax=25.0, remesh_zmin=-10.0, remesh_zmax=10.0,
                    remesh_dx=0.01, remesh_dz=0.01, remesh_sigma=0.05, remesh_circ_threshold=1e-10,
                    merge_threshold=0.0, merge_frequency=0, rng_mode=0)
    defaults.update(kw)
    for k, v in defaults.items():
        setattr(p, k, v)
    return p


def _load(path=None):
    path = path or _LIB_PATH
    if not os.path.exists(path):
        raise RuntimeError(f"oracle library missing: {path} (run `make -C oracle`)")
    lib = ctypes.CDLL(path)
    vp = ctypes.c_void_p
    sig = {
        "orc_rng_create": (vp, [ctypes.c_int]),
        "orc_rng_destroy": (None, [vp]),
        "orc_rng_rand": (ctypes.c_double, [vp]),
        "orc_rng_fill": (None, [vp, _i64, _dp]),
        "orc_rng_raw": (ctypes.c_uint32, [vp]),
        "orc_num_threads": (ctypes.c_int, []),
        "orc_bs_self": (None, [_i64, _dp, _dp, _dp, _dp, _dp, _dp]),
        "orc_bs_self_ld": (None, [_i64, _dp, _dp, _dp, _dp, _dp, _dp]),
        "orc_bs_targets": (None, [_i64, _dp, _dp, _dp, _dp, _i64, _dp, _dp, _dp, _dp]),
        "orc_sheet_create": (vp, [ctypes.c_int, _dp, _dp, ctypes.POINTER(Params)]),
        "orc_sheet_destroy": (None, [vp]),
        "orc_sheet_size": (ctypes.c_int, [vp]),
        "orc_sheet_geometry": (None, [vp] + [_dp] * 9),
        "orc_sheet_influence": (None, [vp, _dp]),
        "orc_sheet_get_gamma": (None, [vp, _dp]),
        "orc_sheet_set_gamma": (None, [vp, _dp]),
        "orc_sheet_solve": (None, [vp, _i64, _dp, _dp, _dp, _dp, _dp, _dp]),
        "orc_sheet_bc": (None, [vp, _i64, _dp, _dp, _dp, _dp]),
        "orc_sheet_inside": (ctypes.c_int, [vp, ctypes.c_double, ctypes.c_double]),
        "orc_sheet_loads": (None, [vp, ctypes.c_double, _dp, _dp, _dp, _dp]),
        "orc_sheet_probe_quirk": (None, [vp, _i64, _dp, _dp, _dp, _dp, _i64, _dp, _dp,
                                         _dp, _dp, _dp, _dp]),
        "orc_sim_create": (vp, [ctypes.c_int, _dp, _dp, ctypes.POINTER(Params)]),
        "orc_sim_destroy": (None, [vp]),
        "orc_sim_size": (_i64, [vp]),
        "orc_sim_step": (ctypes.c_int, [vp]),
        "orc_sim_get_state": (None, [vp, _u32p] + [_dp] * 8),
        "orc_sim_set_state": (None, [vp, _i64, _u32p] + [_dp] * 8),
        "orc_sim_get_gamma": (None, [vp, _dp]),
        "orc_sim_set_gamma": (None, [vp, _dp]),
        "orc_sim_get_loads": (None, [vp, _dp, _dp, _dp, _dp]),
        "orc_sim_step_index": (ctypes.c_int, [vp]),
        "orc_sim_rng_draws": (ctypes.c_uint64, [vp]),
        "orc_sim_set_clock": (None, [vp, ctypes.c_int, ctypes.c_uint64]),
        "orc_philox4x32": (None, [vp, vp, vp]),
        "orc_u01": (ctypes.c_double, [ctypes.c_uint32, ctypes.c_uint32]),
        "orc_sim_convect": (None, [vp]),
        "orc_sim_remesh": (None, [vp]),
        "orc_sim_diffuse": (None, [vp]),
        "orc_sim_release": (ctypes.c_int, [vp]),
        "orc_sim_crossing": (None, [vp]),
        "orc_sim_loads": (None, [vp]),
        "orc_sim_merge": (_i64, [vp]),
        "orc_merge_blobs": (_i64, [ctypes.c_double, _i64, ctypes.POINTER(ctypes.c_uint32)] + [_dp] * 8),
        "orc_remesh": (_i64, [ctypes.POINTER(Params), _i64, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
        "orc_remesh_weight": (ctypes.c_double, [ctypes.c_int, ctypes.c_double]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


_lib = None


def native_lib(path):
    """The oracle built with the reference's flags for the host running it
    (`make -C oracle native`): the CPU baseline's timing build only."""
    return _load(path)


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def philox4x32(ctr, key):
    """Philox4x32-10 words for a 4-word counter and a 2-word key."""
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    o = np.empty(4, dtype=np.uint32)
    lib().orc_philox4x32(c.ctypes.data, k.ctypes.data, o.ctypes.data)
    return o


def u01(a, b):
    """The device stream's 53-bit uniform in (0, 1) from two words."""
    return lib().orc_u01(int(a), int(b))


def bs_self(x, z, circ, sigma):
    x, z, circ, sigma = map(_f64, (x, z, circ, sigma))
    n = x.size
    u = np.empty(n)
    w = np.empty(n)
    lib().orc_bs_self(n, _p(x), _p(z), _p(circ), _p(sigma), _p(u), _p(w))
    return u, w


def bs_self_ld(x, z, circ, sigma):
    x, z, circ, sigma = map(_f64, (x, z, circ, sigma))
    n = x.size
    u = np.empty(n)
    w = np.empty(n)
    lib().orc_bs_self_ld(n, _p(x), _p(z), _p(circ), _p(sigma), _p(u), _p(w))
    return u, w


def bs_targets(sx, sz, scirc, ssigma, tx, tz, using=None):
    """computeInducedAt (src/VelocityKernel.cpp:69-108); `using` picks another
    build of the library (the CPU baseline's)."""
    sx, sz, scirc, ssigma, tx, tz = map(_f64, (sx, sz, scirc, ssigma, tx, tz))
    nt = tx.size
    u = np.empty(nt)
    w = np.empty(nt)
    (using or lib()).orc_bs_targets(sx.size, _p(sx), _p(sz), _p(scirc), _p(ssigma), nt, _p(tx), _p(tz),
                         _p(u), _p(w))
    return u, w


class Rng:
    """src/Random.cpp:3-18."""

    def __init__(self, seed):
        self.h = lib().orc_rng_create(seed)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_rng_destroy(self.h)

    def rand(self):
        return lib().orc_rng_rand(self.h)

    def fill(self, n):
        out = np.empty(n)
        lib().orc_rng_fill(self.h, n, _p(out))
        return out

    def raw(self):
        return lib().orc_rng_raw(self.h)


class Sheet:
    """VortexSheet restated (src/VortexSheet.cpp)."""

    def __init__(self, px, pz, params):
        self.px, self.pz = _f64(px), _f64(pz)
        self.params = params
        self.h = lib().orc_sheet_create(self.px.size, _p(self.px), _p(self.pz), ctypes.byref(params))
        self.nl = lib().orc_sheet_size(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_sheet_destroy(self.h)

    def geometry(self):
        arrs = [np.empty(self.nl) for _ in range(8)]
        aref = ctypes.c_double()
        lib().orc_sheet_geometry(self.h, *[_p(a) for a in arrs], ctypes.byref(aref))
        names = ["xc", "zc", "theta", "ds", "enx", "enz", "etx", "etz"]
        g = dict(zip(names, arrs))
        g["aref"] = aref.value
        return g

    def influence(self):
        A = np.empty((self.nl + 1, self.nl))
        lib().orc_sheet_influence(self.h, _p(A))
        return A

    @property
    def gamma(self):
        g = np.empty(self.nl)
        lib().orc_sheet_get_gamma(self.h, _p(g))
        return g

    @gamma.setter
    def gamma(self, g):
        g = _f64(g)
        lib().orc_sheet_set_gamma(self.h, _p(g))

    def solve(self, x, z, circ, sigma):
        x, z, circ, sigma = map(_f64, (x, z, circ, sigma))
        b = np.empty(self.nl + 1)
        g = np.empty(self.nl)
        lib().orc_sheet_solve(self.h, x.size, _p(x), _p(z), _p(circ), _p(sigma), _p(b), _p(g))
        return b, g

    def bc(self, x, z):
        x, z = _f64(x), _f64(z)
        uvs = np.empty(x.size)
        wvs = np.empty(x.size)
        lib().orc_sheet_bc(self.h, x.size, _p(x), _p(z), _p(uvs), _p(wvs))
        return uvs, wvs

    def inside(self, x, z):
        return bool(lib().orc_sheet_inside(self.h, float(x), float(z)))

    def loads(self, ur):
        v = [ctypes.c_double() for _ in range(4)]
        lib().orc_sheet_loads(self.h, ur, *[ctypes.byref(a) for a in v])
        return tuple(a.value for a in v)

    def probe_quirk(self, x, z, circ, sigma, prx, prz):
        x, z, circ, sigma, prx, prz = map(_f64, (x, z, circ, sigma, prx, prz))
        outs = [np.empty(prx.size) for _ in range(4)]
        lib().orc_sheet_probe_quirk(self.h, x.size, _p(x), _p(z), _p(circ), _p(sigma), prx.size,
                                    _p(prx), _p(prz), *[_p(o) for o in outs])
        return tuple(outs)


STATE_FIELDS = ("x", "z", "circ", "sigma", "u", "w", "uvs", "wvs")


class Sim:
    """DVMBase restated (src/DVMBase.cpp)."""

    def __init__(self, px, pz, params):
        self.px, self.pz = _f64(px), _f64(pz)
        self.params = params
        self.h = lib().orc_sim_create(self.px.size, _p(self.px), _p(self.pz), ctypes.byref(params))
        self.nl = self.px.size - 1

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_sim_destroy(self.h)

    def size(self):
        return lib().orc_sim_size(self.h)

    def step(self):
        return lib().orc_sim_step(self.h)

    def get_state(self):
        n = self.size()
        ids = np.empty(n, dtype=np.uint32)
        arrs = {k: np.empty(n) for k in STATE_FIELDS}
        lib().orc_sim_get_state(self.h, ids.ctypes.data_as(_u32p), *[_p(arrs[k]) for k in STATE_FIELDS])
        arrs["id"] = ids
        return arrs

    def set_state(self, st):
        n = st["x"].size
        ids = np.ascontiguousarray(st["id"], dtype=np.uint32)
        arrs = [_f64(st[k]) for k in STATE_FIELDS]
        lib().orc_sim_set_state(self.h, n, ids.ctypes.data_as(_u32p), *[_p(a) for a in arrs])

    def gamma(self):
        g = np.empty(self.nl)
        lib().orc_sim_get_gamma(self.h, _p(g))
        return g

    def set_gamma(self, g):
        g = _f64(g)
        lib().orc_sim_set_gamma(self.h, _p(g))

    def loads(self):
        v = [ctypes.c_double() for _ in range(4)]
        lib().orc_sim_get_loads(self.h, *[ctypes.byref(a) for a in v])
        return tuple(a.value for a in v)

    def step_index(self):
        return lib().orc_sim_step_index(self.h)

    def rng_draws(self):
        return lib().orc_sim_rng_draws(self.h)

    def set_clock(self, step, draws):
        """Step index and random-stream position of another run."""
        lib().orc_sim_set_clock(self.h, step, draws)

    def convect(self):
        lib().orc_sim_convect(self.h)

    def remesh(self):
        lib().orc_sim_remesh(self.h)

    def diffuse(self):
        lib().orc_sim_diffuse(self.h)

    def release(self):
        return lib().orc_sim_release(self.h)

    def crossing(self):
        lib().orc_sim_crossing(self.h)

    def compute_loads(self):
        lib().orc_sim_loads(self.h)

    def merge(self):
        return lib().orc_sim_merge(self.h)


def merge_blobs(threshold, st):
    """VortexBlobs::merge_blobs on a state dict (id, x, z, circ, sigma, u, w,
    uvs, wvs); returns (merges, new state)."""
    n = st["x"].size
    ids = np.ascontiguousarray(st["id"], dtype=np.uint32).copy()
    f = {k: _f64(st[k]).copy() for k in ("x", "z", "circ", "sigma", "u", "w", "uvs", "wvs")}
    merged = lib().orc_merge_blobs(float(threshold), n,
                                   ids.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                   *[_p(f[k]) for k in ("x", "z", "circ", "sigma", "u", "w",
                                                        "uvs", "wvs")])
    m = n - merged
    out = {k: v[:m].copy() for k, v in f.items()}
    out["id"] = ids[:m].copy()
    return merged, out


def remesh(params, x, z, circ, sigma):
    x, z, circ, sigma = map(_f64, (x, z, circ, sigma))
    n = lib().orc_remesh(ctypes.byref(params), x.size, _p(x), _p(z), _p(circ), _p(sigma),
                         None, None, None, None)
    outs = [np.empty(n) for _ in range(4)]
    lib().orc_remesh(ctypes.byref(params), x.size, _p(x), _p(z), _p(circ), _p(sigma),
                     *[_p(o) for o in outs])
    return tuple(outs)


def remesh_weight(kernel, r):
    return lib().orc_remesh_weight(kernel, float(r))


def cylinder_points(nl, radius=0.5):
    """Closed CCW circle of nl panels; the last point equals the first exactly."""
    k = np.arange(nl + 1)
    th = 2.0 * np.pi * k / nl
    x = radius * np.cos(th)
    z = radius * np.sin(th)
    x[-1] = x[0]
    z[-1] = z[0]
    return x, z
