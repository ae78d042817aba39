"""Python handles for the device-resident vortex sheet and simulation
(dvm_sheet_* / dvm_sim_* / dvm_remesh in include/dvm_hip.h)."""
import ctypes

import numpy as np

from ._lib import F64, RemeshParams, SheetParams, SimParams, check, load

STATE_FIELDS = ("x", "z", "circ", "sigma", "u", "w", "uvs", "wvs")

EULER, RK3 = 0, 1
DELETE, ABSORB, REFLECT = 0, 1, 2
RNG_REFERENCE, RNG_DEVICE = 0, 1
M4PRIME, GAUSSIAN, M6 = 0, 1, 2


def _hp(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def sheet_params(rho=1.0, nu=1e-3, dt=0.005, ux=1.0, uz=0.0, cutoff_exp=0.675, max_npv=3):
    return SheetParams(rho, nu, dt, ux, uz, cutoff_exp, max_npv)


def remesh_params(frequency=0, kernel=M4PRIME, xmin=-5.0, xmax=25.0, zmin=-10.0, zmax=10.0,
                  dx=0.01, dz=0.01, sigma=0.05, circ_threshold=1e-10):
    return RemeshParams(frequency, kernel, xmin, xmax, zmin, zmax, dx, dz, sigma, circ_threshold)


def sim_params(scheme=EULER, crossing=REFLECT, seed=0, prec=0, rng_mode=RNG_REFERENCE,
               remesh=None, merge_threshold=0.0, merge_frequency=0, **sheet_kw):
    return SimParams(sheet_params(**sheet_kw), scheme, crossing, seed, prec, rng_mode,
                     remesh if remesh is not None else remesh_params(), merge_threshold,
                     merge_frequency)


class Sheet:
    """dvm_sheet_t: one body's vortex sheet on the device."""

    def __init__(self, ctx, px, pz, params):
        self.ctx = ctx
        self.px, self.pz = _f64(px), _f64(pz)
        self.params = params
        h = ctypes.c_void_p()
        check(ctx.h, load().dvm_sheet_create(ctx.h, self.px.size, _hp(self.px), _hp(self.pz),
                                             ctypes.byref(params), ctypes.byref(h)))
        self.h = h
        ctx.adopt(self)
        n = ctypes.c_int()
        check(ctx.h, load().dvm_sheet_size(h, ctypes.byref(n)))
        self.nl = n.value

    def close(self):
        if getattr(self, "h", None):
            if getattr(self.ctx, "h", None):  # a closed context already released it
                load().dvm_sheet_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def geometry(self):
        arrs = [np.empty(self.nl) for _ in range(8)]
        aref = ctypes.c_double()
        check(self.ctx.h, load().dvm_sheet_geometry(self.h, *[_hp(a) for a in arrs],
                                                    ctypes.byref(aref)))
        g = dict(zip(["xc", "zc", "theta", "ds", "enx", "enz", "etx", "etz"], arrs))
        g["aref"] = aref.value
        return g

    def influence(self):
        A = np.empty((self.nl + 1, self.nl))
        check(self.ctx.h, load().dvm_sheet_influence(self.h, _hp(A)))
        return A

    @property
    def gamma(self):
        g = np.empty(self.nl)
        check(self.ctx.h, load().dvm_sheet_get_gamma(self.h, _hp(g)))
        return g

    @gamma.setter
    def gamma(self, g):
        g = _f64(g)
        check(self.ctx.h, load().dvm_sheet_set_gamma(self.h, _hp(g)))

    def solve(self, x, z, circ, sigma, prec=0):
        """x.. are torch device tensors; returns the RHS b (host)."""
        b = np.empty(self.nl + 1)
        n = x.numel()
        ptr = (lambda t: ctypes.c_void_p(t.data_ptr()))
        check(self.ctx.h, load().dvm_sheet_solve(self.h, prec, n, ptr(x), ptr(z), ptr(circ),
                                                 ptr(sigma), _hp(b)))
        return b

    def on_blobs(self, x, z, far=False):
        """vortexsheetbc velocities at the blobs; far=True: the fast form
        (Laurent series of the sheet beyond two body radii)."""
        import torch
        n = x.numel()
        uvs = torch.empty(n, dtype=torch.float64, device=x.device)
        wvs = torch.empty(n, dtype=torch.float64, device=x.device)
        ptr = (lambda t: ctypes.c_void_p(t.data_ptr()))
        fn = load().dvm_sheet_on_blobs_far if far else load().dvm_sheet_on_blobs
        check(self.ctx.h, fn(self.h, n, ptr(x), ptr(z), ptr(uvs), ptr(wvs)))
        return uvs, wvs

    def loads(self, ur):
        out = np.empty(4)
        check(self.ctx.h, load().dvm_sheet_loads(self.h, ur, _hp(out)))
        return tuple(out)

    def probe_quirk(self, x, z, circ, sigma, px, pz, prec=F64):
        """probe_velocities in the reference's form: (u, w, uvs, wvs) at the
        probes (device tensors in, device tensors out)."""
        import torch
        n, npr = x.numel(), px.numel()
        outs = [torch.empty(npr, dtype=torch.float64, device=px.device) for _ in range(4)]
        ptr = (lambda t: ctypes.c_void_p(t.data_ptr()))
        check(self.ctx.h, load().dvm_sheet_probe_quirk(
            self.h, prec, n, ptr(x), ptr(z), ptr(circ), ptr(sigma), npr, ptr(px), ptr(pz),
            *[ptr(o) for o in outs]))
        return tuple(outs)

    def inside(self, x, z):
        import torch
        n = x.numel()
        f = torch.empty(n, dtype=torch.int32, device=x.device)
        ptr = (lambda t: ctypes.c_void_p(t.data_ptr()))
        check(self.ctx.h, load().dvm_sheet_inside(self.h, n, ptr(x), ptr(z), ptr(f)))
        return f

    def reflect(self, x, z):
        ptr = (lambda t: ctypes.c_void_p(t.data_ptr()))
        check(self.ctx.h, load().dvm_sheet_reflect(self.h, x.numel(), ptr(x), ptr(z)))


def merge_blobs(ctx, threshold, x, z, circ, sigma, u, w, uvs, wvs, ids=None):
    """dvm_merge_blobs in place on torch device tensors; returns the number of
    merges (the survivors are the first n - merges entries)."""
    ptr = (lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else None)
    k = ctypes.c_int64()
    check(ctx.h, load().dvm_merge_blobs(ctx.h, float(threshold), x.numel(), ptr(ids), ptr(x),
                                        ptr(z), ptr(circ), ptr(sigma), ptr(u), ptr(w), ptr(uvs),
                                        ptr(wvs), ctypes.byref(k)))
    return k.value


def remesh(ctx, params, x, z, circ, sigma):
    """dvm_remesh on torch device tensors; returns four new device tensors."""
    import torch
    ptr = (lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else None)
    n = x.numel()
    n_out = ctypes.c_int64()
    st = load().dvm_remesh(ctx.h, ctypes.byref(params), n, ptr(x), ptr(z), ptr(circ), ptr(sigma),
                           0, None, None, None, None, ctypes.byref(n_out))
    if st not in (0, 2):
        check(ctx.h, st)
    m = n_out.value
    outs = [torch.empty(max(m, 1), dtype=torch.float64, device=x.device) for _ in range(4)]
    check(ctx.h, load().dvm_remesh(ctx.h, ctypes.byref(params), n, ptr(x), ptr(z), ptr(circ),
                                   ptr(sigma), m, *[ptr(o) for o in outs], ctypes.byref(n_out)))
    return tuple(o[:m] for o in outs)


class Sim:
    """dvm_sim_t: one device-resident simulation."""

    def __init__(self, ctx, px, pz, params):
        self.ctx = ctx
        self.px, self.pz = _f64(px), _f64(pz)
        self.params = params
        self.nl = self.px.size - 1
        h = ctypes.c_void_p()
        check(ctx.h, load().dvm_sim_create(ctx.h, self.px.size, _hp(self.px), _hp(self.pz),
                                           ctypes.byref(params), ctypes.byref(h)))
        self.h = h
        ctx.adopt(self)

    def close(self):
        if getattr(self, "h", None):
            if getattr(self.ctx, "h", None):
                load().dvm_sim_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self):
        check(self.ctx.h, load().dvm_sim_step(self.h))

    def run(self, steps):
        check(self.ctx.h, load().dvm_sim_run(self.h, steps))

    def substep(self, which):
        check(self.ctx.h, load().dvm_sim_substep(self.h, which))

    def size(self):
        n = ctypes.c_int64()
        check(self.ctx.h, load().dvm_sim_size(self.h, ctypes.byref(n)))
        return n.value

    def step_index(self):
        n = ctypes.c_int()
        check(self.ctx.h, load().dvm_sim_step_index(self.h, ctypes.byref(n)))
        return n.value

    def rng_draws(self):
        n = ctypes.c_uint64()
        check(self.ctx.h, load().dvm_sim_rng_draws(self.h, ctypes.byref(n)))
        return n.value

    def get_state(self):
        n = self.size()
        ids = np.empty(n, dtype=np.uint32)
        arrs = {k: np.empty(n) for k in STATE_FIELDS}
        check(self.ctx.h, load().dvm_sim_get_state(self.h, _hp(ids),
                                                   *[_hp(arrs[k]) for k in STATE_FIELDS]))
        arrs["id"] = ids
        return arrs

    def set_state(self, st):
        n = st["x"].size
        ids = np.ascontiguousarray(st["id"], dtype=np.uint32)
        arrs = [_f64(st[k]) for k in STATE_FIELDS]
        check(self.ctx.h, load().dvm_sim_set_state(self.h, n, _hp(ids), *[_hp(a) for a in arrs]))

    def gamma(self):
        g = np.empty(self.nl)
        check(self.ctx.h, load().dvm_sim_get_gamma(self.h, _hp(g)))
        return g

    def set_gamma(self, g):
        g = _f64(g)
        check(self.ctx.h, load().dvm_sim_set_gamma(self.h, _hp(g)))

    def loads(self):
        out = np.empty(4)
        check(self.ctx.h, load().dvm_sim_loads(self.h, _hp(out)))
        return tuple(out)

    def merged(self):
        k = ctypes.c_int64()
        check(self.ctx.h, load().dvm_sim_merged(self.h, ctypes.byref(k)))
        return k.value

    def snapshot(self):
        """The whole simulation state as bytes (dvm_sim_snapshot)."""
        nb = ctypes.c_size_t()
        check(self.ctx.h, load().dvm_sim_snapshot_size(self.h, ctypes.byref(nb)))
        buf = np.empty(nb.value, dtype=np.uint8)
        check(self.ctx.h, load().dvm_sim_snapshot(self.h, _hp(buf), nb.value))
        return buf.tobytes()

    def restore(self, blob):
        """Loads a snapshot taken from a simulation of the same body and parameters."""
        buf = np.frombuffer(blob, dtype=np.uint8).copy()
        check(self.ctx.h, load().dvm_sim_restore(self.h, _hp(buf), buf.size))

    def copy_out(self, x, z, circ, gamma, loads4):
        """Enqueues copies of positions, circulations, gamma and loads into
        host buffers (numpy views of dvm_host_alloc memory for asynchronous
        copies); returns the blob count.  Wait on an Event recorded after it."""
        n = ctypes.c_int64()
        cap = x.size if x is not None else 0
        check(self.ctx.h, load().dvm_sim_copy_out(self.h, cap, _hp(x), _hp(z), _hp(circ),
                                                  _hp(gamma), _hp(loads4), ctypes.byref(n)))
        return n.value

    def probe_quirk(self, px, pz, prec=F64):
        """Probe velocities (u, w, uvs, wvs) for the current state, in the
        reference's probe_velocities form (device tensors in and out)."""
        import torch
        npr = px.numel()
        outs = [torch.empty(npr, dtype=torch.float64, device=px.device) for _ in range(4)]
        ptr = (lambda t: ctypes.c_void_p(t.data_ptr()))
        check(self.ctx.h, load().dvm_sim_probe_quirk(self.h, prec, npr, ptr(px), ptr(pz),
                                                     *[ptr(o) for o in outs]))
        return tuple(outs)


class Event:
    """dvm_event_t: a marker on a context's stream."""

    def __init__(self, ctx):
        self.ctx = ctx
        h = ctypes.c_void_p()
        check(ctx.h, load().dvm_event_create(ctx.h, ctypes.byref(h)))
        self.h = h

    def record(self):
        check(self.ctx.h, load().dvm_event_record(self.ctx.h, self.h))

    def wait(self):
        check(None, load().dvm_event_wait(self.h))

    def done(self):
        d = ctypes.c_int()
        check(None, load().dvm_event_query(self.h, ctypes.byref(d)))
        return bool(d.value)

    def close(self):
        if getattr(self, "h", None):
            load().dvm_event_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HostBuffer:
    """Pinned host memory (dvm_host_alloc) viewed as a float64 numpy array."""

    def __init__(self, count):
        p = ctypes.c_void_p()
        check(None, load().dvm_host_alloc(8 * max(count, 1), ctypes.byref(p)))
        self.p = p
        self.array = np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ctypes.c_double)),
                                           shape=(max(count, 1),))[:count]

    def close(self):
        if getattr(self, "p", None):
            self.array = None
            load().dvm_host_free(self.p)
            self.p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
