"""MI355X-native DVM time-step hot path (gfx950 HIP kernels behind a C-ABI).

Python face of libdvm_hip.so for tests and benchmarks.  The product is the
C-ABI in include/dvm_hip.h and the C++17 drop-in classes in dvmpp_amd/host;
this module only moves torch device tensors across that boundary.
"""
import ctypes
import weakref

from ._lib import F32C, F64, FMM, TUNE_MEASURE, TUNE_TABLE, DvmError, check, load  # noqa: F401

__all__ = ["F64", "F32C", "FMM", "TUNE_TABLE", "TUNE_MEASURE", "DvmError", "Context", "bs_self",
           "bs_targets", "load"]

# Compiled F32C variants (source tile, target block, sources per reciprocal;
# 0 = matrix-core far field with 32-target boxes) and occupancy targets, as in
# dvmpp_amd/csrc/bs_kernels.hip.
F32C_VARIANTS = ([(t, b, s) for t in (128, 256) for b in (64, 256) for s in (1, 2, 4)]
                 + [(128, 32, 0), (256, 32, 0)])
F32C_OCCUPANCY = (8, 32, 128)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _dev_f64(t, device, name):
    import torch
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name}: expected a torch tensor")
    if t.dtype != torch.float64 or not t.is_contiguous() or t.device != device:
        raise ValueError(f"{name}: expected a contiguous float64 tensor on {device}")
    return t


class Context:
    """dvm_ctx_t sharing one HIP stream with torch.

    With no stream given, a new torch stream (default priority) is created and
    made torch's current stream, so torch's copies and the library's kernels
    are ordered on one queue.  (torch's default stream has handle 0, which the
    C-ABI reads as "create a private stream".)"""

    def __init__(self, device=0, stream=None):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("no GPU visible: the DVM library has no CPU path")
        lib = load()
        self.device = torch.device("cuda", device)
        torch.cuda.set_device(self.device)
        if stream is None:
            stream = torch.cuda.Stream(device=self.device)
            torch.cuda.set_stream(stream)
        if stream.cuda_stream == 0:
            raise ValueError("pass a non-default torch stream (the default stream has handle 0)")
        self.stream = stream
        self._children = weakref.WeakSet()
        h = ctypes.c_void_p()
        st = lib.dvm_ctx_create(device, ctypes.c_void_p(stream.cuda_stream), ctypes.byref(h))
        if st != 0:
            raise DvmError(st, "dvm_ctx_create failed")
        self.h = h

    def adopt(self, child):
        """Registers a handle (sheet, simulation) that must be destroyed before
        the context: the C-ABI requires children to go first."""
        self._children.add(child)

    def close(self):
        if getattr(self, "h", None):
            for child in list(self._children):
                child.close()
            load().dvm_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        check(self.h, load().dvm_ctx_sync(self.h))

    def prof_enable(self, on=True):
        check(self.h, load().dvm_prof_enable(self.h, 1 if on else 0))

    def prof_reset(self):
        check(self.h, load().dvm_prof_reset(self.h))

    def prof_read(self, kernel):
        """(total_ms, launches) of a named kernel since the last reset."""
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        check(self.h, load().dvm_prof_read(self.h, kernel.encode(), ctypes.byref(ms),
                                           ctypes.byref(n)))
        return ms.value, n.value

    def set_tuning(self, mode):
        """TUNE_TABLE (fixed F32C variant per size) or TUNE_MEASURE (time
        every variant on first use per size bucket, keep the fastest)."""
        check(self.h, load().dvm_set_tuning(self.h, mode))

    def set_f32c_variant(self, tile=0, block=0, share=0, occupancy=0):
        """Force an F32C variant; tile=0 clears the forced choice."""
        check(self.h, load().dvm_set_f32c_variant(self.h, tile, block, share, occupancy))

    def f32c_variant(self):
        """(tile, block, share, occupancy) of the last F32C call."""
        out = (ctypes.c_int32 * 4)()
        check(self.h, load().dvm_get_f32c_variant(self.h, out))
        return tuple(out)

    def set_fmm(self, expansion_order=10, theta=0.5):
        """Parameters of FMM sums (the reference FMMKernel's defaults)."""
        check(self.h, load().dvm_set_fmm(self.h, expansion_order, theta))

    def fmm_last_level(self):
        """Leaf level of the last FMM sum (0: it ran the direct sum)."""
        out = ctypes.c_int()
        check(self.h, load().dvm_fmm_last_level(self.h, ctypes.byref(out)))
        return out.value


def bs_self(ctx, x, z, circ, sigma, prec=F64, out=None):
    """Blob-on-blob velocities (u, w) as float64 device tensors."""
    import torch
    n = x.numel()
    for t, nm in ((x, "x"), (z, "z"), (circ, "circ"), (sigma, "sigma")):
        _dev_f64(t, ctx.device, nm)
        if t.numel() != n:
            raise ValueError("size mismatch")
    if out is None:
        out = (torch.empty(n, dtype=torch.float64, device=ctx.device),
               torch.empty(n, dtype=torch.float64, device=ctx.device))
    u, w = out
    check(ctx.h, load().dvm_bs_self(ctx.h, prec, n, _ptr(x), _ptr(z), _ptr(circ), _ptr(sigma),
                                    _ptr(u), _ptr(w)))
    return u, w


def bs_targets(ctx, sx, sz, scirc, ssigma, tx, tz, prec=F64, guard=True, r2_skip=1e-15, out=None):
    """Velocities induced by sources at targets (computeInducedAt when guard)."""
    import torch
    ns, nt = sx.numel(), tx.numel()
    for t, nm in ((sx, "sx"), (sz, "sz"), (scirc, "scirc"), (ssigma, "ssigma")):
        _dev_f64(t, ctx.device, nm)
        if t.numel() != ns:
            raise ValueError("size mismatch")
    _dev_f64(tx, ctx.device, "tx")
    _dev_f64(tz, ctx.device, "tz")
    if out is None:
        out = (torch.empty(nt, dtype=torch.float64, device=ctx.device),
               torch.empty(nt, dtype=torch.float64, device=ctx.device))
    u, w = out
    check(ctx.h, load().dvm_bs_targets(ctx.h, prec, ns, _ptr(sx), _ptr(sz), _ptr(scirc),
                                       _ptr(ssigma), nt, _ptr(tx), _ptr(tz), 1 if guard else 0,
                                       r2_skip, _ptr(u), _ptr(w)))
    return u, w
