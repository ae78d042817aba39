"""The C-ABI library loads and exports every entry point include/*.h declares
(CPU only: no compute calls)."""
import ctypes
import glob
import os
import re
import subprocess

import pytest

from conftest import ROOT

LIB = os.path.join(ROOT, "dvmpp_amd", "libdvm_hip.so")


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(dvm_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_header_declares_entry_points():
    names = declared_functions()
    assert {"dvm_ctx_create", "dvm_bs_self", "dvm_bs_targets"} <= names


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "libdvm_hip.so not built"
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = sorted(declared_functions() - exported)
    assert not missing, missing


def test_library_loads_and_reports_version():
    lib = ctypes.CDLL(LIB)
    lib.dvm_version.restype = ctypes.c_char_p
    assert b"gfx950" in lib.dvm_version()
    lib.dvm_status_string.restype = ctypes.c_char_p
    lib.dvm_status_string.argtypes = [ctypes.c_int]
    assert lib.dvm_status_string(3) == b"non-finite release counts"


def test_python_binding_covers_header():
    from dvmpp_amd import _lib
    missing = sorted(declared_functions() - set(_lib.SIGNATURES))
    assert not missing, missing


def test_null_context_rejected_without_gpu():
    lib = ctypes.CDLL(LIB)
    lib.dvm_bs_self.restype = ctypes.c_int
    # a null context is rejected before any HIP call
    assert lib.dvm_bs_self(None, 0, ctypes.c_int64(0), None, None, None, None, None, None) == 1


def test_oracle_not_linked_into_product():
    out = subprocess.run(["ldd", LIB], capture_output=True, text=True).stdout
    assert "oracle" not in out
    syms = subprocess.run(["nm", "-D", LIB], capture_output=True, text=True).stdout
    assert "orc_" not in syms


@pytest.mark.parametrize("path", ["dvmpp_amd", "include"])
def test_product_sources_do_not_reference_oracle(path):
    for f in glob.glob(os.path.join(ROOT, path, "**", "*"), recursive=True):
        if os.path.isfile(f) and f.endswith((".py", ".hip", ".hpp", ".cpp", ".h")):
            assert "import oracle" not in open(f).read(), f
            assert "orc_" not in open(f).read(), f


def test_server_rejects_what_cannot_fit_without_gpu():
    """The serving loop rejects on arrival a request whose class region alone
    exceeds the budget (no device work; a fake clock), so run() returns."""
    import numpy as np

    from dvmpp_amd import bodies, sim
    from dvmpp_amd.batch import batch_config
    from dvmpp_amd.serve import STATUS_REJECTED, Request, Server
    now = [0.0]

    def clock():
        now[0] += 0.01
        return now[0]

    srv = Server(None, batch_config(slots=4, slot_capacity=1 << 14, max_panels=64),
                 budget_bytes=1024)
    px, pz = bodies.cylinder(32)
    reqs = [Request(i, 0.02 * i, px, pz, sim.sheet_params(), i, 10) for i in range(3)]
    # one more whose projection exceeds every class
    reqs.append(Request(3, 0.0, px, pz, sim.sheet_params(), 3, 10**6))
    done, wall = srv.run(reqs, clock=clock)
    assert len(done) == 4 and all(r.status == STATUS_REJECTED for r in done)
    assert not srv.batches and wall < 1.0
    assert np.isfinite(wall)


def test_server_fmm_slots_need_euler_f32c_without_gpu():
    """Fast multipole slots are refused for a template the batch cannot run
    them in (FP64 pair sums, or RK3 with three sums inside one graph)."""
    import pytest

    import dvmpp_amd as dvm
    from dvmpp_amd import sim
    from dvmpp_amd.batch import batch_config
    from dvmpp_amd.serve import Server
    for cfg in (batch_config(prec=dvm.F64), batch_config(prec=dvm.F32C, scheme=sim.RK3)):
        with pytest.raises(ValueError):
            Server(None, cfg, budget_bytes=1 << 30, fmm_min=1000)
    Server(None, batch_config(prec=dvm.F32C), budget_bytes=1 << 30, fmm_min=1000)
