"""The dvmpp command (dvmpp_amd/host/main.cpp) on the reference's own test
input (tests/fixtures/test_configs/cylinder_test.xml, paths pointed at this
repository's copy of simple_body.txt): the per-step files it writes in the
reference's formats (src/VortexBlobs.cpp:16-42,310-331, src/VortexSheet.cpp:
38-53,335-343, src/Probe.cpp:12-52) against the oracle's run of the same case,
and a snapshot/restore continuation against an uninterrupted run."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from conftest import GOLDEN, ROOT, load_body

pytestmark = pytest.mark.gpu

CLI = os.path.join(ROOT, "dvmpp_amd", "dvmpp")
STEPS = 10


def _xml(path, out_dir, steps=STEPS, probes=True, scheme="euler", crossing="REFLECT", extra=""):
    probe = ('<x string="2.0 1.5,0.6" />\n<z string="0.1 -0.3 0.05" />' if probes
             else '<x val="20" />\n<z val="1" />')
    path.write_text(f"""<?xml version="1.0"?>
<DVMpp>
  <algorithms><surface_crossing string="{crossing}" />{extra}</algorithms>
  <constants>
    <density val="1.0" /><nu val="0.001" /><max_NumPanelVort val="3" />
    <cutoff_exp val="0.675" /><seed val="42" />
  </constants>
  <io>
    <input_dir string="{GOLDEN}/" /><output_dir string="{out_dir}" />
    <domain_file string="simple_body.txt" />
  </io>
  <flow><ux val="3.0" /><uz val="0.0" /></flow>
  <probe>{probe}</probe>
  <time><scheme string="{scheme}" /><dt val="0.005" /><steps val="{steps}" /></time>
</DVMpp>
""")


def _run(*args):
    print("dvmpp", *args, flush=True)
    r = subprocess.run([CLI, *map(str, args)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    return r.stdout


def _rows(path, skip):
    with open(path) as f:
        lines = f.read().splitlines()
    return lines[:skip], [ln for ln in lines[skip:]]


def _close(got, want, rel=1e-5):
    # six significant digits in the files; the state itself agrees to ~1e-12
    # of the column's scale, which the floor allows for values near zero
    got, want = np.asarray(got, float), np.asarray(want, float)
    floor = 1e-9 * np.max(np.abs(want)) if want.size else 0.0
    return np.all(np.abs(got - want) <= rel * np.abs(want) + floor)


def test_cli_outputs_match_oracle_run(tmp_path):
    out = str(tmp_path) + "/"
    xml = tmp_path / "case.xml"
    _xml(xml, out)
    log = _run(xml, "--stamp", "t")
    assert "Completed DVM computations successfully" in log
    assert f"Step {STEPS}/{STEPS}" in log

    px, pz = load_body("simple_body.txt")
    op = oracle.make_params(scheme=0, crossing=2, seed=42, nu=0.001, rho=1.0, ux=3.0, uz=0.0,
                            dt=0.005, cutoff_exp=0.675, max_npv=3)
    sim = oracle.Sim(px, pz, op)
    sheet = oracle.Sheet(px, pz, op)
    prx, prz = np.array([2.0, 1.5, 0.6]), np.array([0.1, -0.3, 0.05])

    head, vrows = _rows(out + "t_vortex.dat", 4)
    assert head == ["0 # Number of nodes", "0.005 # Time step", "10 # Steps",
                    "Time [s]\tx-position [m]\tz-position [m]\tcirculation"]
    head, nrows = _rows(out + "t_vortex_num.dat", 1)
    assert head == ["Time [s]\tNumber of vortices"]
    nl = px.size - 1
    head, grows = _rows(out + "t_gamma.dat", 4)
    assert head == [f"{nl} # Number of collocation points", "0.005 # Time Step", "10 # Steps",
                    "Time [s] Gamma - Vortex sheet strength"]
    head, lrows = _rows(out + "t_loads.dat", 1)
    assert head == ["Time [s]\tF_x [-]\tF_z [s]"]
    head, prows = _rows(out + "t_probe.dat", 3)
    assert head == ["0.005 # Time step", "10 # Steps", "Time [s]\tu [m/s]\tw [m/s]"]
    assert len(nrows) == STEPS and len(lrows) == STEPS and len(grows) == STEPS * nl
    assert len(prows) == STEPS * 3

    xc = 0.5 * (px[:-1] + px[1:])
    vpos = 0
    for k in range(1, STEPS + 1):
        assert sim.step() == 0
        st = sim.get_state()
        n = st["x"].size
        t = k * 0.005
        assert nrows[k - 1] == f"{k}\t{n}", k
        block = np.array([r.split("\t") for r in vrows[vpos:vpos + n]], dtype=float)
        vpos += n
        assert block.shape == (n, 4)
        assert np.all(block[:, 0] == float(f"{t:.6g}"))
        for col, key in ((1, "x"), (2, "z"), (3, "circ")):
            assert _close(block[:, col], st[key]), (k, key)
        g = np.array([r.split("\t") for r in grows[(k - 1) * nl:k * nl]], dtype=float)
        assert _close(g[:, 1], xc) and _close(g[:, 3], sim.gamma()), k
        assert grows[(k - 1) * nl].split("\t")[1].startswith(" ")  # "time\t xc" as written
        fx, fz = sim.loads()[:2]
        lt, lfx, lfz = map(float, lrows[k - 1].split("\t"))
        assert lt == float(f"{t:.6g}") and _close([lfx, lfz], [fx, fz]), k
        sheet.gamma = sim.gamma()
        u, w, _, _ = sheet.probe_quirk(st["x"], st["z"], st["circ"], st["sigma"], prx, prz)
        p = np.array([r.split("\t") for r in prows[(k - 1) * 3:k * 3]], dtype=float)
        assert _close(p[:, 1], u) and _close(p[:, 2], w), k
    assert vpos == len(vrows)
    # the input is saved beside the outputs
    assert os.path.exists(out + "t_xml_in.xml")


def test_cli_restore_continues_run(tmp_path):
    out = str(tmp_path) + "/"
    half, full = tmp_path / "half.xml", tmp_path / "full.xml"
    _xml(half, out, steps=5, probes=False, scheme="RK3", crossing="DELETE")
    _xml(full, out, steps=STEPS, probes=False, scheme="RK3", crossing="DELETE")
    _run(half, "--stamp", "a", "--snapshot-every", 5, "--quiet")
    _run(full, "--stamp", "b", "--restore", out + "a_snapshot.bin", "--quiet")
    _run(full, "--stamp", "c", "--quiet")
    # steps 6..10 of the restored run are those of the uninterrupted run, byte for byte
    for name, skip in (("vortex_num", 1), ("loads", 1)):
        b = _rows(out + f"b_{name}.dat", skip)[1]
        c = _rows(out + f"c_{name}.dat", skip)[1]
        assert b == c[5:], name
    b = _rows(out + "b_vortex.dat", 4)[1]
    c = _rows(out + "c_vortex.dat", 4)[1]
    assert b == c[-len(b):] and len(b) > 0
    # and --output-every thins the files without changing the run
    _run(full, "--stamp", "d", "--quiet", "--output-every", 5, "--no-blob-output")
    d = _rows(out + "d_vortex_num.dat", 1)[1]
    c = _rows(out + "c_vortex_num.dat", 1)[1]
    assert d == [c[4], c[9]]
    assert _rows(out + "d_vortex.dat", 4)[1] == []


def test_cli_fmm_velocity_kernel_and_merging(tmp_path):
    # <velocity_kernel string="fmm"> with order 20 (error at FP64 rounding):
    # the run follows the direct one; merging on every 2nd step prints the
    # reference's lines and lowers the blob count
    out = str(tmp_path) + "/"
    direct, fmm, merge = tmp_path / "d.xml", tmp_path / "f.xml", tmp_path / "m.xml"
    _xml(direct, out, probes=False)
    _xml(fmm, out, probes=False, extra='<velocity_kernel string="fmm" />'
         '<fmm_expansion_order val="20" /><fmm_theta val="0.5" />')
    _xml(merge, out, probes=False, extra='<merge_threshold val="1e-4" /><merge_frequency val="2" />')
    _run(direct, "--stamp", "d", "--quiet")
    _run(fmm, "--stamp", "f", "--quiet")
    assert _rows(out + "f_vortex_num.dat", 1)[1] == _rows(out + "d_vortex_num.dat", 1)[1]
    ld = np.array([r.split("\t") for r in _rows(out + "d_loads.dat", 1)[1]], dtype=float)
    lf = np.array([r.split("\t") for r in _rows(out + "f_loads.dat", 1)[1]], dtype=float)
    assert _close(lf, ld)
    log = _run(merge, "--stamp", "m")
    assert "Vortex merging enabled: threshold = 0.0001, frequency = 2" in log
    assert "vortex pairs" in log
    nd = [int(r.split("\t")[1]) for r in _rows(out + "d_vortex_num.dat", 1)[1]]
    nm = [int(r.split("\t")[1]) for r in _rows(out + "m_vortex_num.dat", 1)[1]]
    assert nm[0] == nd[0] and nm[-1] < nd[-1]
