"""Input-file reader and output number format of the C++ layer
(dvmpp_amd/host/xml.cpp, output.cpp) through tests/cpp/test_xml.cpp; no GPU."""
import os
import subprocess

from conftest import ROOT

BIN = os.path.join(ROOT, "tests", "cpp", "test_xml.bin")
CLI = os.path.join(ROOT, "dvmpp_amd", "dvmpp")


def test_cpp_xml_suite(tmp_path):
    assert os.path.exists(BIN), "tests/cpp/test_xml.bin not built (make -C dvmpp_amd/host)"
    r = subprocess.run([BIN, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-2000:]
    assert "0 failed" in r.stdout


def test_cli_front_end_without_device(tmp_path):
    # src/main.cpp:43-58: no arguments print the options and fail; --example-xml
    # prints a document holding every option, or saves it under the given name.
    r = subprocess.run([CLI], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "--input-file" in r.stdout
    r = subprocess.run([CLI, "--example-xml"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and r.stdout.startswith('<?xml version="1.0"?>\n<DVMpp>')
    for key in ("surface_crossing", "max_NumPanelVort", "domain_file", "circ_threshold"):
        assert "<" + key + " " in r.stdout
    out = tmp_path / "example.xml"
    r = subprocess.run([CLI, "--example-xml", str(out)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and out.read_text().startswith("<?xml")
    r = subprocess.run([CLI, "--bogus"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "unrecognised option" in r.stderr
    r = subprocess.run([CLI, str(tmp_path / "missing.xml")], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 1 and "Cannot load XML file" in r.stderr
