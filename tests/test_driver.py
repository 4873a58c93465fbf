"""CPU checks of the input.dat driver (tpsc_conductivity_amd/tpsc_run): argument and input
errors are reported before any device is touched (tpsc_main.cpp:25-35 reads input.dat first)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUN = os.path.join(ROOT, "tpsc_conductivity_amd", "tpsc_run")


@pytest.fixture(scope="module")
def driver():
    if not os.path.exists(RUN):
        pytest.fail("tpsc_run not built: run __graft_entry__.build()")
    return RUN


def test_usage_error(driver):
    r = subprocess.run([driver, "--bogus"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr


def test_missing_input_reports_like_the_reference(driver, tmp_path):
    r = subprocess.run([driver, str(tmp_path / "nope.dat")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    assert "echec d'ouverture de input.dat" in r.stdout   # tpsc_main.cpp:29


def test_malformed_input(driver, tmp_path):
    p = tmp_path / "input.dat"
    p.write_text("8 -0.24 0.15\n")
    r = subprocess.run([driver, str(p)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "expected U tp tpp t3 n T nk0 Nw0 chijj3 mmin" in r.stderr
