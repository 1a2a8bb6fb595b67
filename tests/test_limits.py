"""CPU unit test of the device kernels' sizing rules (csrc/limits.hpp): a
2^28-entry list on a one-quartet-per-wave class is launched in pieces below
2^32 work-items, and a shell-pair class beyond 2^24 pairs is rejected (the
driver then returns CQX_ERR_UNSUPPORTED from cqx_ctx_create)."""
import os
import subprocess

from conftest import ROOT


def test_launch_and_index_limits(tmp_path):
    exe = str(tmp_path / "limits_test")
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Wextra",
                        "-I", os.path.join(ROOT, "chronusq_public_amd", "csrc"),
                        os.path.join(ROOT, "tests", "cpp", "limits_test.cpp"), "-o", exe],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout
    assert "FAIL" not in r.stdout


def test_driver_checks_ket_index():
    """The context builder refuses a class beyond the 24-bit ket index."""
    src = open(os.path.join(ROOT, "chronusq_public_amd", "csrc", "cqx_api.cpp")).read()
    assert "ket_index_fits(c->cls[k].size())" in src
    assert "catch (const Unsupported &e)" in src
