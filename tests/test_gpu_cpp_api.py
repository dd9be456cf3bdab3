"""The C++ API (hpg_amd/cpp: hpg::GridderState / hpg::Gridder over the C
ABI) runs the restated reference unit tests on the GPU, in a process of its
own (no torch in it)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "hpg_amd", "cpp", "tests", "test_hpg_api")


@pytest.mark.gpu
def test_cpp_api_unit_tests():
    assert os.path.exists(BIN), "build first: python -c 'import __graft_entry__ as g; g.build()'"
    r = subprocess.run([BIN, os.path.join(ROOT, "tests", "golden", "gridone.bin")],
                       capture_output=True, text=True, timeout=300)
    print(r.stdout)
    print(r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failures" in r.stdout


def test_cpp_api_links_the_c_abi():
    lib = os.path.join(ROOT, "hpg_amd", "cpp", "libhpg.so")
    assert os.path.exists(lib)
    out = subprocess.run(["readelf", "-d", lib], capture_output=True, text=True).stdout
    assert "libhpgx.so" in out
