"""The C++ host layer (mesh_amd/host/mesh/*.h) driven by the compiled
programs: the restated reference test and the mesh_hmc driver."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "mesh_amd", "bin")


def test_cpp_lattice_gauge():
    r = subprocess.run([os.path.join(BIN, "test_lattice_gauge")], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "PASSED" in r.stdout


@pytest.mark.parametrize("precision", ["2", "1"])
def test_cpp_hmc_driver(precision):
    r = subprocess.run([os.path.join(BIN, "mesh_hmc"), "--geom", "8", "8", "8", "8", "--count", "6", "--therm", "3",
                        "--scheme", "2mn", "--epsilon", "0.5", "--substeps", "6", "--seed", "3",
                        "--precision", precision, "--landau"],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "HMC summary" in r.stdout and "landau:" in r.stdout and "time in force_update" in r.stdout
