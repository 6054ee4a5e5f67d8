"""Command-line handling of the apps (mesh_hmc, mesh_sweep_beta, mesh_landau):
bad arguments end with status 2 and a usage line before any GPU work, so
these run without a GPU."""
import os
import subprocess

import pytest

BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "mesh_amd", "bin")


def _run(name, args):
    exe = os.path.join(BIN, name)
    if not os.path.exists(exe):
        pytest.skip(f"{name} is not built (python -c 'import __graft_entry__ as g; g.build()')")
    return subprocess.run([exe] + args, capture_output=True, text=True, timeout=60)


@pytest.mark.parametrize("name,args,msg", [
    ("mesh_hmc", ["--bogus"], "unknown option --bogus"),
    ("mesh_hmc", ["--spacing", "0"], "--spacing must be positive"),
    ("mesh_hmc", ["--count"], "missing value for --count"),
    ("mesh_sweep_beta", ["--beta-count", "0"], "must be positive"),
    ("mesh_sweep_beta", ["--nope"], "unknown option --nope"),
    ("mesh_landau", [], "no input configuration"),
    ("mesh_landau", ["a/configs/1", "b/configs/2"], "more than one input"),
    ("mesh_landau", ["--frobnicate", "x"], "unknown option --frobnicate"),
])
def test_usage_errors(name, args, msg):
    r = _run(name, args)
    assert r.returncode == 2, (r.returncode, r.stderr)
    assert msg in r.stderr and "usage:" in r.stderr


def test_landau_missing_ensemble(tmp_path):
    """A configuration name whose ensemble does not exist is an error, not a crash."""
    r = _run("mesh_landau", [str(tmp_path / "none" / "configs" / "1")])
    assert r.returncode == 1 and "error:" in r.stderr
