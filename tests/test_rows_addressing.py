"""The fused MD step's rows-path addressing (src/gauge/utils.h:110-139 staple
links; su3_dev.h rows_site / blk_off): a link displaced along x0, x1 and x3 is
found from its block index and slot exactly where site_off puts the shifted
site.  Host only: the check is compiled with hipcc for the host and runs every
site of lattices that take the path."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not found")
def test_rows_addressing(tmp_path):
    exe = str(tmp_path / "rows_addressing")
    subprocess.run([HIPCC, "-O2", "-std=c++17", "--offload-arch=gfx950",
                    os.path.join(ROOT, "tests", "rows_addressing.cpp"), "-o", exe],
                   check=True, capture_output=True, timeout=300)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert " 0 wrong" in r.stdout
