"""The bench script's contract on the CPU: one GPU only, the algorithmic byte
model of a trajectory, and the measured traffic read from profiles/."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def _launches(**kw):
    return {k: {"avg_ms": 1.0 if kw.get(k) else None, "launches": kw.get(k, 0)} for k in bench.KERNELS}


def test_trajectory_bytes_2mn_12():
    """2mn with 12 substeps: the momenta + first link step, 24 fused steps and
    Delta H; R (180 + 24 * 216 + 180) + 16 B per site with compact momenta,
    R (216 + 24 * 288 + 216) + 16 with 18-real matrices, and the separate
    kernels' R (216 * 49 + 288) + 16 = 86,992 B FP64."""
    fused = _launches(momenta_link=1, force_link=24, delta_h=1)
    assert bench.trajectory_bytes_per_site(fused, 8, 1) == 8 * (180 + 24 * 216 + 180) + 16 == 44368
    assert bench.trajectory_bytes_per_site(fused, 4, 1, compact=True) == 4 * 5544 + 16
    assert bench.trajectory_bytes_per_site(fused, 8, 1, compact=False) == 8 * (216 + 24 * 288 + 216) + 16
    separate = _launches(momenta=1, link_update=25, force_update=24, delta_h=1)
    assert bench.trajectory_bytes_per_site(separate, 8, 1, compact=False) == 86992
    assert bench.dominant(_launches(force_link=24, delta_h=1, momenta_link=1)) == "force_link"


def test_measured_traffic_from_profiles():
    f64, f32 = bench.measured_traffic("f64", 32), bench.measured_traffic("f32", 32)
    assert isinstance(f64, int) and isinstance(f32, int)
    # above the algorithmic bytes of one fused step, below twice them
    V = 32 ** 4
    assert bench.kernel_bytes("force_link", 8) * V < f64 < 2 * bench.kernel_bytes("force_link", 8) * V
    assert bench.kernel_bytes("force_link", 4) * V < f32 < 2 * bench.kernel_bytes("force_link", 4) * V
    assert bench.measured_traffic("f64", 48) is None


def test_cpu_baseline_fields():
    """the CPU baseline reports the fused oracle on all threads as its value
    and, beside it, the one-thread and reference-structured figures, the CPU
    model and the affinity core count"""
    r = bench.cpu_baseline(4, 2, "2mn", 0.5, 2)
    assert r["value"] > 0 and r["unit"] == "site-updates/s" and r["kind"] == "port"
    assert r["cores"] >= 1 and r["affinity_cpus"] >= 1
    assert set(r["modes"]) == {"fused_1_thread", "reference_structured", "reference_structured_1_thread"}
    assert all(m["value"] > 0 and len(m["stage_seconds"]) == 4 for m in r["modes"].values())
    assert r["modes"]["fused_1_thread"]["threads"] == 1


def test_timed_window_is_labelled():
    """the acceptance, plaquette and <exp(-dH)> of the timed trajectories come
    out in their own object, labelled as a chain before equilibrium, so that
    they are not read as the equilibrium figures of the physics object"""
    w = bench.timed_window({"acceptance": 0.5, "plaquette": 0.59, "exp_mdH": 3.0}, 50, 2)
    assert set(w) == {"acceptance", "plaquette", "exp_minus_dH", "note"}
    assert w["exp_minus_dH"] == 3.0 and "52" in w["note"] and "physics" in w["note"]
