"""The C++ host layer (mesh_amd/host/mesh/*.h) driven by the compiled
programs: the restated reference test and the mesh_hmc driver."""
import os
import re
import subprocess

import pytest

from oracle import u1_oracle as U

pytestmark = pytest.mark.gpu
BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "mesh_amd", "bin")


def test_cpp_lattice_gauge():
    r = subprocess.run([os.path.join(BIN, "test_lattice_gauge")], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "PASSED" in r.stdout


def test_cpp_lattice_apply():
    """mesh/lattice_apply.h: the reference's staple, evolve and reunitize
    lambdas and its plaquette, wilsonDeriv and Landau expressions compiled by
    hipcc over the library's layout, against the named kernels (evolve and
    reunitize bit for bit) and against the same lambdas on the host (bits)."""
    r = subprocess.run([os.path.join(BIN, "test_lattice_apply")], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "FAIL" not in r.stdout and r.stdout.count("PASS") >= 40


@pytest.mark.parametrize("precision", ["2", "1"])
def test_cpp_hmc_driver(precision):
    r = subprocess.run([os.path.join(BIN, "mesh_hmc"), "--geom", "8", "8", "8", "8", "--count", "6", "--therm", "3",
                        "--scheme", "2mn", "--epsilon", "0.5", "--substeps", "6", "--seed", "3",
                        "--precision", precision, "--landau"],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "HMC summary" in r.stdout and "landau:" in r.stdout and "time in force_update" in r.stdout


def test_cpp_hmc_driver_landau_ascent():
    """--landau-ascent: the reference's line-search solver (Landau<G>::run_ascent)
    on the final configuration reaches the reference's stopping condition."""
    r = subprocess.run([os.path.join(BIN, "mesh_hmc"), "--geom", "4", "4", "4", "4", "--count", "4", "--therm", "2",
                        "--scheme", "2mn", "--epsilon", "0.5", "--substeps", "6", "--seed", "3", "--precision", "2",
                        "--landau-ascent"],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr
    m = re.search(r"landau ascent: (\d+) iterations, condition = ([0-9.e+-]+), functional = ([0-9.]+)", r.stdout)
    assert m, r.stdout[-2000:]
    assert int(m.group(1)) < 100000 and float(m.group(2)) < 1e-10


@pytest.mark.parametrize("precision", ["2", "1"])
def test_cpp_hmc_driver_u1(precision):
    """--group u1 on a 16x16 lattice (Hmc<U1<T>>): the mean plaquette of one
    chain against the exact torus <cos theta_P> at beta = 2."""
    r = subprocess.run([os.path.join(BIN, "mesh_hmc"), "--group", "u1", "--geom", "16", "16", "--beta", "2",
                        "--count", "400", "--therm", "50", "--scheme", "2lf", "--epsilon", "0.5", "--substeps",
                        "10", "--seed", "4", "--precision", precision],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    m = re.search(r"plaquette = ([0-9.]+) \+- ([0-9.]+)", r.stdout)
    acc = float(re.search(r"acceptance = ([0-9.]+)", r.stdout).group(1))
    p, err = float(m.group(1)), float(m.group(2))
    assert acc > 0.7
    # the printed error ignores autocorrelation: allow a generous multiple
    assert abs(p - U.exact_mean_cos(2.0, 256)) < 10 * err + 2e-3, (p, err)


@pytest.mark.parametrize("precision", ["2", "1"])
def test_cpp_hmc_driver_u1_4d(precision):
    """--group u1 on the reference's default 4-d geometry class (4^4 here) at
    strong coupling: the mean plaquette against I1(beta)/I0(beta)."""
    import scipy.special as sp

    r = subprocess.run([os.path.join(BIN, "mesh_hmc"), "--group", "u1", "--geom", "4", "4", "4", "4", "--beta", "0.2",
                        "--count", "300", "--therm", "50", "--scheme", "2lf", "--epsilon", "0.5", "--substeps",
                        "8", "--seed", "6", "--precision", precision],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout[-1500:])
    assert r.returncode == 0, r.stdout[-1500:] + r.stderr
    m = re.search(r"plaquette = ([0-9.]+) \+- ([0-9.]+)", r.stdout)
    p, err = float(m.group(1)), float(m.group(2))
    u = sp.iv(1, 0.2) / sp.iv(0, 0.2)
    assert abs(p - u) < 10 * err + 2e-3, (p, err, u)


def _app(name, args, timeout=120):
    r = subprocess.run([os.path.join(BIN, name)] + args, capture_output=True, text=True, timeout=timeout)
    print(r.stdout[-2000:])
    return r


@pytest.mark.parametrize("precision", ["2", "1"])
def test_cpp_sweep_beta_strong_coupling(precision):
    """mesh_sweep_beta (src/apps/sweep_beta.cpp): at small beta the SU(3) mean
    plaquette follows the strong-coupling series beta/18 + beta^2/216."""
    r = _app("mesh_sweep_beta", ["--geom", "4", "4", "4", "4", "--beta-min", "0", "--beta-max", "1.5",
                                 "--beta-count", "4", "--count", "40", "--seed", "3", "--precision", precision])
    assert r.returncode == 0, r.stdout + r.stderr
    rows = re.findall(r"beta = ([0-9.]+), acc = ([0-9.]+), plaq = (-?[0-9.]+)", r.stdout)
    assert len(rows) == 4
    for b, acc, p in rows:
        b, acc, p = float(b), float(acc), float(p)
        assert acc > 0.5
        assert abs(p - (b / 18 + b * b / 216)) < 0.02, (b, p)


def test_cpp_sweep_beta_u1():
    """--group u1 on 8x8: the mean plaquette against the exact torus value."""
    r = _app("mesh_sweep_beta", ["--group", "u1", "--geom", "8", "8", "--beta-min", "0.5", "--beta-max", "2",
                                 "--beta-count", "2", "--count", "300", "--scheme", "2lf", "--epsilon", "0.5",
                                 "--substeps", "10", "--seed", "8"])
    assert r.returncode == 0, r.stdout + r.stderr
    rows = re.findall(r"beta = ([0-9.]+), acc = ([0-9.]+), plaq = (-?[0-9.]+)", r.stdout)
    assert [float(b) for b, _, _ in rows] == [0.5, 2.0]
    for b, _, p in rows:
        assert abs(float(p) - U.exact_mean_cos(float(b), 64)) < 0.03, (b, p)


def test_cpp_landau_app_on_stored_config(tmp_path):
    """mesh_landau (src/apps/landau.cpp) on a configuration that mesh_hmc
    stored: overrelaxation reaches theta < 1e-12 and the reference's ascent its
    own stopping condition, in both precisions; the plaquette is gauge
    invariant."""
    ens = str(tmp_path / "e")
    r = _app("mesh_hmc", ["--geom", "4", "4", "4", "4", "--count", "3", "--scheme", "2mn", "--epsilon", "0.5",
                          "--substeps", "4", "--seed", "9", "--filename", ens])
    assert r.returncode == 0, r.stdout + r.stderr
    cfg = os.path.join(ens, "configs", "3")
    for precision in ("2", "1"):
        r = _app("mesh_landau", ["--precision", precision, cfg])
        assert r.returncode == 0, r.stdout + r.stderr
        m = re.search(r"landau: (\d+) sweeps, theta = ([0-9.e+-]+), functional = ([0-9.]+)", r.stdout)
        assert m and float(m.group(2)) < 1e-12, r.stdout[-500:]
        # theta of U^g rebuilt from the reunitarised g: the FP32 rounding of g
        # and U^g sets its floor
        t = float(re.search(r"reunitarised g: theta = ([0-9.e+-]+)", r.stdout).group(1))
        assert t < (1e-12 if precision == "2" else 1e-9), t
        pb, pa = map(float, re.search(r"plaquette: before ([0-9.]+), after ([0-9.]+)", r.stdout).groups())
        assert abs(pb - pa) < (1e-12 if precision == "2" else 1e-6)
    r = _app("mesh_landau", ["--ascent", "--max-iter", "100000", cfg])
    assert r.returncode == 0, r.stdout + r.stderr
    m = re.search(r"landau ascent: (\d+) iterations, condition = ([0-9.e+-]+)", r.stdout)
    assert m and float(m.group(2)) < 1e-10
    r = _app("mesh_landau", ["--group", "u1", cfg])
    assert r.returncode != 0 and "group mismatch" in r.stderr
    r = _app("mesh_landau", [])
    assert r.returncode == 2


@pytest.mark.parametrize("group,geom,beta", [("su2", ["4", "4", "4", "4"], "2.2"), ("u1", ["6", "6", "6", "6"], "1.0"),
                                             ("u1", ["16", "16"], "2.0")])
def test_cpp_landau_other_groups(tmp_path, group, geom, beta):
    """Landau<SU2<T>> and Landau<U1<T>> (the instantiations of
    src/tests/lattice_gauge.cpp:46-53) through mesh_hmc --landau and
    mesh_landau on a stored configuration: overrelaxation to theta < 1e-12
    (FP32 links: 1e-9), the reference's ascent to its own stopping condition,
    and a gauge-invariant plaquette.  (On rough 3-d U(1) configurations the
    reference's ascent stalls: the parabola vertex of its line search lands at
    a negative step, after which every search returns 0; hence 4-d and 2-d.)"""
    ens = str(tmp_path / "e")
    r = _app("mesh_hmc", ["--group", group, "--geom", *geom, "--beta", beta, "--count", "6", "--scheme", "2lf",
                          "--epsilon", "0.5", "--substeps", "8", "--seed", "5", "--filename", ens, "--landau"])
    assert r.returncode == 0, r.stdout + r.stderr
    assert "HMC summary" in r.stdout and "landau:" in r.stdout
    cfg = os.path.join(ens, "configs", "6")
    for precision in ("2", "1"):
        r = _app("mesh_landau", ["--group", group, "--precision", precision, cfg])
        assert r.returncode == 0, r.stdout + r.stderr
        m = re.search(r"landau: (\d+) sweeps, theta = ([0-9.e+-]+), functional = ([0-9.]+)", r.stdout)
        assert m and float(m.group(2)) < (1e-12 if precision == "2" or group == "u1" else 1e-9), r.stdout[-500:]
        F0 = float(re.search(r"start: functional = (-?[0-9.]+)", r.stdout).group(1))
        assert float(m.group(3)) > F0
        pb, pa = map(float, re.search(r"plaquette: before (-?[0-9.]+), after (-?[0-9.]+)", r.stdout).groups())
        assert abs(pb - pa) < (1e-12 if precision == "2" and group == "su2" else 2e-6)
    r = _app("mesh_landau", ["--group", group, "--ascent", "--max-iter", "100000", cfg])
    assert r.returncode == 0, r.stdout + r.stderr
    m = re.search(r"landau ascent: (\d+) iterations, condition = ([0-9.e+-]+)", r.stdout)
    assert m and float(m.group(2)) < 1e-10, r.stdout[-800:]


def test_cpp_hmc_driver_float_sums():
    """--precision 1 --float-sums: the reference's float-build Delta H runs
    through the driver (the option reaches the library)."""
    r = _app("mesh_hmc", ["--geom", "8", "8", "8", "8", "--count", "4", "--therm", "2", "--scheme", "2mn",
                          "--epsilon", "0.5", "--substeps", "6", "--seed", "3", "--precision", "1", "--float-sums"])
    assert r.returncode == 0, r.stdout + r.stderr
    assert "HMC summary" in r.stdout
