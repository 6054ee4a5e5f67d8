"""Configurations through ensemble files on the GPU: write/read round trips
and bit-exact continuation of a chain from a stored configuration, through the
Python Hmc and through the mesh_hmc driver (the reference's --filename and
--start, src/apps/hmc.cpp:63-98)."""
import os
import subprocess

import numpy as np
import pytest

import mesh_amd
from mesh_amd import F32, F64, configio, gauge
from oracle import oracle as O

pytestmark = pytest.mark.gpu
BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "mesh_amd", "bin", "mesh_hmc")


@pytest.mark.parametrize("prec", [F64, F32])
def test_config_round_trip(ctx, tmp_path, prec):
    sh = [4, 2, 3, 6]
    g = mesh_amd.Grid(sh)
    U = mesh_amd.GaugeField(g, prec, ctx)
    U.upload(O.random_gauge(sh, 9, "f64" if prec == F64 else "f32"))
    e = configio.Ensemble.create(str(tmp_path / "e"))
    e.set_attribute("group", "su3")
    e.set_attribute("geometry", sh)
    configio.write_config(e, "1", U)
    a = np.load(e.config_path("1"), allow_pickle=False)
    assert a.shape == (4, 4, 2, 3, 6, 18) and a.dtype == (np.float64 if prec == F64 else np.float32)
    # element order: [mu, x0..x3, 3x3 row-major, re/im]
    h = U.download()
    assert a[1, 3, 1, 2, 5, 2 * 5 + 1] == h[1, ((3 * 2 + 1) * 3 + 2) * 6 + 5, 1, 2].imag
    V = mesh_amd.GaugeField(g, prec, ctx)
    configio.read_config(os.path.join(e.path, "configs", "1"), V)
    np.testing.assert_array_equal(V.download(), h)
    if prec == F64:
        # an FP64 configuration loads into an FP32 field rounded to nearest
        W = mesh_amd.GaugeField(g, F32, ctx)
        configio.read_config(os.path.join(e.path, "configs", "1"), W)
        np.testing.assert_array_equal(W.download(), h.astype(np.complex64))
    with pytest.raises(mesh_amd.MeshError):
        configio.read_config(os.path.join(e.path, "configs", "1"), mesh_amd.GaugeField(mesh_amd.Grid([4, 2, 3, 4]),
                                                                                       prec, ctx))
    with pytest.raises(mesh_amd.MeshError):
        configio.read_config(os.path.join(e.path, "configs", "1"), V, group="u1")


@pytest.mark.parametrize("prec", [F64, F32])
def test_python_chain_resumes_bit_exact(ctx, tmp_path, prec):
    g = mesh_amd.Grid([4, 4, 4, 4])
    deltas = gauge.makeDeltas("2mn", 0.5, 4)
    straight = gauge.Hmc(g, prec, seed=11, ctx=ctx)
    straight.randomizeGaugeField()
    for _ in range(4):
        straight.runHmcUpdate(6.0, deltas)

    first = gauge.Hmc(g, prec, seed=11, ctx=ctx)
    first.randomizeGaugeField()
    e = configio.Ensemble.create(str(tmp_path / "c"))
    e.set_attribute("group", "su3")
    e.set_attribute("geometry", list(g.shape))
    e.set_attribute("philox_key", str(first.seed))
    e.set_attribute("philox_traj0", first.traj)
    for n in range(2):
        first.runHmcUpdate(6.0, deltas)
        configio.write_config(e, str(n + 1), first.U)

    name = os.path.join(e.path, "configs", "2")
    second = gauge.Hmc(g, prec, seed=0, ctx=ctx).load(name)
    assert (second.seed, second.traj) == (11, 2)
    for _ in range(2):
        second.runHmcUpdate(6.0, deltas)
    np.testing.assert_array_equal(second.U.download(), straight.U.download())
    assert second.deltaH_history == straight.deltaH_history[2:]
    assert second.plaq_history == straight.plaq_history[2:]


def _run(args):
    r = subprocess.run([BIN] + args, capture_output=True, text=True, timeout=120)
    print(r.stdout[-1500:])
    return r


@pytest.mark.parametrize("precision", ["2", "1"])
def test_cpp_chain_resumes_bit_exact(tmp_path, precision):
    common = ["--scheme", "2mn", "--epsilon", "0.5", "--substeps", "4", "--precision", precision]
    a, b, c = (str(tmp_path / n) for n in ("a", "b", "c"))
    r = _run(["--geom", "4", "4", "4", "4", "--count", "5", "--seed", "17", "--filename", a] + common)
    assert r.returncode == 0, r.stdout + r.stderr
    r = _run(["--geom", "4", "4", "4", "4", "--count", "2", "--seed", "17", "--filename", b] + common)
    assert r.returncode == 0, r.stdout + r.stderr
    # no --geom, no --seed: both come from the stored chain
    r = _run(["--start", os.path.join(b, "configs", "2"), "--count", "3", "--filename", c] + common)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "continuing the chain" in r.stdout
    A, C = configio.Ensemble.open(a), configio.Ensemble.open(c)
    for n in range(1, 4):
        x = np.load(A.config_path(str(n + 2)), allow_pickle=False)
        y = np.load(C.config_path(str(n)), allow_pickle=False)
        assert x.shape == (4, 4, 4, 4, 4, 18) and x.dtype == (np.float64 if precision == "2" else np.float32)
        np.testing.assert_array_equal(x, y)
    for h in ("plaq_history", "deltaH_history", "accept_history"):
        np.testing.assert_array_equal(A.read_data(h)[2:], C.read_data(h))
    at = C.attributes
    assert at["geometry"] == [4, 4, 4, 4] and at["philox_traj0"] == 2 and at["group"] == "su3"
    assert at["markov_start"] == os.path.join(b, "configs", "2") and at["hmc_scheme"] == "2mn"
    # readConfig's checks: group and geometry
    r = _run(["--group", "u1", "--start", os.path.join(b, "configs", "2"), "--count", "1"] + common)
    assert r.returncode != 0 and "group mismatch" in r.stderr
    r = _run(["--geom", "4", "4", "4", "2", "--start", os.path.join(b, "configs", "2"), "--count", "1"] + common)
    assert r.returncode != 0 and "grid mismatch" in r.stderr
    # --spacing: every second configuration, recorded in markov_spacing
    s = str(tmp_path / "s")
    r = _run(["--geom", "4", "4", "4", "4", "--count", "5", "--seed", "17", "--spacing", "2", "--filename", s] + common)
    assert r.returncode == 0, r.stdout + r.stderr
    assert sorted(os.listdir(os.path.join(s, "configs"))) == ["2.npy", "4.npy"]
    assert configio.Ensemble.open(s).attributes["markov_spacing"] == 2
    np.testing.assert_array_equal(np.load(os.path.join(s, "configs", "4.npy")), np.load(A.config_path("4")))
    # an existing ensemble is not overwritten without --force
    r = _run(["--geom", "4", "4", "4", "4", "--count", "1", "--seed", "1", "--filename", a] + common)
    assert r.returncode != 0 and "exists" in r.stderr


def test_cpp_u1_chain_resumes_bit_exact(tmp_path):
    common = ["--group", "u1", "--beta", "2", "--scheme", "2lf", "--epsilon", "0.5", "--substeps", "8"]
    a, b, c = (str(tmp_path / n) for n in ("a", "b", "c"))
    assert _run(["--geom", "8", "8", "--count", "4", "--seed", "5", "--filename", a] + common).returncode == 0
    assert _run(["--geom", "8", "8", "--count", "1", "--seed", "5", "--filename", b] + common).returncode == 0
    r = _run(["--start", os.path.join(b, "configs", "1"), "--count", "3", "--filename", c] + common)
    assert r.returncode == 0, r.stdout + r.stderr
    A, C = configio.Ensemble.open(a), configio.Ensemble.open(c)
    x = np.load(A.config_path("4"), allow_pickle=False)
    assert x.shape == (2, 8, 8, 1) and x.dtype == np.float32
    np.testing.assert_array_equal(x, np.load(C.config_path("3"), allow_pickle=False))
    np.testing.assert_array_equal(A.read_data("deltaH_history")[1:], C.read_data("deltaH_history"))


def test_cpp_resume_inside_thermalisation(tmp_path):
    """A configuration stored during --therm continues with the rest of the
    thermalisation (no Metropolis) and reproduces the uninterrupted chain bit
    for bit; a precision different from the stored one is reported."""
    common = ["--beta", "6", "--scheme", "2mn", "--epsilon", "0.5", "--substeps", "4"]
    a, c = str(tmp_path / "a"), str(tmp_path / "c")
    r = _run(["--geom", "4", "4", "4", "4", "--count", "6", "--therm", "4", "--seed", "23", "--filename", a] + common)
    assert r.returncode == 0, r.stdout + r.stderr
    r = _run(["--start", os.path.join(a, "configs", "2"), "--count", "4", "--filename", c] + common)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "2 thermalisation trajectories left" in r.stdout
    A, Cn = configio.Ensemble.open(a), configio.Ensemble.open(c)
    np.testing.assert_array_equal(np.load(A.config_path("6")), np.load(Cn.config_path("4")))
    # both histories start after the thermalisation: the chain's trajectories 5 and 6
    np.testing.assert_array_equal(A.read_data("deltaH_history"), Cn.read_data("deltaH_history"))
    r = _run(["--start", os.path.join(a, "configs", "2"), "--count", "1", "--precision", "1"] + common)
    assert r.returncode == 0 and "does not continue bit for bit" in r.stderr
