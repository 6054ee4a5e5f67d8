"""Ensemble files (mesh/io.h, mesh_amd/configio.py): the C++ writer read by
numpy and json, the Python writer read by the C++ reader, and the checks of
readConfig (src/gauge/utils.h:188-219).  Host only."""
import json
import os
import subprocess

import numpy as np
import pytest

from mesh_amd import configio

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "mesh_amd", "bin", "test_io")


@pytest.fixture(scope="module")
def test_io(tmp_path_factory):
    if os.path.exists(BIN):
        return BIN
    exe = str(tmp_path_factory.mktemp("bin") / "test_io")
    subprocess.run(["g++", "-std=c++20", "-O1", f"-I{ROOT}/include", f"-I{ROOT}/mesh_amd/host",
                    os.path.join(ROOT, "mesh_amd", "apps", "test_io.cpp"), "-o", exe], check=True)
    return exe


def test_cpp_writer_read_by_numpy(test_io, tmp_path):
    d = str(tmp_path / "e")
    r = subprocess.run([test_io, "write", d], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    with open(os.path.join(d, "attributes.json")) as f:
        at = json.load(f)
    assert at == {"group": "su3", "beta": 5.75, "geometry": [2, 3, 4], "markov_seed": 'a "quoted" seed\\path',
                  "markov_count": 7, "philox_key": "18446744073709551615", "philox_traj0": 4294967295}
    e = configio.Ensemble.open(d)
    np.testing.assert_array_equal(e.read_data("plaq_history"), [0.5, 0.25, 1.0 / 3.0])
    u = np.load(e.config_path("1"), allow_pickle=False)
    assert u.shape == (3, 2, 3, 4, 18) and u.dtype == np.float64
    np.testing.assert_array_equal(u.reshape(-1), 0.5 * np.arange(u.size) - 1.0 / 7.0)
    t = np.load(e.config_path("2"), allow_pickle=False)
    assert t.shape == (2, 5, 6, 1) and t.dtype == np.float32
    np.testing.assert_array_equal(t.reshape(-1), np.arange(t.size, dtype=np.float32) / np.float32(3))


def test_python_writer_read_by_cpp(test_io, tmp_path):
    d = str(tmp_path / "p")
    e = configio.Ensemble.create(d)
    e.set_attribute("group", "u1")
    e.set_attribute("beta", 0.1)
    e.set_attribute("geometry", (7, 3))
    e.set_attribute("markov_seed", "tab\there")
    e.set_attribute("markov_count", 12)
    e.write_data("plaq_history", [1.0, 2.0, 0.125])
    u = np.linspace(-1, 1, 2 * 7 * 3 * 18).reshape(2, 7, 3, 18)
    configio._save(e.config_path("1"), u)
    t = (np.arange(2 * 7 * 3, dtype=np.float32) * np.float32(0.1)).reshape(2, 7, 3, 1)
    configio._save(e.config_path("2"), t)
    r = subprocess.run([test_io, "read", d], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    out = dict(l.split(" = ", 1) for l in r.stdout.splitlines())
    assert out["group"] == "u1" and float(out["beta"]) == 0.1 and out["geometry"] == "7 3"
    assert out["markov_seed"] == "tab\there" and out["markov_count"] == "12"
    assert out["plaq_history"] == f"3 {float(np.sum([1.0, 2.0, 0.125])):.17g}"
    f8 = out["config 1"].split()
    assert f8[:5] == ["<f8", "2", "7", "3", "18"] and abs(float(f8[6]) - u.sum()) < 1e-9
    f4 = out["config 2"].split()
    assert f4[:5] == ["<f4", "2", "7", "3", "1"] and abs(float(f4[6]) - t.astype(np.float64).sum()) < 1e-9


def test_create_refuses_existing(tmp_path):
    d = str(tmp_path / "x")
    e = configio.Ensemble.create(d)
    configio._save(e.config_path("1"), np.zeros(3))
    with pytest.raises(FileExistsError):
        configio.Ensemble.create(d)
    configio.Ensemble.create(d, allow_overwrite=True)
    assert os.listdir(os.path.join(d, "configs")) == []
    # a directory that is not an ensemble is not touched either
    other = tmp_path / "notes"
    (other / "configs").mkdir(parents=True)
    (other / "configs" / "keep.txt").write_text("x")
    with pytest.raises(FileExistsError):
        configio.Ensemble.create(str(other))
    assert (other / "configs" / "keep.txt").exists()
    configio.Ensemble.create(str(tmp_path / "empty"))
    # allow_overwrite replaces an ensemble, histories included, and still
    # refuses a directory that is not one
    e = configio.Ensemble.create(d, allow_overwrite=True)
    e.write_data("plaq_history", [1.0])
    configio._save(e.config_path("1"), np.zeros(3))
    configio.Ensemble.create(d, allow_overwrite=True)
    assert not os.path.exists(os.path.join(d, "plaq_history.npy")) and os.listdir(os.path.join(d, "configs")) == []
    with pytest.raises(FileExistsError):
        configio.Ensemble.create(str(other), allow_overwrite=True)
    assert (other / "configs" / "keep.txt").exists()


def test_cpp_create_overwrite(test_io, tmp_path):
    r = subprocess.run([test_io, "overwrite", str(tmp_path)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "overwrite ok" in r.stdout, r.stdout + r.stderr


def test_resume_info(tmp_path):
    """a configuration stored inside the thermalisation leaves the rest of it
    to run; the stored precision is reported"""
    e = configio.Ensemble.create(str(tmp_path / "r"))
    e.set_attribute("philox_key", "77")
    e.set_attribute("philox_traj0", 3)
    e.set_attribute("markov_discard", 10)
    e.set_attribute("precision", 1)
    i4 = configio.resume_info(os.path.join(e.path, "configs", "4"))
    assert i4 == dict(key=77, traj=7, therm_left=6, precision=1)
    assert configio.resume_info(os.path.join(e.path, "configs", "12"))["therm_left"] == 0


def test_config_names(tmp_path):
    assert configio.split_config_name("runs/a.mesh/configs/100") == ("runs/a.mesh", "100")
    with pytest.raises(ValueError):
        configio.split_config_name("runs/a.mesh/100")
    e = configio.Ensemble.create(str(tmp_path / "r"))
    e.set_attribute("philox_key", str(2**64 - 1))
    e.set_attribute("philox_traj0", 5)
    assert configio.resume_state(os.path.join(e.path, "configs", "7")) == (2**64 - 1, 12)
    assert configio.resume_state(os.path.join(e.path, "configs", "last")) is None


def test_read_config_checks(tmp_path):
    """readConfig's group and geometry checks, before any device transfer."""
    from types import SimpleNamespace

    import mesh_amd

    e = configio.Ensemble.create(str(tmp_path / "k"))
    e.set_attribute("group", "su3")
    e.set_attribute("geometry", [2, 2, 2, 2])
    configio._save(e.config_path("1"), np.zeros((4, 2, 2, 2, 2, 18)))
    name = os.path.join(e.path, "configs", "1")
    U = SimpleNamespace(grid=mesh_amd.Grid([2, 2, 2, 4]), nf=4, prec=mesh_amd.F64)
    with pytest.raises(mesh_amd.MeshError, match="grid mismatch"):
        configio.read_config(name, U)
    with pytest.raises(mesh_amd.MeshError, match="group mismatch"):
        configio.read_config(name, U, group="u1")
    assert configio.config_geometry(name) == [2, 2, 2, 2]
