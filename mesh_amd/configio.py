"""Ensemble and configuration files, the format of mesh/io.h.

The reference keeps an ensemble in one HDF5 file: attributes group / beta /
geometry / hmc_* / markov_* (src/apps/hmc.cpp:65-86), a dataset
"/configs/<n>" of shape [Nd, L0, ..., L(nd-1), ncomp] per configuration
(writeToFile, src/lattice/lattice.h:205-233) and the histories
(hmc.cpp:119-124); readConfig opens "file.h5/configs/<n>" and checks group and
geometry (src/gauge/utils.h:188-219).  Without HDF5 an ensemble here is a
directory holding the same things:

    <dir>/attributes.json    flat JSON object: the reference's attribute names
                             plus precision, philox_key (decimal string of the
                             64-bit Philox key) and philox_traj0 (the counter
                             of the chain's first trajectory)
    <dir>/configs/<n>.npy    SU(3): [Nd, L0.., 18] reals, each link's 3x3
                             complex matrix row-major with re/im interleaved,
                             FP64 or FP32 as the field; U(1): [Nd, L0.., 1]
                             FP32 angles
    <dir>/<history>.npy      one-dimensional FP64 arrays

Configuration n of a chain continues bit for bit from Philox counter
philox_traj0 + n.
"""
import json
import os
import shutil

import numpy as np

from ._lib import F64, MESH_E_ARG, MESH_E_SIZE, MeshError

_ATTR = "attributes.json"


class Ensemble:
    """An ensemble directory (the reference's util::Hdf5File of an HMC run)."""

    def __init__(self, path, attributes):
        self.path = str(path)
        self.attributes = attributes

    @classmethod
    def create(cls, path, allow_overwrite=False):
        """New ensemble (Hdf5File::create(filename, allow_overwrite)) in a
        missing or empty directory, or -- with allow_overwrite -- in place of
        an existing ensemble (a directory holding attributes.json), whose
        configurations and data files are removed first.  Anything else is
        refused, so that a stray path cannot be emptied."""
        path = str(path)
        if os.path.exists(path) and not (os.path.isdir(path) and not os.listdir(path)):
            if not allow_overwrite:
                raise FileExistsError(f"ensemble '{path}' exists (allow_overwrite to replace it)")
            if not os.path.isdir(path) or not os.path.isfile(os.path.join(path, _ATTR)):
                raise FileExistsError(f"'{path}' is not an ensemble (no {_ATTR}); not overwritten")
            shutil.rmtree(os.path.join(path, "configs"), ignore_errors=True)
            for f in os.listdir(path):
                if f.endswith(".npy") and os.path.isfile(os.path.join(path, f)):
                    os.remove(os.path.join(path, f))
        cfg = os.path.join(path, "configs")
        os.makedirs(cfg, exist_ok=True)
        e = cls(path, {})
        e._flush()
        return e

    @classmethod
    def open(cls, path):
        with open(os.path.join(str(path), _ATTR)) as f:
            return cls(path, json.load(f))

    def _flush(self):
        tmp = os.path.join(self.path, _ATTR + ".tmp")
        with open(tmp, "w") as f:
            json.dump(dict(sorted(self.attributes.items())), f, indent=2)
            f.write("\n")
        os.replace(tmp, os.path.join(self.path, _ATTR))

    def set_attribute(self, name, value):
        if isinstance(value, (tuple, list)):
            value = [int(v) for v in value]
        self.attributes[name] = value
        self._flush()

    def config_path(self, name):
        return os.path.join(self.path, "configs", f"{name}.npy")

    def write_data(self, name, values):
        _save(os.path.join(self.path, f"{name}.npy"), np.asarray(values, np.float64).reshape(-1))

    def read_data(self, name):
        a = np.load(os.path.join(self.path, f"{name}.npy"), allow_pickle=False)
        if a.dtype != np.float64 or a.ndim != 1:
            raise ValueError(f"'{name}' is not a one-dimensional FP64 array")
        return a


def _save(path, a):
    tmp = path + ".tmp"
    with open(tmp, "wb") as f:
        np.lib.format.write_array(f, np.ascontiguousarray(a), version=(1, 0), allow_pickle=False)
    os.replace(tmp, path)


def split_config_name(name):
    """"dir/configs/<n>" -> (ensemble directory, configuration name)."""
    head, n = os.path.split(os.path.normpath(str(name)))
    ens, cfg = os.path.split(head)
    if cfg != "configs" or not n:
        raise ValueError(f"unknown config name '{name}' (expected <ensemble>/configs/<n>)")
    return ens, n


def _shape(grid, ncomp):
    return (grid.ndim(),) + tuple(grid.shape) + (ncomp,)


def write_config(ens, name, U):
    """Store gauge field U (nf = Nd matrices per site) as configuration `name`."""
    if U.nf != U.grid.ndim():
        raise ValueError("a configuration holds one matrix per site and direction")
    h = U.download()
    real = np.float64 if U.prec == F64 else np.float32
    _save(ens.config_path(name), h.view(real).reshape(_shape(U.grid, 18)))


def config_geometry(config_name):
    return list(Ensemble.open(split_config_name(config_name)[0]).attributes["geometry"])


def read_config(config_name, U, group="su3"):
    """Load "dir/configs/<n>" into U after checking the group, the geometry and
    the array's shape (readConfig, src/gauge/utils.h:188-219).  FP64 files load
    into FP32 fields rounded to nearest, FP32 files into FP64 fields exactly."""
    ens_dir, n = split_config_name(config_name)
    e = Ensemble.open(ens_dir)
    at = e.attributes
    if "group" in at and at["group"] != group:
        raise MeshError(MESH_E_ARG, f"group mismatch on load. Expected {group}, got {at['group']}")
    if "geometry" in at and list(at["geometry"]) != list(U.grid.shape):
        raise MeshError(MESH_E_SIZE, f"grid mismatch on load ('{config_name}')")
    a = np.load(e.config_path(n), allow_pickle=False)
    if a.shape != _shape(U.grid, 18) or a.dtype not in (np.float64, np.float32):
        raise MeshError(MESH_E_SIZE, f"shape mismatch on load ('{config_name}')")
    real = np.float64 if U.prec == F64 else np.float32
    c = np.ascontiguousarray(a, real).view(U.cdtype).reshape(U.nf, U.grid.size(), 3, 3)
    U.upload(c)
    return U


def resume_info(config_name):
    """What continuing the chain that wrote "dir/configs/<n>" needs: dict of
    key and traj (the Philox key and trajectory counter; None when the
    ensemble does not record them), therm_left (thermalisation trajectories
    of that chain still to run without Metropolis) and precision (1 or 2, or
    None)."""
    ens_dir, n = split_config_name(config_name)
    at = Ensemble.open(ens_dir).attributes
    st = resume_state(config_name)
    left = max(0, int(at["markov_discard"]) - int(n)) if "markov_discard" in at and n.isdigit() else 0
    return dict(key=None if st is None else st[0], traj=None if st is None else st[1], therm_left=left,
                precision=at.get("precision"))


def resume_state(config_name):
    """(philox key, trajectory counter) that continue the chain which wrote
    "dir/configs/<n>", or None when the ensemble does not record them."""
    ens_dir, n = split_config_name(config_name)
    at = Ensemble.open(ens_dir).attributes
    if "philox_key" not in at or not n.isdigit():
        return None
    return int(at["philox_key"]), int(at.get("philox_traj0", 0)) + int(n)
