"""C ABI checks that need no GPU: libmesh_hip.so loads, exports every
function include/mesh_hip.h declares, and its host-only entry points agree
with the oracle."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import mesh_amd
from mesh_amd import _lib
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "mesh_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mesh_\w+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    L = C.CDLL(_lib.LIB_PATH)
    names = declared()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert set(declared()) <= set(_lib.exported_symbols())


@pytest.mark.parametrize("scheme,eps,n", [("2lf", 0.5, 10), ("2mn", 0.5, 12), ("4mn", 1.0, 8), ("2mn", 0.3, 1)])
def test_make_deltas_matches_oracle(scheme, eps, n):
    np.testing.assert_array_equal(mesh_amd.gauge.makeDeltas(scheme, eps, n), O.make_deltas(scheme, eps, n))


def test_make_deltas_errors():
    L = _lib.lib()
    buf = (C.c_double * 4)()
    assert L.mesh_make_deltas(b"bogus", 0.5, 4, buf, 4) == _lib.MESH_E_ARG
    assert L.mesh_make_deltas(b"2mn", 0.5, 12, buf, 4) == _lib.MESH_E_SIZE
    assert L.mesh_make_deltas(b"2mn", -1.0, 12, buf, 4) == _lib.MESH_E_ARG


def test_geometry_queries():
    L = _lib.lib()
    g = mesh_amd.Grid((32, 32, 32, 32))
    assert L.mesh_volume(g.ref) == 32 ** 4
    assert L.mesh_site_stride(g.ref) == 32 ** 4
    g = mesh_amd.Grid((3, 4, 5, 2))
    assert L.mesh_site_stride(g.ref) == 256
    b = L.mesh_site_block()
    assert b == 0 or (b & (b - 1) == 0 and 16 <= b <= 256)
    assert L.mesh_su3_field_bytes(g.ref, 4, _lib.F64) == 4 * 9 * 256 * 16
    assert L.mesh_su3_field_bytes(g.ref, 1, _lib.F32) == 9 * 256 * 8
    bad = _lib.mesh_grid()
    bad.nd = 5
    assert L.mesh_volume(C.byref(bad)) == -1
    bad.nd = 2
    bad.dims[0], bad.dims[1] = 1, 4
    assert L.mesh_su3_field_bytes(C.byref(bad), 1, _lib.F64) == 0


def test_invalid_context_is_rejected():
    L = _lib.lib()
    assert L.mesh_synchronize(None) == _lib.MESH_E_STATE
    assert L.mesh_last_error(None) == b"invalid mesh context"
    out = C.c_double()
    g = mesh_amd.Grid((4, 4))
    assert L.mesh_su3_plaquette(None, g.ref, 0, None, C.byref(out)) == _lib.MESH_E_STATE


def test_context_creation_without_gpu_fails_cleanly():
    import torch  # noqa: F401  (only to ask whether a GPU is present)

    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is present")
    h = C.c_void_p()
    assert _lib.lib().mesh_ctx_create(C.byref(h), 0) != 0
    assert not h.value


def test_cpp_programs_built_and_reject_bad_options():
    import subprocess

    exe = os.path.join(ROOT, "mesh_amd", "bin", "mesh_hmc")
    assert os.path.exists(exe) and os.path.exists(os.path.join(ROOT, "mesh_amd", "bin", "test_lattice_gauge"))
    r = subprocess.run([exe, "--bogus"], capture_output=True, text=True, timeout=30)
    assert r.returncode == 2 and "unknown option --bogus" in r.stderr
