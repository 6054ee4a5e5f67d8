"""The no-allocation plaquette entry point against mesh_su3_plaquette: the
same value bit for bit, through the C ABI, on two calls in a row with the
second on a smaller lattice (the shared scratch keeps the first call's
partial sums), in both precisions; lattices beyond its fixed scratch are
refused."""
import ctypes as C

import numpy as np
import pytest

import mesh_amd
from mesh_amd import F32, F64, gauge
from mesh_amd.lattice import DeviceArray

pytestmark = pytest.mark.gpu


def _noalloc(ctx, g, U):
    out = DeviceArray(1, ctx)
    out.upload(np.array([np.nan]))
    ctx.call("mesh_su3_plaquette_noalloc", g.ref, U.prec, U.ptr, out.ptr)
    return out.download()[0]


@pytest.mark.parametrize("prec", [F64, F32])
def test_plaquette_noalloc_matches_plaquette(ctx, prec):
    for shape, seed in (([16, 16, 16, 16], 11), ([8, 6, 4, 4], 12)):
        g = mesh_amd.Grid(shape)
        U = mesh_amd.GaugeField(g, prec, ctx)
        gauge.randomGaugeField(U, seed)
        ref = gauge.plaquette(U)
        got = _noalloc(ctx, g, U)
        assert got == ref, (shape, got, ref)


def test_plaquette_noalloc_unit_field(ctx):
    g = mesh_amd.Grid([4, 4, 4, 4])
    U = mesh_amd.GaugeField(g, F64, ctx)
    U.fill_unit()
    assert _noalloc(ctx, g, U) == 1.0


def test_plaquette_noalloc_refuses_large_lattices(ctx):
    small = mesh_amd.GaugeField(mesh_amd.Grid([4, 4, 4, 4]), F64, ctx)
    big = mesh_amd.Grid([66, 66, 66, 66])  # 18,974,736 sites > 2^24
    out = DeviceArray(1, ctx)
    with pytest.raises(mesh_amd.MeshError) as e:
        ctx.call("mesh_su3_plaquette_noalloc", big.ref, F64, small.ptr, out.ptr)
    assert e.value.code == mesh_amd._lib.MESH_E_SIZE
    with pytest.raises(mesh_amd.MeshError):
        ctx.call("mesh_su3_plaquette_noalloc", small.grid.ref, F64, small.ptr, None)
