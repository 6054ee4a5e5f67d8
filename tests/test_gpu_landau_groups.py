"""Landau gauge fixing for SU(2) and U(1) on the GPU (Landau<SU2<T>>,
Landau<U1<T>>; src/gauge/landau.h, instantiated by
src/tests/lattice_gauge.cpp:46-53) against the oracle: SU(2) derivative,
gauge transformation and overrelaxation half-sweeps bit for bit in both
precisions, functional and condition to 1e-12; U(1) on angles within a few
FP32 ulps (the GPU's and numpy's cos/sin/atan2 differ in the last FP64 bit);
overrelaxation to theta < 1e-12 and the reference's ascent to its stopping
condition on thermalised configurations."""
import math

import numpy as np
import pytest

import mesh_amd
from mesh_amd import F32, F64, gauge
from mesh_amd import su2 as S2
from mesh_amd import u1 as U1m
from oracle import oracle as O
from oracle import u1_oracle as U

pytestmark = pytest.mark.gpu

SHAPES = [(4, 4, 4, 4), (6, 4, 2, 8), (8, 6, 4), (10, 12)]


def _p(prec):
    return "f64" if prec == F64 else "f32"


def _haar_h(shape, seed, prec):
    """a Haar-random gauge transformation: the first direction of random links"""
    return O.su2_random(list(shape), seed, _p(prec))[:1].copy()


def _up(ctx, g, a, prec, nf=None):
    return S2.SU2Field(g, nf, prec, ctx).upload(a)


@pytest.mark.parametrize("prec", [F64, F32])
@pytest.mark.parametrize("shape", SHAPES)
def test_su2_landau_bitexact(ctx, shape, prec):
    sh = list(shape)
    g = mesh_amd.Grid(shape)
    U = O.su2_random(sh, 31, _p(prec))
    h = _haar_h(shape, 32, prec)
    fU, fh = _up(ctx, g, U, prec), _up(ctx, g, h, prec, 1)
    lan = S2.Landau(fU)
    lan.g.upload(h)
    assert lan() == pytest.approx(O.su2_landau_functional(U, h, sh), rel=1e-13)
    assert lan.divsum() == pytest.approx(O.su2_landau_divsum(U, h, sh), rel=1e-12)
    assert lan.condition() == pytest.approx(O.su2_landau_divsum(U, h, sh) / (len(sh) * U.shape[1]), rel=1e-12)
    np.testing.assert_array_equal(lan.deriv().download(), O.su2_landau_deriv(U, h, sh))
    Ug = S2.gauge_transform(fU, fh)
    ref = O.su2_gauge_transform(U, h, sh)
    np.testing.assert_array_equal(Ug.download(), ref)
    if all(s % 2 == 0 for s in shape):
        for parity, omega in ((0, 1.7), (1, 1.0), (0, 1.3)):
            lan._or_half(Ug, parity, omega)
            O.su2_landau_or_half(ref, h, sh, parity, omega)
            np.testing.assert_array_equal(Ug.download(), ref)
            np.testing.assert_array_equal(lan.g.download(), h)
    A, B = O.su2_random(sh, 33, _p(prec)), O.su2_random(sh, 34, _p(prec))
    fA, fB = _up(ctx, g, A, prec), _up(ctx, g, B, prec)
    want = 2 * np.sum(A[..., 0].astype(float) * B[..., 0] - np.sum(A[..., 1:].astype(float) * B[..., 1:], axis=-1))
    assert lan._dot(fA, fB) == pytest.approx(want, rel=1e-12, abs=1e-12)
    lan._axpy(fA, 0.3, fB)
    ty = np.float64 if prec == F64 else np.float32
    if prec == F32:
        # fma(a, x, y) in FP32: the FP64 product is exact, and rounding the
        # FP64 sum to FP32 differs from one rounding only on exact FP32 ties
        want = (np.float64(np.float32(0.3)) * B.astype(np.float64) + A.astype(np.float64)).astype(np.float32)
        np.testing.assert_array_equal(fA.download(), want)
    else:
        np.testing.assert_allclose(fA.download(), A + 0.3 * B, rtol=0, atol=3e-16)


def _thermalised_su2(ctx, shape, beta, prec, n=30):
    h = S2.Hmc(mesh_amd.Grid(shape), prec, seed=41, ctx=ctx)
    h.randomizeGaugeField()
    d = gauge.makeDeltas("2lf", 0.5, 8)
    for _ in range(n):
        h.runHmcUpdate(beta, d, metropolis=False)
    return h.U


@pytest.mark.parametrize("prec", [F64, F32])
def test_su2_landau_run_thermalised(ctx, prec):
    """8^4 at beta 2.3: overrelaxation to theta < 1e-12 (FP64) / 1e-9 (FP32 links),
    the functional above its start, the plaquette of U^g unchanged."""
    U = _thermalised_su2(ctx, (8, 8, 8, 8), 2.3, prec)
    lan = S2.Landau(U)
    F0 = lan()
    it, theta = lan.run(5000, omega=1.7, tol=1e-12 if prec == F64 else 1e-9)
    print(prec, it, theta, F0, lan())
    assert theta < (1e-12 if prec == F64 else 1e-9)
    assert lan() > F0 + 0.05
    assert S2.plaquette(lan.gauge_fixed()) == pytest.approx(S2.plaquette(U), abs=1e-12 if prec == F64 else 2e-6)
    # the transformed links that run() carried agree with U^g rebuilt from g
    np.testing.assert_allclose(lan.Ug.download(), lan.gauge_fixed().download(), atol=1e-12 if prec == F64 else 1e-5)


def test_su2_landau_ascent(ctx):
    """the reference's solver (landau.h:100-136) to its stopping condition on 4^4"""
    U = _thermalised_su2(ctx, (4, 4, 4, 4), 2.3, F64, n=10)
    lan = S2.Landau(U)
    F0 = lan()
    it, cond = lan.run_ascent(20000)
    print(it, cond)
    assert cond < 1e-10 and lan() > F0


# ---- U(1)

def _wd(a, b):
    return np.abs(U.wrap(np.asarray(a, np.float64) - np.asarray(b, np.float64)))


def _u1_thermalised(ctx, dims, beta, n=40):
    ens = U1m.U1Ensemble(None, None, 1, ctx, dims=dims)
    ens.random_start(51)
    d = gauge.makeDeltas("2lf", 0.5, 8)
    for t in range(n):
        ens.hmc(beta, d, 52, t, metropolis=False)
    return ens.angles()[0]


@pytest.mark.parametrize("dims", [(16, 16), (8, 6, 4), (4, 4, 4, 6)])
def test_u1_landau_matches_oracle(ctx, dims):
    g = mesh_amd.Grid(dims)
    rng = np.random.default_rng(53)
    th = rng.uniform(-math.pi, math.pi, size=(len(dims),) + dims).astype(np.float32)
    phi = rng.uniform(-math.pi, math.pi, size=dims).astype(np.float32)
    fU = U1m.U1Field(g, None, ctx).upload(th)
    lan = U1m.Landau(fU)
    lan.g.upload(phi)
    assert lan() == pytest.approx(U.landau_functional(th, phi), rel=1e-12, abs=1e-14)
    assert lan.divsum() == pytest.approx(U.landau_divsum(th, phi), rel=1e-12)
    D = lan.deriv().download()[0]
    np.testing.assert_allclose(D, U.landau_deriv(th, phi), rtol=2e-7, atol=1e-12)
    Ug = lan.gauge_fixed()
    ref = U.gauge_transform(th, phi)
    assert _wd(Ug.download(), ref).max() < 1e-6
    ph = phi.copy()
    for parity, omega in ((0, 1.7), (1, 1.0)):
        lan._or_half(Ug, parity, omega)
        ref, ph = U.landau_or_half(ref, ph, parity, omega)
        assert _wd(Ug.download(), ref).max() < 2e-6
        assert _wd(lan.g.download()[0], ph).max() < 2e-6
    F = U1m.U1Field(g, 1, ctx).upload(rng.normal(size=dims).astype(np.float32))
    lan._step(lan.g, F, 0.7)
    assert _wd(lan.g.download()[0], U.gauge_step(ph, F.download()[0], 0.7)).max() < 1e-6


@pytest.mark.parametrize("dims,beta", [((32, 32), 3.0), ((8, 8, 8, 8), 1.0)])
def test_u1_landau_run_and_ascent(ctx, dims, beta):
    th = _u1_thermalised(ctx, dims, beta)
    fU = U1m.U1Field(mesh_amd.Grid(dims), None, ctx).upload(th)
    lan = U1m.Landau(fU)
    F0 = lan()
    it, theta = lan.run(20000, omega=1.8, tol=1e-12)
    print(dims, it, theta, F0, lan())
    assert theta < 1e-12 and lan() > F0
    lan2 = U1m.Landau(fU)
    it2, cond = lan2.run_ascent(50000)
    print(it2, cond)
    assert cond < 1e-10
    # both solvers find a maximum of the functional: a gauge copy with F close
    # to (at most a Gribov-copy spread away from) the overrelaxation result
    assert lan2() > F0 and abs(lan2() - lan()) < 0.05


def test_u1_landau_full_size_half_sweep(ctx):
    """32^4: a thermalised U(1) field, its gauge transform by random phases and
    two overrelaxation half-sweeps against the numpy oracle (a few FP32 ulps)"""
    dims = (32, 32, 32, 32)
    th = _u1_thermalised(ctx, dims, 1.0, n=10)
    rng = np.random.default_rng(61)
    phi = rng.uniform(-math.pi, math.pi, size=dims).astype(np.float32)
    fU = U1m.U1Field(mesh_amd.Grid(dims), None, ctx).upload(th)
    lan = U1m.Landau(fU)
    lan.g.upload(phi)
    assert lan.divsum() == pytest.approx(U.landau_divsum(th, phi), rel=1e-10)
    Ug = lan.gauge_fixed()
    ref = U.gauge_transform(th, phi)
    ph = phi.copy()
    for parity in (0, 1):
        lan._or_half(Ug, parity, 1.8)
        ref, ph = U.landau_or_half(ref, ph, parity, 1.8)
    assert _wd(Ug.download(), ref).max() < 4e-6
    assert _wd(lan.g.download()[0], ph).max() < 4e-6
