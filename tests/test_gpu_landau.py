"""GPU parity of the Landau gauge kernels against the CPU oracle."""
import ctypes as C

import numpy as np
import pytest

import mesh_amd
from mesh_amd import F32, F64, gauge
from oracle import oracle as O

pytestmark = pytest.mark.gpu

SHAPES = [(4, 4, 4, 4), (6, 4, 2, 8), (8, 8, 8), (6, 10)]


def _pname(prec):
    return "f64" if prec == F64 else "f32"


def _up(ctx, shape, prec, host):
    g = mesh_amd.Grid(shape)
    f = mesh_amd.Field(g, host.shape[0], prec, ctx)
    f.upload(host)
    return g, f


@pytest.mark.parametrize("prec", [F64, F32])
@pytest.mark.parametrize("shape", SHAPES)
def test_functional_and_divsum(ctx, shape, prec):
    sh = list(shape)
    U = O.random_gauge(sh, 21, _pname(prec))
    h = O.random_transform(sh, 22)
    hp = h.astype(U.dtype)
    g, fu = _up(ctx, shape, prec, U)
    _, fh = _up(ctx, shape, prec, hp)
    lan = gauge.Landau(fu)
    assert lan() == pytest.approx(O.landau_functional(U, sh), rel=1e-12, abs=1e-14)  # g = 1
    lan.g.upload(hp)
    assert lan() == pytest.approx(O.landau_functional(U, sh, hp), rel=1e-12, abs=1e-14)
    assert lan.divsum() == pytest.approx(O.landau_divsum(U, sh, hp), rel=1e-12)
    assert lan.condition() == pytest.approx(O.landau_divsum(U, sh, hp) / (len(sh) * g.size()), rel=1e-12)


@pytest.mark.parametrize("shape", SHAPES)
def test_deriv_f64(ctx, shape):
    sh = list(shape)
    U = O.random_gauge(sh, 31)
    h = O.random_transform(sh, 32)
    g, fu = _up(ctx, shape, F64, U)
    lan = gauge.Landau(fu)
    lan.g.upload(h)
    np.testing.assert_allclose(lan.deriv().download()[0], O.landau_deriv(U, h, sh), rtol=0, atol=1e-17)


@pytest.mark.parametrize("prec", [F64, F32])
@pytest.mark.parametrize("shape", SHAPES + [(8, 2, 16, 64), (16, 2, 8, 32)])
@pytest.mark.parametrize("track", [False, True])
def test_or_halfsweep_bitexact(ctx, shape, prec, track):
    """both half-sweep kernels (256-site spans in block order when the fast
    extent divides 256 and 256 divides V, else one thread per parity site)"""
    sh = list(shape)
    U = O.random_gauge(sh, 41, _pname(prec))
    h = np.zeros((1,) + U.shape[1:], U.dtype)
    h[:, :] = np.eye(3)
    g, fu = _up(ctx, shape, prec, U)
    _, fh = _up(ctx, shape, prec, h)
    for sweep in range(3):
        for parity in (0, 1):
            ctx.call("mesh_landau_or_halfsweep", g.ref, prec, fu.ptr, fh.ptr if track else None, parity,
                     C.c_double(1.7))
            O.landau_or_halfsweep(U, sh, parity, 1.7, h if track else None)
    np.testing.assert_array_equal(fu.download(), U)
    if track:
        np.testing.assert_array_equal(fh.download(), h)


def _thermalised(ctx, L, ntherm, seed=2, haar=True):
    g = mesh_amd.Grid([L] * 4)
    h = gauge.Hmc(g, F64, seed=seed, ctx=ctx)
    if haar:
        h.randomizeGaugeField()
    else:
        h.U.fill_unit()
    d = gauge.makeDeltas("2mn", 0.5, 12)
    for _ in range(ntherm):
        h.runHmcUpdate(6.0, d, metropolis=False)
    return g, h.U


@pytest.mark.parametrize("prec", [F64, F32])
def test_or_halfsweep_bitexact_32(ctx, prec):
    """Full size (the Landau workload's lattice): an even and an odd
    overrelaxation half-sweep with the gauge transformation tracked, on the
    32^4 configuration thermalised by 50 trajectories from the seed-2 Haar
    start, against the oracle bit for bit."""
    sh = [32, 32, 32, 32]
    g, U64 = _thermalised(ctx, 32, 50)
    fu = U64 if prec == F64 else U64.convert(prec)
    U = fu.download()
    h = np.zeros((1,) + U.shape[1:], U.dtype)
    h[:, :] = np.eye(3)
    fh = mesh_amd.Field(g, 1, prec, ctx)
    fh.upload(h)
    for parity in (0, 1):
        ctx.call("mesh_landau_or_halfsweep", g.ref, prec, fu.ptr, fh.ptr, parity, C.c_double(1.7))
        O.landau_or_halfsweep(U, sh, parity, 1.7, h)
    np.testing.assert_array_equal(fu.download(), U)
    np.testing.assert_array_equal(fh.download(), h)


def test_run_thermalised_16_both_precisions(ctx):
    """Landau gauge of a thermalised 16^4 configuration (cold start, 40
    trajectories at beta 6): FP64 and FP32 overrelaxation (omega 1.7) both
    reach theta < 1e-12; their functionals agree to 1e-6 and their sweep
    counts within 10%; the plaquette of the fixed field is unchanged."""
    g, U64 = _thermalised(ctx, 16, 40, haar=False)
    res = {}
    for prec in (F64, F32):
        U = U64 if prec == F64 else U64.convert(prec)
        lan = gauge.Landau(U)
        f0 = lan()
        it, theta = lan.run(20000, omega=1.7, tol=1e-12)
        Ug = mesh_amd.GaugeField(g, prec, ctx)
        ctx.call("mesh_su3_gauge_transform", g.ref, prec, Ug.ptr, U.ptr, lan.g.ptr)
        res[prec] = (it, theta, lan(), f0, gauge.plaquette(U), gauge.plaquette(Ug))
        print(prec, res[prec], flush=True)
        assert theta < 1e-12, (prec, it, theta)
        assert lan() > f0
    (i64, _, F64v, _, p64, pg64), (i32, _, F32v, _, p32, pg32) = res[F64], res[F32]
    assert abs(F64v - F32v) < 1e-6, (F64v, F32v)
    assert abs(i64 - i32) <= 0.1 * i64, (i64, i32)
    assert abs(pg64 - p64) < 1e-12 and abs(pg32 - p32) < 1e-6


def test_gauge_transform_bitexact(ctx):
    sh = [4, 6, 2, 4]
    U = O.random_gauge(sh, 51)
    h = O.random_transform(sh, 52)
    g, fu = _up(ctx, sh, F64, U)
    _, fh = _up(ctx, sh, F64, h)
    out = mesh_amd.GaugeField(g, F64, ctx)
    ctx.call("mesh_su3_gauge_transform", g.ref, F64, out.ptr, fu.ptr, fh.ptr)
    np.testing.assert_array_equal(out.download(), O.gauge_transform(U, h, sh))
    assert gauge.plaquette(out) == pytest.approx(gauge.plaquette(fu), abs=1e-14)


def test_or_odd_extent_rejected(ctx):
    g = mesh_amd.Grid((4, 3, 4, 4))
    U = mesh_amd.GaugeField(g, F64, ctx)
    with pytest.raises(mesh_amd.MeshError):
        ctx.call("mesh_landau_or_halfsweep", g.ref, F64, U.ptr, None, 0, C.c_double(1.7))


@pytest.mark.parametrize("prec", [F64, F32])
def test_run_pure_gauge_converges(ctx, prec):
    """U = h(x) h(x+mu)^+ is a gauge copy of the unit field: the solver must
    reach F = 1 and theta ~ 0, and U^g must equal the transformed field."""
    sh = [8, 8, 8, 8]
    g = mesh_amd.Grid(sh)
    one = mesh_amd.GaugeField(g, prec, ctx).fill_unit()
    h = mesh_amd.Lattice(g, prec, ctx)
    gauge.randomGaugeField(h, 5, stream=O.TAG_GAUGE)
    U = mesh_amd.GaugeField(g, prec, ctx)
    ctx.call("mesh_su3_gauge_transform", g.ref, prec, U.ptr, one.ptr, h.ptr)
    lan = gauge.Landau(U)
    tol = 1e-24 if prec == F64 else 1e-10
    it, theta = lan.run(2000, omega=1.7, tol=tol)
    assert theta < tol, (it, theta)
    assert lan() > 1 - (1e-12 if prec == F64 else 1e-6)


def test_landau_finite_difference_gpu(ctx):
    """The reference test's derivative check (src/tests/lattice_gauge.cpp:31-40)
    through the GPU entry points: functional, deriv, exp_mul."""
    sh = [4, 4, 4, 4]
    g = mesh_amd.Grid(sh)
    U = mesh_amd.GaugeField(g, F64, ctx)
    gauge.randomGaugeField(U, 61)
    lan = gauge.Landau(U)
    gauge.randomGaugeField(lan.g, 62, stream=O.TAG_GAUGE)
    f = mesh_amd.Lattice(g, F64, ctx)
    P = mesh_amd.GaugeField(g, F64, ctx)
    gauge.randomMomenta(P, 63, 0)
    f.copy_from(P[0])
    D = lan.deriv()
    eps = 1e-6
    expected = -2 * eps * gauge.sum_real_trace_prod(f, D)
    old = lan()
    gauge.evolve(lan.g, f, eps)
    new = lan()
    assert abs((new - old) / expected - 1) < 1e-4


@pytest.mark.parametrize("prec", [F64, F32])
def test_add_mul_bitexact(ctx, prec):
    """lhs += a * rhs (mesh_su3_axpy) against the oracle's y + (x a)."""
    sh = [4, 2, 6, 2]
    x = O.random_gauge(sh, 31, _pname(prec))
    y = O.random_gauge(sh, 32, _pname(prec))
    _, fx = _up(ctx, sh, prec, x)
    _, fy = _up(ctx, sh, prec, y)
    a = 0.73125
    gauge.add_mul(fy, a, fx)
    if prec == F64:
        np.testing.assert_array_equal(fy.download(), O.axpy(y, a, x))
    else:
        ref = y + (x * np.float32(a)).astype(np.complex64)
        np.testing.assert_array_equal(fy.download(), ref.astype(np.complex64))


@pytest.mark.parametrize("case", ["pure_gauge", "thermalised"])
def test_ascent_against_oracle(ctx, case):
    """The reference's solver on the GPU (Landau.run_ascent over the
    functional, derivative, exp(eps F) g, add_mul and reunitize entry points)
    against its CPU restatement on the same field: the same iterates to
    rounding level for the first iterations, and on a gauge copy of the unit
    field both reach the closed-form maximum F = 1."""
    sh = [4, 4, 4, 4]
    V = O.volume(sh)
    if case == "pure_gauge":
        h = O.random_transform(sh, 10)
        Uu = np.zeros((4, V, 3, 3), np.complex128)
        Uu[:, :] = np.eye(3)
        Uh = O.gauge_transform(Uu, h, sh)
        n_iter = 200
    else:
        Uh = O.random_gauge(sh, 33)
        for t in range(3):
            O.hmc_trajectory(Uh, sh, 6.0, O.make_deltas("2lf", 0.5, 10), 33, t, metropolis=False)
        n_iter = 30
    _, U = _up(ctx, sh, F64, Uh)
    lan = gauge.Landau(U)
    it, cond = lan.run_ascent(n_iter)
    _, hist = O.landau_ascent(Uh, sh, n_iter)
    k = min(len(hist), len(lan.history), 25)
    for i in range(k):
        c_g, f_g, e_g = lan.history[i]
        c_o, f_o, e_o = hist[i]
        assert abs(f_g - f_o) <= 1e-9 * abs(f_o), (i, f_g, f_o)
        assert abs(c_g - c_o) <= 1e-6 * abs(c_o) + 1e-14, (i, c_g, c_o)
        assert abs(e_g - e_o) <= 1e-6 * abs(e_o), (i, e_g, e_o)
    if case == "pure_gauge":
        assert cond < 1e-10 and hist[-1][0] < 1e-10, (cond, hist[-1])
        assert lan() > 1 - 1e-10
