"""SU(2) HMC on the GPU against the oracle (bit for bit in both precisions:
random links, momenta, momentum and link steps, whole trajectories with
Delta H to 1e-12) and ensembles against the exact 2-d plaquette
I_2(beta)/I_1(beta), Creutz <exp(-dH)> = 1 and 4-d strong coupling."""
import math

import numpy as np
import pytest
import scipy.special as sp

import mesh_amd
from mesh_amd import F32, F64, gauge
from mesh_amd import su2 as S2
from oracle import oracle as O

pytestmark = pytest.mark.gpu

SHAPES = [(4, 4, 4, 4), (6, 4, 2, 8), (8, 6, 4), (10, 12)]


def _p(prec):
    return "f64" if prec == F64 else "f32"


def _field(ctx, shape, prec, host=None, nf=None):
    g = mesh_amd.Grid(shape)
    f = S2.SU2Field(g, nf, prec, ctx)
    if host is not None:
        f.upload(host)
    return g, f


@pytest.mark.parametrize("prec", [F64, F32])
@pytest.mark.parametrize("shape", SHAPES)
def test_su2_random_bitexact(ctx, shape, prec):
    g, f = _field(ctx, shape, prec)
    f.random(17)
    np.testing.assert_array_equal(f.download(), O.su2_random(list(shape), 17, _p(prec)))


@pytest.mark.parametrize("prec", [F64, F32])
@pytest.mark.parametrize("shape", SHAPES)
def test_su2_steps_bitexact(ctx, shape, prec):
    sh = list(shape)
    U = O.su2_random(sh, 18, _p(prec))
    P, K = O.su2_momenta(sh, 19, 3, _p(prec))
    g, fu = _field(ctx, shape, prec, U)
    _, fp = _field(ctx, shape, prec)
    Kd = mesh_amd.DeviceArray(g.size(), ctx)
    S2.momenta(fp, 19, 3, Kd)
    np.testing.assert_array_equal(fp.download(), P)
    np.testing.assert_array_equal(Kd.download(), K)
    S2.force_update(fu, fp, 2.2, 0.1)
    P1 = O.su2_force_update(U, P, sh, 2.2, 0.1)
    np.testing.assert_array_equal(fp.download(), P1)
    for t, reunit in ((0.05, False), (2.7, True)):
        _, fo = _field(ctx, shape, prec)
        S2.link_update(fo, fu, fp, t, reunit)
        np.testing.assert_array_equal(fo.download(), O.su2_link_update(U, P1, sh, t, reunit))
    assert S2.plaquette(fu) == pytest.approx(O.su2_plaquette(U, sh), rel=1e-13)


@pytest.mark.parametrize("prec", [F64, F32])
@pytest.mark.parametrize("scheme,n", [("2lf", 6), ("2mn", 3), ("4mn", 1)])
def test_su2_trajectories_match_oracle(ctx, prec, scheme, n):
    sh = [4, 4, 4, 4]
    h = S2.Hmc(mesh_amd.Grid(sh), prec, seed=21, ctx=ctx)
    h.randomizeGaugeField()
    U = O.su2_random(sh, 21, _p(prec))
    d = gauge.makeDeltas(scheme, 0.5, n)
    for traj in range(6):
        r = h.runHmcUpdate(2.0, d, metropolis=traj > 1)
        o = O.su2_hmc_trajectory(U, sh, 2.0, d, 21, traj, metropolis=traj > 1)
        assert r["accepted"] == o["accepted"]
        assert abs(r["dH"] - o["dH"]) <= 1e-12 * max(1.0, abs(o["dH"])), (traj, r["dH"], o["dH"])
        assert r["plaq"] == pytest.approx(o["plaq"], rel=1e-13)
    np.testing.assert_array_equal(h.U.download(), U)


@pytest.mark.parametrize("prec", [F64, F32])
@pytest.mark.parametrize("beta", [1.0, 3.0])
def test_su2_2d_exact_plaquette_and_creutz(ctx, prec, beta):
    """2-d 16x16: <Re tr P / 2> = I_2(beta)/I_1(beta) (finite-volume terms
    (I_2/I_1)^256 negligible), <exp(-dH)> = 1, from one chain of 3000
    trajectories after 200 of thermalisation."""
    g = mesh_amd.Grid([16, 16])
    h = S2.Hmc(g, prec, seed=23 + int(beta), ctx=ctx)
    h.randomizeGaugeField()
    d = gauge.makeDeltas("2lf", 0.5, 12)
    for _ in range(200):
        h.runHmcUpdate(beta, d)
    pl, ex = [], []
    for _ in range(3000):
        r = h.runHmcUpdate(beta, d)
        pl.append(r["plaq"])
        ex.append(math.exp(-r["dH"]))
    b = np.array(pl).reshape(30, -1).mean(axis=1)
    est, err = b.mean(), b.std(ddof=1) / math.sqrt(30)
    exact = sp.iv(2, beta) / sp.iv(1, beta)
    print(prec, beta, est, err, exact, np.mean(h.accept_history))
    assert abs(est - exact) < 5 * err + 2e-4, (est, exact, err)
    be = np.array(ex).reshape(30, -1).mean(axis=1)
    assert abs(be.mean() - 1) < 5 * be.std(ddof=1) / math.sqrt(30) + 0.01


def test_su2_4d_strong_coupling(ctx):
    """4-d 6^4 at beta 0.5: <Re tr P / 2> = I_2(beta)/I_1(beta) + O(beta^5)
    (strong-coupling expansion; beta^5 / 4^5 ~ 3e-5 here)."""
    beta = 0.5
    h = S2.Hmc(mesh_amd.Grid([6, 6, 6, 6]), F64, seed=29, ctx=ctx)
    h.randomizeGaugeField()
    d = gauge.makeDeltas("2lf", 0.5, 8)
    for _ in range(50):
        h.runHmcUpdate(beta, d)
    pl = [h.runHmcUpdate(beta, d)["plaq"] for _ in range(400)]
    b = np.array(pl).reshape(20, -1).mean(axis=1)
    est, err = b.mean(), b.std(ddof=1) / math.sqrt(20)
    exact = sp.iv(2, beta) / sp.iv(1, beta)
    assert abs(est - exact) < 5 * err + 1e-4, (est, exact, err)


def test_su2_errors(ctx):
    g = mesh_amd.Grid([4, 4])
    f = S2.SU2Field(g, None, F64, ctx)
    with pytest.raises(ValueError):
        f.upload(np.zeros(10))
    h = S2.Hmc(g, F64, seed=1, ctx=ctx)
    with pytest.raises(mesh_amd.MeshError):
        h.runHmcUpdate(2.0, [0.1, 0.2])  # even number of steps


@pytest.mark.parametrize("prec", [F64, F32])
@pytest.mark.parametrize("nsub", [1, 2, 5])
def test_su2_fused_trajectory_same_bits(ctx, prec, nsub):
    """each momentum step fused with the next link step (the default) gives the
    unfused trajectory's bits, for both parities of the field ping-pong"""
    from conftest import options

    sh = (4, 4, 2, 6)
    d = gauge.makeDeltas("2mn", 0.5, nsub)
    res = []
    for fuse in (0, 1):
        h = S2.Hmc(mesh_amd.Grid(sh), prec, seed=31, ctx=ctx)
        h.randomizeGaugeField()
        with options(ctx, fuse_md=fuse):
            rs = [h.runHmcUpdate(2.1, d, metropolis=t > 0) for t in range(3)]
        res.append(([r["dH"] for r in rs], h.U.download()))
    assert res[0][0] == res[1][0]
    np.testing.assert_array_equal(res[0][1], res[1][1])


def test_su2_misaligned_field_refused(ctx):
    g = mesh_amd.Grid([4, 4])
    f = S2.SU2Field(g, None, F64, ctx)
    import ctypes as C

    bad = C.c_void_p(f.ptr.value + 8)
    with pytest.raises(mesh_amd.MeshError):
        ctx.call("mesh_su2_unit", g.ref, 2, F64, bad)


@pytest.mark.parametrize("prec", [F64, F32])
def test_su2_full_size_trajectory_bitexact(ctx, prec):
    """32^4: one 2mn, 12-substep trajectory of the default fused path (momentum
    steps fused with the link steps, ping-pong through the spare field) from a
    field thermalised by 5 trajectories, against the oracle: the links bit for
    bit, Delta H to 1e-12."""
    sh = [32, 32, 32, 32]
    O.set_threads(O.max_threads())
    d = gauge.makeDeltas("2mn", 0.5, 12)
    h = S2.Hmc(mesh_amd.Grid(sh), prec, seed=37, ctx=ctx)
    h.randomizeGaugeField()
    U = O.su2_random(sh, 37, _p(prec))
    for t in range(6):
        r = h.runHmcUpdate(2.3, d, metropolis=False)
        o = O.su2_hmc_trajectory(U, sh, 2.3, d, 37, t, metropolis=False)
        print(t, r["dH"], o["dH"], flush=True)
        assert abs(r["dH"] - o["dH"]) <= 1e-12 * max(1.0, abs(o["dH"]))
    np.testing.assert_array_equal(h.U.download(), U)
