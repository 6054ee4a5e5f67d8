"""GPU parity of the SU(3) HMC kernels against the CPU oracle.

FP64: the kernels and the oracle perform the same sequence of IEEE
operations, so link updates, momenta, staples and forces are compared bit for
bit; global sums (different summation order) to 1e-12 relative.  FP32: same
bit-for-bit comparison against the oracle's float instantiation.
"""
import numpy as np
import pytest

import mesh_amd
from mesh_amd import F32, F64, gauge
from conftest import options
from oracle import oracle as O

pytestmark = pytest.mark.gpu

SHAPES = [(4, 4, 4, 4), (6, 4, 2, 8), (8, 8, 8), (3, 4, 5, 2), (6, 10)]


def _field(ctx, shape, prec, host):
    g = mesh_amd.Grid(shape)
    f = mesh_amd.Field(g, host.shape[0], prec, ctx)
    f.upload(host)
    return g, f


def _cdt(prec):
    return np.complex128 if prec == F64 else np.complex64


def _pname(prec):
    return "f64" if prec == F64 else "f32"


@pytest.mark.parametrize("prec", [F64, F32])
@pytest.mark.parametrize("shape", SHAPES)
def test_random_gauge_bitexact(ctx, shape, prec):
    g = mesh_amd.Grid(shape)
    U = mesh_amd.GaugeField(g, prec, ctx)
    gauge.randomGaugeField(U, 1234)
    ref = O.random_gauge(list(shape), 1234, _pname(prec))
    got = U.download()
    assert got.dtype == ref.dtype
    np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("prec", [F64, F32])
def test_upload_download_roundtrip(ctx, prec):
    shape = (3, 4, 5, 2)
    U = O.random_gauge(list(shape), 5, _pname(prec))
    g, f = _field(ctx, shape, prec, U)
    np.testing.assert_array_equal(f.download(), U)


@pytest.mark.parametrize("prec", [F64, F32])
@pytest.mark.parametrize("shape", SHAPES[:3])
def test_momenta_bitexact(ctx, shape, prec):
    g = mesh_amd.Grid(shape)
    P = mesh_amd.GaugeField(g, prec, ctx)
    K = mesh_amd.DeviceArray(g.size(), ctx)
    gauge.randomMomenta(P, 77, 5, K)
    Pr, Kr = O.momenta(list(shape), 77, 5, _pname(prec))
    np.testing.assert_array_equal(P.download(), Pr)
    np.testing.assert_array_equal(K.download(), Kr)


@pytest.mark.parametrize("prec", [F64, F32])
@pytest.mark.parametrize("shape", SHAPES)
def test_plaquette(ctx, shape, prec):
    U = O.random_gauge(list(shape), 3, _pname(prec))
    g, f = _field(ctx, shape, prec, U)
    ref = O.plaquette(U, list(shape))
    got = gauge.plaquette(f)
    assert abs(got - ref) <= 1e-12 * max(abs(ref), 1e-3)
    ff = mesh_amd.GaugeField(g, prec, ctx).fill_unit()
    assert gauge.plaquette(ff) == pytest.approx(1.0, abs=1e-14)
    beta = 6.0
    nP = g.size() * len(shape) * (len(shape) - 1) / 2
    assert gauge.wilsonAction(f, beta) == pytest.approx(beta * nP * (1 - ref), rel=1e-12)


@pytest.mark.parametrize("prec", [F64, F32])
@pytest.mark.parametrize("shape", SHAPES)
def test_staple_bitexact(ctx, shape, prec):
    U = O.random_gauge(list(shape), 11, _pname(prec))
    g, f = _field(ctx, shape, prec, U)
    for mu in range(len(shape)):
        S = gauge.stapleSum(f, mu).download()[0]
        np.testing.assert_array_equal(S, O.staple(U, list(shape), mu))


@pytest.mark.parametrize("shape", SHAPES)
def test_force_bitexact_f64(ctx, shape):
    U = O.random_gauge(list(shape), 12, "f64")
    g, f = _field(ctx, shape, F64, U)
    for mu in range(len(shape)):
        F = gauge.wilsonDeriv(f, mu, 5.7).download()[0]
        np.testing.assert_array_equal(F, O.force(U, list(shape), mu, 5.7))


@pytest.mark.parametrize("prec", [F64, F32])
@pytest.mark.parametrize("shape", SHAPES)
def test_force_update_bitexact(ctx, shape, prec):
    U = O.random_gauge(list(shape), 21, _pname(prec))
    P, _ = O.momenta(list(shape), 21, 0, _pname(prec))
    g, fu = _field(ctx, shape, prec, U)
    _, fp = _field(ctx, shape, prec, P)
    gauge.momentumStep(fu, fp, 6.0, 0.0833)
    np.testing.assert_array_equal(fp.download(), O.force_update(U, P, list(shape), 6.0, 0.0833))


@pytest.mark.parametrize("variant", [0, 1, 2])
@pytest.mark.parametrize("prec", [F64, F32])
@pytest.mark.parametrize("shape", [(4, 4, 2, 32), (2, 3, 4, 16), (2, 2, 2, 64), (3, 2, 6, 32), (4, 4, 4, 4)])
@pytest.mark.parametrize("remap", [0, 1, 2, 3, 4])
def test_force_variants_bitexact(ctx, shape, prec, variant, remap):
    """Every force-kernel variant (wavefront per direction, thread per site,
    LDS tile; tile-eligible shapes included) gives the oracle's bits."""
    U = O.random_gauge(list(shape), 22, _pname(prec))
    P, _ = O.momenta(list(shape), 22, 0, _pname(prec))
    g, fu = _field(ctx, shape, prec, U)
    _, fp = _field(ctx, shape, prec, P)
    with options(ctx, force_variant=variant, xcd_remap=remap):
        gauge.momentumStep(fu, fp, 6.0, 0.0833)
    np.testing.assert_array_equal(fp.download(), O.force_update(U, P, list(shape), 6.0, 0.0833))


# 4-d shapes with V % 64 == 0 whose fastest extent does not divide 64: the
# wavefront-per-direction kernel stages the block but reads x3-neighbours from
# global memory (its "own links" path, with the shared upper-staple products)
OWN_SHAPES = [(2, 4, 4, 6), (4, 2, 8, 12), (4, 4, 4, 3)]
# ... and shapes whose blocks lie within one (x0, x1) plane ((L2 L3) % 64 == 0)
OWN_UNI_SHAPES = [(4, 4, 16, 12), (2, 6, 16, 20)]


@pytest.mark.parametrize("remap", [0, 3])
@pytest.mark.parametrize("variant", [-1, 1])
@pytest.mark.parametrize("prec", [F64, F32])
@pytest.mark.parametrize("shape", OWN_SHAPES + OWN_UNI_SHAPES)
def test_force_own_path_bitexact(ctx, shape, prec, variant, remap):
    U = O.random_gauge(list(shape), 23, _pname(prec))
    P, _ = O.momenta(list(shape), 23, 0, _pname(prec))
    g, fu = _field(ctx, shape, prec, U)
    _, fp = _field(ctx, shape, prec, P)
    with options(ctx, force_variant=variant, xcd_remap=remap):
        gauge.momentumStep(fu, fp, 6.0, 0.0833)
    np.testing.assert_array_equal(fp.download(), O.force_update(U, P, list(shape), 6.0, 0.0833))


@pytest.mark.parametrize("prec", [F64, F32])
@pytest.mark.parametrize("shape", OWN_SHAPES[:2] + OWN_UNI_SHAPES)
def test_fused_trajectory_own_path_bitexact(ctx, shape, prec):
    shape = list(shape)
    hmc = gauge.Hmc(mesh_amd.Grid(shape), prec, seed=8, ctx=ctx)
    hmc.randomizeGaugeField()
    U = O.random_gauge(shape, 8, _pname(prec))
    deltas = gauge.makeDeltas("2mn", 0.4, 2)
    for traj in range(2):
        r = hmc.runHmcUpdate(5.5, deltas, metropolis=True)
        o = O.hmc_trajectory(U, shape, 5.5, deltas, 8, traj, metropolis=True)
        assert r["accepted"] == o["accepted"]
        np.testing.assert_array_equal(hmc.U.download(), U)


@pytest.mark.parametrize("order", [3, 4])
@pytest.mark.parametrize("tile", [1, 2, 4])
@pytest.mark.parametrize("variant", [0, 2])
@pytest.mark.parametrize("prec", [F64, F32])
@pytest.mark.parametrize("shape", [(8, 2, 4, 32), (16, 2, 8, 16), (8, 3, 8, 32), (4, 4, 4, 32)])
def test_slab_order_bitexact(ctx, shape, prec, variant, tile, order):
    """Block orders 3 (x0 slabs per XCD) and 4 (x0 quarters x x1 halves per
    XCD), with row-block tiles of every size that divides the row-blocks:
    force update bit-exact, plaquette to 1e-12."""
    U = O.random_gauge(list(shape), 23, _pname(prec))
    P, _ = O.momenta(list(shape), 23, 0, _pname(prec))
    g, fu = _field(ctx, shape, prec, U)
    _, fp = _field(ctx, shape, prec, P)
    with options(ctx, force_variant=variant, xcd_remap=order, rb_tile=tile):
        gauge.momentumStep(fu, fp, 6.0, 0.0833)
        pl = gauge.plaquette(fu)
    np.testing.assert_array_equal(fp.download(), O.force_update(U, P, list(shape), 6.0, 0.0833))
    ref = O.plaquette(U, list(shape))
    assert abs(pl - ref) <= 1e-12 * max(abs(ref), 1e-3)


@pytest.mark.parametrize("order", [3, 4])
def test_slab_order_trajectory(ctx, order):
    """A fused trajectory in block orders 3 and 4 gives the oracle's field bits."""
    shape = [8, 2, 4, 32]
    with options(ctx, xcd_remap=order, rb_tile=2):
        hmc = gauge.Hmc(mesh_amd.Grid(shape), F64, seed=7, ctx=ctx)
        hmc.randomizeGaugeField()
        U = O.random_gauge(shape, 7, "f64")
        d = gauge.makeDeltas("2mn", 0.5, 2)
        r = hmc.runHmcUpdate(6.0, d, metropolis=True)
    o = O.hmc_trajectory(U, shape, 6.0, d, 7, 0, metropolis=True)
    assert r["accepted"] == o["accepted"]
    assert abs(r["dH"] - o["dH"]) <= 1e-12 * max(1.0, abs(o["dH"]))
    np.testing.assert_array_equal(hmc.U.download(), U)


@pytest.mark.parametrize("prec", [F64, F32])
@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("reunit", [False, True])
@pytest.mark.parametrize("t", [0.04, 0.7, 3.0])
def test_link_update_bitexact(ctx, shape, prec, reunit, t):
    U = O.random_gauge(list(shape), 31, _pname(prec))
    P, _ = O.momenta(list(shape), 31, 2, _pname(prec))
    g, fu = _field(ctx, shape, prec, U)
    _, fp = _field(ctx, shape, prec, P)
    out = mesh_amd.GaugeField(g, prec, ctx)
    gauge.linkStep(out, fu, fp, t, reunit)
    np.testing.assert_array_equal(out.download(), O.link_update(U, P, list(shape), t, reunit))
    # in place gives the same
    gauge.linkStep(fu, fu, fp, t, reunit)
    np.testing.assert_array_equal(fu.download(), out.download())


@pytest.mark.parametrize("prec", [F64, F32])
def test_reunit_bitexact(ctx, prec):
    shape = (4, 4, 2, 2)
    U = O.random_gauge(list(shape), 41, _pname(prec))
    rng = np.random.default_rng(0)
    U = (U + 1e-3 * (rng.normal(size=U.shape) + 1j * rng.normal(size=U.shape))).astype(U.dtype)
    g, f = _field(ctx, shape, prec, U)
    gauge.reunitize(f)
    np.testing.assert_array_equal(f.download(), O.reunit_field(U))


@pytest.mark.parametrize("prec", [F64, F32])
@pytest.mark.parametrize("shape", SHAPES[:3])
def test_delta_h(ctx, shape, prec):
    sh = list(shape)
    U0 = O.random_gauge(sh, 51, _pname(prec))
    P, K = O.momenta(sh, 51, 1, _pname(prec))
    U1 = O.link_update(U0, P, sh, 0.05, True)
    P1 = O.force_update(U1, P, sh, 6.0, 0.1)
    ref = O.delta_h(U0, U1, P1, K, sh, 6.0)
    g, fu0 = _field(ctx, shape, prec, U0)
    _, fu1 = _field(ctx, shape, prec, U1)
    _, fp = _field(ctx, shape, prec, P1)
    Kd = mesh_amd.DeviceArray(g.size(), ctx).upload(K)
    got = gauge.deltaH(fu0, fu1, fp, Kd, 6.0)
    scale = np.abs(K).sum() + 6.0 * g.size()
    assert abs(got["dH"] - ref[0]) <= 1e-13 * scale
    for i, k in enumerate(["ps_old", "ps_new", "K_old", "K_new"]):
        assert got[k] == pytest.approx(ref[i + 1], rel=1e-12, abs=1e-10)


@pytest.mark.parametrize("prec", [F64, F32])
def test_hmc_trajectories_match_oracle(ctx, prec):
    """C1-style run: 4^4 (8^4 in the bench), seeded Haar start, 2lf eps=0.5 n=10;
    three thermalisation trajectories without Metropolis, then five with."""
    shape = [4, 4, 4, 4]
    g = mesh_amd.Grid(shape)
    hmc = gauge.Hmc(g, prec, seed=1, ctx=ctx)
    hmc.randomizeGaugeField()
    U = O.random_gauge(shape, 1, _pname(prec))
    deltas = gauge.makeDeltas("2lf", 0.5, 10)
    np.testing.assert_array_equal(deltas, O.make_deltas("2lf", 0.5, 10))
    for traj in range(8):
        metro = traj >= 3
        r = hmc.runHmcUpdate(6.0, deltas, metropolis=metro)
        o = O.hmc_trajectory(U, shape, 6.0, deltas, 1, traj, metropolis=metro)
        # identical fields, different summation order: |dH| agrees to ~1e-12
        assert abs(r["dH"] - o["dH"]) <= 1e-12 * max(1.0, abs(o["dH"])), (traj, r, o)
        assert r["accepted"] == o["accepted"]
        assert abs(r["plaq"] - o["plaq"]) <= 1e-13
        np.testing.assert_array_equal(hmc.U.download(), U)


@pytest.mark.parametrize("variant", [-1, 0, 2])
@pytest.mark.parametrize("fuse,compact", [(0, 1), (1, 1), (1, 0)])
@pytest.mark.parametrize("scheme,n", [("2lf", 3), ("2mn", 2), ("4mn", 1), ("2lf", 4)])
@pytest.mark.parametrize("prec", [F64, F32])
def test_hmc_fused_md_bitexact(ctx, prec, scheme, n, fuse, compact, variant):
    """The fused momentum+link step and the separate kernels give the oracle's
    bits for every integrator; odd and even numbers of fused pairs exercise
    both parities of the field ping-pong, with both force kernels and with the
    momenta in the compact (9 reals) and the matrix (18 reals) layout."""
    shape = [4, 4, 6, 4]
    g = mesh_amd.Grid(shape)
    ctx.set_option("fuse_md", fuse)
    ctx.set_option("compact_momenta", compact)
    ctx.set_option("force_variant", variant)
    try:
        hmc = gauge.Hmc(g, prec, seed=5, ctx=ctx)
        hmc.randomizeGaugeField()
        U = O.random_gauge(shape, 5, _pname(prec))
        deltas = gauge.makeDeltas(scheme, 0.4, n)
        for traj in range(2):
            r = hmc.runHmcUpdate(5.5, deltas, metropolis=True)
            o = O.hmc_trajectory(U, shape, 5.5, deltas, 5, traj, metropolis=True)
            assert r["accepted"] == o["accepted"]
            np.testing.assert_array_equal(hmc.U.download(), U)
    finally:
        ctx.set_option("fuse_md", 1)
        ctx.set_option("compact_momenta", 1)
        ctx.set_option("force_variant", -1)


@pytest.mark.parametrize("prec", [F64, F32])
def test_compact_momenta_same_bits_16(ctx, prec):
    """At 16^4 (2mn, 12 substeps, three trajectories from a Haar start) the
    trajectory with momenta in 9 reals per link and the one with the 18-real
    matrix layout give the same fields, Delta H and decisions bit for bit."""
    shape = [16, 16, 16, 16]
    g = mesh_amd.Grid(shape)
    d = gauge.makeDeltas("2mn", 0.5, 12)
    runs = {}
    for compact in (1, 0):
        with options(ctx, compact_momenta=compact):
            h = gauge.Hmc(g, prec, seed=4, ctx=ctx)
            h.randomizeGaugeField()
            rs = [h.runHmcUpdate(6.0, d, metropolis=t > 0) for t in range(3)]
            runs[compact] = ([(r["dH"], r["accepted"], r["plaq"]) for r in rs], h.U.download())
    assert runs[1][0] == runs[0][0]
    np.testing.assert_array_equal(runs[1][1], runs[0][1])


def test_f32_vs_f64_same_config(ctx):
    """Mixed precision against FP64 on one seeded trajectory."""
    shape = [4, 4, 4, 4]
    g = mesh_amd.Grid(shape)
    U64 = mesh_amd.GaugeField(g, F64, ctx)
    gauge.randomGaugeField(U64, 9)
    res = {}
    for prec in (F64, F32):
        h = gauge.Hmc(g, prec, seed=9, ctx=ctx)
        h.U = U64.convert(prec)
        r = h.runHmcUpdate(6.0, gauge.makeDeltas("2mn", 0.5, 12), metropolis=False)
        res[prec] = (r, gauge.plaquette(h.U))
    assert abs(res[F32][1] - res[F64][1]) < 1e-5
    assert abs(res[F32][0]["dH"] - res[F64][0]["dH"]) < 1e-2 * (1 + abs(res[F64][0]["dH"]))


def _cold_thermalised(ctx, shape, ntherm, seed=2):
    """FP64 chain from the unit field, ntherm trajectories without Metropolis
    (beta 6, 2mn, 12 substeps)"""
    g = mesh_amd.Grid(shape)
    h = gauge.Hmc(g, F64, seed=seed, ctx=ctx)
    h.U.fill_unit()
    d = gauge.makeDeltas("2mn", 0.5, 12)
    for _ in range(ntherm):
        h.runHmcUpdate(6.0, d, metropolis=False)
    return g, h.U, d


@pytest.mark.parametrize("L,draws", [(16, 6), (32, 2)])
def test_f32_vs_f64_delta_h_thermalised(ctx, L, draws):
    """FP32-mixed trajectories against FP64 ones from the same thermalised
    configuration and the same momenta (beta 6, 2mn, 12 substeps): Delta H
    differs by rounding noise only, with no term proportional to the volume
    (the earlier FP32 link step and reunitarisation gave -1.0e-6 per site,
    -1.02 at 32^4; a force rounded to FP32 before the momentum update gives
    +4.6e-8 per site), and the plaquettes after the trajectory agree to
    better than 1e-6."""
    shape = [L] * 4
    g, U, d = _cold_thermalised(ctx, shape, 40)
    V = g.size()
    diffs = []
    for k in range(draws):
        res = {}
        for prec in (F64, F32):
            h = gauge.Hmc(g, prec, seed=100 + k, ctx=ctx)
            h.U = U.convert(prec)
            r = h.runHmcUpdate(6.0, d, metropolis=False)
            res[prec] = (r["dH"], r["plaq"])
        print(L, k, res[F64], res[F32][0] - res[F64][0], flush=True)
        assert abs(res[F32][1] - res[F64][1]) < 1e-6
        diffs.append(res[F32][0] - res[F64][0])
    assert abs(np.mean(diffs)) / V < 1e-8, diffs


def test_ensemble_f32_vs_f64_16(ctx):
    """Thermalised 16^4 at beta 6 (2mn, 12 substeps), 300 Metropolis
    trajectories in each precision: acceptance, <exp(-dH)> (= 1, Creutz;
    src/gauge/hmc.h:98-99), <dH> and the mean plaquette of the FP32-mixed
    chain agree with the FP64 chain's."""
    import math

    g, U, d = _cold_thermalised(ctx, [16] * 4, 40)
    res = {}
    for prec in (F64, F32):
        h = gauge.Hmc(g, prec, seed=7, ctx=ctx)
        h.U = U.convert(prec)
        rows = [h.runHmcUpdate(6.0, d) for _ in range(300)]
        res[prec] = dict(plaq=[r["plaq"] for r in rows], acc=[float(r["accepted"]) for r in rows],
                         dH=np.array([r["dH"] for r in rows]))
    a64, a32 = np.mean(res[F64]["acc"]), np.mean(res[F32]["acc"])
    ea = math.sqrt(a64 * (1 - a64) / 300 + a32 * (1 - a32) / 300) + 1e-3
    print("acceptance", a64, a32, "dH", res[F64]["dH"].mean(), res[F32]["dH"].mean(), flush=True)
    assert abs(a64 - a32) < 4 * ea, (a64, a32)
    for prec in (F64, F32):
        c, ec = _binned(np.exp(-res[prec]["dH"]))
        print("creutz", prec, c, ec, flush=True)
        assert abs(c - 1) < 5 * ec + 0.01, (prec, c, ec)
    m64, e64 = _binned(res[F64]["dH"])
    m32, e32 = _binned(res[F32]["dH"])
    assert abs(m64 - m32) < 5 * math.hypot(e64, e32) + 0.01, (m64, m32)
    p64, ep64 = _binned(res[F64]["plaq"])
    p32, ep32 = _binned(res[F32]["plaq"])
    assert abs(p64 - p32) < 5 * math.hypot(ep64, ep32) + 1e-5, (p64, p32)


@pytest.mark.parametrize("prec", [F64, F32])
def test_c2_fused_trajectory_32(ctx, prec):
    """The benchmarked path at full size: one 2mn, 12-substep trajectory of
    the default fused kernels with compact momenta (k_momenta_link, 24 x the
    fused momentum + link step, the last with reunitarisation, Delta H) from
    the 32^4 configuration thermalised by 50 trajectories, against the
    oracle: the new field bit for bit, Delta H to 1e-12 relative."""
    shape = [32, 32, 32, 32]
    g = mesh_amd.Grid(shape)
    hmc = gauge.Hmc(g, F64, seed=2, ctx=ctx)
    hmc.randomizeGaugeField()
    d = gauge.makeDeltas("2mn", 0.5, 12)
    for _ in range(50):
        hmc.runHmcUpdate(6.0, d, metropolis=False)
    h = gauge.Hmc(g, prec, seed=2, ctx=ctx)
    h.U = hmc.U if prec == F64 else hmc.U.convert(prec)
    h.traj = 50
    U = h.U.download()
    r = h.runHmcUpdate(6.0, d, metropolis=False)
    o = O.hmc_trajectory(U, shape, 6.0, d, 2, 50, metropolis=False)
    np.testing.assert_array_equal(h.U.download(), U)
    assert abs(r["dH"] - o["dH"]) <= 1e-12 * max(1.0, abs(o["dH"])), (r["dH"], o["dH"])


@pytest.mark.parametrize("prec", [F64, F32])
def test_fused_step_48(ctx, prec):
    """48^4 (the larger size of the mixed-precision workload; the fastest
    extent does not divide 64, so the own-links staple path): a trajectory
    of one leapfrog substep runs the momenta with the first link step and one
    fused momentum + link step with reunitarisation; the field bit for bit
    and Delta H to 1e-12 relative against the oracle."""
    shape = [48, 48, 48, 48]
    g = mesh_amd.Grid(shape)
    h = gauge.Hmc(g, prec, seed=5, ctx=ctx)
    h.randomizeGaugeField()
    d = gauge.makeDeltas("2lf", 0.02, 1)
    U = h.U.download()
    r = h.runHmcUpdate(6.0, d, metropolis=False)
    o = O.hmc_trajectory(U, shape, 6.0, d, 5, 0, metropolis=False)
    np.testing.assert_array_equal(h.U.download(), U)
    assert abs(r["dH"] - o["dH"]) <= 1e-12 * max(1.0, abs(o["dH"])), (r["dH"], o["dH"])


def test_errors(ctx):
    g = mesh_amd.Grid((4, 4, 4, 4))
    U = mesh_amd.GaugeField(g, F64, ctx)
    with pytest.raises(mesh_amd.MeshError):
        gauge.stapleSum(U, 7)
    with pytest.raises(mesh_amd.MeshError):
        gauge.makeDeltas("3xx", 1.0, 4)
    with pytest.raises(ValueError):
        gauge.runHmd(U, U, 6.0, [0.1, 0.2])


def test_nan_momenta_rejects(ctx):
    """A NaN Delta H is rejected (u < exp(NaN) is false, src/gauge/hmc.cpp:93)
    and the field is left bit-unchanged."""
    shape = [4, 4, 4, 4]
    g = mesh_amd.Grid(shape)
    hmc = gauge.Hmc(g, F64, seed=3, ctx=ctx)
    hmc.randomizeGaugeField()
    before = hmc.U.download()
    r = hmc.runHmcUpdate(6.0, np.array([np.nan, 0.1, np.nan]), metropolis=True)
    assert not r["accepted"]
    np.testing.assert_array_equal(hmc.U.download(), before)


def test_c1_workload_matches_oracle(ctx):
    """The small reference workload (8^4, beta 6, FP64, Haar start seed 1,
    2lf eps 0.5 x 10, 20 Metropolis trajectories): every trajectory's
    accept decision, plaquette and Delta H against the oracle, and the final
    field bit for bit."""
    shape = [8, 8, 8, 8]
    hmc = gauge.Hmc(mesh_amd.Grid(shape), F64, seed=1, ctx=ctx)
    hmc.randomizeGaugeField()
    U = O.random_gauge(shape, 1, "f64")
    d = gauge.makeDeltas("2lf", 0.5, 10)
    for traj in range(20):
        r = hmc.runHmcUpdate(6.0, d, metropolis=True)
        o = O.hmc_trajectory(U, shape, 6.0, d, 1, traj, metropolis=True)
        assert r["accepted"] == o["accepted"], traj
        assert abs(r["dH"] - o["dH"]) <= 1e-12 * max(1.0, abs(o["dH"])), (traj, r["dH"], o["dH"])
        assert abs(r["plaq"] - o["plaq"]) <= 1e-13
    np.testing.assert_array_equal(hmc.U.download(), U)


@pytest.mark.parametrize("prec", [F64, F32])
def test_c2_force_parity_32(ctx, prec):
    """Full-size force parity: the momentum step on the 32^4 configuration
    thermalised by 50 HMC trajectories from the seed-2 Haar start, against
    the oracle element by element (bit-exact; the required bound is 1e-12
    relative)."""
    shape = [32, 32, 32, 32]
    g = mesh_amd.Grid(shape)
    hmc = gauge.Hmc(g, F64, seed=2, ctx=ctx)
    hmc.randomizeGaugeField()
    d = gauge.makeDeltas("2mn", 0.5, 12)
    for _ in range(50):
        hmc.runHmcUpdate(6.0, d, metropolis=False)
    U = hmc.U if prec == F64 else hmc.U.convert(prec)
    P = mesh_amd.GaugeField(g, prec, ctx)
    gauge.randomMomenta(P, 2, 99, None)
    Uh, Ph = U.download(), P.download()
    gauge.momentumStep(U, P, 6.0, 1.0 / 12)
    ref = O.force_update(Uh, Ph, shape, 6.0, 1.0 / 12)
    np.testing.assert_array_equal(P.download(), ref)


@pytest.mark.parametrize("prec", [F64, F32])
@pytest.mark.parametrize("shape", [(4, 3, 6, 2), (2, 5, 8)])
def test_cshift_and_sum(ctx, shape, prec):
    """cshift (src/lattice/lattice.h:366-389) against numpy roll, bit for bit,
    for every direction and offsets of both signs; lattice_sum against numpy."""
    sh = list(shape)
    U = O.random_gauge(sh, 61, _pname(prec))
    g, f = _field(ctx, sh, prec, U)
    V = O.volume(sh)
    host = U.reshape((len(sh),) + tuple(sh) + (3, 3))
    for d in range(len(sh)):
        for off in (1, -1, sh[d] - 1):
            if off == 0 or abs(off) >= sh[d]:
                continue
            out = gauge.cshift(f, d, off).download()
            ref = np.roll(host, -off, axis=1 + d).reshape(len(sh), V, 3, 3)
            np.testing.assert_array_equal(out, ref)
    s = gauge.lattice_sum(f)
    ref = U.astype(np.complex128).sum(axis=(0, 1))
    # FP64 sums of O(1) entries: a few ulp per addition at most
    assert np.abs(s - ref).max() < 1e-12 * V * len(sh)
    with pytest.raises(mesh_amd.MeshError):
        gauge.cshift(f, len(sh), 1)
    with pytest.raises(mesh_amd.MeshError):
        gauge.cshift(f, 0, 0)
    with pytest.raises(mesh_amd.MeshError):
        gauge.cshift(f, 0, sh[0])


def _binned(x, nbin=10):
    """mean and error of the mean from nbin consecutive blocks (autocorrelation)."""
    x = np.asarray(x, np.float64)
    b = x[: len(x) // nbin * nbin].reshape(nbin, -1).mean(axis=1)
    return x.mean(), b.std(ddof=1) / np.sqrt(nbin)


def test_ensemble_f32_vs_f64_creutz():
    """Ensemble statistics on 6^4 at beta 5.7 (2mn, eps 0.5, 6 substeps): the
    FP32-mixed chain's mean plaquette and acceptance agree with the FP64
    chain's, <exp(-dH)> = 1 (Creutz; src/gauge/hmc.h:98-99) and the acceptance
    follows erfc(sqrt(<dH>)/2)."""
    import math

    ctx = mesh_amd.default_context()
    g = mesh_amd.Grid([6, 6, 6, 6])
    deltas = gauge.makeDeltas("2mn", 0.5, 6)
    res = {}
    for prec in (F64, F32):
        hmc = gauge.Hmc(g, prec, seed=3, ctx=ctx)
        hmc.randomizeGaugeField()
        for _ in range(60):
            hmc.runHmcUpdate(5.7, deltas, metropolis=False)
        rows = [hmc.runHmcUpdate(5.7, deltas) for _ in range(400)]
        res[prec] = dict(plaq=[r["plaq"] for r in rows], acc=[float(r["accepted"]) for r in rows],
                         dH=[r["dH"] for r in rows])
    p64, e64 = _binned(res[F64]["plaq"])
    p32, e32 = _binned(res[F32]["plaq"])
    print("plaquette", p64, e64, p32, e32)
    assert abs(p64 - p32) < 5 * math.hypot(e64, e32) + 1e-4
    a64, a32 = np.mean(res[F64]["acc"]), np.mean(res[F32]["acc"])
    assert 0.5 < a64 < 1.0 and abs(a64 - a32) < 0.1, (a64, a32)
    for prec in (F64, F32):
        c, ec = _binned(np.exp(-np.asarray(res[prec]["dH"])))
        print("creutz", prec, c, ec)
        assert abs(c - 1) < 5 * ec + 0.02, (prec, c, ec)
    dh = np.mean(res[F64]["dH"])
    assert abs(a64 - math.erfc(math.sqrt(max(dh, 0.0)) / 2)) < 0.1, (a64, dh)


def test_integrator_order():
    """rms dH over eight momentum draws from one 8^4 configuration, short
    trajectories (eps 0.02, tau 0.2), per halving of the step: slope 2 for
    leapfrog and 4 for 4mn (src/gauge/hmc.cpp:7-51).  FP64: at these steps
    dH is below the FP32 storage resolution of the energies."""
    import math

    ctx = mesh_amd.default_context()
    g = mesh_amd.Grid([8, 8, 8, 8])
    hmc = gauge.Hmc(g, F64, seed=4, ctx=ctx)
    hmc.randomizeGaugeField()
    for _ in range(30):
        hmc.runHmcUpdate(6.0, gauge.makeDeltas("2mn", 0.5, 6), metropolis=False)
    start = mesh_amd.GaugeField(g, F64, ctx).copy_from(hmc.U)

    def dh(scheme, n):
        out = []
        for t in range(8):
            hmc.U.copy_from(start)
            hmc.traj = 1000 + t  # the same eight momentum draws for every step size
            out.append(hmc.runHmcUpdate(6.0, gauge.makeDeltas(scheme, 0.02, n), metropolis=False)["dH"])
        return math.sqrt(np.mean(np.square(out)))

    slope = lambda scheme, n: math.log2(dh(scheme, n) / dh(scheme, 2 * n))
    s = slope("2lf", 8)
    assert 1.58 < s < 2.33, s
    s = slope("4mn", 4)
    assert 3.3 < s < 4.5, s


@pytest.mark.parametrize("shape", [(4, 4, 4, 4), (6, 4, 2, 8), (5, 3, 7), (16, 16, 8, 8)])
def test_float_energy_bitexact(ctx, shape):
    """mesh_su3_float_energy (the reference's float-build sums: FP32 pair
    trees of 256 and a sequential FP32 sum of the blocks) against the oracle,
    bit for bit, including a lattice smaller than one block."""
    sh = list(shape)
    g = mesh_amd.Grid(shape)
    U = mesh_amd.GaugeField(g, F32, ctx)
    gauge.randomGaugeField(U, 61)
    P = mesh_amd.GaugeField(g, F32, ctx)
    gauge.randomMomenta(P, 62, 0)
    S, K = gauge.float_energy(U, P, 5.5)
    So, Ko = O.su3_float_energy(U.download(), P.download(), sh, 5.5)
    assert (S, K) == (So, Ko)
    assert gauge.float_energy(U, None, 5.5) == (So, 0.0)


def test_float_sums_trajectory_energies(ctx):
    """option float_sums: H_old and H_new of an FP32 trajectory are the
    float-build energies of (U, P_initial) and (U_new, P_final), Delta H their
    difference; the fused (compact momenta) and unfused paths give the same bits."""
    g, U64, d = _cold_thermalised(ctx, [8, 8, 8, 8], 10)
    res = {}
    for fuse in (0, 1):
        h = gauge.Hmc(g, F32, seed=13, ctx=ctx)
        h.U = U64.convert(F32)
        U0 = h.U.download()
        with options(ctx, float_sums=1, fuse_md=fuse):
            r = h.runHmcUpdate(6.0, d, metropolis=False)
        res[fuse] = (r, h.U.download())
        if fuse == 0:
            P0 = mesh_amd.GaugeField(g, F32, ctx)
            gauge.randomMomenta(P0, 13, 0)
            So, Ko = O.su3_float_energy(U0, P0.download(), [8] * 4, 6.0)
            Sn, Kn = gauge.float_energy(h.U, h.P, 6.0)
            assert r["H_old"] == So + Ko and r["H_new"] == Sn + Kn
            assert r["dH"] == (Sn + Kn) - (So + Ko)
    assert res[0][0]["dH"] == res[1][0]["dH"]
    np.testing.assert_array_equal(res[0][1], res[1][1])


def test_float_sums_error_grows_with_volume(ctx):
    """The reference's float build against FP64 on the same thermalised
    configurations: its Delta H error grows with the volume and is far above
    that of the library's FP32-mixed trajectories (FP64 per-site differenced
    sums)."""
    err = {}
    for L in (8, 16):
        g, U64, d = _cold_thermalised(ctx, [L] * 4, 15)
        e = {"fsum": [], "mixed": []}
        for traj in range(3):
            dh = {}
            for key, prec, fs in (("f64", F64, 0), ("mixed", F32, 0), ("fsum", F32, 1)):
                h = gauge.Hmc(g, prec, seed=17, ctx=ctx)
                h.traj = traj
                h.U = U64.convert(prec)
                with options(ctx, float_sums=fs):
                    dh[key] = h.runHmcUpdate(6.0, d, metropolis=False)["dH"]
            e["fsum"].append(dh["fsum"] - dh["f64"])
            e["mixed"].append(dh["mixed"] - dh["f64"])
        err[L] = {k: float(np.sqrt(np.mean(np.square(v)))) for k, v in e.items()}
        print(L, err[L])
    assert err[16]["fsum"] > 10 * err[16]["mixed"]
    assert err[16]["fsum"] > 4 * err[8]["fsum"]


def test_ensemble_f32_vs_f64_32(ctx):
    """32^4 at beta 6 (2mn, 20 substeps: acceptance ~0.6), 400 Metropolis
    trajectories in each precision from the same thermalised field:
    acceptance, <exp(-dH)> (= 1) and the mean plaquette of the FP32-mixed chain
    agree with the FP64 chain's within their binned errors."""
    import math

    g, U, _ = _cold_thermalised(ctx, [32] * 4, 60)
    d = gauge.makeDeltas("2mn", 0.5, 20)
    res = {}
    for prec in (F64, F32):
        h = gauge.Hmc(g, prec, seed=11, ctx=ctx)
        h.U = U.convert(prec)
        rows = [h.runHmcUpdate(6.0, d) for _ in range(400)]
        res[prec] = dict(plaq=[r["plaq"] for r in rows], acc=[float(r["accepted"]) for r in rows],
                         dH=np.array([r["dH"] for r in rows]))
    a64, a32 = np.mean(res[F64]["acc"]), np.mean(res[F32]["acc"])
    ea = math.sqrt(a64 * (1 - a64) / 400 + a32 * (1 - a32) / 400) + 1e-3
    print("acceptance", a64, a32, "dH", res[F64]["dH"].mean(), res[F32]["dH"].mean(), flush=True)
    assert abs(a64 - a32) < 4 * ea, (a64, a32)
    for prec in (F64, F32):
        c, ec = _binned(np.exp(-res[prec]["dH"]))
        print("creutz", prec, c, ec, flush=True)
        assert abs(c - 1) < 5 * ec + 0.02, (prec, c, ec)
    p64, ep64 = _binned(res[F64]["plaq"])
    p32, ep32 = _binned(res[F32]["plaq"])
    print("plaquette", p64, ep64, p32, ep32, flush=True)
    assert abs(p64 - p32) < 5 * math.hypot(ep64, ep32) + 1e-6, (p64, p32)
