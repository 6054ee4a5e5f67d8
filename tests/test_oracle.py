"""Pins the CPU oracle (oracle/) without trusting its own formulas:
published known-answer vectors, scipy/numpy fixtures (tests/golden), closed
forms, invariants and finite differences.  CPU only."""
import math
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "su3_golden.npz"))


# ---------------------------------------------------------------- random numbers

def test_philox_known_answers():
    """Philox4x32-10 known-answer vectors of the Random123 distribution
    (kat_vectors: philox4x32_10 with zero, all-ones and pi-digit inputs)."""
    assert list(O.philox([0, 0, 0, 0], [0, 0])) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert list(O.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2)) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6,
                                                                   0x6D5451FD]
    assert list(O.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0])) == [
        0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


@pytest.mark.parametrize("x", [2.0 ** -53, 1e-300, 1e-16, 1e-3, 0.1, 0.5, 0.70710678, 0.9999999, 1.0, 1.5, 7.0])
def test_det_log(x):
    assert abs(O.det_log(x) - math.log(x)) <= 4e-16 * max(1.0, abs(math.log(x)))


def test_det_sincos_grid():
    for u in np.linspace(0, 1, 1001):
        s, c = O.det_sincos2pi(u)
        # libm's value carries the rounding of 2*pi*u itself: a few ulp
        assert abs(s - math.sin(2 * math.pi * u)) < 1e-15
        assert abs(c - math.cos(2 * math.pi * u)) < 1e-15
        assert abs(s * s + c * c - 1) < 4e-16


def test_normals_moments():
    z = np.concatenate([O.link_normals(11, O.TAG_MOM, 0, site, 0, 8) for site in range(25000)])
    n = z.size
    assert abs(z.mean()) < 5 / math.sqrt(n)
    assert abs(z.var() - 1) < 5 * math.sqrt(2 / n)
    assert abs((z ** 4).mean() - 3) < 5 * math.sqrt(96 / n)
    # different streams differ
    assert not np.array_equal(O.link_normals(11, O.TAG_MOM, 0, 0, 0, 8), O.link_normals(11, O.TAG_MOM, 1, 0, 0, 8))


# ---------------------------------------------------------------- step tables

def test_make_deltas_known_answers():
    d = O.make_deltas("2lf", 0.5, 10)
    assert len(d) == 21
    np.testing.assert_allclose(d, [0.05] + [0.1] * 19 + [0.05], rtol=0, atol=1e-15)
    for scheme, n, ln in (("2mn", 12, 49), ("4mn", 8, 65), ("2lf", 1, 3)):
        d = O.make_deltas(scheme, 0.5, n)
        assert len(d) == ln
        # each half (position, momentum) integrates to tau = sqrt(2 eps)
        assert abs(d[0::2].sum() - 1.0) < 1e-14 and abs(d[1::2].sum() - 1.0) < 1e-14
    with pytest.raises(ValueError):
        O.make_deltas("3xx", 0.5, 2)


def test_make_deltas_2mn_lambda_closed_form():
    """The 2mn table is (lambda, 1/2, 1 - 2 lambda, 1/2, lambda) tau with the
    minimum-norm lambda of Omelyan, Mryglod and Folk (2003):
    lambda = 1/2 - c/12 + 1/(6c), c = (2 sqrt(326) + 36)^(1/3)."""
    c = (2 * math.sqrt(326) + 36) ** (1 / 3)
    lam = 0.5 - c / 12 + 1 / (6 * c)
    d = O.make_deltas("2mn", 0.5, 1)  # tau = sqrt(2 eps) = 1
    np.testing.assert_allclose(d, [lam, 0.5, 1 - 2 * lam, 0.5, lam], rtol=0, atol=1e-15)
    d = O.make_deltas("2mn", 2.0, 4)  # tau = 2, four merged substeps
    h = 2.0 / 4
    assert d[0] == pytest.approx(lam * h, abs=1e-15) and d[2] == pytest.approx((1 - 2 * lam) * h, abs=1e-15)
    assert d[4] == pytest.approx(2 * lam * h, abs=1e-15)


# ---------------------------------------------------------------- element algebra

def test_exp_against_scipy_expm():
    for X, E in zip(GOLD["exp_X"], GOLD["exp_E"]):
        got = O.su3_exp(X)
        scale = max(1.0, np.abs(X).max())
        assert np.abs(got - E).max() < 2e-15 * scale ** 2
        assert np.abs(got @ got.conj().T - np.eye(3)).max() < 4e-15 * scale
        assert abs(O.su3_det(got) - 1) < 4e-15 * scale


def test_exp_f32():
    for X, E in zip(GOLD["exp_X"][:24], GOLD["exp_E"][:24]):
        got = O.su3_exp(X.astype(np.complex64))
        assert np.abs(got - E).max() < 1e-6


def test_exp_small_argument_terms():
    """The series stops after fewer terms where ||X||_F <= 0.2: against scipy's
    expm on both sides of that switch, as accurate as the full series."""
    import scipy.linalg as sl

    rng = np.random.default_rng(5)
    for r in (0.01, 0.1, 0.1999, 0.2001, 0.5, 0.99):
        for _ in range(50):
            M = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
            X = (M - M.conj().T) / 2
            X -= np.trace(X) / 3 * np.eye(3)
            X *= r / np.linalg.norm(X)
            E = sl.expm(X)
            assert np.abs(O.su3_exp(X) - E).max() < 5e-16
            assert np.abs(O.su3_exp(X.astype(np.complex64)) - E).max() < 2.5e-7


def test_ta_closed_form():
    rng = np.random.default_rng(1)
    for _ in range(20):
        M = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
        T = O.su3_ta(M)
        ref = (M - M.conj().T) / 2
        ref -= np.trace(ref) / 3 * np.eye(3)
        assert np.abs(T - ref).max() < 1e-15
        assert np.abs(O.su3_ta(T) - T).max() < 1e-15  # idempotent on the algebra (to rounding)
        assert abs(np.trace(T)) < 1e-15
        assert np.abs(T + T.conj().T).max() == 0.0


def test_reunit_against_numpy_and_convergence():
    for a, r in zip(GOLD["reunit_in"], GOLD["reunit_out"]):
        got = O.su3_reunit(a)
        assert np.abs(got - r).max() < 1e-14
        e0 = np.abs(a @ a.conj().T - np.eye(3)).max()
        e1 = np.abs(got @ got.conj().T - np.eye(3)).max()
        assert e1 < 10 * e0 ** 2 + 1e-15  # second-order projection


def test_haar_elements():
    rng = np.random.default_rng(2)
    tr = []
    for _ in range(3000):
        H = O.haar_from_normals(rng.normal(size=18))
        assert np.abs(H @ H.conj().T - np.eye(3)).max() < 2e-15
        assert abs(O.su3_det(H) - 1) < 2e-15
        tr.append(np.trace(H))
    tr = np.array(tr)
    # Haar moments on SU(3): E[tr U] = 0, E[|tr U|^2] = 1
    assert abs(tr.mean()) < 5 / math.sqrt(len(tr))
    assert abs((abs(tr) ** 2).mean() - 1) < 5 * math.sqrt(2 / len(tr))


def test_algebra_normalisation():
    """Random algebra fields: ||P||^2 per link averages dim/2 = 4 and
    sum Re tr(P P) = -||P||^2 (the printed ratios of the reference test,
    src/tests/lattice_gauge.cpp:23-29, asserted)."""
    dims = [4, 4, 4, 4]
    P, K = O.momenta(dims, 5, 0)
    V, nd = O.volume(dims), len(dims)
    n2 = np.sum(np.abs(P) ** 2)
    tol = 5 * math.sqrt(2 / (V * nd * 8))  # 5 sigma of a chi^2 mean
    assert abs(n2 / (V * nd * 8) / 0.5 - 1) < tol
    rtr = np.einsum("mxij,mxji->", P, P).real
    assert abs(rtr / (V * nd * 8) / (-0.5) - 1) < tol
    np.testing.assert_allclose(K.sum(), n2, rtol=1e-13)
    assert np.abs(P + np.conj(np.swapaxes(P, -1, -2))).max() == 0.0
    assert np.abs(np.trace(P, axis1=2, axis2=3)).max() < 1e-15


# ---------------------------------------------------------------- lattice ops

@pytest.mark.parametrize("name", ["l4", "l2"])
def test_plaquette_and_staple_against_numpy_roll(name):
    shape = list(GOLD[f"{name}_shape"])
    U = np.ascontiguousarray(GOLD[f"{name}_U"])
    np.testing.assert_allclose(O.plaq_sites(U, shape), GOLD[f"{name}_plaq_sites"], rtol=0, atol=1e-13)
    for mu in range(len(shape)):
        np.testing.assert_allclose(O.staple(U, shape, mu), GOLD[f"{name}_staple"][mu], rtol=0, atol=1e-13)


def test_plaquette_closed_forms():
    dims = [4, 4, 4, 4]
    U = np.zeros((4, O.volume(dims), 3, 3), np.complex128)
    U[:, :] = np.eye(3)
    assert O.plaquette(U, dims) == 1.0
    U = O.random_gauge(dims, 1)
    h = O.random_transform(dims, 3)
    Ug = O.gauge_transform(U, h, dims)
    np.testing.assert_allclose(O.plaq_sites(Ug, dims), O.plaq_sites(U, dims), rtol=0, atol=1e-13)
    # Haar links: <Re tr P / 3> = 0 +- 1/sqrt(6 V * 18)
    assert abs(O.plaquette(U, dims)) < 5 / math.sqrt(6 * O.volume(dims) * 18)


def test_force_finite_difference():
    """dS/deps of S(exp(eps X) U) at eps = 0 equals -2 Re tr(X F) with
    F = (beta/6) TA(U S) (the pattern of src/tests/lattice_gauge.cpp:31-40)."""
    dims = [3, 4, 2, 4]
    beta = 5.5
    U = O.random_gauge(dims, 4)
    V, nd = O.volume(dims), len(dims)
    rng = np.random.default_rng(9)
    S0 = lambda W: beta * (V * nd * (nd - 1) / 2 - O.plaq_sites(W, dims).sum() / 3)
    for mu, x in ((0, 5), (2, 17), (3, V - 1)):
        X = O.algebra_from_normals(rng.normal(size=8))
        F = O.force(U, dims, mu, beta)[x]
        eps = 1e-5
        Up, Um = U.copy(), U.copy()
        Up[mu, x] = O.su3_exp(eps * X) @ U[mu, x]
        Um[mu, x] = O.su3_exp(-eps * X) @ U[mu, x]
        fd = (S0(Up) - S0(Um)) / (2 * eps)
        an = -2 * np.trace(X @ F).real
        assert abs(fd - an) < 1e-6 * max(1.0, abs(an))


def _trajectory(U, P, dims, beta, deltas):
    U, P = U.copy(), P.copy()
    for i, d in enumerate(deltas):
        if i % 2 == 0:
            U = O.link_update(U, P, dims, d)
        else:
            P = O.force_update(U, P, dims, beta, d)
    return U, P


def test_md_order_and_reversibility():
    dims = [4, 4, 4, 4]
    beta = 6.0
    U0 = O.random_gauge(dims, 6)
    # a smoother start: a few MD steps from the Haar field without Metropolis
    for t in range(3):
        O.hmc_trajectory(U0, dims, beta, O.make_deltas("2lf", 0.5, 10), 6, t, metropolis=False)
    P0, K0 = O.momenta(dims, 6, 99)
    dh = {}
    for scheme, ns in (("2lf", (8, 16)), ("4mn", (4, 8))):
        for n in ns:
            d = O.make_deltas(scheme, 0.02, n)
            U1, P1 = _trajectory(U0, P0, dims, beta, d)
            dh[(scheme, n)] = abs(O.delta_h(U0, U1, P1, K0, dims, beta)[0])
    assert 3.0 < dh[("2lf", 8)] / dh[("2lf", 16)] < 5.0    # O(delta^2)
    assert 10.0 < dh[("4mn", 4)] / dh[("4mn", 8)] < 22.0   # O(delta^4)
    # reversibility: integrate, flip momenta, integrate back
    d = O.make_deltas("2mn", 0.02, 6)
    U1, P1 = _trajectory(U0, P0, dims, beta, d)
    U2, P2 = _trajectory(U1, -P1, dims, beta, d)
    assert np.abs(U2 - U0).max() < 1e-12
    assert np.abs(-P2 - P0).max() < 1e-12


def test_trajectory_metropolis_and_nan():
    dims = [4, 4, 4, 4]
    U = O.random_gauge(dims, 7)
    before = U.copy()
    r = O.hmc_trajectory(U, dims, 6.0, np.array([np.nan, 0.1, np.nan]), 7, 0)
    assert not r["accepted"]
    np.testing.assert_array_equal(U, before)
    assert r["u"] == O.metropolis_uniform(7, 0)


# ---------------------------------------------------------------- Landau gauge

def test_landau_deriv_finite_difference():
    """Reference test pattern src/tests/lattice_gauge.cpp:31-40, asserted:
    (F(exp(eps f) g) - F(g)) / (-2 eps sum Re tr(f deriv)) -> 1."""
    dims = [4, 4, 2, 4]
    U = O.random_gauge(dims, 8)
    g = O.random_transform(dims, 9)
    rng = np.random.default_rng(3)
    f = np.array([O.algebra_from_normals(rng.normal(size=8)) for _ in range(O.volume(dims))])
    D = O.landau_deriv(U, g, dims)
    eps = 1e-6
    gp = np.array([O.su3_exp(eps * fx) @ gx for fx, gx in zip(f, g[0])])[None]
    gm = np.array([O.su3_exp(-eps * fx) @ gx for fx, gx in zip(f, g[0])])[None]
    fd = (O.landau_functional(U, dims, gp) - O.landau_functional(U, dims, gm)) / 2
    expected = -2 * eps * np.einsum("xij,xji->", f, D).real
    assert abs(fd / expected - 1) < 1e-6


def test_landau_divsum_is_the_derivative_norm():
    """theta on a generic field: sum |TA(g d)|^2 = |deriv|^2 (2 Nc Nd V)^2,
    with deriv pinned by the finite-difference test above."""
    dims = [4, 4, 2, 4]
    U = O.random_gauge(dims, 18)
    g = O.random_transform(dims, 19)
    D = O.landau_deriv(U, g, dims)
    c = 2 * 3 * len(dims) * O.volume(dims)
    assert O.landau_divsum(U, dims, g) == pytest.approx(np.sum(np.abs(D) ** 2) * c * c, rel=1e-12)
    assert O.landau_divsum(U, dims, g) > 1.0


def test_landau_or_pure_gauge_and_monotone():
    dims = [4, 4, 4, 4]
    V = O.volume(dims)
    # pure gauge: U = h(x) h(x+mu)^+ -> OR reaches F = 1, theta = 0
    h = O.random_transform(dims, 10)
    Uu = np.zeros((4, V, 3, 3), np.complex128)
    Uu[:, :] = np.eye(3)
    U = O.gauge_transform(Uu, h, dims)
    for _ in range(300):
        O.landau_or_halfsweep(U, dims, 0, 1.7)
        O.landau_or_halfsweep(U, dims, 1, 1.7)
    assert O.landau_functional(U, dims) > 1 - 1e-12
    assert O.landau_divsum(U, dims) / V < 1e-20
    # generic field: omega = 1 never lowers F; the plaquette is unchanged
    U = O.random_gauge(dims, 11)
    for t in range(2):
        O.hmc_trajectory(U, dims, 6.0, O.make_deltas("2lf", 0.5, 10), 11, t, metropolis=False)
    p0 = O.plaq_sites(U, dims)
    F = [O.landau_functional(U, dims)]
    for _ in range(10):
        O.landau_or_halfsweep(U, dims, 0, 1.0)
        O.landau_or_halfsweep(U, dims, 1, 1.0)
        F.append(O.landau_functional(U, dims))
    assert all(b >= a - 1e-15 for a, b in zip(F, F[1:]))
    np.testing.assert_allclose(O.plaq_sites(U, dims), p0, rtol=0, atol=1e-12)


def test_landau_tracking_transform():
    """Links transformed in place equal U^h of the tracked transformation."""
    dims = [4, 2, 4, 2]
    U = O.random_gauge(dims, 12)
    Ug = U.copy()
    h = np.zeros((1, O.volume(dims), 3, 3), np.complex128)
    h[:, :] = np.eye(3)
    for _ in range(5):
        O.landau_or_halfsweep(Ug, dims, 0, 1.7, h)
        O.landau_or_halfsweep(Ug, dims, 1, 1.7, h)
    np.testing.assert_allclose(O.gauge_transform(U, h, dims), Ug, rtol=0, atol=1e-13)
    assert abs(O.landau_functional(U, dims, h) - O.landau_functional(Ug, dims)) < 1e-14


def test_landau_solve_quadratic_known_answer():
    """Vertex of the parabola through three points (src/gauge/landau.h:68-75)."""
    f = lambda x: -3.0 * (x - 1.7) ** 2 + 2.0
    xs = (0.5, 0.6, 0.4)
    assert abs(O.landau_solve_quadratic(*xs, *(f(x) for x in xs)) - 1.7) < 1e-12


def test_landau_ascent_pure_gauge():
    """The reference's solver (line-search ascent, Polak-Ribiere steps after
    iteration 20) restated on the CPU: on a gauge copy of the unit field it
    reaches the closed-form maximum F = 1 with a functional that never
    decreases, and the transformation it finds maps the field back to the
    unit field (U^g = 1 up to rounding, since g = c h^+ for a constant c)."""
    dims = [4, 4, 4, 4]
    V = O.volume(dims)
    h = O.random_transform(dims, 10)
    Uu = np.zeros((4, V, 3, 3), np.complex128)
    Uu[:, :] = np.eye(3)
    U = O.gauge_transform(Uu, h, dims)
    g, hist = O.landau_ascent(U, dims, 200)
    cond, F, _ = zip(*hist)
    assert len(hist) > 22 and cond[-1] < 1e-10, (len(hist), cond[-1])
    assert F[-1] > 1 - 1e-10
    assert all(b >= a for a, b in zip(F, F[1:]))
    np.testing.assert_allclose(O.gauge_transform(U, g, dims), Uu, rtol=0, atol=1e-5)


def test_axpy_rounds_product_then_sum():
    rng = np.random.default_rng(3)
    x = rng.normal(size=(2, 5, 3, 3)) + 1j * rng.normal(size=(2, 5, 3, 3))
    y = rng.normal(size=(2, 5, 3, 3)) + 1j * rng.normal(size=(2, 5, 3, 3))
    a = 0.3712
    r = O.axpy(y, a, x)
    assert np.array_equal(r.real, y.real + x.real * a) and np.array_equal(r.imag, y.imag + x.imag * a)


def test_reference_structured_trajectory_matches_fused():
    """The reference's pass structure (cshift copies, separate passes, global
    action sums; oracle_refpass.cpp) and the fused per-site restatement give
    the same trajectory to rounding: the products associate differently, so
    the fields agree to ~1e-13 and Delta H to ~1e-10, not bit for bit."""
    dims = [4, 4, 4, 4]
    d = O.make_deltas("2mn", 0.5, 4)
    U1 = O.random_gauge(dims, 7)
    U2 = U1.copy()
    for traj in range(3):
        a = O.hmc_trajectory(U1, dims, 6.0, d, 7, traj, metropolis=False)
        b = O.refpass_trajectory(U2, dims, 6.0, d, 7, traj, metropolis=False)
        assert abs(a["dH"] - b["dH"]) < 1e-9 * max(1.0, abs(a["dH"])), (traj, a["dH"], b["dH"])
        assert abs(a["H_old"] - b["H_old"]) < 1e-9 * abs(a["H_old"])
    assert np.abs(U1 - U2).max() < 1e-11


def test_stage_times_report_every_stage():
    dims = [4, 4, 4, 4]
    U = O.random_gauge(dims, 8)
    for refpass in (False, True):
        t = O.stage_times(U, dims, 6.0, 0.05, refpass=refpass)
        assert t.shape == (4,) and (t > 0).all()


def test_float_energy_of_the_reference_float_build():
    """orc_su3_float_energy_f32 (the reference's precision-1 sums): the unit
    field is exact (3.0 per site, S = 0), and on random fields the FP32 sums
    stay within their rounding bound of the FP64 sums of the same per-site
    values while differing from them."""
    for dims in ([4, 4, 4, 4], [6, 5, 3]):
        V, nd = O.volume(dims), len(dims)
        unit = np.zeros((nd, V, 3, 3), np.complex64)
        unit[..., [0, 1, 2], [0, 1, 2]] = 1.0
        S, K = O.su3_float_energy(unit, None, dims, 6.0)
        assert S == 0.0 and K == 0.0
    dims = [16, 16, 8, 8]
    V, nd = O.volume(dims), len(dims)
    U = O.random_gauge(dims, 3, "f32")
    P, K64 = O.momenta(dims, 4, 0, "f32")
    S, K = O.su3_float_energy(U, P, dims, 6.0)
    S64 = 6.0 * (nd * (nd - 1) / 2 * V - O.plaq_sites(U, dims).sum() / 3.0)
    K64 = K64.sum()
    # per-site FP32 products differ from the FP64 ones by ~1e-6 relative; the
    # FP32 block sums add a few ulp of each block and the host loop one ulp of
    # the running sum per block
    nb = V // 256
    assert abs(S - S64) < 6.0 / 3.0 * (6 * V * 3 * 1e-6 + nb * 6 * 3 * V * 6e-8)
    assert abs(K - K64) < nd * nb * K64 * 6e-8 + 1e-3
    assert S != S64 and K != K64
