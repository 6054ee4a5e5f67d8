"""The SU(2) and U(1) Landau restatements of the oracle (oracle_su2.cpp,
u1_oracle.py) pinned by mathematics: the derivative check of the reference's
own test (src/tests/lattice_gauge.cpp:30-39: (F[exp(eps f) g] - F[g]) equals
-2 eps sum Re tr(f deriv) to first order), the condition as |deriv|^2 times
(2 Nc Nd V)^2, the closed forms of a pure-gauge field (F = 1 at h = g0^+,
condition 0), gauge covariance (F[U, h] = F[U^h, 1]), and the overrelaxation
step at omega = 1 as the exact local maximum (no small rotation of an updated
site raises the functional; for U(1) Im(g d) = 0 at those sites)."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from oracle import u1_oracle as U1


# ---- numpy quaternions, independent of the oracle's C code

def qmul(a, b):
    w = a[..., 0] * b[..., 0] - a[..., 1] * b[..., 1] - a[..., 2] * b[..., 2] - a[..., 3] * b[..., 3]
    x = a[..., 0] * b[..., 1] + a[..., 1] * b[..., 0] + a[..., 2] * b[..., 3] - a[..., 3] * b[..., 2]
    y = a[..., 0] * b[..., 2] - a[..., 1] * b[..., 3] + a[..., 2] * b[..., 0] + a[..., 3] * b[..., 1]
    z = a[..., 0] * b[..., 3] + a[..., 1] * b[..., 2] - a[..., 2] * b[..., 1] + a[..., 3] * b[..., 0]
    return np.stack([w, x, y, z], axis=-1)


def qconj(a):
    return a * np.array([1, -1, -1, -1])


def qexp(v):
    """exp of the algebra element (0, v), rows of v"""
    al = np.linalg.norm(v, axis=-1, keepdims=True)
    s = np.where(al > 0, np.sin(al) / np.where(al > 0, al, 1), 1.0)
    return np.concatenate([np.cos(al), v * s], axis=-1)


def shift(a, dims, mu, s):
    """a[x + s mu] for a[V, ...] on a row-major lattice"""
    return np.roll(a.reshape(tuple(dims) + a.shape[1:]), -s, axis=mu).reshape(a.shape)


def random_su2(rng, n):
    q = rng.normal(size=(n, 4))
    return q / np.linalg.norm(q, axis=-1, keepdims=True)


DIMS = [[4, 4, 4, 4], [6, 4, 8], [8, 6]]


@pytest.mark.parametrize("dims", DIMS)
def test_su2_functional_and_derivative(dims):
    rng = np.random.default_rng(1)
    V, nd = O.volume(dims), len(dims)
    U = O.su2_random(dims, 3)
    h = random_su2(rng, V)[None].copy()
    F0 = O.su2_landau_functional(U, h, dims)
    # direct sum Re tr (h U h(x+mu)^+) / (2 Nd V)
    ref = sum(qmul(qmul(h[0], U[mu]), qconj(shift(h[0], dims, mu, 1)))[:, 0].sum() for mu in range(nd)) / (nd * V)
    assert F0 == pytest.approx(ref, rel=1e-13)
    D = O.su2_landau_deriv(U, h, dims)[0]
    assert np.all(D[:, 0] == 0)
    f = rng.normal(size=(V, 3)) * 0.5
    # Re tr (f D) as 2x2 matrices = 2 (f.w D.w - f.v . D.v), f.w = 0
    tr_fD = -2 * np.sum(f * D[:, 1:])
    for eps in (1e-4, 1e-5):
        h2 = qmul(qexp(eps * f), h[0])[None].copy()
        ratio = (O.su2_landau_functional(U, h2, dims) - F0) / (-2 * eps * tr_fD)
        assert abs(ratio - 1) < 50 * eps, (eps, ratio)
    # condition = |deriv|^2 (2 Nc Nd V)^2 / (Nd V): divsum = sum |D|^2 / c^2
    c = -0.5 / (2 * nd * V)
    assert O.su2_landau_divsum(U, h, dims) == pytest.approx(np.sum(D ** 2) / c ** 2, rel=1e-12)


@pytest.mark.parametrize("dims", DIMS)
def test_su2_pure_gauge_and_covariance(dims):
    rng = np.random.default_rng(2)
    V, nd = O.volume(dims), len(dims)
    g0 = random_su2(rng, V)
    Upg = np.stack([qmul(g0, qconj(shift(g0, dims, mu, 1))) for mu in range(nd)])
    hinv = qconj(g0)[None].copy()
    assert O.su2_landau_functional(Upg, hinv, dims) == pytest.approx(1.0, abs=1e-14)
    assert O.su2_landau_divsum(Upg, hinv, dims) < 1e-26
    U = O.su2_random(dims, 4)
    h = random_su2(rng, V)[None].copy()
    Uh = O.su2_gauge_transform(U, h, dims)
    assert O.su2_landau_functional(U, h, dims) == pytest.approx(O.su2_landau_functional(Uh, None, dims), rel=1e-13)
    assert O.su2_landau_divsum(U, h, dims) == pytest.approx(O.su2_landau_divsum(Uh, None, dims), rel=1e-11)
    assert O.su2_plaquette(Uh, dims) == pytest.approx(O.su2_plaquette(U, dims), rel=1e-13)


def _local_term(Ug, dims, x, G):
    """sum_mu Re w(G Ug_mu(x)) + Re w(Ug_mu(x-mu) G^+): the functional's terms at x after rotating by G"""
    s = 0.0
    for mu in range(len(dims)):
        xm = shift(np.arange(O.volume(dims)), dims, mu, -1)[x]
        s += qmul(G, Ug[mu, x])[0] + qmul(Ug[mu, xm], qconj(G))[0]
    return s


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_su2_or_step_is_local_maximum_and_converges(prec):
    dims = [4, 4, 4]
    rng = np.random.default_rng(3)
    V = O.volume(dims)
    Ug = O.su2_random(dims, 5, prec)
    h = np.zeros((1, V, 4), Ug.dtype)
    h[..., 0] = 1
    F = [O.su2_landau_functional(Ug, None, dims)]
    par = np.indices(dims).sum(axis=0).reshape(-1) % 2
    U0 = Ug.copy()
    O.su2_landau_or_half(Ug, h, dims, 0, 1.0)
    if prec == "f64":
        # h tracks the rotation: U^h of the original equals the updated links
        np.testing.assert_allclose(O.su2_gauge_transform(U0, h, dims), Ug, atol=1e-14)
        for x in np.flatnonzero(par == 0)[:6]:
            base = _local_term(Ug, dims, x, np.array([1.0, 0, 0, 0]))
            for _ in range(20):
                G = qexp(rng.normal(size=3) * 1e-3)
                assert _local_term(Ug, dims, x, G) <= base + 1e-13
    F.append(O.su2_landau_functional(Ug, None, dims))
    for it in range(300):
        for p in (1, 0):
            O.su2_landau_or_half(Ug, h, dims, p, 1.7)
        F.append(O.su2_landau_functional(Ug, None, dims))
    assert all(b > a - 1e-6 for a, b in zip(F[:3], F[1:4]))
    theta = O.su2_landau_divsum(Ug, None, dims) / V
    assert theta < (1e-20 if prec == "f64" else 1e-9), theta


# ---- U(1)

def _u1_cfg(dims, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(-math.pi, math.pi, size=(len(dims),) + tuple(dims)).astype(np.float32)


@pytest.mark.parametrize("dims", [(8, 8), (4, 6, 4), (4, 4, 4, 4)])
def test_u1_functional_and_derivative(dims):
    rng = np.random.default_rng(6)
    th = _u1_cfg(dims, 7)
    phi = rng.uniform(-math.pi, math.pi, size=dims).astype(np.float32)
    nd, V = len(dims), int(np.prod(dims))
    # direct: cos of each transformed link angle
    ref = np.mean([np.cos(phi.astype(float) + th[mu] - np.roll(phi, -1, axis=mu)) for mu in range(nd)])
    F0 = U1.landau_functional(th, phi)
    assert F0 == pytest.approx(ref, rel=1e-12)
    D = U1.landau_deriv(th, phi).astype(np.float64)
    f = rng.normal(size=dims)
    # algebra elements i f, i D: Re tr (i f i D) = -f D
    tr_fD = -np.sum(f * D)
    for eps in (1e-3, 1e-4):
        ratio = (U1.landau_functional(th, phi.astype(np.float64) + eps * f) - F0) / (-2 * eps * tr_fD)
        assert abs(ratio - 1) < 30 * eps + 2e-3, ratio  # D is float32
    c = -0.5 / (nd * V)
    assert U1.landau_divsum(th, phi) == pytest.approx(np.sum(U1.landau_gd(th, phi) ** 2), rel=1e-15)
    np.testing.assert_allclose(D, c * U1.landau_gd(th, phi), rtol=1e-6, atol=1e-12)


def test_u1_pure_gauge_covariance_and_or():
    dims = (8, 6)
    rng = np.random.default_rng(8)
    p0 = rng.uniform(-math.pi, math.pi, size=dims)
    th_pg = np.stack([U1.wrap(p0 - np.roll(p0, -1, axis=mu)) for mu in range(2)]).astype(np.float32)
    assert U1.landau_functional(th_pg, (-p0).astype(np.float32)) == pytest.approx(1.0, abs=1e-12)
    assert U1.landau_divsum(th_pg, (-p0).astype(np.float32)) < 1e-11
    th = _u1_cfg(dims, 9)
    phi = rng.uniform(-math.pi, math.pi, size=dims).astype(np.float32)
    tg = U1.gauge_transform(th, phi)
    assert U1.landau_functional(th, phi) == pytest.approx(U1.landau_functional(tg, None), abs=1e-6)
    # one omega = 1 half-sweep: Im(g d) vanishes at the updated sites and phi tracks the rotation
    t1, ph1 = U1.landau_or_half(tg, np.zeros(dims, np.float32), 0, 1.0)
    par = np.indices(dims).sum(axis=0) % 2
    gd = U1.landau_gd(t1, None)
    assert np.abs(gd[par == 0]).max() < 2e-6
    assert np.abs(gd[par == 1]).max() > 1e-2
    assert np.abs(U1.wrap(U1.gauge_transform(tg, ph1).astype(float) - t1)).max() < 2e-6
    assert U1.landau_functional(t1, None) > U1.landau_functional(tg, None)
