"""Pins of the U(1) 2-d oracle and of the host-side autocorrelation analysis
(CPU only): limits of the exact finite-volume formulas, integrator order,
the flow's log-det against a brute-force Jacobian, gauge equivariance,
identity initialisation, and the Gamma-method on a process with known
autocorrelation."""
import math

import numpy as np
import pytest
import scipy.special as sp
from scipy.integrate import quad

from mesh_amd.u1 import tau_int
from oracle import oracle as O
from oracle import u1_oracle as U


def test_exact_mean_cos_limits():
    for beta in (0.5, 2.0, 6.0):
        # one plaquette on a torus is the identity: cos = 1
        assert abs(U.exact_mean_cos(beta, 1) - 1.0) < 1e-12
        # large volume: independent von Mises plaquettes, I1/I0
        assert abs(U.exact_mean_cos(beta, 100000) - sp.iv(1, beta) / sp.iv(0, beta)) < 1e-8


def test_exact_chi_limits():
    for beta in (1.0, 4.0):
        # large volume: <phi^2> / (2 pi)^2 of one von Mises plaquette
        m2 = quad(lambda p: p * p * math.exp(beta * math.cos(p)), -math.pi, math.pi)[0] / \
            quad(lambda p: math.exp(beta * math.cos(p)), -math.pi, math.pi)[0]
        assert abs(U.exact_chi_top(beta, 4096) / (m2 / (4 * math.pi ** 2)) - 1) < 2e-3


def test_exact_results_against_brute_force_2x2():
    """A 2x2 torus has 8 links; integrate exp(-S) over the gauge-invariant
    content by Monte Carlo with many samples and compare <cos> (pins the
    Bessel-sum formula's indexing and normalisation)."""
    rng = np.random.default_rng(0)
    beta = 1.3
    th = rng.uniform(-np.pi, np.pi, size=(400000, 2, 2, 2))
    P = U.plaq(th)
    w = np.exp(-beta * np.sum(1 - np.cos(P), axis=(-2, -1)))
    mc = np.sum(w * np.mean(np.cos(P), axis=(-2, -1))) / np.sum(w)
    assert abs(mc - U.exact_mean_cos(beta, 4)) < 5e-3
    Q = np.rint(np.sum(U.wrap(P), axis=(-2, -1)) / (2 * np.pi))
    chi = np.sum(w * Q * Q) / np.sum(w) / 4
    assert abs(chi / U.exact_chi_top(beta, 4) - 1) < 0.03


def test_u1_hmc_oracle_order_and_reversibility():
    rng = np.random.default_rng(1)
    th = rng.uniform(-np.pi, np.pi, size=(2, 8, 8))
    q = U.momenta(3, 0, 0, 8, 8)
    dh = []
    for n in (10, 20):
        d = O.make_deltas("2lf", 0.5, n)
        dh.append(abs(U.hmc_trajectory(th, q, 2.0, d)[2]))
    assert 3.0 < dh[0] / dh[1] < 5.0
    d = O.make_deltas("2lf", 0.5, 10)
    t1, q1, _ = U.hmc_trajectory(th, q, 2.0, d)
    t2, q2, _ = U.hmc_trajectory(t1, -q1, 2.0, d)
    assert np.abs(U.wrap(t2 - th)).max() < 1e-12
    # momenta normalisation: q ~ N(0, 1/2) (src/gauge/groups.h:86-91)
    qq = np.concatenate([U.momenta(5, t, 0, 8, 8).ravel() for t in range(20)])
    assert abs(qq.var() - 0.5) < 5 * 0.5 * math.sqrt(2 / qq.size)


def test_charge_is_integer_and_gauge_invariant():
    rng = np.random.default_rng(2)
    th = rng.uniform(-np.pi, np.pi, size=(16, 2, 6, 8))
    s = np.sum(U.wrap(U.plaq(th)), axis=(-2, -1)) / (2 * np.pi)
    np.testing.assert_allclose(s, np.rint(s), atol=1e-9)
    alpha = rng.uniform(-np.pi, np.pi, size=(16, 6, 8))
    g0 = th[:, 0] + alpha - np.roll(alpha, -1, axis=-2)
    g1 = th[:, 1] + alpha - np.roll(alpha, -1, axis=-1)
    thg = np.stack([g0, g1], axis=1)
    np.testing.assert_array_equal(U.charge(thg), U.charge(th))


def _random_params(seed, scale=0.3):
    rng = np.random.default_rng(seed)
    p = rng.normal(size=U.NL * U.NP_LAYER) * scale
    for layer in range(U.NL):
        p[layer * U.NP_LAYER + U.NP_LAYER - U.NO + 2 * U.NK] += 1.0
    return p


def _identity_params():
    p = np.zeros(U.NL * U.NP_LAYER)
    for layer in range(U.NL):
        p[layer * U.NP_LAYER + U.NP_LAYER - U.NO + 2 * U.NK] = 1.0
    return p


def test_flow_identity_initialisation():
    torch = pytest.importorskip("torch")
    z = torch.tensor(np.random.default_rng(3).uniform(-np.pi, np.pi, size=(2, 2, 8, 8)))
    th, ld = U.flow_forward(torch.tensor(_identity_params()), z)
    assert torch.max(torch.abs(U._twrap(th - z))) < 1e-12
    assert torch.max(torch.abs(ld)) < 1e-12


def test_flow_logdet_against_jacobian():
    """log det of the flow map equals log|det J| of the full 2V -> 2V
    Jacobian computed by autograd on a 4x4 lattice."""
    torch = pytest.importorskip("torch")
    # weights mild enough that the 32 x 32 Jacobian stays well conditioned
    # (with large weights cond(J) reaches 1e14 and slogdet itself is inexact)
    p = torch.tensor(_random_params(4, scale=0.2))
    z = torch.tensor(np.random.default_rng(5).uniform(-np.pi, np.pi, size=(1, 2, 4, 4)))
    _, ld = U.flow_forward(p, z)
    J = torch.autograd.functional.jacobian(lambda zz: U.flow_forward(p, zz.view(1, 2, 4, 4))[0].reshape(-1),
                                           z.reshape(-1))
    assert torch.linalg.cond(J).item() < 1e8
    sign, logabs = torch.linalg.slogdet(J)
    assert sign.item() > 0
    assert abs(ld.item()) > 1.0  # a non-trivial transform
    assert abs(logabs.item() - ld.item()) < 1e-8


def test_flow_gauge_equivariance():
    torch = pytest.importorskip("torch")
    p = torch.tensor(_random_params(6))
    rng = np.random.default_rng(7)
    z = torch.tensor(rng.uniform(-np.pi, np.pi, size=(3, 2, 8, 8)))
    alpha = torch.tensor(rng.uniform(-np.pi, np.pi, size=(3, 8, 8)))

    def gt(t):
        return torch.stack([U._twrap(t[:, 0] + alpha - torch.roll(alpha, -1, dims=1)),
                            U._twrap(t[:, 1] + alpha - torch.roll(alpha, -1, dims=2))], dim=1)

    th1, ld1 = U.flow_forward(p, gt(z))
    th2, ld2 = U.flow_forward(p, z)
    assert torch.max(torch.abs(U._twrap(th1 - gt(th2)))) < 1e-10
    assert torch.max(torch.abs(ld1 - ld2)) < 1e-10


def test_tau_int_ar1():
    """x_t = rho x_{t-1} + e_t has tau_int = (1 + rho) / (2 (1 - rho))."""
    rng = np.random.default_rng(8)
    for rho in (0.0, 0.5, 0.9):
        R, n = 32, 20000
        x = np.zeros((R, n))
        e = rng.normal(size=(R, n))
        for t in range(1, n):
            x[:, t] = rho * x[:, t - 1] + e[:, t]
        tau, err = tau_int(x)
        exact = (1 + rho) / (2 * (1 - rho))
        assert abs(tau - exact) < max(4 * err, 0.05 * exact), (rho, tau, err, exact)
