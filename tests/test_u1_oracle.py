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


def test_tau_int_frozen_series():
    """a charge that never changes gives no effective samples"""
    assert tau_int(np.zeros((4, 100)))[0] == math.inf



def _pflow_params(seed, scale):
    rng = np.random.default_rng(seed)
    p = rng.normal(size=U.NPF) * scale
    p[2 * U.PK] += 1.0
    return p


def test_pflow_links_have_the_drawn_plaquettes():
    """Tree-gauge construction: every free plaquette of the built links is the
    drawn one, the last is minus their sum, and gauge/holonomy draws do not
    touch any plaquette."""
    rng = np.random.default_rng(20)
    for L0, L1 in ((4, 4), (3, 5), (6, 2)):
        V = L0 * L1
        Pf = rng.uniform(-np.pi, np.pi, V - 1)
        th = U.pflow_links(Pf, rng.uniform(-3, 3), rng.uniform(-3, 3), rng.uniform(-3, 3, V), L0, L1)
        P = U.plaq(th).reshape(-1)
        assert np.abs(U.wrap(P[:V - 1] - Pf)).max() < 1e-12
        assert abs(U.wrap(P[V - 1] + Pf.sum())) < 1e-12


def test_pflow_map_is_unimodular():
    """(free plaquettes, h0, h1, gauge a(x != 0)) -> links is integer-linear
    with |det| = 1, so the density of the links is the product of the
    plaquette densities times (2 pi)^-(V+1) (pins the log q normalisation)."""
    L0, L1 = 3, 4
    V = L0 * L1
    eps = 1e-3
    cols = []
    for k in range(2 * V):
        Pf, h, a = np.zeros(V - 1), np.zeros(2), np.zeros(V)
        if k < V - 1:
            Pf[k] = eps
        elif k < V + 1:
            h[k - (V - 1)] = eps
        else:
            a[k - (V + 1) + 1] = eps
        cols.append(U.pflow_links(Pf, h[0], h[1], a, L0, L1).reshape(-1) / eps)
    M = np.array(cols).T
    np.testing.assert_allclose(M, np.rint(M), atol=1e-9)
    assert abs(abs(np.linalg.det(M)) - 1) < 1e-9


def test_pflow_logq_identity_and_inverse():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(21)
    L0, L1 = 4, 6
    V = L0 * L1
    th = rng.uniform(-np.pi, np.pi, size=(3, 2, L0, L1))
    ident = np.zeros(U.NPF)
    ident[2 * U.PK] = 1.0
    lq, _ = U.pflow_logq(ident, th)
    np.testing.assert_allclose(lq, -2 * V * math.log(2 * math.pi), atol=1e-12)
    # log q of built links = the forward map's own -log g' (inverse consistency)
    p = _pflow_params(22, 0.5)
    u = torch.tensor(rng.uniform(-np.pi, np.pi, size=(3, V - 1)))
    thP, nlg = U.pflow_h(torch.tensor(p), u)
    for c in range(3):
        links = U.pflow_links(thP[c].numpy(), 0.3, -1.1, rng.uniform(-3, 3, V), L0, L1)
        lq, _ = U.pflow_logq(p, links)
        assert abs(lq - (nlg[c].sum().item() - 2 * V * math.log(2 * math.pi))) < 1e-9


def test_pflow_loss_gradient_finite_difference():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(23)
    p = torch.tensor(_pflow_params(24, 0.5), requires_grad=True)
    u = torch.tensor(rng.uniform(-np.pi, np.pi, size=(4, 15)))
    loss = U.pflow_loss(p, u, 2.0)[0].mean()
    loss.backward()
    g = p.grad.numpy()
    h = 1e-6
    for i in range(U.NPF):
        pp, pm = p.detach().clone(), p.detach().clone()
        pp[i] += h
        pm[i] -= h
        fd = (U.pflow_loss(pp, u, 2.0)[0].mean() - U.pflow_loss(pm, u, 2.0)[0].mean()).item() / (2 * h)
        assert abs(fd - g[i]) < 1e-6 * (1 + abs(g[i])), (i, fd, g[i])


def _pc_params(seed, scale, layers=8):
    rng = np.random.default_rng(seed)
    p = rng.normal(size=layers * U.PC_NP_LAYER) * scale
    for layer in range(layers):
        p[(layer + 1) * U.PC_NP_LAYER - U.NO + 2 * U.NK] += 1.0
    return p


def _pc_jacobian(p, z, L0, L1):
    import torch

    return torch.autograd.functional.jacobian(
        lambda v: U.pcflow_forward(p, v.unsqueeze(0), L0, L1)[0][0], z).numpy()


def test_pcflow_logdet_against_jacobian():
    """log det of the plaquette coupling flow z -> phi (the V - 1 free
    plaquettes) equals the sum of log g' that the flow reports, for a brute
    force Jacobian of the whole 8-layer map on 4x4 and 4x6."""
    torch = pytest.importorskip("torch")
    for L0, L1, seed in ((4, 4, 30), (4, 6, 31)):
        V = L0 * L1
        p = torch.tensor(_pc_params(seed, 0.3))
        z = torch.tensor(np.random.default_rng(seed).uniform(-np.pi, np.pi, V - 1))
        _, ld = U.pcflow_forward(p, z.unsqueeze(0), L0, L1)
        J = _pc_jacobian(p, z, L0, L1)
        sign, logabs = np.linalg.slogdet(J)
        assert sign > 0
        assert abs(logabs - ld.item()) < 1e-9, (logabs, ld.item())
        assert abs(ld.item()) > 0.5  # a non-trivial map
        assert np.abs(J - np.diag(np.diag(J))).max() > 1e-2  # with couplings


def test_pcflow_layer_structure():
    """One layer of class m: the plaquettes of the other classes do not move
    (unit rows), and an active plaquette depends on itself and on its 8
    surrounding plaquettes only; the last plaquette is never a variable. The
    four classes cover every free plaquette once."""
    torch = pytest.importorskip("torch")
    L0, L1 = 6, 4
    V = L0 * L1
    z = torch.tensor(np.random.default_rng(32).uniform(-np.pi, np.pi, V - 1))
    seen = np.zeros(V - 1, int)
    for m in range(4):
        p = torch.tensor(_pc_params(33 + m, 0.4, layers=m + 1))
        # layers 0 .. m - 1 with zero weights (identity), layer m non-trivial
        p[:m * U.PC_NP_LAYER] = 0.0
        for layer in range(m):
            p[(layer + 1) * U.PC_NP_LAYER - U.NO + 2 * U.NK] = 1.0
        J = _pc_jacobian(p, z, L0, L1)
        mask = U._pc_class_mask(L0, L1, m).reshape(-1)[:V - 1]
        seen += mask
        for x in range(V - 1):
            row = J[x]
            if not mask[x]:
                np.testing.assert_allclose(row, np.eye(V - 1)[x], atol=1e-14)
                continue
            x0, x1 = divmod(x, L1)
            ring = {((x0 + d0) % L0) * L1 + (x1 + d1) % L1 for d0, d1 in U.PC_RING}
            allowed = (ring | {x}) - {V - 1}
            for y in range(V - 1):
                if y not in allowed:
                    assert row[y] == 0.0, (m, x, y)
            assert row[x] > 0
            assert max(abs(row[y]) for y in allowed - {x}) > 1e-4
    np.testing.assert_array_equal(seen, np.ones(V - 1, int))


def test_pcflow_identity_inverse_and_gauge_invariance():
    """Zero weights with the identity bias are the identity (log q = the
    uniform density); the inverse recovers the prior draw and the same log
    det; log q of links built from a sample (any holonomies and gauge) is the
    sample's own log q."""
    torch = pytest.importorskip("torch")
    L0, L1 = 4, 6
    V = L0 * L1
    rng = np.random.default_rng(34)
    z = torch.tensor(rng.uniform(-np.pi, np.pi, size=(3, V - 1)))
    ident = torch.zeros(8 * U.PC_NP_LAYER, dtype=torch.float64)
    for layer in range(8):
        ident[(layer + 1) * U.PC_NP_LAYER - U.NO + 2 * U.NK] = 1.0
    phi, ld = U.pcflow_forward(ident, z, L0, L1)
    assert torch.abs(U._twrap(phi - z)).max().item() < 1e-12
    assert torch.abs(ld).max().item() < 1e-12
    p = _pc_params(35, 0.3)
    phi, ld = U.pcflow_forward(torch.tensor(p), z, L0, L1)
    zi, ldi = U.pcflow_forward(torch.tensor(p), phi, L0, L1, inverse=True)
    assert torch.abs(U._twrap(zi - z)).max().item() < 1e-10
    assert torch.abs(ldi - ld).max().item() < 1e-9
    for c in range(3):
        th = U.pflow_links(phi[c].numpy(), rng.uniform(-3, 3), rng.uniform(-3, 3), rng.uniform(-3, 3, V), L0, L1)
        lq, S, zz = U.pcflow_logq(p, th[None])
        assert abs(lq[0] - (-ld[c].item() - 2 * V * math.log(2 * math.pi))) < 1e-9
        assert np.abs(U.wrap(zz[0] - z[c].numpy())).max() < 1e-9
        assert abs(S[0] - U.action_over_beta(th)) < 1e-9


def test_pcflow_loss_gradient_finite_difference():
    torch = pytest.importorskip("torch")
    L0, L1 = 4, 4
    rng = np.random.default_rng(36)
    p = torch.tensor(_pc_params(37, 0.3, layers=4), requires_grad=True)
    z = torch.tensor(rng.uniform(-np.pi, np.pi, size=(3, 15)))
    U.pcflow_loss(p, z, L0, L1, 2.0)[0].mean().backward()
    g = p.grad.numpy()
    h = 1e-6
    for i in rng.choice(p.numel(), 40, replace=False):
        pp, pm = p.detach().clone(), p.detach().clone()
        pp[i] += h
        pm[i] -= h
        fd = (U.pcflow_loss(pp, z, L0, L1, 2.0)[0].mean() - U.pcflow_loss(pm, z, L0, L1, 2.0)[0].mean()).item()
        fd /= 2 * h
        assert abs(fd - g[i]) < 1e-6 * (1 + abs(g[i])), (i, fd, g[i])


def test_u1_nd_oracle_reduces_to_2d():
    """The any-dimension oracle on a 2-d lattice is the 2-d oracle: the same
    random angles and momenta (the same Philox streams), plaquettes, force and
    trajectory."""
    L0, L1 = 4, 6
    th = U.nd_random_angles(7, 1, (L0, L1))
    np.testing.assert_array_equal(th, U.random_angles(7, 1, L0, L1))
    q = U.nd_momenta(7, 3, 1, (L0, L1))
    np.testing.assert_array_equal(q, U.momenta(7, 3, 1, L0, L1))
    t64 = th.astype(np.float64)
    np.testing.assert_allclose(U.nd_plaquettes(t64)[0], U.plaq(t64), rtol=0, atol=1e-15)
    np.testing.assert_allclose(U.nd_dS(t64, 1.7), U.dS(t64, 1.7), rtol=0, atol=1e-14)
    d = O.make_deltas("2lf", 0.5, 5)
    a, b = U.nd_hmc_trajectory(th, q, 2.0, d), U.hmc_trajectory(th, q, 2.0, d)
    assert abs(a[2] - b[2]) < 1e-12
    np.testing.assert_allclose(a[0], b[0], rtol=0, atol=1e-12)


def test_u1_nd_force_finite_difference_and_integrator():
    """4-d: dS/dtheta against central differences of the action; Delta H of the
    leapfrog scales as the square of the step; the trajectory reverses."""
    dims = (3, 4, 2, 3)
    rng = np.random.default_rng(4)
    th = rng.uniform(-np.pi, np.pi, size=(4,) + dims)
    beta = 1.3
    g = U.nd_dS(th, beta)
    for (mu, x) in [(0, (0, 1, 1, 2)), (2, (2, 3, 0, 1)), (3, (1, 0, 1, 0))]:
        h = 1e-6
        tp, tm = th.copy(), th.copy()
        tp[(mu,) + x] += h
        tm[(mu,) + x] -= h
        fd = beta * (U.nd_action_over_beta(tp) - U.nd_action_over_beta(tm)) / (2 * h)
        assert abs(fd - g[(mu,) + x]) < 1e-6 * max(1.0, abs(fd))
    q = U.nd_momenta(5, 0, 0, dims).astype(np.float64)
    dh = [abs(U.nd_hmc_trajectory(th, q, beta, O.make_deltas("2lf", 0.5, n))[2]) for n in (10, 20)]
    assert 3.0 < dh[0] / dh[1] < 5.0
    d = O.make_deltas("2lf", 0.5, 10)
    t1, q1, _ = U.nd_hmc_trajectory(th, q, beta, d)
    t2, _, _ = U.nd_hmc_trajectory(t1, -q1, beta, d)
    assert np.abs(U.wrap(t2 - th)).max() < 1e-12


def test_factorized_plaquette_proposal_acceptance_ceiling():
    """An independence sampler that proposes the V - 1 free plaquettes i.i.d.
    from the exact one-plaquette law e^{beta cos}/(2 pi I0(beta)) -- the best a
    proposal can do that does not see the constraint on their sum -- has
    log-weights beta cos(P_last) + const, and accepts at
    E_{u ~ vM(beta), v ~ U(-pi, pi)} min(1, e^{beta (cos v - cos u)})
    (0.426 at beta 2, 0.272 at beta 4): the ceiling of the plaquette flows.
    The integral against an MH simulation with this proposal on an 8x8 torus."""
    from scipy import integrate

    rng = np.random.default_rng(7)
    V, n = 64, 40000
    for beta in (2.0, 4.0):
        z = 2 * np.pi * sp.i0(beta)
        ceil, _ = integrate.dblquad(
            lambda v, u: np.exp(beta * np.cos(u)) / z / (2 * np.pi) * min(1.0, np.exp(beta * (np.cos(v) - np.cos(u)))),
            -np.pi, np.pi, -np.pi, np.pi, epsabs=1e-7)
        lw = beta * np.cos(-rng.vonmises(0.0, beta, size=(n, V - 1)).sum(axis=1))
        logu = np.log(rng.uniform(size=n))
        acc, cur = 0, lw[0]
        for i in range(1, n):
            if logu[i] < lw[i] - cur:
                acc += 1
                cur = lw[i]
        assert abs(acc / (n - 1) - ceil) < 0.012, (beta, acc / (n - 1), ceil)
    assert abs(ceil - 0.272) < 1e-3
