"""The U(1) plaquette coupling flow (pcflow.hip) on the GPU against the torch
float64 oracle: identity initialisation, samples and log q of the forward
map, log q of given configurations by the inverse, gauge invariance, the
reverse-KL gradient against autograd, and training against the exact
free-energy bound with MH chains reproducing the exact torus <cos theta_P>."""
import math

import numpy as np
import pytest

import mesh_amd
from mesh_amd import gauge
from mesh_amd import u1 as U1
from oracle import u1_oracle as U

pytestmark = pytest.mark.gpu


def _wrapdiff(a, b):
    return np.abs(U.wrap(np.asarray(a, np.float64) - np.asarray(b, np.float64)))


def _params(seed, scale, layers=8):
    rng = np.random.default_rng(seed)
    p = rng.normal(size=layers * U.PC_NP_LAYER) * scale
    for layer in range(layers):
        p[(layer + 1) * U.PC_NP_LAYER - U.NO + 2 * U.NK] += 1.0
    return p.astype(np.float32)


def _log_z(beta, V):
    import scipy.special as sp

    terms = [V * math.log(sp.ive(n, beta)) for n in range(-30, 31)]
    m = max(terms)
    return 2 * V * math.log(2 * math.pi) + m + math.log(sum(math.exp(t - m) for t in terms))


def _free_plaquettes(th):
    B, _, L0, L1 = th.shape
    return U.wrap(U.plaq(np.asarray(th, np.float64))).reshape(B, L0 * L1)[:, :L0 * L1 - 1]


def test_pcflow_identity(ctx):
    """the initial weights with the last layer zeroed are the identity: the
    free plaquettes of a proposal are the prior draws and log q the uniform
    density; the initial last layer is small"""
    L, B, nl = 8, 4, 8
    f = U1.PlaqCouplingFlow(L, L, ctx)
    p = f.get_params()
    for layer in range(nl):
        o = (layer + 1) * U.PC_NP_LAYER - U.NO - U.NH * U.NO
        w3 = p[o:o + U.NH * U.NO]
        assert 0 < np.abs(w3).max() < 0.01
        p[o:o + U.NH * U.NO] = 0
    f.set_params(p)
    ens = U1.U1Ensemble(L, L, B, ctx)
    out = f.sample(ens, 3, 0).download().reshape(B, 2)
    P = _free_plaquettes(ens.angles())
    for c in range(B):
        z, _, _, _ = U.pcflow_draws(3, 0, c, L, L)
        assert _wrapdiff(P[c], z).max() < 1e-5, c
    np.testing.assert_allclose(out[:, 0], U.log_prior(L, L), atol=1e-4)


@pytest.mark.parametrize("L0,L1", [(4, 4), (8, 8), (6, 10), (16, 16), (32, 32)])
def test_pcflow_sample_against_oracle(ctx, L0, L1):
    """links, plaquettes and log q of samples against the oracle's forward
    map of the same Philox draws (log q sums one FP32 log g' per active
    plaquette per layer: its bound grows with V)"""
    torch = pytest.importorskip("torch")
    B, seed, step = 3, 91, 5
    V = L0 * L1
    p = _params(92, 0.2)
    f = U1.PlaqCouplingFlow(L0, L1, ctx)
    f.set_params(p)
    ens = U1.U1Ensemble(L0, L1, B, ctx)
    out = f.sample(ens, seed, step).download().reshape(B, 2)
    th = ens.angles()
    draws = [U.pcflow_draws(seed, step, c, L0, L1) for c in range(B)]
    z = torch.tensor(np.stack([d[0] for d in draws]).astype(np.float64))
    phi, ld = U.pcflow_forward(torch.tensor(p.astype(np.float64)), z, L0, L1)
    phi, ld = phi.numpy(), ld.numpy()
    assert np.abs(ld).max() > 0.5  # a non-trivial map
    P = _free_plaquettes(th)
    for c in range(B):
        assert _wrapdiff(P[c], phi[c]).max() < 1e-3, c
        _, a, h0, h1 = draws[c]
        ref = U.pflow_links(phi[c], h0, h1, a, L0, L1)
        assert _wrapdiff(th[c], ref).max() < 2e-3, c
    tol = 2e-3 * max(1.0, V / 64)
    np.testing.assert_allclose(out[:, 0], U.log_prior(L0, L1) - ld, atol=tol)
    np.testing.assert_allclose(out[:, 1], U.action_over_beta(th.astype(np.float64)), rtol=1e-4)


def test_pcflow_inverse_roundtrip(ctx):
    """log q by the inverse of a sample equals the sampler's log q and
    recovers its prior draws"""
    L, B, seed, step = 8, 4, 93, 2
    f = U1.PlaqCouplingFlow(L, L, ctx)
    f.set_params(_params(94, 0.2))
    ens = U1.U1Ensemble(L, L, B, ctx)
    fwd = f.sample(ens, seed, step).download().reshape(B, 2)
    z_out = U1.Buffer(B * L * L, np.float32, ctx)
    inv = f.logq(ens, z_out=z_out).download().reshape(B, 2)
    z = z_out.download().reshape(B, L * L)
    for c in range(B):
        assert _wrapdiff(z[c, :-1], U.pcflow_draws(seed, step, c, L, L)[0]).max() < 1e-3, c
    np.testing.assert_allclose(inv[:, 0], fwd[:, 0], atol=2e-3)
    np.testing.assert_allclose(inv[:, 1], fwd[:, 1], rtol=1e-5)


def test_pcflow_inverse_of_a_steep_map(ctx):
    """a steeper map (the s_k biases spread over +-1; log det -22 .. -12 on
    these draws): the inverse still recovers the prior draws and the
    sampler's log q.  Steeper still, the composition itself becomes
    ill-conditioned: with the last layers' weights doubled as well, the exact
    float64 inverse of outputs moved by one FP32 ulp misses the draws by up
    to 1.5e-2, so no FP32 inverse could pass this check there."""
    L, B, seed, step = 8, 8, 109, 6
    p = _params(110, 0.2)
    rng = np.random.default_rng(111)
    for layer in range(8):
        o = (layer + 1) * U.PC_NP_LAYER - U.NO - U.NH * U.NO
        p[o + U.NH * U.NO:o + U.NH * U.NO + U.NK] = rng.uniform(-1, 1, U.NK)
    f = U1.PlaqCouplingFlow(L, L, ctx)
    f.set_params(p.astype(np.float32))
    ens = U1.U1Ensemble(L, L, B, ctx)
    fwd = f.sample(ens, seed, step).download().reshape(B, 2)
    z_out = U1.Buffer(B * L * L, np.float32, ctx)
    inv = f.logq(ens, z_out=z_out).download().reshape(B, 2)
    z = z_out.download().reshape(B, L * L)
    assert np.abs(fwd[:, 0] - U.log_prior(L, L)).max() > 5  # a steep map
    for c in range(B):
        assert _wrapdiff(z[c, :-1], U.pcflow_draws(seed, step, c, L, L)[0]).max() < 2e-3, c
    np.testing.assert_allclose(inv[:, 0], fwd[:, 0], atol=5e-3)


@pytest.mark.parametrize("L0,L1", [(8, 8), (4, 12)])
def test_pcflow_logq_of_given_configurations(ctx, L0, L1):
    """log q of uniform random links (not the flow's samples) against the
    oracle's inverse, and invariance under a gauge transformation"""
    B = 4
    p = _params(95, 0.2)
    f = U1.PlaqCouplingFlow(L0, L1, ctx)
    f.set_params(p)
    ens = U1.U1Ensemble(L0, L1, B, ctx)
    ens.random_start(96)
    th = ens.angles().astype(np.float64)
    out = f.logq(ens).download().reshape(B, 2)
    lq, S, _ = U.pcflow_logq(p, th)
    assert np.abs(lq - U.log_prior(L0, L1)).max() > 0.5
    np.testing.assert_allclose(out[:, 0], lq, atol=3e-3)
    np.testing.assert_allclose(out[:, 1], S, rtol=1e-5)
    a = np.random.default_rng(97).uniform(-np.pi, np.pi, size=(B, L0, L1))
    g = th.copy()
    g[:, 0] += a - np.roll(a, -1, axis=1)
    g[:, 1] += a - np.roll(a, -1, axis=2)
    ens.set_angles(U.wrap(g).astype(np.float32))
    out_g = f.logq(ens).download().reshape(B, 2)
    np.testing.assert_allclose(out_g[:, 0], out[:, 0], atol=2e-3)


def test_pcflow_gradient_against_autograd(ctx):
    """mesh_pcflow_gradient against torch float64 autograd of
    mean(beta S - log det) on the same prior draws, in total and per layer"""
    torch = pytest.importorskip("torch")
    L, B, beta, seed, step, nl = 8, 6, 2.0, 98, 1, 8
    p = _params(99, 0.2)
    f = U1.PlaqCouplingFlow(L, L, ctx)
    f.set_params(p)
    loss, logq, S = f.gradient(B, beta, seed, step)
    g = f.grad.download().astype(np.float64)
    pt = torch.tensor(p.astype(np.float64), requires_grad=True)
    z = torch.tensor(np.stack([U.pcflow_draws(seed, step, c, L, L)[0] for c in range(B)]).astype(np.float64))
    ref, ref_lq, ref_S, _ = U.pcflow_loss(pt, z, L, L, beta)
    ref.mean().backward()
    gr = pt.grad.numpy()
    np.testing.assert_allclose(logq, ref_lq.detach().numpy(), atol=2e-3)
    np.testing.assert_allclose(S, ref_S.detach().numpy(), rtol=1e-4)
    assert loss == pytest.approx(ref.mean().item(), abs=5e-3)
    rel = np.linalg.norm(g - gr) / np.linalg.norm(gr)
    assert rel < 1e-3, rel
    for layer in range(nl):
        sl = slice(layer * U.PC_NP_LAYER, (layer + 1) * U.PC_NP_LAYER)
        assert np.linalg.norm(g[sl] - gr[sl]) <= 2e-3 * np.linalg.norm(gr[sl]) + 1e-6, layer


def test_pcflow_deeper_stack(ctx):
    """a 12-layer stack: samples, inverse log q and gradient against the oracle"""
    torch = pytest.importorskip("torch")
    L, B, seed, step, nl, beta = 8, 3, 100, 4, 12, 3.0
    p = _params(101, 0.15, nl)
    f = U1.PlaqCouplingFlow(L, L, ctx, layers=nl)
    assert f.n == nl * U.PC_NP_LAYER
    f.set_params(p)
    ens = U1.U1Ensemble(L, L, B, ctx)
    out = f.sample(ens, seed, step).download().reshape(B, 2)
    z = torch.tensor(np.stack([U.pcflow_draws(seed, step, c, L, L)[0] for c in range(B)]).astype(np.float64))
    phi, ld = U.pcflow_forward(torch.tensor(p.astype(np.float64)), z, L, L)
    P = _free_plaquettes(ens.angles())
    for c in range(B):
        assert _wrapdiff(P[c], phi[c].numpy()).max() < 2e-3, c
    np.testing.assert_allclose(out[:, 0], U.log_prior(L, L) - ld.numpy(), atol=5e-3)
    inv = f.logq(ens).download().reshape(B, 2)
    np.testing.assert_allclose(inv[:, 0], out[:, 0], atol=3e-3)
    f.gradient(B, beta, seed, step)
    g = f.grad.download().astype(np.float64)
    pt = torch.tensor(p.astype(np.float64), requires_grad=True)
    U.pcflow_loss(pt, z, L, L, beta)[0].mean().backward()
    gr = pt.grad.numpy()
    assert np.linalg.norm(g - gr) / np.linalg.norm(gr) < 2e-3


def test_pcflow_training_and_chains(ctx):
    """Trained at beta 2 on 8x8: the reverse KL falls from its identity value
    and stays above -log Z; HMC + pcflow MH chains reproduce the exact torus
    <cos theta_P> with an MH acceptance well on the way to the 0.426 that any
    proposal blind to the plaquette-sum constraint is limited to."""
    L, beta = 8, 2.0
    f = U1.PlaqCouplingFlow(L, L, ctx, seed=102)
    hist = f.train(beta, 600, batch=128, lr=2e-3, lr_final=2e-4, seed=103)
    bound = -_log_z(beta, L * L)
    start = U.log_prior(L, L) + beta * L * L
    assert abs(hist[:3].mean() - start) < 3.0
    assert hist[-30:].mean() < start - 10.0
    assert hist[-30:].mean() > bound - 0.5, (hist[-30:].mean(), bound)
    B = 256
    ens = U1.U1Ensemble(L, L, B, ctx)
    ens.random_start(104)
    r = U1.run_chains(ens, beta, gauge.makeDeltas("2lf", 0.5, 10), 60, 105, flow=f, n_therm=20, record_every=2)
    assert r["acc_flow"] > 0.25, r["acc_flow"]
    chain_means = r["plaq"].mean(axis=1)
    est, err = chain_means.mean(), chain_means.std(ddof=1) / math.sqrt(B)
    exact = U.exact_mean_cos(beta, L * L)
    assert abs(est - exact) < 5 * err + 1e-4, (est, exact, err)


def test_pcflow_errors(ctx):
    with pytest.raises(mesh_amd.MeshError):  # odd extent
        U1.PlaqCouplingFlow(6, 7, ctx).sample(U1.U1Ensemble(6, 7, 2, ctx), 0, 0)
    with pytest.raises(mesh_amd.MeshError):  # V > 4096
        U1.PlaqCouplingFlow(128, 64, ctx).sample(U1.U1Ensemble(128, 64, 1, ctx), 0, 0)
    with pytest.raises(mesh_amd.MeshError):
        U1.PlaqCouplingFlow(8, 8, ctx, layers=0)
    with pytest.raises(mesh_amd.MeshError):
        U1.PlaqCouplingFlow(8, 8, ctx, layers=65)


def test_pcflow_propose_is_logq_and_sample(ctx):
    """mesh_pcflow_propose = mesh_pcflow_logq of the current configurations
    and mesh_pcflow_sample of the proposals, bit for bit, on a lattice where
    the two halves share CUs (8x8) and on one where they do not (32x32)"""
    for L in (8, 32):
        B, seed, step = 16, 106, 3
        f = U1.PlaqCouplingFlow(L, L, ctx)
        f.set_params(_params(107, 0.2))
        ens = U1.U1Ensemble(L, L, B, ctx)
        ens.random_start(108)
        st = U1.Buffer(2 * B, np.float64, ctx)
        prop = U1.Buffer(B * 2 * L * L, np.float32, ctx)
        pout = U1.Buffer(2 * B, np.float64, ctx)
        f.propose(ens, st, seed, step, prop, pout)
        ref_st = f.logq(ens).download()
        ref_prop = U1.Buffer(B * 2 * L * L, np.float32, ctx)
        ref_out = f.sample(ens, seed, step, theta_out=ref_prop).download()
        np.testing.assert_array_equal(st.download(), ref_st)
        np.testing.assert_array_equal(pout.download(), ref_out)
        np.testing.assert_array_equal(prop.download(), ref_prop.download())
        with pytest.raises(mesh_amd.MeshError):
            f.propose(ens, st, seed, step, ens.theta, pout)
        with pytest.raises(mesh_amd.MeshError):
            f.propose(ens, st, seed, step, prop, st)


def test_pcflow_gain_over_hmc_L16(ctx):
    """At 16^2, beta 4: HMC + MH steps proposed by the trained 8-layer
    plaquette coupling flow decorrelate Q within a few steps where plain HMC
    needs tens of trajectories, with the exact <cos theta_P> and chi_t."""
    L, beta, B = 16, 4.0, 256
    d = gauge.makeDeltas("2lf", 0.5, 20)
    f = U1.PlaqCouplingFlow(L, L, ctx, seed=112)
    hist = f.train(beta, 4000, batch=B, lr=1e-3, lr_final=1e-4, seed=113)
    assert hist[-50:].mean() > -_log_z(beta, L * L) - 0.5
    runs = {}
    for name, flow, n in (("hmc", None, 2000), ("flow", f, 500)):
        e = U1.U1Ensemble(L, L, B, ctx)
        e.random_start(114)
        runs[name] = U1.run_chains(e, beta, d, n, 115, flow=flow, n_therm=300)
    t_hmc, _ = U1.tau_int(runs["hmc"]["Q"])
    t_flow, _ = U1.tau_int(runs["flow"]["Q"])
    r = runs["flow"]
    print(f"16^2 beta 4: tau_int(Q) HMC {t_hmc:.1f}, HMC + pcflow {t_flow:.2f}, flow acceptance {r['acc_flow']:.3f}")
    assert r["acc_flow"] > 0.15, r["acc_flow"]
    assert t_flow < t_hmc / 8, (t_hmc, t_flow)
    chain_means = r["plaq"].mean(axis=1)
    est, err = chain_means.mean(), chain_means.std(ddof=1) / math.sqrt(B)
    exact = U.exact_mean_cos(beta, L * L)
    assert abs(est - exact) < 5 * err + 1e-4, (est, exact, err)
    chi = float(np.mean(r["Q"] ** 2)) / (L * L)
    chi_exact = U.exact_chi_top(beta, L * L)
    assert abs(chi - chi_exact) < 0.1 * chi_exact, (chi, chi_exact)
