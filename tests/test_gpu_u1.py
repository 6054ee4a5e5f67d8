"""U(1) 2-d HMC and the learned flow proposal on the GPU against the oracle:
bit-exact random starts and Metropolis draws, FP32 HMC against the FP64
oracle trajectory, the HMC ensemble against the exact torus <cos theta_P>,
the flow's forward map / log-det / inverse / reverse-KL gradient against the
torch float64 oracle, training against the exact free-energy bound, the
independence Metropolis-Hastings rule, the trained flow on HMC-generated
configurations (forward KL with the exact log Z) and its gain over plain HMC
(tau_int of the topological charge)."""
import math

import numpy as np
import pytest

import mesh_amd
from conftest import options
from mesh_amd import gauge
from mesh_amd import u1 as U1
from oracle import oracle as O
from oracle import u1_oracle as U

pytestmark = pytest.mark.gpu

TAG_FLOWMH = 0x464C4D48


def _wrapdiff(a, b):
    return np.abs(U.wrap(np.asarray(a, np.float64) - np.asarray(b, np.float64)))


def _params(seed, scale, layers=U.NL):
    rng = np.random.default_rng(seed)
    p = rng.normal(size=layers * U.NP_LAYER) * scale
    for layer in range(layers):
        p[layer * U.NP_LAYER + U.NP_LAYER - U.NO + 2 * U.NK] += 1.0
    return p.astype(np.float32)


def _log_z(beta, V):
    """log of int prod dtheta exp(-beta S) on a periodic 2-d torus:
    (2 pi)^{2V} e^{-beta V} sum_n I_n(beta)^V (character expansion)."""
    import scipy.special as sp

    terms = [V * math.log(sp.ive(n, beta)) for n in range(-30, 31)]
    m = max(terms)
    return 2 * V * math.log(2 * math.pi) + m + math.log(sum(math.exp(t - m) for t in terms))


@pytest.mark.parametrize("L0,L1", [(8, 8), (4, 12), (6, 10)])
def test_u1_random_bitexact(ctx, L0, L1):
    ens = U1.U1Ensemble(L0, L1, 3, ctx)
    ens.random_start(11)
    got = ens.angles()
    for c in range(3):
        np.testing.assert_array_equal(got[c], U.random_angles(11, c, L0, L1))


def test_u1_observables(ctx):
    L0, L1, B = 8, 12, 6
    ens = U1.U1Ensemble(L0, L1, B, ctx)
    ens.random_start(12)
    th = ens.angles().astype(np.float64)
    o = ens.observables()
    P = U.plaq(th)
    np.testing.assert_allclose(o["S"], np.sum(1 - np.cos(P), axis=(-2, -1)), rtol=1e-5)
    np.testing.assert_allclose(o["cos"], np.mean(np.cos(P), axis=(-2, -1)), atol=1e-5)
    np.testing.assert_array_equal(o["Q"], U.charge(th))


@pytest.mark.parametrize("scheme,n", [("2lf", 6), ("2mn", 3)])
def test_u1_hmc_against_oracle(ctx, scheme, n):
    """One FP32 trajectory per chain against the FP64 oracle from the same
    angles and the same (bit-identical) momenta; Metropolis uniforms exact."""
    L0, L1, B, beta, seed, traj = 8, 8, 4, 2.0, 21, 3
    ens = U1.U1Ensemble(L0, L1, B, ctx)
    ens.random_start(seed)
    th0 = ens.angles().astype(np.float64)
    d = gauge.makeDeltas(scheme, 0.5, n)
    ens.hmc(beta, d, seed, traj, metropolis=False)
    raw = ens._hmc.download().reshape(B, 4)
    th1 = ens.angles()
    for c in range(B):
        q = U.momenta(seed, traj, c, L0, L1)
        t_ref, _, dh_ref = U.hmc_trajectory(th0[c], q, beta, d)
        assert _wrapdiff(th1[c], t_ref).max() < 2e-4
        assert abs(raw[c, 0] - dh_ref) < 2e-3 * (1 + abs(dh_ref)), (c, raw[c, 0], dh_ref)
        assert raw[c, 2] == U.metropolis_uniform(seed, traj, c)
        assert raw[c, 3] == pytest.approx(U.action_over_beta(t_ref), rel=1e-4)


def test_u1_hmc_rejection_keeps_field(ctx):
    """A huge step size makes Delta H large: every chain rejects and the
    angles stay bit-unchanged; acceptance follows u < exp(-dH) exactly."""
    ens = U1.U1Ensemble(8, 8, 8, ctx)
    ens.random_start(5)
    before = ens.angles().copy()
    r = ens.hmc(4.0, gauge.makeDeltas("2lf", 3.0, 2), 5, 0)
    raw = ens._hmc.download().reshape(8, 4)
    after = ens.angles()
    for c in range(8):
        a = raw[c, 2] < math.exp(-raw[c, 0])
        assert bool(r["acc"][c]) == a
        if not a:
            np.testing.assert_array_equal(after[c], before[c])


def test_u1_hmc_ensemble_matches_exact(ctx):
    """<cos theta_P> of 512 independent chains (8x8, beta 2) after
    thermalisation against the exact torus value, within 5 sigma."""
    L, B, beta = 8, 512, 2.0
    ens = U1.U1Ensemble(L, L, B, ctx)
    ens.random_start(31)
    d = gauge.makeDeltas("2mn", 1.0, 5)
    acc = []
    vals = []
    for t in range(160):
        r = ens.hmc(beta, d, 31, t)
        if t >= 40:
            acc.append(r["acc"].mean())
            if t % 4 == 0:
                vals.append(ens.observables()["cos"])
    v = np.array(vals)          # (measurements, chains)
    chain_means = v.mean(axis=0)
    est = chain_means.mean()
    err = chain_means.std(ddof=1) / math.sqrt(B)
    exact = U.exact_mean_cos(beta, L * L)
    assert np.mean(acc) > 0.6
    assert abs(est - exact) < 5 * err + 1e-4, (est, exact, err)


def test_u1_hmc_ensemble_matches_exact_32(ctx):
    """The 32x32 lattice of the learned-proposal workload: <cos theta_P> of
    256 chains (beta 2, 2lf with 20 steps, tau = 1, the workload's HMC) after
    thermalisation from a random start against the exact torus value."""
    L, B, beta = 32, 256, 2.0
    ens = U1.U1Ensemble(L, L, B, ctx)
    ens.random_start(33)
    d = gauge.makeDeltas("2lf", 0.5, 20)
    acc, vals = [], []
    for t in range(120):
        r = ens.hmc(beta, d, 33, t)
        if t >= 40:
            acc.append(r["acc"].mean())
            if t % 4 == 0:
                vals.append(ens.observables()["cos"])
    chain_means = np.array(vals).mean(axis=0)
    est, err = chain_means.mean(), chain_means.std(ddof=1) / math.sqrt(B)
    exact = U.exact_mean_cos(beta, L * L)
    print("32x32 beta 2:", est, err, exact, np.mean(acc))
    assert np.mean(acc) > 0.8
    assert abs(est - exact) < 5 * err + 1e-4, (est, exact, err)


def test_flow_identity_initialisation(ctx):
    """the initial weights with the last layer zeroed are the identity map;
    the initial last layer is small (it only breaks the components' symmetry)"""
    L, B = 8, 4
    f = U1.Flow(L, L, ctx)
    p = f.get_params()
    for layer in range(U.NL):
        o = layer * U.NP_LAYER + U.NP_LAYER - U.NO - U.NH * U.NO
        w3 = p[o:o + U.NH * U.NO]
        assert 0 < np.abs(w3).max() < 0.01
        p[o:o + U.NH * U.NO] = 0
    f.set_params(p)
    ens = U1.U1Ensemble(L, L, B, ctx)
    out = f.sample(ens, 3, 0).download().reshape(B, 2)
    th = ens.angles()
    for c in range(B):
        z = U.flow_z(3, 0, c, L, L)
        assert _wrapdiff(th[c], z).max() < 1e-5
    assert np.abs(out[:, 0] - U.log_prior(L, L)).max() < 1e-3


@pytest.mark.parametrize("L0,L1", [(8, 8), (4, 8), (16, 16), (32, 32)])
def test_flow_forward_against_oracle(ctx, L0, L1):
    """coupling-layer flow samples and log q against the torch float64 oracle,
    up to the 32x32 lattice of the learned-proposal workload (log q sums one
    FP32 log-det term per plaquette per layer: its bound grows with V)"""
    torch = pytest.importorskip("torch")
    B, seed, step = 4, 41, 7
    tol_lq = 5e-3 * max(1.0, L0 * L1 / 64)
    p = _params(42, 0.2)
    f = U1.Flow(L0, L1, ctx)
    f.set_params(p)
    ens = U1.U1Ensemble(L0, L1, B, ctx)
    out = f.sample(ens, seed, step).download().reshape(B, 2)
    th = ens.angles()
    z = torch.tensor(np.stack([U.flow_z(seed, step, c, L0, L1) for c in range(B)]).astype(np.float64))
    t_ref, ld = U.flow_forward(torch.tensor(p.astype(np.float64)), z)
    t_ref, ld = t_ref.numpy(), ld.numpy()
    assert np.abs(ld).max() > 0.5  # a non-trivial map
    for c in range(B):
        assert _wrapdiff(th[c], t_ref[c]).max() < 2e-3, c
    np.testing.assert_allclose(out[:, 0], U.log_prior(L0, L1) - ld, atol=tol_lq)
    np.testing.assert_allclose(out[:, 1], U.action_over_beta(th.astype(np.float64)), rtol=1e-4)


def test_flow_deeper_stack_against_oracle(ctx):
    """a 20-layer stack (the 8-layer pattern repeated): samples, log q by the
    inverse flow and the reverse-KL gradient against the torch float64 oracle"""
    torch = pytest.importorskip("torch")
    L, B, seed, step, nl, beta = 8, 4, 51, 3, 20, 2.0
    p = _params(52, 0.15, nl)
    f = U1.Flow(L, L, ctx, layers=nl)
    assert f.n == nl * U.NP_LAYER
    f.set_params(p)
    ens = U1.U1Ensemble(L, L, B, ctx)
    out = f.sample(ens, seed, step).download().reshape(B, 2)
    th = ens.angles()
    z = torch.tensor(np.stack([U.flow_z(seed, step, c, L, L) for c in range(B)]).astype(np.float64))
    t_ref, ld = U.flow_forward(torch.tensor(p.astype(np.float64)), z)
    for c in range(B):
        assert _wrapdiff(th[c], t_ref[c].numpy()).max() < 2e-3, c
    np.testing.assert_allclose(out[:, 0], U.log_prior(L, L) - ld.numpy(), atol=1e-2)
    inv = f.logq(ens).download().reshape(B, 2)
    np.testing.assert_allclose(inv[:, 0], out[:, 0], atol=3e-3)
    f.gradient(B, beta, seed, step)
    g = f.grad.download().astype(np.float64)
    pt = torch.tensor(p.astype(np.float64), requires_grad=True)
    ref_loss, _, _, _ = U.flow_loss(pt, z, beta)
    ref_loss.backward()
    gr = pt.grad.numpy()
    assert np.linalg.norm(g - gr) / np.linalg.norm(gr) < 2e-3
    with pytest.raises(mesh_amd.MeshError):
        U1.Flow(L, L, ctx, layers=65)


def test_flow_inverse_roundtrip(ctx):
    """log q by the inverse flow of a sample equals the sampler's log q and
    recovers its prior draw z."""
    L, B, seed, step = 8, 4, 43, 2
    f = U1.Flow(L, L, ctx)
    f.set_params(_params(44, 0.2))
    ens = U1.U1Ensemble(L, L, B, ctx)
    fwd = f.sample(ens, seed, step).download().reshape(B, 2)
    z_out = U1.Buffer(B * 2 * L * L, np.float32, ctx)
    inv = f.logq(ens, z_out=z_out).download().reshape(B, 2)
    z = z_out.download().reshape(B, 2, L, L)
    for c in range(B):
        assert _wrapdiff(z[c], U.flow_z(seed, step, c, L, L)).max() < 1e-3
    np.testing.assert_allclose(inv[:, 0], fwd[:, 0], atol=2e-3)
    np.testing.assert_allclose(inv[:, 1], fwd[:, 1], rtol=1e-6)


def test_flow_gradient_against_autograd(ctx):
    """mesh_flow_gradient against torch float64 autograd of
    mean(beta S(f(z)) - log det f'(z)) on the same prior draws."""
    torch = pytest.importorskip("torch")
    L, B, beta, seed, step = 8, 8, 2.0, 45, 1
    p = _params(46, 0.2)
    f = U1.Flow(L, L, ctx)
    f.set_params(p)
    loss, logq, S = f.gradient(B, beta, seed, step)
    g = f.grad.download().astype(np.float64)
    pt = torch.tensor(p.astype(np.float64), requires_grad=True)
    z = torch.tensor(np.stack([U.flow_z(seed, step, c, L, L) for c in range(B)]).astype(np.float64))
    ref_loss, _, ld, Sref = U.flow_loss(pt, z, beta)
    ref_loss.backward()
    gr = pt.grad.numpy()
    assert loss == pytest.approx(U.log_prior(L, L) + ref_loss.item(), abs=5e-3)
    np.testing.assert_allclose(S, Sref.detach().numpy(), rtol=1e-4)
    rel = np.linalg.norm(g - gr) / np.linalg.norm(gr)
    assert rel < 1e-3, rel
    # per layer as well, so that no layer's gradient hides under another's
    for layer in range(U.NL):
        sl = slice(layer * U.NP_LAYER, (layer + 1) * U.NP_LAYER)
        assert np.linalg.norm(g[sl] - gr[sl]) <= 2e-3 * np.linalg.norm(gr[sl]) + 1e-6, layer


def test_flow_training_respects_free_energy_bound(ctx):
    """Adam on the reverse KL lowers E[log q + beta S] from its identity value
    and never below -log Z (the KL is non-negative)."""
    L, beta = 8, 2.0
    f = U1.Flow(L, L, ctx)
    hist = f.train(beta, 300, batch=128, lr=2e-3, seed=47)
    bound = -_log_z(beta, L * L)
    start = U.log_prior(L, L) + beta * L * L  # identity map: E[S] = V under the uniform prior
    assert abs(hist[:3].mean() - start) < 3.0
    assert hist[-30:].mean() < start - 10.0, (hist[:3].mean(), hist[-30:].mean())
    assert hist[-30:].mean() > bound - 0.5, (hist[-30:].mean(), bound)


def test_flow_mh_rule(ctx):
    """Independence MH: acceptance exactly u < exp(lw' - lw) with the Philox
    uniform of (step, chain); accepted chains take the proposal bits, the
    others keep theirs, and the state's (log q, S/beta) follows."""
    L, B, beta, seed, step = 8, 16, 2.0, 48, 9
    f = U1.Flow(L, L, ctx)
    f.set_params(_params(49, 0.1))
    ens = U1.U1Ensemble(L, L, B, ctx)
    ens.random_start(50)
    state = U1.Buffer(2 * B, np.float64, ctx)
    f.logq(ens, out=state)
    prop = U1.Buffer(B * 2 * L * L, np.float32, ctx)
    pout = f.sample(ens, seed + 1, step, theta_out=prop)
    st0, pr = state.download().reshape(B, 2), pout.download().reshape(B, 2)
    th0, pth = ens.angles().copy(), prop.download().reshape(B, 2, L, L)
    acc = U1.Buffer(B, np.float64, ctx)
    ctx.call("mesh_flow_mh", ens.grid.ref, B, ens.theta.ptr, state.ptr, prop.ptr, pout.ptr, U1._d(beta),
             U1.C.c_uint64(seed), U1.C.c_uint32(step), acc.ptr)
    a = acc.download()
    st1, th1 = state.download().reshape(B, 2), ens.angles()
    for c in range(B):
        u = O.philox_draw(seed, step, c, 1, TAG_FLOWMH)[0]
        lw_new = -beta * pr[c, 1] - pr[c, 0]
        lw_old = -beta * st0[c, 1] - st0[c, 0]
        expect = u < math.exp(lw_new - lw_old)
        assert bool(a[c]) == expect, c
        np.testing.assert_array_equal(th1[c], pth[c] if expect else th0[c])
        np.testing.assert_array_equal(st1[c], pr[c] if expect else st0[c])


def test_flow_errors(ctx):
    with pytest.raises(mesh_amd.MeshError):
        U1.Flow(6, 8, ctx).sample(U1.U1Ensemble(6, 8, 2, ctx), 0, 0)
    with pytest.raises((mesh_amd.MeshError, ValueError)):  # U(1) lattices have 2 to 4 dimensions
        U1.U1Ensemble(0, 0, 1, ctx, dims=(16,)).hmc(1.0, [0.1], 0, 0)
    with pytest.raises(mesh_amd.MeshError):  # step tables have odd length
        U1.U1Ensemble(128, 128, 1, ctx).hmc(1.0, [0.1, 0.2], 0, 0)


def _pparams(seed, scale):
    rng = np.random.default_rng(seed)
    p = rng.normal(size=U.NPF) * scale
    p[2 * U.PK] += 1.0
    return p.astype(np.float32)


def test_pflow_identity(ctx):
    """identity map: the free plaquettes of a proposal are the uniform draws
    and log q is the uniform density (2 pi)^-2V"""
    L, B = 8, 3
    f = U1.PlaqFlow(L, L, ctx)
    p = f.get_params()
    np.testing.assert_allclose(p[:U.PK], np.linspace(-3, 1, U.PK), atol=1e-6)  # distinct components
    p[:U.PK] = 0
    f.set_params(p)
    ens = U1.U1Ensemble(L, L, B, ctx)
    out = f.sample(ens, 5, 1).download().reshape(B, 2)
    th = ens.angles().astype(np.float64)
    for c in range(B):
        u, _, _, _ = U.pflow_draws(5, 1, c, L, L)
        P = U.plaq(th[c]).reshape(-1)
        assert _wrapdiff(P[:-1], u).max() < 2e-5
    np.testing.assert_allclose(out[:, 0], -2 * L * L * math.log(2 * math.pi), atol=1e-3)


@pytest.mark.parametrize("L0,L1", [(8, 8), (6, 10), (16, 16), (32, 32)])
def test_pflow_sample_against_oracle(ctx, L0, L1):
    """plaquette-flow samples and log q against the torch float64 oracle, up
    to the 32x32 lattice of the learned-proposal workload"""
    torch = pytest.importorskip("torch")
    B, seed, step = 3, 61, 4
    tol_lq = 5e-3 * max(1.0, L0 * L1 / 64)
    p = _pparams(62, 0.5)
    f = U1.PlaqFlow(L0, L1, ctx)
    f.set_params(p)
    ens = U1.U1Ensemble(L0, L1, B, ctx)
    out = f.sample(ens, seed, step).download().reshape(B, 2)
    th = ens.angles().astype(np.float64)
    pt = torch.tensor(p.astype(np.float64))
    for c in range(B):
        u, a, h0, h1 = U.pflow_draws(seed, step, c, L0, L1)
        thP, nlg = U.pflow_h(pt, torch.tensor(u.astype(np.float64)))
        ref = U.pflow_links(thP.numpy(), h0, h1, a, L0, L1)
        assert _wrapdiff(th[c], ref).max() < 2e-4, c
        lq, S = U.pflow_logq(p, th[c])
        assert abs(out[c, 0] - lq) < tol_lq, (c, out[c, 0], lq)
        assert abs(lq - (nlg.sum().item() - 2 * L0 * L1 * math.log(2 * math.pi))) < tol_lq
        assert out[c, 1] == pytest.approx(S, rel=1e-5)


def test_pflow_logq_of_given_configurations(ctx):
    L0, L1, B = 8, 12, 4
    p = _pparams(63, 0.7)
    f = U1.PlaqFlow(L0, L1, ctx)
    f.set_params(p)
    ens = U1.U1Ensemble(L0, L1, B, ctx)
    ens.random_start(64)
    out = f.logq(ens).download().reshape(B, 2)
    lq, S = U.pflow_logq(p, ens.angles().astype(np.float64))
    np.testing.assert_allclose(out[:, 0], lq, atol=5e-3)
    np.testing.assert_allclose(out[:, 1], S, rtol=1e-5)


def test_pflow_propose(ctx):
    """mesh_pflow_propose: the states' (log q, S/beta) are mesh_pflow_logq's
    bit for bit, the proposals mesh_pflow_sample's, and the proposals' log q
    (from the draws) that of the inverse map to rounding; aliased outputs are
    refused"""
    for L in (8, 32):
        B, seed, step = 16, 201, 4
        f = U1.PlaqFlow(L, L, ctx)
        f.set_params(_pparams(202, 0.5))
        ens = U1.U1Ensemble(L, L, B, ctx)
        ens.random_start(203)
        st = U1.Buffer(2 * B, np.float64, ctx)
        prop = U1.Buffer(B * 2 * L * L, np.float32, ctx)
        pout = U1.Buffer(2 * B, np.float64, ctx)
        f.propose(ens, st, seed, step, prop, pout)
        np.testing.assert_array_equal(st.download(), f.logq(ens).download())
        ref = U1.Buffer(B * 2 * L * L, np.float32, ctx)
        ctx.call("mesh_pflow_sample", ens.grid.ref, B, f.params.ptr, U1.C.c_uint64(seed), U1.C.c_uint32(step),
                 ref.ptr)
        np.testing.assert_array_equal(prop.download(), ref.download())
        e2 = U1.U1Ensemble(L, L, B, ctx)
        e2.set_angles(prop.download().reshape(B, 2, L, L))
        inv = f.logq(e2).download().reshape(B, 2)
        po = pout.download().reshape(B, 2)
        np.testing.assert_allclose(po[:, 0], inv[:, 0], atol=2e-3)
        np.testing.assert_array_equal(po[:, 1], inv[:, 1])
        with pytest.raises(mesh_amd.MeshError):
            f.propose(ens, st, seed, step, ens.theta, pout)
        with pytest.raises(mesh_amd.MeshError):
            f.propose(ens, st, seed, step, prop, st)


def test_pflow_gradient_against_autograd(ctx):
    torch = pytest.importorskip("torch")
    L, B, beta, seed, step = 8, 6, 2.0, 65, 3
    p = _pparams(66, 0.5)
    f = U1.PlaqFlow(L, L, ctx)
    f.set_params(p)
    loss, logq, S = f.gradient(B, beta, seed, step)
    g = f.grad.download().astype(np.float64)
    pt = torch.tensor(p.astype(np.float64), requires_grad=True)
    u = torch.tensor(np.stack([U.pflow_draws(seed, step, c, L, L)[0] for c in range(B)]).astype(np.float64))
    ref, ref_lq, ref_S = U.pflow_loss(pt, u, beta)
    ref.mean().backward()
    np.testing.assert_allclose(logq, ref_lq.detach().numpy(), atol=2e-3)
    np.testing.assert_allclose(S, ref_S.detach().numpy(), rtol=1e-5)
    assert loss == pytest.approx(ref.mean().item(), abs=5e-3)
    gr = pt.grad.numpy()
    assert np.linalg.norm(g - gr) <= 1e-3 * np.linalg.norm(gr), (g, gr)


def test_pflow_training_and_chains(ctx):
    """Trained at beta 2 on 8x8: the reverse KL stays above -log Z, the MH
    acceptance of its proposals is high, and HMC + pflow MH chains reproduce
    the exact torus <cos theta_P>."""
    L, beta = 8, 2.0
    f = U1.PlaqFlow(L, L, ctx)
    hist = f.train(beta, 400, batch=256, lr=2e-2, lr_final=2e-3, seed=67)
    bound = -_log_z(beta, L * L)
    assert hist[-40:].mean() > bound - 0.3, (hist[-40:].mean(), bound)
    assert hist[-40:].mean() < hist[:5].mean() - 10
    B = 512
    ens = U1.U1Ensemble(L, L, B, ctx)
    ens.random_start(68)
    r = U1.run_chains(ens, beta, gauge.makeDeltas("2lf", 0.5, 10), 60, 69, flow=f, n_therm=20, record_every=2)
    assert r["acc_flow"] > 0.3, r["acc_flow"]
    chain_means = r["plaq"].mean(axis=1)
    est, err = chain_means.mean(), chain_means.std(ddof=1) / math.sqrt(B)
    exact = U.exact_mean_cos(beta, L * L)
    assert abs(est - exact) < 5 * err + 1e-4, (est, exact, err)


def test_pflow_validated_on_hmc_configurations(ctx):
    """The plaquette flow evaluated on configurations HMC produced (not on its
    own samples): the forward KL E_p[log p - log q], with the exact log Z of
    the torus, is non-negative within errors, small once trained, and far
    below the untrained flow's."""
    L, beta, B = 8, 2.0, 512
    lz = _log_z(beta, L * L)
    ens = U1.U1Ensemble(L, L, B, ctx)
    ens.random_start(71)
    U1.run_chains(ens, beta, gauge.makeDeltas("2lf", 0.5, 10), 1, 72, n_therm=100)
    f = U1.PlaqFlow(L, L, ctx)

    def forward_kl():
        o = f.logq(ens).download().reshape(B, 2)
        x = -beta * o[:, 1] - lz - o[:, 0]
        return x.mean(), x.std(ddof=1) / math.sqrt(B)

    k0, _ = forward_kl()
    f.train(beta, 400, batch=256, lr=2e-2, lr_final=2e-3, seed=73)
    k1, e1 = forward_kl()
    print(f"forward KL on HMC configurations: untrained {k0:.3f}, trained {k1:.4f} +- {e1:.4f}")
    assert k1 > -5 * e1 - 1e-3, (k1, e1)
    assert k1 < 1.0, k1
    assert k1 < k0 - 10, (k0, k1)


def test_pflow_gain_over_hmc(ctx):
    """What the trained network adds to the sampler: at beta 6 on 8x8 plain
    HMC rarely changes the topological charge; interleaved 1:1 with MH steps
    proposed by the trained plaquette flow it decorrelates Q within a few
    steps, and with the untrained flow the proposals are not accepted."""
    L, beta, B = 8, 6.0, 256
    d = gauge.makeDeltas("2lf", 0.5, 20)
    f = U1.PlaqFlow(L, L, ctx)
    ens = U1.U1Ensemble(L, L, B, ctx)
    ens.random_start(81)
    r0 = U1.run_chains(ens, beta, d, 20, 82, flow=f, n_therm=100)
    f.train(beta, 1000, batch=256, lr=1e-2, lr_final=1e-3, seed=83)
    runs = {}
    for name, flow in (("hmc", None), ("flow", f)):
        e = U1.U1Ensemble(L, L, B, ctx)
        e.random_start(84)
        runs[name] = U1.run_chains(e, beta, d, 400, 85, flow=flow, n_therm=100)
    t_hmc, _ = U1.tau_int(runs["hmc"]["Q"])
    t_flow, _ = U1.tau_int(runs["flow"]["Q"])
    print(f"tau_int(Q): HMC {t_hmc:.1f}, HMC + flow {t_flow:.2f}; acceptance untrained "
          f"{r0['acc_flow']:.4f}, trained {runs['flow']['acc_flow']:.3f}")
    assert r0["acc_flow"] < 0.01, r0["acc_flow"]
    assert runs["flow"]["acc_flow"] > 0.05, runs["flow"]["acc_flow"]
    assert t_flow < t_hmc / 5, (t_hmc, t_flow)


# ---------------------------------------------------------------- any dimension

@pytest.mark.parametrize("dims", [(4, 2, 3, 2), (3, 4, 2), (96, 100)])
def test_u1_nd_random_bitexact(ctx, dims):
    """Site-parallel random start (3-d, 4-d, and a 2-d lattice too large for
    the LDS kernels) against the oracle's Philox streams."""
    B = 2
    ens = U1.U1Ensemble(0, 0, B, ctx, dims=dims)
    ens.random_start(41)
    th = ens.angles()
    for c in range(B) if np.prod(dims) < 1000 else range(1):
        np.testing.assert_array_equal(th[c], U.nd_random_angles(41, c, dims))


@pytest.mark.parametrize("dims", [(4, 4, 4, 4), (4, 6, 4)])
def test_u1_nd_hmc_against_oracle(ctx, dims):
    """One FP32 trajectory of the site-parallel kernels against the FP64
    oracle from the same angles and momenta; observables of the result."""
    B, beta, seed, traj = 2, 1.5, 43, 2
    ens = U1.U1Ensemble(0, 0, B, ctx, dims=dims)
    ens.random_start(seed)
    th0 = ens.angles().astype(np.float64)
    d = gauge.makeDeltas("2lf", 0.5, 6)
    ens.hmc(beta, d, seed, traj, metropolis=False)
    raw = ens._hmc.download().reshape(B, 4)
    th1 = ens.angles()
    o = ens.observables()
    for c in range(B):
        q = U.nd_momenta(seed, traj, c, dims)
        t_ref, _, dh_ref = U.nd_hmc_trajectory(th0[c], q, beta, d)
        assert _wrapdiff(th1[c], t_ref).max() < 2e-4
        assert abs(raw[c, 0] - dh_ref) < 2e-3 * (1 + abs(dh_ref)), (c, raw[c, 0], dh_ref)
        assert raw[c, 2] == U.metropolis_uniform(seed, traj, c)
        S = U.nd_action_over_beta(th1[c].astype(np.float64))
        assert raw[c, 3] == pytest.approx(S, rel=1e-5)
        assert o["S"][c] == pytest.approx(S, rel=1e-5)
        P = U.nd_plaquettes(th1[c].astype(np.float64))
        assert o["cos"][c] == pytest.approx(np.mean([np.cos(p).mean() for p in P]), abs=1e-6)
        assert o["Q"][c] == 0.0


@pytest.mark.parametrize("L0,L1,B", [(8, 12, 6), (32, 32, 8), (64, 64, 4)])
def test_u1_general_path_matches_lds_path_2d(ctx, L0, L1, B):
    """On a 2-d lattice the site-parallel kernels give the LDS kernels' bits:
    the same fields after three Metropolis trajectories, the same decisions.
    At 32^2 and 64^2 each of the LDS kernel's four waves handles many sites,
    so a wave that starts moving angles before the others have summed their
    start energy would change Delta H and the decisions."""
    d = gauge.makeDeltas("2mn", 0.5, 4)
    res = {}
    for general in (0, 1):
        with options(ctx, u1_general=general):
            ens = U1.U1Ensemble(L0, L1, B, ctx)
            ens.random_start(47)
            rs = [ens.hmc(2.0, d, 47, t) for t in range(3)]
            res[general] = (ens.angles(), [r["acc"] for r in rs], [r["dH"] for r in rs])
    np.testing.assert_array_equal(res[0][0], res[1][0])
    for a, b in zip(res[0][1], res[1][1]):
        np.testing.assert_array_equal(a, b)
    for a, b in zip(res[0][2], res[1][2]):
        np.testing.assert_allclose(a, b, rtol=0, atol=1e-9)


def test_u1_4d_strong_coupling(ctx):
    """4-d U(1) at beta = 0.2: <cos theta_P> = I1(beta)/I0(beta) + O(u^5)
    (u^5 = 1e-5 here) from 64 chains of HMC."""
    import scipy.special as sp

    beta, B = 0.2, 64
    ens = U1.U1Ensemble(0, 0, B, ctx, dims=(4, 4, 4, 4))
    ens.random_start(51)
    d = gauge.makeDeltas("2lf", 0.5, 8)
    means = []
    for t in range(160):
        r = ens.hmc(beta, d, 52, t)
        if t >= 40:
            means.append(ens.observables()["cos"])
    m = np.array(means)  # (steps, chains)
    chain_means = m.mean(axis=0)
    est, err = chain_means.mean(), chain_means.std(ddof=1) / math.sqrt(B)
    u = sp.iv(1, beta) / sp.iv(0, beta)
    assert r["acc"].mean() > 0.5
    assert abs(est - u) < 5 * err + 3e-5, (est, u, err)
