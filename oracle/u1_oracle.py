"""CPU oracle of the U(1) 2-d path: observables, HMC, exact finite-volume
results, Landau gauge on angles, and the learned proposals (the coupling-layer
flow on links, the plaquette flow, the coupling-layer flow on plaquettes).

TEST INFRASTRUCTURE ONLY (imported by tests/ and bench.py's cpu_baseline;
never by mesh_amd).  Conventions follow mesh_amd/csrc/u1.hip, flow.hip,
pflow.hip and pcflow.hip,
which follow the reference's U(1) algebra (src/gauge/groups.h:29-108):
angles theta[mu][x0][x1], theta_P = theta_0(x) + theta_1(x+0) - theta_0(x+1)
- theta_1(x), S = beta sum (1 - cos theta_P), momenta q ~ N(0, 1/2)
(groups.h:86-91), force (beta/2) Im(U S).

Numerics: numpy float64 for the lattice parts; the flow is restated in torch
(CPU, float64) so that its gradients come from autograd.
"""
import math

import numpy as np

from . import oracle as O

TAG_U1MOM = 0x55314D4F
TAG_U1METRO = 0x55314D45
TAG_U1RAND = 0x55315244
TAG_FLOWZ = 0x464C575A


def wrap(a):
    return np.remainder(np.asarray(a) + np.pi, 2 * np.pi) - np.pi


def plaq(th):
    """th: (..., 2, L0, L1) -> theta_P (..., L0, L1) (unwrapped)."""
    t0, t1 = th[..., 0, :, :], th[..., 1, :, :]
    return t0 + np.roll(t1, -1, axis=-2) - np.roll(t0, -1, axis=-1) - t1


def action_over_beta(th):
    return np.sum(1 - np.cos(plaq(th)), axis=(-2, -1))


def charge(th):
    return np.rint(np.sum(wrap(plaq(th)), axis=(-2, -1)) / (2 * np.pi))


def dS(th, beta):
    """dS/dtheta, shape of th."""
    s = np.sin(plaq(th))
    g0 = beta * (s - np.roll(s, 1, axis=-1))
    g1 = beta * (np.roll(s, 1, axis=-2) - s)
    return np.stack([g0, g1], axis=-3)


def _draw(seed, c0, c1, c2, c3):
    return O.philox_draw(seed, c0, c1, c2, c3)


def random_angles(seed, chain, L0, L1):
    V = L0 * L1
    out = np.zeros((2, V), np.float32)
    for mu in range(2):
        for x in range(V):
            u = _draw(seed, x, (chain << 1) | mu, 0, TAG_U1RAND)[0] - 2.0 ** -53
            out[mu, x] = np.float32(6.283185307179586477 * u - 3.141592653589793238)
    return out.reshape(2, L0, L1)


def momenta(seed, traj, chain, L0, L1):
    V = L0 * L1
    q = np.zeros((2, V), np.float32)
    for mu in range(2):
        for x in range(V):
            q[mu, x] = np.float32(_draw(seed, x, (chain << 1) | mu, traj << 5, TAG_U1MOM)[1] * 0.70710678118654752440)
    return q.reshape(2, L0, L1)


def metropolis_uniform(seed, traj, chain):
    return _draw(seed, traj, chain, 0, TAG_U1METRO)[0]


def hmc_trajectory(th, q, beta, deltas):
    """float64 leapfrog/Omelyan over a step table (hmc.h:38-59); returns
    (theta', q', Delta H)."""
    th = np.asarray(th, np.float64).copy()
    q = np.asarray(q, np.float64).copy()
    e0 = beta * action_over_beta(th) + np.sum(q * q)
    for i, d in enumerate(deltas):
        if i % 2 == 0:
            th = wrap(th + d * q)
        else:
            q = q - d * 0.5 * dS(th, beta)
    e1 = beta * action_over_beta(th) + np.sum(q * q)
    return th, q, e1 - e0


# ------------------------------------------------- any dimension (2 <= nd <= 4)
# th: (nd, L0, ..., L_{nd-1}); th[mu] shifted along axis mu is th[mu](x + mu).

def _ctr(chain, mu, step):
    """Philox counter words 1 and 2 of link (chain, mu): bit 0 of mu in word 1,
    bit 1 of mu in bit 4 of word 2 (mesh_amd/csrc/u1.hip u1_ctr)."""
    return (chain << 1) | (mu & 1), (step << 5) | ((mu >> 1) << 4)


def nd_random_angles(seed, chain, dims):
    nd, V = len(dims), int(np.prod(dims))
    out = np.zeros((nd, V), np.float32)
    for mu in range(nd):
        c1, c2 = _ctr(chain, mu, 0)
        for x in range(V):
            u = _draw(seed, x, c1, c2, TAG_U1RAND)[0] - 2.0 ** -53
            out[mu, x] = np.float32(6.283185307179586477 * u - 3.141592653589793238)
    return out.reshape(nd, *dims)


def nd_momenta(seed, traj, chain, dims):
    nd, V = len(dims), int(np.prod(dims))
    q = np.zeros((nd, V), np.float32)
    for mu in range(nd):
        c1, c2 = _ctr(chain, mu, traj)
        for x in range(V):
            q[mu, x] = np.float32(_draw(seed, x, c1, c2, TAG_U1MOM)[1] * 0.70710678118654752440)
    return q.reshape(nd, *dims)


def nd_plaquettes(th):
    """theta_{mu nu}(x) for the planes mu < nu in the order (0,1), (0,2), ...:
    theta_mu(x) + theta_nu(x+mu) - theta_mu(x+nu) - theta_nu(x)."""
    nd = th.shape[0]
    return [th[mu] + np.roll(th[nu], -1, axis=mu) - np.roll(th[mu], -1, axis=nu) - th[nu]
            for mu in range(nd) for nu in range(mu + 1, nd)]


def nd_action_over_beta(th):
    return float(sum(np.sum(1 - np.cos(p)) for p in nd_plaquettes(th)))


def nd_dS(th, beta):
    """dS/dtheta_mu(x) = beta sum_{nu != mu} [sin theta_{mu nu}(x) - sin theta_{mu nu}(x - nu)]."""
    nd = th.shape[0]
    P = nd_plaquettes(th)
    k = 0
    s = {}
    for mu in range(nd):
        for nu in range(mu + 1, nd):
            s[mu, nu] = np.sin(P[k])
            s[nu, mu] = -s[mu, nu]
            k += 1
    g = np.zeros_like(th, dtype=np.float64)
    for mu in range(nd):
        for nu in range(nd):
            if nu != mu:
                g[mu] += s[mu, nu] - np.roll(s[mu, nu], 1, axis=nu)
    return beta * g


def nd_hmc_trajectory(th, q, beta, deltas):
    """float64 trajectory over a step table (hmc.h:38-59) in any dimension;
    returns (theta', q', Delta H)."""
    th = np.asarray(th, np.float64).copy()
    q = np.asarray(q, np.float64).copy()
    e0 = beta * nd_action_over_beta(th) + np.sum(q * q)
    for i, d in enumerate(deltas):
        if i % 2 == 0:
            th = wrap(th + d * q)
        else:
            q = q - d * 0.5 * nd_dS(th, beta)
    e1 = beta * nd_action_over_beta(th) + np.sum(q * q)
    return th, q, e1 - e0


# ---------------------------------------------------------------- exact results

def _c_nu(beta, nu):
    from scipy.integrate import quad

    return quad(lambda p: math.exp(beta * math.cos(p)) * math.cos(nu * p), 0, math.pi, limit=200)[0] / math.pi


def exact_mean_cos(beta, V, nmax=30):
    """<cos theta_P> on a periodic torus of V plaquettes: Z = sum_n I_n(beta)^V."""
    from scipy.special import ive

    n = np.arange(-nmax, nmax + 1)
    # scaled Bessel functions: I_n = ive * e^beta; ratios to I_0
    r = ive(n, beta) / ive(0, beta)
    dr = 0.5 * (ive(n - 1, beta) + ive(n + 1, beta)) / ive(0, beta)
    return float(np.sum(r ** (V - 1) * dr) / np.sum(r ** V))


def exact_chi_top(beta, V, nmax=12):
    """<Q^2>/V for the geometric charge on a periodic torus:
    Z(theta) = sum_n c(n - theta/2pi)^V, c(nu) = (1/pi) int_0^pi e^{beta cos p} cos(nu p) dp."""
    from scipy.integrate import quad

    num = den = 0.0
    scale = _c_nu(beta, 0)
    for n in range(-nmax, nmax + 1):
        c = _c_nu(beta, n) / scale
        c1 = -quad(lambda p: p * math.exp(beta * math.cos(p)) * math.sin(n * p), 0, math.pi, limit=200)[0] / math.pi
        c2 = -quad(lambda p: p * p * math.exp(beta * math.cos(p)) * math.cos(n * p), 0, math.pi, limit=200)[0] / math.pi
        c1 /= scale
        c2 /= scale
        # d/dtheta = -(1/2pi) d/dnu
        ct1, ct2 = -c1 / (2 * math.pi), c2 / (2 * math.pi) ** 2
        den += c ** V
        num += V * (V - 1) * c ** (V - 2) * ct1 ** 2 + V * c ** (V - 1) * ct2
    return -num / den / V


# ---------------------------------------------------------------- flow (torch)

NL, NK, NI, NH = 8, 4, 12, 64
NO = 2 * NK + 2
NP_LAYER = NI * NH + NH + NH * NH + NH + NH * NO + NO


def flow_z(seed, step, chain, L0, L1):
    """prior sample the GPU draws for (seed, step, chain)."""
    V = L0 * L1
    z = np.zeros((2, V), np.float32)
    for mu in range(2):
        for x in range(V):
            u = _draw(seed, x, (chain << 1) | mu, step << 5, TAG_FLOWZ)[0] - 2.0 ** -53
            z[mu, x] = np.float32(6.283185307179586477 * u - 3.141592653589793238)
    return z.reshape(2, L0, L1)


def _layer_params(p, layer):
    o = layer * NP_LAYER
    W1 = p[o:o + NI * NH].reshape(NI, NH)
    o += NI * NH
    b1 = p[o:o + NH]
    o += NH
    W2 = p[o:o + NH * NH].reshape(NH, NH)
    o += NH * NH
    b2 = p[o:o + NH]
    o += NH
    W3 = p[o:o + NH * NO].reshape(NH, NO)
    o += NH * NO
    b3 = p[o:o + NO]
    return W1, b1, W2, b2, W3, b3


def _tplaq(th):
    import torch

    t0, t1 = th[:, 0], th[:, 1]
    return t0 + torch.roll(t1, -1, dims=1) - torch.roll(t0, -1, dims=2) - t1


def _twrap(a):
    import torch

    return torch.remainder(a + math.pi, 2 * math.pi) - math.pi


def flow_forward(params, z, inverse=False):
    """torch float64 flow on a batch z (B, 2, L0, L1).  Returns (theta,
    logdet) with logdet = sum over layers and active plaquettes of log g'(u).
    The same definitions as mesh_amd/csrc/flow.hip (layer masks, features,
    NCP mixture, shift by atan2)."""
    import torch

    th = z.clone()
    B, _, L0, L1 = th.shape
    x0 = torch.arange(L0).view(L0, 1).expand(L0, L1)
    x1 = torch.arange(L1).view(1, L1).expand(L0, L1)
    logdet = torch.zeros(B, dtype=th.dtype)
    nl = params.numel() // NP_LAYER  # 8 in the specified design
    for layer in range(nl):
        mu, a = layer & 1, (layer >> 1) & 3
        W1, b1, W2, b2, W3, b3 = _layer_params(params, layer)
        P = _tplaq(th)
        feats = []
        for jj in range(2):
            for i in (-1, 0, 1):
                if mu == 0:
                    sh = (-i, -(1 + jj))       # feature(y) = P(y + (i, 1 + jj))
                else:
                    sh = (1 + jj, -i)          # feature(y) = P(y + (-1 - jj, i))
                F = torch.roll(P, shifts=sh, dims=(1, 2))
                feats += [torch.cos(F), torch.sin(F)]
        X = torch.stack(feats, dim=-1)                          # (B, L0, L1, 12)
        mask = ((x1 if mu == 0 else x0) % 4) == a
        Xa = X[:, mask]                                         # (B, NA, 12)
        H1 = torch.tanh(Xa @ W1 + b1)
        H2 = torch.tanh(H1 @ W2 + b2)
        Y = H2 @ W3 + b3
        s = torch.exp(torch.clamp(Y[..., :NK], -8.0, 8.0))
        w = torch.softmax(Y[..., NK:2 * NK], dim=-1)
        t = torch.atan2(Y[..., 2 * NK + 1], Y[..., 2 * NK])
        tp = P[:, mask]
        u = _twrap(tp - t)
        c, sn = torch.cos(0.5 * u), torch.sin(0.5 * u)
        G = torch.sum(w * 2 * torch.atan2(s * sn.unsqueeze(-1), c.unsqueeze(-1)), dim=-1)
        Gp = torch.sum(w * s / (c.unsqueeze(-1) ** 2 + s ** 2 * sn.unsqueeze(-1) ** 2), dim=-1)
        delta = G + t - tp
        logdet = logdet + torch.sum(torch.log(Gp), dim=-1)
        D = torch.zeros(B, L0, L1, dtype=th.dtype)
        D[:, mask] = delta
        if mu == 0:
            new0 = _twrap(th[:, 0] + D)
            th = torch.stack([new0, th[:, 1]], dim=1)
        else:
            new1 = _twrap(th[:, 1] + torch.roll(D, 1, dims=1))
            th = torch.stack([th[:, 0], new1], dim=1)
    return th, logdet


def flow_loss(params, z, beta):
    """mean over the batch of beta S(f(z)) - log det (reverse KL up to constants)."""
    import torch

    th, ld = flow_forward(params, z)
    S = torch.sum(1 - torch.cos(_tplaq(th)), dim=(1, 2))
    return torch.mean(beta * S - ld), th, ld, S


def log_prior(L0, L1):
    return -2 * L0 * L1 * math.log(2 * math.pi)



# ---------------------------------------------------------------- plaquette flow
#
# Restatement of mesh_amd/csrc/pflow.hip: V - 1 plaquettes theta_P = h(u),
# h(u) = g(wrap(u - t)) + t, g = sum_k w_k 2 atan(s_k tan(v/2)); links in the
# maximal-tree gauge (theta_1(x0, x1 < L1-1) = 0, theta_0(x0 < L0-1, L1-1) = 0)
# from the plaquettes and holonomies h0, h1, then a gauge transformation a(x).

PK = 8
NPF = 2 * PK + 2
TAG_PFLOW = 0x50464C57


def _uniform_angle(seed, step, a, b):
    u = O.philox_draw(seed, a, b, step << 5, TAG_PFLOW)[0] - 2.0 ** -53
    return np.float32(6.283185307179586477 * u - 3.141592653589793238)


def pflow_draws(seed, step, chain, L0, L1):
    """(u[V-1], a[V], h0, h1) the GPU draws for (seed, step, chain), float32."""
    V = L0 * L1
    cc = chain << 1
    u = np.array([_uniform_angle(seed, step, x, cc) for x in range(V - 1)], np.float32)
    a = np.array([_uniform_angle(seed, step, x, cc | 1) for x in range(V)], np.float32)
    return u, a, _uniform_angle(seed, step, V, cc), _uniform_angle(seed, step, V + 1, cc)


def pflow_links(P_free, h0, h1, a, L0, L1):
    """links theta[2, L0, L1] (float64, wrapped) with free plaquettes P_free
    (V-1, row-major, the last plaquette omitted), holonomies and gauge a[V]."""
    V = L0 * L1
    P = np.zeros(V)
    P[:V - 1] = P_free
    P = P.reshape(L0, L1)
    t0 = np.zeros((L0, L1))
    t1 = np.zeros((L0, L1))
    for x0 in range(L0):
        acc = float(h0) if x0 == L0 - 1 else 0.0
        t0[x0, L1 - 1] = acc
        for x1 in range(L1 - 2, -1, -1):
            acc += P[x0, x1]
            t0[x0, x1] = acc
    acc = float(h1)
    t1[0, L1 - 1] = acc
    for x0 in range(L0 - 1):
        acc += P[x0, L1 - 1] - t0[x0, L1 - 1] + t0[x0, 0]
        t1[x0 + 1, L1 - 1] = acc
    a = np.asarray(a, np.float64).reshape(L0, L1)
    t0 = t0 + a - np.roll(a, -1, axis=0)
    t1 = t1 + a - np.roll(a, -1, axis=1)
    return wrap(np.stack([t0, t1]))


def _pflow_map(params):
    import torch

    y = params
    s = torch.exp(torch.clamp(y[:PK], -8.0, 8.0))
    w = torch.softmax(y[PK:2 * PK], dim=0)
    t = torch.atan2(y[2 * PK + 1], y[2 * PK])
    return s, w, t


def _pflow_g(s, w, v):
    """g(v), g'(v) for v of any shape (torch)"""
    import torch

    c, sn = torch.cos(0.5 * v).unsqueeze(-1), torch.sin(0.5 * v).unsqueeze(-1)
    G = torch.sum(w * 2 * torch.atan2(s * sn, c), dim=-1)
    Gp = torch.sum(w * s / (c * c + s * s * sn * sn), dim=-1)
    return G, Gp


def pflow_h(params, u):
    """theta = h(u) (unwrapped) and -log g'(v) for u (torch float64)"""
    import torch

    s, w, t = _pflow_map(params)
    v = _twrap(u - t)
    G, Gp = _pflow_g(s, w, v)
    return G + t, -torch.log(Gp)


def pflow_logq(params, th):
    """(log q, S/beta) of configurations th[..., 2, L0, L1] by inverting g by
    bisection in float64 (numpy in, numpy out)."""
    import torch

    p = torch.as_tensor(np.asarray(params, np.float64))
    s, w, t = _pflow_map(p)
    th = np.asarray(th, np.float64)
    P = plaq(th)
    V = P.shape[-1] * P.shape[-2]
    flat = P.reshape(P.shape[:-2] + (V,))
    target = torch.as_tensor(wrap(flat[..., :V - 1] - t.item()))
    lo = torch.full_like(target, -math.pi)
    hi = torch.full_like(target, math.pi)
    for _ in range(80):
        mid = 0.5 * (lo + hi)
        G, _ = _pflow_g(s, w, mid)
        less = G < target
        lo = torch.where(less, mid, lo)
        hi = torch.where(less, hi, mid)
    _, Gp = _pflow_g(s, w, 0.5 * (lo + hi))
    logq = -torch.sum(torch.log(Gp), dim=-1).numpy() - 2 * V * math.log(2 * math.pi)
    return logq, np.sum(1 - np.cos(flat), axis=-1)


def pflow_loss(params, u, beta):
    """per-configuration log q + beta S for draws u[..., V-1] (torch float64),
    with the last plaquette = -sum of the others; returns (loss, log q, S)."""
    import torch

    th, nlg = pflow_h(params, u)
    V = u.shape[-1] + 1
    last = -torch.sum(th, dim=-1)
    S = torch.sum(1 - torch.cos(th), dim=-1) + (1 - torch.cos(last))
    logq = torch.sum(nlg, dim=-1) - 2 * V * math.log(2 * math.pi)
    return logq + beta * S, logq, S


# ---------------------------------------------------------------- plaquette coupling flow
#
# Restatement of mesh_amd/csrc/pcflow.hip: the V - 1 free plaquettes (the last,
# (L0-1, L1-1), enters contexts as 0) start uniform; layer l moves the
# plaquettes of parity class m = l % 4, (x0 % 2, x1 % 2) = (m % 2, m // 2), by
# h(a) = g(wrap(a - t)) + t, the parameters from an MLP of (cos, sin) of the 8
# ring plaquettes, the 8 1x2 loops of neighbouring ring pairs and the 4
# L-shaped loops of ring positions (2j-1, 2j, 2j+1) (each 2x2 square of the
# 3x3 neighbourhood without its centre); the links follow as for the
# plaquette flow (pflow_links).

PC_NI = 40
PC_NP_LAYER = PC_NI * NH + NH + NH * NH + NH + NH * NO + NO
TAG_PCFZ = 0x5043465A
# ring offsets (d0, d1), in order around the plaquette
PC_RING = ((-1, -1), (-1, 0), (-1, 1), (0, 1), (1, 1), (1, 0), (1, -1), (0, -1))


def pcflow_draws(seed, step, chain, L0, L1):
    """(z[V-1], a[V], h0, h1) the GPU draws for (seed, step, chain), float32."""
    V = L0 * L1
    cc = chain << 1

    def ang(a, b):
        u = O.philox_draw(seed, a, b, step << 5, TAG_PCFZ)[0] - 2.0 ** -53
        return np.float32(6.283185307179586477 * u - 3.141592653589793238)

    z = np.array([ang(x, cc) for x in range(V - 1)], np.float32)
    a = np.array([ang(x, cc | 1) for x in range(V)], np.float32)
    return z, a, ang(V, cc), ang(V + 1, cc)


def _pc_layer_params(p, layer):
    o = layer * PC_NP_LAYER
    sizes = [(PC_NI, NH), (NH,), (NH, NH), (NH,), (NH, NO), (NO,)]
    out = []
    for sh in sizes:
        n = int(np.prod(sh))
        out.append(p[o:o + n].reshape(sh))
        o += n
    return out


def _pc_class_mask(L0, L1, m):
    x0 = np.arange(L0).reshape(L0, 1)
    x1 = np.arange(L1).reshape(1, L1)
    mask = ((x0 % 2) == (m % 2)) & ((x1 % 2) == (m // 2))
    mask[L0 - 1, L1 - 1] = False
    return mask


def _pc_mlp_y(W, phi, mask):
    """MLP outputs (B, NA, NO) of the active plaquettes for contexts phi (B, L0, L1)
    whose last plaquette is 0."""
    import torch

    W1, b1, W2, b2, W3, b3 = W
    ring = [torch.roll(phi, shifts=(-d0, -d1), dims=(1, 2))[:, mask] for d0, d1 in PC_RING]
    feats = []
    for k in range(8):
        feats += [torch.cos(ring[k]), torch.sin(ring[k])]
    for k in range(8):
        lp = ring[k] + ring[(k + 1) % 8]
        feats += [torch.cos(lp), torch.sin(lp)]
    for j in range(4):
        lp = ring[(2 * j + 7) % 8] + ring[2 * j] + ring[2 * j + 1]
        feats += [torch.cos(lp), torch.sin(lp)]
    X = torch.stack(feats, dim=-1)
    H1 = torch.tanh(X @ W1 + b1)
    H2 = torch.tanh(H1 @ W2 + b2)
    return H2 @ W3 + b3


def _pc_ncp(Y):
    import torch

    s = torch.exp(torch.clamp(Y[..., :NK], -8.0, 8.0))
    w = torch.softmax(Y[..., NK:2 * NK], dim=-1)
    t = torch.atan2(Y[..., 2 * NK + 1], Y[..., 2 * NK])
    return s, w, t


def pcflow_forward(params, z, L0, L1, inverse=False):
    """torch float64 flow on free plaquettes z (B, V-1).  Forward: (phi, logdet)
    with phi the plaquettes after the last layer and logdet = sum of log g'(u);
    inverse: z given the flow's output (the same logdet, at the inverse's
    points, by bisection to float64 precision)."""
    import torch

    B = z.shape[0]
    V = L0 * L1
    nl = params.numel() // PC_NP_LAYER
    phi = torch.cat([z, torch.zeros(B, 1, dtype=z.dtype)], dim=1).reshape(B, L0, L1)
    logdet = torch.zeros(B, dtype=z.dtype)
    layers = range(nl - 1, -1, -1) if inverse else range(nl)
    for layer in layers:
        m = layer % 4
        mask = torch.as_tensor(_pc_class_mask(L0, L1, m))
        s, w, t = _pc_ncp(_pc_mlp_y(_pc_layer_params(params, layer), phi, mask))
        a = phi[:, mask]
        if inverse:
            target = _twrap(a - t)
            lo = torch.full_like(target, -math.pi)
            hi = torch.full_like(target, math.pi)
            for _ in range(80):
                mid = 0.5 * (lo + hi)
                G, _ = _pflow_g(s, w, mid)
                less = G < target
                lo = torch.where(less, mid, lo)
                hi = torch.where(less, hi, mid)
            u = 0.5 * (lo + hi)
            _, Gp = _pflow_g(s, w, u)
            new = _twrap(u + t)
        else:
            u = _twrap(a - t)
            G, Gp = _pflow_g(s, w, u)
            new = _twrap(G + t)
        logdet = logdet + torch.sum(torch.log(Gp), dim=-1)
        phi = phi.clone()
        phi[:, mask] = new
    return phi.reshape(B, V)[:, :V - 1], logdet


def pcflow_loss(params, z, L0, L1, beta):
    """per-configuration log q + beta S for prior draws z (B, V-1) (torch
    float64), the last plaquette = -sum of the others; returns (loss, log q,
    S, phi)."""
    import torch

    phi, ld = pcflow_forward(params, z, L0, L1)
    V = L0 * L1
    S = torch.sum(1 - torch.cos(phi), dim=-1) + (1 - torch.cos(torch.sum(phi, dim=-1)))
    logq = -ld - 2 * V * math.log(2 * math.pi)
    return logq + beta * S, logq, S, phi


def pcflow_logq(params, th):
    """(log q, S/beta, z) of configurations th (B, 2, L0, L1) (numpy in, numpy out)."""
    import torch

    th = np.asarray(th, np.float64)
    B, _, L0, L1 = th.shape
    V = L0 * L1
    P = wrap(plaq(th)).reshape(B, V)
    S = np.sum(1 - np.cos(P), axis=-1)
    p = torch.as_tensor(np.asarray(params, np.float64))
    z, ld = pcflow_forward(p, torch.as_tensor(P[:, :V - 1]), L0, L1, inverse=True)
    return -ld.numpy() - 2 * V * math.log(2 * math.pi), S, z.numpy()


# ---- Landau gauge on angles (src/gauge/landau.h for U1<T>): links
# th[mu][x...] and the gauge field g = exp(i phi), phi[x...]; float64 from
# the float32 angles, in the operation order of mesh_amd/csrc/u1.hip

def wrap_d(a):
    """to [-pi, pi) as a - 2 pi floor((a + pi) / (2 pi)), then the end fix-ups"""
    tp = 6.283185307179586
    a = np.asarray(a, np.float64)
    a = a - tp * np.floor((a + 3.141592653589793) * (1.0 / tp))
    a = np.where(a >= 3.141592653589793, a - tp, a)
    return np.where(a < -3.141592653589793, a + tp, a)


def landau_angles(th, phi):
    """angle of g(x) U_mu(x) g(x+mu)^*: (phi(x) + theta_mu(x)) - phi(x+mu)"""
    t = np.asarray(th, np.float32).astype(np.float64)
    if phi is None:
        return t
    p = np.asarray(phi, np.float32).astype(np.float64)
    return np.stack([(p + t[mu]) - np.roll(p, -1, axis=mu) for mu in range(t.shape[0])])


def landau_functional(th, phi):
    """sum cos(angle) / (Nd V) (landau.h:23-34 with Nc = 1)"""
    a = landau_angles(th, phi)
    return float(np.cos(a).sum() / a.size)


def landau_gd(th, phi):
    """Im(g d)(x) = sum_mu sin(angle_mu(x)) - sin(angle_mu(x - mu)) (landau.h:41-48)"""
    a = landau_angles(th, phi)
    s = np.zeros(a.shape[1:])
    for mu in range(a.shape[0]):
        sn = np.sin(a[mu])
        s = s + sn
        s = s - np.roll(sn, 1, axis=mu)
    return s


def landau_divsum(th, phi):
    """sum_x Im(g d)^2; condition() = divsum / (Nd V)"""
    return float(np.sum(landau_gd(th, phi) ** 2))


def landau_deriv(th, phi):
    """-0.5 / (Nd V) Im(g d), the algebra element i F (landau.h:52-65), as float32"""
    nd = np.asarray(th).shape[0]
    return ((-0.5 / (nd * np.asarray(th)[0].size)) * landau_gd(th, phi)).astype(np.float32)


def gauge_transform(th, phi):
    return wrap_d(landau_angles(th, phi)).astype(np.float32)


def parity_mask(shape):
    return np.indices(shape).sum(axis=0) % 2


def landau_or_half(th, phi, parity, omega):
    """overrelaxation of one parity on the transformed angles: K(x) = sum_mu
    exp(i theta_mu(x)) + exp(-i theta_mu(x - mu)), turn by -omega arg K.
    Sites of one parity share no link, so the vectorised update is exact.
    Returns (th, phi) as new float32 arrays (phi None stays None)."""
    t = np.asarray(th, np.float32).astype(np.float64)
    nd = t.shape[0]
    kr = np.zeros(t.shape[1:])
    ki = np.zeros(t.shape[1:])
    for mu in range(nd):
        b = np.roll(t[mu], 1, axis=mu)
        kr = kr + np.cos(t[mu])
        ki = ki + np.sin(t[mu])
        kr = kr + np.cos(b)
        ki = ki - np.sin(b)
    m = (parity_mask(t.shape[1:]) == parity) & ~((kr == 0.0) & (ki == 0.0))
    al = np.where(m, -omega * np.arctan2(ki, kr), 0.0)
    out = np.asarray(th, np.float32).copy()
    for mu in range(nd):
        fw = wrap_d(t[mu] + al).astype(np.float32)
        out[mu] = np.where(m, fw, out[mu])
        mb = np.roll(m, -1, axis=mu)
        bw = wrap_d(t[mu] - np.roll(al, -1, axis=mu)).astype(np.float32)
        out[mu] = np.where(mb, bw, out[mu])
    ph = None
    if phi is not None:
        p = np.asarray(phi, np.float32)
        ph = np.where(m, wrap_d(p.astype(np.float64) + al).astype(np.float32), p)
    return out, ph


def gauge_step(phi, F, t):
    """wrap(phi + t F): g = exp(t i F) g (landau.h:85)"""
    return wrap_d(t * np.asarray(F, np.float32).astype(np.float64) +
                  np.asarray(phi, np.float32).astype(np.float64)).astype(np.float32)
