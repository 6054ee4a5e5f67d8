"""U(1) 2-d ensembles on the GPU: batched HMC, the coupling-layer flow
(training by reverse KL, sampling, inverse), independence Metropolis-Hastings
interleaved with HMC, and the autocorrelation analysis that compares them.

Everything numerical runs in libmesh_hip.so (u1.hip, flow.hip); this module
only sequences launches and reads back per-chain scalars.
"""
import ctypes as C
import math
import time

import numpy as np

from ._lib import MeshError, lib
from .gauge import LandauSolver
from .lattice import Grid, default_context

_d = C.c_double


class Buffer:
    """Device buffer of `n` elements of a numpy dtype."""

    def __init__(self, n, dtype, ctx):
        self.ctx, self.n, self.dtype = ctx, int(n), np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        p = C.c_void_p()
        ctx.call("mesh_malloc", C.byref(p), self.nbytes)
        self.ptr = p

    def download(self):
        out = np.empty(self.n, self.dtype)
        self.ctx.call("mesh_memcpy_to_host", out.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes)
        return out

    def upload(self, a):
        a = np.ascontiguousarray(a, self.dtype).reshape(-1)
        if a.size != self.n:
            raise ValueError("size mismatch")
        self.ctx.call("mesh_memcpy_from_host", self.ptr, a.ctypes.data_as(C.c_void_p), self.nbytes)
        return self

    def copy_from(self, other):
        self.ctx.call("mesh_memcpy", self.ptr, other.ptr, self.nbytes)
        return self

    def zero(self):
        self.ctx.call("mesh_memclear", self.ptr, self.nbytes)
        return self

    def __del__(self):
        try:
            if self.ptr and self.ctx.handle:
                self.ctx._L.mesh_free(self.ctx.handle, self.ptr)
                self.ptr = None
        except Exception:
            pass


class U1Ensemble:
    """`batch` independent U(1) configurations of an L0 x L1 lattice, or of a
    lattice of extents `dims` (2 to 4 dimensions) when given."""

    def __init__(self, L0, L1, batch, ctx=None, dims=None):
        self.ctx = ctx or default_context()
        self.dims = tuple(dims) if dims is not None else (L0, L1)
        self.grid = Grid(self.dims)
        self.nd = len(self.dims)
        self.V = int(np.prod(self.dims))
        self.np = self.nd * (self.nd - 1) // 2
        self.L0, self.L1, self.batch = self.dims[0], self.dims[1], batch
        self.theta = Buffer(batch * self.nd * self.V, np.float32, self.ctx)
        self._obs = Buffer(4 * batch, np.float64, self.ctx)
        self._hmc = Buffer(4 * batch, np.float64, self.ctx)

    def random_start(self, seed):
        self.ctx.call("mesh_u1_random", self.grid.ref, self.batch, self.theta.ptr, C.c_uint64(seed))

    def cold_start(self):
        self.theta.zero()

    def angles(self):
        return self.theta.download().reshape(self.batch, self.nd, *self.dims)

    def set_angles(self, a):
        self.theta.upload(np.asarray(a, np.float32))

    def observables(self):
        """dict of per-chain arrays: action/beta, Q (2-d), mean cos theta_P."""
        self.ctx.call("mesh_u1_observables", self.grid.ref, self.batch, self.theta.ptr, self._obs.ptr)
        o = self._obs.download().reshape(self.batch, 4)
        return dict(S=o[:, 0], Q=o[:, 1], cos=o[:, 2] / (self.V * self.np))

    def hmc(self, beta, deltas, seed, traj, metropolis=True):
        """One trajectory of every chain; returns dict(dH, acc, S) arrays."""
        d = np.ascontiguousarray(deltas, np.float64)
        self.ctx.call("mesh_u1_hmc", self.grid.ref, self.batch, self.theta.ptr,
                      d.ctypes.data_as(C.POINTER(C.c_double)), len(d), _d(beta), C.c_uint64(seed),
                      C.c_uint32(traj), int(bool(metropolis)), self._hmc.ptr)
        o = self._hmc.download().reshape(self.batch, 4)
        return dict(dH=o[:, 0], acc=o[:, 1], S=o[:, 3])


class Flow:
    """Gauge-equivariant coupling-layer flow on an L0 x L1 lattice (8 layers
    in the specified design; `layers` up to 64)."""

    _api = "mesh_flow"

    def __init__(self, L0, L1, ctx=None, seed=0, layers=8):
        self.ctx = ctx or default_context()
        self.grid = Grid((L0, L1))
        self.L0, self.L1, self.V = L0, L1, L0 * L1
        self.layers = int(layers)
        self.n = getattr(lib(), f"{self._api}_num_params")(self.layers)
        if self.n <= 0:
            raise MeshError(-1, "the flow has 1 to 64 layers")
        self.params = Buffer(self.n, np.float32, self.ctx)
        self.m = Buffer(self.n, np.float32, self.ctx).zero()
        self.v = Buffer(self.n, np.float32, self.ctx).zero()
        self.grad = Buffer(self.n, np.float32, self.ctx)
        self.t = 0
        self._ws = None
        self._ws_batch = 0
        self.ctx.call(f"{self._api}_init", self.layers, self.params.ptr, C.c_uint64(seed))

    def get_params(self):
        return self.params.download()

    def set_params(self, p):
        self.params.upload(p)

    def _workspace(self, batch):
        if self._ws is None or self._ws_batch != batch:
            nb = getattr(lib(), f"{self._api}_workspace_bytes")(self.grid.ref, batch, self.layers)
            if nb == 0:
                raise MeshError(-2, "unsupported flow geometry")
            self._ws = Buffer(nb, np.uint8, self.ctx)
            self._out = Buffer(2 * batch, np.float64, self.ctx)
            self._ws_batch = batch
        return self._ws

    def gradient(self, batch, beta, seed, step):
        """Reverse-KL gradient over a fresh batch; returns (loss, log q, S/beta)."""
        ws = self._workspace(batch)
        self.ctx.call(f"{self._api}_gradient", self.grid.ref, batch, self.layers, self.params.ptr, C.c_uint64(seed),
                      C.c_uint32(step), _d(beta), ws.ptr, self.grad.ptr, self._out.ptr)
        o = self._out.download().reshape(batch, 2)
        logq, S = o[:, 0], o[:, 1]
        return float(np.mean(logq + beta * S)), logq, S

    def train(self, beta, steps, batch=256, lr=1e-3, seed=1, log_every=0, lr_final=None):
        """Adam on the reverse KL; returns the history of the mean loss
        E[log q + beta S] (= KL(q||p) - log Z)."""
        hist = []
        for i in range(steps):
            loss, _, _ = self.gradient(batch, beta, seed, self.t)
            self.t += 1
            lr_i = lr if lr_final is None else lr * (lr_final / lr) ** (i / max(1, steps - 1))
            self.ctx.call(f"{self._api}_adam", self.layers, self.params.ptr, self.grad.ptr, self.m.ptr, self.v.ptr,
                          self.t, _d(lr_i))
            hist.append(loss)
            if log_every and (i % log_every == 0 or i == steps - 1):
                print(f"flow step {i}: loss {loss:.4f}", flush=True)
        return np.array(hist)

    def sample(self, ens, seed, step, theta_out=None, out=None):
        """Fill `theta_out` (or the ensemble) with proposals; returns
        per-chain (log q, S/beta) as a device buffer of 2*batch doubles."""
        tgt = theta_out if theta_out is not None else ens.theta
        out = out or Buffer(2 * ens.batch, np.float64, self.ctx)
        self.ctx.call("mesh_flow_sample", self.grid.ref, ens.batch, self.layers, self.params.ptr, C.c_uint64(seed),
                      C.c_uint32(step), tgt.ptr, None, out.ptr)
        return out

    def logq(self, ens, out=None, z_out=None):
        out = out or Buffer(2 * ens.batch, np.float64, self.ctx)
        self.ctx.call(f"{self._api}_logq", self.grid.ref, ens.batch, self.layers, self.params.ptr, ens.theta.ptr,
                      None if z_out is None else z_out.ptr, out.ptr)
        return out


class PlaqCouplingFlow(Flow):
    """Coupling layers on the plaquettes of an L0 x L1 lattice (pcflow.hip):
    layer l moves the plaquettes of parity class l % 4 through a circle map
    whose parameters an MLP computes from the surrounding plaquettes, 1x2
    loops and L-shaped loops; the links follow by a maximal-tree gauge with
    uniform holonomies and gauge.  Same interface as Flow (even extents >= 4)."""

    _api = "mesh_pcflow"

    def sample(self, ens, seed, step, theta_out=None, out=None):
        tgt = theta_out if theta_out is not None else ens.theta
        out = out or Buffer(2 * ens.batch, np.float64, self.ctx)
        self.ctx.call("mesh_pcflow_sample", self.grid.ref, ens.batch, self.layers, self.params.ptr, C.c_uint64(seed),
                      C.c_uint32(step), tgt.ptr, out.ptr)
        return out

    def propose(self, ens, state, seed, step, theta_out, out):
        """logq(ens) into `state` and sample(...) into `theta_out` / `out`,
        in one launch (mesh_pcflow_propose)."""
        self.ctx.call("mesh_pcflow_propose", self.grid.ref, ens.batch, self.layers, self.params.ptr, ens.theta.ptr,
                      state.ptr, C.c_uint64(seed), C.c_uint32(step), theta_out.ptr, out.ptr)


class PlaqFlow:
    """Plaquette flow on an L0 x L1 lattice: V - 1 independent plaquettes
    through a learned circle map, links by a maximal-tree gauge, uniform
    holonomies and gauge; exact log q of any configuration from its
    plaquettes (pflow.hip).  Same interface as Flow for run_chains."""

    def __init__(self, L0, L1, ctx=None):
        self.ctx = ctx or default_context()
        self.grid = Grid((L0, L1))
        self.L0, self.L1, self.V = L0, L1, L0 * L1
        self.n = lib().mesh_pflow_num_params()
        self.params = Buffer(self.n, np.float32, self.ctx)
        self.m = Buffer(self.n, np.float32, self.ctx).zero()
        self.v = Buffer(self.n, np.float32, self.ctx).zero()
        self.grad = Buffer(self.n, np.float32, self.ctx)
        self.t = 0
        self._ws = None
        self._ws_batch = 0
        self.ctx.call("mesh_pflow_init", self.params.ptr)

    def get_params(self):
        return self.params.download()

    def set_params(self, p):
        self.params.upload(p)

    def gradient(self, batch, beta, seed, step):
        """Reverse-KL gradient over a fresh batch; returns (loss, log q, S/beta)."""
        if self._ws is None or self._ws_batch != batch:
            self._ws = Buffer(lib().mesh_pflow_workspace_bytes(self.grid.ref, batch), np.uint8, self.ctx)
            self._out = Buffer(2 * batch, np.float64, self.ctx)
            self._ws_batch = batch
        self.ctx.call("mesh_pflow_gradient", self.grid.ref, batch, self.params.ptr, C.c_uint64(seed),
                      C.c_uint32(step), _d(beta), self._ws.ptr, self.grad.ptr, self._out.ptr)
        o = self._out.download().reshape(batch, 2)
        return float(np.mean(o[:, 0] + beta * o[:, 1])), o[:, 0], o[:, 1]

    def train(self, beta, steps, batch=256, lr=1e-2, seed=1, log_every=0, lr_final=None):
        """Adam on the reverse KL; returns the history of E[log q + beta S]."""
        hist = []
        for i in range(steps):
            loss, _, _ = self.gradient(batch, beta, seed, self.t)
            self.t += 1
            lr_i = lr if lr_final is None else lr * (lr_final / lr) ** (i / max(1, steps - 1))
            self.ctx.call("mesh_pflow_adam", self.params.ptr, self.grad.ptr, self.m.ptr, self.v.ptr, self.t,
                          _d(lr_i))
            hist.append(loss)
            if log_every and (i % log_every == 0 or i == steps - 1):
                print(f"pflow step {i}: loss {loss:.4f}", flush=True)
        return np.array(hist)

    def sample(self, ens, seed, step, theta_out=None, out=None):
        """Proposals into `theta_out` (or the ensemble) and their per-chain
        (log q, S/beta) as a device buffer of 2*batch doubles."""
        tgt = theta_out if theta_out is not None else ens.theta
        out = out or Buffer(2 * ens.batch, np.float64, self.ctx)
        self.ctx.call("mesh_pflow_sample", self.grid.ref, ens.batch, self.params.ptr, C.c_uint64(seed),
                      C.c_uint32(step), tgt.ptr)
        self.ctx.call("mesh_pflow_logq", self.grid.ref, ens.batch, self.params.ptr, tgt.ptr, out.ptr)
        return out

    def logq(self, ens, out=None, z_out=None):
        out = out or Buffer(2 * ens.batch, np.float64, self.ctx)
        self.ctx.call("mesh_pflow_logq", self.grid.ref, ens.batch, self.params.ptr, ens.theta.ptr, out.ptr)
        return out

    def propose(self, ens, state, seed, step, theta_out, out):
        """logq(ens) into `state` and proposals into `theta_out` with their
        (log q, S/beta) in `out`, in one launch (mesh_pflow_propose)."""
        self.ctx.call("mesh_pflow_propose", self.grid.ref, ens.batch, self.params.ptr, ens.theta.ptr, state.ptr,
                      C.c_uint64(seed), C.c_uint32(step), theta_out.ptr, out.ptr)


def tau_int(series, S=1.5):
    """Integrated autocorrelation time by the Gamma method with Wolff's
    automatic window (U. Wolff, Comput. Phys. Commun. 156 (2004) 143).
    series: (n,) or (replicas, n).  Returns (tau_int, error); a series that
    never changes (a charge frozen in every chain) carries no information:
    tau_int = inf."""
    x = np.atleast_2d(np.asarray(series, np.float64))
    R, n = x.shape
    xm = x - x.mean()
    var = np.mean(xm * xm)
    if var == 0:
        return math.inf, 0.0
    csum = 0.0
    W = 0
    tau = 0.5
    for W in range(1, n // 2):
        g = np.mean(xm[:, :-W] * xm[:, W:]) / var
        csum += g
        tau = 0.5 + csum
        if tau <= 0.5:
            tau = 0.5
            break
        tauW = S / math.log((2 * tau + 1) / (2 * tau - 1))
        if math.exp(-W / tauW) - tauW / math.sqrt(W * n * R) < 0:
            break
    err = tau * math.sqrt(4 * (W + 0.5 - tau) / (n * R))
    return tau, err


def run_chains(ens, beta, deltas, n_steps, seed, flow=None, n_therm=0, record_every=1):
    """Markov chains over the ensemble.  Without `flow`: plain HMC.  With it:
    each step is one HMC trajectory followed by one independence MH step
    with a flow proposal (1:1), exact accept/reject against log q computed
    by the inverse flow for the state HMC produced.

    Per-step results stay on the device until the end, so the timed loop
    only enqueues launches.  Returns dict(Q (chains x records), plaq,
    acc_hmc, acc_flow, seconds)."""
    ctx = ens.ctx
    B = ens.batch
    d = np.ascontiguousarray(deltas, np.float64)
    dptr = d.ctypes.data_as(C.POINTER(C.c_double))
    n_tot = n_therm + n_steps
    nrec = (n_steps + record_every - 1) // record_every
    obs = Buffer(max(1, nrec) * 4 * B, np.float64, ctx)
    hout = Buffer(n_tot * 4 * B, np.float64, ctx)
    if flow is not None:
        prop = Buffer(B * 2 * ens.V, np.float32, ctx)
        state = Buffer(2 * B, np.float64, ctx)
        pout = Buffer(2 * B, np.float64, ctx)
        accb = Buffer(n_tot * B, np.float64, ctx)
    ctx.synchronize()
    t0 = time.perf_counter()
    for it in range(n_tot):
        if it == n_therm:
            ctx.synchronize()
            t0 = time.perf_counter()
        ctx.call("mesh_u1_hmc", ens.grid.ref, B, ens.theta.ptr, dptr, len(d), _d(beta), C.c_uint64(seed),
                 C.c_uint32(it), 1, hout.ptr.value + it * 4 * B * 8)
        if flow is not None:
            if hasattr(flow, "propose"):
                flow.propose(ens, state, seed + 7919, it, prop, pout)
            else:
                flow.logq(ens, out=state)
                flow.sample(ens, seed + 7919, it, theta_out=prop, out=pout)
            ctx.call("mesh_flow_mh", ens.grid.ref, B, ens.theta.ptr, state.ptr, prop.ptr, pout.ptr,
                     _d(beta), C.c_uint64(seed), C.c_uint32(it), accb.ptr.value + it * B * 8)
        if it >= n_therm and (it - n_therm) % record_every == 0:
            k = (it - n_therm) // record_every
            ctx.call("mesh_u1_observables", ens.grid.ref, B, ens.theta.ptr, obs.ptr.value + k * 4 * B * 8)
    ctx.synchronize()
    dt = time.perf_counter() - t0
    o = obs.download().reshape(-1, B, 4)[:nrec]
    h = hout.download().reshape(n_tot, B, 4)
    return dict(Q=o[:, :, 1].T.copy(), plaq=(o[:, :, 2].T / ens.V).copy(),
                acc_hmc=float(h[n_therm:, :, 1].mean()),
                acc_flow=float(accb.download().reshape(n_tot, B)[n_therm:].mean()) if flow is not None else None,
                seconds=dt)


def ess_per_second(res, obs="Q"):
    """effective samples per second of an observable over all chains."""
    x = res[obs]
    tau, err = tau_int(x)
    n = x.shape[0] * x.shape[1]
    return n / (2 * tau) / res["seconds"], tau, err


class U1Field:
    """One U(1) configuration as FP32 angles theta[f][x] (nf = Nd: links, nf
    = 1: a gauge transformation g = exp(i phi)) on a grid of 2 to 4 dimensions."""

    def __init__(self, grid, nf=None, ctx=None):
        self.grid = grid
        self.ctx = ctx or default_context()
        self.nf = grid.ndim() if nf is None else nf
        self.buf = Buffer(self.nf * grid.size(), np.float32, self.ctx)
        self.ptr = self.buf.ptr

    def upload(self, a):
        self.buf.upload(a)
        return self

    def download(self):
        return self.buf.download().reshape((self.nf,) + tuple(self.grid.shape))

    def fill_unit(self):
        self.buf.zero()
        return self


class Landau(LandauSolver):
    """Landau gauge fixing of a U(1) field (Landau<U1<T>>, src/gauge/landau.h)
    on angles: g = exp(i phi), phi one FP32 angle per site."""

    def __init__(self, U):
        self.U = U
        self.g = U1Field(U.grid, 1, U.ctx)
        self.verbose = False
        self.reset()

    def reset(self):
        self.g.fill_unit()

    def _call(self, name, *args):
        self.U.ctx.call(name, self.U.grid.ref, *args)

    def __call__(self, h=None):
        """F[h] = sum cos(phi(x) + theta_mu(x) - phi(x+mu)) / (Nd V) (landau.h:23-34); default h = g"""
        h = self.g if h is None else h
        o = (C.c_double * 1)()
        self._call("mesh_u1_landau_functional", self.U.ptr, h.ptr, o)
        return o[0]

    def divsum(self, Ug=None):
        if Ug is not None:
            return self._divsum_of(Ug)
        o = (C.c_double * 1)()
        self._call("mesh_u1_landau_divsum", self.U.ptr, self.g.ptr, o)
        return o[0]

    def deriv(self):
        """-0.5 / (Nd V) Im(g d), the algebra element i F (landau.h:52-65)"""
        F = U1Field(self.U.grid, 1, self.U.ctx)
        self._call("mesh_u1_landau_deriv", self.U.ptr, self.g.ptr, F.ptr)
        return F

    def _links(self):
        return U1Field(self.U.grid, None, self.U.ctx)

    def _transform_into(self, Ug):
        self._call("mesh_u1_gauge_transform", Ug.ptr, self.U.ptr, self.g.ptr)

    def _or_half(self, Ug, parity, omega):
        self._call("mesh_u1_landau_or_halfsweep", Ug.ptr, self.g.ptr, parity, _d(omega))

    def _divsum_of(self, Ug):
        o = (C.c_double * 1)()
        self._call("mesh_u1_landau_divsum", Ug.ptr, None, o)
        return o[0]

    def _clone(self, h):
        c = U1Field(h.grid, h.nf, h.ctx)
        c.buf.copy_from(h.buf)
        return c

    def _assign(self, dst, src):
        dst.buf.copy_from(src.buf)

    def _step(self, h, F, e):
        self._call("mesh_u1_gauge_step", h.ptr, h.ptr, F.ptr, _d(e))

    def _reunit(self, h):
        pass  # angles stay on the group

    def _dot(self, A, B):
        o = (C.c_double * 1)()
        A.ctx.call("mesh_u1_dot", C.c_long(A.buf.n), A.ptr, B.ptr, o)
        return o[0]

    def _axpy(self, Y, a, X):
        Y.ctx.call("mesh_u1_axpy", C.c_long(Y.buf.n), Y.ptr, _d(a), X.ptr)
