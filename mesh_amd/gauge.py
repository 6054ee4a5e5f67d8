"""Gauge-theory layer: the Python mirror of krox/mesh's src/gauge call surface
(utils.h, hmc.h/hmc.cpp, landau.h), executed by libmesh_hip.so on the GPU.

Names and argument meaning follow the reference; each function cites the
reference function it stands for.  Errors raise ``MeshError`` (the reference
throws ``std::runtime_error``).
"""
import ctypes as C
import math
import warnings
import time

import numpy as np

from ._lib import F32, F64, MeshError, lib
from .lattice import DeviceArray, Field, GaugeField, Grid, Lattice, _prec, default_context

_d = C.c_double


def _out(n=1):
    return (C.c_double * n)()


def makeDeltas(scheme, epsilon, substeps):
    """Integrator step table (src/gauge/hmc.cpp:7-51)."""
    buf = (C.c_double * 65536)()
    n = lib().mesh_make_deltas(scheme.encode(), float(epsilon), int(substeps), buf, len(buf))
    if n < 0:
        raise MeshError(n, f"unknown scheme '{scheme}'")
    return np.array(buf[:n])


def plaquette(U):
    """Average plaquette normalised to [0,1] (src/gauge/utils.h:93-108)."""
    o = _out()
    U.ctx.call("mesh_su3_plaquette", U.grid.ref, U.prec, U.ptr, o)
    return o[0]


def wilsonAction(U, beta):
    """beta * sum_P (1 - Re tr P / 3) (src/gauge/utils.h:141-146)."""
    o = _out()
    U.ctx.call("mesh_su3_wilson_action", U.grid.ref, U.prec, U.ptr, _d(beta), o)
    return o[0]


def stapleSum(U, mu):
    """Staple sum of direction mu (src/gauge/utils.h:110-139)."""
    S = Lattice(U.grid, U.prec, U.ctx)
    U.ctx.call("mesh_su3_staple", U.grid.ref, U.prec, U.ptr, int(mu), S.ptr)
    return S


def wilsonDeriv(U, mu, beta):
    """(beta/6) TA(U_mu * stapleSum(U, mu)) (src/gauge/utils.h:148-153)."""
    F = Lattice(U.grid, U.prec, U.ctx)
    U.ctx.call("mesh_su3_force", U.grid.ref, U.prec, U.ptr, int(mu), _d(beta), F.ptr)
    return F


def reunitize(U):
    """project_on_group_fast on every link, in place (src/gauge/utils.h:80-91)."""
    U.ctx.call("mesh_su3_reunit", U.grid.ref, U.nf, U.prec, U.ptr)
    return U


def norm2(P):
    """sum ||P||_F^2 over sites and directions (src/gauge/utils.h:28, 35-41)."""
    o = _out()
    P.ctx.call("mesh_su3_norm2", P.grid.ref, P.nf, P.prec, P.ptr, o)
    return o[0]


def sum_real_trace_prod(A, B):
    """sum_real_trace(A * B) (src/gauge/utils.h:29)."""
    o = _out()
    A.ctx.call("mesh_su3_sum_real_trace_prod", A.grid.ref, A.nf, A.prec, A.ptr, B.ptr, o)
    return o[0]


def evolve(U, P, t):
    """U = exp(t TA(P)) U (src/gauge/hmc.h:26-36); U, P are Lattice or GaugeField."""
    if U.nf != P.nf:
        raise ValueError("U and P must have the same number of directions")
    U.ctx.call("mesh_su3_exp_mul", U.grid.ref, U.nf, U.prec, U.ptr, P.ptr, _d(t))
    return U


def cshift(a, dir, offset):
    """New field b[x] = a[x + offset e_dir], periodic (src/lattice/lattice.h:366-389)."""
    out = Field(a.grid, a.nf, a.prec, a.ctx)
    a.ctx.call("mesh_su3_cshift", a.grid.ref, a.nf, a.prec, out.ptr, a.ptr, int(dir), int(offset))
    return out


def lattice_sum(a):
    """Sum of every matrix of the field as a complex 3x3 array, summed in FP64
    (lattice_sum, src/lattice/lattice.h:270-273)."""
    o = _out(18)
    a.ctx.call("mesh_su3_sum", a.grid.ref, a.nf, a.prec, a.ptr, o)
    v = np.array([o[k] for k in range(18)])
    return (v[0::2] + 1j * v[1::2]).reshape(3, 3)


def add_mul(lhs, a, rhs):
    """lhs += a * rhs, the product rounded before the sum (add_mul and
    `lhs += rhs * a`, src/lattice/lattice.h:302-309, 349-363)."""
    if lhs.nf != rhs.nf:
        raise ValueError("fields differ in the number of matrices per site")
    lhs.ctx.call("mesh_su3_axpy", lhs.grid.ref, lhs.nf, lhs.prec, lhs.ptr, _d(a), rhs.ptr)
    return lhs


def momentumStep(U, P, beta, delta):
    """P[mu] -= wilsonDeriv(U, mu, beta) * delta for every mu, fused (hmc.h:54-58)."""
    U.ctx.call("mesh_su3_force_update", U.grid.ref, U.prec, U.ptr, P.ptr, _d(beta), _d(delta))


def linkStep(U_out, U_in, P, t, reunit=False):
    """U_out = exp(t TA(P)) U_in for every mu (hmc.h:49-52), optionally reunitised."""
    U_in.ctx.call("mesh_su3_link_update", U_in.grid.ref, U_in.prec, U_out.ptr, U_in.ptr, P.ptr, _d(t),
                  int(reunit))


def runHmd(U, P, beta, deltas, reunit_last=False):
    """Hamiltonian dynamics over a step table (src/gauge/hmc.h:38-59): even
    index = link step, odd index = momentum step."""
    if len(deltas) % 2 != 1:
        raise ValueError("deltas must have odd length")
    n = len(deltas)
    for i, d in enumerate(deltas):
        if i % 2 == 0:
            linkStep(U, U, P, d, reunit_last and i == n - 1)
        else:
            momentumStep(U, P, beta, d)


def randomGaugeField(U, seed, stream=0):
    """Haar-random links (src/gauge/utils.h:43-52, 64-70)."""
    U.ctx.call("mesh_su3_random_gauge", U.grid.ref, U.nf, U.prec, U.ptr, C.c_uint64(seed), C.c_uint32(stream))
    return U


def randomMomenta(P, seed, traj, K=None):
    """Gaussian su(3) momenta (src/gauge/utils.h:54-62, 72-78)."""
    P.ctx.call("mesh_su3_momenta", P.grid.ref, P.prec, P.ptr, None if K is None else K.ptr, C.c_uint64(seed),
               C.c_uint32(traj))
    return P


def deltaH(U_old, U_new, P, K_old, beta):
    """Per-site differenced Delta H: returns dict(dH, ps_old, ps_new, K_old, K_new)."""
    o = _out(5)
    U_old.ctx.call("mesh_su3_delta_h", U_old.grid.ref, U_old.prec, U_old.ptr, U_new.ptr, P.ptr, K_old.ptr,
                   _d(beta), o)
    return dict(dH=o[0], ps_old=o[1], ps_new=o[2], K_old=o[3], K_new=o[4])


def float_energy(U, P=None, beta=6.0, p_compact=False):
    """(wilsonAction(U, beta), norm2(P)) of FP32 fields summed as the
    reference's float build does (mesh_su3_float_energy): a comparison
    baseline, not the energies the library's trajectories use."""
    o = _out(2)
    U.ctx.call("mesh_su3_float_energy", U.grid.ref, U.ptr, None if P is None else P.ptr, int(bool(p_compact)),
               _d(beta), o)
    return o[0], o[1]


class Hmc:
    """HMC state and observables (src/gauge/hmc.h:61-115, hmc.cpp:53-107).

    ``seed`` keys the Philox streams; trajectory ``n`` draws its momenta from
    counter stream n and its Metropolis uniform from Philox(seed; n)."""

    def __init__(self, grid, prec=F64, seed=1, ctx=None):
        self.g = grid
        self.prec = _prec(prec)
        self.ctx = ctx or default_context()
        self.seed = int(seed)
        self.U = GaugeField(grid, self.prec, self.ctx)
        self.U_new = GaugeField(grid, self.prec, self.ctx)
        # workspace of the trajectory: after runHmcUpdate P holds the final
        # momenta in the compact antihermitian layout, after randomizeMomenta
        # the matrix layout
        self.P = GaugeField(grid, self.prec, self.ctx)
        self.K = DeviceArray(grid.size(), self.ctx)
        self.traj = 0
        self.reset_observables()

    def reset_observables(self):
        self.plaq_history, self.deltaH_history, self.accept_history, self.time_history = [], [], [], []

    def randomizeGaugeField(self):
        randomGaugeField(self.U, self.seed)

    def randomizeMomenta(self):
        randomMomenta(self.P, self.seed, self.traj, self.K)

    def load(self, config_name, resume=True):
        """U from a stored configuration "<ensemble>/configs/<n>" (readConfig,
        src/gauge/utils.h:188-219); with resume, and where the ensemble records
        them, the Philox key and trajectory counter continue the chain that
        stored it, bit for bit."""
        from . import configio

        configio.read_config(config_name, self.U)
        self.therm_left = 0
        if resume:
            info = configio.resume_info(config_name)
            if info["key"] is not None:
                self.seed, self.traj = info["key"], info["traj"]
                # trajectories of the stored chain's thermalisation still to
                # run with metropolis=False for the chain to continue bit for bit
                self.therm_left = info["therm_left"]
            want = 2 if self.prec == F64 else 1
            if info["precision"] is not None and int(info["precision"]) != want:
                warnings.warn(f"{config_name} was written at precision {info['precision']}, this chain runs at "
                              f"{want}: it does not continue bit for bit", stacklevel=2)
        return self

    def runHmcUpdate(self, beta, deltas, metropolis=True):
        """momenta -> trajectory -> accept/reject -> measure (hmc.cpp:75-107)."""
        t0 = time.perf_counter()
        d = np.ascontiguousarray(deltas, np.float64)
        out = _out(6)
        pu, pw = C.c_void_p(self.U.ptr.value), C.c_void_p(self.U_new.ptr.value)
        self.ctx.call("mesh_su3_hmc_trajectory", self.g.ref, self.prec, C.byref(pu), C.byref(pw), self.P.ptr,
                      self.K.ptr, _d(beta), d.ctypes.data_as(C.POINTER(C.c_double)), len(d),
                      C.c_uint64(self.seed), C.c_uint32(self.traj), int(bool(metropolis)), out)
        if pu.value != self.U.ptr.value:
            self.U, self.U_new = self.U_new, self.U
        self.traj += 1
        dt = time.perf_counter() - t0
        self.accept_history.append(float(out[1]))
        self.plaq_history.append(out[2])
        self.deltaH_history.append(out[0])
        self.time_history.append(dt)
        return dict(dH=out[0], accepted=bool(out[1]), plaq=out[2], H_old=out[3], H_new=out[4], u=out[5])

    def summary(self):
        """Same quantities as Hmc::print_summary (hmc.h:88-100)."""
        p = np.array(self.plaq_history)
        err = float(p.std(ddof=1) / math.sqrt(len(p))) if len(p) > 1 else 0.0
        return dict(time_per_step=float(np.mean(self.time_history)),
                    plaquette=float(p.mean()), plaquette_err=err,
                    acceptance=float(np.mean(self.accept_history)),
                    exp_mdH=float(np.mean(np.exp(-np.array(self.deltaH_history)))))

    def print_summary(self):
        s = self.summary()
        print("========== HMC summary ==========")
        print(f"time per step = {s['time_per_step']:.3f} s")
        print(f"plaquette = {s['plaquette']:.4f} +- {s['plaquette_err']:.4f}")
        print(f"acceptance = {s['acceptance']:.2f}")
        print(f"<exp(-dH)> = {s['exp_mdH']:.4f}")


class LandauSolver:
    """The solvers of src/gauge/landau.h:67-136 (and the overrelaxation that
    replaces its ascent) over a group's field operations, which a subclass
    supplies: ``g`` (the gauge transformation), ``__call__(h)`` (the
    functional), ``divsum()``, ``deriv()``, ``_links()`` (a new field shaped
    like U), ``_transform_into(Ug)``, ``_or_half(Ug, parity, omega)``,
    ``_divsum_of(Ug)``, ``_clone(h)``, ``_assign(dst, src)``,
    ``_step(h, F, e)`` (h = exp(e F) h), ``_reunit(h)``, ``_dot(A, B)``
    (sum Re tr A B) and ``_axpy(Y, a, X)``."""

    def condition(self):
        """(1/(Nd V)) sum_x ||project_on_algebra(g d)||^2 (landau.h:37-49)."""
        return self.divsum() / (self.U.grid.ndim() * self.U.grid.size())

    def theta(self):
        """(1/V) sum_x |div A(x)|^2 = Nd * condition()."""
        return self.divsum() / self.U.grid.size()

    def gauge_fixed(self):
        """U^g_mu(x) = g(x) U_mu(x) g(x+mu)^+"""
        Ug = self._links()
        self._transform_into(Ug)
        return Ug

    def run(self, max_iter, omega=1.7, tol=1e-12, check_every=10, reunit_every=100):
        """Overrelaxation sweeps until theta < tol; returns (sweeps, theta).
        The sweeps update U^g and g together; every reunit_every sweeps g is
        reunitarised and U^g rebuilt from it."""
        Ug = self.gauge_fixed()
        theta = math.inf
        it = 0
        while it < max_iter:
            for _ in range(check_every):
                for parity in (0, 1):
                    self._or_half(Ug, parity, omega)
                it += 1
                if reunit_every and it % reunit_every == 0:
                    self._reunit(self.g)
                    self._transform_into(Ug)
            theta = self._divsum_of(Ug) / self.U.grid.size()
            if self.verbose:
                print(f"iter = {it}, theta = {theta}")
            if theta < tol:
                break
        self.Ug = Ug
        return it, theta

    @staticmethod
    def solve_quadratic(x1, x2, x3, y1, y2, y3):
        """abscissa of the vertex of the parabola through three points (landau.h:68-75)"""
        return 0.5 * ((x2 * x2 - x3 * x3) * y1 - (x1 * x1 - x3 * x3) * y2 + (x1 * x1 - x2 * x2) * y3) / \
            ((x2 - x3) * y1 - (x1 - x3) * y2 + (x1 - x2) * y3)

    def line_search(self, F, eps):
        """step length along F (landau.h:77-98): halve eps until the functional
        of exp(eps F) g exceeds that of g, try 1.2 eps, else the vertex of the
        parabola through eps, 1.2 eps and 0.8 eps."""
        base = self()
        h = self._clone(self.g)

        def trial(e):
            self._assign(h, self.g)
            self._step(h, F, e)
            return self(h)

        y1 = trial(eps)
        while y1 <= base:
            eps *= 0.5
            y1 = trial(eps)
            if eps < 1e-10:
                return 0.0
        y2 = trial(eps * 1.2)
        if y2 > y1:
            return 1.2 * eps
        y3 = trial(eps * 0.8)
        return self.solve_quadratic(eps, 1.2 * eps, 0.8 * eps, y1, y2, y3)

    def run_ascent(self, max_iter, tol=1e-10):
        """The reference's solver (landau.h:100-136): ascent along deriv() with
        the line search, Polak-Ribiere steps after iteration 20 (as there,
        `last = F` precedes `F += last * beta`, so the step is (1 + beta) F),
        then g = reunitize(exp(eps F) g), until condition() < tol.  Returns
        (iterations, condition); ``self.history`` holds (condition,
        functional, eps) per iteration."""
        eps, last_norm, last = 3000.0, 0.0, None
        self.history = []
        cond = math.inf
        for it in range(max_iter):
            cond = self.condition()
            self.history.append((cond, self(), eps))
            if self.verbose:
                print(f"iter = {it + 1}, cond = {cond}, term = {self.history[-1][1]}, eps = {eps}")
            if cond < tol:
                return it, cond
            F = self.deriv()
            norm = self._dot(F, F)
            if it > 20:
                beta = (norm - self._dot(F, last)) / last_norm
            last = self._clone(F)
            if it > 20 and beta > 0:
                self._axpy(F, beta, last)
            eps = self.line_search(F, eps)
            self._step(self.g, F, eps)
            self._reunit(self.g)
            last_norm = norm
        return max_iter, cond


class Landau(LandauSolver):
    """Landau gauge fixing of an SU(3) gauge field (src/gauge/landau.h:9-136).

    Holds the gauge transformation ``g`` (identity after ``reset``, as
    landau.h:14) and, for the overrelaxation solver, the transformed links
    ``Ug`` = U^g.  ``run`` replaces the reference's line-search ascent by
    even/odd overrelaxation and stops at theta = (1/V) sum|div A|^2 < tol."""

    def __init__(self, U):
        self.U = U
        self.g = Lattice(U.grid, U.prec, U.ctx)
        self.verbose = False
        self.reset()

    def reset(self):
        self.g.fill_unit()

    def __call__(self, h=None):
        """functional F[h] (landau.h:23-34); default h = g."""
        h = self.g if h is None else h
        o = _out()
        self.U.ctx.call("mesh_landau_functional", self.U.grid.ref, self.U.prec, self.U.ptr, h.ptr, o)
        return o[0]

    def divsum(self, Ug=None):
        if Ug is not None:
            return self._divsum_of(Ug)
        o = _out()
        self.U.ctx.call("mesh_landau_divsum", self.U.grid.ref, self.U.prec, self.U.ptr, self.g.ptr, o)
        return o[0]

    def deriv(self):
        """-0.5/(Nc Nd V) TA(g d) (landau.h:52-65)."""
        F = Lattice(self.U.grid, self.U.prec, self.U.ctx)
        self.U.ctx.call("mesh_landau_deriv", self.U.grid.ref, self.U.prec, self.U.ptr, self.g.ptr, F.ptr)
        return F

    def _links(self):
        return Field(self.U.grid, self.U.grid.ndim(), self.U.prec, self.U.ctx)

    def _transform_into(self, Ug):
        self.U.ctx.call("mesh_su3_gauge_transform", self.U.grid.ref, self.U.prec, Ug.ptr, self.U.ptr, self.g.ptr)

    def _or_half(self, Ug, parity, omega):
        self.U.ctx.call("mesh_landau_or_halfsweep", self.U.grid.ref, self.U.prec, Ug.ptr, self.g.ptr, parity,
                        _d(omega))

    def _divsum_of(self, Ug):
        o = _out()
        Ug.ctx.call("mesh_landau_divsum", Ug.grid.ref, Ug.prec, Ug.ptr, None, o)
        return o[0]

    def _clone(self, h):
        return h.clone()

    def _assign(self, dst, src):
        dst.copy_from(src)

    def _step(self, h, F, e):
        evolve(h, F, e)

    def _reunit(self, h):
        reunitize(h)

    def _dot(self, A, B):
        return sum_real_trace_prod(A, B)

    def _axpy(self, Y, a, X):
        add_mul(Y, a, X)


