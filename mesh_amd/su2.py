"""SU(2) pure-gauge HMC and Landau gauge fixing: the Python mirror of the
reference's Hmc<SU2<T>> (src/gauge/hmc.cpp:75-114, src/gauge/groups.h:110-192)
and Landau<SU2<T>> (src/gauge/landau.h) over the SU(2) entry points of
libmesh_hip.so.

Links are unit quaternions (w, x, y, z); host arrays are [Nd][V][4].  The
kinetic energy is the 2x2 matrix norm of the momenta, 2 (x^2 + y^2 + z^2)
(the norm that the force (beta/4) project_on_algebra(U Sigma) conserves; the
reference adds the quaternion norm, half of it).
"""
import ctypes as C
import math
import time

import numpy as np

from ._lib import F64, lib
from .gauge import LandauSolver
from .lattice import DeviceArray, _prec, default_context

_d = C.c_double


class SU2Field:
    """nf quaternions per site on the device (nf = Nd: the gauge field)."""

    def __init__(self, grid, nf=None, prec=F64, ctx=None):
        self.grid = grid
        self.nf = grid.ndim() if nf is None else nf
        self.prec = _prec(prec)
        self.ctx = ctx or default_context()
        self.dtype = np.float64 if self.prec == F64 else np.float32
        self.bytes = lib().mesh_su2_field_bytes(grid.ref, self.nf, self.prec)
        p = C.c_void_p()
        self.ctx.call("mesh_malloc", C.byref(p), self.bytes)
        self.ptr = p

    def __del__(self):
        try:
            if self.ptr and self.ctx.handle:
                self.ctx._L.mesh_free(self.ctx.handle, self.ptr)
                self.ptr = None
        except Exception:
            pass

    def upload(self, a):
        a = np.ascontiguousarray(a, self.dtype)
        if a.size != self.nf * self.grid.size() * 4:
            raise ValueError("upload: shape mismatch")
        self.ctx.call("mesh_su2_transfer", self.grid.ref, self.nf, self.prec, a.ctypes.data_as(C.c_void_p),
                      self.ptr, 1)
        return self

    def download(self):
        out = np.empty((self.nf, self.grid.size(), 4), self.dtype)
        self.ctx.call("mesh_su2_transfer", self.grid.ref, self.nf, self.prec, out.ctypes.data_as(C.c_void_p),
                      self.ptr, 0)
        return out

    def fill_unit(self):
        self.ctx.call("mesh_su2_unit", self.grid.ref, self.nf, self.prec, self.ptr)
        return self

    def random(self, seed):
        """Haar-random links (random_gauge_field, src/gauge/utils.h:64-70)"""
        self.ctx.call("mesh_su2_random", self.grid.ref, self.nf, self.prec, self.ptr, C.c_uint64(seed))
        return self


def plaquette(U):
    """mean Re tr P / 2 (plaquette, src/gauge/utils.h:93-108)"""
    o = (C.c_double * 1)()
    U.ctx.call("mesh_su2_plaquette", U.grid.ref, U.prec, U.ptr, o)
    return o[0]


def momenta(P, seed, traj, K=None):
    P.ctx.call("mesh_su2_momenta", P.grid.ref, P.prec, P.ptr, K.ptr if K is not None else None, C.c_uint64(seed),
               C.c_uint32(traj))


def force_update(U, P, beta, delta):
    U.ctx.call("mesh_su2_force_update", U.grid.ref, U.prec, U.ptr, P.ptr, _d(beta), _d(delta))


def link_update(Uout, Uin, P, t, reunit=False):
    Uin.ctx.call("mesh_su2_link_update", Uin.grid.ref, Uin.nf, Uin.prec, Uout.ptr, Uin.ptr, P.ptr, _d(t),
                 int(bool(reunit)))


def reunitize(U):
    U.ctx.call("mesh_su2_reunit", U.grid.ref, U.nf, U.prec, U.ptr)


class Hmc:
    """HMC state and observables for SU(2) (src/gauge/hmc.h:61-115)."""

    def __init__(self, grid, prec=F64, seed=1, ctx=None):
        self.g = grid
        self.prec = _prec(prec)
        self.ctx = ctx or default_context()
        self.seed = int(seed)
        self.U = SU2Field(grid, None, self.prec, self.ctx)
        self.U_new = SU2Field(grid, None, self.prec, self.ctx)
        self.P = SU2Field(grid, None, self.prec, self.ctx)
        self.K = DeviceArray(grid.size(), self.ctx)
        self.traj = 0
        self.plaq_history, self.deltaH_history, self.accept_history, self.time_history = [], [], [], []

    def randomizeGaugeField(self):
        self.U.random(self.seed)

    def runHmcUpdate(self, beta, deltas, metropolis=True):
        t0 = time.perf_counter()
        d = np.ascontiguousarray(deltas, np.float64)
        out = (C.c_double * 6)()
        pu, pw = C.c_void_p(self.U.ptr.value), C.c_void_p(self.U_new.ptr.value)
        self.ctx.call("mesh_su2_hmc_trajectory", self.g.ref, self.prec, C.byref(pu), C.byref(pw), self.P.ptr,
                      self.K.ptr, _d(beta), d.ctypes.data_as(C.POINTER(C.c_double)), len(d),
                      C.c_uint64(self.seed), C.c_uint32(self.traj), int(bool(metropolis)), out)
        if pu.value != self.U.ptr.value:
            self.U, self.U_new = self.U_new, self.U
        self.traj += 1
        self.accept_history.append(float(out[1]))
        self.plaq_history.append(out[2])
        self.deltaH_history.append(out[0])
        self.time_history.append(time.perf_counter() - t0)
        return dict(dH=out[0], accepted=bool(out[1]), plaq=out[2], H_old=out[3], H_new=out[4], u=out[5])

    def summary(self):
        p = np.array(self.plaq_history)
        return dict(plaquette=float(p.mean()),
                    plaquette_err=float(p.std(ddof=1) / math.sqrt(len(p))) if len(p) > 1 else 0.0,
                    acceptance=float(np.mean(self.accept_history)),
                    exp_mdH=float(np.mean(np.exp(-np.array(self.deltaH_history)))))


def gauge_transform(U, h):
    """U^h_mu(x) = h(x) U_mu(x) h(x+mu)^+ as a new field"""
    Ug = SU2Field(U.grid, None, U.prec, U.ctx)
    U.ctx.call("mesh_su2_gauge_transform", U.grid.ref, U.prec, Ug.ptr, U.ptr, h.ptr)
    return Ug


class Landau(LandauSolver):
    """Landau gauge fixing of an SU(2) field (Landau<SU2<T>>, src/gauge/landau.h;
    src/tests/lattice_gauge.cpp:46-53 instantiates it): the gauge
    transformation g is one quaternion per site."""

    def __init__(self, U):
        self.U = U
        self.g = SU2Field(U.grid, 1, U.prec, U.ctx)
        self.verbose = False
        self.reset()

    def reset(self):
        self.g.fill_unit()

    def _call(self, name, *args):
        self.U.ctx.call(name, self.U.grid.ref, self.U.prec, *args)

    def __call__(self, h=None):
        """F[h] = sum Re tr (h U h(x+mu)^+) / (Nc Nd V) (landau.h:23-34); default h = g"""
        h = self.g if h is None else h
        o = (C.c_double * 1)()
        self._call("mesh_su2_landau_functional", self.U.ptr, h.ptr, o)
        return o[0]

    def divsum(self, Ug=None):
        if Ug is not None:
            return self._divsum_of(Ug)
        o = (C.c_double * 1)()
        self._call("mesh_su2_landau_divsum", self.U.ptr, self.g.ptr, o)
        return o[0]

    def deriv(self):
        """-0.5 / (Nc Nd V) project_on_algebra(g d) (landau.h:52-65)"""
        F = SU2Field(self.U.grid, 1, self.U.prec, self.U.ctx)
        self._call("mesh_su2_landau_deriv", self.U.ptr, self.g.ptr, F.ptr)
        return F

    def _links(self):
        return SU2Field(self.U.grid, None, self.U.prec, self.U.ctx)

    def _transform_into(self, Ug):
        self._call("mesh_su2_gauge_transform", Ug.ptr, self.U.ptr, self.g.ptr)

    def _or_half(self, Ug, parity, omega):
        self._call("mesh_su2_landau_or_halfsweep", Ug.ptr, self.g.ptr, parity, _d(omega))

    def _divsum_of(self, Ug):
        o = (C.c_double * 1)()
        self._call("mesh_su2_landau_divsum", Ug.ptr, None, o)
        return o[0]

    def _clone(self, h):
        c = SU2Field(h.grid, h.nf, h.prec, h.ctx)
        self._assign(c, h)
        return c

    def _assign(self, dst, src):
        dst.ctx.call("mesh_memcpy", dst.ptr, src.ptr, dst.bytes)

    def _step(self, h, F, e):
        link_update(h, h, F, e)

    def _reunit(self, h):
        reunitize(h)

    def _dot(self, A, B):
        o = (C.c_double * 1)()
        A.ctx.call("mesh_su2_dot", A.grid.ref, A.nf, A.prec, A.ptr, B.ptr, o)
        return o[0]

    def _axpy(self, Y, a, X):
        Y.ctx.call("mesh_su2_axpy", Y.grid.ref, Y.nf, Y.prec, Y.ptr, _d(a), X.ptr)
