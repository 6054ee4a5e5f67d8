"""ctypes front end of the CPU oracle (oracle/liboracle_mesh.so).

TEST INFRASTRUCTURE ONLY.  The oracle restates krox/mesh's lattice arithmetic
on the CPU (see oracle_mesh.cpp for the reference file:line each function
follows).  Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline``
leg of ``bench.py`` import this module, and only as a checker; the product
path (``mesh_amd``) never does.

Arrays use the host layout U[Nd][V][3][3] complex (numpy complex128 or
complex64 of shape (Nd, V, 3, 3)), sites row-major with dimension 0 slowest.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_mesh.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")


def _declare(L):
    L.orc_make_deltas.argtypes = [C.c_char_p, C.c_double, C.c_int, _dp, C.c_int]
    L.orc_make_deltas.restype = C.c_int
    L.orc_philox4x32_10.argtypes = [_up, _up, _up]
    L.orc_det_log.argtypes = [C.c_double]
    L.orc_det_log.restype = C.c_double
    L.orc_det_sincos2pi.argtypes = [C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.orc_link_normals.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, _dp]
    L.orc_metropolis_uniform.argtypes = [C.c_uint64, C.c_uint32]
    L.orc_metropolis_uniform.restype = C.c_double
    L.orc_philox_draw.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, _dp]
    for n in ["orc_su3_mul_nn"]:
        getattr(L, n).argtypes = [_dp, _dp, _dp]
    for n in ["orc_su3_ta", "orc_su3_exp", "orc_su3_reunit", "orc_su3_det",
              "orc_su3_haar_from_normals", "orc_su3_algebra_from_normals"]:
        getattr(L, n).argtypes = [_dp, _dp]
    L.orc_su3_exp_f32.argtypes = [_fp, _fp]
    for sfx, rp in (("f64", _dp), ("f32", _fp)):
        g = lambda n: getattr(L, n + "_" + sfx)
        g("orc_su3_random_gauge").argtypes = [rp, _ip, C.c_int, C.c_uint64]
        g("orc_su3_momenta").argtypes = [rp, _dp, _ip, C.c_int, C.c_uint64, C.c_uint32]
        g("orc_su3_plaq_sites").argtypes = [rp, _ip, C.c_int, _dp]
        g("orc_su3_plaquette").argtypes = [rp, _ip, C.c_int]
        g("orc_su3_plaquette").restype = C.c_double
        g("orc_su3_staple").argtypes = [rp, _ip, C.c_int, C.c_int, rp]
        g("orc_su3_force_update").argtypes = [rp, rp, _ip, C.c_int, C.c_double, C.c_double]
        g("orc_su3_link_update").argtypes = [rp, rp, rp, _ip, C.c_int, C.c_double, C.c_int]
        g("orc_su3_reunit_field").argtypes = [rp, C.c_int64]
        g("orc_su3_kinetic_sites").argtypes = [rp, _ip, C.c_int, _dp]
        g("orc_su3_delta_h").argtypes = [rp, rp, rp, _dp, _ip, C.c_int, C.c_double, _dp]
        g("orc_su3_hmc_trajectory").argtypes = [rp, _ip, C.c_int, C.c_double, _dp, C.c_int,
                                                C.c_uint64, C.c_uint32, C.c_int, _dp]
        g("orc_landau_functional").argtypes = [rp, C.c_void_p, _ip, C.c_int]
        g("orc_landau_functional").restype = C.c_double
        g("orc_landau_divsum").argtypes = [rp, C.c_void_p, _ip, C.c_int]
        g("orc_landau_divsum").restype = C.c_double
        g("orc_landau_or_halfsweep").argtypes = [rp, C.c_void_p, _ip, C.c_int, C.c_int, C.c_double]
    L.orc_su3_force_f64.argtypes = [_dp, _ip, C.c_int, C.c_int, C.c_double, _dp]
    L.orc_landau_deriv_f64.argtypes = [_dp, _dp, _ip, C.c_int, _dp]
    L.orc_su3_exp_mul_field_f64.argtypes = [_dp, _dp, C.c_int64, C.c_double]
    L.orc_axpy_f64.argtypes = [_dp, C.c_double, _dp, C.c_int64]
    L.orc_su3_random_transform_f64.argtypes = [_dp, _ip, C.c_int, C.c_uint64]
    L.orc_su3_gauge_transform_f64.argtypes = [_dp, _dp, _ip, C.c_int]
    L.orc_su3_stage_times_f64.argtypes = [_dp, _ip, C.c_int, C.c_double, C.c_double, _dp]
    L.orc_refpass_trajectory_f64.argtypes = [_dp, _ip, C.c_int, C.c_double, _dp, C.c_int, C.c_uint64,
                                             C.c_uint32, C.c_int, _dp]
    L.orc_refpass_stage_times_f64.argtypes = [_dp, _ip, C.c_int, C.c_double, C.c_double, _dp]
    L.orc_set_threads.argtypes = [C.c_int]
    L.orc_su3_float_energy_f32.argtypes = [_fp, C.c_void_p, _ip, C.c_int, C.c_double, _dp]
    for sfx, rp in (("f64", _dp), ("f32", _fp)):
        g = lambda n: getattr(L, n + "_" + sfx)
        g("orc_su2_random").argtypes = [rp, _ip, C.c_int, C.c_uint64]
        g("orc_su2_momenta").argtypes = [rp, _dp, _ip, C.c_int, C.c_uint64, C.c_uint32]
        g("orc_su2_force_update").argtypes = [rp, rp, _ip, C.c_int, C.c_double, C.c_double]
        g("orc_su2_link_update").argtypes = [rp, rp, rp, _ip, C.c_int, C.c_double, C.c_int]
        g("orc_su2_plaquette").argtypes = [rp, _ip, C.c_int]
        g("orc_su2_plaquette").restype = C.c_double
        g("orc_su2_staple").argtypes = [rp, _ip, C.c_int, C.c_int, rp]
        g("orc_su2_hmc_trajectory").argtypes = [rp, _ip, C.c_int, C.c_double, _dp, C.c_int, C.c_uint64,
                                                C.c_uint32, C.c_int, _dp]
        g("orc_su2_landau_functional").argtypes = [rp, C.c_void_p, _ip, C.c_int]
        g("orc_su2_landau_functional").restype = C.c_double
        g("orc_su2_landau_divsum").argtypes = [rp, C.c_void_p, _ip, C.c_int]
        g("orc_su2_landau_divsum").restype = C.c_double
        g("orc_su2_landau_deriv").argtypes = [rp, C.c_void_p, _ip, C.c_int, rp]
        g("orc_su2_gauge_transform").argtypes = [rp, rp, _ip, C.c_int, rp]
        g("orc_su2_landau_or_half").argtypes = [rp, C.c_void_p, _ip, C.c_int, C.c_int, C.c_double]
    L.orc_su2_qmul.argtypes = [_dp, _dp, _dp]
    L.orc_su2_expm1.argtypes = [_dp, _dp]
    L.orc_max_threads.restype = C.c_int


# ---------------------------------------------------------------- helpers

def _real(a):
    """complex (..., 3, 3) array -> contiguous real view (float64/float32)."""
    a = np.ascontiguousarray(a)
    return a.view(np.float64 if a.dtype == np.complex128 else np.float32)


def _sfx(a):
    return "f64" if a.dtype in (np.complex128, np.float64) else "f32"


def _dims(dims):
    return np.ascontiguousarray(dims, dtype=np.int32)


def volume(dims):
    return int(np.prod(dims))


def make_deltas(scheme, epsilon, substeps):
    out = np.zeros(4096)
    n = lib().orc_make_deltas(scheme.encode(), float(epsilon), int(substeps), out, out.size)
    if n < 0:
        raise ValueError(f"unknown scheme '{scheme}'")
    return out[:n].copy()


def philox(ctr, key):
    out = np.zeros(4, np.uint32)
    lib().orc_philox4x32_10(np.asarray(ctr, np.uint32), np.asarray(key, np.uint32), out)
    return out


def det_log(x):
    return lib().orc_det_log(float(x))


def det_sincos2pi(u):
    s, c = C.c_double(), C.c_double()
    lib().orc_det_sincos2pi(float(u), C.byref(s), C.byref(c))
    return s.value, c.value


def link_normals(seed, tag, ctr2, site, mu, n):
    z = np.zeros(n)
    lib().orc_link_normals(seed, tag, ctr2, site, mu, n, z)
    return z


def metropolis_uniform(seed, traj):
    return lib().orc_metropolis_uniform(seed, traj)


def philox_draw(seed, c0, c1, c2, c3):
    """(u01 in (0,1], normal z0, normal z1) from one Philox4x32-10 block."""
    out = np.zeros(3)
    lib().orc_philox_draw(seed, c0, c1, c2, c3, out)
    return out


def _unary(name, a):
    a = np.ascontiguousarray(a, np.complex128)
    out = np.zeros(a.shape[:-2] + (3, 3), np.complex128) if name != "orc_su3_det" else np.zeros(1, np.complex128)
    getattr(lib(), name)(_real(a), _real(out))
    return out


def su3_mul(a, b):
    out = np.zeros((3, 3), np.complex128)
    lib().orc_su3_mul_nn(_real(np.asarray(a, np.complex128)), _real(np.asarray(b, np.complex128)), _real(out))
    return out


def su3_ta(m):
    return _unary("orc_su3_ta", m)


def su3_exp(x):
    if np.asarray(x).dtype == np.complex64:
        x = np.ascontiguousarray(x)
        out = np.zeros((3, 3), np.complex64)
        lib().orc_su3_exp_f32(_real(x), _real(out))
        return out
    return _unary("orc_su3_exp", x)


def su3_reunit(a):
    return _unary("orc_su3_reunit", a)


def su3_det(a):
    return _unary("orc_su3_det", a)[0]


def haar_from_normals(z):
    out = np.zeros((3, 3), np.complex128)
    lib().orc_su3_haar_from_normals(np.ascontiguousarray(z, np.float64), _real(out))
    return out


def algebra_from_normals(p):
    out = np.zeros((3, 3), np.complex128)
    lib().orc_su3_algebra_from_normals(np.ascontiguousarray(p, np.float64), _real(out))
    return out


def _cdt(prec):
    return np.complex128 if prec in ("f64", 64, "double") else np.complex64


def random_gauge(dims, seed, prec="f64"):
    nd, V = len(dims), volume(dims)
    U = np.zeros((nd, V, 3, 3), _cdt(prec))
    getattr(lib(), "orc_su3_random_gauge_" + _sfx(U))(_real(U), _dims(dims), nd, seed)
    return U


def momenta(dims, seed, traj, prec="f64"):
    nd, V = len(dims), volume(dims)
    P = np.zeros((nd, V, 3, 3), _cdt(prec))
    K = np.zeros(V)
    getattr(lib(), "orc_su3_momenta_" + _sfx(P))(_real(P), K, _dims(dims), nd, seed, traj)
    return P, K


def plaq_sites(U, dims):
    out = np.zeros(volume(dims))
    getattr(lib(), "orc_su3_plaq_sites_" + _sfx(U))(_real(U), _dims(dims), len(dims), out)
    return out


def plaquette(U, dims):
    return getattr(lib(), "orc_su3_plaquette_" + _sfx(U))(_real(U), _dims(dims), len(dims))


def staple(U, dims, mu):
    S = np.zeros((volume(dims), 3, 3), U.dtype)
    getattr(lib(), "orc_su3_staple_" + _sfx(U))(_real(U), _dims(dims), len(dims), mu, _real(S))
    return S


def force(U, dims, mu, beta):
    F = np.zeros((volume(dims), 3, 3), np.complex128)
    lib().orc_su3_force_f64(_real(U), _dims(dims), len(dims), mu, beta, _real(F))
    return F


def force_update(U, P, dims, beta, delta):
    P = P.copy()
    getattr(lib(), "orc_su3_force_update_" + _sfx(U))(_real(U), _real(P), _dims(dims), len(dims), beta, delta)
    return P


def link_update(U, P, dims, t, reunit=False):
    out = np.zeros_like(U)
    getattr(lib(), "orc_su3_link_update_" + _sfx(U))(_real(U), _real(out), _real(P), _dims(dims), len(dims),
                                                     t, int(reunit))
    return out


def reunit_field(U):
    U = np.ascontiguousarray(U).copy()
    getattr(lib(), "orc_su3_reunit_field_" + _sfx(U))(_real(U), U.size // 9)
    return U


def kinetic_sites(P, dims):
    out = np.zeros(volume(dims))
    getattr(lib(), "orc_su3_kinetic_sites_" + _sfx(P))(_real(P), _dims(dims), len(dims), out)
    return out


def delta_h(Uo, Un, P, Ko, dims, beta):
    out = np.zeros(5)
    getattr(lib(), "orc_su3_delta_h_" + _sfx(Uo))(_real(Uo), _real(Un), _real(P), np.ascontiguousarray(Ko),
                                                  _dims(dims), len(dims), beta, out)
    return out


def hmc_trajectory(U, dims, beta, deltas, seed, traj, metropolis=True):
    """In place on U. Returns dict(dH, accepted, plaq, H_old, H_new, u)."""
    out = np.zeros(6)
    d = np.ascontiguousarray(deltas, np.float64)
    getattr(lib(), "orc_su3_hmc_trajectory_" + _sfx(U))(_real(U), _dims(dims), len(dims), beta, d, d.size,
                                                        seed, traj, int(metropolis), out)
    return dict(dH=out[0], accepted=bool(out[1]), plaq=out[2], H_old=out[3], H_new=out[4], u=out[5])


def _hptr(h):
    return None if h is None else _real(h).ctypes.data


def landau_functional(U, dims, h=None):
    return getattr(lib(), "orc_landau_functional_" + _sfx(U))(_real(U), _hptr(h), _dims(dims), len(dims))


def landau_divsum(U, dims, h=None):
    return getattr(lib(), "orc_landau_divsum_" + _sfx(U))(_real(U), _hptr(h), _dims(dims), len(dims))


def landau_deriv(U, h, dims):
    F = np.zeros((volume(dims), 3, 3), np.complex128)
    lib().orc_landau_deriv_f64(_real(U), _real(h), _dims(dims), len(dims), _real(F))
    return F


def landau_or_halfsweep(U, dims, parity, omega, h=None):
    """In place on U (and h if given)."""
    assert U.flags.c_contiguous and (h is None or h.flags.c_contiguous)
    getattr(lib(), "orc_landau_or_halfsweep_" + _sfx(U))(_real(U), _hptr(h), _dims(dims), len(dims),
                                                         parity, omega)


def exp_mul_field(h, F, t):
    """exp(t TA(F)) h per matrix (new array; F, h complex128 of equal shape)."""
    h = np.ascontiguousarray(h, np.complex128).copy()
    F = np.ascontiguousarray(F, np.complex128)
    assert F.shape == h.shape
    lib().orc_su3_exp_mul_field_f64(_real(h), _real(F), h.size // 9, t)
    return h


def axpy(y, a, x):
    """y + (x * a), product and sum rounded separately (new array)."""
    y = np.ascontiguousarray(y, np.complex128).copy()
    x = np.ascontiguousarray(x, np.complex128)
    assert x.shape == y.shape
    lib().orc_axpy_f64(_real(y), a, _real(x), 2 * y.size)
    return y


def sum_real_trace_prod(A, B):
    """sum_x Re tr(A_x B_x) (src/gauge/utils.h:29)."""
    return float(np.einsum("xij,xji->", A.reshape(-1, 3, 3), B.reshape(-1, 3, 3)).real)


def landau_solve_quadratic(x1, x2, x3, y1, y2, y3):
    """Abscissa of the vertex of the parabola through three points
    (src/gauge/landau.h:68-75)."""
    return 0.5 * ((x2 * x2 - x3 * x3) * y1 - (x1 * x1 - x3 * x3) * y2 + (x1 * x1 - x2 * x2) * y3) / \
        ((x2 - x3) * y1 - (x1 - x3) * y2 + (x1 - x2) * y3)


def landau_line_search(U, dims, h, F, eps):
    """Step length along F from the transformation h (src/gauge/landau.h:77-98):
    halve eps until F[exp(eps F) h] exceeds F[h], try 1.2 eps, else the vertex
    of the parabola through eps, 1.2 eps and 0.8 eps."""
    base = landau_functional(U, dims, h)
    trial = lambda e: landau_functional(U, dims, exp_mul_field(h, F, e))
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
    return landau_solve_quadratic(eps, 1.2 * eps, 0.8 * eps, y1, y2, y3)


def landau_ascent(U, dims, max_iter, tol=1e-10, h=None):
    """The reference's Landau solver (src/gauge/landau.h:100-136): steepest
    ascent with the line search, Polak-Ribiere steps after iteration 20 (as
    written there, `last = F` precedes `F += last * beta`, so the step is
    (1 + beta) F), then g = reunitize(exp(eps F) g); stops when condition() <
    tol.  Returns (g, history of (condition, functional, eps) per iteration)."""
    nd, V = len(dims), volume(dims)
    g = np.tile(np.eye(3, dtype=np.complex128), (1, V, 1, 1)) if h is None else \
        np.ascontiguousarray(h, np.complex128).reshape(1, V, 3, 3).copy()
    eps, last_norm, hist = 3000.0, 0.0, []
    last = None
    for it in range(max_iter):
        cond = landau_divsum(U, dims, g) / (nd * V)
        hist.append((cond, landau_functional(U, dims, g), eps))
        if cond < tol:
            break
        F = landau_deriv(U, g, dims).reshape(1, V, 3, 3)
        norm = sum_real_trace_prod(F, F)
        if it > 20:
            beta = (norm - sum_real_trace_prod(F, last)) / last_norm
        last = F.copy()
        if it > 20 and beta > 0:
            F = axpy(F, beta, last)
        eps = landau_line_search(U, dims, g, F, eps)
        g = reunit_field(exp_mul_field(g, F, eps))
        last_norm = norm
    return g, hist


def random_transform(dims, seed):
    h = np.zeros((1, volume(dims), 3, 3), np.complex128)
    lib().orc_su3_random_transform_f64(_real(h), _dims(dims), len(dims), seed)
    return h


def gauge_transform(U, h, dims):
    U = np.ascontiguousarray(U, np.complex128).copy()
    lib().orc_su3_gauge_transform_f64(_real(U), _real(h), _dims(dims), len(dims))
    return U


TAG_HAAR = 0x48414152
TAG_MOM = 0x4D4F4D30
TAG_METRO = 0x4D455452
TAG_GAUGE = 0x47415547


# ---------------------------------------------------------------- CPU baselines

def stage_times(U, dims, beta, delta, refpass=False):
    """Seconds of one pass of each trajectory stage on U (FP64):
    (momenta, momentum step, link step, last link step with Delta H); with
    refpass the reference's pass structure (cshift copies, separate passes,
    oracle_refpass.cpp), whose last stage is reunitarisation + H_new."""
    t = np.zeros(4)
    f = lib().orc_refpass_stage_times_f64 if refpass else lib().orc_su3_stage_times_f64
    f(_real(U), _dims(dims), len(dims), beta, delta, t)
    return t


def refpass_trajectory(U, dims, beta, deltas, seed, traj, metropolis=True):
    """One trajectory in the reference's pass structure, in place on U.
    Returns dict(dH, accepted, H_old, H_new)."""
    out = np.zeros(4)
    d = np.ascontiguousarray(deltas, np.float64)
    lib().orc_refpass_trajectory_f64(_real(U), _dims(dims), len(dims), beta, d, d.size, seed, traj,
                                     int(metropolis), out)
    return dict(dH=out[0], accepted=bool(out[1]), H_old=out[2], H_new=out[3])


def set_threads(n):
    lib().orc_set_threads(int(n))


def max_threads():
    return lib().orc_max_threads()


# ---------------------------------------------------------------- SU(2)
# quaternion fields U[Nd][V][4] (w, x, y, z), float64 or float32

def _q(a):
    return "f64" if a.dtype == np.float64 else "f32"


def su2_random(dims, seed, prec="f64"):
    U = np.zeros((len(dims), volume(dims), 4), np.float64 if prec == "f64" else np.float32)
    getattr(lib(), "orc_su2_random_" + _q(U))(U, _dims(dims), len(dims), seed)
    return U


def su2_momenta(dims, seed, traj, prec="f64"):
    P = np.zeros((len(dims), volume(dims), 4), np.float64 if prec == "f64" else np.float32)
    K = np.zeros(volume(dims))
    getattr(lib(), "orc_su2_momenta_" + _q(P))(P, K, _dims(dims), len(dims), seed, traj)
    return P, K


def su2_force_update(U, P, dims, beta, delta):
    P = np.ascontiguousarray(P).copy()
    getattr(lib(), "orc_su2_force_update_" + _q(U))(np.ascontiguousarray(U), P, _dims(dims), len(dims), beta, delta)
    return P


def su2_link_update(U, P, dims, t, reunit=False):
    out = np.zeros_like(U)
    getattr(lib(), "orc_su2_link_update_" + _q(U))(np.ascontiguousarray(U), out, np.ascontiguousarray(P),
                                                   _dims(dims), len(dims), t, int(reunit))
    return out


def su2_plaquette(U, dims):
    return getattr(lib(), "orc_su2_plaquette_" + _q(U))(np.ascontiguousarray(U), _dims(dims), len(dims))


def su2_staple(U, dims, mu):
    S = np.zeros((1, volume(dims), 4), U.dtype)
    getattr(lib(), "orc_su2_staple_" + _q(U))(np.ascontiguousarray(U), _dims(dims), len(dims), mu, S)
    return S[0]


def su2_hmc_trajectory(U, dims, beta, deltas, seed, traj, metropolis=True):
    """In place on U. Returns dict(dH, accepted, plaq, H_old, H_new, u)."""
    out = np.zeros(6)
    d = np.ascontiguousarray(deltas, np.float64)
    getattr(lib(), "orc_su2_hmc_trajectory_" + _q(U))(U, _dims(dims), len(dims), beta, d, d.size, seed, traj,
                                                      int(metropolis), out)
    return dict(dH=out[0], accepted=bool(out[1]), plaq=out[2], H_old=out[3], H_new=out[4], u=out[5])


def su2_qmul(a, b):
    out = np.zeros(4)
    lib().orc_su2_qmul(np.ascontiguousarray(a, np.float64), np.ascontiguousarray(b, np.float64), out)
    return out


def su2_expm1(v):
    out = np.zeros(4)
    lib().orc_su2_expm1(np.ascontiguousarray(v, np.float64), out)
    return out


# SU(2) Landau gauge: h[1][V][4] (or None for h = 1)

def _opt(h, U):
    """pointer of an optional array of U's dtype (None -> NULL); h must be C-contiguous"""
    if h is None:
        return None
    if h.dtype != U.dtype or not h.flags.c_contiguous:
        raise ValueError("h must be a C-contiguous array of the links' dtype")
    return h.ctypes.data

def su2_landau_functional(U, h, dims):
    return getattr(lib(), "orc_su2_landau_functional_" + _q(U))(np.ascontiguousarray(U), _opt(h, U), _dims(dims),
                                                                len(dims))


def su2_landau_divsum(U, h, dims):
    return getattr(lib(), "orc_su2_landau_divsum_" + _q(U))(np.ascontiguousarray(U), _opt(h, U), _dims(dims),
                                                            len(dims))


def su2_landau_deriv(U, h, dims):
    F = np.zeros((1, volume(dims), 4), U.dtype)
    getattr(lib(), "orc_su2_landau_deriv_" + _q(U))(np.ascontiguousarray(U), _opt(h, U), _dims(dims), len(dims), F)
    return F


def su2_gauge_transform(U, h, dims):
    out = np.zeros_like(U)
    getattr(lib(), "orc_su2_gauge_transform_" + _q(U))(np.ascontiguousarray(U), np.ascontiguousarray(h),
                                                       _dims(dims), len(dims), out)
    return out


def su2_landau_or_half(U, h, dims, parity, omega):
    """in place on U (transformed links) and h (may be None)"""
    getattr(lib(), "orc_su2_landau_or_half_" + _q(U))(U, _opt(h, U), _dims(dims), len(dims), parity, omega)


def su3_float_energy(U, P, dims, beta):
    """(wilsonAction, norm2(P)) of complex64 fields U, P [Nd][V][3][3] (P may
    be None) as the reference's float build sums them"""
    out = np.zeros(2)
    Ur = _real(np.ascontiguousarray(U, np.complex64))
    Pr = None if P is None else _real(np.ascontiguousarray(P, np.complex64))
    lib().orc_su3_float_energy_f32(Ur, None if Pr is None else Pr.ctypes.data, _dims(dims), len(dims), beta, out)
    return out
