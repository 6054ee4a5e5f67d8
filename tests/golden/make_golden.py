"""Generates tests/golden/su3_golden.npz: fixtures that pin the CPU oracle
independently of its own code.

* exp:   random su(3) elements X (several scales) and scipy.linalg.expm(X)
* lattice: a seeded Haar-ish SU(3) field on a 3x4x5x2 and a 4x4 lattice with
  per-site plaquette sums and staple sums computed here by whole-array
  numpy.roll shifts (the reference's cshift convention,
  src/lattice/lattice.h:366-389: cshift(a, mu, +1)[x] = a[x + mu])
* reunit: perturbed unitary matrices and project_on_group_fast evaluated with
  plain numpy matrix algebra (src/gauge/groups.h:237-243)

Run from the repository root: python tests/golden/make_golden.py
"""
import os

import numpy as np
import scipy.linalg as sl

HERE = os.path.dirname(os.path.abspath(__file__))
rng = np.random.default_rng(20261202)


def rand_su3():
    z = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
    q, r = np.linalg.qr(z)
    q = q * (np.diag(r) / abs(np.diag(r)))
    return q / np.linalg.det(q) ** (1 / 3)


def rand_alg(scale):
    z = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
    a = (z - z.conj().T) / 2
    a -= np.trace(a) / 3 * np.eye(3)
    return a * scale


def field(shape):
    nd = len(shape)
    V = int(np.prod(shape))
    return np.array([[rand_su3() for _ in range(V)] for _ in range(nd)])


def shift(a, shape, mu, s):
    """cshift: out[x] = a[x + s*mu] for a of shape (V, 3, 3)."""
    b = a.reshape(tuple(shape) + (3, 3))
    return np.roll(b, -s, axis=mu).reshape(a.shape)


def mm(a, b):
    return np.einsum("xij,xjk->xik", a, b)


def dag(a):
    return np.conj(np.swapaxes(a, -1, -2))


def plaq_sites(U, shape):
    nd = len(shape)
    s = np.zeros(U.shape[1])
    for mu in range(nd):
        for nu in range(mu + 1, nd):
            P = mm(mm(mm(U[mu], shift(U[nu], shape, mu, 1)), dag(shift(U[mu], shape, nu, 1))), dag(U[nu]))
            s += np.trace(P, axis1=1, axis2=2).real
    return s


def staple(U, shape, mu):
    nd = len(shape)
    S = np.zeros_like(U[0])
    for nu in range(nd):
        if nu == mu:
            continue
        up = mm(mm(shift(U[nu], shape, mu, 1), dag(shift(U[mu], shape, nu, 1))), dag(U[nu]))
        lo = mm(mm(dag(shift(U[nu], shape, mu, 1)), dag(U[mu])), U[nu])
        S += up + shift(lo, shape, nu, -1)
    return S


def reunit(a):
    r = 1.5 * a - 0.5 * a @ a.conj().T @ a
    return r * (4 / 3 - np.linalg.det(r) / 3)


def main():
    out = {}
    scales = np.repeat([0.01, 0.1, 0.5, 1.0, 2.0, 5.0, 10.0], 6)
    X = np.array([rand_alg(s) for s in scales])
    out["exp_X"] = X
    out["exp_E"] = np.array([sl.expm(x) for x in X])
    for name, shape in (("l4", (3, 4, 5, 2)), ("l2", (4, 4))):
        U = field(shape)
        out[f"{name}_shape"] = np.array(shape)
        out[f"{name}_U"] = U
        out[f"{name}_plaq_sites"] = plaq_sites(U, shape)
        out[f"{name}_staple"] = np.array([staple(U, shape, mu) for mu in range(len(shape))])
    A = np.array([rand_su3() + 1e-3 * (rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3)))
                  for _ in range(20)])
    out["reunit_in"] = A
    out["reunit_out"] = np.array([reunit(a) for a in A])
    np.savez_compressed(os.path.join(HERE, "su3_golden.npz"), **out)
    print("wrote", os.path.join(HERE, "su3_golden.npz"))


if __name__ == "__main__":
    main()
