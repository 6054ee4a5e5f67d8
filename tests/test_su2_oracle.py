"""The SU(2) restatement of the oracle (oracle_su2.cpp) pinned by mathematics:
the quaternion product is the 2x2 matrix product under q -> w - i (x, y, z).sigma,
exp of an algebra element is scipy's expm of that matrix, random links are
Haar (uniform on S^3), the force is the derivative of the Wilson action with
the kinetic normalisation 2 |q|^2 (finite differences), the leapfrog error
scales as delta^2, and HMC chains on a 2-d lattice reproduce the exact
plaquette I_2(beta) / I_1(beta) with <exp(-dH)> = 1."""
import math

import numpy as np
import pytest
import scipy.linalg as sl
import scipy.special as sp

from oracle import oracle as O

SIG = [np.array([[0, 1], [1, 0]], complex), np.array([[0, -1j], [1j, 0]]), np.array([[1, 0], [0, -1]], complex)]


def mat(q):
    return q[0] * np.eye(2) - 1j * sum(q[k + 1] * SIG[k] for k in range(3))


def test_quaternion_product_is_matrix_product():
    rng = np.random.default_rng(1)
    for _ in range(50):
        a, b = rng.normal(size=4), rng.normal(size=4)
        np.testing.assert_allclose(mat(O.su2_qmul(a, b)), mat(a) @ mat(b), atol=1e-14)
        assert np.trace(mat(a)).real == pytest.approx(2 * a[0])


@pytest.mark.parametrize("r", [0.0, 1e-9, 0.05, 0.5, 0.99, 1.5, 7.0])
def test_exp_of_algebra_element(r):
    rng = np.random.default_rng(2)
    for _ in range(10):
        v = rng.normal(size=3)
        v *= r / max(np.linalg.norm(v), 1e-300)
        e = O.su2_expm1(v)
        e[0] += 1.0
        np.testing.assert_allclose(mat(e), sl.expm(mat(np.r_[0.0, v])), atol=2e-15 * max(1.0, r))
        assert abs(np.sum(e * e) - 1) < 4e-15 * max(1.0, r)


def test_haar_links():
    dims = [64, 64]
    U = O.su2_random(dims, 5)
    assert np.abs((U ** 2).sum(-1) - 1).max() < 1e-15
    w = U[..., 0].ravel()
    n = w.size
    # uniform on S^3: E[w] = 0, E[w^2] = 1/4, E[w^4] = 1/8
    assert abs(w.mean()) < 5 / math.sqrt(4 * n)
    assert abs((w ** 2).mean() - 0.25) < 5 * math.sqrt((0.125 - 0.0625) / n)
    assert abs(U[..., 3].mean()) < 5 / math.sqrt(4 * n)


def test_force_is_the_action_derivative():
    """dS/de of S(exp(e X) U_mu(x)) = 4 v . F with X = (0, v) and F the force
    (beta/4) Im(U Sigma) whose momentum step P -= delta F the oracle takes:
    the normalisation that conserves S + 2 |q|^2."""
    dims, beta = [4, 4, 4], 2.3
    U = O.su2_random(dims, 7)
    V, nd = O.volume(dims), len(dims)
    npl = V * nd * (nd - 1) / 2
    S = lambda u: beta * npl * (1 - O.su2_plaquette(u, dims))
    F = -O.su2_force_update(U, np.zeros_like(U), dims, beta, 1.0)  # P = -F
    rng = np.random.default_rng(8)
    for mu, x in ((0, 5), (2, 37)):
        v = rng.normal(size=3)
        eps = 1e-5

        def moved(e):
            u = U.copy()
            d = O.su2_expm1(e * v)
            u[mu, x] = u[mu, x] + O.su2_qmul(d, u[mu, x])
            return u

        num = (S(moved(eps)) - S(moved(-eps))) / (2 * eps)
        ana = 4 * np.dot(v, F[mu, x, 1:])
        assert num == pytest.approx(ana, rel=1e-7, abs=1e-9)


def test_integrator_order_and_reversible_energy():
    dims, beta = [6, 6, 6, 6], 2.0
    U0 = O.su2_random(dims, 9)
    for _ in range(5):
        O.su2_hmc_trajectory(U0, dims, beta, O.make_deltas("2lf", 0.5, 10), 9, _, metropolis=False)
    rms = []
    for n in (8, 16):
        d = O.make_deltas("2lf", 0.02, n)  # tau = 0.2
        dh = [O.su2_hmc_trajectory(U0.copy(), dims, beta, d, 10, t, metropolis=False)["dH"] for t in range(6)]
        rms.append(math.sqrt(np.mean(np.square(dh))))
    slope = math.log(rms[0] / rms[1], 2)
    assert 1.7 < slope < 2.3, (rms, slope)


def test_two_dimensional_plaquette_exact():
    """2-d SU(2) on 8x8 at beta 2: <Re tr P / 2> = I_2(beta)/I_1(beta) (the
    finite-volume correction (I_2/I_1)^64 is negligible), and <exp(-dH)> = 1."""
    dims, beta = [8, 8], 2.0
    U = O.su2_random(dims, 11)
    d = O.make_deltas("2lf", 0.5, 10)
    pl, ex = [], []
    for t in range(800):
        r = O.su2_hmc_trajectory(U, dims, beta, d, 11, t)
        if t >= 100:
            pl.append(O.su2_plaquette(U, dims))
            ex.append(math.exp(-r["dH"]))
    pl, ex = np.array(pl), np.array(ex)
    b = pl.reshape(20, -1).mean(axis=1)
    est, err = b.mean(), b.std(ddof=1) / math.sqrt(20)
    exact = sp.iv(2, beta) / sp.iv(1, beta)
    assert abs(est - exact) < 5 * err + 1e-4, (est, exact, err)
    be = ex.reshape(20, -1).mean(axis=1)
    assert abs(be.mean() - 1) < 5 * be.std(ddof=1) / math.sqrt(20) + 0.01, be.mean()


def test_quaternion_norm_kinetic_term_is_not_conserved():
    """With the momenta and force above, H = S + 2|q|^2 is conserved as the
    step shrinks, while S + |q|^2 (the quaternion's own norm, as the
    reference's norm2(P) adds it) is not: its change stays O(1)."""
    dims, beta = [6, 6], 2.0
    U = O.su2_random(dims, 12)
    npl = O.volume(dims)
    d = O.make_deltas("2lf", 0.005, 20)  # tau = 0.1, small steps
    for t in range(4):
        u = U.copy()
        s_old = beta * npl * (1 - O.su2_plaquette(u, dims))
        r = O.su2_hmc_trajectory(u, dims, beta, d, 13, t, metropolis=False)
        s_new = beta * npl * (1 - O.su2_plaquette(u, dims))
        k_old, k_new = r["H_old"] - s_old, r["H_new"] - s_new
        dh_half = (s_new + k_new / 2) - (s_old + k_old / 2)
        assert abs(r["dH"]) < 1e-3
        assert abs(dh_half) > 20 * abs(r["dH"]), (r["dH"], dh_half)
