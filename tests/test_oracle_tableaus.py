"""Pin the oracle's method coefficients by the mathematics that fixes them.

Tsit5 (Tsitouras 2011) and Rodas5 (Di Marzo 1993 / Hairer-Wanner II) are
not vendored in the reference (they live in OrdinaryDiffEq, Project.toml:13).
A wrong or mistyped coefficient breaks one of these conditions.
"""
import numpy as np
import pytest

from butcher_trees import density, rk_residuals, row_residuals, trees


def test_tree_counts():
    # number of rooted trees of order 1..6 (OEIS A000081)
    assert [len(trees(p)) for p in range(1, 7)] == [1, 1, 2, 4, 9, 20]
    assert density(((), ())) == 3 and density((((),),)) == 6


def tsit5(oracle):
    c, A, bt, R = oracle.tsit5_tableau()
    b = A[6].copy()  # FSAL: u_{n+1} uses row 7, b7 = 0
    bhat = b + bt
    return c, A, b, bhat, bt, R


def test_tsit5_row_sums(oracle):
    c, A, *_ = tsit5(oracle)
    np.testing.assert_allclose(A.sum(axis=1), c, atol=2e-15)


def test_tsit5_order5(oracle):
    c, A, b, *_ = tsit5(oracle)
    res = rk_residuals(A, b, 5)
    worst = max(abs(r) for *_, r in res)
    assert worst < 5e-14, [r for r in res if abs(r[2]) > 5e-14]


def test_tsit5_not_order6(oracle):
    c, A, b, *_ = tsit5(oracle)
    res = rk_residuals(A, b, 6)
    assert max(abs(r) for p, t, r in res if p == 6) > 1e-6


def test_tsit5_embedded_order4_not5(oracle):
    c, A, b, bhat, bt, _ = tsit5(oracle)
    res = rk_residuals(A, bhat, 5)
    assert max(abs(r) for p, t, r in res if p <= 4) < 5e-14
    assert max(abs(r) for p, t, r in res if p == 5) > 1e-5
    assert abs(bt.sum()) < 1e-15


def test_tsit5_fsal_and_bt7(oracle):
    c, A, b, bhat, bt, _ = tsit5(oracle)
    assert b[6] == 0.0 and c[5] == 1.0 and c[6] == 1.0
    assert bt[6] == pytest.approx(1.0 / 66.0, abs=1e-17)


def test_tsit5_interpolant(oracle):
    """b_i(theta): b(1) = b, sum b_i(theta) = theta, and the continuous
    conditions of order 4: sum b_i(th) Phi_i(t) = th^|t| / gamma(t)."""
    c, A, b, bhat, bt, R = tsit5(oracle)

    def bth(th):
        return np.array([sum(R[i, j] * th ** (j + 1) for j in range(4)) for i in range(7)])

    np.testing.assert_allclose(bth(1.0), b, atol=3e-15)
    from butcher_trees import rk_phi, size

    for th in (0.1, 0.37, 0.5, 0.83):
        w = bth(th)
        # coefficients reach |r| ~ 88, so cancellation leaves ~1e-14
        assert abs(w.sum() - th) < 3e-14
        for p in range(1, 5):
            for t in trees(p):
                lhs = w @ rk_phi(t, A)
                assert abs(lhs - th ** size(t) / density(t)) < 3e-13, (th, t)
        # not a 5th-order continuous extension (degree-4 polynomials)
        worst5 = max(abs(w @ rk_phi(t, A) - th ** 5 / density(t)) for t in trees(5))
        assert worst5 > 1e-6


ROSENBROCK = pytest.mark.parametrize("which", [0, 1], ids=["rodas5", "rodas5p"])


def rodas5(oracle, which=0):
    g, a, Cm, c, d = oracle.rodas5_tableau(which)
    Ginv = np.diag(np.full(8, 1.0 / g)) - Cm
    Gam = np.linalg.inv(Ginv)
    alpha = a @ Gam
    m = a[7].copy()
    m[7] = 1.0  # u = U8 + k8
    mhat = a[7].copy()  # embedded: U8
    b = m @ Gam
    bhat = mhat @ Gam
    return g, a, Cm, c, d, Gam, alpha, b, bhat


@ROSENBROCK
def test_rodas5_gamma_diagonal(oracle, which):
    g, a, Cm, c, d, Gam, *_ = rodas5(oracle, which)
    np.testing.assert_allclose(np.diag(Gam), g, rtol=1e-14)
    assert np.allclose(np.triu(Gam, 1), 0.0)


@ROSENBROCK
def test_rodas5_c_and_d_are_row_sums(oracle, which):
    g, a, Cm, c, d, Gam, alpha, b, bhat = rodas5(oracle, which)
    np.testing.assert_allclose(alpha.sum(axis=1), c, atol=2e-13)
    np.testing.assert_allclose(Gam.sum(axis=1), d, atol=2e-13)


@ROSENBROCK
def test_rodas5_order5(oracle, which):
    g, a, Cm, c, d, Gam, alpha, b, bhat = rodas5(oracle, which)
    res = row_residuals(alpha, Gam, b, 5)
    bad = [r for r in res if abs(r[2]) > 1e-11]
    assert not bad, bad


@ROSENBROCK
def test_rodas5_embedded_order4(oracle, which):
    g, a, Cm, c, d, Gam, alpha, b, bhat = rodas5(oracle, which)
    res = row_residuals(alpha, Gam, bhat, 5)
    assert max(abs(r) for p, t, r in res if p <= 4) < 1e-11
    assert max(abs(r) for p, t, r in res if p == 5) > 1e-6


@ROSENBROCK
def test_rodas5_stiffly_accurate_and_Rinf(oracle, which):
    g, a, Cm, c, d, Gam, alpha, b, bhat = rodas5(oracle, which)
    beta = alpha + Gam
    # stiff accuracy: b equals the last row of beta
    np.testing.assert_allclose(b, beta[7], atol=1e-12)
    np.testing.assert_allclose(bhat, beta[6], atol=1e-12)
    # stability function R(z) = 1 + z b (I - z beta)^-1 1 ; R(inf) = 1 - b beta^-1 1
    one = np.ones(8)
    assert abs(1.0 - b @ np.linalg.solve(beta, one)) < 1e-12
    assert abs(1.0 - bhat @ np.linalg.solve(beta, one)) < 1e-12

    def R(z, w):
        return 1.0 + z * (w @ np.linalg.solve(np.eye(8) - z * beta, one))

    # R approximates exp(z) to order 5 near 0
    for z in (1e-2, -1e-2):
        assert abs(R(z, b) - np.exp(z)) < 1e-11
    # A-stability sample: |R(iy)| <= 1 and |R(x)| <= 1 for x < 0
    for y in np.logspace(-2, 4, 40):
        assert abs(R(1j * y, b)) <= 1 + 1e-12
    for x in -np.logspace(-2, 6, 40):
        assert abs(R(x, b)) <= 1 + 1e-12


def test_planted_mistake_is_caught(oracle):
    """A single flipped digit in a coefficient must break an order condition."""
    c, A, b, *_ = tsit5(oracle)
    A2 = A.copy()
    A2[4, 2] += 1e-9
    assert max(abs(r) for *_, r in rk_residuals(A2, A2[6], 5)) > 1e-12
    A3 = A.copy()
    A3[6, 3] *= -1.0  # a sign error in a weight
    assert max(abs(r) for *_, r in rk_residuals(A3, A3[6], 5)) > 1e-3
    g, a, Cm, c8, d8, Gam, alpha, bb, _ = rodas5(oracle)
    Cm2 = Cm.copy()
    Cm2[5, 2] *= 1.0 + 1e-9
    Gam2 = np.linalg.inv(np.diag(np.full(8, 1.0 / g)) - Cm2)
    m = a[7].copy()
    m[7] = 1.0
    res = row_residuals(a @ Gam2, Gam2, m @ Gam2, 5)
    assert max(abs(r) for *_, r in res) > 1e-10


def test_rodas5p_tableau_values(oracle):
    """Rodas5P's gamma and node/derivative vectors (Steinebach 2023) and that
    it is a different method from Rodas5."""
    g, a, Cm, c, d = oracle.rodas5_tableau(1)
    assert g == 0.21193756319429014
    np.testing.assert_allclose(c[:5], [0.0, 0.6358126895828704, 0.4095798393397535,
                                       0.9769306725060716, 0.4288403609558664], rtol=1e-15)
    np.testing.assert_allclose(d[:5], [0.21193756319429014, -0.42387512638858027,
                                       -0.3384627126235924, 1.8046452872882734,
                                       2.325825639765069], rtol=1e-15)
    g5, a5, *_ = oracle.rodas5_tableau(0)
    assert g5 == 0.19 and not np.allclose(a, a5)


def test_rosenbrock_linear_order(oracle):
    """R(z) = 1 + z b (I - z beta)^-1 1 has Taylor coefficients b beta^k 1:
    both methods match exp(z) through z^5; Rodas5P also through z^6 (order 6
    on linear problems), Rodas5 not."""
    import math
    one = np.ones(8)
    for which, k_exact in ((0, 5), (1, 6)):
        g, a, Cm, c, d, Gam, alpha, b, bhat = rodas5(oracle, which)
        beta = alpha + Gam
        dev = [abs(b @ np.linalg.matrix_power(beta, k) @ one - 1 / math.factorial(k + 1))
               for k in range(8)]
        assert max(dev[:k_exact]) < 1e-13, (which, dev)
        assert dev[k_exact] > 1e-6, (which, dev)

