"""Pin the oracle's Rosenbrock dense output (saveat inside a Rodas5 or
Rodas5P step) by the conditions that define it and by measurement.

  u(th) = (1-th) uprev + th (u1 + (1-th)(K1 + th (K2 + th K3))),
  K_m = sum_j H[m][j] k_j   (k_j the stage increments of the W-form)

Upstream's own Rodas interpolation table is not vendored, so these values
are this library's derivation; against upstream they are parity unpinned.
A mistyped coefficient breaks the recomputation, an order condition or the
measured local order."""
import numpy as np
import pytest
from scipy.integrate import solve_ivp

from butcher_trees import density, row_phi, size, trees
from test_oracle_tableaus import ROSENBROCK, rodas5


def dense_system(oracle, which):
    """The 9 x 7 linear system for each m (coefficients of k_1..k_7 in the
    W-form variables): 8 tree conditions for every th, then the stiff limit."""
    g, a, Cm, c, d, Gam, alpha, b, bhat = rodas5(oracle, which)
    ts = [t for p in range(1, 5) for t in trees(p)]
    # k_HW = Gam^-1 u  =>  a condition row r on HW weights is r @ Gam^T on
    # the W-form weights
    Mhw = np.array([row_phi(t, alpha, Gam) for t in ts])
    Mw = Mhw @ Gam.T
    # stage increments of y' = lambda y, y0 = 1, as lambda dt -> -inf
    uinf = np.zeros(8)
    for i in range(8):
        uinf[i] = -(1.0 + a[i, :i] @ uinf[:i])
    A = np.vstack([Mw[:, :7], uinf[None, :7]])
    rhs = []
    for m in range(3):
        r = [(-1.0 / density(t)) if size(t) >= m + 2 else 0.0 for t in ts]
        rhs.append(np.array(r + [[-3.0, 3.0, -1.0][m]]))
    return A, rhs, Mw, uinf, ts


@ROSENBROCK
def test_dense_coefficients_recomputed(oracle, which):
    """H is the unique solution of its defining system (rank 7, consistent)."""
    A, rhs, *_ = dense_system(oracle, which)
    assert np.linalg.matrix_rank(A, tol=1e-8) == 7
    H = oracle.rosenbrock_dense(which)
    assert np.all(H[:, 7] == 0.0)  # no weight on the error stage
    for m in range(3):
        x, *_ = np.linalg.lstsq(A, rhs[m], rcond=None)
        assert np.linalg.norm(A @ x - rhs[m]) < 1e-11
        np.testing.assert_allclose(H[m, :7], x, rtol=1e-9, atol=1e-9)


@ROSENBROCK
def test_dense_order4_for_every_theta(oracle, which):
    """sum_j b_j(th) Phi_j(t) = th^rho(t) / gamma(t) for all trees up to
    order 4 (HW weights b(th) = th b + th (1-th)(h1 + th h2 + th^2 h3))."""
    g, a, Cm, c, d, Gam, alpha, b, bhat = rodas5(oracle, which)
    H = oracle.rosenbrock_dense(which)
    Hhw = H @ Gam  # K_m = H u = H Gam k_HW
    for th in (0.0, 0.13, 0.5, 0.81, 1.0):
        bth = th * b + th * (1 - th) * (Hhw[0] + th * Hhw[1] + th * th * Hhw[2])
        for p in range(1, 5):
            for t in trees(p):
                lhs = bth @ row_phi(t, alpha, Gam)
                assert abs(lhs - th ** p / density(t)) < 1e-11, (th, t)
        # not a 5th-order extension inside the step
        if 0 < th < 1:
            worst = max(abs(bth @ row_phi(t, alpha, Gam) - th ** 5 / density(t))
                        for t in trees(5))
            assert worst > 1e-6


@ROSENBROCK
def test_dense_stiff_limit(oracle, which):
    """y' = lambda y, lambda dt -> -inf: u(th)/uprev -> (1-th)^4; the step
    itself -> 0 (R(inf) = 0, stiffly accurate)."""
    g, a, *_ = rodas5(oracle, which)
    A, rhs, Mw, uinf, ts = dense_system(oracle, which)
    H = oracle.rosenbrock_dense(which)
    y1 = 1.0 + a[7] @ uinf + uinf[7]
    assert abs(y1) < 1e-12
    for th in (0.2, 0.5, 0.9):
        K = H @ uinf
        val = (1 - th) + th * (y1 + (1 - th) * (K[0] + th * (K[1] + th * K[2])))
        assert abs(val - (1 - th) ** 4) < 1e-11


def _lorenz(t, u, p=(10.0, 28.0, 8.0 / 3.0)):
    s, r, bb = p
    return [s * (u[1] - u[0]), u[0] * (r - u[2]) - u[1], u[0] * u[1] - bb * u[2]]


@ROSENBROCK
def test_dense_local_order_and_end_points(oracle, which):
    """One step of the oracle from a Lorenz state: the interpolant's error
    at interior th scales as dt^5 (local order 4); th = 0 and th = 1 give
    uprev and u1 exactly."""
    u = np.array([1.0, 2.0, 20.0])
    p = np.array([10.0, 28.0, 8.0 / 3.0])
    thetas = [0.0, 0.3, 0.77, 1.0]
    hs = [0.02, 0.01, 0.005, 0.0025]
    errs = {0.3: [], 0.77: []}
    for h in hs:
        u1, vals = oracle.rosenbrock_interp(which, "lorenz", u, p, 0.0, h, thetas)
        np.testing.assert_array_equal(vals[0], u)
        np.testing.assert_array_equal(vals[3], u1)
        for k, th in ((1, 0.3), (2, 0.77)):
            ref = solve_ivp(_lorenz, (0.0, th * h), u, method="DOP853", rtol=2.3e-14,
                            atol=1e-14).y[:, -1]
            errs[th].append(np.linalg.norm(vals[k] - ref))
    for th, e in errs.items():
        slopes = np.diff(np.log(e)) / np.diff(np.log(hs))
        assert slopes[-1] > 4.6 and slopes.min() > 4.3, (th, e, slopes)


@pytest.mark.parametrize("alg", ["rodas5", "rodas5p", "auto", "auto5p"])
def test_saveat_with_rosenbrock_matches_radau(oracle, alg):
    """Robertson with saveat through whole solves: every saved value within a
    few tolerance units of Radau at 1e-12; the t0 point is u0 exactly."""
    k = np.array([0.04, 3e7, 1e4])
    N = 3
    u0 = np.zeros((3, N))
    u0[0] = 1.0
    p = np.tile(k[:, None], (1, N)) * np.array([1.0, 0.5, 2.0])[None, :]
    sa = np.concatenate([[0.0], np.logspace(-5, 4, 40)])
    r = oracle.ensemble("robertson", u0, p, 0.0, 1e4, alg=alg, saveat=sa, abstol=1e-10,
                        reltol=1e-8)
    assert (r["retcode"] == 1).all()
    np.testing.assert_array_equal(r["saveat"][0], u0)
    for i in range(N):
        kk = p[:, i]
        ref = solve_ivp(lambda t, y: [-kk[0] * y[0] + kk[2] * y[1] * y[2],
                                      kk[0] * y[0] - kk[1] * y[1] ** 2 - kk[2] * y[1] * y[2],
                                      kk[1] * y[1] ** 2], (0.0, 1e4), u0[:, i], method="Radau",
                        t_eval=sa, rtol=1e-12, atol=1e-16)
        w = np.abs(r["saveat"][:, :, i] - ref.y.T) / (1e-10 + 1e-8 * np.abs(ref.y.T))
        assert np.median(w) < 0.2 and w.max() < 5.0, (alg, i, np.median(w), w.max())
