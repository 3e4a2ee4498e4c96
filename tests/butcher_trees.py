"""Rooted trees and order conditions for Runge-Kutta and Rosenbrock methods.

Test infrastructure.  Trees are nested tuples of children (sorted, so each
tree has one canonical form).  The RK condition for tree t is
    sum_i b_i Phi_i(t) = 1 / gamma(t),
with Phi_i(leaf) = 1 and Phi_i(t) = prod_children sum_j a_ij Phi_j(child)
(Hairer-Norsett-Wanner I, Thm II.2.13).

For a Rosenbrock (ROW) method k_i = h f(y0 + sum alpha_ij k_j)
+ h J sum_{j<=i} gamma_ij k_j the same density appears, but a vertex with
exactly one child carries the weight beta~_ij = alpha_ij + gamma_ij summed
over j <= i (the Jacobian term includes the diagonal gamma), while a vertex
with two or more children carries alpha_ij (Hairer-Wanner II, Sec. IV.7,
Table 7.1 is the expanded form of this rule).
"""
from fractions import Fraction
from functools import lru_cache

import numpy as np


@lru_cache(None)
def trees(order):
    """All rooted trees with `order` vertices, canonical nested tuples."""
    if order == 1:
        return ((),)
    out = set()
    # children multiset: partitions of order-1 into subtree orders
    def parts(rem, maxpart):
        if rem == 0:
            yield ()
            return
        for k in range(min(rem, maxpart), 0, -1):
            for rest in parts(rem - k, k):
                yield (k,) + rest

    for part in parts(order - 1, order - 1):
        # choose a tree for each part (with repetition, order-insensitive)
        def choose(idx):
            if idx == len(part):
                yield ()
                return
            for t in trees(part[idx]):
                for rest in choose(idx + 1):
                    yield (t,) + rest

        for ch in choose(0):
            out.add(tuple(sorted(ch)))
    return tuple(sorted(out))


def size(t):
    return 1 + sum(size(c) for c in t)


def density(t):
    d = size(t)
    for c in t:
        d *= density(c)
    return d


def rk_phi(t, A):
    s = A.shape[0]
    v = np.ones(s)
    for c in t:
        v = v * (A @ rk_phi(c, A))
    return v


def row_phi(t, alpha, gam):
    """Rosenbrock elementary weights with the single-child beta~ rule."""
    s = alpha.shape[0]
    if len(t) == 0:
        return np.ones(s)
    if len(t) == 1:
        return (alpha + gam) @ row_phi(t[0], alpha, gam)
    v = np.ones(s)
    for c in t:
        v = v * (alpha @ row_phi(c, alpha, gam))
    return v


def rk_residuals(A, b, max_order):
    res = []
    for p in range(1, max_order + 1):
        for t in trees(p):
            res.append((p, t, float(b @ rk_phi(t, A) - 1.0 / density(t))))
    return res


def row_residuals(alpha, gam, b, max_order):
    res = []
    for p in range(1, max_order + 1):
        for t in trees(p):
            res.append((p, t, float(b @ row_phi(t, alpha, gam) - 1.0 / density(t))))
    return res
