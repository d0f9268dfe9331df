"""Shared problem definitions of the tests (restated from test/runtests.jl)."""
import numpy as np

from edg_westervelt_amd.coefficients import Coefficients

MIXED = dict(bot="absorbing_2beta", top="dirichlet", lft="wall", rgt="absorbing")      # test/runtests.jl:240
MIXED2 = dict(bot="wall", top="absorbing", lft="dirichlet", rgt="absorbing_2beta")      # test/runtests.jl:241


def _z0(x, y):
    return 0.0 * x


def quadrant(xs, ys, a):
    """a[1 + (x >= xs) + 2 (y >= ys)] (test/runtests.jl:50)."""
    a = np.asarray(a, dtype=float)
    return lambda x, y: a[(np.asarray(x) >= xs).astype(int) + 2 * (np.asarray(y) >= ys).astype(int)]


def quadrant_coefficients(xs, ys, c=(1.0, 0.6, 0.8, 0.45), alpha=(0.01, 0.004, 0.02, 0.001),
                          beta=(0.0, 0.03, 0.01, 0.02), smooth=False):
    """Piecewise coefficients on quadrants (test/runtests.jl:51-62)."""
    cq = quadrant(xs, ys, c)
    if smooth:
        pi = np.pi
        return Coefficients(lambda x, y: cq(x, y) * (1 + 0.1 * np.cos(pi * x) * np.cos(pi * y)),
                            lambda x, y: -0.1 * pi * cq(x, y) * np.sin(pi * x) * np.cos(pi * y),
                            lambda x, y: -0.1 * pi * cq(x, y) * np.cos(pi * x) * np.sin(pi * y),
                            quadrant(xs, ys, alpha), _z0, _z0, quadrant(xs, ys, beta))
    return Coefficients(cq, _z0, _z0, quadrant(xs, ys, alpha), _z0, _z0, quadrant(xs, ys, beta))


def rand_state(rng, N, pre, su=1.0, sv=0.05):
    u = np.asfortranarray(su * rng.standard_normal((N, N, pre.qu1, pre.qu1)))
    v = np.asfortranarray(sv * rng.standard_normal((N, N, pre.qv1, pre.qv1)))
    return u, v
