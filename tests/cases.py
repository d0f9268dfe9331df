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


# ---- the reference's result tables, copied by tests/golden/make_fixtures.py
import csv as _csv
import os as _os

_GOLDEN = _os.path.join(_os.path.dirname(_os.path.abspath(__file__)), "golden")


def convergence_row(q, pairing, flux, N):
    """erru, errv, errux (= errgradu), errE of data/convergence/results.csv."""
    with open(_os.path.join(_GOLDEN, "convergence_results.csv")) as f:
        for r in _csv.DictReader(f):
            if (int(r["q"]), r["pairing"], r["flux"], int(r["N"])) == (q, pairing, flux, N):
                return dict(erru=float(r["erru"]), errv=float(r["errv"]), errgradu=float(r["errux"]),
                            errE=float(r["errE"]))
    raise KeyError((q, pairing, flux, N))


def alpha_limit_rows():
    """(alpha, erru, errv, errE) of data/alpha_limit/order.csv."""
    with open(_os.path.join(_GOLDEN, "alpha_limit_order.csv")) as f:
        return [(float(r["alpha"]), dict(erru=float(r["erru"]), errv=float(r["errv"]), errE=float(r["errE"])))
                for r in _csv.DictReader(f)]


# The bump of the alpha -> 0 study (experiments/alpha_limit/run.jl:17-33):
# u0 = g(s), s = |x - x0|^2 / R^2, g(s) = exp(1 - 1/(1 - s)) for s < 1, else 0,
# x0 = (1/2, 1/2), R = 1/4, with its first and second derivatives.
BUMP_X0, BUMP_R = (0.5, 0.5), 0.25


def _bump_s(x, y):
    return ((x - BUMP_X0[0]) ** 2 + (y - BUMP_X0[1]) ** 2) / BUMP_R ** 2


def _bump_g(s, order):
    s = np.asarray(s, dtype=float)
    inside = s < 1
    om = np.where(inside, 1 - s, 1.0)
    g = np.where(inside, np.exp(1 - 1 / om), 0.0)
    if order == 0:
        return g
    if order == 1:
        return np.where(inside, -g / om ** 2, 0.0)
    return np.where(inside, g * (1 / om ** 4 - 2 / om ** 3), 0.0)


def _bump_d(x, y, ax):
    d = (x - BUMP_X0[0]) if ax == 0 else (y - BUMP_X0[1])
    return 2 * d / BUMP_R ** 2


BUMP = dict(
    u0=lambda x, y: _bump_g(_bump_s(x, y), 0),
    u0_x=lambda x, y: _bump_g(_bump_s(x, y), 1) * _bump_d(x, y, 0),
    u0_y=lambda x, y: _bump_g(_bump_s(x, y), 1) * _bump_d(x, y, 1),
    u0_xx=lambda x, y: _bump_g(_bump_s(x, y), 2) * _bump_d(x, y, 0) ** 2 + _bump_g(_bump_s(x, y), 1) * 2 / BUMP_R ** 2,
    u0_yy=lambda x, y: _bump_g(_bump_s(x, y), 2) * _bump_d(x, y, 1) ** 2 + _bump_g(_bump_s(x, y), 1) * 2 / BUMP_R ** 2,
)


# The wave-steepening study of experiments/steepening/run.jl:20-34: q_u = 4,
# q_v = 3, N = 64, T = 1.8, upwind, c = 1, alpha = 1e-6, zero initial data, the
# Gaussian source at (0.15, 0.5) with r0 = 0.05, omega = 16 pi, T_s = 1,
# record_every = 32, receivers (0.45, 0.5), (0.7, 0.5), (0.85, 0.5), snapshots
# at t = 0.5, 0.8, 1.1; case -> (beta, boundary).
STEEPENING = {"nl": (8.0, "absorbing_2beta"), "lin": (0.0, "absorbing_2beta"), "nl_dbc": (8.0, "dirichlet")}
STEEPENING_RUN = dict(q=4, N=64, T=1.8, record_every=32, receivers=((0.45, 0.5), (0.7, 0.5), (0.85, 0.5)),
                      snap_times=(0.5, 0.8, 1.1))


def _rows(name):
    with open(_os.path.join(_GOLDEN, name)) as f:
        return [[float(x) for x in line.split(",")] for line in f.read().splitlines() if line]


def steepening_series(case):
    """(t, energy, max|v_h|, min(1 - 2 beta v_h)) rows of data/steepening/<case>/series.csv."""
    return _rows(f"steepening_{case}_series.csv")


def steepening_recv(case):
    """(step, t, v_1, v_2, v_3) rows: every 8th step of data/steepening/<case>/recv.csv."""
    return _rows(f"steepening_{case}_recv.csv")


def steepening_grid():
    """(grid_x, grid_y) rows of data/steepening/nl/grid.csv."""
    return _rows("steepening_nl_grid.csv")


def steepening_snapshot():
    """(rows, cols, values) of the every-8th subsample of data/steepening/nl/snap_0.80.csv."""
    with open(_os.path.join(_GOLDEN, "steepening_nl_snap_0.80.csv")) as f:
        lines = f.read().splitlines()
    cols = [int(c) for c in lines[0].split(",")[1:]]
    r = [[float(x) for x in line.split(",")] for line in lines[1:] if line]
    return [int(x[0]) for x in r], cols, [x[1:] for x in r]
