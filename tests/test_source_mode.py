"""Source mode (src/solver.jl:856-958) on the CPU: the oracle's diagnostics
against closed forms, and the host logic of the device path (locate, vrow, the
record/snapshot schedule, the coordinates of the slice and of the snapshot
grid) against the oracle, the reference's loop and data/steepening/nl/grid.csv.
"""
import math

import numpy as np
import pytest
from numpy.polynomial import legendre

from edg_westervelt_amd import constant_coefficients, precompute
from edg_westervelt_amd.quadrature import JacobiP
from oracle import diagnostics as od
from tests.cases import STEEPENING_RUN, steepening_grid, steepening_recv, steepening_series
from tests.test_oracle import _project


def _poly_case(mixed):
    """u = a x + b y and v = p + r x + s y^2: both lie in the discrete spaces
    (q_v >= 2), so every diagnostic has a closed form."""
    q, N, c, beta = 3, 4, 1.3, 0.7
    pre, geo = precompute(q, N, mixed=mixed, mode="source",
                          coefficients=constant_coefficients(c=c, alpha=0.0, beta=beta))
    a, b, p0, r, s = 0.8, -1.7, 0.1, -0.3, 0.25
    ufun = lambda x, y: a * x + b * y
    vfun = lambda x, y: p0 + r * x + s * y * y
    u = _project(ufun, pre, N)
    v = _project(vfun, pre, N, P=pre.P_v)
    return pre, geo, N, c, beta, (a, b), vfun, u, v


@pytest.mark.parametrize("mixed", [True, False])
def test_oracle_diagnostics_closed_forms(mixed):
    pre, geo, N, c, beta, (a, b), vfun, u, v = _poly_case(mixed)
    h = 1.0 / N
    # energy: the integral over the unit square with an 8-point Gauss rule per
    # direction (exact for the degree-6 integrand)
    z, wq = legendre.leggauss(8)
    x = (z + 1) / 2
    wx = wq / 2
    X, Y = np.meshgrid(x, x, indexing="ij")
    V = vfun(X, Y)
    dens = 0.5 * c * c * (a * a + b * b) + (0.5 - (2 / 3) * beta * V) * V * V
    E_exact = float(np.sum(wx[:, None] * wx[None, :] * dens))
    E, vmax, nd = od.diagnostics(u, v, N, h, pre, geo)
    assert abs(E / E_exact - 1) < 1e-13, (E, E_exact)
    # max |v| and min (1 - 2 beta v) over the standard points of every element
    Xp = geo.xloc[None, :, None, :]
    Yp = geo.yloc[:, None, :, None]
    Vp = vfun(Xp, Yp) + 0 * Xp * Yp
    assert abs(vmax - np.abs(Vp).max()) < 1e-14
    assert abs(nd - (1 - 2 * beta * Vp).min()) < 1e-14
    # slice along y = 0.43 at the x points, receivers, the snapshot grid
    sl = od.slice_values(v, N, pre, 0.43)
    xs = np.concatenate([geo.xloc[k] for k in range(N)])
    assert np.abs(sl - vfun(xs, 0.43)).max() < 1e-14
    rc = [(0.37, 0.81), (0.9, 0.12), (0.0, 1.0)]
    rv = od.receiver_values(v, N, pre, rc)
    assert np.abs(rv - np.array([vfun(x, y) for x, y in rc])).max() < 1e-14
    G = od.grid(v, N, pre)
    ys = np.concatenate([geo.yloc[j] for j in range(N)])
    assert np.abs(G - vfun(xs[None, :], ys[:, None])).max() < 1e-14


def test_locate_and_vrow():
    """locate/vrow of the device path against the oracle's (an independent
    Legendre evaluation) over interior points, element edges and the clamps."""
    from edg_westervelt_amd.solver import locate, vrow
    for N in (1, 4, 64):
        for s in (0.0, 1e-17, 0.15, 0.25, 0.5, 0.45, 0.7, 0.85, 1.0, 1.0 + 1e-9, -0.1):
            assert locate(s, N) == od.locate(s, N), (s, N)
    for r in (-1.0, -0.3, 0.0, 0.4, 1.0):
        for qv1 in (1, 3, 6):
            assert np.abs(vrow(r, qv1) - od.vrow(r, qv1)).max() < 1e-14
    # locate maps into the right element: the midpoint of element i has local coordinate 0
    assert od.locate((3 - 0.5) / 8, 8) == (3, 0.0)


def test_record_schedule():
    """_events against the reference's step loop (src/solver.jl:936-951)."""
    from edg_westervelt_amd.solver import _events
    for dt, nsteps, every, snaps in ((0.01, 37, 5, (0.05, 0.051, 0.2, 5.0)), (1 / 3, 7, 2, ()),
                                     (0.1, 10, 20, (0.0, 0.3, 0.3))):
        rec, snap = [], []
        left = sorted(snaps)
        t = 0.0
        for it in range(1, nsteps + 1):
            t += dt
            if it % every == 0 or it == nsteps:
                rec.append(it)
            if left and t >= left[0]:
                snap.append(it)
                left = left[1:]
        assert _events(dt, nsteps, every, snaps) == (rec, snap)


def test_steepening_fixtures_and_grid():
    """The fixtures hold the reference's time axis for this run (t += dt with
    dt = T / round(T / (h / (10 q^2)))), and the snapshot coordinates
    (grid_x, grid_y = the x / y points element by element) match grid.csv."""
    q, N, T = STEEPENING_RUN["q"], STEEPENING_RUN["N"], STEEPENING_RUN["T"]
    h = 1.0 / N
    nsteps = int(round(T / (h / (10 * q * q))))
    dt = T / nsteps
    ts = []
    t = 0.0
    for _ in range(nsteps):
        t += dt
        ts.append(t)
    for case in ("nl", "lin", "nl_dbc"):
        for row in steepening_recv(case):
            assert row[1] == ts[int(row[0]) - 1]
        ser = steepening_series(case)
        assert len(ser) == nsteps // STEEPENING_RUN["record_every"]
        for i, row in enumerate(ser):
            assert row[0] == ts[(i + 1) * STEEPENING_RUN["record_every"] - 1]
    pre, geo = precompute(q, N, mode="source", coefficients=constant_coefficients(c=1.0, alpha=1e-6, beta=8.0))
    g = np.array(steepening_grid())
    xs = np.concatenate([geo.xloc[k] for k in range(N)])
    ys = np.concatenate([geo.yloc[j] for j in range(N)])
    assert np.abs(g[:, 0] - xs).max() < 1e-16 and np.abs(g[:, 1] - ys).max() < 1e-16
