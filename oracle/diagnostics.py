"""Loop restatements of the source-mode diagnostics (test infrastructure).

Follows src/solver.jl:856-958: the element values of the energy diagnostics
(``_energy_diag!``, src/solver.jl:866-889) and their serial accumulation in the
order row j, column k (src/solver.jl:907-918), the line slice (893-896), the
receiver values (898-905) and the snapshot grid (919-923), with ``locate`` and
``vrow`` of src/solver.jl:880-881. The modal rows come from numpy's Legendre
series, scaled to the L2(-1, 1)-orthonormal basis, independent of the
three-term recurrence the product uses.

Pinned in tests/test_oracle.py by closed forms: states that represent a
linear u and a polynomial v exactly, whose energy, max |v|, min(1 - 2 beta v),
slice, receiver values and grid are known in closed form; and the reference's
series of data/steepening/*/series.csv through the device run, which these
functions check on the same device state.
"""
from __future__ import annotations

import math

import numpy as np
from numpy.polynomial import legendre


def locate(s, N):
    i = min(max(int(math.ceil(s * N)), 1), N)
    return i, min(max(2.0 * (s * N - (i - 1)) - 1.0, -1.0), 1.0)


def vrow(r, qv1):
    out = np.zeros(qv1)
    for d in range(qv1):
        c = np.zeros(d + 1)
        c[d] = 1.0
        out[d] = legendre.legval(r, c) * math.sqrt((2 * d + 1) / 2.0)
    return out


def energy_elements(u, v, N, h, pre, geo):
    """Per element e = j + N k: (h^2/4) sum W (1/2 c^2 |grad u_h|^2 + (1/2 - 2/3 beta v_h) v_h^2),
    max |v_h| and min (1 - 2 beta v_h) at the standard points."""
    E = np.zeros(N * N); vmax = np.zeros(N * N); nd = np.zeros(N * N)
    P_u, P_v, S_u, W2d = pre.P_u, pre.P_v, pre.S_u, geo.W2d
    for j in range(N):
        for k in range(N):
            e = j + N * k
            U, V = u[j, k], v[j, k]
            ux = (2 / h) * (P_u @ U @ S_u.T)
            uy = (2 / h) * (S_u @ U @ P_u.T)
            vh = P_v @ V @ P_v.T
            c2q, bq = geo.err_c2[j, k], geo.err_beta[j, k]
            Z = W2d * (0.5 * c2q * (ux ** 2 + uy ** 2) + (0.5 - (2 / 3) * bq * vh) * vh ** 2)
            E[e] = (h ** 2 / 4) * np.sum(Z)
            vmax[e] = np.max(np.abs(vh))
            nd[e] = np.min(1 - 2 * bq * vh)
    return E, vmax, nd


def diagnostics(u, v, N, h, pre, geo):
    """(energy, max |v_h|, min (1 - 2 beta v_h)) accumulated over j outer, k inner."""
    Ee, ve, ne = energy_elements(u, v, N, h, pre, geo)
    E, vmax, nd = 0.0, 0.0, math.inf
    for j in range(N):
        for k in range(N):
            e = j + N * k
            E += Ee[e]
            vmax = max(vmax, ve[e])
            nd = min(nd, ne[e])
    return E, vmax, nd


def slice_values(v, N, pre, slice_y):
    j0, eta0 = locate(slice_y, N)
    r = vrow(eta0, pre.qv1)
    return np.concatenate([pre.P_v @ (r @ v[j0 - 1, k]) for k in range(N)])


def receiver_values(v, N, pre, receivers):
    out = []
    for (xr, yr) in receivers:
        kr, xir = locate(xr, N)
        jr, etar = locate(yr, N)
        out.append(vrow(etar, pre.qv1) @ v[jr - 1, kr - 1] @ vrow(xir, pre.qv1))
    return np.array(out)


def grid(v, N, pre):
    nq = pre.qu1
    G = np.zeros((N * nq, N * nq))
    for j in range(N):
        for k in range(N):
            G[j * nq:(j + 1) * nq, k * nq:(k + 1) * nq] = pre.P_v @ v[j, k] @ pre.P_v.T
    return G


def run_source(pre, geo, N, dt, nsteps, *, record_every, slice_y, receivers, snap_times, nthreads=0):
    """The loop of _run_source (src/solver.jl:936-951) on the C oracle's RK4 steps."""
    from . import batched as ob
    op = ob.OraclePre(pre, N)
    h = 1.0 / N
    u = np.zeros((N, N, pre.qu1, pre.qu1), order="F")
    v = np.zeros((N, N, pre.qv1, pre.qv1), order="F")
    t = 0.0
    out = dict(t=[], slices=[], energy=[], vmax=[], ndmin=[], recv_t=[], recv=[], snaps=[])
    left = sorted(snap_times)
    for it in range(1, nsteps + 1):
        u, v, _, _ = ob.rk4_steps(op, u, v, t, dt, 1, nthreads=nthreads)
        t += dt
        out["recv_t"].append(t)
        out["recv"].append(receiver_values(v, N, pre, receivers))
        if it % record_every == 0 or it == nsteps:
            E, vm, nd = diagnostics(u, v, N, h, pre, geo)
            out["t"].append(t)
            out["slices"].append(slice_values(v, N, pre, slice_y))
            out["energy"].append(E)
            out["vmax"].append(vm)
            out["ndmin"].append(nd)
        if left and t >= left[0]:
            out["snaps"].append((t, grid(v, N, pre)))
            left.pop(0)
    return out
