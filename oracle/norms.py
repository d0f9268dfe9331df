"""Loop restatements of the error norms (test infrastructure).

``mms_errors`` follows src/solver.jl:774-810 and ``difference_norms``
src/solver.jl:819-836: quadrature sums over the (q_u+1)^2 standard points of
every element, scaled by h/2.
"""
from __future__ import annotations

import math

import numpy as np


def mms_errors(u, v, T, N, h, pre, geo, kx=1, ky=1, phase=(0.0, 0.0)):
    px, py = phase
    P_u, P_v, S_u, W2d = pre.P_u, pre.P_v, geo.S_u, geo.W2d
    erru = errv = errg = errE = 0.0
    ct = math.cos(pre.omega * T)
    st = math.sin(pre.omega * T)
    tpx, tpy = 2 * math.pi * kx, 2 * math.pi * ky
    for j in range(N):
        for k in range(N):
            Xq = np.broadcast_to(geo.xloc[k][None, :], W2d.shape)
            Yq = np.broadcast_to(geo.yloc[j][:, None], W2d.shape)
            sp = np.cos(tpx * Xq + px) * np.sin(tpy * Yq + py)
            spx = -tpx * np.sin(tpx * Xq + px) * np.sin(tpy * Yq + py)
            spy = tpy * np.cos(tpx * Xq + px) * np.cos(tpy * Yq + py)
            ut, vt_ = (ct - 1.0) * sp, -pre.omega * st * sp
            uxt, uyt = (ct - 1.0) * spx, (ct - 1.0) * spy
            U, V = u[j, k], v[j, k]
            uh = P_u @ U @ P_u.T
            vh = P_v @ V @ P_v.T
            uxh = (2 / h) * (P_u @ U @ S_u.T)
            uyh = (2 / h) * (S_u @ U @ P_u.T)
            c2q = geo.err_c2[j, k]
            erru += np.sum(W2d * (ut - uh) ** 2)
            errv += np.sum(W2d * (vt_ - vh) ** 2)
            errg += np.sum(W2d * ((uxt - uxh) ** 2 + (uyt - uyh) ** 2))
            errE += np.sum(W2d * (0.5 * c2q * ((uxt - uxh) ** 2 + (uyt - uyh) ** 2) + 0.5 * (vt_ - vh) ** 2))
    s = h / 2
    return dict(erru=s * math.sqrt(erru), errv=s * math.sqrt(errv),
                errgradu=s * math.sqrt(errg), errE=s * math.sqrt(errE))


def difference_norms(uA, vA, uB, vB, N, h, pre, geo):
    P_u, P_v, S_u, W2d = pre.P_u, pre.P_v, geo.S_u, geo.W2d
    eu_ = ev_ = eE = 0.0
    for j in range(N):
        for k in range(N):
            eu = uA[j, k] - uB[j, k]
            ev = vA[j, k] - vB[j, k]
            du = P_u @ eu @ P_u.T
            dux = (2 / h) * (P_u @ eu @ S_u.T)
            duy = (2 / h) * (S_u @ eu @ P_u.T)
            dv = P_v @ ev @ P_v.T
            c2q = geo.err_c2[j, k]
            eu_ += np.sum(W2d * du ** 2)
            ev_ += np.sum(W2d * dv ** 2)
            eE += np.sum(W2d * (0.5 * c2q * (dux ** 2 + duy ** 2) + 0.5 * dv ** 2))
    s = h / 2
    return dict(erru=s * math.sqrt(eu_), errv=s * math.sqrt(ev_), errE=s * math.sqrt(eE))
