"""Orthonormal Jacobi/Legendre polynomials and Gauss-Legendre rules on [-1, 1].

Follows src/quadrature.jl:11-66 of the reference (``JacobiP``, ``GradJacobiP``,
``GaussQCofs``). These produce the 1D tables that the device keeps in constant
memory; they run once on the host.
"""
from __future__ import annotations

import math

import numpy as np

__all__ = ["JacobiP", "GradJacobiP", "GaussQCofs"]


def JacobiP(x, alpha: float, beta: float, N: int):
    """Orthonormal Jacobi polynomial P_N^(alpha,beta) at the points ``x``.

    Three-term forward recurrence of src/quadrature.jl:11-35; alpha = beta = 0
    gives the L2(-1,1)-orthonormal Legendre polynomials. A scalar ``x`` gives a
    scalar (src/quadrature.jl:37).
    """
    scalar = np.ndim(x) == 0
    xp = np.atleast_1d(np.asarray(x, dtype=np.float64))
    PL = np.zeros((N + 1, xp.size))
    gamma0 = (2.0 ** (alpha + beta + 1) / (alpha + beta + 1)
              * math.gamma(alpha + 1) * math.gamma(beta + 1) / math.gamma(alpha + beta + 1))
    PL[0, :] = 1.0 / math.sqrt(gamma0)
    if N == 0:
        out = PL[0, :]
    else:
        gamma1 = (alpha + 1) * (beta + 1) / (alpha + beta + 3) * gamma0
        PL[1, :] = ((alpha + beta + 2) * xp / 2 + (alpha - beta) / 2) / math.sqrt(gamma1)
        if N > 1:
            aold = 2 / (2 + alpha + beta) * math.sqrt((alpha + 1) * (beta + 1) / (alpha + beta + 3))
            for i in range(1, N):
                h1 = 2 * i + alpha + beta
                anew = 2 / (h1 + 2) * math.sqrt((i + 1) * (i + 1 + alpha + beta)
                                                * (i + 1 + alpha) * (i + 1 + beta) / (h1 + 1) / (h1 + 3))
                bnew = -(alpha ** 2 - beta ** 2) / h1 / (h1 + 2)
                PL[i + 1, :] = (1 / anew) * (-aold * PL[i - 1, :] + (xp - bnew) * PL[i, :])
                aold = anew
        out = PL[N, :]
    return float(out[0]) if scalar else out.copy()


def GradJacobiP(r, alpha: float, beta: float, N: int):
    """First derivative of the orthonormal P_N^(alpha,beta) (src/quadrature.jl:44-49)."""
    scalar = np.ndim(r) == 0
    rr = np.atleast_1d(np.asarray(r, dtype=np.float64))
    if N == 0:
        out = np.zeros(rr.size)
    else:
        out = math.sqrt(N * (N + alpha + beta + 1)) * JacobiP(rr, alpha + 1, beta + 1, N - 1)
    return float(out[0]) if scalar else out


def GaussQCofs(q: int):
    """Nodes and weights of the q-point Gauss-Legendre rule (Golub-Welsch).

    Eigen-decomposition of the symmetric Jacobi matrix, as src/quadrature.jl:57-66;
    nodes sorted ascending, weights 2*v[0]^2.
    """
    if q < 2:
        raise ValueError("GaussQCofs needs q >= 2")
    kk = np.arange(1, q, dtype=np.float64) ** 2
    bv = np.sqrt(kk / (4 * kk - 1))
    J = np.diag(bv, 1) + np.diag(bv, -1)
    vals, vecs = np.linalg.eigh(J)
    w = 2 * vecs[0, :] ** 2
    idx = np.argsort(vals, kind="stable")
    return vals[idx].copy(), w[idx].copy()
