"""PDE coefficients c(x,y), alpha(x,y), beta(x,y) of the Westervelt equation.

Mirrors src/coefficients.jl:28-84 of the reference. Coefficients are only
evaluated by ``precompute`` on the host, never by the device kernels. The
callables take ``(x, y)``; they may be written for numpy arrays (fast path) or
for scalars (evaluated point by point).
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Callable

import numpy as np

__all__ = ["Coefficients", "paper_coefficients", "replace_coefficients",
           "constant_coefficients", "element_coefficients", "evaluate"]

PI = math.pi


@dataclass(frozen=True)
class Coefficients:
    """Sound speed ``c``, diffusivity ``alpha`` and nonlinearity ``beta`` with the
    first derivatives of ``c`` and ``alpha`` (src/coefficients.jl:28-36)."""
    c: Callable
    c_x: Callable
    c_y: Callable
    alpha: Callable
    alpha_x: Callable
    alpha_y: Callable
    beta: Callable


def evaluate(f: Callable, X, Y) -> np.ndarray:
    """``f`` at the points (X, Y), broadcast like Julia's ``f.(X, Y)``.

    Tries one vectorized call first; a callable that returns a scalar or fails on
    arrays (Python ``if``/``and`` on the coordinates) is applied point by point.
    """
    X = np.asarray(X, dtype=np.float64)
    Y = np.asarray(Y, dtype=np.float64)
    X, Y = np.broadcast_arrays(X, Y)
    try:
        with np.errstate(all="ignore"):
            r = f(X, Y)
        r = np.asarray(r, dtype=np.float64)
        if r.shape == X.shape:
            return np.array(r, dtype=np.float64, copy=True)
        if r.shape == () and _is_constant(f):
            return np.full(X.shape, float(r))
    except Exception:
        pass
    out = np.empty(X.shape)
    fx, fy, fo = X.ravel(), Y.ravel(), out.reshape(-1)
    for i in range(fx.size):
        fo[i] = float(f(float(fx[i]), float(fy[i])))
    return out


def _is_constant(f) -> bool:
    return getattr(f, "_wedg_const", None) is not None


def _const(val: float) -> Callable:
    def f(x, y):
        return val
    f._wedg_const = val
    return f


def _zero(x, y):
    return 0.0


_zero._wedg_const = 0.0


def paper_coefficients() -> Coefficients:
    """Coefficients of the convergence test (src/coefficients.jl:47-55):
    c = 1 + 0.1 cos(pi x) cos(pi y), alpha = 1e-5 + 1e-6 cos(pi x) cos(pi y),
    beta = 1e-4 (1 + 0.1 sin(pi x) sin(pi y))."""
    return Coefficients(
        lambda x, y: 1 + 0.1 * np.cos(PI * x) * np.cos(PI * y),
        lambda x, y: -0.1 * PI * np.sin(PI * x) * np.cos(PI * y),
        lambda x, y: -0.1 * PI * np.cos(PI * x) * np.sin(PI * y),
        lambda x, y: 1e-5 + 1e-6 * np.cos(PI * x) * np.cos(PI * y),
        lambda x, y: -1e-6 * PI * np.sin(PI * x) * np.cos(PI * y),
        lambda x, y: -1e-6 * PI * np.cos(PI * x) * np.sin(PI * y),
        lambda x, y: 1e-4 * (1 + 0.1 * np.sin(PI * x) * np.sin(PI * y)),
    )


def replace_coefficients(coef: Coefficients, c=None, alpha=None, beta=None) -> Coefficients:
    """Copy of ``coef`` with the coefficients given as numbers replaced by constants
    and their derivatives by zero (src/coefficients.jl:66-76)."""
    return Coefficients(
        coef.c if c is None else _const(float(c)),
        coef.c_x if c is None else _zero,
        coef.c_y if c is None else _zero,
        coef.alpha if alpha is None else _const(float(alpha)),
        coef.alpha_x if alpha is None else _zero,
        coef.alpha_y if alpha is None else _zero,
        coef.beta if beta is None else _const(float(beta)),
    )


def constant_coefficients(*, c, alpha, beta) -> Coefficients:
    """Constant coefficients (src/coefficients.jl:83-84)."""
    return replace_coefficients(paper_coefficients(), c=c, alpha=alpha, beta=beta)


def _element_field(A) -> Callable:
    """f(x, y) = A[j, k] on element (j, k) of the N x N mesh (row j in y, column k
    in x, the reference's element indexing), for numpy arrays or scalars."""
    A = np.asarray(A, dtype=np.float64)
    N = A.shape[0]
    if A.shape != (N, N):
        raise ValueError(f"per-element fields are (N, N) arrays (got shape {A.shape})")

    def f(x, y):
        k = np.clip(np.floor(np.asarray(x) * N).astype(int), 0, N - 1)
        j = np.clip(np.floor(np.asarray(y) * N).astype(int), 0, N - 1)
        return A[j, k]

    return f


def element_coefficients(*, c, alpha, beta) -> Coefficients:
    """Coefficients constant on every element, given as (N, N) arrays of
    per-element values (or numbers, constant everywhere), with zero derivatives
    inside the elements. They jump across element faces, so ``precompute``
    takes one-sided face values from each element (src/solver.jl:371-408,
    742); the point evaluations of the arrays replace closures."""
    def field(a):
        return _const(float(a)) if np.ndim(a) == 0 else _element_field(a)
    return Coefficients(field(c), _zero, _zero, field(alpha), _zero, _zero, field(beta))
