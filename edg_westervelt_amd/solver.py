"""The reference's hot-path API on the device: ``Workspace``, ``edg_rhs_``
(``edg_rhs!``), ``edg_rhs``, ``rk4_step_`` (``rk4_step!``), ``rk4_step``,
``solve`` and ``difference_norms``.

Mirrors src/rhs.jl:166-220, 475-571 and src/solver.jl:88-339, 772-836 of the
reference (names, argument meaning and error behaviour). Julia's mutating
``f!`` is spelled ``f_`` here. Every right-hand side and RK4 step runs in the
HIP kernels of libwedg.so; the state and work arrays stay on the device during
``solve``.
"""
from __future__ import annotations

import ctypes as C
import math
from types import SimpleNamespace

import numpy as np

from . import _lib
from ._lib import DimensionMismatch, WedgError, lib
from .precompute import GaussianSource, precompute

__all__ = ["Workspace", "edg_rhs_", "edg_rhs", "rk4_step_", "rk4_step", "solve", "difference_norms",
           "DimensionMismatch", "WedgError"]

_dp = C.POINTER(C.c_double)


def _ptr(a):
    return a.ctypes.data_as(_dp)


class Workspace:
    """Device-resident arrays for ``edg_rhs_`` and ``rk4_step_`` on an N x N mesh
    (src/rhs.jl:184-220).

    ``nblocks`` is accepted for the reference's signature; the device grid
    replaces the element blocks, and results do not depend on it. ``cheb``
    selects the Chebyshev step count: ``"global"`` (the reference: every element
    takes the largest count) or ``"element"`` (each element its own a-priori
    count to the same 1e-14 residual bound).
    """

    def __init__(self, pre, N: int, nblocks=None, cheb: str = "global"):
        if nblocks is not None and nblocks < 1:
            raise ValueError(f"nblocks must be ≥ 1 (got {nblocks})")
        if pre.Ainv_u.shape[1] != N * N:
            raise DimensionMismatch(f"state, workspace and pre do not match (N = {N}, q_u + 1 = {pre.qu1}, "
                                    f"q_v + 1 = {pre.qv1})")
        self.pre = pre
        self.N = N
        self.nb = 1
        self.qu1, self.qv1, self.q_high = pre.qu1, pre.qv1, pre.q_high
        self.problem = _lib.Problem(pre, N)
        self.dev = _lib.DeviceWorkspace(self.problem, cheb)
        self.ut = np.zeros((N, N, pre.qu1, pre.qu1), order="F")
        self.vt = np.zeros((N, N, pre.qv1, pre.qv1), order="F")
        self.fin = [True]
        self.cheb_steps = 0

    def _check_sizes(self, u, v, N, pre):
        nu, nv = pre.qu1, pre.qv1
        if (u.shape != (N, N, nu, nu) or v.shape != (N, N, nv, nv) or N != self.N
                or self.qu1 != nu or self.qv1 != nv or self.q_high != pre.q_high
                or pre.Ainv_u.shape[1] != N * N):
            raise DimensionMismatch(f"state, workspace and pre do not match (N = {N}, q_u + 1 = {nu}, "
                                    f"q_v + 1 = {nv})")


def _state(a):
    a = np.asarray(a, dtype=np.float64)
    if not a.flags.f_contiguous:
        a = np.asfortranarray(a)
    return a


def edg_rhs_(ws: Workspace, u, v, t: float, N: int, h: float, pre):
    """(u_t, v_t) at time t into ``ws.ut``, ``ws.vt``, overwritten by the next
    call (src/rhs.jl:490-494)."""
    u = _state(u); v = _state(v)
    ws._check_sizes(u, v, N, pre)
    st = lib.wedg_rhs(ws.dev.h, _ptr(u), _ptr(v), float(t), int(N), float(h), _ptr(ws.ut), _ptr(ws.vt))
    if st == _lib.ERR_DIMENSION:
        raise DimensionMismatch(f"state, workspace and pre do not match (N = {N}, q_u + 1 = {pre.qu1}, "
                                f"q_v + 1 = {pre.qv1})")
    ws.dev.check(st)
    ws.cheb_steps = ws.dev.last_error().cheb_steps
    return ws.ut, ws.vt


def edg_rhs(u, v, t: float, N: int, h: float, pre, nblocks=None, cheb: str = "global"):
    """Allocating version of ``edg_rhs_`` (src/rhs.jl:560-571): copies of (u_t, v_t)."""
    u = _state(u); v = _state(v)
    nu, nv = pre.qu1, pre.qv1
    if (u.shape != (N, N, nu, nu) or v.shape != (N, N, nv, nv) or pre.Ainv_u.shape[1] != N * N):
        raise DimensionMismatch(f"state, workspace and pre do not match (N = {N}, q_u + 1 = {nu}, "
                                f"q_v + 1 = {nv})")
    ws = Workspace(pre, N, nblocks=nblocks, cheb=cheb)
    ut, vt = edg_rhs_(ws, u, v, t, N, h, pre)
    return ut.copy(order="F"), vt.copy(order="F")


def rk4_step_(u, v, t: float, dt: float, N: int, h: float, pre, ws: Workspace):
    """One classic RK4 step in place (src/solver.jl:289-321); ``all(ws.fin)``
    tells whether every entry stayed finite."""
    if not (isinstance(u, np.ndarray) and u.dtype == np.float64 and u.flags.f_contiguous
            and isinstance(v, np.ndarray) and v.dtype == np.float64 and v.flags.f_contiguous):
        raise TypeError("rk4_step_ updates u and v in place: pass Fortran-ordered float64 arrays")
    ws._check_sizes(u, v, N, pre)
    fin = C.c_int(1)
    st = lib.wedg_rk4_step(ws.dev.h, _ptr(u), _ptr(v), float(t), float(dt), int(N), float(h), C.byref(fin))
    ws.dev.check(st)
    ws.fin = [bool(fin.value)]
    return u, v


def rk4_step(u, v, t: float, dt: float, N: int, h: float, pre, nblocks=None, cheb: str = "global"):
    """Allocating version of ``rk4_step_`` (src/solver.jl:331-339)."""
    ws = Workspace(pre, N, nblocks=nblocks, cheb=cheb)
    uu = np.array(_state(u), order="F", copy=True)
    vv = np.array(_state(v), order="F", copy=True)
    return rk4_step_(uu, vv, t, dt, N, h, pre, ws)


def time_step(q: int, N: int, T: float, dt_fac=10):
    """dt = h / (dt_fac q^2), nsteps = round(T / dt), dt = T / nsteps (src/solver.jl:200-205)."""
    h = 1.0 / N
    dt = h / (dt_fac * q ** 2)
    nsteps = int(round(T / dt))
    return T / nsteps, nsteps


def solve(q: int, N: int, T: float, *, mixed: bool = True, flux: str = "upwind", flux_params=None,
          sigma: float = 0.0, boundary="absorbing_2beta", sigma_dirichlet=None, kappa: float = 0.0,
          phase=(0.0, 0.0), coefficients=None, mode: str = "mms", kx: int = 1, ky: int = 1,
          dt_fac=10, mass_solver: str = "chebyshev", nblocks=None, initial_profile=None,
          return_state: bool = False, source: GaussianSource = None, cheb: str = "global"):
    """Integrate the semi-discrete scheme to time T on the device (src/solver.jl:167-228).

    Returns the errors (erru, errv, errgradu, errE) at T in :mms mode, or the
    final state with the quadrature operators (``return_state=True``). The
    time loop runs ``nsteps`` RK4 steps on the device-resident state with
    t += dt after every step, as the reference's loop.
    """
    mode = str(mode).lstrip(":")
    if mode not in ("mms", "initial_value", "source"):
        raise ValueError(f"unknown mode {mode} (expected :mms, :initial_value or :source)")
    if mode == "source" and not return_state:
        raise NotImplementedError("source-mode time series are not produced by this build; "
                                  "use return_state=True for the final state")
    pre, geo = precompute(q, N, mixed=mixed, flux=flux, flux_params=flux_params, sigma=sigma,
                          boundary=boundary, sigma_dirichlet=sigma_dirichlet, kappa=kappa, phase=phase,
                          coefficients=coefficients, mode=mode, kx=kx, ky=ky, source=source,
                          mass_solver=mass_solver, initial_profile=initial_profile)
    h = 1.0 / N
    dt, nsteps = time_step(q, N, T, dt_fac)
    ws = Workspace(pre, N, nblocks=nblocks, cheb=cheb)
    u = np.zeros((N, N, pre.qu1, pre.qu1), order="F")
    v = np.zeros((N, N, pre.qv1, pre.qv1), order="F")
    st = lib.wedg_set_state(ws.dev.h, _ptr(u), _ptr(v))
    ws.dev.check(st)
    tend = C.c_double(0.0)
    st = lib.wedg_advance(ws.dev.h, 0.0, float(dt), int(nsteps), int(mode == "source"), C.byref(tend))
    ws.dev.check(st)
    st = lib.wedg_get_state(ws.dev.h, _ptr(u), _ptr(v))
    ws.dev.check(st)
    if return_state:
        return SimpleNamespace(u=u, v=v, N=N, h=h, P_u=pre.P_u, P_v=pre.P_v, S_u=geo.S_u, W2d=geo.W2d,
                               err_c2=geo.err_c2, t=tend.value, dt=dt, nsteps=nsteps)
    if mode != "mms":
        raise ValueError("errors are only available in :mms mode; use return_state = true")
    return mms_errors(u, v, T, N, h, pre, geo, kx, ky, phase)


def mms_errors(u, v, T, N, h, pre, geo, kx=1, ky=1, phase=(0.0, 0.0)):
    """||e_u||, ||e_v||, ||grad_h e_u||, ||e||_en at T against u = cos(wT) Phi
    (src/solver.jl:774-810), vectorized over the elements."""
    px, py = phase
    W = geo.W2d
    ct, st = math.cos(pre.omega * T), math.sin(pre.omega * T)
    tpx, tpy = 2 * math.pi * kx, 2 * math.pi * ky
    X = geo.xloc[None, :, None, :]          # [j, k, iy, ix]
    Y = geo.yloc[:, None, :, None]
    sp = np.cos(tpx * X + px) * np.sin(tpy * Y + py)
    spx = -tpx * np.sin(tpx * X + px) * np.sin(tpy * Y + py)
    spy = tpy * np.cos(tpx * X + px) * np.cos(tpy * Y + py)
    P_u, P_v, S_u = pre.P_u, pre.P_v, geo.S_u
    uh = np.einsum("pa,jkab,qb->jkpq", P_u, u, P_u, optimize=True)
    vh = np.einsum("pa,jkab,qb->jkpq", P_v, v, P_v, optimize=True)
    uxh = (2 / h) * np.einsum("pa,jkab,qb->jkpq", P_u, u, S_u, optimize=True)
    uyh = (2 / h) * np.einsum("pa,jkab,qb->jkpq", S_u, u, P_u, optimize=True)
    eu = (ct - 1.0) * sp - uh
    ev = -pre.omega * st * sp - vh
    ex = (ct - 1.0) * spx - uxh
    ey = (ct - 1.0) * spy - uyh
    s = h / 2
    return SimpleNamespace(
        erru=s * math.sqrt(np.sum(W * eu ** 2)),
        errv=s * math.sqrt(np.sum(W * ev ** 2)),
        errgradu=s * math.sqrt(np.sum(W * (ex ** 2 + ey ** 2))),
        errE=s * math.sqrt(np.sum(W * (0.5 * geo.err_c2 * (ex ** 2 + ey ** 2) + 0.5 * ev ** 2))),
    )


def difference_norms(A, B):
    """L2 norms of u_A - u_B, v_A - v_B and the energy norm of the difference of
    two final states of ``solve(...; return_state=True)`` (src/solver.jl:819-836)."""
    h = B.h
    W = B.W2d
    eu = A.u - B.u
    ev = A.v - B.v
    du = np.einsum("pa,jkab,qb->jkpq", B.P_u, eu, B.P_u, optimize=True)
    dux = (2 / h) * np.einsum("pa,jkab,qb->jkpq", B.P_u, eu, B.S_u, optimize=True)
    duy = (2 / h) * np.einsum("pa,jkab,qb->jkpq", B.S_u, eu, B.P_u, optimize=True)
    dv = np.einsum("pa,jkab,qb->jkpq", B.P_v, ev, B.P_v, optimize=True)
    s = h / 2
    return SimpleNamespace(
        erru=s * math.sqrt(np.sum(W * du ** 2)),
        errv=s * math.sqrt(np.sum(W * dv ** 2)),
        errE=s * math.sqrt(np.sum(W * (0.5 * B.err_c2 * (dux ** 2 + duy ** 2) + 0.5 * dv ** 2))),
    )
