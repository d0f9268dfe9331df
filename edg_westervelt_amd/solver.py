"""The reference's hot-path API on the device: ``Workspace``, ``edg_rhs_``
(``edg_rhs!``), ``edg_rhs``, ``rk4_step_`` (``rk4_step!``), ``rk4_step``,
``solve`` (with the source-mode time series of ``_run_source``) and
``difference_norms``.

Mirrors src/rhs.jl:166-220, 475-571 and src/solver.jl:88-339, 772-958 of the
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
from .quadrature import JacobiP

__all__ = ["Workspace", "edg_rhs_", "edg_rhs", "rk4_step_", "rk4_step", "solve", "difference_norms",
           "DimensionMismatch", "WedgError"]

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def _ptr(a):
    return a.ctypes.data_as(_dp)


class Workspace:
    """Device-resident arrays for ``edg_rhs_`` and ``rk4_step_`` on an N x N mesh
    (src/rhs.jl:184-220).

    ``nblocks`` is accepted for the reference's signature; the device grid
    replaces the element blocks, and results do not depend on it. ``cheb``
    selects the Chebyshev step count: ``"global"`` (the reference: every element
    takes the largest count) or ``"element"`` (each element its own a-priori
    count to the same 1e-14 residual bound). ``lanes`` is the lanes per element
    of the device stage kernels: 1, or 2 (a pair of lanes per element: twice
    the waves, half the registers per lane) where the problem allows it
    (q_u <= 6, coefficients constant on each element, Chebyshev mass solve);
    results agree to rounding. ``"auto"`` takes two lanes where they were
    measured faster: two-lane problems with q_u >= 3 whose one-lane grid
    leaves SIMDs idle (fewer than 65,536 elements) or with q_u = 6, where
    one lane spills registers.
    """

    def __init__(self, pre, N: int, nblocks=None, cheb: str = "global", lanes=1):
        if nblocks is not None and nblocks < 1:
            raise ValueError(f"nblocks must be ≥ 1 (got {nblocks})")
        if pre.Ainv_u.shape[1] != N * N:
            raise DimensionMismatch(f"state, workspace and pre do not match (N = {N}, q_u + 1 = {pre.qu1}, "
                                    f"q_v + 1 = {pre.qv1})")
        self.pre = pre
        self.N = N
        self.nb = 1
        self.qu1, self.qv1, self.q_high = pre.qu1, pre.qv1, pre.q_high
        if lanes not in (1, 2, "auto"):
            raise ValueError(f"lanes must be 1, 2 or 'auto' (got {lanes!r})")
        self.problem = _lib.Problem(pre, N)
        if lanes == "auto":
            q_u = pre.qu1 - 1
            lanes = 2 if (self.problem.info()["two_lane"] and q_u >= 3 and (N * N < 65536 or q_u == 6)) else 1
        self.dev = _lib.DeviceWorkspace(self.problem, cheb, lanes)
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


def edg_rhs(u, v, t: float, N: int, h: float, pre, nblocks=None, cheb: str = "global", lanes: int = 1):
    """Allocating version of ``edg_rhs_`` (src/rhs.jl:560-571): copies of (u_t, v_t)."""
    u = _state(u); v = _state(v)
    nu, nv = pre.qu1, pre.qv1
    if (u.shape != (N, N, nu, nu) or v.shape != (N, N, nv, nv) or pre.Ainv_u.shape[1] != N * N):
        raise DimensionMismatch(f"state, workspace and pre do not match (N = {N}, q_u + 1 = {nu}, "
                                f"q_v + 1 = {nv})")
    ws = Workspace(pre, N, nblocks=nblocks, cheb=cheb, lanes=lanes)
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


def rk4_step(u, v, t: float, dt: float, N: int, h: float, pre, nblocks=None, cheb: str = "global",
             lanes: int = 1):
    """Allocating version of ``rk4_step_`` (src/solver.jl:331-339)."""
    ws = Workspace(pre, N, nblocks=nblocks, cheb=cheb, lanes=lanes)
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
          return_state: bool = False, source: GaussianSource = None, cheb: str = "global",
          lanes: int = 1, record_every: int = 20, slice_y: float = 0.5,
          receivers=((0.45, 0.5), (0.7, 0.5), (0.85, 0.5)), snap_times=()):
    """Integrate the semi-discrete scheme to time T on the device (src/solver.jl:167-228).

    Returns the errors (erru, errv, errgradu, errE) at T in :mms mode, or the
    final state with the quadrature operators (``return_state=True``). The
    time loop runs ``nsteps`` RK4 steps on the device-resident state with
    t += dt after every step, as the reference's loop. :source mode returns
    the time series of ``_run_source`` (src/solver.jl:872-958) whatever
    ``return_state`` says, as the reference does. ``cheb`` and ``lanes``
    choose the device kernels as for ``Workspace``.
    """
    mode = str(mode).lstrip(":")
    if mode not in ("mms", "initial_value", "source"):
        raise ValueError(f"unknown mode {mode} (expected :mms, :initial_value or :source)")
    pre, geo = precompute(q, N, mixed=mixed, flux=flux, flux_params=flux_params, sigma=sigma,
                          boundary=boundary, sigma_dirichlet=sigma_dirichlet, kappa=kappa, phase=phase,
                          coefficients=coefficients, mode=mode, kx=kx, ky=ky, source=source,
                          mass_solver=mass_solver, initial_profile=initial_profile)
    h = 1.0 / N
    dt, nsteps = time_step(q, N, T, dt_fac)
    ws = Workspace(pre, N, nblocks=nblocks, cheb=cheb, lanes=lanes)
    u = np.zeros((N, N, pre.qu1, pre.qu1), order="F")
    v = np.zeros((N, N, pre.qv1, pre.qv1), order="F")
    st = lib.wedg_set_state(ws.dev.h, _ptr(u), _ptr(v))
    ws.dev.check(st)
    if mode == "source":
        return _run_source(ws, dt, nsteps, N, h, pre, geo, record_every=record_every, slice_y=slice_y,
                           receivers=receivers, snap_times=snap_times)
    tend = C.c_double(0.0)
    st = lib.wedg_advance(ws.dev.h, 0.0, float(dt), int(nsteps), 0, C.byref(tend))
    ws.dev.check(st)
    if return_state:
        st = lib.wedg_get_state(ws.dev.h, _ptr(u), _ptr(v))
        ws.dev.check(st)
        _set_diag(ws, geo)
        return SimpleNamespace(u=u, v=v, N=N, h=h, P_u=pre.P_u, P_v=pre.P_v, S_u=geo.S_u, W2d=geo.W2d,
                               err_c2=geo.err_c2, t=tend.value, dt=dt, nsteps=nsteps, workspace=ws)
    if mode != "mms":
        raise ValueError("errors are only available in :mms mode; use return_state = true")
    return _device_mms_errors(ws, T, N, h, pre, geo, kx, ky, phase)


def _set_diag(ws: Workspace, geo):
    """The point coefficients of the diagnostics and norms on the device, once per workspace."""
    if getattr(ws, "_diag_set", False):
        return
    c2 = np.asfortranarray(geo.err_c2, dtype=np.float64)
    bq = np.asfortranarray(geo.err_beta, dtype=np.float64)
    W2d = np.asfortranarray(geo.W2d, dtype=np.float64)
    ws.dev.check(lib.wedg_problem_set_diag(ws.problem.h, _ptr(c2), _ptr(bq), _ptr(W2d)))
    ws._diag_set = True


def _serial_sums(elem, N):
    """Sums of the element values elem[c, e] (e = j + N k) over j outer, k inner,
    one after the other, as the reference's loops add them (src/solver.jl:781, 825)."""
    order = (np.arange(N)[:, None] + N * np.arange(N)[None, :]).ravel()
    return [float(np.add.accumulate(np.concatenate(([0.0], elem[c, order])))[-1]) for c in range(elem.shape[0])]


def _device_mms_errors(ws: Workspace, T, N, h, pre, geo, kx=1, ky=1, phase=(0.0, 0.0)):
    """_mms_errors (src/solver.jl:774-810) of the workspace state on the device."""
    _set_diag(ws, geo)
    px, py = phase
    tpx, tpy = 2 * math.pi * kx, 2 * math.pi * ky
    cx = np.ascontiguousarray(np.cos(tpx * geo.xloc + px))
    sx = np.ascontiguousarray(np.sin(tpx * geo.xloc + px))
    cy = np.ascontiguousarray(np.cos(tpy * geo.yloc + py))
    sy = np.ascontiguousarray(np.sin(tpy * geo.yloc + py))
    ct, st = math.cos(pre.omega * T), math.sin(pre.omega * T)
    elem = np.zeros((4, N * N))
    ws.dev.check(lib.wedg_mms_errors(ws.dev.h, ct - 1.0, -pre.omega * st, tpx, tpy, _ptr(cx), _ptr(sx), _ptr(cy),
                                     _ptr(sy), _ptr(elem)))
    eu, ev, eg, eE = _serial_sums(elem, N)
    s = h / 2
    return SimpleNamespace(erru=s * math.sqrt(eu), errv=s * math.sqrt(ev), errgradu=s * math.sqrt(eg),
                           errE=s * math.sqrt(eE))


def locate(s: float, N: int):
    """Element index (1-based) and local coordinate in [-1, 1] of the physical
    coordinate s on an N-element line (src/solver.jl:880)."""
    i = min(max(int(math.ceil(s * N)), 1), N)
    return i, min(max(2 * (s * N - (i - 1)) - 1, -1.0), 1.0)


def vrow(r: float, qv1: int):
    """The modal basis of v at the local coordinate r, P_{d-1}(r) for d = 1..q_v+1
    (src/solver.jl:881)."""
    return np.array([JacobiP(float(r), 0.0, 0.0, d) for d in range(qv1)])


def _events(dt: float, nsteps: int, record_every: int, snap_times):
    """The steps (1-based) after which _run_source records diagnostics or takes a
    snapshot, with t after each step accumulated as t += dt (src/solver.jl:936-951):
    at most one snapshot per step, the earliest open snapshot time at or before t."""
    left = sorted(float(x) for x in snap_times)
    rec, snap, t = [], [], 0.0
    for it in range(1, nsteps + 1):
        t += dt
        if it % record_every == 0 or it == nsteps:
            rec.append(it)
        if left and t >= left[0]:
            snap.append(it)
            left.pop(0)
    return rec, snap


def _source_setup(ws: Workspace, dt: float, nsteps: int, N: int, pre, geo, *, record_every: int, slice_y: float,
                  receivers, snap_times):
    """The state of one source-mode run (src/solver.jl:872-933): the diagnostics
    coefficients on the device, the slice row, the receivers' elements and basis
    rows, the record and snapshot steps, and the series collected so far."""
    if record_every < 1:
        raise ValueError(f"record_every must be ≥ 1 (got {record_every})")
    qu1, qv1 = pre.qu1, pre.qv1
    c2 = np.asfortranarray(geo.err_c2, dtype=np.float64)
    bq = np.asfortranarray(geo.err_beta, dtype=np.float64)
    W2d = np.asfortranarray(geo.W2d, dtype=np.float64)
    ws.dev.check(lib.wedg_problem_set_diag(ws.problem.h, _ptr(c2), _ptr(bq), _ptr(W2d)))
    S = SimpleNamespace(ws=ws, dt=dt, nsteps=nsteps, N=N, pre=pre, geo=geo)
    S.j0, eta0 = locate(slice_y, N)
    S.Pveta0 = np.ascontiguousarray(vrow(eta0, qv1))
    S.slice_x = np.concatenate([geo.xloc[k, :] for k in range(N)])
    S.receivers = [tuple(map(float, r)) for r in receivers]
    nr = S.nr = len(S.receivers)
    S.rj = np.zeros(max(nr, 1), dtype=np.int32)
    S.rk = np.zeros(max(nr, 1), dtype=np.int32)
    rows = np.zeros((max(nr, 1), 2 * qv1))
    for i, (xr, yr) in enumerate(S.receivers):
        kr, xir = locate(xr, N)
        jr, etar = locate(yr, N)
        S.rj[i], S.rk[i] = jr - 1, kr - 1
        rows[i, :qv1] = vrow(etar, qv1)
        rows[i, qv1:] = vrow(xir, qv1)
    S.rows = np.ascontiguousarray(rows)
    rec_steps, snap_steps = _events(dt, nsteps, record_every, snap_times)
    S.stops = sorted(set(rec_steps) | set(snap_steps))
    S.rec_set, S.snap_set = set(rec_steps), set(snap_steps)
    S.E_el = np.zeros(N * N); S.vmax_el = np.zeros(N * N); S.nd_el = np.zeros(N * N)
    S.sl = np.zeros(N * qu1)
    S.order = (np.arange(N)[:, None] + N * np.arange(N)[None, :]).ravel()   # e for j outer, k inner
    S.rec_t, S.rec_slice, S.rec_E, S.rec_vmax, S.rec_nd = [], [], [], [], []
    S.recv_all = np.zeros((nsteps, nr))
    S.snaps = []
    S.t = 0.0
    S.done = 0
    return S


def _blow_up(S):
    """The reference's message for a non-finite step (src/solver.jl:940): t
    accumulated over the steps of the failing call."""
    tb = S.t
    for _ in range(S.ws.dev.last_error().step):
        tb += S.dt
    return WedgError(f"blow-up at t={tb!r}")


def _source_record(S, stop: int):
    """The diagnostics and snapshot of a record/snapshot step (src/solver.jl:907-923)."""
    dh = S.ws.dev.h
    N, qu1 = S.N, S.pre.qu1
    if stop in S.rec_set:
        S.ws.dev.check(lib.wedg_diagnostics(dh, _ptr(S.E_el), _ptr(S.vmax_el), _ptr(S.nd_el), S.j0 - 1,
                                            _ptr(S.Pveta0), _ptr(S.sl)))
        S.rec_t.append(S.t)
        S.rec_slice.append(S.sl.copy())
        S.rec_E.append(float(np.add.accumulate(np.concatenate(([0.0], S.E_el[S.order])))[-1]))
        S.rec_vmax.append(float(max(0.0, S.vmax_el.max())))
        S.rec_nd.append(float(min(np.inf, S.nd_el.min())))
    if stop in S.snap_set:
        G = np.zeros((N * qu1, N * qu1), order="F")
        S.ws.dev.check(lib.wedg_snapshot(dh, _ptr(G)))
        S.snaps.append((S.t, G))


def _source_result(S):
    N, qu1, geo = S.N, S.pre.qu1, S.geo
    recv_t = np.zeros(S.nsteps)
    tt = 0.0
    for i in range(S.nsteps):   # t after each step, accumulated as the time loop does
        tt += S.dt
        recv_t[i] = tt
    grid_x = np.concatenate([geo.xloc[k, :] for k in range(N)])
    grid_y = np.concatenate([geo.yloc[j, :] for j in range(N)])
    return SimpleNamespace(t=np.array(S.rec_t), slice_x=S.slice_x,
                           slices=np.array(S.rec_slice).T if S.rec_slice else np.zeros((N * qu1, 0)),
                           energy=np.array(S.rec_E), vmax=np.array(S.rec_vmax), ndmin=np.array(S.rec_nd),
                           recv_t=recv_t, recv=S.recv_all, receivers=S.receivers, snaps=S.snaps,
                           grid_x=grid_x, grid_y=grid_y, dt=S.dt)


def _run_source(ws: Workspace, dt: float, nsteps: int, N: int, h: float, pre, geo, *, record_every: int,
                slice_y: float, receivers, snap_times):
    """Source mode (src/solver.jl:872-958): RK4 on the device with v_h at the
    receivers after every step, the energy diagnostics and the slice along
    y = slice_y every ``record_every`` steps and after the last, and snapshots of
    v_h on the point grid at the first step with t >= each snapshot time.

    The receivers, the element values of the diagnostics, the slice and the
    snapshot grids are formed on the device (``wedg_advance_receivers``,
    ``wedg_diagnostics``, ``wedg_snapshot``); the host accumulates the element
    values serially in the order row j, column k, as the reference does."""
    S = _source_setup(ws, dt, nsteps, N, pre, geo, record_every=record_every, slice_y=slice_y,
                      receivers=receivers, snap_times=snap_times)
    nr = S.nr
    tend = C.c_double(0.0)
    for stop in S.stops:
        n = stop - S.done
        buf = np.zeros(max(n * nr, 1))
        st = lib.wedg_advance_receivers(ws.dev.h, S.t, float(dt), int(n), nr, S.rj.ctypes.data_as(_ip),
                                        S.rk.ctypes.data_as(_ip), _ptr(S.rows), 1, _ptr(buf), C.byref(tend))
        if st == _lib.ERR_NONFINITE:
            raise _blow_up(S)
        ws.dev.check(st)
        S.t = tend.value
        S.recv_all[S.done:stop] = buf[:n * nr].reshape(n, nr)
        S.done = stop
        _source_record(S, stop)
    return _source_result(S)


def mms_errors(u, v, T, N, h, pre, geo, kx=1, ky=1, phase=(0.0, 0.0)):
    """||e_u||, ||e_v||, ||grad_h e_u||, ||e||_en at T against u = cos(wT) Phi
    (src/solver.jl:774-810) of the host state (u, v), on the device."""
    u = _state(u); v = _state(v)
    ws = Workspace(pre, N)
    ws._check_sizes(u, v, N, pre)
    ws.dev.check(lib.wedg_set_state(ws.dev.h, _ptr(u), _ptr(v)))
    return _device_mms_errors(ws, T, N, h, pre, geo, kx, ky, phase)


def difference_norms(A, B):
    """L2 norms of u_A - u_B, v_A - v_B and the energy norm of the difference of
    two final states of ``solve(...; return_state=True)`` (src/solver.jl:819-836),
    on the device of B's workspace (B's tables and c^2, as the reference)."""
    ws = getattr(B, "workspace", None)
    if ws is None:
        raise ValueError("difference_norms needs B from solve(...; return_state = true) (its device workspace)")
    N = B.N
    uA = _state(A.u); vA = _state(A.v); uB = _state(B.u); vB = _state(B.v)
    if uA.shape != uB.shape or vA.shape != vB.shape:
        raise DimensionMismatch("difference_norms: the two states do not have the same shape")
    elem = np.zeros((4, N * N))
    ws.dev.check(lib.wedg_difference_norms(ws.dev.h, _ptr(uA), _ptr(vA), _ptr(uB), _ptr(vB), _ptr(elem)))
    eu, ev, _, eE = _serial_sums(elem, N)
    s = B.h / 2
    return SimpleNamespace(erru=s * math.sqrt(eu), errv=s * math.sqrt(ev), errE=s * math.sqrt(eE))
