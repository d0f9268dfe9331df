"""Gradients of a source-mode boundary misfit with respect to per-element β and
c², by the discrete adjoint of the device RK4 run (``wedg_adj_*`` of libwedg.so).

The forward map is the reference's ``solve(...; mode = :source)``
(src/solver.jl:167-228, 892-958) with coefficients constant on every element:
zero initial state, ``rk4_step!`` (src/solver.jl:289-321) of ``edg_rhs!``
(src/rhs.jl:490-550) with the Gaussian source (src/solver.jl:80-86, 647-662),
t += dt after every step. The objective, which the reference does not have, is
the L² misfit of u_h on the bottom side y = 0 after every step,

    J = ½ Σ_n Σ_k Σ_i (h/2) w_i (u_h(x_ki, 0, t_n) − g_n[k, i])²,

with g the target traces ([n, k, i]); ``SourceAdjoint.traces`` gives the traces
of a run, so a target is the traces of a run with other coefficients.

``SourceAdjoint.gradient`` returns J, dJ/dc² and dJ/dβ ((N, N) arrays, row j in
y, column k in x) by the adjoint of every stage and of RK4, with binomial
checkpointing of the RK4 states; ``misfit_rrule`` packages it as a ChainRules
``rrule`` does: the value and a pullback. ``rhs_vjp`` / ``rhs_tl`` are the
vector-Jacobian and Jacobian-vector products of one right-hand side.
"""
from __future__ import annotations

import ctypes as C
from types import SimpleNamespace

import numpy as np

from . import _lib
from ._lib import WedgError, lib
from .coefficients import constant_coefficients
from .precompute import GaussianSource, precompute

__all__ = ["SourceAdjoint", "AdjointBatch", "misfit_rrule", "beta_field_workload", "steepening_batch_workload"]

_dp = C.POINTER(C.c_double)


def _ptr(a):
    return a.ctypes.data_as(_dp)


def _field(a, N, name):
    a = np.asarray(a, dtype=np.float64)
    if a.ndim == 0:
        return np.full((N, N), float(a), order="F")
    if a.shape != (N, N):
        raise ValueError(f"{name} must be a number or an (N, N) array (got shape {a.shape})")
    return np.asfortranarray(a)


class SourceAdjoint:
    """A source-mode problem with per-element coefficients on the device, its
    forward runs, misfit and gradient.

    Arguments as ``solve`` (src/solver.jl:167-188) where they apply; ``alpha``
    is a number or an (N, N) array and stays fixed, c² and β are set with
    ``set_params``. ``cheb`` is the Chebyshev step-count policy of the forward
    steps ("global": the reference's; "element"). ``lanes`` is the lanes per
    element of the forward stage kernels (``Workspace``): 1, or 2 where the
    two-lane kernels admit the problem; the adjoint kernels keep one."""

    def __init__(self, q: int, N: int, *, mixed: bool = False, flux: str = "upwind", flux_params=None,
                 sigma: float = 0.0, boundary="absorbing_2beta", sigma_dirichlet=None, kappa: float = 0.0,
                 alpha=1e-5, source: GaussianSource = None, cheb: str = "global", lanes: int = 1):
        source = source or GaussianSource()
        if sigma_dirichlet is None:
            sigma_dirichlet = sigma
        pre, _ = precompute(q, N, mixed=mixed, flux=flux, flux_params=flux_params, sigma=sigma,
                            boundary=boundary, sigma_dirichlet=sigma_dirichlet, kappa=kappa, mode="source",
                            source=source, coefficients=constant_coefficients(c=1.0, alpha=1e-5, beta=0.0))
        self.q, self.N, self.h = q, N, 1.0 / N
        self.qu1, self.qv1 = pre.qu1, pre.qv1
        self.pre = pre
        self.alpha = _field(alpha, N, "alpha")
        d = _lib.WedgAdjDesc()
        d.N, d.qu1, d.qv1 = N, pre.qu1, pre.qv1
        d.theta, d.tau, d.lambda_u, d.lambda_v = pre.theta, pre.tau, pre.lambda_u, pre.lambda_v
        for i, side in enumerate(_lib.SIDES):
            b = pre.bc[side]
            d.bc[i].kind = _lib.KINDS[b["kind"]]
            d.bc[i].gamma_abc, d.bc[i].eta_abc, d.bc[i].sigma = b["gamma_abc"], b["eta_abc"], b["sigma"]
            d.bc[i].kb, d.bc[i].kappa = b["kb"], b["kappa"]
        d.mass_tol = pre.mass_tol
        d.src_omega, d.src_Ts = source.omega, source.Ts
        keep = {}
        for n in ("w", "P_u", "S_u", "Ebl_u", "Ebr_u", "E1bl_u", "E1br_u", "P_h_v", "W2d_h", "src_projE"):
            keep[n] = np.asfortranarray(getattr(pre, n), dtype=np.float64)
            setattr(d, n, _ptr(keep[n]))
        d.alpha = _ptr(self.alpha)
        h = C.c_void_p()
        pol = {"global": _lib.CHEB_GLOBAL, "element": _lib.CHEB_ELEMENT}[cheb]
        st = lib.wedg_adj_create(C.byref(d), pol, C.byref(h))
        if st != _lib.OK:
            raise WedgError(f"wedg_adj_create failed with status {st}")
        self.h_ = h
        self.c2 = np.ones((N, N), order="F")
        self.beta = np.zeros((N, N), order="F")
        self.target_steps = 0
        if lanes != 1:
            self.set_lanes(lanes)

    def set_lanes(self, lanes: int):
        """Lanes per element of the forward stage kernels (1 or 2)."""
        self._check(lib.wedg_adj_set_lanes_per_element(self.h_, int(lanes)), "set_lanes")

    @property
    def lanes(self) -> int:
        """The lanes per element the forward steps use (2 only where admitted)."""
        n = C.c_int(0)
        self._check(lib.wedg_adj_lanes_per_element(self.h_, C.byref(n)), "lanes")
        return n.value

    def __del__(self):
        h = getattr(self, "h_", None)
        if h is not None and h.value:
            lib.wedg_adj_destroy(h)
            self.h_ = None

    def _check(self, st, what):
        if st == _lib.OK:
            return
        msg = {_lib.ERR_MASS: "mass matrix not positive definite: 1 - 2βv_h ≤ 0",
               _lib.ERR_NONFINITE: "non-finite state (blow-up)",
               _lib.ERR_BOUNDARY_TRACE: "boundary weight 1 - kβv_h ≤ 0 at a face point",
               _lib.ERR_ARGUMENT: "invalid argument", _lib.ERR_HIP: "HIP runtime error"}.get(st, f"status {st}")
        raise WedgError(f"{what}: {msg}")

    def set_params(self, c2, beta):
        """Per-element c² (> 0) and β, numbers or (N, N) arrays."""
        self.c2 = _field(c2, self.N, "c2")
        self.beta = _field(beta, self.N, "beta")
        self._check(lib.wedg_adj_set_params(self.h_, _ptr(self.c2), _ptr(self.beta)), "set_params")

    def set_target(self, target):
        """The target traces g, (nsteps, N, q_u+1): [n, k, i] after step n + 1."""
        g = np.ascontiguousarray(target, dtype=np.float64)
        if g.ndim != 3 or g.shape[1:] != (self.N, self.qu1):
            raise ValueError(f"target must be (nsteps, N, q_u+1) = (*, {self.N}, {self.qu1})")
        self._check(lib.wedg_adj_set_target(self.h_, g.shape[0], _ptr(g)), "set_target")
        self.target_steps = g.shape[0]

    def traces(self, dt: float, nsteps: int):
        """u_h at the bottom face points after every step, (nsteps, N, q_u+1)."""
        tr = np.zeros((nsteps, self.N, self.qu1))
        self._check(lib.wedg_adj_forward(self.h_, float(dt), int(nsteps), None, _ptr(tr)), "forward")
        return tr

    def misfit(self, dt: float, nsteps: int) -> float:
        J = C.c_double(0.0)
        self._check(lib.wedg_adj_forward(self.h_, float(dt), int(nsteps), C.byref(J), None), "forward")
        return J.value

    def gradient(self, dt: float, nsteps: int, snapshots: int = 32):
        """(J, dJ/dc², dJ/dβ) with binomial checkpointing in `snapshots` state slots."""
        J = C.c_double(0.0)
        gc2 = np.zeros((self.N, self.N), order="F")
        gb = np.zeros((self.N, self.N), order="F")
        st = lib.wedg_adj_gradient(self.h_, float(dt), int(nsteps), int(snapshots), C.byref(J), _ptr(gc2), _ptr(gb))
        self._check(st, "gradient")
        return J.value, gc2, gb

    def stats(self):
        s = _lib.WedgAdjStats()
        lib.wedg_adj_last_stats(self.h_, C.byref(s))
        return SimpleNamespace(forward_steps=s.forward_steps, taped_steps=s.taped_steps,
                               adjoint_steps=s.adjoint_steps, snapshots_taken=s.snapshots_taken,
                               snapshots_used=s.snapshots_used, tapes=s.tapes, enqueue_ms=s.enqueue_ms,
                               total_ms=s.total_ms)

    def _state(self, u, v):
        u = np.asfortranarray(u, dtype=np.float64)
        v = np.asfortranarray(v, dtype=np.float64)
        if u.shape != (self.N, self.N, self.qu1, self.qu1) or v.shape != (self.N, self.N, self.qv1, self.qv1):
            raise ValueError("state shapes do not match the problem")
        return u, v

    def rhs_vjp(self, u, v, t, ub, vb):
        """Cotangents of (u, v, c², β) from those of (u_t, v_t) of one right-hand side at time t."""
        u, v = self._state(u, v)
        ub, vb = self._state(ub, vb)
        Ub = np.zeros_like(u, order="F"); Vb = np.zeros_like(v, order="F")
        c2b = np.zeros((self.N, self.N), order="F"); bb = np.zeros((self.N, self.N), order="F")
        st = lib.wedg_adj_rhs_vjp(self.h_, _ptr(u), _ptr(v), float(t), _ptr(ub), _ptr(vb), _ptr(Ub), _ptr(Vb),
                                  _ptr(c2b), _ptr(bb))
        self._check(st, "rhs_vjp")
        return Ub, Vb, c2b, bb

    def rhs_tl(self, u, v, t, du, dv, dc2, dbeta):
        """(du_t, dv_t) of one right-hand side at time t along (du, dv, dc², dβ)."""
        u, v = self._state(u, v)
        du, dv = self._state(du, dv)
        dc2 = _field(dc2, self.N, "dc2"); dbeta = _field(dbeta, self.N, "dbeta")
        ou = np.zeros_like(u, order="F"); ov = np.zeros_like(v, order="F")
        st = lib.wedg_adj_rhs_tl(self.h_, _ptr(u), _ptr(v), float(t), _ptr(du), _ptr(dv), _ptr(dc2), _ptr(dbeta),
                                 _ptr(ou), _ptr(ov))
        self._check(st, "rhs_tl")
        return ou, ov


def misfit_rrule(adj: SourceAdjoint, c2, beta, dt: float, nsteps: int, snapshots: int = 32):
    """The value J(c², β) and its pullback J̄ -> (J̄ dJ/dc², J̄ dJ/dβ), as a
    ChainRules ``rrule`` returns them (the gradient is computed once, here)."""
    adj.set_params(c2, beta)
    J, gc2, gb = adj.gradient(dt, nsteps, snapshots)

    def pullback(Jbar=1.0):
        return Jbar * gc2, Jbar * gb

    return J, pullback


def beta_field_workload(N: int = 64, q: int = 4, steps: int = 2000, seed: int = 1, cheb: str = "global",
                        x0=(0.5, 0.7), lanes: int = 1):
    """The receiver-misfit gradient workload: q_u = q_v = q on an N x N mesh in
    source mode, a Gaussian source of width 0.05 at (0.5, 0.7) (amplitude 1,
    omega 16 pi, Ts = 1), c = 1, alpha = 1e-5, beta per element
    4 + 2 sin(2 pi x) cos(2 pi y) + N(0, 0.1) at the element centres (numpy
    generator of `seed`), the bottom side absorbing_2beta and the other three
    walls with kappa = 0.5, `steps` RK4 steps of dt = h / (10 q^2); the target
    is the bottom trace of the same run with beta = 4. `x0` moves the source;
    `lanes`: the forward kernels' lanes per element."""
    import math
    h = 1.0 / N
    xc = (np.arange(N) + 0.5) * h
    X, Y = xc[None, :], xc[:, None]                     # column k in x, row j in y
    rng = np.random.default_rng(seed)
    beta = 4.0 + 2.0 * np.sin(2 * math.pi * X) * np.cos(2 * math.pi * Y) + 0.1 * rng.standard_normal((N, N))
    adj = SourceAdjoint(q, N, mixed=False, boundary=dict(bot="absorbing_2beta", top="wall", lft="wall", rgt="wall"),
                        kappa=0.5, alpha=1e-5, source=GaussianSource(x0=tuple(x0), r0=0.05), cheb=cheb, lanes=lanes)
    dt = h / (10 * q * q)
    adj.set_params(1.0, 4.0)
    adj.set_target(adj.traces(dt, steps))
    c2 = np.ones((N, N), order="F")
    adj.set_params(c2, beta)
    return SimpleNamespace(adj=adj, c2=c2, beta=np.asfortranarray(beta), dt=dt, steps=steps, N=N, q=q)


class AdjointBatch:
    """Forward runs and gradients of several ``SourceAdjoint`` instances
    together (``wedg_adj_batch_*``): lockstep rounds on one stream, one launch
    per RK stage and per adjoint kernel for each group of instances with the
    same degrees. Every instance gets exactly its own run's results. The
    instances keep their coefficients and targets."""

    def __init__(self, instances):
        self.instances = list(instances)
        n = len(self.instances)
        if n < 1:
            raise ValueError("an adjoint batch needs at least one instance")
        arr = (C.c_void_p * n)(*[a.h_.value for a in self.instances])
        h = C.c_void_p()
        st = lib.wedg_adj_batch_create(arr, n, C.byref(h))
        if st != _lib.OK:
            raise WedgError(f"wedg_adj_batch_create failed with status {st}")
        self.h = h

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value:
            lib.wedg_adj_batch_destroy(h)
            self.h = None

    def _steps(self, dts, nsteps):
        n = len(self.instances)
        dts = np.ascontiguousarray(np.broadcast_to(np.asarray(dts, dtype=np.float64), (n,)))
        ns = (C.c_long * n)(*[int(k) for k in np.broadcast_to(np.asarray(nsteps), (n,))])
        return dts, ns

    def _raise(self, st, status, what):
        for a, si in zip(self.instances, status):
            if si != _lib.OK:
                a._check(int(si), what)
        if st != _lib.OK:
            raise WedgError(f"{what} failed with status {st}")

    def forward(self, dts, nsteps):
        """J of every instance after nsteps[i] steps of dts[i]."""
        n = len(self.instances)
        dts, ns = self._steps(dts, nsteps)
        J = np.zeros(n)
        status = np.zeros(n, dtype=np.int32)
        st = lib.wedg_adj_batch_forward(self.h, _ptr(dts), ns, _ptr(J), status.ctypes.data_as(C.POINTER(C.c_int)))
        self._raise(st, status, "batch forward")
        return J

    def gradient(self, dts, nsteps, snapshots: int = 32):
        """[(J, dJ/dc², dJ/dβ)] of every instance."""
        n = len(self.instances)
        dts, ns = self._steps(dts, nsteps)
        J = np.zeros(n)
        gc2 = [np.zeros((a.N, a.N), order="F") for a in self.instances]
        gb = [np.zeros((a.N, a.N), order="F") for a in self.instances]
        pc = (_dp * n)(*[_ptr(g) for g in gc2])
        pb = (_dp * n)(*[_ptr(g) for g in gb])
        status = np.zeros(n, dtype=np.int32)
        st = lib.wedg_adj_batch_gradient(self.h, _ptr(dts), ns, int(snapshots), _ptr(J), pc, pb,
                                         status.ctypes.data_as(C.POINTER(C.c_int)))
        self._raise(st, status, "batch gradient")
        return [(float(J[i]), gc2[i], gb[i]) for i in range(n)]

    def time(self):
        """(device ms, host issue ms) of the last batch call."""
        d, e = C.c_double(0.0), C.c_double(0.0)
        lib.wedg_adj_batch_time(self.h, C.byref(d), C.byref(e))
        return d.value, e.value


def steepening_batch_workload(count: int = 16, seed: int = 2, T: float = 1.0, x0=(0.5, 0.3), r0: float = 0.04,
                              margin: float = 0.3, cheb: str = "global", lanes: int = 1):
    """The batched-adjoint workload: `count` source-mode instances drawn from
    numpy's generator of `seed`, each with N in {64, 96, 128}, q_u = q_v in
    {5, 6}, c = 1, alpha log-uniform in [1e-6, 1e-5], beta0 uniform in [6, 8],
    the bottom side absorbing_2beta and each other side absorbing_2beta,
    absorbing or a wall (kappa 0.5). The Gaussian source has width r0 at x0
    (omega 16 pi, Ts = 1); its amplitude is the largest for which the run with
    beta = beta0 on a 32 x 32 mesh of the same degree keeps
    min(1 - 2 beta0 v_h) >= margin + 0.05 at its record steps (bisection), so
    that the instance keeps about `margin`. The run takes round(T / dt) steps of
    dt = h / (10 q^2). The gradient is taken at beta = beta0 + 0.5 sin(2 pi x)
    cos(2 pi y) per element (c^2 = 1); the target is the bottom trace of the
    run with beta = beta0. `lanes`: the forward kernels' lanes per element."""
    import math
    rng = np.random.default_rng(seed)
    kinds = ("absorbing_2beta", "absorbing", "wall")
    out = []
    for _ in range(count):
        N = int(rng.choice([64, 96, 128]))
        q = int(rng.choice([5, 6]))
        alpha = float(10 ** rng.uniform(-6, -5))
        beta0 = float(rng.uniform(6, 8))
        sides = {s: kinds[int(rng.integers(0, 3))] for s in ("bot", "top", "lft", "rgt")}
        sides["bot"] = "absorbing_2beta"   # the receivers' side absorbs
        amp = _calibrated_amplitude(q, alpha, beta0, sides, T, x0, r0, margin + 0.05)
        h = 1.0 / N
        dt = h / (10 * q * q)
        nsteps = max(1, int(round(T / dt)))
        dt = T / nsteps
        xc = (np.arange(N) + 0.5) * h
        X, Y = xc[None, :], xc[:, None]
        beta = np.asfortranarray(beta0 + 0.5 * np.sin(2 * math.pi * X) * np.cos(2 * math.pi * Y))
        for _try in range(8):   # the finer mesh may steepen a little more than the calibration mesh
            src = GaussianSource(x0=tuple(x0), r0=r0, amplitude=amp)
            adj = SourceAdjoint(q, N, mixed=False, boundary=sides, kappa=0.5, alpha=alpha, source=src, cheb=cheb,
                                lanes=lanes)
            try:
                adj.set_params(1.0, beta0)
                adj.set_target(adj.traces(dt, nsteps))
                adj.set_params(1.0, beta)
                adj.misfit(dt, nsteps)
                break
            except WedgError:
                amp *= 0.9
        else:
            raise WedgError("no source amplitude keeps the instance's mass matrices positive definite")
        out.append(SimpleNamespace(adj=adj, N=N, q=q, alpha=alpha, beta0=beta0, beta=beta, sides=sides,
                                   amplitude=amp, dt=dt, steps=nsteps))
    return out


def _calibrated_amplitude(q, alpha, beta0, sides, T, x0, r0, floor, Ncal: int = 32, iters: int = 10):
    """The largest source amplitude (bisection) for which the source-mode run
    of degree q with beta = beta0 on an Ncal x Ncal mesh to T keeps
    min(1 - 2 beta0 v_h) >= floor at its record steps."""
    from .solver import solve

    def ndmin(amp):
        try:
            H = solve(q, Ncal, T, mixed=False, mode="source", boundary=sides, kappa=0.5,
                      coefficients=_constant(alpha, beta0), source=GaussianSource(x0=tuple(x0), r0=r0, amplitude=amp),
                      record_every=4)
            return float(np.min(H.ndmin))
        except WedgError:
            return -np.inf

    H = solve(q, Ncal, T, mixed=False, mode="source", boundary=sides, kappa=0.5, coefficients=_constant(alpha, 0.0),
              source=GaussianSource(x0=tuple(x0), r0=r0, amplitude=1.0), record_every=4)
    vlin = float(max(np.max(H.vmax), 1e-300))
    hi = (1.0 - floor) / (2.0 * beta0 * vlin)   # the linear estimate: the nonlinear peak is higher
    if ndmin(hi) >= floor:
        return hi
    lo = 0.0
    for _ in range(iters):
        mid = 0.5 * (lo + hi)
        if ndmin(mid) >= floor:
            lo = mid
        else:
            hi = mid
    return lo


def _constant(alpha, beta):
    from .coefficients import constant_coefficients as cc
    return cc(c=1.0, alpha=alpha, beta=beta)
