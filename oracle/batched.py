"""ctypes front end of the C oracle (oracle/wedg_oracle.c). Test infrastructure.

``OraclePre(pre, N)`` binds the reference-layout tables of ``precompute`` to
the C struct; ``rhs``, ``rk4_steps``, ``mass_cheb`` and ``mass_cg`` follow
src/rhs.jl:490-550, src/solver.jl:289-321 and src/rhs.jl:862-951.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "_build", "libwedg_oracle.so")
_KINDS = {"absorbing_2beta": 0, "absorbing": 1, "dirichlet": 2, "wall": 3}
_SIDES = ("bot", "top", "lft", "rgt")

_PTRS = ("w", "P_u", "S_u", "Ebl_u", "Ebr_u", "E1bl_u", "E1br_u", "P_h_v", "W2d_h",
         "Cw", "Aw", "Ainv_u", "beta_hq", "beta_hE",
         "hc2", "haf", "hcf", "hbf", "vc2", "vaf", "vcf", "vbf",
         "hc2m", "hafm", "vc2m", "vafm", "hJC", "vJC",
         "bcB_sp", "bcB_spy", "bcT_sp", "bcT_spy", "bcL_sp", "bcL_spx", "bcR_sp", "bcR_spx",
         "PiSP", "PiBSP2", "PiCdiv", "PiDdiv", "src_projE")

_dp = C.POINTER(C.c_double)


class _Pre(C.Structure):
    _fields_ = ([("N", C.c_int), ("qu1", C.c_int), ("qv1", C.c_int), ("qh", C.c_int),
                 ("omega", C.c_double), ("theta", C.c_double), ("tau", C.c_double),
                 ("lambda_u", C.c_double), ("lambda_v", C.c_double),
                 ("mms_on", C.c_int), ("src_on", C.c_int), ("one_sided", C.c_int), ("use_cg", C.c_int),
                 ("mass_tol", C.c_double), ("src_omega", C.c_double), ("src_Ts", C.c_double),
                 ("kind", C.c_int * 4), ("gam", C.c_double * 4), ("eta", C.c_double * 4),
                 ("sig", C.c_double * 4), ("kb", C.c_double * 4), ("kap", C.c_double * 4)]
                + [(n, _dp) for n in _PTRS])


class Status(C.Structure):
    _fields_ = [("code", C.c_int), ("side", C.c_int), ("key", C.c_long), ("elem", C.c_long),
                ("j", C.c_long), ("k", C.c_long), ("step", C.c_long), ("cheb_steps", C.c_int),
                ("cg_unconverged", C.c_long)]


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = C.CDLL(_SO)
        for fn in ("orc_rhs", "orc_rk4_steps", "orc_mass_cheb", "orc_mass_cg"):
            getattr(L, fn).restype = C.c_int
        _lib = L
    return _lib


def _f(a) -> np.ndarray:
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


class OraclePre:
    """The C view of one problem's tables (the arrays are kept alive here)."""

    def __init__(self, pre, N: int):
        self.pre = pre
        self.N = N
        self._keep = {}
        s = _Pre()
        s.N, s.qu1, s.qv1, s.qh = N, pre.qu1, pre.qv1, pre.q_high
        s.omega, s.theta, s.tau = pre.omega, pre.theta, pre.tau
        s.lambda_u, s.lambda_v = pre.lambda_u, pre.lambda_v
        s.mms_on, s.src_on, s.one_sided = int(pre.mms_on), int(pre.src_on), int(pre.one_sided)
        s.use_cg = int(pre.mass_solver == "cg")
        s.mass_tol = pre.mass_tol
        s.src_omega, s.src_Ts = pre.source.omega, pre.source.Ts
        for i, side in enumerate(_SIDES):
            b = pre.bc[side]
            s.kind[i] = _KINDS[b["kind"]]
            s.gam[i], s.eta[i], s.sig[i] = b["gamma_abc"], b["eta_abc"], b["sigma"]
            s.kb[i], s.kap[i] = b["kb"], b["kappa"]
        for n in _PTRS:
            a = _f(getattr(pre, n))
            self._keep[n] = a
            setattr(s, n, a.ctypes.data_as(_dp))
        self.s = s

    def status(self, st: Status, what: str = "rhs"):
        raise_for_status(self.pre, self.N, st)


def raise_for_status(pre, N, st: Status):
    """The reference's error for a failed evaluation (src/rhs.jl:112-116, 768-776,
    911-918, 845; src/solver.jl:940)."""
    if st.code == 0:
        return
    if st.code in (1, 2):
        bs = pre.bc[_SIDES[st.side]]
        k = "2" if bs["kb"] == 2.0 else ""
        if st.code == 1:
            raise RuntimeError(f"boundary weight 1 - {k}βv_h ≤ 0 at a face point ({bs['kind']} side)")
        raise RuntimeError(f"boundary weight 1 - {k}βg ≤ 0 for the boundary data g ({bs['kind']} side)")
    if st.code == 3:
        raise RuntimeError(f"mass matrix not positive definite on element {st.elem}: 1 - 2βv_h ≤ 0")
    if st.code == 4:
        raise RuntimeError(f"mass matrix not positive definite on element ({st.j}, {st.k}): 1 - 2βv_h ≤ 0")
    if st.code == 5:
        raise RuntimeError(f"blow-up at step {st.step}")
    raise RuntimeError(f"oracle failure code {st.code}")


def _ptr(a):
    return a.ctypes.data_as(_dp)


def rhs(op: OraclePre, u, v, t: float, cheb_mode: int = 0, nthreads: int = 0, check: bool = True):
    """(u_t, v_t) of one evaluation (src/rhs.jl:490-550)."""
    u = _f(u); v = _f(v)
    ut = np.zeros_like(u, order="F"); vt = np.zeros_like(v, order="F")
    st = Status()
    lib().orc_rhs(C.byref(op.s), _ptr(u), _ptr(v), C.c_double(t), _ptr(ut), _ptr(vt),
                  C.c_int(cheb_mode), C.c_int(nthreads), C.byref(st))
    if check:
        raise_for_status(op.pre, op.N, st)
    return ut, vt, st


def rk4_steps(op: OraclePre, u, v, t0: float, dt: float, nsteps: int, cheb_mode: int = 0,
              nthreads: int = 0, check: bool = True):
    """nsteps classic RK4 steps in place on copies of (u, v), t += dt each step."""
    u = _f(u).copy(order="F"); v = _f(v).copy(order="F")
    st = Status()
    tend = C.c_double(0.0)
    lib().orc_rk4_steps(C.byref(op.s), _ptr(u), _ptr(v), C.c_double(t0), C.c_double(dt),
                        C.c_long(nsteps), C.c_int(cheb_mode), C.c_int(nthreads), C.byref(tend), C.byref(st))
    if check:
        raise_for_status(op.pre, op.N, st)
    return u, v, tend.value, st


def mass_cheb(op: OraclePre, V, R, cheb_mode: int = 0, nthreads: int = 0):
    """X = M(V)^-1 R by Chebyshev iteration; V, R in the element-last layout (nv, nv, N, N)."""
    V = _f(V); R = _f(R); X = np.zeros_like(R, order="F")
    st = Status()
    lib().orc_mass_cheb(C.byref(op.s), _ptr(V), _ptr(R), _ptr(X), C.c_int(cheb_mode), C.c_int(nthreads),
                        C.byref(st))
    raise_for_status(op.pre, op.N, st)
    return X, st


def mass_cg(op: OraclePre, V, R, nthreads: int = 0):
    V = _f(V); R = _f(R); X = np.zeros_like(R, order="F")
    st = Status()
    lib().orc_mass_cg(C.byref(op.s), _ptr(V), _ptr(R), _ptr(X), C.c_int(nthreads), C.byref(st))
    raise_for_status(op.pre, op.N, st)
    return X, st


def solve_state(pre, N: int, T: float, q: int, dt_fac: float = 10, cheb_mode: int = 0, nthreads: int = 0):
    """The time loop of ``solve`` (src/solver.jl:200-217) on the oracle: returns (u, v, t, dt, nsteps)."""
    h = 1.0 / N
    dt = h / (dt_fac * q ** 2)
    nsteps = int(round(T / dt))
    dt = T / nsteps
    op = OraclePre(pre, N)
    u = np.zeros((N, N, pre.qu1, pre.qu1), order="F")
    v = np.zeros((N, N, pre.qv1, pre.qv1), order="F")
    u, v, t, st = rk4_steps(op, u, v, 0.0, dt, nsteps, cheb_mode=cheb_mode, nthreads=nthreads)
    return u, v, t, dt, nsteps
