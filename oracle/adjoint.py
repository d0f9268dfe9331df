"""CPU oracle of the source-mode right-hand side with per-element coefficients,
its hand-written vector-Jacobian product (VJP), RK4 trajectories and the
boundary-receiver misfit with its gradient — test infrastructure only.

Only ``tests/`` import this module, as the checker; the product path never does.

What it restates. One evaluation of ``edg_rhs!`` in ``mode = :source``
(src/rhs.jl:490-550) for coefficients that are constant on every element: c²,
α, β given as (N, N) arrays (row j in y, column k in x). Source mode has zero
boundary data and no u₀ terms, so the interface terms K vanish
(src/solver.jl:647-662, 748-751); faces use the one-sided formulas of
``_interior_faces_os!`` (src/rhs.jl:663-700), which reduce to
``_interior_faces!`` where the coefficients agree. Boundary sides follow
``_abc_face`` / ``_wall_face`` (src/rhs.jl:74-96) with the side parameters of
``_boundary_side`` (src/solver.jl:39-45), the face value of c taken as
sqrt(c²) from inside the element (src/solver.jl:546-552). Lifts, volume terms
and the source follow src/rhs.jl:317-350, 421-473; the u-equation solves the
pinned c²-stiffness (src/solver.jl:528-531) and the mass system
M = Φᵀ diag((1 − 2βv_q) W_h) Φ (src/rhs.jl:778-790) densely with LAPACK.
RK4 follows src/solver.jl:289-321 and the time factor of the source
src/solver.jl:662.

The objective (no counterpart in the reference; the adjoint extension of the
hot path): J = ½ Σ_n Σ_k Σ_i (h/2) w_i (u_h(x_ki, 0, t_n) − g_n[k, i])², the
L² misfit of u_h on the bottom side y = 0 at its face quadrature points, over
the states after every RK4 step n = 1..S.

Independent tables. The 1D Legendre tables come from numpy's Legendre series
and Gauss rule (``numpy.polynomial.legendre``), not from the product's
quadrature module.

Pinning:
* ``rhs`` against the C oracle (oracle/wedg_oracle.c, itself pinned by the
  reference's fixtures) on precompute's tables for the same per-element
  coefficients (tests/test_adjoint_oracle.py), ≤ 1e-12;
* ``rhs_vjp`` against complex-step derivatives of ``rhs`` (exact to rounding:
  Im F(p + i h d) / h, h = 1e-30) through the dot-product identity
  ⟨ȳ, F'(p) d⟩ = ⟨F'(p)ᵀ ȳ, d⟩ in every argument (state, c², β);
* ``gradient`` against complex-step directional derivatives of ``misfit`` and
  against central differences of the C oracle's RK4.
"""
from __future__ import annotations

import math
from types import SimpleNamespace

import numpy as np
from numpy.polynomial import legendre as npleg

_KINDS = ("absorbing_2beta", "absorbing", "dirichlet", "wall")


def _onb(n_modes, x):
    """Orthonormal Legendre values L_d(x) = sqrt((2d+1)/2) P_d(x) and derivatives,
    (len(x), n_modes)."""
    x = np.atleast_1d(np.asarray(x, dtype=np.float64))
    V = np.zeros((x.size, n_modes))
    D = np.zeros((x.size, n_modes))
    for d in range(n_modes):
        c = np.zeros(d + 1)
        c[d] = math.sqrt((2 * d + 1) / 2.0)
        V[:, d] = npleg.legval(x, c)
        D[:, d] = npleg.legval(x, npleg.legder(c))
    return V, D


def side_params(kind, sigma=0.0, sigma_dirichlet=None, kappa=0.0):
    """(kind, gamma, eta, sigma, k_b, kappa) of one side (src/solver.jl:39-45)."""
    if sigma_dirichlet is None:
        sigma_dirichlet = sigma
    if kind == "wall":
        return dict(kind=kind, gam=0.0, eta=1.0, sig=0.0, kb=1.0, kap=float(kappa))
    g = 1.0 if kind == "dirichlet" else 1.0 / math.sqrt(2.0)
    return dict(kind=kind, gam=g, eta=math.sqrt(1.0 - g * g),
                sig=float(sigma_dirichlet if kind == "dirichlet" else sigma),
                kb=2.0 if kind == "absorbing_2beta" else 1.0, kap=float(kappa))


class Problem:
    """Tables and fixed data of one source-mode problem on an N x N mesh.

    sides: four side-parameter dicts (bot, top, lft, rgt) from ``side_params``;
    alpha: (N, N) or a number; source: (x0, r0, amplitude, omega, Ts) of the
    GaussianSource (src/solver.jl:80-86)."""

    def __init__(self, qu, qv, N, *, flux=(1.0, 1.0, 1.0, 1.0), sides, alpha, source):
        self.qu, self.qv, self.N = qu, qv, N
        nu, nv = qu + 1, qv + 1
        nh = max(int(math.ceil(3 * qv / 2) + 1), qu + 1)
        self.nu, self.nv, self.nq, self.nh = nu, nv, nu, nh
        self.h = 1.0 / N
        self.two_h, self.four_hh = 2.0 / self.h, 4.0 / self.h ** 2
        self.theta, self.tau, self.lu, self.lv = map(float, flux)
        z, w = npleg.leggauss(nu)
        self.z, self.w = z, w
        self.P, self.S = _onb(nu, z)                      # (nq, nu)
        El, E1l = _onb(nu, [-1.0])
        Er, E1r = _onb(nu, [1.0])
        self.El, self.Er, self.E1l, self.E1r = El[0], Er[0], E1l[0], E1r[0]
        zh, wh = npleg.leggauss(nh)
        self.Ph = _onb(nv, zh)[0]                          # (nh, nv)
        self.Wh = np.outer(wh, wh)
        self.W2 = np.outer(w, w)
        self.sides = [dict(s) for s in sides]
        self.alpha = np.broadcast_to(np.asarray(alpha, dtype=np.float64), (N, N)).copy()
        # the source projection P_h^T (W_h o g) P_h per element (src/solver.jl:656-660)
        x0, r0, amp, som, Ts = source
        self.src_omega, self.src_Ts = float(som), float(Ts)
        ii = np.arange(1, N + 1)[:, None]
        xh = ((ii - 1) * self.h + ii * self.h + self.h * zh[None, :]) / 2     # [element, point]
        X = xh[None, :, None, :]        # [j, k, p, q]: x of column k, point q
        Y = xh[:, None, :, None]        # y of row j, point p
        g = amp * np.exp(-((X - x0[0]) ** 2 + (Y - x0[1]) ** 2) / r0 ** 2)
        self.src = np.einsum("pa,jkpq,qb->jkab", self.Ph, self.Wh[None, None] * g, self.Ph)
        # the pinned stiffness with c^2 = 1 without the pin, (nu^2 x nu^2), C-order vec index a nu + b
        P, S, W = self.P, self.S, self.W2
        self.K2 = (np.einsum("pq,pa,qb,pc,qd->abcd", W, S, P, S, P)
                   + np.einsum("pq,pa,qb,pc,qd->abcd", W, P, S, P, S)).reshape(nu * nu, nu * nu)
        self.Phi_h = np.einsum("pa,qb->pqab", self.Ph, self.Ph).reshape(nh * nh, nv * nv)
        self.wgt = 0.5 * self.h * w                          # face quadrature weights of the misfit

    def srct(self, t):
        """Time factor of the source, sin(ωt) sin²(πt/Ts) for t < Ts (src/solver.jl:662)."""
        if t < self.src_Ts:
            return math.sin(self.src_omega * t) * math.sin(math.pi * t / self.src_Ts) ** 2
        return 0.0


# ---------------------------------------------------------------------------
# Traces: B/T contract the first (y) mode with the side's row, L/R the second
# (src/rhs.jl:388-410, 575-580); o[i] = sum_b P[i, b] sum_a E[a] X[a, b].

def _trace_h(X, E, P):
    return np.einsum("ib,a,jkab->jki", P[:, :X.shape[3]], E[:X.shape[2]], X)


def _trace_v(X, E, P):
    return np.einsum("ib,a,jkba->jki", P[:, :X.shape[2]], E[:X.shape[3]], X)


def _traces(pb, X, n):
    P = pb.P[:, :n]
    return dict(B=_trace_h(X, pb.El, P), T=_trace_h(X, pb.Er, P), L=_trace_v(X, pb.El, P), R=_trace_v(X, pb.Er, P))


def _dtraces(pb, X, n):
    P = pb.P[:, :n]
    return dict(B=_trace_h(X, pb.E1l, P), T=_trace_h(X, pb.E1r, P), L=_trace_v(X, pb.E1l, P),
                R=_trace_v(X, pb.E1r, P))


def _trace_h_T(o, E, P, n):
    """Adjoint of _trace_h: X[a, b] = E[a] sum_i P[i, b] o[i]."""
    return np.einsum("a,ib,jki->jkab", E[:n], P[:, :n], o)


def _trace_v_T(o, E, P, n):
    return np.einsum("a,ib,jki->jkba", E[:n], P[:, :n], o)


def _traces_T(pb, bar, n, der):
    El, Er = (pb.E1l, pb.E1r) if der else (pb.El, pb.Er)
    return (_trace_h_T(bar["B"], El, pb.P, n) + _trace_h_T(bar["T"], Er, pb.P, n)
            + _trace_v_T(bar["L"], El, pb.P, n) + _trace_v_T(bar["R"], Er, pb.P, n))


# ---------------------------------------------------------------------------

def _interior(pb, vm, vp, dum, dup, dvm, dvp, c2m, c2p, afm, afp):
    """Slots of the "+" and "-" elements of interior face points (src/rhs.jl:683-696), K = 0."""
    th, tau, two_h = pb.theta, pb.tau, pb.two_h
    J = two_h * (c2m * dum - c2p * dup + afm * dvm - afp * dvp)
    G = two_h * ((1 - th) * (c2m * dum + afm * dvm) + th * (c2p * dup + afp * dvp))
    pen = (0.5 * (c2m + c2p)) * pb.lu + (0.5 * (afm + afp)) * pb.lv
    s = vm - vp
    dvsp = th * s - tau * J
    dvsm = (1 - th) * (vp - vm) - tau * J
    return (pen * s - G, -afp * dvsp, c2p * dvsp), (G - pen * s, afm * dvsm, -c2m * dvsm)


def _interior_T(pb, vm, vp, dum, dup, dvm, dvp, c2m, c2p, afm, afp, Abp, Dbp, Fbp, Abm, Dbm, Fbm):
    """Adjoint of _interior: from the slot adjoints of both sides, the adjoints of
    the six traces and of c2m, c2p."""
    th, tau, two_h = pb.theta, pb.tau, pb.two_h
    J = two_h * (c2m * dum - c2p * dup + afm * dvm - afp * dvp)
    pen = (0.5 * (c2m + c2p)) * pb.lu + (0.5 * (afm + afp)) * pb.lv
    s = vm - vp
    dvsp = th * s - tau * J
    dvsm = (1 - th) * (vp - vm) - tau * J
    sb = (Abp - Abm) * pen
    penb = (Abp - Abm) * s
    Gb = Abm - Abp
    dvspb = -afp * Dbp + c2p * Fbp
    dvsmb = afm * Dbm - c2m * Fbm
    c2pb = dvsp * Fbp
    c2mb = -dvsm * Fbm
    sb = sb + th * dvspb - (1 - th) * dvsmb
    Jb = -tau * (dvspb + dvsmb)
    c2mb = c2mb + 0.5 * pb.lu * penb
    c2pb = c2pb + 0.5 * pb.lu * penb
    dumb = two_h * c2m * Jb + two_h * (1 - th) * c2m * Gb
    dupb = -two_h * c2p * Jb + two_h * th * c2p * Gb
    dvmb = two_h * afm * Jb + two_h * (1 - th) * afm * Gb
    dvpb = -two_h * afp * Jb + two_h * th * afp * Gb
    c2mb = c2mb + two_h * dum * Jb + two_h * (1 - th) * dum * Gb
    c2pb = c2pb - two_h * dup * Jb + two_h * th * dup * Gb
    return sb, -sb, dumb, dupb, dvmb, dvpb, c2mb, c2pb


def _boundary(pb, side, sgn, vt, du, dv, c2, af, bf):
    """Slots of a physical boundary side with zero data (src/rhs.jl:74-96, 738-763)."""
    sp = pb.sides[side]
    cf = np.sqrt(c2)
    gun = sgn * (pb.two_h * du)
    gvn = sgn * (pb.two_h * dv)
    if sp["kind"] == "wall":
        rho = c2 * gun + af * gvn
        F1 = np.zeros_like(rho)
        dvv = -(sp["kap"] / cf) * rho
    else:
        wgt = 1.0 - sp["kb"] * bf * vt
        if np.any(np.real(wgt) <= 0.0):
            raise RuntimeError("boundary weight <= 0 at a face point")
        sq = np.sqrt(wgt)
        rho = sp["eta"] * (c2 * gun + af * gvn) + sp["gam"] * cf * sq * vt
        F1 = c2 * gun + af * gvn - (sp["eta"] + sp["sig"] * sp["gam"]) * rho
        dvv = -(sp["gam"] - sp["sig"] * sp["eta"]) / (cf * sq) * rho
    return F1, sgn * (af * dvv), -sgn * (c2 * dvv)


def _boundary_T(pb, side, sgn, vt, du, dv, c2, af, bf, Ab, Db, Fb):
    """Adjoint of _boundary: adjoints of (vt, du, dv, c2, bf)."""
    sp = pb.sides[side]
    cf = np.sqrt(c2)
    gun = sgn * (pb.two_h * du)
    gvn = sgn * (pb.two_h * dv)
    F1b = Ab
    dvvb = sgn * af * Db - sgn * c2 * Fb
    c2b = np.zeros_like(vt) if np.isrealobj(vt) else np.zeros(vt.shape, dtype=complex)
    vtb = np.zeros_like(c2b)
    bfb = np.zeros_like(c2b)
    cfb = np.zeros_like(c2b)
    if sp["kind"] == "wall":
        rho = c2 * gun + af * gvn
        dvv = -(sp["kap"] / cf) * rho
        c2b = c2b - sgn * dvv * Fb
        rhob = -(sp["kap"] / cf) * dvvb
        cfb = cfb + (sp["kap"] / cf ** 2) * rho * dvvb
        gunb = c2 * rhob
        gvnb = af * rhob
        c2b = c2b + gun * rhob
    else:
        eta, gam, sig, kb = sp["eta"], sp["gam"], sp["sig"], sp["kb"]
        sq = np.sqrt(1.0 - kb * bf * vt)
        rho = eta * (c2 * gun + af * gvn) + gam * cf * sq * vt
        kp = gam - sig * eta
        dvv = -kp / (cf * sq) * rho
        c2b = c2b - sgn * dvv * Fb
        rhob = -(eta + sig * gam) * F1b - kp / (cf * sq) * dvvb
        cfb = cfb + kp * rho / (cf ** 2 * sq) * dvvb
        sqb = kp * rho / (cf * sq ** 2) * dvvb
        c2b = c2b + gun * F1b + eta * gun * rhob
        gunb = c2 * F1b + eta * c2 * rhob
        gvnb = af * F1b + eta * af * rhob
        cfb = cfb + gam * sq * vt * rhob
        sqb = sqb + gam * cf * vt * rhob
        vtb = vtb + gam * cf * sq * rhob
        vtb = vtb - kb * bf / (2 * sq) * sqb
        bfb = bfb - kb * vt / (2 * sq) * sqb
    c2b = c2b + cfb / (2 * cf)
    return vtb, sgn * pb.two_h * gunb, sgn * pb.two_h * gvnb, c2b, bfb


# side tables of the lifts: trace row of v (A), derivative row of v (D), derivative row of u (F)
def _side_rows(pb, s):
    return {"B": (pb.El, pb.E1l), "T": (pb.Er, pb.E1r), "L": (pb.El, pb.E1l), "R": (pb.Er, pb.E1r)}[s]


def _faces(pb, c2, uu, vv, dt):
    """Traces and the slots (A, D, F) of every side of every element."""
    N, nu, nv = pb.N, pb.nu, pb.nv
    af = pb.alpha
    du = _dtraces(pb, uu, nu)
    vt = _traces(pb, vv, nv)
    dv = _dtraces(pb, vv, nv)
    slots = {s: [np.zeros((N, N, pb.nq), dtype=dt) for _ in range(3)] for s in "BTLR"}
    C2 = c2[:, :, None]
    AF = af[:, :, None]
    if N > 1:
        # horizontal faces: "-" = (j-1, k) (its T side), "+" = (j, k) (its B side)
        p, m = _interior(pb, vt["T"][:-1], vt["B"][1:], du["T"][:-1], du["B"][1:], dv["T"][:-1], dv["B"][1:],
                         C2[:-1], C2[1:], AF[:-1], AF[1:])
        for q in range(3):
            slots["B"][q][1:] = p[q]
            slots["T"][q][:-1] = m[q]
        # vertical faces: "-" = (j, k-1) (its R side), "+" = (j, k) (its L side)
        p, m = _interior(pb, vt["R"][:, :-1], vt["L"][:, 1:], du["R"][:, :-1], du["L"][:, 1:], dv["R"][:, :-1],
                         dv["L"][:, 1:], C2[:, :-1], C2[:, 1:], AF[:, :-1], AF[:, 1:])
        for q in range(3):
            slots["L"][q][:, 1:] = p[q]
            slots["R"][q][:, :-1] = m[q]
    return du, vt, dv, slots


def _boundary_slots(pb, c2, beta, du, vt, dv, slots):
    af = pb.alpha
    for side, (s, sgn, sel) in enumerate((("B", -1.0, np.s_[0, :]), ("T", 1.0, np.s_[-1, :]),
                                          ("L", -1.0, np.s_[:, 0]), ("R", 1.0, np.s_[:, -1]))):
        r = _boundary(pb, side, sgn, vt[s][sel], du[s][sel], dv[s][sel], c2[sel][:, None], af[sel][:, None],
                      beta[sel][:, None])
        for q in range(3):
            slots[s][q][sel] = r[q]


def _lifts(pb, slots, dt):
    """Rv (v-test, lifts of A and D) and Fu (u-test, lifts of F) (src/rhs.jl:421-473)."""
    N, nu, nv = pb.N, pb.nu, pb.nv
    Rv = np.zeros((N, N, nv, nv), dtype=dt)
    Fu = np.zeros((N, N, nu, nu), dtype=dt)
    Pv, Pu, w = pb.P[:, :nv], pb.P, pb.w
    for s in "BTLR":
        A, D, F = slots[s]
        YA = np.einsum("ib,i,jki->jkb", Pv, w, A)
        YD = np.einsum("ib,i,jki->jkb", Pv, w, D)
        YF = np.einsum("ib,i,jki->jkb", Pu, w, F)
        E, E1 = _side_rows(pb, s)
        if s in "BT":
            Rv += np.einsum("a,jkb->jkab", pb.two_h * E[:nv], YA) + np.einsum("a,jkb->jkab", pb.four_hh * E1[:nv], YD)
            Fu += np.einsum("a,jkb->jkab", E1, YF)
        else:
            Rv += np.einsum("jka,b->jkab", YA, pb.two_h * E[:nv]) + np.einsum("jka,b->jkab", YD, pb.four_hh * E1[:nv])
            Fu += np.einsum("jka,b->jkab", YF, E1)
    return Rv, Fu


def _grads(pb, X, n):
    """(d/dy, d/dx) of the modal X (n x n) at the standard points, [j, k, p, q]."""
    P, S = pb.P[:, :n], pb.S[:, :n]
    return np.einsum("pa,jkab,qb->jkpq", S, X, P), np.einsum("pa,jkab,qb->jkpq", P, X, S)


def _grads_T(pb, Hy, Hx, n):
    """Adjoint of _grads: S^T Hy P + P^T Hx S, (n x n)."""
    P, S = pb.P[:, :n], pb.S[:, :n]
    return np.einsum("pa,jkpq,qb->jkab", S, Hy, P) + np.einsum("pa,jkpq,qb->jkab", P, Hx, S)


def _solve(A, b):
    """Batched dense solve over the leading (N, N) axes."""
    return np.linalg.solve(A, b[..., None])[..., 0]


def _forward_parts(pb, c2, beta, u, v, t):
    """One evaluation with every intermediate the VJP needs."""
    N, nu, nv, nh = pb.N, pb.nu, pb.nv, pb.nh
    dt = np.result_type(u, v, c2, beta)
    du, vt, dv, slots = _faces(pb, c2, u, v, dt)
    _boundary_slots(pb, c2, beta, du, vt, dv, slots)
    Rlift, Fu = _lifts(pb, slots, dt)
    W = pb.W2[None, None]
    C2 = c2[:, :, None, None]
    AF = pb.alpha[:, :, None, None]
    guy, gux = _grads(pb, u, nu)
    gvy, gvx = _grads(pb, v, nv)
    Rvol = _grads_T(pb, C2 * W * guy + AF * W * gvy, C2 * W * gux + AF * W * gvx, nv)
    VB = _grads_T(pb, C2 * W * gvy, C2 * W * gvx, nu)
    Rv = -pb.four_hh * Rvol + pb.srct(t) * pb.src + Rlift
    b = VB - Fu
    b[:, :, 0, 0] = b[:, :, 0, 0] + v[:, :, 0, 0]
    e0 = np.zeros(nu * nu)
    e0[0] = 1.0
    A = c2[:, :, None, None] * pb.K2[None, None] + np.outer(e0, e0)[None, None]
    ut = _solve(A, b.reshape(N, N, nu * nu)).reshape(N, N, nu, nu)
    vq = np.einsum("pa,jkab,qb->jkpq", pb.Ph, v, pb.Ph)
    wq = (1.0 - 2.0 * beta[:, :, None, None] * vq) * pb.Wh[None, None]
    if np.any(np.real(1.0 - 2.0 * beta[:, :, None, None] * vq) <= 0.0):
        raise RuntimeError("mass matrix not positive definite")
    M = np.einsum("xa,jkx,xb->jkab", pb.Phi_h, wq.reshape(N, N, nh * nh), pb.Phi_h)
    X = _solve(M, Rv.reshape(N, N, nv * nv)).reshape(N, N, nv, nv)
    return SimpleNamespace(du=du, vt=vt, dv=dv, slots=slots, ut=ut, vt_=X, A=A, M=M, vq=vq,
                           guy=guy, gux=gux, gvy=gvy, gvx=gvx)


def rhs(pb, c2, beta, u, v, t):
    """(u_t, v_t) of one evaluation at time t; c2, beta: (N, N) per element."""
    f = _forward_parts(pb, np.asarray(c2), np.asarray(beta), np.asarray(u), np.asarray(v), t)
    return f.ut, f.vt_


def rhs_vjp(pb, c2, beta, u, v, t, ub, vb):
    """Vector-Jacobian product of ``rhs``: given the cotangents (ub, vb) of
    (u_t, v_t), the cotangents of (u, v, c2, beta)."""
    N, nu, nv = pb.N, pb.nu, pb.nv
    f = _forward_parts(pb, c2, beta, u, v, t)
    W = pb.W2[None, None]
    C2 = c2[:, :, None, None]
    AF = pb.alpha[:, :, None, None]
    # mass system: X = M^-1 Rv, M symmetric; dM = Phi^T diag(-2 (dbeta v_q + beta dv_q) W_h) Phi
    lam = _solve(f.M, vb.reshape(N, N, nv * nv)).reshape(N, N, nv, nv)
    Lq = np.einsum("pa,jkab,qb->jkpq", pb.Ph, lam, pb.Ph)
    Xq = np.einsum("pa,jkab,qb->jkpq", pb.Ph, f.vt_, pb.Ph)
    betab = np.einsum("jkpq->jk", 2.0 * pb.Wh[None, None] * f.vq * Lq * Xq)
    vbar_q = 2.0 * pb.Wh[None, None] * beta[:, :, None, None] * Lq * Xq
    Vb = np.einsum("pa,jkpq,qb->jkab", pb.Ph, vbar_q, pb.Ph)
    Rb = lam
    # u-equation: ut = A^-1 b, A = c^2 K2 + e0 e0^T
    bb = _solve(f.A, ub.reshape(N, N, nu * nu))
    c2b = -np.einsum("jkx,xy,jky->jk", bb, pb.K2, f.ut.reshape(N, N, nu * nu))
    bb = bb.reshape(N, N, nu, nu)
    VBb = bb
    Fub = -bb
    Vb[:, :, 0, 0] += bb[:, :, 0, 0]
    # volume terms
    Rvolb = -pb.four_hh * Rb
    Hyb, Hxb = _grads(pb, Rvolb, nv)                  # adjoints of the v-test point fields
    Ub = _grads_T(pb, C2 * W * Hyb, C2 * W * Hxb, nu)
    c2b = c2b + np.einsum("jkpq->jk", W * (Hyb * f.guy + Hxb * f.gux))
    gvyb = AF * W * Hyb
    gvxb = AF * W * Hxb
    Gyb, Gxb = _grads(pb, VBb, nu)
    gvyb = gvyb + C2 * W * Gyb
    gvxb = gvxb + C2 * W * Gxb
    c2b = c2b + np.einsum("jkpq->jk", W * (Gyb * f.gvy + Gxb * f.gvx))
    Vb = Vb + _grads_T(pb, gvyb, gvxb, nv)
    # lifts -> slot adjoints
    Pv, Pu, w = pb.P[:, :nv], pb.P, pb.w
    sb = {}
    for s in "BTLR":
        E, E1 = _side_rows(pb, s)
        if s in "BT":
            YAb = np.einsum("a,jkab->jkb", pb.two_h * E[:nv], Rb)
            YDb = np.einsum("a,jkab->jkb", pb.four_hh * E1[:nv], Rb)
            YFb = np.einsum("a,jkab->jkb", E1, Fub)
        else:
            YAb = np.einsum("jkab,b->jka", Rb, pb.two_h * E[:nv])
            YDb = np.einsum("jkab,b->jka", Rb, pb.four_hh * E1[:nv])
            YFb = np.einsum("jkab,b->jka", Fub, E1)
        sb[s] = (np.einsum("i,ib,jkb->jki", w, Pv, YAb), np.einsum("i,ib,jkb->jki", w, Pv, YDb),
                 np.einsum("i,ib,jkb->jki", w, Pu, YFb))
    # slots -> traces and coefficients
    z = lambda: np.zeros((N, N, pb.nq))
    dub = {s: z() for s in "BTLR"}
    vtb = {s: z() for s in "BTLR"}
    dvb = {s: z() for s in "BTLR"}
    du, vt, dv = f.du, f.vt, f.dv
    C2f = c2[:, :, None]
    AFf = pb.alpha[:, :, None]
    if N > 1:
        r = _interior_T(pb, vt["T"][:-1], vt["B"][1:], du["T"][:-1], du["B"][1:], dv["T"][:-1], dv["B"][1:],
                        C2f[:-1], C2f[1:], AFf[:-1], AFf[1:], *(x[1:] for x in sb["B"]), *(x[:-1] for x in sb["T"]))
        vtb["T"][:-1] += r[0]; vtb["B"][1:] += r[1]; dub["T"][:-1] += r[2]; dub["B"][1:] += r[3]
        dvb["T"][:-1] += r[4]; dvb["B"][1:] += r[5]
        c2b[:-1] += r[6].sum(axis=-1); c2b[1:] += r[7].sum(axis=-1)
        r = _interior_T(pb, vt["R"][:, :-1], vt["L"][:, 1:], du["R"][:, :-1], du["L"][:, 1:], dv["R"][:, :-1],
                        dv["L"][:, 1:], C2f[:, :-1], C2f[:, 1:], AFf[:, :-1], AFf[:, 1:],
                        *(x[:, 1:] for x in sb["L"]), *(x[:, :-1] for x in sb["R"]))
        vtb["R"][:, :-1] += r[0]; vtb["L"][:, 1:] += r[1]; dub["R"][:, :-1] += r[2]; dub["L"][:, 1:] += r[3]
        dvb["R"][:, :-1] += r[4]; dvb["L"][:, 1:] += r[5]
        c2b[:, :-1] += r[6].sum(axis=-1); c2b[:, 1:] += r[7].sum(axis=-1)
    for side, (s, sgn, sel) in enumerate((("B", -1.0, np.s_[0, :]), ("T", 1.0, np.s_[-1, :]),
                                          ("L", -1.0, np.s_[:, 0]), ("R", 1.0, np.s_[:, -1]))):
        r = _boundary_T(pb, side, sgn, vt[s][sel], du[s][sel], dv[s][sel], c2[sel][:, None],
                        pb.alpha[sel][:, None], beta[sel][:, None], *(x[sel] for x in sb[s]))
        vtb[s][sel] += r[0]; dub[s][sel] += r[1]; dvb[s][sel] += r[2]
        c2b[sel] += r[3].sum(axis=-1); betab[sel] += r[4].sum(axis=-1)
    Ub = Ub + _traces_T(pb, dub, nu, True)
    Vb = Vb + _traces_T(pb, vtb, nv, False) + _traces_T(pb, dvb, nv, True)
    return Ub, Vb, c2b, betab


# ---------------------------------------------------------------------------
# RK4 trajectories and the misfit

def bottom_trace(pb, u):
    """u_h at the bottom side's face points of the elements of row j = 0, [k, i]."""
    return np.einsum("ib,a,kab->ki", pb.P, pb.El, u[0])


def _rk4(pb, c2, beta, u, v, t, dt, keep=False):
    k1 = rhs(pb, c2, beta, u, v, t)
    Y2 = (u + 0.5 * dt * k1[0], v + 0.5 * dt * k1[1])
    k2 = rhs(pb, c2, beta, *Y2, t + 0.5 * dt)
    Y3 = (u + 0.5 * dt * k2[0], v + 0.5 * dt * k2[1])
    k3 = rhs(pb, c2, beta, *Y3, t + 0.5 * dt)
    Y4 = (u + dt * k3[0], v + dt * k3[1])
    k4 = rhs(pb, c2, beta, *Y4, t + dt)
    c = dt / 6.0
    un = u + c * (k1[0] + 2.0 * k2[0] + 2.0 * k3[0] + k4[0])
    vn = v + c * (k1[1] + 2.0 * k2[1] + 2.0 * k3[1] + k4[1])
    return (un, vn, [(u, v), Y2, Y3, Y4]) if keep else (un, vn)


def misfit(pb, c2, beta, dt, nsteps, target=None, traces=False):
    """J over nsteps RK4 steps from the zero state (t += dt after every step);
    with traces=True also the bottom traces after every step [n, k, i]."""
    N = pb.N
    dtp = np.result_type(np.asarray(c2), np.asarray(beta))
    u = np.zeros((N, N, pb.nu, pb.nu), dtype=dtp)
    v = np.zeros((N, N, pb.nv, pb.nv), dtype=dtp)
    J = 0.0
    tr = []
    t = 0.0
    for n in range(nsteps):
        u, v = _rk4(pb, c2, beta, u, v, t, dt)
        t += dt
        b = bottom_trace(pb, u)
        if traces:
            tr.append(b)
        if target is not None:
            r = b - target[n]
            J = J + 0.5 * np.sum(pb.wgt[None, :] * r * r)
    return (J, np.array(tr)) if traces else J


def gradient(pb, c2, beta, dt, nsteps, target):
    """J and (dJ/dc2, dJ/dbeta) by the discrete adjoint of RK4 with every stage state kept."""
    N = pb.N
    u = np.zeros((N, N, pb.nu, pb.nu))
    v = np.zeros((N, N, pb.nv, pb.nv))
    steps = []
    t = 0.0
    J = 0.0
    for n in range(nsteps):
        un, vn, Y = _rk4(pb, c2, beta, u, v, t, dt, keep=True)
        steps.append((t, Y))
        u, v = un, vn
        t += dt
        r = bottom_trace(pb, u) - target[n]
        J += 0.5 * np.sum(pb.wgt[None, :] * r * r)
    gc2 = np.zeros((N, N))
    gb = np.zeros((N, N))
    ub = np.zeros_like(u)
    vb = np.zeros_like(v)
    for n in range(nsteps - 1, -1, -1):
        # dJ/dy_{n+1}: the misfit of the state after step n
        yn = u if n == nsteps - 1 else steps[n + 1][1][0][0]
        r = bottom_trace(pb, yn) - target[n]
        ub[0] += np.einsum("a,ib,ki->kab", pb.El, pb.P, pb.wgt[None, :] * r)
        t0, Y = steps[n]
        ts = (t0, t0 + 0.5 * dt, t0 + 0.5 * dt, t0 + dt)
        # stage cotangents, last stage first: k4 = (dt/6) y_bar, k3 = (dt/3) y_bar + dt Y4_bar, ...
        acc_u, acc_v = ub.copy(), vb.copy()
        kbu, kbv = (dt / 6.0) * ub, (dt / 6.0) * vb
        nxt = ((dt / 3.0, dt), (dt / 3.0, 0.5 * dt), (dt / 6.0, 0.5 * dt))
        for s in (3, 2, 1, 0):
            Yu, Yv, c2s, bs = rhs_vjp(pb, c2, beta, Y[s][0], Y[s][1], ts[s], kbu, kbv)
            acc_u += Yu
            acc_v += Yv
            gc2 += c2s
            gb += bs
            if s > 0:
                ca, cb = nxt[3 - s]
                kbu, kbv = ca * ub + cb * Yu, ca * vb + cb * Yv
        ub, vb = acc_u, acc_v
    return J, gc2, gb
