"""Host-side table producer: ``precompute(q, N; ...) -> (pre, geo)``.

Restates src/solver.jl:1-86 (flux and boundary parameters, source model) and
src/solver.jl:341-770 (``_side_values!``, ``precompute``, ``_rhs_data``) of the
reference. Every array keeps the reference's name, shape and column-major
(Fortran) layout, so ``pre`` is the same set of tables a Julia caller would hand
to the device library; index (j, k, a, b) here is (j+1, k+1, a+1, b+1) there.
The device library (``_lib.WedgProblem``) only reads these tables; nothing here
runs inside a time step.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass
from types import SimpleNamespace
from typing import Callable, Optional

import numpy as np

from .coefficients import Coefficients, evaluate, paper_coefficients
from .quadrature import GaussQCofs, GradJacobiP, JacobiP

__all__ = ["flux_parameters", "BOUNDARY_KINDS", "GaussianSource", "precompute", "Pre",
           "boundary_side", "boundary_sides"]

PI = math.pi
FACE_SHIFT = 2.0 ** -40          # src/solver.jl:361
JUMP_TOL = 1e-8                  # src/solver.jl:362
MASS_TOL = 1e-14                 # src/solver.jl:767

BOUNDARY_KINDS = ("absorbing_2beta", "absorbing", "dirichlet", "wall")   # src/solver.jl:37
SIDE_NAMES = ("bot", "top", "lft", "rgt")


class Pre(SimpleNamespace):
    """Attribute bag with the field names of the reference's ``pre`` NamedTuple."""

    def merge(self, **kw) -> "Pre":
        d = dict(self.__dict__)
        d.update(kw)
        return Pre(**d)


def F(shape) -> np.ndarray:
    return np.zeros(shape, dtype=np.float64, order="F")


def flux_parameters(flux: str):
    """(theta, tau, lambda_u, lambda_v) of the flux families (src/solver.jl:16-21)."""
    flux = str(flux).lstrip(":")
    if flux == "upwind":
        return dict(theta=1.0, tau=1.0, lambda_u=1.0, lambda_v=1.0)
    if flux == "alternating":
        return dict(theta=1.0, tau=0.0, lambda_u=0.0, lambda_v=0.0)
    if flux == "central":
        return dict(theta=0.5, tau=0.0, lambda_u=0.0, lambda_v=0.0)
    raise ValueError(f"unknown flux {flux} (expected :upwind, :alternating or :central)")


def boundary_side(kind: str, sigma: float, sigma_dirichlet: float, kappa: float) -> dict:
    """Parameters (gamma, eta, sigma, k_b, kappa) of one side (src/solver.jl:39-45)."""
    if kind == "wall":
        return dict(kind=kind, gamma_abc=0.0, eta_abc=1.0, sigma=0.0, kb=1.0, kappa=kappa)
    gamma_abc = 1.0 if kind == "dirichlet" else 1.0 / math.sqrt(2.0)
    return dict(kind=kind, gamma_abc=gamma_abc, eta_abc=math.sqrt(1.0 - gamma_abc ** 2),
                sigma=sigma_dirichlet if kind == "dirichlet" else sigma,
                kb=2.0 if kind == "absorbing_2beta" else 1.0, kappa=kappa)


def boundary_sides(boundary, sigma: float, sigma_dirichlet: float, kappa: float) -> dict:
    """Per-side parameters from a kind or a mapping bot/top/lft/rgt, with the
    argument checks of src/solver.jl:50-68 (same messages)."""
    if isinstance(boundary, str):
        kinds = (boundary.lstrip(":"),) * 4
    else:
        if len(boundary) != 4 or set(boundary.keys()) != set(SIDE_NAMES):
            raise ValueError("boundary must be a kind or a NamedTuple with the keys bot, top, lft, rgt")
        kinds = tuple(str(boundary[s]).lstrip(":") for s in SIDE_NAMES)
    for kind in kinds:
        if kind not in BOUNDARY_KINDS:
            raise ValueError(f"unknown boundary {kind} (expected one of {BOUNDARY_KINDS})")
    if any(k in ("absorbing", "absorbing_2beta") for k in kinds) and not (-1.0 <= sigma <= 1.0):
        raise ValueError(f"absorbing sides need -1 ≤ sigma ≤ 1 (got {sigma})")
    if "dirichlet" in kinds and not (sigma_dirichlet >= 0.0):
        raise ValueError(f"Dirichlet sides need sigma_dirichlet ≥ 0 (got {sigma_dirichlet}; "
                         "it defaults to sigma)")
    if not (kappa >= 0.0):
        raise ValueError(f"kappa must be ≥ 0 (got {kappa})")
    return {s: boundary_side(k, sigma, sigma_dirichlet, kappa) for s, k in zip(SIDE_NAMES, kinds)}


@dataclass(frozen=True)
class GaussianSource:
    """f(t, x) = amplitude exp(-|x - x0|^2 / r0^2) sin(omega t) sin^2(pi t / Ts) for
    t < Ts, 0 after (src/solver.jl:80-86)."""
    x0: tuple = (0.15, 0.5)
    r0: float = 0.05
    amplitude: float = 1.0
    omega: float = 16 * PI
    Ts: float = 1.0

    def time_factor(self, t: float) -> float:
        """src_time(t) of src/solver.jl:662."""
        if t < self.Ts:
            return math.sin(self.omega * t) * math.sin(PI * t / self.Ts) ** 2
        return 0.0


def _side_values(bot, top, lft, rgt, f, xloc, yloc, N, h):
    """One-sided values of ``f`` on the sides of every element (src/solver.jl:371-408).

    bot[j, k, i] is the value at point i of the bottom side of element (j, k) seen
    from inside it; a face point takes f(x_f) on both sides unless the values
    shifted by 2^-40 into either element disagree with it beyond 1e-8 (relative,
    plus 1e-8 times the largest finite shifted value).
    """
    nq = xloc.shape[1]
    d = FACE_SHIFT
    hz = bot is not None
    vt = lft is not None
    if hz:
        jf = np.arange(N + 1, dtype=np.float64)
        Xh = np.broadcast_to(xloc.T[:, None, :], (nq, N + 1, N))          # xloc[k, i] at [i, jf, k]
        Yh = np.broadcast_to((jf * h)[None, :, None], (nq, N + 1, N))
        H0 = evaluate(f, Xh, Yh)
        HM = H0.copy()
        HP = H0.copy()
        HM[:, 1:, :] = evaluate(f, Xh[:, 1:, :], Yh[:, 1:, :] - d)
        HP[:, :N, :] = evaluate(f, Xh[:, :N, :], Yh[:, :N, :] + d)
    if vt:
        kf = np.arange(N + 1, dtype=np.float64)
        Xv = np.broadcast_to((kf * h)[None, None, :], (nq, N, N + 1))
        Yv = np.broadcast_to(yloc.T[:, :, None], (nq, N, N + 1))           # yloc[j, i] at [i, j, kf]
        V0 = evaluate(f, Xv, Yv)
        VM = V0.copy()
        VP = V0.copy()
        VM[:, :, 1:] = evaluate(f, Xv[:, :, 1:] - d, Yv[:, :, 1:])
        VP[:, :, :N] = evaluate(f, Xv[:, :, :N] + d, Yv[:, :, :N])

    def amax(A):
        A = np.abs(A[np.isfinite(A)])
        return float(A.max()) if A.size else 0.0

    sh = max(amax(HM[:, 1:, :]), amax(HP[:, :N, :])) if hz else 0.0
    sv = max(amax(VM[:, :, 1:]), amax(VP[:, :, :N])) if vt else 0.0
    s = JUMP_TOL * max(sh, sv)

    def sides(f0, fm, fp):
        with np.errstate(invalid="ignore", over="ignore"):
            def near(a, b):
                return np.abs(a - b) <= JUMP_TOL * np.maximum(np.abs(a), np.abs(b)) + s
            keep = np.isfinite(f0) & near(fm, f0) & near(fp, f0)
        return np.where(keep, f0, fm), np.where(keep, f0, fp)

    if hz:
        fm, fp = sides(H0, HM, HP)
        top[:, :, :] = np.transpose(fm[:, 1:, :], (1, 2, 0))              # top[jf-1, k, i] = fm
        bot[:, :, :] = np.transpose(fp[:, :N, :], (1, 2, 0))              # bot[jf, k, i] = fp
    if vt:
        fm, fp = sides(V0, VM, VP)
        rgt[:, :, :] = np.transpose(fm[:, :, 1:], (1, 2, 0))              # rgt[j, kf-1, i] = fm
        lft[:, :, :] = np.transpose(fp[:, :, :N], (1, 2, 0))              # lft[j, kf, i] = fp


def _tensor_points(loc_row, loc_col, n):
    """(Xq, Yq) on one element: rows vary in y, columns in x (src/solver.jl:519-520)."""
    return (np.broadcast_to(loc_col[None, :], (n, n)), np.broadcast_to(loc_row[:, None], (n, n)))


def precompute(q: int, N: int, *, mixed: bool = True, flux: str = "upwind", flux_params=None,
               sigma: float = 0.0, boundary="absorbing_2beta", sigma_dirichlet: Optional[float] = None,
               kappa: float = 0.0, phase=(0.0, 0.0), coefficients: Optional[Coefficients] = None,
               mode: str = "mms", kx: int = 1, ky: int = 1, source: Optional[GaussianSource] = None,
               mass_solver: str = "chebyshev", initial_profile=None):
    """All element operators, face coefficients and data arrays of one problem
    (src/solver.jl:428-694 followed by ``_rhs_data``, src/solver.jl:716-770).

    Returns ``(pre, geo)`` with the reference's field names. Raises the
    reference's errors for bad arguments and for jumps of c or alpha in :mms mode.
    """
    mode = str(mode).lstrip(":")
    mass_solver = str(mass_solver).lstrip(":")
    if sigma_dirichlet is None:
        sigma_dirichlet = sigma
    if coefficients is None:
        coefficients = paper_coefficients()
    if source is None:
        source = GaussianSource()
    if mode not in ("mms", "initial_value", "source"):
        raise ValueError(f"unknown mode {mode} (expected :mms, :initial_value or :source)")
    if mass_solver not in ("cg", "chebyshev"):
        raise ValueError(f"unknown mass_solver {mass_solver} (expected :cg or :chebyshev)")
    if initial_profile is not None and mode != "initial_value":
        raise ValueError("initial_profile is only available in :initial_value mode")
    sigma = float(sigma)
    sigma_dirichlet = float(sigma_dirichlet)
    kappa = float(kappa)
    q_u = int(q)
    q_v = q_u - 1 if mixed else q_u
    qu1 = q_u + 1
    qv1 = q_v + 1
    h = 1.0 / N
    omega = 2 * PI * math.sqrt(kx ** 2 + ky ** 2)

    if flux_params is None:
        fp = flux_parameters(flux)
    else:
        g = (lambda k: flux_params[k]) if isinstance(flux_params, dict) else (lambda k: getattr(flux_params, k))
        fp = dict(theta=float(g("theta")), tau=float(g("tau")),
                  lambda_u=float(g("lambda_u")), lambda_v=float(g("lambda_v")))
    bc = boundary_sides(boundary, sigma, sigma_dirichlet, kappa)
    legacy = isinstance(boundary, str) and boundary.lstrip(":") in ("absorbing", "dirichlet")
    gamma_abc = bc["bot"]["gamma_abc"] if legacy else math.nan
    px, py = float(phase[0]), float(phase[1])

    cC, cxC, cyC = coefficients.c, coefficients.c_x, coefficients.c_y
    aC, axC, ayC = coefficients.alpha, coefficients.alpha_x, coefficients.alpha_y
    bC = coefficients.beta

    # ---- 1D bases at the (q_u+1)-point Gauss rule (src/solver.jl:462-480) ----
    z, w = GaussQCofs(qu1)
    W2d = np.asfortranarray(np.outer(w, w))
    P_u = F((qu1, qu1)); S_u = F((qu1, qu1))
    Ebl_u = F((1, qu1)); Ebr_u = F((1, qu1)); E1bl_u = F((1, qu1)); E1br_u = F((1, qu1))
    for d in range(qu1):
        P_u[:, d] = JacobiP(z, 0, 0, d)
        S_u[:, d] = GradJacobiP(z, 0, 0, d)
        Ebl_u[0, d] = JacobiP(-1.0, 0, 0, d)
        Ebr_u[0, d] = JacobiP(1.0, 0, 0, d)
        E1bl_u[0, d] = GradJacobiP(-1.0, 0, 0, d)
        E1br_u[0, d] = GradJacobiP(1.0, 0, 0, d)
    P_v = np.asfortranarray(P_u[:, :qv1])
    Ebl_v = np.asfortranarray(Ebl_u[:, :qv1]); Ebr_v = np.asfortranarray(Ebr_u[:, :qv1])
    E1bl_v = np.asfortranarray(E1bl_u[:, :qv1]); E1br_v = np.asfortranarray(E1br_u[:, :qv1])

    # ---- high-order rule (src/solver.jl:482-490) ----
    q_high = max(int(math.ceil(3 * q_v / 2) + 1), q_u + 1)
    z_h, w_h = GaussQCofs(q_high)
    W2d_h = np.asfortranarray(np.outer(w_h, w_h))
    P_h_v = F((q_high, qv1))
    for d in range(qv1):
        P_h_v[:, d] = JacobiP(z_h, 0, 0, d)
    Phi2d_h_v = np.asfortranarray(np.kron(P_h_v, P_h_v))
    Phi_xi_u = np.kron(P_u, S_u)
    Phi_eta_u = np.kron(S_u, P_u)
    S_v = np.asfortranarray(S_u[:, :qv1])

    # ---- quadrature coordinates per element row / column (src/solver.jl:497-507) ----
    ii = np.arange(1, N + 1, dtype=np.float64)[:, None]
    xloc = ((ii - 1) * h + ii * h + h * z[None, :]) / 2
    yloc = xloc.copy()
    xloc_h = ((ii - 1) * h + ii * h + h * z_h[None, :]) / 2
    yloc_h = xloc_h.copy()

    # all standard points of all elements: [j, k, iy, ix]
    Xq = np.broadcast_to(xloc[None, :, None, :], (N, N, qu1, qu1))
    Yq = np.broadcast_to(yloc[:, None, :, None], (N, N, qu1, qu1))
    Xh = np.broadcast_to(xloc_h[None, :, None, :], (N, N, q_high, q_high))
    Yh = np.broadcast_to(yloc_h[:, None, :, None], (N, N, q_high, q_high))

    # ---- c^2 W, alpha W and the packed inverse of the pinned u-stiffness
    # (src/solver.jl:509-536) ----
    cq = evaluate(cC, Xq, Yq)
    aq = evaluate(aC, Xq, Yq)
    cw4 = cq ** 2 * W2d[None, None, :, :]               # [j, k, iy, ix]
    aw4 = aq * W2d[None, None, :, :]
    E = N * N
    # element e = j + N k; Cw[iy, e, ix]
    Cw = _elem_mid(cw4, N)
    Aw = _elem_mid(aw4, N)
    nu2 = qu1 * qu1
    npk = (nu2 * (nu2 + 1)) // 2
    Ainv_u = F((npk, E))
    cwv = np.transpose(cw4, (1, 0, 3, 2)).reshape(E, nu2)   # row e, vec index a + n*b (a = iy)
    uniq, inv_idx = np.unique(cwv, axis=0, return_inverse=True)
    inv_idx = np.asarray(inv_idx).reshape(-1)
    iu, ju = np.triu_indices(nu2)                           # packed order: column j, rows i <= j
    order = np.lexsort((iu, ju))
    iu, ju = iu[order], ju[order]
    packs = np.empty((uniq.shape[0], npk))

    def classes(c0, c1):
        # the classes' matrices, factors and inverses as stacks (LAPACK per
        # matrix, as the loop over the elements calls it)
        cw_c = uniq[c0:c1]
        Mj = (np.matmul(Phi_xi_u.T[None, :, :], cw_c[:, :, None] * Phi_xi_u[None, :, :])
              + np.matmul(Phi_eta_u.T[None, :, :], cw_c[:, :, None] * Phi_eta_u[None, :, :]))
        Mj[:, 0, 0] += 1.0
        Linv = np.linalg.inv(np.linalg.cholesky(Mj))
        packs[c0:c1] = np.matmul(np.swapaxes(Linv, 1, 2), Linv)[:, iu, ju]

    nc = uniq.shape[0]
    chunk = 1024
    if nc <= chunk:
        classes(0, nc)
    else:   # numpy's linalg releases the GIL: chunks on the host cores
        import concurrent.futures as cf
        with cf.ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
            list(ex.map(lambda c0: classes(c0, min(c0 + chunk, nc)), range(0, nc, chunk)))
    Ainv_u[:, :] = packs[inv_idx].T

    # ---- face coefficients seen from inside every element (src/solver.jl:538-552) ----
    faces = {s: {k: F((N, N, qu1)) for k in ("c2", "af", "cf", "bf")} for s in SIDE_NAMES}
    for fun, key in ((cC, "cf"), (aC, "af"), (bC, "bf")):
        _side_values(faces["bot"][key], faces["top"][key], faces["lft"][key], faces["rgt"][key],
                     fun, xloc, yloc, N, h)
    for s in SIDE_NAMES:
        faces[s]["c2"][...] = faces[s]["cf"] ** 2

    # ---- d_e u0 one-sided (src/solver.jl:554-568) ----
    g0 = {s: F((N, N, qu1)) for s in SIDE_NAMES}
    if mode != "source":
        if initial_profile is None:
            def u0_x(x, y):
                return -2 * PI * kx * np.sin(2 * PI * kx * x + px) * np.sin(2 * PI * ky * y + py)

            def u0_y(x, y):
                return 2 * PI * ky * np.cos(2 * PI * kx * x + px) * np.cos(2 * PI * ky * y + py)
        else:
            u0_x = _prof(initial_profile, "u0_x")
            u0_y = _prof(initial_profile, "u0_y")
        _side_values(g0["bot"], g0["top"], None, None, u0_y, xloc, yloc, N, h)
        _side_values(None, None, g0["lft"], g0["rgt"], u0_x, xloc, yloc, N, h)

    # ---- manufactured-solution data at the high-order points (src/solver.jl:570-602) ----
    if initial_profile is None:
        tpx = 2 * PI * kx
        tpy = 2 * PI * ky
        sp = np.cos(tpx * Xh + px) * np.sin(tpy * Yh + py)
        sp_x = -tpx * np.sin(tpx * Xh + px) * np.sin(tpy * Yh + py)
        sp_y = tpy * np.cos(tpx * Xh + px) * np.cos(tpy * Yh + py)
        sp_xx = -(tpx) ** 2 * sp
        sp_yy = -(tpy) ** 2 * sp
    else:
        ip = initial_profile
        sp = evaluate(_prof(ip, "u0"), Xh, Yh)
        sp_x = evaluate(_prof(ip, "u0_x"), Xh, Yh)
        sp_y = evaluate(_prof(ip, "u0_y"), Xh, Yh)
        sp_xx = evaluate(_prof(ip, "u0_xx"), Xh, Yh)
        sp_yy = evaluate(_prof(ip, "u0_yy"), Xh, Yh)
    cfh = evaluate(cC, Xh, Yh)
    c2h = cfh ** 2
    cxh = evaluate(cxC, Xh, Yh)
    cyh = evaluate(cyC, Xh, Yh)
    c2x = 2 * cfh * cxh
    c2y = 2 * cfh * cyh
    afh = evaluate(aC, Xh, Yh)
    axh = evaluate(axC, Xh, Yh)
    ayh = evaluate(ayC, Xh, Yh)
    mms_SP = np.asfortranarray(sp)
    mms_Cdiv = np.asfortranarray(c2x * sp_x + c2y * sp_y + c2h * (sp_xx + sp_yy))
    mms_Ddiv = np.asfortranarray(axh * sp_x + ayh * sp_y + afh * (sp_xx + sp_yy))
    mms_BF = np.asfortranarray(evaluate(bC, Xh, Yh))

    # ---- Phi and grad Phi on the physical boundary (src/solver.jl:604-635) ----
    bc_bot_sp = F((N, qu1)); bc_bot_spy = F((N, qu1)); bc_top_sp = F((N, qu1)); bc_top_spy = F((N, qu1))
    bc_lft_sp = F((N, qu1)); bc_lft_spx = F((N, qu1)); bc_rgt_sp = F((N, qu1)); bc_rgt_spx = F((N, qu1))
    tpx = 2 * PI * kx
    tpy = 2 * PI * ky
    bc_bot_sp[...] = np.cos(tpx * xloc + px) * math.sin(tpy * 0 + py)
    bc_bot_spy[...] = tpy * np.cos(tpx * xloc + px) * math.cos(tpy * 0 + py)
    bc_top_sp[...] = np.cos(tpx * xloc + px) * math.sin(tpy * 1 + py)
    bc_top_spy[...] = tpy * np.cos(tpx * xloc + px) * math.cos(tpy * 1 + py)
    bc_lft_sp[...] = math.cos(tpx * 0 + px) * np.sin(tpy * yloc + py)
    bc_lft_spx[...] = -tpx * math.sin(tpx * 0 + px) * np.sin(tpy * yloc + py)
    bc_rgt_sp[...] = math.cos(tpx * 1 + px) * np.sin(tpy * yloc + py)
    bc_rgt_spx[...] = -tpx * math.sin(tpx * 1 + px) * np.sin(tpy * yloc + py)
    if initial_profile is not None:
        ip = initial_profile
        zer = np.zeros_like(xloc)
        one = np.ones_like(xloc)
        bc_bot_sp[...] = evaluate(_prof(ip, "u0"), xloc, zer)
        bc_bot_spy[...] = evaluate(_prof(ip, "u0_y"), xloc, zer)
        bc_top_sp[...] = evaluate(_prof(ip, "u0"), xloc, one)
        bc_top_spy[...] = evaluate(_prof(ip, "u0_y"), xloc, one)
        bc_lft_sp[...] = evaluate(_prof(ip, "u0"), zer, yloc)
        bc_lft_spx[...] = evaluate(_prof(ip, "u0_x"), zer, yloc)
        bc_rgt_sp[...] = evaluate(_prof(ip, "u0"), one, yloc)
        bc_rgt_spx[...] = evaluate(_prof(ip, "u0_x"), one, yloc)

    # ---- c^2 and beta at the standard points (src/solver.jl:637-645) ----
    err_c2 = np.asfortranarray(cq ** 2)
    err_beta = np.asfortranarray(evaluate(bC, Xq, Yq))

    # ---- source mode (src/solver.jl:647-662) ----
    src_on = mode == "source"
    src_proj = F((N, N, qv1, qv1))
    if src_on:
        for A in (mms_SP, mms_Cdiv, mms_Ddiv, bc_bot_sp, bc_bot_spy, bc_top_sp, bc_top_spy,
                  bc_lft_sp, bc_lft_spx, bc_rgt_sp, bc_rgt_spx):
            A.fill(0.0)
        g = source.amplitude * np.exp(-((Xh - source.x0[0]) ** 2 + (Yh - source.x0[1]) ** 2) / source.r0 ** 2)
        src_proj[...] = np.einsum("pa,jkpq,qb->jkab", P_h_v, W2d_h[None, None] * g, P_h_v, optimize=True)

    pre = Pre(
        qu1=qu1, qv1=qv1, q_high=q_high, omega=omega, N=N, mode=mode,
        w=w, P_u=P_u, P_v=P_v, S_u=S_u, S_v=S_v,
        Ebl_u=Ebl_u, Ebr_u=Ebr_u, E1bl_u=E1bl_u, E1br_u=E1br_u,
        Ebl_v=Ebl_v, Ebr_v=Ebr_v, E1bl_v=E1bl_v, E1br_v=E1br_v,
        W2d_h=W2d_h, P_h_v=P_h_v, Phi2d_h_v=Phi2d_h_v,
        Cw=Cw, Aw=Aw, Ainv_u=Ainv_u,
        mms_SP=mms_SP, mms_Cdiv=mms_Cdiv, mms_Ddiv=mms_Ddiv, mms_BF=mms_BF,
        bc_bot_sp=bc_bot_sp, bc_bot_spy=bc_bot_spy, bc_top_sp=bc_top_sp, bc_top_spy=bc_top_spy,
        bc_lft_sp=bc_lft_sp, bc_lft_spx=bc_lft_spx, bc_rgt_sp=bc_rgt_sp, bc_rgt_spx=bc_rgt_spx,
        theta=fp["theta"], tau=fp["tau"], lambda_u=fp["lambda_u"], lambda_v=fp["lambda_v"],
        gamma_abc=gamma_abc, eta_abc=math.sqrt(1.0 - gamma_abc ** 2) if legacy else math.nan,
        sigma=bc["bot"]["sigma"] if legacy else math.nan,
        bc=bc, src_on=src_on, src_proj=src_proj, source=source,
        mms_on=(mode == "mms"), mass_solver=mass_solver,
    )
    for s in SIDE_NAMES:
        for k in ("c2", "af", "cf", "bf"):
            setattr(pre, f"{s}_{k}", faces[s][k])
        setattr(pre, f"{s}_g0", g0[s])
    pre = _rhs_data(pre, N)
    geo = Pre(S_u=S_u, W2d=W2d, xloc=xloc, yloc=yloc, err_c2=err_c2, err_beta=err_beta)
    return pre, geo


def _prof(ip, name) -> Callable:
    return ip[name] if isinstance(ip, dict) else getattr(ip, name)


def _elem_mid(A4, N):
    """[j, k, iy, ix] -> (iy, e, ix) with e = j + N k (the reference's Cw layout)."""
    n = A4.shape[2]
    out = F((n, N * N, n))
    out[...] = np.transpose(A4, (2, 1, 0, 3)).reshape(n, N * N, n)
    return out


def src_time(pre, t: float) -> float:
    """Time factor of the source term; 0 outside :source mode (src/solver.jl:662)."""
    return pre.source.time_factor(t) if pre.src_on else 0.0


def _rhs_data(pre: Pre, N: int) -> Pre:
    """Face arrays by face, interface jumps, projections and the Chebyshev layout of
    beta (src/solver.jl:716-770)."""
    nq, nv, qh = pre.qu1, pre.qv1, pre.q_high

    def hface(bot, top):
        A = F((nq, N + 1, N))
        A[:, :N, :] = np.transpose(bot, (2, 0, 1))
        A[:, N, :] = top[N - 1, :, :].T
        return A

    def vface(lft, rgt):
        A = F((nq, N, N + 1))
        A[:, :, :N] = np.transpose(lft, (2, 0, 1))
        A[:, :, N] = rgt[:, N - 1, :].T
        return A

    def hfaceM(bot, top):
        A = F((nq, N + 1, N))
        A[:, 0, :] = bot[0, :, :].T
        A[:, 1:, :] = np.transpose(top, (2, 0, 1))
        return A

    def vfaceM(lft, rgt):
        A = F((nq, N, N + 1))
        A[:, :, 0] = lft[:, 0, :].T
        A[:, :, 1:] = np.transpose(rgt, (2, 0, 1))
        return A

    PhT = pre.P_h_v.T

    def proj(X):
        out = F((nv, nv, N, N))
        out[...] = np.einsum("pa,jkpq,qb->abjk", pre.P_h_v, pre.W2d_h[None, None] * X, pre.P_h_v,
                             optimize=True)
        return out

    def elast(X):
        return np.asfortranarray(np.transpose(X, (2, 3, 0, 1)))

    hc2 = hface(pre.bot_c2, pre.top_c2); hc2m = hfaceM(pre.bot_c2, pre.top_c2)
    haf = hface(pre.bot_af, pre.top_af); hafm = hfaceM(pre.bot_af, pre.top_af)
    vc2 = vface(pre.lft_c2, pre.rgt_c2); vc2m = vfaceM(pre.lft_c2, pre.rgt_c2)
    vaf = vface(pre.lft_af, pre.rgt_af); vafm = vfaceM(pre.lft_af, pre.rgt_af)
    hg0 = hface(pre.bot_g0, pre.top_g0); hg0m = hfaceM(pre.bot_g0, pre.top_g0)
    vg0 = vface(pre.lft_g0, pre.rgt_g0); vg0m = vfaceM(pre.lft_g0, pre.rgt_g0)

    def differ(a, b):
        return not np.array_equal(a, b)    # NaN != NaN, like Julia's !=

    jumps = differ(hc2m, hc2) or differ(hafm, haf) or differ(vc2m, vc2) or differ(vafm, vaf)
    if pre.mms_on and jumps:
        raise ValueError("mode = :mms needs c and α continuous across the element faces; "
                         "use :initial_value or :source for coefficients that jump")
    with np.errstate(invalid="ignore", over="ignore"):
        hJC = np.asfortranarray(hc2m * hg0m - hc2 * hg0)
        vJC = np.asfortranarray(vc2m * vg0m - vc2 * vg0)
    hJC[:, 0, :] = 0.0; hJC[:, N, :] = 0.0
    vJC[:, :, 0] = 0.0; vJC[:, :, N] = 0.0
    one_sided = bool(jumps or np.any(hJC != 0) or np.any(vJC != 0))
    beta_hE = elast(pre.mms_BF)
    beta_hq = F((qh, N * N, qh))
    beta_hq[...] = np.transpose(beta_hE.reshape(qh, qh, N * N, order="F"), (0, 2, 1))
    return pre.merge(
        hc2=hc2, haf=haf, hcf=hface(pre.bot_cf, pre.top_cf), hbf=hface(pre.bot_bf, pre.top_bf),
        vc2=vc2, vaf=vaf, vcf=vface(pre.lft_cf, pre.rgt_cf), vbf=vface(pre.lft_bf, pre.rgt_bf),
        hc2m=hc2m, hafm=hafm, vc2m=vc2m, vafm=vafm, hJC=hJC, vJC=vJC, one_sided=one_sided,
        bcB_sp=np.asfortranarray(pre.bc_bot_sp.T), bcB_spy=np.asfortranarray(pre.bc_bot_spy.T),
        bcT_sp=np.asfortranarray(pre.bc_top_sp.T), bcT_spy=np.asfortranarray(pre.bc_top_spy.T),
        bcL_sp=np.asfortranarray(pre.bc_lft_sp.T), bcL_spx=np.asfortranarray(pre.bc_lft_spx.T),
        bcR_sp=np.asfortranarray(pre.bc_rgt_sp.T), bcR_spx=np.asfortranarray(pre.bc_rgt_spx.T),
        PiSP=proj(pre.mms_SP), PiBSP2=proj(pre.mms_BF * pre.mms_SP ** 2),
        PiCdiv=proj(pre.mms_Cdiv), PiDdiv=proj(pre.mms_Ddiv),
        src_projE=elast(pre.src_proj), beta_hE=beta_hE, beta_hq=beta_hq,
        W_hq=np.asfortranarray(pre.W2d_h.reshape(qh, 1, qh, order="F")),
        P_h_vT=np.asfortranarray(PhT), mass_tol=MASS_TOL,
        P_uT=np.asfortranarray(pre.P_u.T), S_uT=np.asfortranarray(pre.S_u.T),
        P_vT=np.asfortranarray(pre.P_v.T), S_vT=np.asfortranarray(pre.S_v.T),
    )
