"""Element-by-element dense restatement of the RHS (test infrastructure).

Follows test/reference_rhs.jl:48-352 (the v0.1.0 formulation: dense volume
operators, per-face formulas with one-sided coefficients, dense Cholesky
u-solve and a dense solve of every mass system) with the operator set-up of
test/runtests.jl:11-20 (``dense_volume``) and the boundary mapping of
test/runtests.jl:233-239 (``reference_pre``: uniform boundary kinds only).
Structurally different from oracle/wedg_oracle.c, so the two pin each other.
Python loops: small meshes only.
"""
from __future__ import annotations

import numpy as np


def _vec(X):
    return np.asarray(X).ravel(order="F")


def dense_volume(pre, N):
    """Mc2_uv, Ma_vv and the Cholesky factors of the pinned u-stiffness per element
    (test/runtests.jl:11-20)."""
    Pxu = np.kron(pre.P_u, pre.S_u); Peu = np.kron(pre.S_u, pre.P_u)
    Pxv = np.kron(pre.P_v, pre.S_v); Pev = np.kron(pre.S_v, pre.P_v)
    Mc2, Ma, Ku = {}, {}, {}
    for k in range(N):
        for j in range(N):
            e = j + N * k
            cw = _vec(pre.Cw[:, e, :]); aw = _vec(pre.Aw[:, e, :])
            Mc2[j, k] = Pxv.T @ (cw[:, None] * Pxu) + Pev.T @ (cw[:, None] * Peu)
            Ma[j, k] = Pxv.T @ (aw[:, None] * Pxv) + Pev.T @ (aw[:, None] * Pev)
            K = Pxu.T @ (cw[:, None] * Pxu) + Peu.T @ (cw[:, None] * Peu)
            K[0, 0] += 1.0
            Ku[j, k] = K
    return Mc2, Ma, Ku


def reference_bc(pre):
    """(gamma, eta, sigma, k) of the uniform boundary kind (test/runtests.jl:233-239)."""
    b = pre.bc["bot"]
    kinds = {pre.bc[s]["kind"] for s in ("bot", "top", "lft", "rgt")}
    if len(kinds) != 1:
        raise ValueError("the dense restatement takes one boundary kind for all sides")
    if b["kind"] == "wall":
        return 0.0, 1.0, -b["kappa"], 0.0
    return b["gamma_abc"], b["eta_abc"], b["sigma"], b["kb"]


def _abc_face(g, et, sg, c2f, af, cf, bf, v_fp, gun, gvn, gu_ex, gv_ex, v_ex):
    """test/reference_rhs.jl:33-46 (the weight sqrt(1 - bf v) with bf pre-scaled by k)."""
    rho_num = et * (c2f * gun + af * gvn) + g * cf * np.sqrt(1.0 - bf * v_fp) * v_fp
    rho_ex = et * (c2f * gu_ex + af * gv_ex) + g * cf * np.sqrt(1.0 - bf * v_ex) * v_ex
    rho_net = rho_num - rho_ex
    sqf = np.sqrt(1.0 - bf * v_fp)
    F1 = c2f * gun + af * gvn - (et + sg * g) * rho_net
    dv = -(g - sg * et) / (cf * sqf) * rho_net
    return F1, dv


def edg_rhs_dense(u, v, t, N, h, pre):
    """(u_t, v_t) by the element-by-element formulation (test/reference_rhs.jl:48-352)."""
    qu1, qv1, om = pre.qu1, pre.qv1, pre.omega
    w = pre.w
    P_u, P_v = pre.P_u, pre.P_v
    Ebl_u, Ebr_u = pre.Ebl_u[0], pre.Ebr_u[0]
    E1bl_u, E1br_u = pre.E1bl_u[0], pre.E1br_u[0]
    Ebl_v, Ebr_v = pre.Ebl_v[0], pre.Ebr_v[0]
    E1bl_v, E1br_v = pre.E1bl_v[0], pre.E1br_v[0]
    W2d_h, P_h_v, Phi = pre.W2d_h, pre.P_h_v, pre.Phi2d_h_v
    th, tau, lu, lv = pre.theta, pre.tau, pre.lambda_u, pre.lambda_v
    g, et, sg, kref = reference_bc(pre)
    two_h = 2.0 / h
    four_hh = 4.0 / (h * h)
    ct = np.cos(om * t) if pre.mms_on else 0.0
    st = np.sin(om * t) if pre.mms_on else 0.0
    srct = pre.source.time_factor(t) if pre.src_on else 0.0
    Mc2, Ma, Ku = dense_volume(pre, N)

    def pen_w(c2m, c2p, afm, afp):
        return 0.5 * (c2m + c2p) * lu + 0.5 * (afm + afp) * lv

    def Y(P, vals):              # P' (w .* vals)
        return P.T @ (w * vals)

    ut = np.zeros_like(u); vt = np.zeros_like(v)
    rhs_all = {}
    for j in range(N):
        for k in range(N):
            U = u[j, k]; V = v[j, k]
            vr = -four_hh * (Mc2[j, k] @ _vec(U) + Ma[j, k] @ _vec(V)).reshape(qv1, qv1, order="F")
            SP, Cd, Dd, BF = pre.mms_SP[j, k], pre.mms_Cdiv[j, k], pre.mms_Ddiv[j, k], pre.mms_BF[j, k]
            vh_mms = -om * st * SP
            vth_mms = -om ** 2 * ct * SP
            fv = (1.0 - 2.0 * BF * vh_mms) * vth_mms - ct * Cd + om * st * Dd + Cd
            vr = vr + P_h_v.T @ (W2d_h * fv) @ P_h_v
            if pre.src_on:
                vr = vr + srct * pre.src_proj[j, k]
            f = np.zeros((qu1, qu1))
            # bottom face, n = (0, -1)
            if j == 0:
                c2f, af, cf, bf = pre.bot_c2[j, k], pre.bot_af[j, k], pre.bot_cf[j, k], kref * pre.bot_bf[j, k]
                v_fp = P_v @ (Ebl_v @ V)
                uy = P_u @ (two_h * (E1bl_u @ U)); vy = P_v @ (two_h * (E1bl_v @ V))
                sp_, spy = pre.bc_bot_sp[k], pre.bc_bot_spy[k]
                F1, dv = _abc_face(g, et, sg, c2f, af, cf, bf, v_fp, -uy, -vy,
                                   -ct * spy + spy, om * st * spy, -om * st * sp_)
                vr = vr + two_h * np.outer(Ebl_v, Y(P_v, F1)) - four_hh * np.outer(E1bl_v, Y(P_v, af * dv))
                f = f + np.outer(E1bl_u, Y(P_u, c2f * dv))
            else:
                Uo, Vo = u[j - 1, k], v[j - 1, k]
                c2p, afp = pre.bot_c2[j, k], pre.bot_af[j, k]
                c2m, afm = pre.top_c2[j - 1, k], pre.top_af[j - 1, k]
                K = c2m * pre.top_g0[j - 1, k] - c2p * pre.bot_g0[j, k]
                vj = P_v @ (Ebr_v @ Vo - Ebl_v @ V)
                gum, gup = P_u @ (E1br_u @ Uo), P_u @ (E1bl_u @ U)
                gvm, gvp = P_v @ (E1br_v @ Vo), P_v @ (E1bl_v @ V)
                jw = two_h * (c2m * gum - c2p * gup + afm * gvm - afp * gvp) + K
                fl = th * (c2p * gup + afp * gvp) + (1 - th) * (c2m * gum + afm * gvm)
                dvp = th * vj - tau * jw
                pen = pen_w(c2m, c2p, afm, afp) * vj - (1 - th) * K
                f = f + np.outer(E1bl_u, Y(P_u, c2p * dvp))
                vr = (vr - four_hh * np.outer(Ebl_v, Y(P_v, fl)) - four_hh * np.outer(E1bl_v, Y(P_v, afp * dvp))
                      + two_h * np.outer(Ebl_v, Y(P_v, pen)))
            # top face, n = (0, +1)
            if j == N - 1:
                c2f, af, cf, bf = pre.top_c2[j, k], pre.top_af[j, k], pre.top_cf[j, k], kref * pre.top_bf[j, k]
                v_fp = P_v @ (Ebr_v @ V)
                uy = P_u @ (two_h * (E1br_u @ U)); vy = P_v @ (two_h * (E1br_v @ V))
                sp_, spy = pre.bc_top_sp[k], pre.bc_top_spy[k]
                F1, dv = _abc_face(g, et, sg, c2f, af, cf, bf, v_fp, uy, vy,
                                   ct * spy - spy, -om * st * spy, -om * st * sp_)
                vr = vr + two_h * np.outer(Ebr_v, Y(P_v, F1)) + four_hh * np.outer(E1br_v, Y(P_v, af * dv))
                f = f - np.outer(E1br_u, Y(P_u, c2f * dv))
            else:
                Uo, Vo = u[j + 1, k], v[j + 1, k]
                c2m, afm = pre.top_c2[j, k], pre.top_af[j, k]
                c2p, afp = pre.bot_c2[j + 1, k], pre.bot_af[j + 1, k]
                K = c2m * pre.top_g0[j, k] - c2p * pre.bot_g0[j + 1, k]
                vj = P_v @ (Ebl_v @ Vo - Ebr_v @ V)
                gum, gup = P_u @ (E1br_u @ U), P_u @ (E1bl_u @ Uo)
                gvm, gvp = P_v @ (E1br_v @ V), P_v @ (E1bl_v @ Vo)
                jw = two_h * (c2m * gum - c2p * gup + afm * gvm - afp * gvp) + K
                fl = (1 - th) * (c2m * gum + afm * gvm) + th * (c2p * gup + afp * gvp)
                dvp = (1 - th) * vj - tau * jw
                pen = pen_w(c2m, c2p, afm, afp) * vj - th * K
                f = f - np.outer(E1br_u, Y(P_u, c2m * dvp))
                vr = (vr + four_hh * np.outer(Ebr_v, Y(P_v, fl)) + four_hh * np.outer(E1br_v, Y(P_v, afm * dvp))
                      + two_h * np.outer(Ebr_v, Y(P_v, pen)))
            # left face, n = (-1, 0): the y-face pattern on transposes
            if k == 0:
                c2f, af, cf, bf = pre.lft_c2[j, k], pre.lft_af[j, k], pre.lft_cf[j, k], kref * pre.lft_bf[j, k]
                v_fp = P_v @ (Ebl_v @ V.T)
                ux = P_u @ (two_h * (E1bl_u @ U.T)); vx = P_v @ (two_h * (E1bl_v @ V.T))
                sp_, spx = pre.bc_lft_sp[j], pre.bc_lft_spx[j]
                F1, dv = _abc_face(g, et, sg, c2f, af, cf, bf, v_fp, -ux, -vx,
                                   -ct * spx + spx, om * st * spx, -om * st * sp_)
                va = two_h * np.outer(Ebl_v, Y(P_v, F1)) - four_hh * np.outer(E1bl_v, Y(P_v, af * dv))
                fa = np.outer(E1bl_u, Y(P_u, c2f * dv))
                vr = vr + va.T
                f = f + fa.T
            else:
                Uo, Vo = u[j, k - 1].T, v[j, k - 1].T
                c2p, afp = pre.lft_c2[j, k], pre.lft_af[j, k]
                c2m, afm = pre.rgt_c2[j, k - 1], pre.rgt_af[j, k - 1]
                K = c2m * pre.rgt_g0[j, k - 1] - c2p * pre.lft_g0[j, k]
                vj = P_v @ (Ebr_v @ Vo - Ebl_v @ V.T)
                gum, gup = P_u @ (E1br_u @ Uo), P_u @ (E1bl_u @ U.T)
                gvm, gvp = P_v @ (E1br_v @ Vo), P_v @ (E1bl_v @ V.T)
                jw = two_h * (c2m * gum - c2p * gup + afm * gvm - afp * gvp) + K
                fl = th * (c2p * gup + afp * gvp) + (1 - th) * (c2m * gum + afm * gvm)
                dvp = th * vj - tau * jw
                pen = pen_w(c2m, c2p, afm, afp) * vj - (1 - th) * K
                f = f + np.outer(Y(P_u, c2p * dvp), E1bl_u)
                vr = (vr - four_hh * np.outer(Y(P_v, fl), Ebl_v) - four_hh * np.outer(Y(P_v, afp * dvp), E1bl_v)
                      + two_h * np.outer(Y(P_v, pen), Ebl_v))
            # right face, n = (+1, 0)
            if k == N - 1:
                c2f, af, cf, bf = pre.rgt_c2[j, k], pre.rgt_af[j, k], pre.rgt_cf[j, k], kref * pre.rgt_bf[j, k]
                v_fp = P_v @ (Ebr_v @ V.T)
                ux = P_u @ (two_h * (E1br_u @ U.T)); vx = P_v @ (two_h * (E1br_v @ V.T))
                sp_, spx = pre.bc_rgt_sp[j], pre.bc_rgt_spx[j]
                F1, dv = _abc_face(g, et, sg, c2f, af, cf, bf, v_fp, ux, vx,
                                   ct * spx - spx, -om * st * spx, -om * st * sp_)
                va = two_h * np.outer(Ebr_v, Y(P_v, F1)) + four_hh * np.outer(E1br_v, Y(P_v, af * dv))
                fa = -np.outer(E1br_u, Y(P_u, c2f * dv))
                vr = vr + va.T
                f = f + fa.T
            else:
                Uo, Vo = u[j, k + 1].T, v[j, k + 1].T
                c2m, afm = pre.rgt_c2[j, k], pre.rgt_af[j, k]
                c2p, afp = pre.lft_c2[j, k + 1], pre.lft_af[j, k + 1]
                K = c2m * pre.rgt_g0[j, k] - c2p * pre.lft_g0[j, k + 1]
                vj = P_v @ (Ebl_v @ Vo - Ebr_v @ V.T)
                gum, gup = P_u @ (E1br_u @ U.T), P_u @ (E1bl_u @ Uo)
                gvm, gvp = P_v @ (E1br_v @ V.T), P_v @ (E1bl_v @ Vo)
                jw = two_h * (c2m * gum - c2p * gup + afm * gvm - afp * gvp) + K
                fl = (1 - th) * (c2m * gum + afm * gvm) + th * (c2p * gup + afp * gvp)
                dvp = (1 - th) * vj - tau * jw
                pen = pen_w(c2m, c2p, afm, afp) * vj - th * K
                f = f - np.outer(Y(P_u, c2m * dvp), E1br_u)
                vr = (vr + four_hh * np.outer(Y(P_v, fl), Ebr_v) + four_hh * np.outer(Y(P_v, afm * dvp), E1br_v)
                      + two_h * np.outer(Y(P_v, pen), Ebr_v))
            rhs_all[j, k] = vr
            rhs_u = Mc2[j, k].T @ _vec(V) - _vec(f)
            rhs_u[0] += V[0, 0]
            ut[j, k] = np.linalg.solve(Ku[j, k], rhs_u).reshape(qu1, qu1, order="F")
    for j in range(N):
        for k in range(N):
            vq = P_h_v @ v[j, k] @ P_h_v.T
            wv = _vec((1.0 - 2.0 * pre.mms_BF[j, k] * vq) * W2d_h)
            M = Phi.T @ (wv[:, None] * Phi)
            vt[j, k] = np.linalg.solve(M, _vec(rhs_all[j, k])).reshape(qv1, qv1, order="F")
    return ut, vt


def mass_dense(pre, N, V, R):
    """X = M(V)^-1 R element by element by a dense solve (test/runtests.jl:423-434);
    V, R in the element-last layout (nv, nv, N, N)."""
    B = pre.Phi2d_h_v
    Wv = _vec(pre.W2d_h)
    X = np.zeros_like(R)
    for k in range(N):
        for j in range(N):
            a = 1 - 2 * _vec(pre.beta_hE[:, :, j, k]) * (B @ _vec(V[:, :, j, k]))
            M = B.T @ ((a * Wv)[:, None] * B)
            X[:, :, j, k] = np.linalg.solve(M, _vec(R[:, :, j, k])).reshape(pre.qv1, pre.qv1, order="F")
    return X
