"""The adjoint oracle (oracle/adjoint.py) pinned on the CPU.

* its right-hand side against the C oracle (itself pinned by the reference's
  fixtures, tests/test_oracle.py) on precompute's tables for the same
  per-element coefficients;
* its hand-written vector-Jacobian product against complex-step derivatives
  of its right-hand side, Im F(p + i h d) / h (exact to rounding, no
  cancellation), through <ybar, F'(p) d> = <F'(p)^T ybar, d> in the state, c^2
  and beta separately;
* its trajectory gradient against complex-step derivatives of its misfit and
  against central differences (eps = 1e-6) of the C oracle's RK4.
"""
import math

import numpy as np
import pytest

import edg_westervelt_amd as w
from oracle import adjoint as A

SIDES = dict(bot="absorbing_2beta", top="wall", lft="dirichlet", rgt="absorbing")
FLUX = dict(theta=0.7, tau=0.4, lambda_u=0.9, lambda_v=1.1)
SRC = ((0.5, 0.35), 0.2, 1.0, 16 * math.pi, 1.0)


def problem(q, qv, N, rng):
    sides = [A.side_params(SIDES[s], 0.3, 0.6, 0.5) for s in ("bot", "top", "lft", "rgt")]
    al = 1e-3 * (1 + rng.random((N, N)))
    pb = A.Problem(q, qv, N, flux=(FLUX["theta"], FLUX["tau"], FLUX["lambda_u"], FLUX["lambda_v"]), sides=sides,
                   alpha=al, source=SRC)
    c2 = 1 + 0.3 * rng.random((N, N))
    beta = 2 + rng.random((N, N))
    return pb, c2, beta, al


def c_oracle_pre(q, qv, N, c2, beta, al):
    x0, r0, amp, om, Ts = SRC
    pre, _ = w.precompute(q, N, mixed=(qv == q - 1), mode="source", boundary=SIDES, sigma=0.3, sigma_dirichlet=0.6,
                          kappa=0.5, flux_params=FLUX,
                          coefficients=w.element_coefficients(c=np.sqrt(c2), alpha=al, beta=beta),
                          source=w.GaussianSource(x0=x0, r0=r0, amplitude=amp, omega=om, Ts=Ts))
    return pre


def test_tables_against_precompute():
    pb = A.Problem(4, 4, 3, sides=[A.side_params("wall")] * 4, alpha=1e-3, source=SRC)
    pre, _ = w.precompute(4, 3, mixed=False, mode="source")
    assert np.abs(pb.P - pre.P_u).max() < 1e-14
    assert np.abs(pb.S - pre.S_u).max() < 1e-13
    assert np.abs(pb.Ph - pre.P_h_v).max() < 1e-14
    assert np.abs(pb.El - pre.Ebl_u[0]).max() < 1e-14 and np.abs(pb.E1r - pre.E1br_u[0]).max() < 1e-13


@pytest.mark.parametrize("q,mixed", [(3, False), (3, True), (2, False)])
def test_rhs_against_c_oracle(oracle_built, q, mixed):
    rng = np.random.default_rng(11 + q)
    N, qv = 6, q - 1 if mixed else q
    pb, c2, beta, al = problem(q, qv, N, rng)
    pre = c_oracle_pre(q, qv, N, c2, beta, al)
    assert pre.one_sided
    u = np.asfortranarray(1e-2 * rng.standard_normal((N, N, pb.nu, pb.nu)))
    v = np.asfortranarray(1e-2 * rng.standard_normal((N, N, pb.nv, pb.nv)))
    for t in (0.02, 0.3):
        ou, ov, _ = oracle_built.rhs(oracle_built.OraclePre(pre, N), u, v, t)
        au, av = A.rhs(pb, c2, beta, u, v, t)
        assert np.abs(ou - au).max() <= 1e-12 * np.abs(ou).max()
        assert np.abs(ov - av).max() <= 1e-12 * np.abs(ov).max()


@pytest.mark.parametrize("q,mixed", [(3, False), (3, True)])
def test_vjp_against_complex_step(q, mixed):
    rng = np.random.default_rng(5)
    N, qv = 5, q - 1 if mixed else q
    pb, c2, beta, _ = problem(q, qv, N, rng)
    u = 1e-2 * rng.standard_normal((N, N, pb.nu, pb.nu))
    v = 1e-2 * rng.standard_normal((N, N, pb.nv, pb.nv))
    t = 0.013
    ut, vt = A.rhs(pb, c2, beta, u, v, t)
    ub = rng.standard_normal(ut.shape)
    vb = rng.standard_normal(vt.shape)
    bars = dict(zip(("u", "v", "c2", "beta"), A.rhs_vjp(pb, c2, beta, u, v, t, ub, vb)))
    base = dict(u=u, v=v, c2=c2, beta=beta)
    h = 1e-30
    for name in base:
        d = rng.standard_normal(base[name].shape)
        arg = {k: (x + 1j * h * d if k == name else x) for k, x in base.items()}
        ut2, vt2 = A.rhs(pb, arg["c2"], arg["beta"], arg["u"], arg["v"], t)
        tl = (ub * ut2.imag / h).sum() + (vb * vt2.imag / h).sum()
        ad = (bars[name] * d).sum()
        assert abs(tl - ad) <= 1e-12 * abs(tl), (name, tl, ad)


def _target(pb, dt, ns):
    N = pb.N
    return A.misfit(pb, np.ones((N, N)), 3.0 * np.ones((N, N)), dt, ns, traces=True)[1]


def test_gradient_against_complex_step():
    rng = np.random.default_rng(7)
    q, N = 3, 5
    pb, c2, beta, _ = problem(q, q, N, rng)
    dt, ns = 2e-3, 25
    tgt = _target(pb, dt, ns)
    J, gc2, gb = A.gradient(pb, c2, beta, dt, ns, tgt)
    assert J > 0
    assert abs(A.misfit(pb, c2, beta, dt, ns, tgt) - J) <= 1e-14 * J
    for _ in range(3):
        dc, db = rng.standard_normal((N, N)), rng.standard_normal((N, N))
        Jc = A.misfit(pb, c2 + 1j * 1e-30 * dc, beta + 1j * 1e-30 * db, dt, ns, tgt)
        cs = Jc.imag / 1e-30
        ad = (gc2 * dc).sum() + (gb * db).sum()
        assert abs(cs - ad) <= 1e-10 * abs(cs)


def test_gradient_against_central_differences_of_the_c_oracle(oracle_built):
    rng = np.random.default_rng(8)
    q, N = 2, 4
    pb, c2, beta, al = problem(q, q, N, rng)
    dt, ns = 2e-3, 15
    tgt = _target(pb, dt, ns)
    J, gc2, gb = A.gradient(pb, c2, beta, dt, ns, tgt)

    def J_c(c2x, bx):
        pre = c_oracle_pre(q, q, N, c2x, bx, al)
        op = oracle_built.OraclePre(pre, N)
        u = np.zeros((N, N, pre.qu1, pre.qu1), order="F")
        v = np.zeros((N, N, pre.qv1, pre.qv1), order="F")
        t, s = 0.0, 0.0
        for n in range(ns):
            u, v, _, _ = oracle_built.rk4_steps(op, u, v, t, dt, 1)
            t += dt
            r = A.bottom_trace(pb, u) - tgt[n]
            s += 0.5 * np.sum(pb.wgt[None, :] * r * r)
        return s

    assert abs(J_c(c2, beta) - J) <= 1e-10 * J
    eps = 1e-6
    for _ in range(2):
        dc, db = rng.standard_normal((N, N)), rng.standard_normal((N, N))
        fd = (J_c(c2 + eps * dc, beta + eps * db) - J_c(c2 - eps * dc, beta - eps * db)) / (2 * eps)
        ad = (gc2 * dc).sum() + (gb * db).sum()
        assert abs(fd - ad) <= 1e-6 * abs(ad), (fd, ad)
