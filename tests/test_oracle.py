"""The CPU oracle pinned against the reference's own tests and data (no GPU).

Each check follows a test of test/runtests.jl or a fixture of data/, restated;
together they pin every function of oracle/wedg_oracle.c:
* the element-by-element dense restatement (oracle/dense.py) at <= 1e-12 over
  the reference's matrix (test/runtests.jl:189-293, 716-732);
* the golden errors of data/convergence/results.csv (:360, :364, and every
  q / pairing / flux of the table at N = 5) and of test/runtests.jl:207-217;
* the :initial_value mode against data/alpha_limit/order.csv;
* mass solvers against dense solves (test/runtests.jl:418-444);
* boundary flux closed forms (test/runtests.jl:295-317);
* locality and superposition of the boundary sides (test/runtests.jl:319-338);
* the semi-discrete energy identity with jumps (test/runtests.jl:739-758);
* exact consistency at an interface (test/runtests.jl:790-811).
"""
import ctypes as C
import math

import numpy as np
import pytest

from edg_westervelt_amd import (GaussianSource, constant_coefficients, paper_coefficients, precompute,
                                replace_coefficients)
from edg_westervelt_amd.coefficients import Coefficients
from edg_westervelt_amd.quadrature import GaussQCofs, GradJacobiP, JacobiP
from oracle import batched as ob
from oracle import dense, norms
from tests.cases import BUMP, MIXED, alpha_limit_rows, convergence_row, quadrant_coefficients, rand_state
from tests.conftest import rel


@pytest.fixture(scope="module", autouse=True)
def _built(oracle_built):
    return oracle_built


@pytest.mark.parametrize("mode", ["mms", "initial_value", "source"])
def test_batched_equals_dense(mode):
    """test/runtests.jl:189-202, 266-278 on the oracle."""
    rng = np.random.default_rng(2)
    worst = 0.0
    for boundary in ("absorbing", "dirichlet", "absorbing_2beta", "wall"):
        for q in (1, 3, 4):
            for N in (1, 2, 4):
                for flux in ("upwind", "central", "alternating"):
                    for mixed in (True, False):
                        pre, _ = precompute(q, N, flux=flux, mixed=mixed, mode=mode, boundary=boundary,
                                            sigma=0.3, sigma_dirichlet=0.6, kappa=0.7, phase=(0.4, 0.3))
                        u, v = rand_state(rng, N, pre)
                        a = dense.edg_rhs_dense(u, v, 0.037, N, 1.0 / N, pre)
                        b = ob.rhs(ob.OraclePre(pre, N), u, v, 0.037)
                        worst = max(worst, rel(a[0], b[0]), rel(a[1], b[1]))
    assert worst < 1e-12


@pytest.mark.parametrize("mode", ["initial_value", "source"])
def test_batched_equals_dense_one_sided(mode):
    """test/runtests.jl:716-732: jumps of c, alpha, beta; custom flux parameters."""
    rng = np.random.default_rng(6)
    co = quadrant_coefficients(0.5, 0.5, smooth=True)
    worst = 0.0
    for q in (1, 3):
        for N in (2, 4):
            for mixed in (True, False):
                for boundary in ("absorbing_2beta", "dirichlet", "wall"):
                    for fp in (dict(theta=1.0, tau=1.0, lambda_u=1.0, lambda_v=1.0),
                               dict(theta=0.5, tau=0.0, lambda_u=0.0, lambda_v=0.0),
                               dict(theta=1.0, tau=0.0, lambda_u=0.0, lambda_v=0.0),
                               dict(theta=0.3, tau=0.7, lambda_u=0.4, lambda_v=0.9)):
                        pre, _ = precompute(q, N, flux_params=fp, mixed=mixed, mode=mode, boundary=boundary,
                                            coefficients=co, sigma=0.3, sigma_dirichlet=0.6, kappa=0.7,
                                            phase=(0.4, 0.3))
                        assert pre.one_sided
                        u, v = rand_state(rng, N, pre)
                        a = dense.edg_rhs_dense(u, v, 0.037, N, 1.0 / N, pre)
                        b = ob.rhs(ob.OraclePre(pre, N), u, v, 0.037)
                        worst = max(worst, rel(a[0], b[0]), rel(a[1], b[1]))
    assert worst < 1e-12


GOLD_N6 = dict(erru=5.9847841944e-04, errv=1.9998237454e-02, errgradu=2.4571643407e-02, errE=2.2400357448e-02)
GOLD_N6_ABS = dict(erru=5.9847813824e-04, errv=1.9998249758e-02, errgradu=2.4571634466e-02,
                   errE=2.2400354434e-02)
GOLD_CFG1 = dict(erru=1.0717765588e-04, errv=4.5772517217e-03, errgradu=7.1299988239e-03,
                 errE=5.9884034343e-03)


@pytest.mark.parametrize("solver", ["cg", "chebyshev"])
@pytest.mark.parametrize("case", [(6, "absorbing_2beta", GOLD_N6), (6, "absorbing", GOLD_N6_ABS),
                                  (10, "absorbing_2beta", GOLD_CFG1)])
def test_golden_errors(solver, case):
    """data/convergence/results.csv:360, :364 (cfg1) and test/runtests.jl:207-217 at rtol 1e-9."""
    N, boundary, gold = case
    pre, geo = precompute(3, N, flux="upwind", mixed=True, boundary=boundary, mass_solver=solver)
    u, v, t, dt, ns = ob.solve_state(pre, N, 0.1, 3)
    assert ns == round(0.1 / (1.0 / N / 90))
    e = norms.mms_errors(u, v, 0.1, N, 1.0 / N, pre, geo)
    for k, g in gold.items():
        assert abs(e[k] / g - 1) < 1e-9, (k, e[k], g)


def test_convergence_order():
    """test/runtests.jl:220-225."""
    es = []
    for N in (4, 8):
        pre, geo = precompute(2, N, flux="upwind", mixed=True)
        u, v, t, dt, ns = ob.solve_state(pre, N, 0.1, 2)
        es.append(norms.mms_errors(u, v, 0.1, N, 1.0 / N, pre, geo))
    assert math.log2(es[0]["erru"] / es[1]["erru"]) > 2.5
    assert math.log2(es[0]["errE"] / es[1]["errE"]) > 1.5


def test_mass_solvers_against_dense():
    """test/runtests.jl:418-444."""
    rng = np.random.default_rng(7)
    for (q, mixed, N) in ((1, True, 3), (3, True, 4), (4, False, 3), (6, True, 2)):
        pre, _ = precompute(q, N, mixed=mixed)
        op = ob.OraclePre(pre, N)
        nv = pre.qv1
        B = pre.Phi2d_h_v
        V0 = np.asfortranarray(rng.standard_normal((nv, nv, N, N)))
        R = np.asfortranarray(rng.standard_normal((nv, nv, N, N)))
        g0 = max(np.abs(2 * pre.beta_hE[:, :, j, k].ravel(order="F") * (B @ V0[:, :, j, k].ravel(order="F"))).max()
                 for j in range(N) for k in range(N))
        for gmax in (1e-3, 0.5, 0.9):
            V = np.asfortranarray(V0 * (gmax / g0))
            Xref = dense.mass_dense(pre, N, V, R)
            for mode in (0, 1):
                Xc, _ = ob.mass_cheb(op, V, R, cheb_mode=mode)
                assert np.abs(Xc - Xref).max() <= 1e-12 * np.abs(Xref).max()
            Xg, _ = ob.mass_cg(op, V, R)
            assert np.abs(Xg - Xref).max() <= 1e-12 * np.abs(Xref).max()
        X0, st = ob.mass_cheb(op, np.zeros_like(V0), R)
        assert np.array_equal(X0, R) and st.cheb_steps == 1
        gpos = max((2 * pre.beta_hE[:, :, j, k].ravel(order="F") * (B @ V0[:, :, j, k].ravel(order="F"))).max()
                   for j in range(N) for k in range(N))
        with pytest.raises(RuntimeError, match="mass matrix not positive definite"):
            ob.mass_cheb(op, np.asfortranarray(V0 * (1.5 / gpos)), R)


def test_boundary_flux_closed_forms():
    """Consistency (data = trace => no correction) and the boundary energy term
    against its closed form, every kind (test/runtests.jl:295-317)."""
    from edg_westervelt_amd.precompute import boundary_side
    rng = np.random.default_rng(4)
    n = 1000
    c = 0.7 + 0.5 * rng.random(n); c2 = c ** 2; a = 1e-2 * rng.random(n); b = 0.3 * rng.random(n)
    v = 2 * rng.random(n) - 1; gu = rng.standard_normal(n); gv = rng.standard_normal(n)
    psi = c2 * gu + a * gv
    scale = np.max(psi ** 2 + v ** 2)
    L = ob.lib()
    L.orc_side_flux.restype = None
    kinds = {"absorbing_2beta": 0, "absorbing": 1, "dirichlet": 2, "wall": 3}
    d = C.c_double
    for kind in kinds:
        for s in (-1.0, 0.0, 0.5, 1.0):
            for kap in (0.0, 1.0):
                bs = boundary_side(kind, s, abs(s), kap)
                p = ob._Pre()
                p.kind[0] = kinds[kind]
                p.gam[0], p.eta[0], p.sig[0], p.kb[0], p.kap[0] = (bs["gamma_abc"], bs["eta_abc"], bs["sigma"],
                                                                   bs["kb"], bs["kappa"])
                F1 = np.empty(n); dv = np.empty(n); F1h = np.empty(n); dvh = np.empty(n)
                for i in range(n):
                    o1, o2 = d(), d()
                    L.orc_side_flux(C.byref(p), 0, d(c2[i]), d(a[i]), d(c[i]), d(b[i]), d(v[i]), d(gu[i]),
                                    d(gv[i]), d(gu[i]), d(gv[i]), d(v[i]), C.byref(o1), C.byref(o2))
                    F1[i], dv[i] = o1.value, o2.value
                    L.orc_side_flux(C.byref(p), 0, d(c2[i]), d(a[i]), d(c[i]), d(b[i]), d(v[i]), d(gu[i]),
                                    d(gv[i]), d(0.0), d(0.0), d(0.0), C.byref(o1), C.byref(o2))
                    F1h[i], dvh[i] = o1.value, o2.value
                assert np.array_equal(F1, psi) and np.all(dv == 0)
                Bq = psi * dvh + v * F1h
                if kind == "wall":
                    Bx = -(kap / c) * psi ** 2
                else:
                    g, e, sg = bs["gamma_abc"], bs["eta_abc"], bs["sigma"]
                    Wt = np.sqrt(1 - bs["kb"] * b * v)
                    Bx = (-e * (g - sg * e) * psi ** 2 - g * (e + sg * g) * (c * Wt * v) ** 2) / (c * Wt)
                assert np.abs(Bq - Bx).max() <= 1e-12 * scale
                assert Bq.max() <= 1e-12 * scale


def test_boundary_sides_are_local():
    """test/runtests.jl:319-338."""
    N = 4
    A4 = dict(bot="absorbing", top="absorbing", lft="absorbing", rgt="absorbing")
    S = dict(bot="absorbing_2beta", top="dirichlet", lft="wall", rgt="wall")

    def prep(bd):
        return precompute(3, N, boundary=bd, sigma=0.3, kappa=0.5, phase=(0.4, 0.3))[0]

    p0 = prep(A4)
    rng = np.random.default_rng(9)
    u, v = rand_state(rng, N, p0)

    def r(bd):
        pre = prep(bd)
        a, b, _ = ob.rhs(ob.OraclePre(pre, N), u, v, 0.037)
        return a, b

    r0 = r(A4)
    su, sv = r0[0].copy(), r0[1].copy()
    adj = {"bot": lambda j, k: j == 0, "top": lambda j, k: j == N - 1,
           "lft": lambda j, k: k == 0, "rgt": lambda j, k: k == N - 1}
    for s in ("bot", "top", "lft", "rgt"):
        bd = dict(A4)
        bd[s] = S[s]
        r1 = r(bd)
        for j in range(N):
            for k in range(N):
                dd = np.abs(r1[0][j, k] - r0[0][j, k]).max() + np.abs(r1[1][j, k] - r0[1][j, k]).max()
                assert (dd > 0) if adj[s](j, k) else (dd == 0)
        su += r1[0] - r0[0]
        sv += r1[1] - r0[1]
    rS = r(S)
    assert max(rel(rS[0], su), rel(rS[1], sv)) < 1e-12


def _energy_terms(u, v, ut, vt, pre, coef, N):
    """Terms of the semi-discrete energy identity (test/runtests.jl:71-106)."""
    h = 1.0 / N
    nu, nv, qh = pre.qu1, pre.qv1, pre.q_high
    z, w = GaussQCofs(nu)
    zh, wh = GaussQCofs(qh)
    W = np.outer(w, w); Wh = np.outer(wh, wh)
    Pu, Su, Pv, Sv, Ph = pre.P_u, pre.S_u, pre.P_v, pre.S_v, pre.P_h_v

    def crd(i, r):
        return ((i - 1) * h + i * h + h * r) / 2

    T1 = T2 = Dv = scale = 0.0
    for k in range(1, N + 1):
        for j in range(1, N + 1):
            X = np.broadcast_to(crd(k, z)[None, :], (nu, nu)); Y = np.broadcast_to(crd(j, z)[:, None], (nu, nu))
            Xh = np.broadcast_to(crd(k, zh)[None, :], (qh, qh)); Yh = np.broadcast_to(crd(j, zh)[:, None], (qh, qh))
            U = u[j - 1, k - 1]; Ud = ut[j - 1, k - 1]; V = v[j - 1, k - 1]; Vd = vt[j - 1, k - 1]
            c2 = np.vectorize(coef.c)(X, Y) ** 2
            al = np.vectorize(coef.alpha)(X, Y)
            bh = np.vectorize(coef.beta)(Xh, Yh)
            vq = Ph @ V @ Ph.T
            t1 = W * c2 * ((Pu @ U @ Su.T) * (Pu @ Ud @ Su.T) + (Su @ U @ Pu.T) * (Su @ Ud @ Pu.T))
            t2 = (h ** 2 / 4) * Wh * (1 - 2 * bh * vq) * vq * (Ph @ Vd @ Ph.T)
            T1 += t1.sum(); T2 += t2.sum(); scale += np.abs(t1).sum() + np.abs(t2).sum()
            Dv += np.sum(W * al * ((Pv @ V @ Sv.T) ** 2 + (Sv @ V @ Pv.T) ** 2))

    def E(n, x):
        return np.array([JacobiP(x, 0, 0, d) for d in range(n)])

    def E1(n, x):
        return np.array([GradJacobiP(x, 0, 0, d) for d in range(n)])

    DI = 0.0
    wf = (h / 2) * w
    for k in range(1, N + 1):
        for j in range(1, N + 1):
            for hz in (True, False):
                j2, k2 = (j + 1, k) if hz else (j, k + 1)
                if j2 > N or k2 > N:
                    continue

                def tr(Z, P, e):
                    return P @ (Z.T @ e) if hz else P @ (Z @ e)

                U1 = u[j - 1, k - 1]; V1 = v[j - 1, k - 1]; U2 = u[j2 - 1, k2 - 1]; V2 = v[j2 - 1, k2 - 1]
                jv = tr(V1, Pv, E(nv, 1.0)) - tr(V2, Pv, E(nv, -1.0))
                du1 = (2 / h) * tr(U1, Pu, E1(nu, 1.0)); du2 = (2 / h) * tr(U2, Pu, E1(nu, -1.0))
                dv1 = (2 / h) * tr(V1, Pv, E1(nv, 1.0)); dv2 = (2 / h) * tr(V2, Pv, E1(nv, -1.0))
                x1, y1, x2, y2 = crd(k, 0.0), crd(j, 0.0), crd(k2, 0.0), crd(j2, 0.0)
                a1 = coef.c(x1, y1) ** 2; a2 = coef.c(x2, y2) ** 2
                b1 = coef.alpha(x1, y1); b2 = coef.alpha(x2, y2)
                J = a1 * du1 - a2 * du2 + b1 * dv1 - b2 * dv2
                pen = 0.5 * (a1 + a2) * pre.lambda_u + 0.5 * (b1 + b2) * pre.lambda_v
                DI += np.sum(wf * (pen * jv ** 2 + pre.tau * J ** 2))
    return T1 + T2, Dv, DI, scale + Dv + DI


def test_energy_identity_with_jumps():
    """dE/dt = -int alpha |grad v|^2 - D_I (test/runtests.jl:739-758)."""
    co = quadrant_coefficients(0.5, 0.5, beta=(0.0, 0.3, 0.1, 0.2))
    N = 4
    worst = 0.0
    rise = -np.inf
    for q in (1, 2, 3, 4):
        for mixed in (True, False):
            for fp in (dict(theta=0.5, tau=0.0, lambda_u=0.0, lambda_v=0.0),
                       dict(theta=1.0, tau=0.0, lambda_u=0.0, lambda_v=0.0),
                       dict(theta=0.3, tau=0.0, lambda_u=0.0, lambda_v=0.0),
                       dict(theta=1.0, tau=1.0, lambda_u=1.0, lambda_v=1.0),
                       dict(theta=0.5, tau=1.0, lambda_u=0.0, lambda_v=0.0),
                       dict(theta=0.3, tau=0.7, lambda_u=0.4, lambda_v=0.9)):
                pre, _ = precompute(q, N, flux_params=fp, mixed=mixed, mode="source", boundary="wall", kappa=0.0,
                                    coefficients=co, source=GaussianSource(amplitude=0.0))
                rng = np.random.default_rng(10 * q + mixed)
                u, v = rand_state(rng, N, pre)
                ut, vt, _ = ob.rhs(ob.OraclePre(pre, N), u, v, 0.3)
                dEdt, Dv, DI, scale = _energy_terms(u, v, ut, vt, pre, co, N)
                if fp["tau"] == 0 and fp["lambda_u"] == 0 and fp["lambda_v"] == 0:
                    assert DI == 0
                worst = max(worst, abs(dEdt + Dv + DI) / scale)
                rise = max(rise, dEdt / scale)
    assert worst < 1e-12
    assert rise < 1e-12


def _project(f, pre, N, P=None):
    """L2 projection on the u- (or v-) space (test/runtests.jl:111-118)."""
    P = pre.P_u if P is None else P
    h = 1.0 / N
    z, w = GaussQCofs(pre.qu1)
    n = P.shape[1]
    U = np.zeros((N, N, n, n), order="F")
    for k in range(1, N + 1):
        for j in range(1, N + 1):
            xs = ((k - 1) * h + k * h + h * z) / 2
            ys = ((j - 1) * h + j * h + h * z) / 2
            F = np.vectorize(f)(xs[None, :], ys[:, None])
            U[j - 1, k - 1] = (P.T * w[None, :]) @ F @ (w[:, None] * P)
    return U


class _Step:
    """lo where the coordinate dir is below s, hi above (test/runtests.jl:28-32)."""

    def __init__(self, d, s, lo, hi):
        self.d, self.s, self.lo, self.hi = d, s, lo, hi

    def __call__(self, x, y):
        t = x if self.d == 1 else y
        return self.lo if t < self.s else self.hi


def test_exact_consistency_at_interface():
    """Piecewise-quadratic states that satisfy the transmission conditions give the
    exact time derivatives (test/runtests.jl:790-811); no reference implementation."""
    c1, c2, a1, a2 = 1.0, 0.6, 0.02, 0.005

    def prof(k1, k2):
        B = -k1 / k2
        A = 0.25 * (1 - B)
        return (lambda s: s ** 2 if s < 0.5 else A + B * (1 - s) ** 2,
                lambda s: 2 * k1 if s < 0.5 else 2 * k2 * B)

    wu, du = prof(c1 ** 2, c2 ** 2)
    wv, dv = prof(a1, a2)
    worst = 0.0
    z0 = lambda x, y: 0.0  # noqa: E731
    for d in (1, 2):
        for (q, N, mixed) in ((2, 4, False), (3, 4, True), (3, 6, False), (4, 8, True)):
            for fp in (dict(theta=1.0, tau=1.0, lambda_u=1.0, lambda_v=1.0),
                       dict(theta=0.5, tau=0.0, lambda_u=0.0, lambda_v=0.0),
                       dict(theta=1.0, tau=0.0, lambda_u=0.0, lambda_v=0.0),
                       dict(theta=0.0, tau=0.0, lambda_u=0.0, lambda_v=0.0),
                       dict(theta=0.3, tau=0.8, lambda_u=0.5, lambda_v=2.0)):
                co = Coefficients(_Step(d, 0.5, c1, c2), z0, z0, _Step(d, 0.5, a1, a2), z0, z0, z0)
                pre, _ = precompute(q, N, mixed=mixed, flux_params=fp, boundary="wall", kappa=0.7, coefficients=co,
                                    mode="source", source=GaussianSource(amplitude=0.0))

                def sel(x, y):
                    return x if d == 1 else y

                U = _project(lambda x, y: wu(sel(x, y)), pre, N)
                V = _project(lambda x, y: wv(sel(x, y)), pre, N, P=pre.P_v)
                Ue = _project(lambda x, y: wv(sel(x, y)), pre, N)
                Ve = _project(lambda x, y: du(sel(x, y)) + dv(sel(x, y)), pre, N, P=pre.P_v)
                ut, vt, _ = ob.rhs(ob.OraclePre(pre, N), U, V, 0.3)
                worst = max(worst, np.abs(ut - Ue).max() / np.abs(Ue).max(), np.abs(vt - Ve).max() / np.abs(Ve).max())
    assert worst < 1e-9


def test_errors_first_in_serial_order():
    """test/runtests.jl:548-577 on the oracle (both mass solvers)."""
    N = 8
    co = replace_coefficients(paper_coefficients(), alpha=1e-6, beta=1.0)
    bd = dict(bot="absorbing_2beta", top="absorbing", lft="absorbing", rgt="wall")
    for solver, mass_msg in (("chebyshev", "mass matrix not positive definite on element 13: 1 - 2βv_h ≤ 0"),
                             ("cg", "mass matrix not positive definite on element (5, 2): 1 - 2βv_h ≤ 0")):
        pre, _ = precompute(3, N, coefficients=co, boundary=bd, mass_solver=solver)
        op = ob.OraclePre(pre, N)
        u = np.zeros((N, N, pre.qu1, pre.qu1), order="F")
        v = np.zeros((N, N, pre.qv1, pre.qv1), order="F")
        v[0, N - 1, 0, 0] = 2.5
        v[3, 0, 0, 0] = 2.5
        wv = np.zeros_like(v)
        wv[3, 5, 0, 0] = 2.5
        wv[4, 1, 0, 0] = 2.5
        with pytest.raises(RuntimeError) as ei:
            ob.rhs(op, u, v, 0.0)
        assert str(ei.value) == "boundary weight 1 - 2βv_h ≤ 0 at a face point (absorbing_2beta side)"
        with pytest.raises(RuntimeError) as ei:
            ob.rhs(op, u, wv, 0.0)
        assert str(ei.value) == mass_msg


TABLE = [(q, pairing, flux) for q in range(1, 7) for pairing in ("mixed", "matched")
         for flux in ("upwind", "central", "alternating")]


@pytest.mark.parametrize("q,pairing,flux", TABLE)
def test_convergence_table(q, pairing, flux):
    """Every (q, pairing, flux) of data/convergence/results.csv at N = 5 (CG,
    T = 0.1): rtol 1e-9 plus 2e-13 absolute. The absolute part is the rounding
    accumulated by the O(1) solution over the run (the q = 6 errors, 1e-9..1e-6,
    differ from the table by up to 7e-14 for any summation order); most rows
    agree to 1e-11."""
    N = 5
    gold = convergence_row(q, pairing, flux, N)
    pre, geo = precompute(q, N, flux=flux, mixed=pairing == "mixed", mass_solver="cg")
    u, v, t, dt, ns = ob.solve_state(pre, N, 0.1, q)
    e = norms.mms_errors(u, v, 0.1, N, 1.0 / N, pre, geo)
    for k, g in gold.items():
        assert abs(e[k] - g) <= 1e-9 * g + 2e-13, (k, e[k], g)


def test_alpha_limit_golden():
    """:initial_value mode with the bump of experiments/alpha_limit/run.jl: the
    difference norms of alpha = 1e-5 against alpha = 0 (q_u = 6, q_v = 5, N = 20,
    T = 0.5) against data/alpha_limit/order.csv, rtol 1e-9."""
    N, q, T = 20, 6, 0.5
    res = {}
    for alpha in (0.0, 1e-5):
        pre, geo = precompute(q, N, flux="upwind", mixed=True, mode="initial_value", initial_profile=BUMP,
                              coefficients=replace_coefficients(paper_coefficients(), alpha=alpha))
        u, v, t, dt, ns = ob.solve_state(pre, N, T, q, nthreads=8)
        res[alpha] = (u, v)
    d = norms.difference_norms(*res[1e-5], *res[0.0], N, 1.0 / N, pre, geo)
    gold = dict(alpha_limit_rows())[1e-5]
    for k, g in gold.items():
        assert abs(d[k] / g - 1) < 1e-9, (k, d[k], g)
