"""Host table production (precompute) against closed forms and the reference's
known answers (no GPU): quadrature (test/runtests.jl:156-163), one-sided face
values (test/runtests.jl:644-698), argument checks (test/runtests.jl:359-366),
and the layouts the device library reads."""
import math

import numpy as np
import pytest

from edg_westervelt_amd import constant_coefficients, paper_coefficients, precompute
from edg_westervelt_amd.coefficients import Coefficients
from edg_westervelt_amd.quadrature import GaussQCofs, GradJacobiP, JacobiP
from tests.cases import MIXED, quadrant_coefficients


def test_quadrature_closed_forms():
    x, w = GaussQCofs(5)
    assert abs(w.sum() - 2) < 1e-14
    assert abs(np.sum(w * x ** 8) - 2 / 9) < 1e-14
    G = np.array([[np.sum(w * JacobiP(x, 0, 0, m) * JacobiP(x, 0, 0, n)) for n in range(5)] for m in range(5)])
    assert np.abs(G - np.eye(5)).max() < 1e-13
    for q in range(2, 11):
        xg, wg = GaussQCofs(q)
        xl, wl = np.polynomial.legendre.leggauss(q)
        assert np.abs(xg - xl).max() < 1e-14 and np.abs(wg - wl).max() < 1e-14


def test_legendre_values_and_derivatives():
    """L_n(1) = sqrt((2n+1)/2), L_n(-1) = (-1)^n L_n(1), L_n'(1) = sqrt((2n+1)/2) n(n+1)/2."""
    for n in range(8):
        a = math.sqrt((2 * n + 1) / 2)
        assert abs(JacobiP(1.0, 0, 0, n) - a) < 1e-13
        assert abs(JacobiP(-1.0, 0, 0, n) - (-1) ** n * a) < 1e-13
        assert abs(GradJacobiP(1.0, 0, 0, n) - a * n * (n + 1) / 2) < 1e-12
    # derivative against a central difference
    xs = np.linspace(-0.9, 0.9, 7)
    for n in range(1, 7):
        fd = (JacobiP(xs + 1e-6, 0, 0, n) - JacobiP(xs - 1e-6, 0, 0, n)) / 2e-6
        assert np.abs(GradJacobiP(xs, 0, 0, n) - fd).max() < 1e-7


def test_q_high_and_shapes():
    for q in range(1, 7):
        for mixed in (True, False):
            pre, geo = precompute(q, 3, mixed=mixed)
            qv = q - 1 if mixed else q
            assert pre.q_high == max(math.ceil(3 * qv / 2) + 1, q + 1)
            assert pre.Cw.shape == (q + 1, 9, q + 1) and pre.Cw.flags.f_contiguous
            nu2 = (q + 1) ** 2
            assert pre.Ainv_u.shape == (nu2 * (nu2 + 1) // 2, 9)


def test_packed_inverse_is_the_inverse():
    """Ainv_u holds the upper triangle of (pinned c^2 stiffness)^-1 (src/solver.jl:528-535)."""
    pre, _ = precompute(3, 4)
    n = pre.qu1 ** 2
    Pxu = np.kron(pre.P_u, pre.S_u); Peu = np.kron(pre.S_u, pre.P_u)
    for e in (0, 5, 15):
        cw = pre.Cw[:, e, :].ravel(order="F")
        K = Pxu.T @ (cw[:, None] * Pxu) + Peu.T @ (cw[:, None] * Peu)
        K[0, 0] += 1.0
        A = np.zeros((n, n))
        for jj in range(n):
            for i in range(jj + 1):
                A[i, jj] = A[jj, i] = pre.Ainv_u[jj * (jj + 1) // 2 + i, e]
        assert np.abs(A @ K - np.eye(n)).max() < 1e-10


def test_face_values_continuous_are_point_values():
    """Continuous coefficients: the plain point values, bit for bit (test/runtests.jl:645-663)."""
    for (q, N, kw) in ((2, 4, {}), (3, 5, dict(mode="initial_value", phase=(0.4, 0.3))),
                       (1, 3, dict(coefficients=constant_coefficients(c=1.0, alpha=0.0, beta=0.0)))):
        pre, _ = precompute(q, N, **kw)
        co = kw.get("coefficients", paper_coefficients())
        h = 1.0 / N
        z, _ = GaussQCofs(q + 1)

        def crd(i):
            return ((i - 1) * h + i * h + h * z) / 2

        for k in range(1, N + 1):
            for jf in range(1, N + 2):
                xf = crd(k); yf = np.full_like(xf, (jf - 1) * h)
                ev = lambda f, x, y: np.array([float(f(a, b)) for a, b in zip(x, y)])  # noqa: E731
                assert np.array_equal(pre.hc2[:, jf - 1, k - 1], ev(co.c, xf, yf) ** 2)
                assert np.array_equal(pre.haf[:, jf - 1, k - 1], ev(co.alpha, xf, yf))
                assert np.array_equal(pre.hbf[:, jf - 1, k - 1], ev(co.beta, xf, yf))
                assert np.array_equal(pre.vc2[:, k - 1, jf - 1], ev(co.c, yf, xf) ** 2)
        assert np.array_equal(pre.hc2m, pre.hc2) and np.array_equal(pre.vc2m, pre.vc2)
        assert np.all(pre.hJC == 0) and np.all(pre.vJC == 0) and not pre.one_sided


class _Step:
    """test/runtests.jl:28-32: lo below s, hi above, on on the line, edge at 0."""

    def __init__(self, d, s, lo, hi, on=None, edge=None):
        self.d, self.s, self.lo, self.hi = d, s, lo, hi
        self.on = hi if on is None else on
        self.edge = lo if edge is None else edge

    def __call__(self, x, y):
        t = x if self.d == 1 else y
        return self.edge if t == 0 else (self.lo if t < self.s else (self.on if t == self.s else self.hi))


class _Spike:
    def __init__(self, g):
        self.g = g

    def __call__(self, x, y):
        return self.g if x == 0.5 else (1.0 if x < 0.25 else 1.5)


def _z0(x, y):
    return 0.0


def test_face_values_at_jumps():
    """test/runtests.jl:664-691: one-sided values whichever piece owns the line."""
    def step_coef(c, beta=None):
        beta = beta or _Step(1, 2.0, 0.0, 0.0)
        return Coefficients(c, _z0, _z0, _Step(1, 2.0, 1e-3, 1e-3), _z0, _z0, beta)

    for (N, kf, cfun) in ((4, 3, _Step(1, 0.5, 1.0, 2.0)), (4, 3, _Step(1, 0.5, 1.0, 2.0, on=1.0)),
                          (4, 3, _Step(1, 0.5, 1.0, 2.0, on=7.0)), (10, 4, _Step(1, 0.3, 1.0, 2.0)),
                          (10, 4, _Step(1, 0.3, 1.0, 2.0, on=1.0))):
        pre, _ = precompute(2, N, mode="source", coefficients=step_coef(cfun))
        rest = [i for i in range(N + 1) if i != kf - 1]
        assert np.all(pre.vc2m[:, :, kf - 1] == 1.0) and np.all(pre.vc2[:, :, kf - 1] == 4.0)
        assert np.array_equal(pre.vc2m[:, :, rest], pre.vc2[:, :, rest]) and np.array_equal(pre.hc2m, pre.hc2)
        assert pre.one_sided
    pre, _ = precompute(2, 4, mode="source", coefficients=step_coef(_Step(1, 0.5, 1.0, 2.0, edge=2.0),
                                                                     beta=_Step(2, 2.0, 0.1, 0.1, edge=5.0)))
    assert np.all(pre.vc2[:, :, 0] == 1.0) and np.all(pre.vcf[:, :, 0] == 1.0) and np.all(pre.hbf[:, 0, :] == 0.1)
    for g in (1e8, np.inf, np.nan):
        pre, _ = precompute(2, 8, mode="source", coefficients=step_coef(_Spike(g)))
        assert np.all(pre.vc2m[:, :, 2] == 1.0) and np.all(pre.vc2[:, :, 2] == 2.25)
        assert np.all(pre.vc2m[:, :, 4] == 2.25) and np.all(pre.vc2[:, :, 4] == 2.25)
        assert np.all(np.isfinite(pre.vc2m)) and np.all(np.isfinite(pre.vc2)) and pre.one_sided
    pre, _ = precompute(2, 4, mode="source", coefficients=quadrant_coefficients(0.5, 0.5, c=(1.0, 1.0, 1.0, 1.0)))
    assert pre.one_sided and np.array_equal(pre.hc2m, pre.hc2) and not np.array_equal(pre.hafm, pre.haf)
    pre, _ = precompute(2, 4, mode="source", coefficients=quadrant_coefficients(
        0.5, 0.5, c=(1.0, 1.0, 1.0, 1.0), alpha=(1e-3, 1e-3, 1e-3, 1e-3)))
    assert not pre.one_sided


def test_argument_checks():
    """test/runtests.jl:359-366, 826-828 (same messages)."""
    with pytest.raises(ValueError, match="sigma_dirichlet"):
        precompute(2, 2, boundary="dirichlet", sigma=-0.5)
    with pytest.raises(ValueError):
        precompute(2, 2, boundary="dirichlet", sigma_dirichlet=-0.1)
    with pytest.raises(ValueError, match="-1 ≤ sigma ≤ 1"):
        precompute(2, 2, boundary="absorbing", sigma=1.5)
    with pytest.raises(ValueError, match="kappa"):
        precompute(2, 2, boundary="wall", kappa=-1.0)
    with pytest.raises(ValueError, match="unknown boundary"):
        precompute(2, 2, boundary="neumann")
    with pytest.raises(ValueError, match="NamedTuple"):
        precompute(2, 2, boundary=dict(bot="wall", top="wall", lft="wall"))
    with pytest.raises(ValueError, match="unknown mass_solver"):
        precompute(2, 2, mass_solver="lu")
    with pytest.raises(ValueError, match="continuous across the element faces"):
        precompute(2, 6, coefficients=quadrant_coefficients(0.5, 0.5, c=(1.0, 0.7, 0.8, 0.6),
                                                            alpha=(1e-3, 1e-3, 2e-3, 2e-3),
                                                            beta=(1e-4, 0.0, 1e-4, 0.0), smooth=True))
    with pytest.raises(ValueError, match="initial_profile"):
        precompute(2, 2, initial_profile=dict(u0=_z0, u0_x=_z0, u0_y=_z0, u0_xx=_z0, u0_yy=_z0))
    pre, _ = precompute(2, 2, boundary=MIXED, sigma=0.3, sigma_dirichlet=0.6, kappa=0.7)
    bc = pre.bc
    assert (bc["bot"]["sigma"], bc["bot"]["kb"], bc["rgt"]["sigma"], bc["rgt"]["kb"]) == (0.3, 2.0, 0.3, 1.0)
    assert (bc["top"]["gamma_abc"], bc["top"]["sigma"], bc["lft"]["kind"], bc["lft"]["kappa"]) == (1.0, 0.6, "wall", 0.7)


def test_initial_profile_equals_default():
    """Passing Phi as the initial profile reproduces the default tables (test/runtests.jl:627-638)."""
    pi2 = 2 * math.pi
    prof = dict(u0=lambda x, y: np.cos(pi2 * x) * np.sin(pi2 * y),
                u0_x=lambda x, y: -pi2 * np.sin(pi2 * x) * np.sin(pi2 * y),
                u0_y=lambda x, y: pi2 * np.cos(pi2 * x) * np.cos(pi2 * y),
                u0_xx=lambda x, y: -(pi2) ** 2 * np.cos(pi2 * x) * np.sin(pi2 * y),
                u0_yy=lambda x, y: -(pi2) ** 2 * np.cos(pi2 * x) * np.sin(pi2 * y))
    A, _ = precompute(3, 6, mode="initial_value")
    B, _ = precompute(3, 6, mode="initial_value", initial_profile=prof)
    for name in ("PiCdiv", "bcB_sp", "bcB_spy", "bcL_spx", "hJC"):
        assert np.allclose(getattr(A, name), getattr(B, name), rtol=1e-13, atol=1e-12), name


def test_element_coefficient_arrays():
    """Per-element arrays: Cw, Aw and beta_hq carry each element's own value, every
    interior face takes the value of the element on each side (bot/lft = the
    element above/right, the M arrays the one below/left) and the jumps make
    the faces one-sided (src/solver.jl:371-408, 716-770)."""
    from edg_westervelt_amd import element_coefficients
    rng = np.random.default_rng(4)
    N, q = 6, 3
    c = rng.uniform(0.8, 1.2, (N, N))
    alpha = 10.0 ** rng.uniform(-6, -3, (N, N))
    beta = rng.uniform(0, 8, (N, N))
    pre, _ = precompute(q, N, mixed=True, mode="initial_value", coefficients=element_coefficients(
        c=c, alpha=alpha, beta=beta))
    assert pre.one_sided
    nq = pre.qu1
    w = pre.w
    Cw = np.asarray(pre.Cw).reshape((nq, N, N, nq), order="F")
    Aw = np.asarray(pre.Aw).reshape((nq, N, N, nq), order="F")
    nh = pre.q_high
    Bq = np.asarray(pre.beta_hq).reshape((nh, N, N, nh), order="F")
    for j in range(N):
        for k in range(N):
            assert np.allclose(Cw[:, j, k, :], c[j, k] ** 2 * np.outer(w, w), rtol=1e-14, atol=0)
            assert np.allclose(Aw[:, j, k, :], alpha[j, k] * np.outer(w, w), rtol=1e-14, atol=0)
            assert (Bq[:, j, k, :] == beta[j, k]).all()
    hc2, hc2m = np.asarray(pre.hc2), np.asarray(pre.hc2m)
    vc2, vc2m = np.asarray(pre.vc2), np.asarray(pre.vc2m)
    for jf in range(1, N):
        for k in range(N):
            assert np.allclose(hc2[:, jf, k], c[jf, k] ** 2, rtol=1e-14, atol=0)
            assert np.allclose(hc2m[:, jf, k], c[jf - 1, k] ** 2, rtol=1e-14, atol=0)
    for j in range(N):
        for kf in range(1, N):
            assert np.allclose(vc2[:, j, kf], c[j, kf] ** 2, rtol=1e-14, atol=0)
            assert np.allclose(vc2m[:, j, kf], c[j, kf - 1] ** 2, rtol=1e-14, atol=0)
    # numbers are constants, the same as constant_coefficients
    A, _ = precompute(q, N, mode="initial_value", coefficients=element_coefficients(c=1.0, alpha=1e-4, beta=2.0))
    B, _ = precompute(q, N, mode="initial_value", coefficients=constant_coefficients(c=1.0, alpha=1e-4, beta=2.0))
    assert np.array_equal(A.Cw, B.Cw) and np.array_equal(A.hc2, B.hc2) and A.one_sided == B.one_sided
    with pytest.raises(ValueError):
        element_coefficients(c=np.ones((3, 4)), alpha=0.0, beta=0.0)
