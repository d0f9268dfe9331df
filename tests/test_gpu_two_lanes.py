"""The two-lane stage kernels (a pair of lanes per element, wedg_stage2.h)
against the CPU oracle and against the one-lane kernels, through the C ABI.

They cover every degree q_u = 1..6 with both pairings (q_v = q_u, q_u - 1),
coefficients constant on each element (the spectral u-solve) and the
Chebyshev mass solve; every test checks that the workspace really runs them
(``lanes == 2``). Odd mode counts leave the parity-1 lane a dummy row, and odd
point counts a middle point held by both lanes: both are covered. Tolerance: 1e-12 relative per
right-hand side against the oracle, as the one-lane parity tests; the two
device kernels group their sums differently, so they agree to rounding, not
bit for bit.
"""
import sys

import numpy as np
import pytest

import edg_westervelt_amd as w
from edg_westervelt_amd import precompute
from edg_westervelt_amd.coefficients import constant_coefficients, element_coefficients
from oracle import batched as ob
from tests.cases import MIXED, MIXED2, quadrant_coefficients, rand_state
from tests.conftest import rel

pytestmark = pytest.mark.gpu
TOL = 1e-12


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def device_rhs(pre, N, u, v, t, cheb="global", lanes=2):
    ws = w.Workspace(pre, N, cheb=cheb, lanes=lanes)
    assert ws.dev.lanes == lanes, ws.problem.info()
    ut, vt = w.edg_rhs_(ws, u, v, t, N, 1.0 / N, pre)
    return ut.copy(order="F"), vt.copy(order="F")


def oracle_rhs(pre, N, u, v, t, cheb="global"):
    return ob.rhs(ob.OraclePre(pre, N), u, v, t, cheb_mode=0 if cheb == "global" else 1)


def random_element_coefficients(rng, N):
    return element_coefficients(c=rng.uniform(0.8, 1.2, (N, N)), alpha=10 ** rng.uniform(-4, -2, (N, N)),
                                beta=rng.uniform(0.0, 0.05, (N, N)))


@pytest.mark.parametrize("mode", ["mms", "initial_value", "source"])
@pytest.mark.parametrize("boundary", ["absorbing_2beta", "absorbing", "dirichlet", "wall", "MIXED", "MIXED2"])
def test_two_lane_rhs_matrix(mode, boundary):
    """Every flux, q = 1..6, both pairings, meshes from one element to 7 x 7
    (element counts that leave part of the last wave empty), constant
    coefficients."""
    bnd = {"MIXED": MIXED, "MIXED2": MIXED2}.get(boundary, boundary)
    rng = np.random.default_rng(21)
    worst = 0.0
    for q in (1, 2, 3, 4, 5, 6):
        for N in (1, 2, 7):
            for flux, mixed in (("upwind", False), ("central", True), ("alternating", q % 2 == 0)):
                say(f"case two-lane rhs mode={mode} boundary={boundary} q={q} N={N} flux={flux} mixed={mixed}")
                pre, _ = precompute(q, N, flux=flux, mixed=mixed, mode=mode, boundary=bnd, sigma=0.3,
                                    sigma_dirichlet=0.6, kappa=0.7, phase=(0.4, 0.3),
                                    coefficients=constant_coefficients(c=0.9, alpha=3e-3, beta=0.02))
                u, v = rand_state(rng, N, pre)
                a = oracle_rhs(pre, N, u, v, 0.037)
                b = device_rhs(pre, N, u, v, 0.037)
                worst = max(worst, rel(a[0], b[0]), rel(a[1], b[1]))
    say(f"worst {worst:.3e}")
    assert worst < TOL


@pytest.mark.parametrize("cheb", ["global", "element"])
@pytest.mark.parametrize("mode", ["initial_value", "source"])
def test_two_lane_jumping_coefficients(mode, cheb):
    """Coefficients constant on each element and jumping across faces (the
    one-sided face values, src/solver.jl:742; :mms mode refuses jumps): per-element
    random fields and the quadrants of test/runtests.jl:51-62 (even N, so the
    jumps fall on element faces)."""
    rng = np.random.default_rng(22)
    worst = 0.0
    for q in (1, 2, 3, 4, 5, 6):
        for N in (4, 8):
            for co in (random_element_coefficients(rng, N), quadrant_coefficients(0.5, 0.5)):
                mixed = (q + N) % 2 == 1
                say(f"case two-lane jumps mode={mode} cheb={cheb} q={q} N={N} mixed={mixed}")
                pre, _ = precompute(q, N, mixed=mixed, mode=mode, boundary=MIXED, sigma=0.3,
                                    sigma_dirichlet=0.6, kappa=0.7, coefficients=co)
                u, v = rand_state(rng, N, pre)
                a = oracle_rhs(pre, N, u, v, 0.021, cheb)
                b = device_rhs(pre, N, u, v, 0.021, cheb)
                worst = max(worst, rel(a[0], b[0]), rel(a[1], b[1]))
    say(f"worst {worst:.3e}")
    assert worst < TOL


def test_two_lane_larger_meshes():
    """Meshes of several waves (32 elements per wave) with the XCD block order
    (grids of a multiple of 8 waves), against the oracle and the one-lane kernel."""
    rng = np.random.default_rng(23)
    worst = worst1 = 0.0
    for q, N, mixed in ((5, 16, False), (3, 32, False), (5, 40, True), (1, 64, False), (4, 24, False),
                        (6, 16, False), (6, 24, True), (2, 40, True)):
        say(f"case two-lane large q={q} N={N} mixed={mixed}")
        pre, _ = precompute(q, N, mixed=mixed, mode="initial_value", boundary=MIXED2, sigma=0.2, kappa=0.5,
                            coefficients=random_element_coefficients(rng, N))
        u, v = rand_state(rng, N, pre, su=1e-2, sv=1e-2)
        a = oracle_rhs(pre, N, u, v, 0.0)
        b = device_rhs(pre, N, u, v, 0.0)
        c = device_rhs(pre, N, u, v, 0.0, lanes=1)
        worst = max(worst, rel(a[0], b[0]), rel(a[1], b[1]))
        worst1 = max(worst1, rel(c[0], b[0]), rel(c[1], b[1]))
    say(f"worst vs oracle {worst:.3e}, vs one lane {worst1:.3e}")
    assert worst < TOL and worst1 < TOL


@pytest.mark.parametrize("cheb", ["global", "element"])
def test_two_lane_rk4_steps(cheb):
    """Several RK4 steps (every stage kind, the epilogue's next-stage counts)
    against the oracle's RK4 and the one-lane kernel."""
    rng = np.random.default_rng(24)
    for q, N, mixed in ((3, 6, False), (5, 9, False), (4, 7, False), (6, 5, True), (2, 6, True)):
        say(f"case two-lane rk4 cheb={cheb} q={q} N={N} mixed={mixed}")
        pre, _ = precompute(q, N, mixed=mixed, mode="source", boundary=MIXED, sigma=0.3, sigma_dirichlet=0.6,
                            kappa=0.7, coefficients=random_element_coefficients(rng, N))
        u, v = rand_state(rng, N, pre, su=1e-3, sv=1e-2)
        dt, nst = 2e-4, 5
        ou, ov, _, _ = ob.rk4_steps(ob.OraclePre(pre, N), u, v, 0.01, dt, nst, cheb_mode=0 if cheb == "global" else 1)
        res = {}
        for lanes in (1, 2):
            ws = w.Workspace(pre, N, cheb=cheb, lanes=lanes)
            assert ws.dev.lanes == lanes
            uu, vv = np.array(u, order="F"), np.array(v, order="F")
            t = 0.01
            for _ in range(nst):
                w.rk4_step_(uu, vv, t, dt, N, 1.0 / N, pre, ws)
                t += dt
            res[lanes] = (uu, vv)
        e2 = max(rel(ou, res[2][0]), rel(ov, res[2][1]))
        e12 = max(rel(res[1][0], res[2][0]), rel(res[1][1], res[2][1]))
        say(f"vs oracle {e2:.3e}, vs one lane {e12:.3e}")
        assert e2 < TOL and e12 < TOL


def _advance(pre, N, u, v, t0, dt, nsteps, lanes, cheb="global"):
    import ctypes as C
    from edg_westervelt_amd._lib import lib
    from edg_westervelt_amd.solver import _ptr
    ws = w.Workspace(pre, N, cheb=cheb, lanes=lanes)
    assert ws.dev.lanes == lanes
    uu, vv = np.array(u, order="F"), np.array(v, order="F")
    ws.dev.check(lib.wedg_set_state(ws.dev.h, _ptr(uu), _ptr(vv)))
    t = C.c_double(0.0)
    ws.dev.check(lib.wedg_advance(ws.dev.h, t0, dt, nsteps, 0, C.byref(t)))
    ws.dev.check(lib.wedg_get_state(ws.dev.h, _ptr(uu), _ptr(vv)))
    mean, stages = C.c_double(0.0), C.c_long(0)
    lib.wedg_cheb_stats(ws.dev.h, C.byref(mean), C.byref(stages))
    return uu, vv, ws.dev.last_error().cheb_steps, mean.value


@pytest.mark.parametrize("cheb", ["global", "element"])
def test_two_lane_advance_matches_one_lane(cheb):
    """wedg_advance over 40 steps with beta = 2 (counts carried from stage to
    stage through the per-wave slots): the same run to rounding with both
    kernels, the same largest Chebyshev count."""
    from edg_westervelt_amd.batch import gaussian_profile
    N, q = 24, 5
    pre, _ = precompute(q, N, mixed=False, mode="initial_value", boundary="absorbing_2beta",
                        coefficients=constant_coefficients(c=1.0, alpha=1e-4, beta=2.0),
                        initial_profile=gaussian_profile(0.5, 0.5, 0.1, 0.05))
    u = np.zeros((N, N, pre.qu1, pre.qu1), order="F")
    v = np.zeros((N, N, pre.qv1, pre.qv1), order="F")
    say(f"case two-lane advance cheb={cheb}")
    a = _advance(pre, N, u, v, 0.0, 1.0 / (N * 250), 40, 1, cheb)
    b = _advance(pre, N, u, v, 0.0, 1.0 / (N * 250), 40, 2, cheb)
    e = max(rel(a[0], b[0]), rel(a[1], b[1]))
    say(f"two-lane vs one-lane after 40 steps {e:.3e}; max counts {a[2]} {b[2]}, mean {a[3]:.3f} {b[3]:.3f}")
    assert e < 1e-11
    assert a[2] == b[2] and a[2] > 1
    assert rel(a[1], v) > 1e-6


def test_two_lane_full_size():
    """The 256 x 256, q = 5 instance of the benchmark: one right-hand side
    against the oracle, and 10 RK4 steps against the one-lane kernel."""
    from edg_westervelt_amd.batch import gaussian_profile
    q, N = 5, 256
    pre, _ = precompute(q, N, mixed=False, mode="initial_value", boundary="absorbing_2beta",
                        coefficients=constant_coefficients(c=1.0, alpha=1e-4, beta=2.0),
                        initial_profile=gaussian_profile(0.5, 0.5, 0.05, 0.05))
    rng = np.random.default_rng(26)
    u, v = rand_state(rng, N, pre, su=1e-3, sv=1e-3)
    say("case two-lane full size rhs")
    a = oracle_rhs(pre, N, u, v, 0.0)
    b = device_rhs(pre, N, u, v, 0.0)
    e = max(rel(a[0], b[0]), rel(a[1], b[1]))
    say(f"rhs vs oracle {e:.3e}")
    assert e < TOL
    say("case two-lane full size 10 steps")
    z = np.zeros_like(u), np.zeros_like(v)
    r1 = _advance(pre, N, z[0], z[1], 0.0, 1.5625e-5, 10, 1)
    r2 = _advance(pre, N, z[0], z[1], 0.0, 1.5625e-5, 10, 2)
    e2 = max(rel(r1[0], r2[0]), rel(r1[1], r2[1]))
    say(f"10 steps two-lane vs one-lane {e2:.3e}")
    assert e2 < 1e-11


def test_two_lane_errors_in_serial_order():
    """The reference's failures with the two-lane kernel: the boundary weight
    of an absorbing_2beta side and the mass matrix (src/rhs.jl:112-116, 916),
    the boundary error first when both fail in one stage (src/rhs.jl:530-531)."""
    from edg_westervelt_amd._lib import WedgError
    N, q = 4, 3
    for boundary, amp, expect in (("absorbing_2beta", 40.0, "boundary weight 1 - 2βv_h ≤ 0"),
                                  ("wall", 40.0, "mass matrix not positive definite")):
        pre, _ = precompute(q, N, mixed=False, mode="source", boundary=boundary,
                            coefficients=constant_coefficients(c=1.0, alpha=1e-4, beta=0.5))
        u = np.zeros((N, N, pre.qu1, pre.qu1), order="F")
        v = np.zeros((N, N, pre.qv1, pre.qv1), order="F")
        v[:, :, 0, 0] = amp
        msgs = []
        for lanes in (1, 2):
            with pytest.raises(WedgError) as ex:
                device_rhs(pre, N, u, v, 0.0, lanes=lanes)
            msgs.append(str(ex.value))
        say(f"case two-lane failure {boundary}: {msgs[1]}")
        assert msgs[0] == msgs[1] and expect in msgs[1]


def test_two_lane_graph_chunks_equal_plain_launches():
    """RK4 chunks replayed as HIP graphs (wedg_graph_enable) with the two-lane
    kernels: bit for bit the plain two-lane launches, over three replays with
    plain work of the same workspace in between."""
    import ctypes as C
    from edg_westervelt_amd import _lib
    from edg_westervelt_amd._lib import lib
    from edg_westervelt_amd.solver import _ptr
    co = constant_coefficients(c=1.0, alpha=1e-4, beta=1e-3)
    for q, N, mixed, mode, chunk, n in ((5, 12, False, "source", 7, 30), (4, 9, True, "initial_value", 4, 17),
                                        (6, 8, False, "source", 16, 16)):
        say(f"case two-lane graph q={q} N={N} mixed={mixed} {mode} chunk={chunk}")
        pre, _ = precompute(q, N, mixed=mixed, mode=mode, boundary=MIXED, sigma=0.3, sigma_dirichlet=0.6,
                            kappa=0.7, coefficients=co)
        u, v = rand_state(np.random.default_rng(70 + q), N, pre, su=1e-2, sv=2e-3)
        ws = w.Workspace(pre, N, lanes=2)
        assert ws.dev.lanes == 2

        def run(k):
            assert lib.wedg_graph_enable(ws.dev.h, k) == _lib.OK
            uu, vv = np.array(u, order="F", copy=True), np.array(v, order="F", copy=True)
            ws.dev.check(lib.wedg_set_state(ws.dev.h, _ptr(uu), _ptr(vv)))
            t = C.c_double(0.0)
            ws.dev.check(lib.wedg_advance(ws.dev.h, 0.01, 1e-4, n, 0, C.byref(t)))
            ws.dev.check(lib.wedg_get_state(ws.dev.h, _ptr(uu), _ptr(vv)))
            return uu, vv, t.value

        ref = run(0)
        for rep in range(3):
            got = run(chunk)
            assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1]) and got[2] == ref[2], rep
            w.edg_rhs_(ws, u, v, 0.3, N, 1.0 / N, pre)
        assert ws.dev.lanes == 2


def test_two_lane_source_mode_solve():
    """solve(mode = :source) with two lanes per element against one lane:
    receivers after every step, the record-step diagnostics, the slice and
    the snapshots agree to 1e-11 relative (the kernels sum in different
    orders); q_u = 4 and 6, the steepening coefficients (beta = 8)."""
    for q, N, mixed in ((4, 16, True), (6, 8, False)):
        say(f"case two-lane source solve q={q} N={N}")
        kw = dict(mixed=mixed, mode="source", boundary={"bot": "absorbing_2beta", "top": "wall",
                                                        "lft": "absorbing", "rgt": "absorbing_2beta"},
                  kappa=0.5, coefficients=constant_coefficients(c=1.0, alpha=1e-6, beta=8.0),
                  source=w.GaussianSource(x0=(0.3, 0.55), r0=0.1, amplitude=0.5),
                  record_every=5, slice_y=0.3, receivers=[(0.45, 0.5), (0.3, 0.55)], snap_times=(0.01,))
        A = w.solve(q, N, 0.03, lanes=1, **kw)
        B = w.solve(q, N, 0.03, lanes=2, **kw)
        assert list(A.t) == list(B.t) and np.abs(A.recv).max() > 0
        assert rel(A.recv, B.recv) < 1e-11 and rel(A.slices, B.slices) < 1e-11
        for k in ("energy", "vmax", "ndmin"):
            assert rel(getattr(A, k), getattr(B, k)) < 1e-11, k
        assert len(A.snaps) == len(B.snaps) == 1 and rel(A.snaps[0][1], B.snaps[0][1]) < 1e-11


def test_auto_lanes():
    """Workspace(lanes="auto"): two lanes for two-lane problems with q_u >= 3
    on meshes that leave SIMDs idle with one lane, and always at q_u = 6;
    one lane otherwise (q_u <= 2, the full-size q_u = 5 mesh, the general
    layout)."""
    co = constant_coefficients(c=1.0, alpha=1e-4, beta=0.5)
    for q, N, coef, expect in ((5, 16, co, 2), (3, 24, co, 2), (2, 16, co, 1), (6, 8, co, 2), (5, 256, co, 1),
                               (5, 8, None, 1)):
        pre, _ = precompute(q, N, mixed=False, mode="initial_value", coefficients=coef)
        ws = w.Workspace(pre, N, lanes="auto")
        say(f"case auto lanes q={q} N={N} general={coef is None}: {ws.dev.lanes}")
        assert ws.dev.lanes == expect
