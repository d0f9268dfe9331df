"""Source mode on the device (src/solver.jl:856-958): the diagnostics kernels
against the oracle on the same state, a short source-mode run against the
oracle's loop, and the three wave-steepening runs of
experiments/steepening/run.jl at full length against the reference's series,
receiver and snapshot data (data/steepening/*).
"""
import ctypes as C
import math
import sys
from types import SimpleNamespace

import numpy as np
import pytest

import edg_westervelt_amd as w
from edg_westervelt_amd import _lib, precompute
from edg_westervelt_amd._lib import lib
from oracle import batched as ob
from oracle import diagnostics as od
from tests.cases import (MIXED, STEEPENING, STEEPENING_RUN, quadrant_coefficients, rand_state,
                         steepening_grid, steepening_recv, steepening_series, steepening_snapshot)
from tests.conftest import rel

pytestmark = pytest.mark.gpu
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _p(a):
    return a.ctypes.data_as(_dp)


@pytest.mark.parametrize("q,N,mixed", [(3, 6, True), (2, 5, False), (5, 4, True)])
def test_diagnostics_device(q, N, mixed):
    """k_diag and k_recv on a random state with piecewise c, beta: the element
    energies, max |v_h|, min (1 - 2 beta v_h), the slice and the snapshot grid
    against the oracle's restatement (1e-13 relative), and the receivers over
    three device RK4 steps against the oracle's steps (1e-12)."""
    say(f"case diagnostics q={q} N={N} mixed={mixed}")
    rng = np.random.default_rng(100 + q)
    co = quadrant_coefficients(0.5, 0.5, c=(1.0, 0.7, 0.8, 0.6), alpha=(1e-3, 1e-3, 2e-3, 2e-3),
                               beta=(0.9, 0.5, 0.3, 0.7), smooth=True)
    pre, geo = precompute(q, N, mixed=mixed, mode="source", boundary=MIXED, coefficients=co)
    u, v = rand_state(rng, N, pre, su=1e-3, sv=5e-3)
    ws = w.Workspace(pre, N)
    dh = ws.dev.h
    ws.dev.check(lib.wedg_set_state(dh, _p(u), _p(v)))
    c2 = np.asfortranarray(geo.err_c2); bq = np.asfortranarray(geo.err_beta); W2d = np.asfortranarray(geo.W2d)
    ws.dev.check(lib.wedg_problem_set_diag(ws.problem.h, _p(c2), _p(bq), _p(W2d)))
    E = np.zeros(N * N); vm = np.zeros(N * N); nd = np.zeros(N * N); sl = np.zeros(N * pre.qu1)
    for slice_y in (0.55, 0.0, 1.0):
        j0, eta0 = od.locate(slice_y, N)
        r0 = od.vrow(eta0, pre.qv1)
        ws.dev.check(lib.wedg_diagnostics(dh, _p(E), _p(vm), _p(nd), j0 - 1, _p(r0), _p(sl)))
        Eo, vmo, ndo = od.energy_elements(u, v, N, 1.0 / N, pre, geo)
        assert rel(Eo, E) < 1e-13 and rel(vmo, vm) < 1e-13 and rel(ndo, nd) < 1e-13
        assert rel(od.slice_values(v, N, pre, slice_y), sl) < 1e-13
    G = np.zeros((N * pre.qu1, N * pre.qu1), order="F")
    ws.dev.check(lib.wedg_snapshot(dh, _p(G)))
    assert rel(od.grid(v, N, pre), G) < 1e-13
    rc = [(0.45, 0.5), (0.05, 0.93), (1.0, 0.0), (0.71, 0.28)]
    rj = np.zeros(len(rc), dtype=np.int32); rk = np.zeros(len(rc), dtype=np.int32)
    rows = np.zeros((len(rc), 2 * pre.qv1))
    for i, (x, y) in enumerate(rc):
        kr, xi = od.locate(x, N); jr, eta = od.locate(y, N)
        rj[i], rk[i] = jr - 1, kr - 1
        rows[i] = np.concatenate([od.vrow(eta, pre.qv1), od.vrow(xi, pre.qv1)])
    dt = 1e-4
    rec = np.zeros(3 * len(rc))
    tend = C.c_double(0.0)
    ws.dev.check(lib.wedg_advance_receivers(dh, 0.2, dt, 3, len(rc), rj.ctypes.data_as(_ip), rk.ctypes.data_as(_ip),
                                            _p(rows), 1, _p(rec), C.byref(tend)))
    op = ob.OraclePre(pre, N)
    uo, vo, t = u, v, 0.2
    for s in range(3):
        uo, vo, _, _ = ob.rk4_steps(op, uo, vo, t, dt, 1)
        t += dt
        assert rel(od.receiver_values(vo, N, pre, rc), rec[s * len(rc):(s + 1) * len(rc)]) < 1e-12, s
    assert tend.value == t
    ud = np.zeros_like(u); vd = np.zeros_like(v)
    ws.dev.check(lib.wedg_get_state(dh, _p(ud), _p(vd)))
    assert rel(uo, ud) < 1e-12 and rel(vo, vd) < 1e-12


def test_source_mode_short_run():
    """solve(mode = :source) against the oracle's restatement of the loop
    (q = 2, N = 8, T = 0.05: 16 steps, record_every = 5, two snapshot times
    inside one step): every series 1e-12 relative."""
    say("case short source run")
    q, N, T = 2, 8, 0.05
    kw = dict(mixed=True, mode="source", boundary="absorbing_2beta",
              coefficients=w.constant_coefficients(c=1.0, alpha=1e-3, beta=4.0),
              source=w.GaussianSource(x0=(0.3, 0.55), r0=0.1, amplitude=2.0))
    run = dict(record_every=5, slice_y=0.3, receivers=[(0.45, 0.5), (0.3, 0.55), (0.99, 0.01)],
               snap_times=(0.011, 0.01, 0.03))
    H = w.solve(q, N, T, **kw, **run)
    pre, geo = precompute(q, N, **kw)
    dt, nsteps = w.time_step(q, N, T)
    O = od.run_source(pre, geo, N, dt, nsteps, **run)
    assert list(H.t) == O["t"] and list(H.recv_t) == O["recv_t"]
    assert rel(np.array(O["recv"]), H.recv) < 1e-12
    assert rel(np.array(O["slices"]).T, H.slices) < 1e-12
    for k in ("energy", "vmax", "ndmin"):
        assert rel(np.array(O[k]), getattr(H, k)) < 1e-12, k
    assert [s[0] for s in H.snaps] == [s[0] for s in O["snaps"]] and len(H.snaps) == 3
    for (_, a), (_, b) in zip(O["snaps"], H.snaps):
        assert rel(a, b) < 1e-12
    assert H.dt == dt and H.receivers == [tuple(r) for r in run["receivers"]]


def steepening_kwargs(case):
    """The arguments of solve for experiments/steepening/run.jl:21-31, case <case>."""
    beta, bnd = STEEPENING[case]
    R = STEEPENING_RUN
    return dict(q=R["q"], N=R["N"], T=R["T"], flux="upwind", mixed=True, mode="source", boundary=bnd,
                coefficients=w.constant_coefficients(c=1.0, alpha=1e-6, beta=beta),
                source=w.GaussianSource(x0=(0.15, 0.5), r0=0.05, omega=16 * math.pi, Ts=1.0),
                record_every=R["record_every"], receivers=list(R["receivers"]), snap_times=list(R["snap_times"]))


def check_steepening(case, H):
    """H against data/steepening/<case>: the time axis exactly; the energy, max
    |v_h| and min (1 - 2 beta v_h) of all 576 records, 1e-12 relative; v_h at
    the three receivers every 8th step, 1e-12 of the series' largest |v_h|;
    for nl the snapshot at t = 0.80 (every 8th point), 1e-12 of its largest
    |v_h|, and the grid coordinates."""
    ser = np.array(steepening_series(case))
    assert np.array_equal(H.t, ser[:, 0])
    dev = {k: float(np.max(np.abs(getattr(H, k) / ser[:, i] - 1))) for i, k in ((1, "energy"), (2, "vmax"), (3, "ndmin"))}
    rv = np.array(steepening_recv(case))
    steps = rv[:, 0].astype(int) - 1
    assert np.array_equal(H.recv_t[steps], rv[:, 1])
    dev["recv"] = float(np.abs(H.recv[steps] - rv[:, 2:]).max() / np.abs(rv[:, 2:]).max())
    if case == "nl":
        rows, cols, vals = steepening_snapshot()
        snap = [G for (t, G) in H.snaps if f"{t:.2f}" == "0.80"]
        assert len(snap) == 1 and len(H.snaps) == 3
        G = snap[0][np.ix_(rows, cols)]
        vals = np.array(vals)
        dev["snap"] = float(np.abs(G - vals).max() / np.abs(vals).max())
        g = np.array(steepening_grid())
        assert np.abs(H.grid_x - g[:, 0]).max() < 1e-16 and np.abs(H.grid_y - g[:, 1]).max() < 1e-16
    say(f"steepening {case}: " + " ".join(f"{k} {x:.2e}" for k, x in dev.items()))
    for k in ("energy", "vmax", "ndmin"):
        assert dev[k] < 1e-12, (k, dev[k])
    assert dev["recv"] < 1e-12
    if case == "nl":
        assert dev["snap"] < 1e-12


@pytest.mark.parametrize("case", ["nl", "lin", "nl_dbc"])
def test_steepening(case):
    """experiments/steepening/run.jl on the device, all 18432 steps, against
    data/steepening/<case> (check_steepening). (Measured on MI355X: 2e-14 or
    better for every series.)"""
    say(f"case steepening {case}")
    kw = steepening_kwargs(case)
    H = w.solve(kw.pop("q"), kw.pop("N"), kw.pop("T"), **kw)
    check_steepening(case, H)


def test_steepening_batched():
    """The three steepening runs in one batch (run_source_many), each against
    the reference's data as the separate runs are."""
    from edg_westervelt_amd import batch
    import time
    cases = ["nl", "lin", "nl_dbc"]
    t0 = time.perf_counter()
    Hs = batch.run_source_many([steepening_kwargs(c) for c in cases])
    say(f"three steepening runs batched: {time.perf_counter() - t0:.2f} s")
    for c, H in zip(cases, Hs):
        check_steepening(c, H)


@pytest.mark.parametrize("q,N,mixed", [(3, 6, True), (5, 4, False), (1, 7, True)])
def test_norms_device(q, N, mixed):
    """k_norms: _mms_errors and difference_norms of random states with varying
    c (the paper's coefficients) against the oracle's loop restatements
    (oracle/norms.py), 1e-13 relative."""
    from oracle import norms
    say(f"case norms q={q} N={N} mixed={mixed}")
    rng = np.random.default_rng(200 + q)
    co = w.paper_coefficients()
    pre, geo = precompute(q, N, mixed=mixed, mode="mms", kx=2, ky=1, phase=(0.3, -0.2), coefficients=co)
    u, v = rand_state(rng, N, pre, su=0.1, sv=0.05)
    h = 1.0 / N
    for T in (0.0, 0.37):
        d = w.mms_errors(u, v, T, N, h, pre, geo, kx=2, ky=1, phase=(0.3, -0.2))
        o = norms.mms_errors(u, v, T, N, h, pre, geo, kx=2, ky=1, phase=(0.3, -0.2))
        for k2 in ("erru", "errv", "errgradu", "errE"):
            assert abs(getattr(d, k2) / o[k2] - 1) < 1e-13, (T, k2, getattr(d, k2), o[k2])
    # difference norms: B from a device solve, A a perturbed copy of it
    B = w.solve(q, N, 0.05, mixed=mixed, mode="mms", kx=2, ky=1, phase=(0.3, -0.2), coefficients=co,
                return_state=True)
    A = SimpleNamespace(u=B.u + 1e-3 * rng.standard_normal(B.u.shape), v=B.v + 1e-3 * rng.standard_normal(B.v.shape))
    d = w.difference_norms(A, B)
    o = norms.difference_norms(A.u, A.v, B.u, B.v, N, h, pre, geo)
    for k2 in ("erru", "errv", "errE"):
        assert abs(getattr(d, k2) / o[k2] - 1) < 1e-13, (k2, getattr(d, k2), o[k2])


@pytest.mark.parametrize("q", [1, 2, 3, 4, 5, 6])
@pytest.mark.parametrize("mixed", [True, False])
def test_energy_diagnostics_all_degrees(q, mixed):
    """The element energies, max |v_h| and min (1 - 2 beta v_h) of k_diag for every
    q and pairing with beta = 6 (test/runtests.jl:603-625's cases) against the
    oracle's loop, 1e-13 relative."""
    say(f"case energy diag q={q} mixed={mixed}")
    N = 5
    pre, geo = precompute(q, N, mixed=mixed, mode="source",
                          coefficients=w.replace_coefficients(w.paper_coefficients(), beta=6.0))
    rng = np.random.default_rng(q)
    u, v = rand_state(rng, N, pre, su=1e-2, sv=1e-2)
    ws = w.Workspace(pre, N)
    dh = ws.dev.h
    ws.dev.check(lib.wedg_set_state(dh, _p(u), _p(v)))
    c2 = np.asfortranarray(geo.err_c2); bq = np.asfortranarray(geo.err_beta); W2d = np.asfortranarray(geo.W2d)
    ws.dev.check(lib.wedg_problem_set_diag(ws.problem.h, _p(c2), _p(bq), _p(W2d)))
    E = np.zeros(N * N); vm = np.zeros(N * N); nd = np.zeros(N * N); sl = np.zeros(N * pre.qu1)
    r0 = od.vrow(0.0, pre.qv1)
    ws.dev.check(lib.wedg_diagnostics(dh, _p(E), _p(vm), _p(nd), 2, _p(r0), _p(sl)))
    Eo, vmo, ndo = od.energy_elements(u, v, N, 1.0 / N, pre, geo)
    assert rel(Eo, E) < 1e-13 and rel(vmo, vm) < 1e-13 and rel(ndo, nd) < 1e-13


@pytest.mark.parametrize("flux", ["upwind", "central"])
def test_source_energy_decay(flux):
    """test/runtests.jl:839-850 on the device: quadrant coefficients, a source that
    ends at t = 0.25, solve(3, 8, 0.75; mode = :source, record_every = 5): after
    the source the recorded energy never rises by more than 1e-10 E[1] and ends
    below 0.95 of its first value."""
    say(f"case energy decay {flux}")
    co = quadrant_coefficients(0.5, 0.5, alpha=(1e-5, 4e-6, 2e-5, 1e-6), beta=(1e-4, 3e-4, 0.0, 2e-4))
    src = w.GaussianSource(x0=(0.45, 0.55), r0=0.08, amplitude=50.0, omega=8 * math.pi, Ts=0.25)
    H = w.solve(3, 8, 0.75, mode="source", coefficients=co, flux=flux, source=src, record_every=5)
    E = H.energy[H.t > 0.25]
    assert np.max(np.diff(E)) <= 1e-10 * E[0]
    assert E[-1] / E[0] < 0.95


def test_norms_after_receivers_on_one_workspace():
    """Error norms, then a receiver run, then the norms again on one workspace:
    the receiver buffers grow without touching the norms' staging, and the
    norms agree with a fresh workspace that took the same steps."""
    import ctypes as C
    from edg_westervelt_amd import _lib
    from edg_westervelt_amd.solver import Workspace, _device_mms_errors, _ptr, time_step, vrow
    lib = _lib.lib
    N, q = 6, 2
    pre, geo = precompute(q, N, mode="mms")
    dt, _ = time_step(q, N, 0.1)

    def fresh():
        ws = Workspace(pre, N)
        u = np.zeros((N, N, pre.qu1, pre.qu1), order="F")
        v = np.zeros((N, N, pre.qv1, pre.qv1), order="F")
        ws.dev.check(lib.wedg_set_state(ws.dev.h, _ptr(u), _ptr(v)))
        return ws

    def errs(ws, T):
        e = _device_mms_errors(ws, T, N, 1.0 / N, pre, geo)
        return (e.erru, e.errv, e.errgradu, e.errE)

    ws = fresh()
    t = C.c_double(0.0)
    ws.dev.check(lib.wedg_advance(ws.dev.h, 0.0, dt, 3, 0, C.byref(t)))
    e1 = errs(ws, t.value)
    rj = np.array([1, 4], dtype=np.int32)
    rk = np.array([2, 3], dtype=np.int32)
    rows = np.ascontiguousarray(np.tile(vrow(0.2, pre.qv1), (2, 2)))
    recv = np.zeros(2 * 2)
    ip = C.POINTER(C.c_int)
    t2 = C.c_double(0.0)
    ws.dev.check(lib.wedg_advance_receivers(ws.dev.h, t.value, dt, 0, 2, rj.ctypes.data_as(ip), rk.ctypes.data_as(ip),
                                            _ptr(rows), 0, _ptr(recv), C.byref(t2)))
    assert errs(ws, t.value) == e1
    ws.dev.check(lib.wedg_advance_receivers(ws.dev.h, t.value, dt, 2, 2, rj.ctypes.data_as(ip), rk.ctypes.data_as(ip),
                                            _ptr(rows), 0, _ptr(recv), C.byref(t2)))
    ref = fresh()
    t3 = C.c_double(0.0)
    ref.dev.check(lib.wedg_advance(ref.dev.h, 0.0, dt, 3, 0, C.byref(t3)))
    ref.dev.check(lib.wedg_advance(ref.dev.h, t3.value, dt, 2, 0, C.byref(t3)))
    assert t3.value == t2.value
    assert errs(ws, t2.value) == errs(ref, t3.value)
    del ws


@pytest.mark.parametrize("cheb", ["global", "element"])
def test_source_blow_up_inside_a_record_interval(cheb):
    """beta = 0 and a time step 200 times the stable one (dt_fac = 0.05): the
    state overflows at a step inside a record interval (record_every = 20). The
    run stops with the reference's "blow-up at t=..." (src/solver.jl:940), t
    after that step as the C oracle's RK4 loop finds it, and not with the mass
    failure the non-finite state gives in the next stage, which the device
    records as it runs on to the end of the interval."""
    from edg_westervelt_amd.solver import time_step
    q, N, T, dtf = 2, 4, 400.0, 0.05
    lin = w.constant_coefficients(c=1.0, alpha=1e-4, beta=0.0)
    pre, _ = precompute(q, N, mixed=True, mode="source", coefficients=lin)
    dt, ns = time_step(q, N, T, dtf)
    u0 = np.zeros((N, N, pre.qu1, pre.qu1), order="F")
    v0 = np.zeros((N, N, pre.qv1, pre.qv1), order="F")
    _, _, _, ost = ob.rk4_steps(ob.OraclePre(pre, N), u0, v0, 0.0, dt, ns, cheb_mode=0 if cheb == "global" else 1,
                                check=False)
    assert ost.code == _lib.ERR_NONFINITE and ost.step % 20 != 0
    tb = 0.0
    for _ in range(ost.step):
        tb += dt
    say(f"case blow-up {cheb}: oracle step {ost.step}, t={tb!r}")
    with pytest.raises(w.WedgError) as e:
        w.solve(q, N, T, mixed=True, mode="source", coefficients=lin, dt_fac=dtf, cheb=cheb)
    assert str(e.value) == f"blow-up at t={tb!r}"


def _bottom_beta(f, b0=20.0, ell=0.04):
    """beta = b0 (1 + f exp(-y / ell)): larger at the bottom face points than at
    the volume points near them, so the boundary weight and the mass matrix
    fail at different stages as f varies."""
    from edg_westervelt_amd.coefficients import Coefficients

    def z0(x, y):
        return 0.0 * x

    return Coefficients(lambda x, y: 1.0 + 0 * x, z0, z0, lambda x, y: 1e-3 + 0 * x, z0, z0,
                        lambda x, y: b0 * (1 + f * np.exp(-y / ell)))


# f: the failure the oracle's RK4 finds first (step, stage found on the CPU)
MIDSTEP = [(0.2, "mass, step 8 stage 2"), (1.0, "boundary, step 5 stage 4"),
           (0.6, "boundary and mass at the same stage, step 7 stage 2")]


@pytest.mark.parametrize("cheb", ["global", "element"])
@pytest.mark.parametrize("f,what", MIDSTEP)
def test_failure_inside_an_advance(f, what, cheb):
    """v grows from rest under the u0 forcing of :initial_value mode (a bump
    on the bottom side) until a boundary weight or a mass matrix fails at a
    stage 2-4 of a step after the first. wedg_advance reports the failure the
    C oracle's rk4 loop reports: status, side, key, element; where both fail
    at one stage, the boundary weight (src/rhs.jl:530-531)."""
    from edg_westervelt_amd.batch import gaussian_profile
    q, N = 3, 6
    pre, _ = precompute(q, N, mixed=True, mode="initial_value", coefficients=_bottom_beta(f),
                        boundary=dict(bot="absorbing_2beta", top="wall", lft="wall", rgt="wall"),
                        initial_profile=gaussian_profile(0.5, 0.0, 0.12, 0.05))
    dt, ns = 1.0 / (N * 10 * q * q), 20
    u0 = np.zeros((N, N, pre.qu1, pre.qu1), order="F")
    v0 = np.zeros((N, N, pre.qv1, pre.qv1), order="F")
    _, _, _, ost = ob.rk4_steps(ob.OraclePre(pre, N), u0, v0, 0.0, dt, ns, cheb_mode=0 if cheb == "global" else 1,
                                check=False)
    say(f"case failure inside an advance f={f} ({what}) {cheb}: oracle code {ost.code}")
    assert ost.code in (_lib.ERR_BOUNDARY_TRACE, _lib.ERR_MASS)
    ws = w.Workspace(pre, N, cheb=cheb)
    ws.dev.check(lib.wedg_set_state(ws.dev.h, u0.ctypes.data_as(C.POINTER(C.c_double)),
                                    v0.ctypes.data_as(C.POINTER(C.c_double))))
    tend = C.c_double(0.0)
    st = lib.wedg_advance(ws.dev.h, 0.0, dt, ns, 0, C.byref(tend))
    info = ws.dev.last_error()
    assert st == ost.code
    if st == _lib.ERR_MASS:
        assert info.element == ost.elem
    else:
        assert (info.side, info.key) == (ost.side, ost.key)
