"""Batched instances on the device (wedg_batch_*, through the C ABI): every
instance of a batch ends bitwise where its own wedg_advance takes it (the same
kernels get the same stage arguments), within the RK4 tolerance of the oracle,
and a failing instance reports its own error without touching the others.

The instances are the uneven workload of edg_westervelt_amd.batch
(initial-value mode, piecewise-constant coefficients with jumps on a 4 x 4
block grid, every flux family and boundary kind, q_u 2..6, both pairings);
the tests run a few steps of each.
"""
import ctypes as C
import sys

import numpy as np
import pytest

import edg_westervelt_amd.batch as B
from edg_westervelt_amd import _lib
from edg_westervelt_amd._lib import lib
from oracle import batched as ob
from tests.conftest import rel

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def solo(b, nsteps, v0=None):
    """The instance alone from the zero state (or v0): state, end time, status, error info."""
    B.set_zero_state(b.ws)
    if v0 is not None:
        u = np.zeros((b.ws.N, b.ws.N, b.ws.qu1, b.ws.qu1), order="F")
        b.ws.dev.check(lib.wedg_set_state(b.ws.dev.h, u.ctypes.data_as(C.POINTER(C.c_double)),
                                          v0.ctypes.data_as(C.POINTER(C.c_double))))
    t = C.c_double(0.0)
    st = lib.wedg_advance(b.ws.dev.h, 0.0, b.dt, int(nsteps), 0, C.byref(t))
    info = b.ws.dev.last_error()
    u, v = B.get_state(b.ws)
    return u, v, t.value, st, (info.status, info.side, info.key, info.element)


def describe(b):
    x = b.inst
    return (f"instance {x.index}: q_u={x.q} mixed={x.mixed} N={x.N} {x.flux} {x.boundary} "
            f"one_sided={b.pre.one_sided}")


@pytest.mark.parametrize("schedule", [0, 1])
def test_batch_bitwise_equals_separate_runs(schedule):
    """Every instance bitwise as its own run, with the instances starting
    together (schedule 0) or ending together (1), over two streams and chunks
    of three steps (so chunks and instance starts cut each other)."""
    insts = B.uneven_instances(seed=0, count=12)
    steps = [3, 7, 0, 5, 1, 9, 4, 2, 6, 8, 3, 5]
    built = [B.build_instance(x) for x in insts]
    ref = []
    for b, n in zip(built, steps):
        say("solo", describe(b), "steps", n)
        ref.append(solo(b, n))
    for b in built:
        B.set_zero_state(b.ws)
    say("batch of", len(built), "schedule", schedule)
    bt = B.Batch([b.ws for b in built])
    bt.configure(2, 3, schedule)
    tend, status = bt.advance(0.0, [b.dt for b in built], steps)
    for i, b in enumerate(built):
        u, v = B.get_state(b.ws)
        assert status[i] == _lib.OK
        assert tend[i] == ref[i][2]
        assert np.array_equal(u, ref[i][0]) and np.array_equal(v, ref[i][1]), describe(b)
    # a second advance of the same batch continues from the states it left
    tend2, _ = bt.advance(tend, [b.dt for b in built], [2] * len(built))
    for i, b in enumerate(built):
        u, v = B.get_state(b.ws)
        B.set_zero_state(b.ws)
        uu, vv, t, _, _ = solo(b, steps[i] + 2)
        assert np.array_equal(u, uu) and np.array_equal(v, vv), describe(b)
        assert tend2[i] == t


def test_batch_matches_oracle():
    """Four instances of different groups, four steps each, against the oracle's RK4."""
    insts = B.uneven_instances(seed=0, count=16)
    pick = [0, 3, 7, 12]   # q_u 5, 6, 3 and 2
    built = [B.build_instance(insts[i]) for i in pick]
    bt = B.Batch([b.ws for b in built])
    bt.advance(0.0, [b.dt for b in built], [4] * len(built))
    for b in built:
        say("oracle", describe(b))
        u, v = B.get_state(b.ws)
        z_u = np.zeros_like(u, order="F")
        z_v = np.zeros_like(v, order="F")
        uo, vo, _, _ = ob.rk4_steps(ob.OraclePre(b.pre, b.inst.N), z_u, z_v, 0.0, b.dt, 4)
        assert np.abs(vo).max() > 0
        assert rel(uo, u) < 1e-11 and rel(vo, v) < 1e-11, describe(b)


def test_batch_failure_stays_with_its_instance():
    """One instance starts with v_h = 1 (1 - 2 beta v_h <= 0 wherever beta >= 1/2):
    it stops with the error (and message) its own run gives; the others are
    unaffected."""
    insts = B.uneven_instances(seed=0, count=6)
    built = [B.build_instance(x) for x in insts]
    bad = 2
    v0 = np.zeros((built[bad].inst.N, built[bad].inst.N, built[bad].ws.qv1, built[bad].ws.qv1), order="F")
    v0[:, :, 0, 0] = 2.0   # v_h = 2 L_0(y) L_0(x) = 1
    steps = [3] * len(built)
    ref = [solo(b, 3, v0 if i == bad else None) for i, b in enumerate(built)]
    assert ref[bad][3] != _lib.OK
    say("bad instance status", ref[bad][3], ref[bad][4])
    for b in built:
        B.set_zero_state(b.ws)
    u = np.zeros((built[bad].ws.N, built[bad].ws.N, built[bad].ws.qu1, built[bad].ws.qu1), order="F")
    built[bad].ws.dev.check(lib.wedg_set_state(built[bad].ws.dev.h, u.ctypes.data_as(C.POINTER(C.c_double)),
                                               v0.ctypes.data_as(C.POINTER(C.c_double))))
    bt = B.Batch([b.ws for b in built])
    _, status = bt.advance(0.0, [b.dt for b in built], steps, raise_errors=False)
    for i, b in enumerate(built):
        info = b.ws.dev.last_error()
        assert status[i] == ref[i][3]
        assert (info.status, info.side, info.key, info.element) == ref[i][4]
        if i != bad:
            uu, vv = B.get_state(b.ws)
            assert np.array_equal(uu, ref[i][0]) and np.array_equal(vv, ref[i][1]), describe(b)
    with pytest.raises(_lib.WedgError) as e:
        bt.advance(0.0, [b.dt for b in built], [1] * len(built))
    assert str(e.value) == _lib.error_message(built[bad].pre, built[bad].ws.dev.last_error())


def test_batch_full_workload_two_steps():
    """All 64 instances of the uneven workload (every group, N up to 64), two steps."""
    built = [B.build_instance(x) for x in B.uneven_instances(seed=0, count=64)]
    ref = [solo(b, 2)[:2] for b in built]
    for b in built:
        B.set_zero_state(b.ws)
    bt = B.Batch([b.ws for b in built])
    _, status = bt.advance(0.0, [b.dt for b in built], [2] * len(built))
    assert (status == _lib.OK).all()
    for i, b in enumerate(built):
        u, v = B.get_state(b.ws)
        assert np.array_equal(u, ref[i][0]) and np.array_equal(v, ref[i][1]), describe(b)


def test_batch_two_lanes_bitwise_equals_separate_runs():
    """Instances with two lanes per element next to one-lane instances of the
    same (q_u, q_v): two groups per pair, every instance bitwise as its own run
    (which uses its lanes too), within the RK4 tolerance of the oracle, over
    two streams, chunks of three steps and a second advance."""
    insts = B.uneven_instances(seed=0, count=16)
    built = [B.build_instance(x) for x in insts]
    lanes = [1 if i % 3 == 1 else 2 for i in range(len(built))]
    steps = [3, 7, 0, 5, 1, 9, 4, 2, 6, 8, 3, 5, 4, 6, 2, 7]
    ref = []
    for b, n, l in zip(built, steps, lanes):
        b.ws.dev.set_lanes(l)
        assert b.ws.dev.lanes == l, describe(b)   # the uneven workload has two-lane kernels throughout
        say("solo", describe(b), "lanes", l, "steps", n)
        ref.append(solo(b, n))
    for b in built:
        B.set_zero_state(b.ws)
    bt = B.Batch([b.ws for b in built])
    assert bt.lanes_per_element() == lanes
    bt.configure(2, 3, 0)
    say("batch of", len(built), "with lanes", lanes)
    tend, status = bt.advance(0.0, [b.dt for b in built], steps)
    for i, b in enumerate(built):
        u, v = B.get_state(b.ws)
        assert status[i] == _lib.OK and tend[i] == ref[i][2]
        assert np.array_equal(u, ref[i][0]) and np.array_equal(v, ref[i][1]), describe(b)
    for i in (3, 12, 14):   # two-lane instances of q_u 6, 2 and 4 (seed 0)
        b = built[i]
        say("oracle", describe(b))
        u, v = B.get_state(b.ws)
        z_u = np.zeros_like(u, order="F")
        z_v = np.zeros_like(v, order="F")
        uo, vo, _, _ = ob.rk4_steps(ob.OraclePre(b.pre, b.inst.N), z_u, z_v, 0.0, b.dt, steps[i])
        assert np.abs(vo).max() > 0
        assert rel(uo, u) < 1e-11 and rel(vo, v) < 1e-11, describe(b)
    tend2, _ = bt.advance(tend, [b.dt for b in built], [2] * len(built))
    for i, b in enumerate(built):
        u, v = B.get_state(b.ws)
        B.set_zero_state(b.ws)
        uu, vv, t, _, _ = solo(b, steps[i] + 2)
        assert np.array_equal(u, uu) and np.array_equal(v, vv), describe(b)
        assert tend2[i] == t
    # the lanes are fixed while the batch exists; set_lanes_per_element forms it again
    built[1].ws.dev.set_lanes(2)
    with pytest.raises(_lib.WedgError, match=f"wedg_batch_advance failed with status {_lib.ERR_ARGUMENT}"):
        bt.advance(0.0, [b.dt for b in built], [1] * len(built), raise_errors=False)
    bt.set_lanes_per_element(2)
    assert bt.lanes_per_element() == [2] * len(built) and bt.config()[:2] == (2, 3)
    for b in built:
        B.set_zero_state(b.ws)
    bt.advance(0.0, [b.dt for b in built], [1] * len(built))


def test_batch_two_lanes_failure_and_tuner():
    """A failing two-lane instance reports its own run's error; the lanes
    tuner chooses per kernel group, restores every state, and the tuned batch
    is bitwise the instances' own runs with the chosen lanes."""
    insts = B.uneven_instances(seed=0, count=8)
    built = [B.build_instance(x) for x in insts]
    for b in built:
        b.ws.dev.set_lanes(2)
    bad = 3
    v0 = np.zeros((built[bad].inst.N, built[bad].inst.N, built[bad].ws.qv1, built[bad].ws.qv1), order="F")
    v0[:, :, 0, 0] = 2.0   # v_h = 1
    ref = [solo(b, 3, v0 if i == bad else None) for i, b in enumerate(built)]
    assert ref[bad][3] != _lib.OK
    for b in built:
        B.set_zero_state(b.ws)
    u = np.zeros((built[bad].ws.N, built[bad].ws.N, built[bad].ws.qu1, built[bad].ws.qu1), order="F")
    built[bad].ws.dev.check(lib.wedg_set_state(built[bad].ws.dev.h, u.ctypes.data_as(C.POINTER(C.c_double)),
                                               v0.ctypes.data_as(C.POINTER(C.c_double))))
    bt = B.Batch([b.ws for b in built])
    _, status = bt.advance(0.0, [b.dt for b in built], [3] * len(built), raise_errors=False)
    for i, b in enumerate(built):
        info = b.ws.dev.last_error()
        assert status[i] == ref[i][3]
        assert (info.status, info.side, info.key, info.element) == ref[i][4]
        if i != bad:
            uu, vv = B.get_state(b.ws)
            assert np.array_equal(uu, ref[i][0]) and np.array_equal(vv, ref[i][1]), describe(b)
    for b in built:
        B.set_zero_state(b.ws)
    times, choice = bt.tune(0.0, [b.dt for b in built], probe_steps=4, lanes=(1, 2), chunks=(0,), repeats=1,
                            lanes_per_element=(1, 2))
    say("lane tuner", bt.lane_times, bt.lane_choice, "streams", times, choice)
    groups = bt.kernel_groups()
    assert set(bt.lane_choice) == set(groups)
    assert set(bt.lane_times) == {k + (l,) for k in groups for l in (1, 2)}
    lpe = bt.lanes_per_element()
    for key, members in groups.items():
        assert all(lpe[i] == bt.lane_choice[key] for i in members)
    for b in built:
        u, v = B.get_state(b.ws)
        assert not u.any() and not v.any()
    steps = [5, 3, 6, 2, 4, 5, 1, 3]
    tend, status = bt.advance(0.0, [b.dt for b in built], steps)
    got = [B.get_state(b.ws) for b in built]
    for i, b in enumerate(built):
        uu, vv, t, _, _ = solo(b, steps[i])
        assert np.array_equal(got[i][0], uu) and np.array_equal(got[i][1], vv) and tend[i] == t, describe(b)


def test_calibrated_amplitudes_keep_the_weight_bound():
    """The 64 instances with their calibrated amplitudes, whole runs to T:
    2 beta_max max|v_h| (after every 16 steps) stays in [0.48, 0.5], so
    1 - 2 beta max|v| >= 0.5 as the workload states, and matches the stored
    calibration to 1e-6 (the runs are deterministic)."""
    import json
    insts = B.uneven_instances(seed=0, count=64)
    assert all(x.amp_rule == "calibrated" for x in insts)
    built = [B.build_instance(x) for x in insts]
    vm, status = B.run_vmax(built, 16)
    assert (status == _lib.OK).all()
    d = 2.0 * np.array([x.beta.max() for x in insts]) * vm
    say("d range", d.min(), d.max())
    with open(B.AMPLITUDE_FILE) as f:
        stored = np.array(json.load(f)["runs"][0]["d"])
    assert ((0.48 <= d) & (d <= 0.5)).all()
    assert np.abs(d - stored).max() < 1e-4


def test_batch_rejects_a_workspace_twice():
    b = B.build_instance(B.uneven_instances(seed=0, count=1)[0])
    arr = (C.c_void_p * 2)(b.ws.dev.h.value, b.ws.dev.h.value)
    h = C.c_void_p()
    assert lib.wedg_batch_create(arr, 2, C.byref(h)) == _lib.ERR_ARGUMENT
    with pytest.raises(ValueError):
        B.Batch([b.ws]).advance(0.0, [b.dt, b.dt], [1, 1])


def test_batch_mixes_every_kernel_layout_and_policy():
    """One batch with every specialisation of the stage kernel — per-element
    classes with the stored inverse (paper coefficients), with the spectral
    u-solve (quadrant jumps on element faces), one class (constant
    coefficients), conjugate gradients — both count policies in one group, a
    source-mode instance and random states: bitwise the solo runs."""
    import edg_westervelt_amd as w
    from edg_westervelt_amd import precompute
    from tests.cases import MIXED, quadrant_coefficients, rand_state
    rng = np.random.default_rng(21)
    cc = w.constant_coefficients(c=1.0, alpha=1e-4, beta=2.0)
    cases = [
        dict(q=3, N=10, kw=dict(mixed=True, mode="mms"), cheb="global"),
        dict(q=3, N=12, kw=dict(mixed=False, mode="mms"), cheb="element"),
        dict(q=3, N=7, kw=dict(mixed=False, mode="mms", flux="central"), cheb="global"),
        dict(q=4, N=16, kw=dict(mode="initial_value", coefficients=cc, boundary=MIXED, sigma_dirichlet=0.3,
                                initial_profile=B.gaussian_profile(0.5, 0.5, 0.08, 0.01)), cheb="global"),
        dict(q=4, N=8, kw=dict(mode="source", coefficients=cc, boundary="wall", kappa=0.5,
                               source=w.GaussianSource(x0=(0.5, 0.7), r0=0.05)), cheb="global"),
        dict(q=2, N=9, kw=dict(mode="mms", mass_solver="cg"), cheb="global"),
        dict(q=5, N=20, kw=dict(mode="initial_value", coefficients=quadrant_coefficients(0.5, 0.5),
                                flux="alternating"), cheb="global"),
    ]
    built = []
    for c in cases:
        pre, _ = precompute(c["q"], c["N"], **c["kw"])
        ws = w.Workspace(pre, c["N"], cheb=c["cheb"])
        u, v = rand_state(rng, c["N"], pre, su=1e-5, sv=1e-3)
        built.append((c, ws, u, v))
    steps = [3, 2, 4, 3, 5, 2, 3]
    dts = [1e-4, 2e-4, 1.5e-4, 1e-4, 2e-4, 1e-4, 5e-5]
    P = C.POINTER(C.c_double)

    def put(ws, u, v):
        ws.dev.check(lib.wedg_set_state(ws.dev.h, u.ctypes.data_as(P), v.ctypes.data_as(P)))

    ref = []
    for (c, ws, u, v), n, dt in zip(built, steps, dts):
        say("solo", c["q"], c["N"], c["kw"].get("mode"), c["kw"].get("mass_solver", "chebyshev"), c["cheb"])
        put(ws, u, v)
        t = C.c_double(0.0)
        ws.dev.check(lib.wedg_advance(ws.dev.h, 0.05, dt, n, 0, C.byref(t)))
        ref.append(B.get_state(ws) + (t.value, ws.dev.last_error().cheb_steps))
    for c, ws, u, v in built:
        put(ws, u, v)
    bt = B.Batch([x[1] for x in built])
    tend, status = bt.advance(0.05, dts, steps)
    assert (status == _lib.OK).all()
    for i, (c, ws, u, v) in enumerate(built):
        uu, vv = B.get_state(ws)
        assert np.array_equal(uu, ref[i][0]) and np.array_equal(vv, ref[i][1]), c
        assert tend[i] == ref[i][2] and ws.dev.last_error().cheb_steps == ref[i][3]
        assert np.abs(vv - v).max() > 0


def _plain(q, N, **kw):
    import edg_westervelt_amd as w
    from edg_westervelt_amd import precompute
    pre, _ = precompute(q, N, **kw)
    return w.Workspace(pre, N)


def _run_both(wss, states, dts, steps):
    """Solo runs, then the batch from the same states; returns both results."""
    P = C.POINTER(C.c_double)

    def put(ws, u, v):
        ws.dev.check(lib.wedg_set_state(ws.dev.h, u.ctypes.data_as(P), v.ctypes.data_as(P)))

    ref = []
    for ws, (u, v), dt, n in zip(wss, states, dts, steps):
        put(ws, u, v)
        t = C.c_double(0.0)
        ws.dev.check(lib.wedg_advance(ws.dev.h, 0.0, dt, n, 0, C.byref(t)))
        ref.append(B.get_state(ws) + (t.value,))
    for ws, (u, v) in zip(wss, states):
        put(ws, u, v)
    tend, status = B.Batch(wss).advance(0.0, dts, steps)
    got = [B.get_state(ws) for ws in wss]
    return ref, got, tend, status


def test_batch_edge_cases():
    """The smallest meshes (N = 1, 2, 3: one partly filled wave), a batch of one,
    and an advance of zero steps (the states stay as they are)."""
    from tests.cases import rand_state
    rng = np.random.default_rng(8)
    wss = [_plain(1, 1, mode="mms"), _plain(2, 2, mode="mms", mixed=False), _plain(3, 3, mode="mms", flux="central"),
           _plain(2, 1, mode="mms", mixed=False)]
    states = [rand_state(rng, ws.N, ws.pre, su=1e-3, sv=1e-3) for ws in wss]
    ref, got, tend, status = _run_both(wss, states, [1e-3] * 4, [3, 2, 4, 1])
    assert (status == _lib.OK).all()
    for r, g in zip(ref, got):
        assert np.array_equal(r[0], g[0]) and np.array_equal(r[1], g[1])
    assert list(tend) == [r[2] for r in ref]
    # one instance
    ref, got, tend, _ = _run_both(wss[2:3], states[2:3], [2e-3], [3])
    assert np.array_equal(ref[0][0], got[0][0]) and np.array_equal(ref[0][1], got[0][1])
    # zero steps: nothing runs, the states and times are unchanged
    ref, got, tend, status = _run_both(wss, states, [1e-3] * 4, [0, 0, 0, 0])
    assert (status == _lib.OK).all() and list(tend) == [0.0] * 4
    for (u, v), g in zip(states, got):
        assert np.array_equal(u, g[0]) and np.array_equal(v, g[1])


def test_batch_with_the_full_size_instance():
    """The 256 x 256, q = 5 instance (1,024 waves: the staggered start of the
    stage kernel) in one batch with a small one, two steps: bitwise the solo runs."""
    import edg_westervelt_amd as w
    cc = w.constant_coefficients(c=1.0, alpha=1e-4, beta=2.0)
    big = _plain(5, 256, mixed=False, mode="initial_value", coefficients=cc,
                 initial_profile=B.gaussian_profile(0.5, 0.5, 0.05, 0.05))
    sb = B.build_instance(B.uneven_instances(seed=0, count=1)[0])
    small = sb.ws
    zero = [(np.zeros((ws.N, ws.N, ws.qu1, ws.qu1), order="F"), np.zeros((ws.N, ws.N, ws.qv1, ws.qv1), order="F"))
            for ws in (big, small)]
    ref, got, tend, status = _run_both([big, small], zero, [1.5625e-5, sb.dt], [2, 3])
    assert (status == _lib.OK).all()
    for r, g in zip(ref, got):
        assert np.array_equal(r[0], g[0]) and np.array_equal(r[1], g[1])
        assert np.abs(g[1]).max() > 0


def test_solve_many_equals_solve():
    """solve_many returns what solve returns for each run: the :mms errors
    (bitwise: the same states and the same norm kernels) and the final states."""
    import edg_westervelt_amd as w
    runs = [dict(q=3, N=10, T=0.1, flux="upwind", mixed=True),
            dict(q=2, N=8, T=0.05, flux="central", mixed=False, boundary="absorbing", sigma=0.3),
            dict(q=4, N=6, T=0.05, mass_solver="cg"),
            dict(q=3, N=8, T=0.02, mode="initial_value", return_state=True,
                 coefficients=w.constant_coefficients(c=1.0, alpha=1e-4, beta=2.0),
                 initial_profile=B.gaussian_profile(0.5, 0.5, 0.1, 0.01))]
    got = B.solve_many(runs)
    for r, g in zip(runs, got):
        r = dict(r)
        ref = w.solve(r.pop("q"), r.pop("N"), r.pop("T"), **r)
        if r.get("return_state"):
            assert np.array_equal(ref.u, g.u) and np.array_equal(ref.v, g.v) and ref.t == g.t
        else:
            assert (ref.erru, ref.errv, ref.errgradu, ref.errE) == (g.erru, g.errv, g.errgradu, g.errE)
    # the cfg1 run's errors are the reference's (data/convergence/results.csv row 364 was made with CG)
    assert abs(got[0].erru / 1.0717765588e-04 - 1) < 1e-3
    with pytest.raises(ValueError, match="source"):
        B.solve_many([dict(q=2, N=4, T=0.01, mode="source")])
    with pytest.raises(ValueError, match="return_state"):
        B.solve_many([dict(q=2, N=4, T=0.01, mode="initial_value")])


def test_batch_tuner_changes_no_result():
    """The autotuner tries stream counts and argument chunks, restores every
    state, and the tuned batch still gives bitwise the solo runs."""
    insts = B.uneven_instances(seed=0, count=8)
    built = [B.build_instance(x) for x in insts]
    steps = [5, 3, 6, 2, 4, 5, 1, 3]
    ref = [solo(b, n)[:3] for b, n in zip(built, steps)]
    for b in built:
        B.set_zero_state(b.ws)
    bt = B.Batch([b.ws for b in built])
    times, choice = bt.tune(0.0, [b.dt for b in built], probe_steps=4, lanes=(1, 2, 4), chunks=(0, 1), repeats=1)
    say("tuner", times, choice)
    assert set(times) == {(l, c) for l in (1, 2, 4) for c in (0, 1)} and choice in times
    assert bt.config()[0] == choice[0]
    for b in built:
        u, v = B.get_state(b.ws)
        assert not u.any() and not v.any()   # the probes' states were put back
    tend, status = bt.advance(0.0, [b.dt for b in built], steps)
    for i, b in enumerate(built):
        u, v = B.get_state(b.ws)
        assert np.array_equal(u, ref[i][0]) and np.array_equal(v, ref[i][1]) and tend[i] == ref[i][2]
    with pytest.raises(_lib.WedgError):
        bt.configure(0)
    with pytest.raises(_lib.WedgError):
        bt.configure(17)


def test_batch_blow_up_instance():
    """An instance that blows up (beta = 0, a time step far above the stable
    one) stops with the error its own run reports — the non-finite state
    fails the next stage's Chebyshev interval — at the same element and step,
    with stop_on_nonfinite; the other instance finishes unaffected."""
    import edg_westervelt_amd as w
    lin = w.constant_coefficients(c=1.0, alpha=1e-4, beta=0.0)
    wss = [_plain(2, 4, mode="initial_value", coefficients=lin,
                  initial_profile=B.gaussian_profile(0.5, 0.5, 0.1, 0.01)),
           _plain(3, 6, mode="mms")]
    dts, steps = [0.5, 1e-3], [400, 5]
    zero = [(np.zeros((ws.N, ws.N, ws.qu1, ws.qu1), order="F"), np.zeros((ws.N, ws.N, ws.qv1, ws.qv1), order="F"))
            for ws in wss]
    P = C.POINTER(C.c_double)

    def reset():
        for ws, (u, v) in zip(wss, zero):
            ws.dev.check(lib.wedg_set_state(ws.dev.h, u.ctypes.data_as(P), v.ctypes.data_as(P)))

    def info(ws):
        e = ws.dev.last_error()
        return (e.status, e.side, e.key, e.element, e.step)

    reset()
    ref = []
    for ws, dt, n in zip(wss, dts, steps):
        t = C.c_double(0.0)
        st = lib.wedg_advance(ws.dev.h, 0.0, dt, n, 1, C.byref(t))
        ref.append((st, info(ws), B.get_state(ws)))
    assert ref[0][0] != _lib.OK and ref[1][0] == _lib.OK
    say("blow-up instance:", ref[0][1])
    reset()
    bt = B.Batch(wss)
    _, status = bt.advance(0.0, dts, steps, stop_on_nonfinite=True, raise_errors=False)
    assert list(status) == [ref[0][0], ref[1][0]]
    assert info(wss[0]) == ref[0][1]
    u, v = B.get_state(wss[1])
    assert np.array_equal(u, ref[1][2][0]) and np.array_equal(v, ref[1][2][1])
    reset()
    with pytest.raises(_lib.WedgError):
        bt.advance(0.0, dts, steps, stop_on_nonfinite=True)


def _same_series(a, b):
    assert np.array_equal(a.t, b.t) and np.array_equal(a.energy, b.energy)
    assert np.array_equal(a.vmax, b.vmax) and np.array_equal(a.ndmin, b.ndmin)
    assert np.array_equal(a.slices, b.slices) and np.array_equal(a.recv, b.recv)
    assert np.array_equal(a.recv_t, b.recv_t) and len(a.snaps) == len(b.snaps)
    for (ta, Ga), (tb, Gb) in zip(a.snaps, b.snaps):
        assert ta == tb and np.array_equal(Ga, Gb)


def test_run_source_many_uneven_schedules():
    """Source-mode runs of different meshes, degrees, end times, record
    intervals and snapshot times in one batch: every series, receiver row and
    snapshot bitwise those of separate solve(..., mode="source") calls."""
    import edg_westervelt_amd as w
    src = w.GaussianSource(x0=(0.4, 0.5), r0=0.1)
    cc = w.constant_coefficients(c=1.0, alpha=1e-4, beta=2.0)
    runs = [dict(q=2, N=8, T=0.05, record_every=5, snap_times=(0.02,), source=src, coefficients=cc,
                 receivers=((0.45, 0.5), (0.7, 0.3))),
            dict(q=3, N=6, T=0.04, record_every=7, source=src, coefficients=cc, boundary="wall", kappa=0.4,
                 mixed=False),
            dict(q=2, N=8, T=0.03, record_every=3, snap_times=(0.01, 0.025), source=src, coefficients=cc,
                 boundary="dirichlet", receivers=((0.2, 0.8),))]
    got = B.run_source_many(runs)
    for r, g in zip(runs, got):
        r = dict(r)
        ref = w.solve(r.pop("q"), r.pop("N"), r.pop("T"), mode="source", **r)
        _same_series(ref, g)
        assert len(g.t) > 1 and np.abs(g.recv).max() > 0
    with pytest.raises(ValueError):
        B.run_source_many([dict(q=2, N=4, T=0.01, mode="mms")])


def test_convergence_study_whole_table():
    """The reference's whole convergence study, data/convergence/results.csv
    (756 rows: q = 1..6, both pairings, upwind/central/alternating, N = 5..25,
    T = 0.1, made with CG), reproduced on the device with solve_many, one batch
    per mass solver. Every error of every row agrees to 1e-9 relative with CG
    and to 4 significant digits with Chebyshev, above an absolute floor of
    1e-12: the rounding of a state of size 1 carried over up to 900 RK4 steps
    (errors below ~1e-9 at q = 6, N >= 17 differ by a few 1e-13 between
    summation orders)."""
    import csv
    import gc
    import os
    import time
    path = os.path.join(os.path.dirname(__file__), "golden", "convergence_results.csv")
    rows = list(csv.DictReader(open(path)))
    assert len(rows) == 756
    keys = (("erru", "erru"), ("errv", "errv"), ("errgradu", "errux"), ("errE", "errE"))
    for solver, rtol in (("cg", 1e-9), ("chebyshev", 5e-5)):
        runs = [dict(q=int(r["q"]), N=int(r["N"]), T=0.1, flux=r["flux"], mixed=r["pairing"] == "mixed",
                     mass_solver=solver) for r in rows]
        t0 = time.perf_counter()
        got = B.solve_many(runs)
        say(f"{solver}: 756 runs in {time.perf_counter() - t0:.1f} s (precompute, upload, batch, norms)")
        dev = [(abs(getattr(e, a) - float(r[b])), float(r[b]), r["q"], r["pairing"], r["flux"], r["N"], a)
               for e, r in zip(got, rows) for a, b in keys]
        say(f"{solver}: largest absolute deviation {max(dev)}")
        say(f"{solver}: largest relative deviation above 1e-9 "
            f"{max((d / g, g, *rest) for d, g, *rest in dev if g > 1e-9)}")
        for d, g, *rest in dev:
            assert d <= rtol * g + 1e-12, (solver, d, g, rest)
        del got
        gc.collect()
