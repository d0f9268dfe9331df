"""The device adjoint (wedg_adj_*, through the C ABI) against the CPU oracle
(oracle/adjoint.py, pinned by complex-step derivatives and by the C oracle):
the forward traces and misfit, one right-hand side's vector-Jacobian and
Jacobian-vector products, the dot-product test between the two, the
checkpointed trajectory gradient, and central differences (eps = 1e-6) of the
forward along random directions, on a small mesh against the C oracle and at
the full receiver-misfit workload (64 x 64, q = 4, 2000 steps, 32 snapshots)
against the device forward."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIDES = dict(bot="absorbing_2beta", top="wall", lft="dirichlet", rgt="absorbing")
FLUX = dict(theta=0.7, tau=0.4, lambda_u=0.9, lambda_v=1.1)
SRC = ((0.5, 0.35), 0.2, 1.0, 16 * math.pi, 1.0)


def rel(a, b):
    a = np.asarray(a); b = np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def setup(q, mixed, N, seed, lanes=1):
    import edg_westervelt_amd as w
    from edg_westervelt_amd.adjoint import SourceAdjoint
    from oracle import adjoint as A
    rng = np.random.default_rng(seed)
    qv = q - 1 if mixed else q
    al = 1e-3 * (1 + rng.random((N, N)))
    c2 = 1 + 0.3 * rng.random((N, N))
    beta = 2 + rng.random((N, N))
    x0, r0, amp, om, Ts = SRC
    adj = SourceAdjoint(q, N, mixed=mixed, flux_params=FLUX, sigma=0.3, boundary=SIDES, sigma_dirichlet=0.6, kappa=0.5,
                        alpha=al, source=w.GaussianSource(x0=x0, r0=r0, amplitude=amp, omega=om, Ts=Ts), lanes=lanes)
    adj.set_params(c2, beta)
    sides = [A.side_params(SIDES[s], 0.3, 0.6, 0.5) for s in ("bot", "top", "lft", "rgt")]
    pb = A.Problem(q, qv, N, flux=tuple(FLUX.values()), sides=sides, alpha=al, source=SRC)
    return adj, pb, c2, beta, al, rng


@pytest.mark.parametrize("q,mixed", [(3, False), (3, True), (4, False)])
def test_forward_traces_and_misfit(q, mixed):
    from oracle import adjoint as A
    adj, pb, c2, beta, _, _ = setup(q, mixed, 6, 1)
    dt, ns = 2e-3, 25
    tr = adj.traces(dt, ns)
    J0, otr = A.misfit(pb, c2, beta, dt, ns, traces=True)
    assert rel(tr, otr) < 1e-10
    tgt = A.misfit(pb, np.ones_like(c2), 3 * np.ones_like(beta), dt, ns, traces=True)[1]
    adj.set_target(tgt)
    J = adj.misfit(dt, ns)
    Jo = A.misfit(pb, c2, beta, dt, ns, tgt)
    assert Jo > 0 and abs(J - Jo) <= 1e-9 * Jo


@pytest.mark.parametrize("q,mixed", [(3, False), (3, True), (2, False), (5, False), (6, True)])
def test_rhs_vjp_and_tl_against_oracle(q, mixed):
    from oracle import adjoint as A
    N = 5
    adj, pb, c2, beta, _, rng = setup(q, mixed, N, 2 + q)
    u = 1e-2 * rng.standard_normal((N, N, pb.nu, pb.nu))
    v = 2e-3 * rng.standard_normal((N, N, pb.nv, pb.nv))      # 1 - 2 beta v_h > 0 at every point
    t = 0.013
    ub = rng.standard_normal(u.shape)
    vb = rng.standard_normal(v.shape)
    got = adj.rhs_vjp(u, v, t, ub, vb)
    ref = A.rhs_vjp(pb, c2, beta, u, v, t, ub, vb)
    for g, r, name in zip(got, ref, ("u", "v", "c2", "beta")):
        assert rel(g, r) < 1e-10, name
    d = [rng.standard_normal(x.shape) for x in (u, v, c2, beta)]
    tu, tv = adj.rhs_tl(u, v, t, *d)
    h = 1e-30
    cu, cv = A.rhs(pb, c2 + 1j * h * d[2], beta + 1j * h * d[3], u + 1j * h * d[0], v + 1j * h * d[1], t)
    assert rel(tu, cu.imag / h) < 1e-10 and rel(tv, cv.imag / h) < 1e-10
    # dot-product test: <kbar, TL d> = <VJP kbar, d>, both on the device
    lhs = (ub * tu).sum() + (vb * tv).sum()
    rhs = sum((g * x).sum() for g, x in zip(got, d))
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), abs(rhs))


def _closed_form_advances(n, snaps):
    """Forward steps of binomial checkpointing with `snaps` slots (the first state
    held in one of them): r n - C(c + r, r - 1), c = snaps, r the least with
    C(c + r, r) >= n."""
    if n == 1:
        return 0
    if snaps == 1:
        return n * (n - 1) // 2
    r = 0
    while math.comb(snaps + r, r) < n:
        r += 1
    return r * n - math.comb(snaps + r, r - 1)


@pytest.mark.parametrize("q,mixed,N,ns", [(3, False, 6, 30), (4, False, 4, 16), (3, True, 5, 12)])
def test_gradient_against_oracle_and_checkpointing(q, mixed, N, ns):
    from oracle import adjoint as A
    adj, pb, c2, beta, _, _ = setup(q, mixed, N, 3)
    dt = 2e-3
    tgt = A.misfit(pb, np.ones_like(c2), 3 * np.ones_like(beta), dt, ns, traces=True)[1]
    adj.set_target(tgt)
    J, gc2, gb = adj.gradient(dt, ns, snapshots=4)
    st = adj.stats()
    assert st.taped_steps == ns and st.adjoint_steps == ns
    assert st.forward_steps == _closed_form_advances(ns, 4)
    assert st.snapshots_used <= 4
    assert st.tapes == min(32, ns)   # small tapes: one per step up to 32
    Jo, oc2, ob = A.gradient(pb, c2, beta, dt, ns, tgt)
    assert abs(J - Jo) <= 1e-9 * Jo
    assert rel(gc2, oc2) < 1e-9 and rel(gb, ob) < 1e-9
    # every schedule gives the same numbers: the recomputed states are bitwise the forward's
    for snaps in (1, 2, ns + 1):
        J2, gc2b, gbb = adj.gradient(dt, ns, snapshots=snaps)
        assert J2 == J and np.array_equal(gc2b, gc2) and np.array_equal(gbb, gb)
        assert adj.stats().forward_steps == _closed_form_advances(ns, snaps)


@pytest.mark.parametrize("q,mixed,N,ns", [(3, False, 6, 30), (5, False, 5, 12), (6, True, 4, 10), (4, True, 5, 14)])
def test_gradient_with_two_lane_forward(q, mixed, N, ns):
    """The gradient with the forward steps on the two-lane stage kernels
    (wedg_adj_set_lanes_per_element): against the oracle's adjoint to 1e-9,
    against the one-lane gradient to 1e-11, and bitwise the same for every
    checkpointing schedule (the recomputed states are the forward's)."""
    from oracle import adjoint as A
    adj, pb, c2, beta, _, _ = setup(q, mixed, N, 3, lanes=2)
    assert adj.lanes == 2
    dt = 2e-3
    tgt = A.misfit(pb, np.ones_like(c2), 3 * np.ones_like(beta), dt, ns, traces=True)[1]
    adj.set_target(tgt)
    J, gc2, gb = adj.gradient(dt, ns, snapshots=4)
    Jo, oc2, ob = A.gradient(pb, c2, beta, dt, ns, tgt)
    assert abs(J - Jo) <= 1e-9 * Jo
    assert rel(gc2, oc2) < 1e-9 and rel(gb, ob) < 1e-9
    for snaps in (1, ns + 1):
        J2, gc2b, gbb = adj.gradient(dt, ns, snapshots=snaps)
        assert J2 == J and np.array_equal(gc2b, gc2) and np.array_equal(gbb, gb)
    assert adj.misfit(dt, ns) == J
    adj.set_lanes(1)
    assert adj.lanes == 1
    J1, gc21, gb1 = adj.gradient(dt, ns, snapshots=4)
    assert abs(J1 - J) <= 1e-11 * J and rel(gc2, gc21) < 1e-11 and rel(gb, gb1) < 1e-11


def test_gradient_against_central_differences_of_the_c_oracle(oracle_built):
    import edg_westervelt_amd as w
    from oracle import adjoint as A
    q, N = 3, 5
    adj, pb, c2, beta, al, rng = setup(q, False, N, 4)
    dt, ns = 2e-3, 20
    tgt = A.misfit(pb, np.ones_like(c2), 3 * np.ones_like(beta), dt, ns, traces=True)[1]
    adj.set_target(tgt)
    J, gc2, gb = adj.gradient(dt, ns, snapshots=5)
    x0, r0, amp, om, Ts = SRC

    def J_c(c2x, bx):
        pre, _ = w.precompute(q, N, mixed=False, mode="source", boundary=SIDES, sigma=0.3, sigma_dirichlet=0.6,
                              kappa=0.5, flux_params=FLUX,
                              coefficients=w.element_coefficients(c=np.sqrt(c2x), alpha=al, beta=bx),
                              source=w.GaussianSource(x0=x0, r0=r0, amplitude=amp, omega=om, Ts=Ts))
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

    assert abs(J_c(c2, beta) - J) <= 1e-9 * J
    eps = 1e-6
    for _ in range(4):
        dc, db = rng.standard_normal((N, N)), rng.standard_normal((N, N))
        fd = (J_c(c2 + eps * dc, beta + eps * db) - J_c(c2 - eps * dc, beta - eps * db)) / (2 * eps)
        ad = (gc2 * dc).sum() + (gb * db).sum()
        assert abs(fd - ad) <= 1e-6 * abs(ad), (fd, ad)


def test_full_workload_gradient_against_central_differences():
    """64 x 64, q = 4, 2000 steps, 32 snapshots: the gradient along 16 random
    directions against central differences (eps = 1e-6) of the device forward.
    The source is at (0.5, 0.15): from (0.5, 0.7) the wave front (c = 1) would
    reach the bottom side only after t = 0.7, about 7,200 steps, and over 2000
    steps the misfit there (about 1e-68) is only the scheme's numerical tail."""
    from edg_westervelt_amd.adjoint import beta_field_workload
    wl = beta_field_workload(x0=(0.5, 0.15))
    adj, N = wl.adj, wl.N
    J, gc2, gb = adj.gradient(wl.dt, wl.steps, snapshots=32)
    st = adj.stats()
    print(f"J = {J:.6e}, |dJ/dc2| = {np.abs(gc2).max():.3e}, |dJ/dbeta| = {np.abs(gb).max():.3e}, stats {st}")
    assert J > 0 and np.abs(gb).max() > 0 and np.abs(gc2).max() > 0
    assert st.forward_steps == _closed_form_advances(wl.steps, 32)
    assert abs(adj.misfit(wl.dt, wl.steps) - J) <= 1e-12 * J
    rng = np.random.default_rng(16)
    eps = 1e-6
    worst = 0.0
    for _ in range(16):
        dc, db = rng.standard_normal((N, N)), rng.standard_normal((N, N))
        adj.set_params(wl.c2 + eps * dc, wl.beta + eps * db)
        Jp = adj.misfit(wl.dt, wl.steps)
        adj.set_params(wl.c2 - eps * dc, wl.beta - eps * db)
        Jm = adj.misfit(wl.dt, wl.steps)
        fd = (Jp - Jm) / (2 * eps)
        ad = (gc2 * dc).sum() + (gb * db).sum()
        worst = max(worst, abs(fd - ad) / abs(ad))
    adj.set_params(wl.c2, wl.beta)
    print(f"worst relative difference over 16 directions: {worst:.3e}")
    assert worst <= 1e-6


def test_full_workload_as_specified():
    """The workload with the source at (0.5, 0.7): the gradient run completes
    with the closed-form number of forward steps, J is the forward's misfit
    bitwise, and a second run repeats it bitwise."""
    from edg_westervelt_amd.adjoint import beta_field_workload
    wl = beta_field_workload()
    J, gc2, gb = wl.adj.gradient(wl.dt, wl.steps, snapshots=32)
    assert wl.adj.stats().forward_steps == _closed_form_advances(wl.steps, 32)
    assert J == wl.adj.misfit(wl.dt, wl.steps) and J > 0
    J2, gc2b, gbb = wl.adj.gradient(wl.dt, wl.steps, snapshots=32)
    assert J2 == J and np.array_equal(gc2b, gc2) and np.array_equal(gbb, gb)


def _batch_cases(lanes=(1, 1, 1, 1)):
    """Instances of two degree groups, uneven meshes and step counts."""
    from oracle import adjoint as A
    out = []
    for k, (q, mixed, N, ns) in enumerate([(3, False, 6, 30), (4, False, 4, 16), (3, True, 5, 12),
                                           (3, False, 8, 21)]):
        adj, pb, c2, beta, _, _ = setup(q, mixed, N, 20 + k, lanes=lanes[k])
        assert adj.lanes == lanes[k]
        dt = 2e-3 if k != 3 else 1.5e-3
        tgt = A.misfit(pb, np.ones_like(c2), 3 * np.ones_like(beta), dt, ns, traces=True)[1]
        adj.set_target(tgt)
        out.append((adj, dt, ns))
    return out


@pytest.mark.parametrize("snaps,lanes", [(4, (1, 1, 1, 1)), (1, (1, 1, 1, 1)), (4, (2, 1, 2, 2)), (1, (2, 2, 1, 2))])
def test_batched_gradient_equals_each_instances_own(snaps, lanes):
    """wedg_adj_batch_gradient over four instances (two degree groups, uneven N
    and step counts, forward kernels with one or two lanes per element: the
    groups split by lanes) gives every instance bitwise its own
    wedg_adj_gradient, with the same checkpointing counts;
    wedg_adj_batch_forward gives bitwise each instance's misfit."""
    from edg_westervelt_amd.adjoint import AdjointBatch
    cases = _batch_cases(lanes)
    own, counts = [], []
    for adj, dt, ns in cases:
        own.append(adj.gradient(dt, ns, snapshots=snaps))
        s = adj.stats()
        counts.append((s.forward_steps, s.taped_steps, s.adjoint_steps, s.snapshots_taken, s.snapshots_used))
        assert s.forward_steps == _closed_form_advances(ns, snaps)
    batch = AdjointBatch([c[0] for c in cases])
    res = batch.gradient([c[1] for c in cases], [c[2] for c in cases], snapshots=snaps)
    for (adj, dt, ns), (J, gc2, gb), (Jb, gc2b, gbb), cnt in zip(cases, own, res, counts):
        assert Jb == J and J > 0
        assert np.array_equal(gc2b, gc2) and np.array_equal(gbb, gb)
        s = adj.stats()
        assert (s.forward_steps, s.taped_steps, s.adjoint_steps, s.snapshots_taken, s.snapshots_used) == cnt
    Jf = batch.forward([c[1] for c in cases], [c[2] for c in cases])
    for (adj, dt, ns), jf in zip(cases, Jf):
        assert jf == adj.misfit(dt, ns)
    dev_ms, enq_ms = batch.time()
    assert dev_ms > 0 and enq_ms > 0


def test_batched_gradient_reports_an_instances_failure():
    """An instance whose coefficients make a mass matrix indefinite during the
    run fails with the single run's error; the others' results are unchanged."""
    from edg_westervelt_amd.adjoint import AdjointBatch
    from edg_westervelt_amd._lib import WedgError
    cases = _batch_cases()[:2]
    good_adj, dt_g, ns_g = cases[1]
    ref = good_adj.gradient(dt_g, ns_g, snapshots=4)
    bad_adj, dt_b, ns_b = cases[0]
    bad_adj.set_params(1.0, 1e6)          # 1 - 2 beta v_h <= 0 once the source has made v
    with pytest.raises(WedgError) as e_single:
        bad_adj.gradient(dt_b, ns_b, snapshots=4)
    batch = AdjointBatch([bad_adj, good_adj])
    with pytest.raises(WedgError) as e_batch:
        batch.gradient([dt_b, dt_g], [ns_b, ns_g], snapshots=4)
    assert str(e_batch.value).split(": ", 1)[1] == str(e_single.value).split(": ", 1)[1]
    J, gc2, gb = good_adj.gradient(dt_g, ns_g, snapshots=4)
    assert J == ref[0] and np.array_equal(gc2, ref[1]) and np.array_equal(gb, ref[2])
