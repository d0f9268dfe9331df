"""Throughput of the device right-hand side + RK4 on the large single instance.

Workload (one MI355X): q_u = q_v = 5 on a uniform 256 x 256 mesh (65,536
elements, 36 modes per field per element, 4,718,592 DOF), mode :initial_value
with the Gaussian u0 = 0.05 exp(-|x - (1/2,1/2)|^2 / (2 0.05^2)), constant
coefficients c = 1, alpha = 1e-4, beta = 2, upwind flux, absorbing_2beta on all
sides, Chebyshev mass solve to 1e-14, FP64 RK4 with dt = h / (10 q_u^2)
(32,000 steps to T = 0.5). The scheme cannot reach T = 0.5 on this instance:
v_h grows until 1 - 2 beta v_h <= 0 at the centre element near t = 0.131
(between steps 8,400 and 8,800; the CPU oracle stops the same way on coarser
meshes), where the mass solve of the reference raises its positivity error.
The timed run is therefore the part of the run the scheme completes: 8,000 RK4
steps after a 20-step warmup (t from 3.1e-4 to 0.1253), over which the global
Chebyshev count rises from 8 to 18. A step is one RK4 step (four right-hand
sides) of the whole mesh; the state stays in HBM for the timed region.

Prints one JSON line: DOF-stages/s, the roofline of the dominant kernel from
HIP events on the workspace stream, and the CPU oracle (C/OpenMP restatement
of the reference's algorithm) timed on a bounded sample on this host.

The line also carries the batched speedup ("batched"): the 64 uneven
instances of edg_westervelt_amd.batch.uneven_instances(seed=0) (N in
{16..64}, q_u in 2..6, both pairings, every flux and boundary kind,
piecewise-constant coefficients with jumps, 240 to 5,760 steps to T = 0.25)
advanced together to T (one launch per RK stage per basis group, the groups on
concurrent streams) against the same 64 solves one after the other
(wedg_advance each), in aggregate DOF-stages/s; the batched states are checked
bitwise against the sequential ones.
"""
from __future__ import annotations

import argparse
import ctypes as C
import gc
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0       # MI355X HBM3E, GB/s
FP64_PEAK_TFS = 78.6        # MI355X FP64 (vector = matrix), TFLOP/s
METRIC = "FP64 DOF·stages/s for edg_rhs!+RK4"
FAMILIES = ("prep", "stage")   # k_prep (first stage of a call); k_stage (one RK stage)


def gaussian_profile(s=0.05, amp=0.05):
    def u0(x, y):
        return amp * np.exp(-((x - 0.5) ** 2 + (y - 0.5) ** 2) / (2 * s * s))

    return dict(u0=u0,
                u0_x=lambda x, y: -(x - 0.5) / s ** 2 * u0(x, y),
                u0_y=lambda x, y: -(y - 0.5) / s ** 2 * u0(x, y),
                u0_xx=lambda x, y: ((x - 0.5) ** 2 / s ** 4 - 1 / s ** 2) * u0(x, y),
                u0_yy=lambda x, y: ((y - 0.5) ** 2 / s ** 4 - 1 / s ** 2) * u0(x, y))


COEFFICIENTS = {"constant": "c=1 alpha=1e-4 beta=2",
                "paper": "paper_coefficients (smooth c, alpha, beta: one u-stiffness inverse per element)"}


def build_problem(q, N, coefficients="constant"):
    import edg_westervelt_amd as w
    coef = (w.constant_coefficients(c=1.0, alpha=1e-4, beta=2.0) if coefficients == "constant"
            else w.paper_coefficients())
    pre, geo = w.precompute(q, N, mixed=False, flux="upwind", boundary="absorbing_2beta",
                            mode="initial_value", initial_profile=gaussian_profile(), coefficients=coef)
    return pre, geo


def cpu_threads():
    """The host threads of the CPU baseline: every CPU this process may run on,
    capped by OMP_NUM_THREADS where the machine sets it (a one-GPU share of a
    shared host)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    cap = os.environ.get("OMP_NUM_THREADS", "")
    if cap.isdigit() and int(cap) > 0:
        n = min(n, int(cap))
    return max(1, n)


def flops_per_element_stage(nu, nv, nh, nc):
    """Algorithmic FP64 flops of one element and stage: traces, volume, u-solve,
    lifts, Chebyshev set-up and nc - 1 Chebyshev steps (the plain sum-factorized
    counts of the reference's algorithm). Returns (total, traces part)."""
    nq = nu
    traces = 2 * (4 * (nu * nu + nq * nu) + 8 * (nv * nv + nq * nv))
    base = (traces
            + 2 * (2 * (nq * nu * nu + nq * nq * nu) + 2 * (nq * nv * nv + nq * nq * nv)
                   + 2 * (nv * nq * nq + nv * nv * nq) + 2 * (nu * nq * nq + nu * nu * nq))
            + 2 * nu ** 4 + 4 * (3 * nq * (2 * nv + nu) + 3 * (2 * nv * nv + nu * nu))
            + 2 * (nh * nv * nv + nh * nh * nv))
    step = 4 * (nh * nv * nv + nh * nh * nv) + nh * nh + 6 * nv * nv
    return base + (nc - 1) * step, traces


def bench_batched(count=64, seed=0, sequential=True, cheb="global", lanes_per_element=(1, 2)):
    """Aggregate DOF·stages/s of the uneven batch: together vs one after the
    other, both with the lanes per element the tuner chose for each group."""
    import edg_westervelt_amd.batch as B
    from edg_westervelt_amd._lib import lib
    t = time.perf_counter()
    built = [B.build_instance(x, cheb=cheb) for x in B.uneven_instances(seed=seed, count=count)]
    t_build = time.perf_counter() - t
    work = sum(4 * b.nsteps * b.dof for b in built)
    bt = B.Batch([b.ws for b in built])
    bt.advance(0.0, [b.dt for b in built], [1] * count)          # warm-up step
    # launch parameters by the autotuner (32-step probes: lanes per element of
    # each kernel group, then every stream count and chunk)
    tune_ms, choice = bt.tune(0.0, [b.dt for b in built], probe_steps=32, lanes_per_element=lanes_per_element)
    lpe = bt.lanes_per_element()
    for b in built:                                              # back to t = 0
        B.set_zero_state(b.ws)
    print(f"# batched: {count} instances together", file=sys.stderr, flush=True)
    c0 = time.perf_counter()
    bt.advance(0.0, [b.dt for b in built], [b.nsteps for b in built])
    wb = time.perf_counter() - c0
    dev_ms = bt.device_ms()
    states = [B.get_state(b.ws) for b in built]
    seq = 0.0
    same = True
    longest = 0.0
    if sequential:
        print(f"# batched: {count} instances one after the other", file=sys.stderr, flush=True)
    for i, b in enumerate(built if sequential else []):
        B.set_zero_state(b.ws)
        tend = C.c_double(0.0)
        c0 = time.perf_counter()
        b.ws.dev.check(lib.wedg_advance(b.ws.dev.h, 0.0, b.dt, b.nsteps, 0, C.byref(tend)))
        d = time.perf_counter() - c0
        seq += d
        longest = max(longest, d)
        u, v = B.get_state(b.ws)
        same = same and np.array_equal(u, states[i][0]) and np.array_equal(v, states[i][1])
    groups = len({(b.inst.q, b.inst.mixed, b.ws.problem.pre.one_sided) for b in built})
    if not sequential:
        return dict(value=work / wb, unit="DOF·stages/s", wall_s=wb, device_s=dev_ms / 1e3, dof_stages=work,
                    instances=count, tuner=dict(lanes=choice[0], chunk_steps=choice[1],
                                                two_lane_instances=sum(1 for x in lpe if x == 2)))
    return dict(value=work / wb, unit="DOF·stages/s", sequential_value=work / seq, speedup=seq / wb,
                wall_s=wb, device_s=dev_ms / 1e3, sequential_s=seq, longest_instance_s=longest,
                dof_stages=work, instances=count, groups=groups, bitwise_equal_to_sequential=bool(same),
                workload=f"edg_westervelt_amd.batch.uneven_instances(seed={seed}, count={count}), T = 0.25, "
                         "initial_value Gaussian, piecewise-constant c, alpha, beta on 4x4 blocks",
                amplitude_rule=sorted({b.inst.amp_rule for b in built}),
                cheb_policy=cheb, build_s=round(t_build, 2),
                tuner=dict(choice=dict(lanes=choice[0], chunk_steps=choice[1]),
                           probe_ms={f"{k[0]}x{k[1]}": round(v, 3) for k, v in tune_ms.items()},
                           lanes_per_element={f"q{k[0]}/{k[1]} mode {k[2]}": v for k, v in bt.lane_choice.items()},
                           lanes_per_element_probe_ms={f"q{k[0]}/{k[1]} mode {k[2]} x{k[3]}": round(v, 3)
                                                       for k, v in bt.lane_times.items()},
                           two_lane_instances=sum(1 for x in lpe if x == 2)))


def bench_adjoint(steps=2000, snapshots=32, N=64, q=4, lanes=1):
    """Wall time of the receiver-misfit gradient (forward sweep, recomputations,
    taped steps and the adjoint of every stage, binomial checkpointing in
    `snapshots` slots) against one forward solve of the same run with the misfit,
    on edg_westervelt_amd.adjoint.beta_field_workload (q = 4, 64 x 64, 2000
    steps, per-element beta, source at (0.5, 0.7)); `lanes`: the forward
    stage kernels' lanes per element (both legs)."""
    from edg_westervelt_amd.adjoint import beta_field_workload
    t = time.perf_counter()
    wl = beta_field_workload(N=N, q=q, steps=steps, lanes=lanes)
    t_build = time.perf_counter() - t
    adj = wl.adj
    adj.misfit(wl.dt, 20)                          # warm-up of both paths
    adj.gradient(wl.dt, 20, snapshots)
    print(f"# adjoint: forward solve of {steps} steps", file=sys.stderr, flush=True)
    c0 = time.perf_counter()
    J = adj.misfit(wl.dt, steps)
    tf = time.perf_counter() - c0
    print(f"# adjoint: gradient with {snapshots} snapshots", file=sys.stderr, flush=True)
    c0 = time.perf_counter()
    Jg, gc2, gb = adj.gradient(wl.dt, steps, snapshots)
    tg = time.perf_counter() - c0
    st = adj.stats()
    dof = N * N * ((q + 1) ** 2 * 2)
    return dict(adjoint_forward_ratio=tg / tf, forward_s=tf, gradient_s=tg, J=J, J_gradient_run=Jg,
                forward_dof_stages_per_s=4 * steps * dof / tf, snapshots=snapshots,
                forward_steps_recomputed=st.forward_steps, taped_steps=st.taped_steps,
                snapshots_used=st.snapshots_used, tapes=st.tapes, build_s=round(t_build, 2), forward_lanes_per_element=adj.lanes,
                host_enqueue_ms=st.enqueue_ms, device_ms=st.total_ms,
                workload=f"edg_westervelt_amd.adjoint.beta_field_workload(N={N}, q={q}, steps={steps}): source "
                         "mode, Gaussian source width 0.05 at (0.5, 0.7), c = 1, alpha = 1e-5, per-element beta "
                         "4 + 2 sin(2 pi x) cos(2 pi y) + N(0, 0.1) (seed 1), bottom absorbing_2beta, walls with "
                         "kappa = 0.5; J = bottom-side L2 misfit after every step against the beta = 4 run")


def bench_batched_adjoint(T=0.02, count=16, seed=2, snapshots=32, sequential=True, lanes=1):
    """The batched adjoint: `count` source-mode instances of
    edg_westervelt_amd.adjoint.steepening_batch_workload (N in {64, 96, 128},
    q in {5, 6}, beta about 6-8), forward and checkpointed gradient of every
    instance together (one launch per RK stage and per adjoint kernel for each
    degree group), against the instances' own gradients one after another.
    The workload is defined to T = 1.0 (16,000-46,000 steps per instance); the
    timed runs stop at T (the same steps for both legs). `lanes`: the lanes
    per element of the forward stage kernels of every instance (both legs)."""
    from edg_westervelt_amd.adjoint import AdjointBatch, steepening_batch_workload
    t = time.perf_counter()
    wl = steepening_batch_workload(count=count, seed=seed, T=T, lanes=lanes)
    t_build = time.perf_counter() - t
    dts = [x.dt for x in wl]
    ns = [x.steps for x in wl]
    batch = AdjointBatch([x.adj for x in wl])
    batch.forward(dts, [min(n, 8) for n in ns])               # warm-up
    print(f"# batched adjoint: forward of {count} instances", file=sys.stderr, flush=True)
    c0 = time.perf_counter()
    Jf = batch.forward(dts, ns)
    tf = time.perf_counter() - c0
    print(f"# batched adjoint: gradients of {count} instances", file=sys.stderr, flush=True)
    c0 = time.perf_counter()
    res = batch.gradient(dts, ns, snapshots)
    tg = time.perf_counter() - c0
    dev_ms, enq_ms = batch.time()
    same = all(r[0] == j for r, j in zip(res, Jf))
    out = dict(batched_gradient_s=tg, batched_forward_s=tf, adjoint_forward_ratio=tg / tf, device_s=dev_ms / 1e3,
               host_enqueue_s=enq_ms / 1e3, instances=count, snapshots=snapshots, T=T,
               steps=ns, N=[x.N for x in wl], q=[x.q for x in wl], build_s=round(t_build, 2),
               J_equal_forward=bool(same), forward_lanes_per_element=[x.adj.lanes for x in wl],
               workload=f"edg_westervelt_amd.adjoint.steepening_batch_workload(count={count}, seed={seed}, T={T}): "
                        "source mode, N in {64, 96, 128}, q_u = q_v in {5, 6}, c = 1, alpha log-uniform "
                        "[1e-6, 1e-5], beta0 uniform [6, 8], source width 0.04 at (0.5, 0.3), amplitude for "
                        "1 - 2 beta0 max|v| about 0.3; gradient at beta0 + 0.5 sin(2 pi x) cos(2 pi y), target the "
                        "beta0 run; T shortened from the workload's 1.0")
    if sequential:
        print(f"# batched adjoint: the {count} gradients one after another", file=sys.stderr, flush=True)
        c0 = time.perf_counter()
        seq = [x.adj.gradient(x.dt, x.steps, snapshots) for x in wl]
        ts = time.perf_counter() - c0
        out.update(sequential_s=ts, speedup=ts / tg,
                   bitwise_equal_to_sequential=bool(all(a[0] == b[0] and np.array_equal(a[1], b[1]) and
                                                        np.array_equal(a[2], b[2]) for a, b in zip(res, seq))))
    return out


def bench_graph(chunk=16):
    """Plain stage launches against HIP-graph chunks of `chunk` RK4 steps
    (wedg_graph_enable) on meshes from launch-bound to the headline size, the
    same steps from the same state (bitwise the same end states)."""
    import edg_westervelt_amd as w
    from edg_westervelt_amd import _lib
    L = _lib.lib
    P = C.POINTER(C.c_double)
    out = []
    for q, N, steps in ((3, 16, 4000), (4, 64, 1000), (5, 128, 400)):
        pre, _ = w.precompute(q, N, mixed=False, flux="upwind", boundary="absorbing_2beta", mode="initial_value",
                              initial_profile=gaussian_profile(),
                              coefficients=w.constant_coefficients(c=1.0, alpha=1e-4, beta=0.0))
        ws = w.Workspace(pre, N)
        h = ws.dev.h
        dof = N * N * (pre.qu1 ** 2 + pre.qv1 ** 2)
        dt = 1.0 / N / (10 * q * q)
        u = np.zeros((N, N, pre.qu1, pre.qu1), order="F")
        v = np.zeros((N, N, pre.qv1, pre.qv1), order="F")
        res = {}
        for k in (0, chunk, 0, chunk):                       # each twice; the second run is timed
            ws.dev.check(L.wedg_graph_enable(h, k))
            ws.dev.check(L.wedg_set_state(h, u.ctypes.data_as(P), v.ctypes.data_as(P)))
            tend = C.c_double(0.0)
            ws.dev.check(L.wedg_sync(h))
            c0 = time.perf_counter()
            ws.dev.check(L.wedg_advance(h, 0.0, dt, steps, 0, C.byref(tend)))
            el = time.perf_counter() - c0
            uo = np.zeros_like(u)
            vo = np.zeros_like(v)
            ws.dev.check(L.wedg_get_state(h, uo.ctypes.data_as(P), vo.ctypes.data_as(P)))
            res[k] = (el, uo, vo)
        same = np.array_equal(res[0][1], res[chunk][1]) and np.array_equal(res[0][2], res[chunk][2])
        out.append(dict(q=q, N=N, steps=steps, plain_dof_stages_per_s=4 * steps * dof / res[0][0],
                        graph_dof_stages_per_s=4 * steps * dof / res[chunk][0], speedup=res[0][0] / res[chunk][0],
                        us_per_stage_plain=1e6 * res[0][0] / (4 * steps),
                        us_per_stage_graph=1e6 * res[chunk][0] / (4 * steps), bitwise_equal=bool(same)))
        print(f"# graph q={q} N={N}: {out[-1]['speedup']:.2f}x", file=sys.stderr, flush=True)
    return dict(chunk_steps=chunk, runs=out, workload="initial_value Gaussian, c=1 alpha=1e-4 beta=0, upwind, "
                                                        "absorbing_2beta (linear: no positivity limit over the steps)")


def reference_runs():
    """Wall times (precompute, upload, run and results, on this machine) of the
    runs the reference times itself, as it states them: README.md:244-245,
    313-315 and data/convergence/results.csv; its figures are on its own
    machine (Apple M4 Pro for the README), given here as context."""
    import csv
    import edg_westervelt_amd as w
    from edg_westervelt_amd import batch as B
    from tests.cases import BUMP, STEEPENING_RUN

    def timed(f):
        t = time.perf_counter()
        f()
        return time.perf_counter() - t

    out = {}
    w.solve(2, 4, 0.01)   # loads the library and the device once
    print("# reference runs: cfg1, solve(4, 32, 0.1), source 64x64", file=sys.stderr, flush=True)
    out["cfg1 solve(3, 10, 0.1; flux=:upwind, mixed=true)"] = dict(
        s=timed(lambda: w.solve(3, 10, 0.1, flux="upwind", mixed=True)),
        reference="0.1 s (data/convergence/results.csv:364, one Julia thread)")
    out["solve(4, 32, 0.1)"] = dict(s=timed(lambda: w.solve(4, 32, 0.1)),
                                    reference="1.2 s with 8 threads, 3.4 s serial (README.md:313-314, M4 Pro)")
    out["solve(4, 64, 0.05; mode=:source)"] = dict(
        s=timed(lambda: w.solve(4, 64, 0.05, mode="source")),
        reference="2.5 s with 8 threads, 13 s serial (README.md:314-315, M4 Pro)")
    R = STEEPENING_RUN
    steep = dict(flux="upwind", mixed=True, mode="source",
                 source=w.GaussianSource(x0=(0.15, 0.5), r0=0.05, omega=16 * math.pi, Ts=1.0),
                 record_every=R["record_every"], receivers=list(R["receivers"]), snap_times=list(R["snap_times"]))
    cases = {"nl": (8.0, "absorbing_2beta"), "lin": (0.0, "absorbing_2beta"), "nl_dbc": (8.0, "dirichlet")}
    print("# reference runs: steepening", file=sys.stderr, flush=True)
    out["steepening nl (experiments/steepening/run.jl)"] = dict(
        s=timed(lambda: w.solve(R["q"], R["N"], R["T"], boundary="absorbing_2beta",
                                coefficients=w.constant_coefficients(c=1.0, alpha=1e-6, beta=8.0), **steep)),
        reference="about 2.5 min with 8 threads, 11 min on one (README.md:245)")
    runs = [dict(q=R["q"], N=R["N"], T=R["T"], boundary=b,
                 coefficients=w.constant_coefficients(c=1.0, alpha=1e-6, beta=beta), **steep)
            for beta, b in cases.values()]
    out["steepening, all three cases in one batch (run_source_many)"] = dict(
        s=timed(lambda: B.run_source_many(runs)), reference="about 2.5 min per case with 8 threads")

    def alpha_table(many):
        alphas = [0.0, 1e-5, 1e-6, 1e-7, 1e-8, 1e-9]
        kw = [dict(q=6, N=20, T=0.5, flux="upwind", mixed=True, mode="initial_value", initial_profile=BUMP,
                   coefficients=w.replace_coefficients(w.paper_coefficients(), alpha=a), return_state=True)
              for a in alphas]
        if many:
            res = B.solve_many(kw)
        else:
            res = []
            for k in kw:
                k = dict(k)
                res.append(w.solve(k.pop("q"), k.pop("N"), k.pop("T"), **k))
        return [w.difference_norms(r, res[0]) for r in res[1:]]

    print("# reference runs: alpha table", file=sys.stderr, flush=True)
    out["alpha -> 0 table (experiments/alpha_limit/run.jl), runs one after the other"] = dict(
        s=timed(lambda: alpha_table(False)), reference="about 1 min with 8 threads, 2-3 min on one (README.md:244)")
    out["alpha -> 0 table, runs in one batch (solve_many)"] = dict(s=timed(lambda: alpha_table(True)),
                                                                  reference="as above")
    rows = list(csv.DictReader(open(os.path.join(ROOT, "tests", "golden", "convergence_results.csv"))))
    conv = [dict(q=int(r["q"]), N=int(r["N"]), T=0.1, flux=r["flux"], mixed=r["pairing"] == "mixed") for r in rows]
    print("# reference runs: convergence study", file=sys.stderr, flush=True)
    out["convergence study, 756 solves (experiments/convergence/run_all.sh) in one batch (solve_many)"] = dict(
        s=timed(lambda: B.solve_many(conv)),
        reference="about 5 min with 8 processes, 36 min on one core (README.md:240); results.csv secs sum "
                  "to 3,426 s")
    for v in out.values():
        v["s"] = round(v["s"], 3)
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8000,
                    help="timed RK4 steps (default: the 8,000 steps the run completes before the positivity "
                         "limit near t = 0.131)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cheb", choices=("global", "element"), default="global")
    ap.add_argument("--lanes", type=int, choices=(1, 2), default=1,
                    help="lanes per element of the stage kernels of the headline run")
    ap.add_argument("--cpu-steps", type=int, default=192, help="RK4 steps of the CPU oracle sample (0: skip)")
    ap.add_argument("--traffic", default=None, help="JSON with measured HBM bytes per launch (rocprofv3 --pmc)")
    ap.add_argument("--q", type=int, default=5)
    ap.add_argument("--coefficients", choices=tuple(COEFFICIENTS), default="constant",
                    help="constant (the headline) or the smooth paper_coefficients (per-element records)")
    ap.add_argument("--N", type=int, default=256)
    ap.add_argument("--batched", choices=("on", "off", "only"), default="on",
                    help="the uneven-batch speedup: with the headline (on), not at all (off), or alone (only)")
    ap.add_argument("--batched-sequential", choices=("on", "off"), default="on",
                    help="off: time only the batched run (for a kernel trace of it)")
    ap.add_argument("--batched-lanes", choices=("tune", "1", "2"), default="tune",
                    help="lanes per element of the batch's instances: the tuner's choice per kernel group, or fixed")
    ap.add_argument("--adjoint", choices=("on", "off", "only"), default="on",
                    help="the receiver-misfit gradient and its cost against one forward solve: with the headline "
                         "(on), not at all (off), or alone (only)")
    ap.add_argument("--batched-adjoint", choices=("on", "off", "only"), default="on",
                    help="the 16-instance batched adjoint against the instances' own gradients: with the headline "
                         "(on), not at all (off), or alone (only)")
    ap.add_argument("--adjoint-lanes", type=int, choices=(1, 2), default=2,
                    help="lanes per element of the forward stage kernels in the adjoint legs (2: the two-lane "
                         "kernels, which the adjoint workloads admit)")
    ap.add_argument("--batched-adjoint-T", type=float, default=0.02,
                    help="end time of the batched-adjoint runs (the workload is defined to 1.0)")
    ap.add_argument("--graph", choices=("on", "off", "only"), default="on",
                    help="plain stage launches against HIP-graph chunks on small to large meshes")
    ap.add_argument("--reference-runs", action="store_true",
                    help="only time the runs the reference times itself (README, data files) and exit")
    ap.add_argument("--probe-copy", type=float, default=0.0,
                    help="only run a plain copy kernel of this many MiB (counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); multi-GPU runs are not part of this build",
              file=sys.stderr)
        sys.exit(2)
    blpe = (1, 2) if args.batched_lanes == "tune" else (int(args.batched_lanes),)

    import edg_westervelt_amd as w
    from edg_westervelt_amd import _lib
    from edg_westervelt_amd.solver import time_step

    if args.probe_copy > 0:
        n = int(args.probe_copy * (1 << 20) / 8)
        ms = C.c_double(0.0)
        st = _lib.lib.wedg_probe_copy(n, 10, C.byref(ms))
        if st != 0:
            raise SystemExit(f"copy probe failed ({st})")
        print(json.dumps(dict(copy_probe_bytes_each_way=8 * n, launches=11, ms_avg=ms.value,
                              gbs=2 * 8 * n / ms.value / 1e6)))
        return
    if args.reference_runs:
        print(json.dumps(dict(metric="wall time of the reference's own timed runs", unit="s", n_gpus=1,
                              higher_is_better=False, dtype="f64", data="synthetic", runs=reference_runs())))
        return
    if args.adjoint == "only":
        ad = bench_adjoint(lanes=args.adjoint_lanes)
        print(json.dumps(dict(metric="adjoint/forward cost ratio", value=ad["adjoint_forward_ratio"], unit="x",
                              n_gpus=1, higher_is_better=False, dtype="f64", data="synthetic", adjoint=ad)))
        return
    if args.graph == "only":
        print(json.dumps(dict(metric="HIP graph speedup of RK4 chunks", value=None, unit="x", n_gpus=1,
                              higher_is_better=True, dtype="f64", data="synthetic", graph=bench_graph())))
        return
    if args.batched_adjoint == "only":
        ba = bench_batched_adjoint(T=args.batched_adjoint_T, lanes=args.adjoint_lanes)
        print(json.dumps(dict(metric="batched adjoint speedup (16 instances)", value=ba.get("speedup"), unit="x",
                              n_gpus=1, higher_is_better=True, dtype="f64", data="synthetic", batched_adjoint=ba)))
        return
    if args.batched == "only":
        bb = bench_batched(sequential=args.batched_sequential == "on", cheb=args.cheb, lanes_per_element=blpe)
        print(json.dumps(dict(metric="batched speedup (uneven instances)", value=bb.get("speedup"), unit="x",
                              n_gpus=1, higher_is_better=True, dtype="f64", data="synthetic", batched=bb)))
        return

    q, N, T = args.q, args.N, 0.5
    t_pre = time.perf_counter()
    pre, geo = build_problem(q, N, args.coefficients)
    t_pre = time.perf_counter() - t_pre
    dt, nsteps_full = time_step(q, N, T)
    steps = args.steps
    E = N * N
    nu, nv, nh = pre.qu1, pre.qv1, pre.q_high
    dof = E * (nu * nu + nv * nv)

    ws = w.Workspace(pre, N, cheb=args.cheb, lanes=args.lanes)
    h = ws.dev.h
    L = _lib.lib
    u = np.zeros((N, N, nu, nu), order="F")
    v = np.zeros((N, N, nv, nv), order="F")
    ws.dev.check(L.wedg_set_state(h, u.ctypes.data_as(C.POINTER(C.c_double)), v.ctypes.data_as(C.POINTER(C.c_double))))
    tend = C.c_double(0.0)
    print(f"# warmup {args.warmup} steps", file=sys.stderr, flush=True)
    ws.dev.check(L.wedg_advance(h, 0.0, dt, args.warmup, 0, C.byref(tend)))
    t0 = tend.value
    uw = np.zeros_like(u)
    vw = np.zeros_like(v)
    ws.dev.check(L.wedg_get_state(h, uw.ctypes.data_as(C.POINTER(C.c_double)), vw.ctypes.data_as(C.POINTER(C.c_double))))

    # timed region: the state is already in HBM; synchronized on both sides
    print(f"# timing {steps} steps", file=sys.stderr, flush=True)
    ws.dev.check(L.wedg_sync(h))
    c0 = time.perf_counter()
    ws.dev.check(L.wedg_advance_async(h, t0, dt, steps, 0, C.byref(tend)))
    ws.dev.check(L.wedg_sync(h))
    c1 = time.perf_counter()
    elapsed = c1 - c0
    mean_nc = C.c_double(0.0)
    nst = C.c_long(0)
    L.wedg_cheb_stats(h, C.byref(mean_nc), C.byref(nst))
    info = ws.dev.last_error()

    # per-kernel durations: HIP events around every launch on the workspace
    # stream, over the same steps from the same state as the timed region (in
    # two halves; the CPU sample starts from the state at the middle)
    ws.dev.check(L.wedg_set_state(h, uw.ctypes.data_as(C.POINTER(C.c_double)), vw.ctypes.data_as(C.POINTER(C.c_double))))
    L.wedg_timing_enable(h, 1)
    h1 = steps // 2
    ws.dev.check(L.wedg_advance(h, t0, dt, h1, 0, C.byref(tend)))
    L.wedg_cheb_stats(h, C.byref(mean_nc), C.byref(nst))
    nc1, ns1 = mean_nc.value, nst.value
    t_mid = tend.value
    um = np.zeros_like(u)
    vm = np.zeros_like(v)
    ws.dev.check(L.wedg_get_state(h, um.ctypes.data_as(C.POINTER(C.c_double)), vm.ctypes.data_as(C.POINTER(C.c_double))))
    ws.dev.check(L.wedg_advance(h, t_mid, dt, steps - h1, 0, C.byref(tend)))
    fam = {}
    for i, name in enumerate(FAMILIES):
        ms = C.c_double(0.0)
        n = C.c_long(0)
        L.wedg_timing_read(h, i, C.byref(ms), C.byref(n))
        fam[name] = dict(ms_total=ms.value, launches=n.value, us_avg=1e3 * ms.value / max(1, n.value))
    L.wedg_timing_enable(h, 0)
    L.wedg_cheb_stats(h, C.byref(mean_nc), C.byref(nst))
    nc_timing = (nc1 * ns1 + mean_nc.value * nst.value) / max(1, ns1 + nst.value)

    # final state is finite (the run is meaningful)
    ws.dev.check(L.wedg_get_state(h, u.ctypes.data_as(C.POINTER(C.c_double)), v.ctypes.data_as(C.POINTER(C.c_double))))
    finite = bool(np.isfinite(u).all() and np.isfinite(v).all())

    stages = 4 * steps
    value = dof * stages / elapsed
    ms_per_step = 1e3 * elapsed / steps

    # roofline of k_stage (one launch per RK stage, FP64-bound): the
    # algorithmic flops per element and stage over its average duration
    f_total, f_traces = flops_per_element_stage(nu, nv, nh, nc_timing)
    st_s = fam["stage"]["us_avg"] * 1e-6
    lanes = ws.dev.lanes
    kname = "k_stage2" if lanes == 2 else "k_stage"
    achieved = E * f_total / st_s / 1e12
    roof = dict(bound="mfma", achieved=achieved, peak=FP64_PEAK_TFS, unit="TFLOP/s",
                frac=achieved / FP64_PEAK_TFS, traffic=None, kernel=kname,
                flops_per_launch=E * f_total, us_avg=fam["stage"]["us_avg"])
    bytes_stage = 8 * 17 * dof / 4
    note = "compulsory bytes of the three-register RK4, 34 B per DOF"
    pmode = ws.problem.info()["mode"]
    if pmode in (0, 3):   # per-element records: Cw, Aw, the packed u-stiffness inverse, beta at the points
        npk = (nu * nu * (nu * nu + 1)) // 2
        bytes_stage += 8 * E * (2 * nu * nu + npk + nh * nh)
        note += f" + the per-element records, {8 * (2 * nu * nu + npk + nh * nh)} B per element"
    if not pre.mms_on and np.any(pre.PiCdiv):   # the static forcing of :initial_value mode, read every stage
        bytes_stage += 8 * E * nv * nv
        note += f" + the forcing projection PiCdiv, {8 * nv * nv} B per element"
    roof_hbm = dict(bound="hbm", achieved=bytes_stage / st_s / 1e9, peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=bytes_stage / st_s / 1e9 / HBM_PEAK_GBS, kernel=kname,
                    bytes_per_launch=bytes_stage, note=note)
    sfx = "" if args.coefficients == "constant" else "_" + args.coefficients
    traffic_file = args.traffic or os.path.join(ROOT, "profiles", f"traffic_q{q}_n{N}{sfx}.json")
    if os.path.exists(traffic_file):
        tj = json.load(open(traffic_file))
        roof["traffic"] = tj.get(kname)
        roof_hbm["traffic"] = tj.get(kname)
    stage_s = elapsed / (4 * steps)
    bytes_step = 8 * 17 * dof
    whole = dict(flops_per_stage=E * f_total, tflops=E * f_total / stage_s / 1e12,
                 fp64_frac=E * f_total / stage_s / 1e12 / FP64_PEAK_TFS,
                 hbm_gbs_model=bytes_step / 4 / stage_s / 1e9,
                 hbm_frac_model=bytes_step / 4 / stage_s / 1e9 / HBM_PEAK_GBS,
                 mean_cheb_steps=nc_timing, max_cheb_steps=info.cheb_steps)

    # The other legs start from a process that holds no workspace of this one:
    # a live workspace keeps its stream, and with the process's four hardware
    # queues (GPU_MAX_HW_QUEUES) two of the gradient's three streams can then
    # share a queue, whose event waits hold up the other stream's launches.
    del ws
    gc.collect()
    batched = bench_batched(cheb=args.cheb, lanes_per_element=blpe) if args.batched == "on" else None
    adjoint = bench_adjoint(lanes=args.adjoint_lanes) if args.adjoint == "on" else None
    batched_adjoint = (bench_batched_adjoint(T=args.batched_adjoint_T, lanes=args.adjoint_lanes)
                       if args.batched_adjoint == "on" else None)
    graph = bench_graph() if args.graph == "on" else None

    # CPU baseline: the oracle (C/OpenMP restatement of the reference) on the same
    # problem, a bounded sample of RK4 steps from the same initial state
    cpu = None
    if args.cpu_steps > 0:
        from oracle import batched as ob
        ob.build() if not os.path.exists(ob._SO) else None
        threads = cpu_threads()
        op = ob.OraclePre(pre, N)
        print(f"# cpu oracle {args.cpu_steps} steps on {threads} threads", file=sys.stderr, flush=True)
        c0 = time.perf_counter()
        ob.rk4_steps(op, um, vm, t_mid, dt, args.cpu_steps, cheb_mode=0 if args.cheb == "global" else 1,
                     nthreads=threads)
        cs = time.perf_counter() - c0
        cpu = dict(value=dof * 4 * args.cpu_steps / cs, unit="DOF·stages/s", cores=threads, kind="port",
                   sample=f"{args.cpu_steps} RK4 steps of the same instance from the state at the middle of the "
                          f"timed window (t = {t_mid:.4f}; {cs:.1f} s, C/OpenMP oracle)")

    out = dict(metric=METRIC, value=value, unit="DOF·stages/s", n_gpus=1, steps=steps, warmup=args.warmup,
               ms_per_step=ms_per_step, higher_is_better=True, scaling="weak", vs_baseline=None, dtype="f64",
               data="synthetic",
               config=dict(workload=f"q_u=q_v={q}, {N}x{N} mesh, initial_value Gaussian, "
                                    f"{COEFFICIENTS[args.coefficients]}, "
                                    f"upwind, absorbing_2beta, Chebyshev 1e-14, RK4 dt=h/(10 q^2)",
                           kernel_mode=["general", "separable", "uniform", "general_cg"][pmode],
                           elements=E, dof=dof, dt=dt, cheb_policy=args.cheb, lanes_per_element=lanes,
                           mean_cheb_steps=nc_timing,
                           max_cheb_steps=info.cheb_steps, full_run_steps=nsteps_full,
                           timed_window=[t0, t0 + steps * dt],
                           note=("the run stops at the positivity limit 1 - 2 beta v_h <= 0 near t = 0.131; the "
                                 "timed window is the part it completes") if args.coefficients == "constant" else None,
                           precompute_s=round(t_pre, 2)),
               roofline=roof, roofline_hbm=roof_hbm, whole_step=whole, kernels_us=fam, cpu_baseline=cpu,
               batched=batched, adjoint=adjoint, batched_adjoint=batched_adjoint, graph=graph,
               adjoint_forward_ratio=adjoint["adjoint_forward_ratio"] if adjoint else None,
               final_state_finite=finite)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
