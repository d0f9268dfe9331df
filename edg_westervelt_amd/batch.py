"""Batched instances on the device: many independent problems (uneven N, q_u,
q_v, coefficients, boundary kinds, time steps and step counts) advanced
together, one kernel launch per RK stage for each group of instances with the
same (q_u, q_v) and coefficient layout (``wedg_batch_*`` of include/wedg.h).

The reference advances one instance per ``solve`` (src/solver.jl:167-228);
every instance here gets the stage arguments its own ``wedg_advance`` would
give it, so a batched run returns bitwise the states of separate runs.

``uneven_instances`` draws the 64-instance workload of uneven problems
(initial-value mode, piecewise-constant coefficients on a 4 x 4 grid of
mesh-aligned blocks); ``solve_batch`` builds and runs such a set;
``solve_many`` is ``solve`` for a list of runs; ``run_source_many`` is
source-mode ``solve`` (time series, receivers, snapshots) for a list of runs.
"""
from __future__ import annotations

import ctypes as C
import json
import math
import os
from types import SimpleNamespace

import numpy as np

from . import _lib
from ._lib import WedgError, lib
from .coefficients import Coefficients, _zero
from .precompute import precompute
from .solver import Workspace, _ptr, time_step

__all__ = ["Batch", "piecewise_constant", "block_coefficients", "gaussian_profile", "uneven_instances",
           "build_instance", "solve_batch", "solve_many", "run_source_many"]


class Batch:
    """Workspaces advanced together (``wedg_batch_create``); each keeps its own
    state, which ``wedg_get_state`` / ``Workspace`` calls read as usual, and
    its own lanes per element (``Workspace(..., lanes=)``), which the batch's
    kernels use too."""

    def __init__(self, workspaces):
        self.workspaces = list(workspaces)
        self.h = None
        self._create()

    def _create(self):
        n = len(self.workspaces)
        arr = (C.c_void_p * n)(*[w.dev.h.value for w in self.workspaces])
        h = C.c_void_p()
        st = lib.wedg_batch_create(arr, n, C.byref(h))
        if st != _lib.OK:
            raise WedgError(f"wedg_batch_create failed with status {st}")
        self.h = h

    def _destroy(self):
        h = getattr(self, "h", None)
        if h is not None and h.value:
            lib.wedg_batch_destroy(h)
            self.h = None

    def __del__(self):
        self._destroy()

    def kernel_groups(self):
        """{(q_u, q_v, mode): [instance indexes]}: the instances that share
        kernels whatever their lanes per element."""
        out = {}
        for i, w in enumerate(self.workspaces):
            key = (w.qu1 - 1, w.qv1 - 1, w.problem.info()["mode"])
            out.setdefault(key, []).append(i)
        return out

    def set_lanes_per_element(self, lanes):
        """Lanes per element (1 or 2; one value, or one per instance) of the
        instances' kernels, the batch's and their own calls'; instances whose
        problem has no two-lane kernels keep one lane. The batch is formed
        again with its launch parameters."""
        n = len(self.workspaces)
        per = [int(lanes)] * n if np.ndim(lanes) == 0 else [int(x) for x in lanes]
        if len(per) != n:
            raise ValueError("one lane count per instance")
        cfg = self.config()
        self._destroy()
        for w, l in zip(self.workspaces, per):
            w.dev.set_lanes(l)
        self._create()
        self.configure(*cfg)

    def lanes_per_element(self):
        """The lanes per element each instance runs with."""
        return [w.dev.lanes for w in self.workspaces]

    def advance(self, t0, dt, nsteps, stop_on_nonfinite: bool = False, raise_errors: bool = True):
        """nsteps[i] RK4 steps of dt[i] from t0[i] for instance i. Returns the
        final times and the per-instance statuses; with ``raise_errors`` the
        first failing instance raises the reference's message."""
        n = len(self.workspaces)
        dt = np.ascontiguousarray(np.asarray(dt, dtype=np.float64))
        if dt.shape != (n,) or len(nsteps) != n or np.ndim(t0) not in (0, 1) or np.size(t0) not in (1, n):
            raise ValueError("t0, dt and nsteps need one entry per instance")
        t0 = np.ascontiguousarray(np.broadcast_to(np.asarray(t0, dtype=np.float64), (n,)))
        ns = (C.c_long * n)(*[int(x) for x in nsteps])
        tend = np.zeros(n)
        status = np.zeros(n, dtype=np.int32)
        st = lib.wedg_batch_advance(self.h, _ptr(t0), _ptr(dt), ns, int(stop_on_nonfinite), _ptr(tend),
                                    status.ctypes.data_as(C.POINTER(C.c_int)))
        if st != _lib.OK and not np.any(status != _lib.OK):   # the call ended before any instance's record
            raise WedgError(f"wedg_batch_advance failed with status {st}")
        if raise_errors:
            for i, w in enumerate(self.workspaces):
                if status[i] != _lib.OK:
                    w.dev.check(int(status[i]))
            if st != _lib.OK:
                raise WedgError(f"wedg_batch_advance failed with status {st}")
        return tend, status

    def configure(self, lanes: int, chunk_steps: int = 0, schedule=None):
        """Streams, argument chunk and schedule (0: every instance starts at the
        first step, 1: all end at the last; None keeps it) of the launches
        (``wedg_batch_configure``)."""
        if schedule is None:
            schedule = self.config()[2]
        st = lib.wedg_batch_configure(self.h, int(lanes), int(chunk_steps), int(schedule))
        if st != _lib.OK:
            raise WedgError(f"wedg_batch_configure failed with status {st}")

    def config(self):
        lanes, chunk, sched = C.c_int(0), C.c_long(0), C.c_int(0)
        lib.wedg_batch_config(self.h, C.byref(lanes), C.byref(chunk), C.byref(sched))
        return lanes.value, chunk.value, sched.value

    def tune(self, t0, dt, probe_steps: int = 32, lanes=(1, 2, 3, 4, 6, 8, 12, 16), chunks=(0, 4), repeats: int = 2,
             lanes_per_element=None):
        """Autotuner of the launch parameters: times ``probe_steps`` steps of the
        batch (device time between the batch's events, best of ``repeats``) for
        every (lanes, chunk) pair, keeps the fastest and puts every state back
        as it was. Returns {(lanes, chunk): ms} and the choice.
        ``lanes_per_element``: None keeps every instance's lanes per element;
        one value sets it for all; with more than one option the tuner first
        chooses, for each kernel group with two-lane kernels, the option whose
        probe of the group alone is fastest (``lane_times``, ``lane_choice``).
        Results then agree with one-lane runs to rounding, and bitwise with
        the instances' own runs after the choice."""
        saved = [get_state(w) for w in self.workspaces]
        n = len(self.workspaces)
        times = {}
        self.lane_times, self.lane_choice = {}, {}
        try:
            opts = [] if lanes_per_element is None else sorted({int(x) for x in np.atleast_1d(lanes_per_element)})
            if len(opts) == 1:
                self.set_lanes_per_element(opts[0])
            elif len(opts) > 1:
                per = self.lanes_per_element()
                for key, members in self.kernel_groups().items():
                    if not self.workspaces[members[0]].problem.info()["two_lane"]:
                        continue
                    steps = [0] * n
                    for i in members:
                        steps[i] = probe_steps
                    for lpe in opts:
                        for i in members:
                            per[i] = lpe
                        self.set_lanes_per_element(per)
                        best = math.inf
                        for _ in range(repeats):
                            for i in members:
                                w = self.workspaces[i]
                                w.dev.check(lib.wedg_set_state(w.dev.h, _ptr(saved[i][0]), _ptr(saved[i][1])))
                            self.advance(t0, dt, steps)
                            best = min(best, self.device_ms())
                        self.lane_times[key + (lpe,)] = best
                    pick = min(opts, key=lambda l: self.lane_times[key + (l,)])
                    self.lane_choice[key] = pick
                    for i in members:
                        per[i] = pick
                self.set_lanes_per_element(per)
            for nl in lanes:
                for ch in chunks:
                    self.configure(nl, ch)
                    best = math.inf
                    for _ in range(repeats):
                        for w, (u, v) in zip(self.workspaces, saved):
                            w.dev.check(lib.wedg_set_state(w.dev.h, _ptr(u), _ptr(v)))
                        self.advance(t0, dt, [probe_steps] * n)
                        best = min(best, self.device_ms())
                    times[(nl, ch)] = best
        finally:
            for w, (u, v) in zip(self.workspaces, saved):
                w.dev.check(lib.wedg_set_state(w.dev.h, _ptr(u), _ptr(v)))
        choice = min(times, key=times.get)
        self.configure(*choice)
        return times, choice

    def device_ms(self) -> float:
        """Device time of the last advance (HIP events on the batch stream), ms."""
        ms = C.c_double(0.0)
        lib.wedg_batch_time(self.h, C.byref(ms))
        return ms.value


def piecewise_constant(values) -> callable:
    """f(x, y) = values[bx, by] on a B x B grid of blocks of the unit square,
    bx = floor(B x) (the last block closed on the right), for numpy arrays or
    scalars."""
    vals = np.asarray(values, dtype=np.float64)
    B = vals.shape[0]

    def f(x, y):
        bx = np.clip(np.floor(np.asarray(x) * B).astype(int), 0, B - 1)
        by = np.clip(np.floor(np.asarray(y) * B).astype(int), 0, B - 1)
        return vals[bx, by]

    return f


def block_coefficients(c, alpha, beta) -> Coefficients:
    """Piecewise-constant c, alpha, beta (B x B block tables) with zero
    derivatives inside the blocks (the jumps are taken by ``precompute``'s
    one-sided face values)."""
    return Coefficients(piecewise_constant(c), _zero, _zero, piecewise_constant(alpha), _zero, _zero,
                        piecewise_constant(beta))


def gaussian_profile(x0=0.5, y0=0.5, s=0.05, amp=0.05) -> dict:
    """u0 = amp exp(-((x-x0)^2 + (y-y0)^2) / (2 s^2)) with its analytic first
    and second derivatives (the ``initial_profile`` of :initial_value mode)."""
    def u0(x, y):
        return amp * np.exp(-((x - x0) ** 2 + (y - y0) ** 2) / (2 * s * s))

    return dict(u0=u0,
                u0_x=lambda x, y: -(x - x0) / s ** 2 * u0(x, y),
                u0_y=lambda x, y: -(y - y0) / s ** 2 * u0(x, y),
                u0_xx=lambda x, y: ((x - x0) ** 2 / s ** 4 - 1 / s ** 2) * u0(x, y),
                u0_yy=lambda x, y: ((y - y0) ** 2 / s ** 4 - 1 / s ** 2) * u0(x, y))


KINDS = ("absorbing_2beta", "absorbing", "dirichlet", "wall")
FLUXES = ("upwind", "central", "alternating")


AMPLITUDE_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "uneven_amplitudes.json")


def _calibrated_amplitudes(seed: int, T: float):
    """The amplitudes of calibrate_amplitudes for (seed, T), or None."""
    try:
        with open(AMPLITUDE_FILE) as f:
            tab = json.load(f)
    except FileNotFoundError:
        return None
    for e in tab["runs"]:
        if e["seed"] == seed and e["T"] == T:
            return e["amp"]
    return None


def uneven_instances(seed: int = 0, count: int = 64, T: float = 0.25, amplitude: str = "auto"):
    """The uneven batched workload, drawn with numpy's default_rng(seed), per
    instance in this order: N in {16, 24, 32, 48, 64}; q_u in {2..6}; q_v =
    q_u - 1 or q_u; the flux (upwind, central, alternating); per 4 x 4 block c
    in U[0.8, 1.2], alpha log-uniform in [1e-6, 1e-3], beta in U[0, 8]; the
    Gaussian's centre in [0.3, 0.7]^2 and width in U[0.03, 0.08]; each side's
    kind from (absorbing_2beta, absorbing, dirichlet, wall); sigma in U[-1, 1]
    (Dirichlet sides take |sigma|, which must be >= 0) and kappa in U[0, 1].
    Every run goes to T.

    The amplitude (``amp_rule`` of each instance): ``"calibrated"`` takes the
    amplitudes of ``calibrate_amplitudes`` stored in uneven_amplitudes.json,
    which bring 1 - 2 beta_max max|v_h| over the whole run to 0.51 (max|v_h|
    over the points of the diagnostics, sampled every 16 steps; >= 0.5);
    ``"bound"`` keeps 1 - 2 beta |v| >= 0.5 with a margin of 4 on the
    free-space bound |v| <= c_max |grad u0|_max = c_max amp / (s sqrt(e)):
    amp = min(0.05, s sqrt(e) / (16 beta_max c_max)), the calibration's
    starting point; ``"auto"`` is "calibrated" for the (seed, T) and instances
    the file holds, else "bound"."""
    if amplitude not in ("auto", "calibrated", "bound"):
        raise ValueError(f"unknown amplitude rule {amplitude} (expected auto, calibrated or bound)")
    cal = None if amplitude == "bound" else _calibrated_amplitudes(seed, T)
    if amplitude == "calibrated" and (cal is None or len(cal) < count):
        raise ValueError(f"no calibrated amplitudes for seed {seed}, T = {T}, {count} instances")
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        N = int(rng.choice([16, 24, 32, 48, 64]))
        q = int(rng.integers(2, 7))
        mixed = bool(rng.integers(0, 2))
        flux = FLUXES[int(rng.integers(0, 3))]
        c = rng.uniform(0.8, 1.2, (4, 4))
        alpha = 10.0 ** rng.uniform(-6.0, -3.0, (4, 4))
        beta = rng.uniform(0.0, 8.0, (4, 4))
        x0, y0 = rng.uniform(0.3, 0.7, 2)
        s = float(rng.uniform(0.03, 0.08))
        kinds = [KINDS[int(k)] for k in rng.integers(0, 4, 4)]
        sigma = float(rng.uniform(-1.0, 1.0))
        kappa = float(rng.uniform(0.0, 1.0))
        amp = min(0.05, s * math.sqrt(math.e) / (16.0 * float(beta.max()) * float(c.max())))
        rule = "bound"
        if cal is not None and i < len(cal):
            amp, rule = float(cal[i]), "calibrated"
        out.append(SimpleNamespace(index=i, N=N, q=q, mixed=mixed, flux=flux, c=c, alpha=alpha, beta=beta,
                                   x0=float(x0), y0=float(y0), s=s, amp=amp, amp_rule=rule,
                                   boundary=dict(zip(("bot", "top", "lft", "rgt"), kinds)),
                                   sigma=sigma, sigma_dirichlet=abs(sigma), kappa=kappa, T=T))
    return out


def build_instance(inst, dt_fac=10, cheb: str = "global"):
    """precompute, a device workspace with the zero initial state (u - u0 = 0,
    v = 0) and the time step of one instance (src/solver.jl:167-213)."""
    pre, geo = precompute(inst.q, inst.N, mixed=inst.mixed, flux=inst.flux, sigma=inst.sigma,
                          boundary=inst.boundary, sigma_dirichlet=inst.sigma_dirichlet, kappa=inst.kappa,
                          coefficients=block_coefficients(inst.c, inst.alpha, inst.beta), mode="initial_value",
                          initial_profile=gaussian_profile(inst.x0, inst.y0, inst.s, inst.amp))
    dt, nsteps = time_step(inst.q, inst.N, inst.T, dt_fac)
    ws = Workspace(pre, inst.N, cheb=cheb)
    set_zero_state(ws)
    return SimpleNamespace(inst=inst, pre=pre, geo=geo, ws=ws, dt=dt, nsteps=nsteps,
                           dof=inst.N * inst.N * (pre.qu1 ** 2 + pre.qv1 ** 2))


def run_vmax(built, sample_every: int = 16):
    """Each built instance's run from the zero state in one batch, with
    max|v_h| at the points of the diagnostics (``wedg_diagnostics``) after
    every ``sample_every`` of its steps and after its last: returns the largest
    value per instance (inf where the run failed) and the statuses."""
    from .solver import _set_diag
    n = len(built)
    for x in built:
        _set_diag(x.ws, x.geo)
        set_zero_state(x.ws)
    b = Batch([x.ws for x in built])
    done = np.zeros(n, dtype=np.int64)
    t = np.zeros(n)
    vm = np.zeros(n)
    status = np.zeros(n, dtype=np.int32)
    bufs = [SimpleNamespace(E=np.zeros(x.ws.N ** 2), vmax=np.zeros(x.ws.N ** 2), nd=np.zeros(x.ws.N ** 2),
                            eta=np.zeros(x.ws.qv1), sl=np.zeros(x.ws.N * x.ws.qu1)) for x in built]
    while True:
        steps = [min(sample_every, x.nsteps - int(d)) if st == _lib.OK else 0
                 for x, d, st in zip(built, done, status)]
        if not any(steps):
            break
        tend, stat = b.advance(t, [x.dt for x in built], steps, raise_errors=False)
        for i, x in enumerate(built):
            if steps[i] == 0:
                continue
            if stat[i] != _lib.OK:
                status[i] = stat[i]
                vm[i] = math.inf
                continue
            t[i] = tend[i]
            done[i] += steps[i]
            f = bufs[i]
            x.ws.dev.check(lib.wedg_diagnostics(x.ws.dev.h, _ptr(f.E), _ptr(f.vmax), _ptr(f.nd), 0, _ptr(f.eta),
                                                _ptr(f.sl)))
            vm[i] = max(vm[i], float(f.vmax.max()))
    return vm, status


def calibrate_amplitudes(seed: int = 0, count: int = 64, T: float = 0.25, target: float = 0.49,
                         updates: int = 3, sample_every: int = 16, log=None):
    """Amplitudes of the uneven workload that bring d = 2 beta_max max|v_h|
    over the whole run (run_vmax) to ``target``, so that 1 - 2 beta_max
    max|v| >= 0.5: from the "bound" amplitudes, ``updates`` rescalings amp <-
    amp target / d (a run that fails takes 0.7 amp; the cap 0.05 stays), then
    runs that confirm d <= 0.5 for every instance, rescaling any that exceed
    it. Returns the amplitudes and the confirmed d."""
    insts = uneven_instances(seed=seed, count=count, T=T, amplitude="bound")
    amp = np.array([x.amp for x in insts])
    bmax = np.array([float(x.beta.max()) for x in insts])
    d = None
    for it in range(updates + 4):
        for x, a in zip(insts, amp):
            x.amp = float(a)
        built = [build_instance(x) for x in insts]
        vm, st = run_vmax(built, sample_every)
        d = 2.0 * bmax * vm
        if log:
            log(f"calibration pass {it}: d in [{d.min():.4f}, {d.max():.4f}], failed {int((st != 0).sum())}")
        del built
        if it >= updates and np.all(d <= 0.5):
            break
        scale = np.where(np.isfinite(d), target / np.maximum(d, 1e-300), 0.7)
        if it >= updates:   # confirmation: only the instances above the bound move
            scale = np.where(d > 0.5, np.minimum(scale, 0.98), 1.0)
        amp = np.minimum(0.05, amp * scale)
    return amp, d


def set_zero_state(ws: Workspace):
    u = np.zeros((ws.N, ws.N, ws.qu1, ws.qu1), order="F")
    v = np.zeros((ws.N, ws.N, ws.qv1, ws.qv1), order="F")
    ws.dev.check(lib.wedg_set_state(ws.dev.h, _ptr(u), _ptr(v)))


def get_state(ws: Workspace):
    u = np.zeros((ws.N, ws.N, ws.qu1, ws.qu1), order="F")
    v = np.zeros((ws.N, ws.N, ws.qv1, ws.qv1), order="F")
    ws.dev.check(lib.wedg_get_state(ws.dev.h, _ptr(u), _ptr(v)))
    return u, v


def solve_batch(instances, dt_fac=10, cheb: str = "global"):
    """Builds the instances and advances them together to their T. Returns the
    built instances (with ``u``, ``v``, ``t`` set) and the Batch."""
    built = [build_instance(x, dt_fac=dt_fac, cheb=cheb) for x in instances]
    b = Batch([x.ws for x in built])
    tend, _ = b.advance(0.0, [x.dt for x in built], [x.nsteps for x in built])
    for x, t in zip(built, tend):
        x.u, x.v = get_state(x.ws)
        x.t = float(t)
    return built, b


def solve_many(runs, cheb: str = "global"):
    """``solve`` for many runs at once: ``runs`` is a list of dicts with the
    arguments of ``solve`` (``q``, ``N``, ``T`` and its keywords); returns, per
    run, what ``solve`` returns for it — the errors at T in :mms mode, the
    final state with ``return_state=True`` — from states advanced together in
    one batch (bitwise the states of separate ``solve`` calls). :source mode,
    whose host loop stops at every record step, is not batched here."""
    from .solver import _device_mms_errors, _set_diag
    setups = []
    for r in runs:
        r = dict(r)
        q, N, T = r.pop("q"), r.pop("N"), r.pop("T")
        return_state = r.pop("return_state", False)
        dt_fac = r.pop("dt_fac", 10)
        nblocks = r.pop("nblocks", None)
        run_cheb = r.pop("cheb", cheb)
        for k in ("record_every", "slice_y", "receivers", "snap_times"):   # :source mode only
            r.pop(k, None)
        mode = str(r.get("mode", "mms")).lstrip(":")
        if mode not in ("mms", "initial_value", "source"):
            raise ValueError(f"unknown mode {mode} (expected :mms, :initial_value or :source)")
        if mode == "source":
            raise ValueError("solve_many does not batch :source mode; call solve for it")
        if mode != "mms" and not return_state:
            raise ValueError("errors are only available in :mms mode; use return_state = true")
        r["mode"] = mode
        pre, geo = precompute(q, N, **r)
        dt, nsteps = time_step(q, N, T, dt_fac)
        ws = Workspace(pre, N, nblocks=nblocks, cheb=run_cheb)
        set_zero_state(ws)
        setups.append(SimpleNamespace(q=q, N=N, T=T, pre=pre, geo=geo, ws=ws, dt=dt, nsteps=nsteps,
                                      return_state=return_state, kx=r.get("kx", 1), ky=r.get("ky", 1),
                                      phase=r.get("phase", (0.0, 0.0))))
    if not setups:
        return []
    b = Batch([s.ws for s in setups])
    tend, _ = b.advance(0.0, [s.dt for s in setups], [s.nsteps for s in setups])
    out = []
    for s, t in zip(setups, tend):
        h = 1.0 / s.N
        if s.return_state:
            u, v = get_state(s.ws)
            _set_diag(s.ws, s.geo)
            out.append(SimpleNamespace(u=u, v=v, N=s.N, h=h, P_u=s.pre.P_u, P_v=s.pre.P_v, S_u=s.geo.S_u,
                                       W2d=s.geo.W2d, err_c2=s.geo.err_c2, t=float(t), dt=s.dt, nsteps=s.nsteps,
                                       workspace=s.ws))
        else:
            out.append(_device_mms_errors(s.ws, s.T, s.N, h, s.pre, s.geo, s.kx, s.ky, s.phase))
    return out


def run_source_many(runs, cheb: str = "global"):
    """``solve(..., mode="source")`` for many runs at once: ``runs`` is a list of
    dicts with the arguments of ``solve`` (``q``, ``N``, ``T``, ``record_every``,
    ``slice_y``, ``receivers``, ``snap_times`` and the ``precompute`` keywords);
    returns per run the time series ``solve`` returns (src/solver.jl:872-958).
    The host loop stops at the union of the runs' record and snapshot steps;
    between stops the runs advance in one batch with their receivers recorded
    after every step (``wedg_batch_advance_receivers``). The series are bitwise
    those of separate runs."""
    from .solver import _blow_up, _source_record, _source_result, _source_setup
    setups = []
    for r in runs:
        r = dict(r)
        q, N, T = r.pop("q"), r.pop("N"), r.pop("T")
        dt_fac = r.pop("dt_fac", 10)
        nblocks = r.pop("nblocks", None)
        run_cheb = r.pop("cheb", cheb)
        r.pop("return_state", None)   # source mode returns the time series either way
        kw = dict(record_every=r.pop("record_every", 20), slice_y=r.pop("slice_y", 0.5),
                  receivers=r.pop("receivers", ((0.45, 0.5), (0.7, 0.5), (0.85, 0.5))),
                  snap_times=r.pop("snap_times", ()))
        mode = str(r.pop("mode", "source")).lstrip(":")
        if mode != "source":
            raise ValueError("run_source_many runs :source mode; use solve_many for the others")
        pre, geo = precompute(q, N, mode="source", **r)
        dt, nsteps = time_step(q, N, T, dt_fac)
        ws = Workspace(pre, N, nblocks=nblocks, cheb=run_cheb)
        set_zero_state(ws)
        S = _source_setup(ws, dt, nsteps, N, pre, geo, **kw)
        ws.dev.check(lib.wedg_set_receivers(ws.dev.h, S.nr, S.rj.ctypes.data_as(C.POINTER(C.c_int)),
                                            S.rk.ctypes.data_as(C.POINTER(C.c_int)), _ptr(S.rows)))
        setups.append(S)
    if not setups:
        return []
    n = len(setups)
    b = Batch([S.ws for S in setups])
    stops = sorted(set().union(*[S.stops for S in setups]))
    dts = np.ascontiguousarray([S.dt for S in setups], dtype=np.float64)
    for stop in stops:
        steps = [max(0, min(stop, S.nsteps) - S.done) for S in setups]
        if not any(steps):
            continue
        bufs = [np.zeros(max(k * S.nr, 1)) for k, S in zip(steps, setups)]
        ptrs = (C.POINTER(C.c_double) * n)(*[_ptr(x) for x in bufs])
        t0 = np.ascontiguousarray([S.t for S in setups], dtype=np.float64)
        tend = np.zeros(n)
        status = np.zeros(n, dtype=np.int32)
        ns = (C.c_long * n)(*steps)
        st = lib.wedg_batch_advance_receivers(b.h, _ptr(t0), _ptr(dts), ns, 1, _ptr(tend),
                                              status.ctypes.data_as(C.POINTER(C.c_int)), ptrs)
        if st != _lib.OK and not np.any(status != _lib.OK):   # the call ended before any instance's record
            raise WedgError(f"wedg_batch_advance_receivers failed with status {st}")
        for i, S in enumerate(setups):
            if status[i] == _lib.ERR_NONFINITE:
                raise _blow_up(S)
            S.ws.dev.check(int(status[i]))
        for k, S, buf, te in zip(steps, setups, bufs, tend):
            if k == 0:
                continue
            S.t = float(te)
            S.recv_all[S.done:S.done + k] = buf[:k * S.nr].reshape(k, S.nr)
            S.done += k
            _source_record(S, S.done)
    return [_source_result(S) for S in setups]


if __name__ == "__main__":
    # python -m edg_westervelt_amd.batch: calibrates the uneven workload's
    # amplitudes on the device and writes uneven_amplitudes.json
    import sys
    amp, d = calibrate_amplitudes(log=lambda m: print(m, file=sys.stderr, flush=True))
    out = dict(rule="amp such that d = 2 beta_max max|v_h| over the run <= 0.5 (target 0.49), max|v_h| at the "
                    "diagnostics points after every 16 steps; from the bound amplitudes, 3 rescalings amp <- "
                    "amp 0.49 / d, then confirmation runs",
               runs=[dict(seed=0, T=0.25, amp=[float(a) for a in amp], d=[round(float(x), 5) for x in d])])
    with open(AMPLITUDE_FILE if len(sys.argv) < 2 else sys.argv[1], "w") as f:
        json.dump(out, f, indent=1)
