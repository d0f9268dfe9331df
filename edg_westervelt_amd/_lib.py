"""ctypes binding of libwedg.so, the gfx950 HIP library behind include/wedg.h.

The library is required: importing this module raises if ``libwedg.so`` has
not been built (``python -c "import __graft_entry__ as g; g.build()"``) or
cannot be loaded; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import warnings

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("WEDG_LIBRARY") or os.path.join(_HERE, "libwedg.so")

# Hardware queues of the process. HIP maps streams onto GPU_MAX_HW_QUEUES
# queues (4 by default), and streams that share a queue run one after the other
# (an event wait on one holds up the other). The batched mode runs a stream per
# kernel group and a gradient three streams that wait on each other, so the
# package raises the setting to 16 when it is lower (WEDG_KEEP_HW_QUEUES=1 keeps
# it as it is). It takes effect when the HIP runtime is loaded, i.e. if nothing
# in the process has loaded it yet.
MIN_HW_QUEUES = 16
_hwq = os.environ.get("GPU_MAX_HW_QUEUES", "")
if not os.environ.get("WEDG_KEEP_HW_QUEUES") and not (_hwq.isdigit() and int(_hwq) >= MIN_HW_QUEUES):
    os.environ["GPU_MAX_HW_QUEUES"] = str(MIN_HW_QUEUES)

KINDS = {"absorbing_2beta": 0, "absorbing": 1, "dirichlet": 2, "wall": 3}
SIDES = ("bot", "top", "lft", "rgt")

(OK, ERR_BOUNDARY_TRACE, ERR_BOUNDARY_DATA, ERR_MASS, ERR_DIMENSION, ERR_NONFINITE, ERR_HIP, ERR_ARGUMENT,
 ERR_SCRATCH) = range(9)
CHEB_GLOBAL, CHEB_ELEMENT = 0, 1

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class WedgSide(C.Structure):
    _fields_ = [("kind", C.c_int), ("gamma_abc", C.c_double), ("eta_abc", C.c_double),
                ("sigma", C.c_double), ("kb", C.c_double), ("kappa", C.c_double)]


_PTRS = ("w", "P_u", "S_u", "Ebl_u", "Ebr_u", "E1bl_u", "E1br_u", "P_h_v", "W2d_h",
         "Cw", "Aw", "Ainv_u", "beta_hq",
         "hc2", "haf", "hcf", "hbf", "vc2", "vaf", "vcf", "vbf",
         "hc2m", "hafm", "vc2m", "vafm", "hJC", "vJC",
         "bcB_sp", "bcB_spy", "bcT_sp", "bcT_spy", "bcL_sp", "bcL_spx", "bcR_sp", "bcR_spx",
         "PiSP", "PiBSP2", "PiCdiv", "PiDdiv", "src_projE")


class WedgPre(C.Structure):
    _fields_ = ([("N", C.c_int), ("qu1", C.c_int), ("qv1", C.c_int), ("q_high", C.c_int),
                 ("omega", C.c_double), ("theta", C.c_double), ("tau", C.c_double),
                 ("lambda_u", C.c_double), ("lambda_v", C.c_double),
                 ("mms_on", C.c_int), ("src_on", C.c_int), ("one_sided", C.c_int), ("mass_solver", C.c_int),
                 ("mass_tol", C.c_double), ("src_omega", C.c_double), ("src_Ts", C.c_double),
                 ("bc", WedgSide * 4)]
                + [(n, _dp) for n in _PTRS])


class WedgAdjDesc(C.Structure):
    _fields_ = ([("N", C.c_int), ("qu1", C.c_int), ("qv1", C.c_int),
                 ("theta", C.c_double), ("tau", C.c_double), ("lambda_u", C.c_double), ("lambda_v", C.c_double),
                 ("bc", WedgSide * 4), ("mass_tol", C.c_double), ("src_omega", C.c_double), ("src_Ts", C.c_double)]
                + [(n, _dp) for n in ("w", "P_u", "S_u", "Ebl_u", "Ebr_u", "E1bl_u", "E1br_u", "P_h_v", "W2d_h",
                                      "src_projE", "alpha")])


class WedgAdjStats(C.Structure):
    _fields_ = [("forward_steps", C.c_long), ("taped_steps", C.c_long), ("adjoint_steps", C.c_long),
                ("snapshots_taken", C.c_long), ("snapshots_used", C.c_int), ("tapes", C.c_int),
                ("enqueue_ms", C.c_double), ("total_ms", C.c_double)]


class WedgErrorInfo(C.Structure):
    _fields_ = [("status", C.c_int), ("side", C.c_int), ("key", C.c_long), ("element", C.c_long),
                ("step", C.c_long), ("cheb_steps", C.c_int), ("cg_unconverged", C.c_int)]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it first (__graft_entry__.build()); "
                          "there is no CPU fallback")
    lib = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    sig = {
        "wedg_problem_create": [C.POINTER(WedgPre), C.POINTER(vp)],
        "wedg_problem_destroy": [vp],
        "wedg_workspace_create": [vp, C.c_int, C.POINTER(vp)],
        "wedg_workspace_destroy": [vp],
        "wedg_rhs": [vp, _dp, _dp, C.c_double, C.c_int, C.c_double, _dp, _dp],
        "wedg_rhs_device": [vp, vp, vp, C.c_double, vp, vp],
        "wedg_rhs_noalloc": [vp, C.c_int, vp, vp, C.c_double, vp, vp, vp, vp],
        "wedg_rk4_step": [vp, _dp, _dp, C.c_double, C.c_double, C.c_int, C.c_double, C.POINTER(C.c_int)],
        "wedg_set_state": [vp, _dp, _dp],
        "wedg_get_state": [vp, _dp, _dp],
        "wedg_state_device": [vp, C.POINTER(vp), C.POINTER(vp)],
        "wedg_advance": [vp, C.c_double, C.c_double, C.c_long, C.c_int, C.POINTER(C.c_double)],
        "wedg_advance_async": [vp, C.c_double, C.c_double, C.c_long, C.c_int, C.POINTER(C.c_double)],
        "wedg_sync": [vp],
        "wedg_graph_enable": [vp, C.c_long],
        "wedg_workspace_set_lanes_per_element": [vp, C.c_int],
        "wedg_problem_info": [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)],
        "wedg_workspace_lanes_per_element": [vp, C.POINTER(C.c_int)],
        "wedg_adj_set_lanes_per_element": [vp, C.c_int],
        "wedg_adj_lanes_per_element": [vp, C.POINTER(C.c_int)],
        "wedg_last_error": [vp, C.POINTER(WedgErrorInfo)],
        "wedg_timing_enable": [vp, C.c_int],
        "wedg_timing_read": [vp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_long)],
        "wedg_device_info": [C.c_char_p, C.c_size_t, C.POINTER(C.c_int), C.POINTER(C.c_double)],
        "wedg_cheb_stats": [vp, C.POINTER(C.c_double), C.POINTER(C.c_long)],
        "wedg_probe_copy": [C.c_size_t, C.c_int, C.POINTER(C.c_double)],
        "wedg_problem_set_diag": [vp, _dp, _dp, _dp],
        "wedg_advance_receivers": [vp, C.c_double, C.c_double, C.c_long, C.c_int, _ip, _ip, _dp, C.c_int, _dp,
                                   C.POINTER(C.c_double)],
        "wedg_diagnostics": [vp, _dp, _dp, _dp, C.c_int, _dp, _dp],
        "wedg_snapshot": [vp, _dp],
        "wedg_mms_errors": [vp, C.c_double, C.c_double, C.c_double, C.c_double, _dp, _dp, _dp, _dp, _dp],
        "wedg_difference_norms": [vp, _dp, _dp, _dp, _dp, _dp],
        "wedg_batch_create": [C.POINTER(vp), C.c_int, C.POINTER(vp)],
        "wedg_batch_destroy": [vp],
        "wedg_batch_advance": [vp, _dp, _dp, C.POINTER(C.c_long), C.c_int, _dp, _ip],
        "wedg_batch_time": [vp, C.POINTER(C.c_double)],
        "wedg_batch_configure": [vp, C.c_int, C.c_long, C.c_int],
        "wedg_set_receivers": [vp, C.c_int, _ip, _ip, _dp],
        "wedg_batch_advance_receivers": [vp, _dp, _dp, C.POINTER(C.c_long), C.c_int, _dp, _ip, C.POINTER(_dp)],
        "wedg_batch_config": [vp, C.POINTER(C.c_int), C.POINTER(C.c_long), C.POINTER(C.c_int)],
        "wedg_adj_create": [C.POINTER(WedgAdjDesc), C.c_int, C.POINTER(vp)],
        "wedg_adj_destroy": [vp],
        "wedg_adj_set_params": [vp, _dp, _dp],
        "wedg_adj_set_target": [vp, C.c_long, _dp],
        "wedg_adj_forward": [vp, C.c_double, C.c_long, _dp, _dp],
        "wedg_adj_gradient": [vp, C.c_double, C.c_long, C.c_int, _dp, _dp, _dp],
        "wedg_adj_last_stats": [vp, C.POINTER(WedgAdjStats)],
        "wedg_adj_rhs_vjp": [vp, _dp, _dp, C.c_double, _dp, _dp, _dp, _dp, _dp, _dp],
        "wedg_adj_rhs_tl": [vp, _dp, _dp, C.c_double, _dp, _dp, _dp, _dp, _dp, _dp],
        "wedg_adj_batch_create": [C.POINTER(vp), C.c_int, C.POINTER(vp)],
        "wedg_adj_batch_destroy": [vp],
        "wedg_adj_batch_forward": [vp, _dp, C.POINTER(C.c_long), _dp, _ip],
        "wedg_adj_batch_gradient": [vp, _dp, C.POINTER(C.c_long), C.c_int, _dp, C.POINTER(_dp), C.POINTER(_dp),
                                    _ip],
        "wedg_adj_batch_time": [vp, C.POINTER(C.c_double), C.POINTER(C.c_double)],
    }
    for name, args in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = C.c_int
    lib.wedg_version.restype = C.c_char_p
    lib.wedg_version.argtypes = []
    return lib


lib = _load()


class DimensionMismatch(ValueError):
    """Sizes of the state, workspace and tables do not match (src/rhs.jl:500-503)."""


class WedgError(RuntimeError):
    """An error of the reference's scheme (its messages), raised from a device run."""


def _f64(a) -> np.ndarray:
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


def error_message(pre, info: WedgErrorInfo) -> str:
    """The reference's message for a failure (src/rhs.jl:112-116, 916; src/solver.jl:940)."""
    st = info.status
    if st in (ERR_BOUNDARY_TRACE, ERR_BOUNDARY_DATA):
        bs = pre.bc[SIDES[info.side]]
        k = "2" if bs["kb"] == 2.0 else ""
        if st == ERR_BOUNDARY_TRACE:
            return f"boundary weight 1 - {k}βv_h ≤ 0 at a face point ({bs['kind']} side)"
        return f"boundary weight 1 - {k}βg ≤ 0 for the boundary data g ({bs['kind']} side)"
    if st == ERR_MASS:
        if getattr(pre, "mass_solver", "chebyshev") == "cg":   # (j, k), src/rhs.jl:845
            N = pre.N
            j, k = (info.element - 1) % N + 1, (info.element - 1) // N + 1
            return f"mass matrix not positive definite on element ({j}, {k}): 1 - 2βv_h ≤ 0"
        return f"mass matrix not positive definite on element {info.element}: 1 - 2βv_h ≤ 0"
    if st == ERR_NONFINITE:
        return f"non-finite state after step {info.step}"
    if st == ERR_HIP:
        return "HIP runtime error"
    if st == ERR_ARGUMENT:
        return "invalid argument"
    return f"wedg status {st}"


class Problem:
    """Device copy of one problem's tables (``wedg_problem_create``)."""

    def __init__(self, pre, N: int):
        if pre.mass_solver not in ("chebyshev", "cg"):
            raise ValueError(f"unknown mass_solver {pre.mass_solver} (expected :cg or :chebyshev)")
        s = WedgPre()
        s.N, s.qu1, s.qv1, s.q_high = N, pre.qu1, pre.qv1, pre.q_high
        s.omega, s.theta, s.tau = pre.omega, pre.theta, pre.tau
        s.lambda_u, s.lambda_v = pre.lambda_u, pre.lambda_v
        s.mms_on, s.src_on, s.one_sided = int(pre.mms_on), int(pre.src_on), int(pre.one_sided)
        s.mass_solver = 1 if pre.mass_solver == "cg" else 0
        s.mass_tol = pre.mass_tol
        s.src_omega, s.src_Ts = pre.source.omega, pre.source.Ts
        for i, side in enumerate(SIDES):
            b = pre.bc[side]
            s.bc[i].kind = KINDS[b["kind"]]
            s.bc[i].gamma_abc, s.bc[i].eta_abc, s.bc[i].sigma = b["gamma_abc"], b["eta_abc"], b["sigma"]
            s.bc[i].kb, s.bc[i].kappa = b["kb"], b["kappa"]
        keep = []
        for n in _PTRS:
            a = _f64(getattr(pre, n))
            keep.append(a)
            setattr(s, n, a.ctypes.data_as(_dp))
        h = C.c_void_p()
        st = lib.wedg_problem_create(C.byref(s), C.byref(h))
        if st != OK:
            raise WedgError(f"wedg_problem_create failed with status {st}")
        self.h = h
        self.N = N
        self.pre = pre

    def info(self) -> dict:
        """The kernel layout of the upload: mode (0 general, 1 separable, 2
        uniform, 3 conjugate gradients), coefficient classes, two-lane capable."""
        m, c, t = C.c_int(0), C.c_int(0), C.c_int(0)
        lib.wedg_problem_info(self.h, C.byref(m), C.byref(c), C.byref(t))
        return dict(mode=m.value, classes=c.value, two_lane=bool(t.value))

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value:
            lib.wedg_problem_destroy(h)
            self.h = None


class DeviceWorkspace:
    """Device-resident work arrays and state (``wedg_workspace_create``)."""

    def __init__(self, problem: Problem, cheb: str = "global", lanes: int = 1):
        self.problem = problem
        pol = {"global": CHEB_GLOBAL, "element": CHEB_ELEMENT}[cheb]
        h = C.c_void_p()
        st = lib.wedg_workspace_create(problem.h, pol, C.byref(h))
        if st != OK:
            raise WedgError(f"wedg_workspace_create failed with status {st}")
        self.h = h
        if lanes != 1:
            self.set_lanes(lanes)

    def set_lanes(self, lanes: int):
        """Lanes per element of the stage kernels (1, or 2 where the problem allows it)."""
        st = lib.wedg_workspace_set_lanes_per_element(self.h, int(lanes))
        if st != OK:
            raise WedgError(f"wedg_workspace_set_lanes_per_element failed with status {st}")

    @property
    def lanes(self) -> int:
        """The lanes per element the workspace's calls use."""
        n = C.c_int(0)
        lib.wedg_workspace_lanes_per_element(self.h, C.byref(n))
        return n.value

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value:
            lib.wedg_workspace_destroy(h)
            self.h = None

    def last_error(self) -> WedgErrorInfo:
        info = WedgErrorInfo()
        lib.wedg_last_error(self.h, C.byref(info))
        return info

    def check(self, st: int):
        if st == OK:
            info = self.last_error()
            if info.cg_unconverged > 0:   # src/rhs.jl:855-858
                warnings.warn(f"mass solve: {info.cg_unconverged} element systems did not reach the tolerance "
                              f"{self.problem.pre.mass_tol}")
            return
        info = self.last_error()
        if info.status == OK:
            info.status = st
        raise WedgError(error_message(self.problem.pre, info))


def device_info():
    name = C.create_string_buffer(64)
    cu = C.c_int(0)
    bw = C.c_double(0.0)
    st = lib.wedg_device_info(name, 64, C.byref(cu), C.byref(bw))
    if st != OK:
        raise WedgError("no HIP device")
    return name.value.decode(), cu.value, bw.value
