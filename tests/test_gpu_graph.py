"""RK4 chunks as HIP graphs (wedg_graph_enable, the loop of solve,
src/solver.jl:213-216): the replayed graph gives bit for bit the states, end
time and failure record of the plain stage launches. Every case replays the
graph several times with plain work and synchronizations of the same
workspace in between, and checks every replay."""
import ctypes as C
import sys

import numpy as np
import pytest

import edg_westervelt_amd as w
from edg_westervelt_amd import _lib, precompute
from edg_westervelt_amd._lib import lib
from tests.cases import MIXED, rand_state

pytestmark = pytest.mark.gpu
P = C.POINTER(C.c_double)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _run(ws, u, v, t0, dt, n, chunk):
    assert lib.wedg_graph_enable(ws.dev.h, chunk) == _lib.OK
    uu = np.array(u, order="F", copy=True)
    vv = np.array(v, order="F", copy=True)
    ws.dev.check(lib.wedg_set_state(ws.dev.h, uu.ctypes.data_as(P), vv.ctypes.data_as(P)))
    tend = C.c_double(0.0)
    st = lib.wedg_advance(ws.dev.h, t0, dt, n, 0, C.byref(tend))
    info = ws.dev.last_error()
    ws.dev.check(lib.wedg_get_state(ws.dev.h, uu.ctypes.data_as(P), vv.ctypes.data_as(P)))
    return st, (info.status, info.side, info.key, info.element), uu, vv, tend.value


@pytest.mark.parametrize("q,N,mixed,mode,chunk,n", [(3, 8, True, "mms", 4, 30), (5, 16, False, "source", 7, 30),
                                                    (2, 5, False, "initial_value", 1, 9),
                                                    (4, 12, True, "source", 16, 16)])
def test_graph_chunks_equal_plain_launches(q, N, mixed, mode, chunk, n):
    say(f"case graph q={q} N={N} mixed={mixed} {mode} chunk={chunk} steps={n}")
    pre, _ = precompute(q, N, mixed=mixed, mode=mode, boundary=MIXED, sigma=0.3, sigma_dirichlet=0.6, kappa=0.7)
    u, v = rand_state(np.random.default_rng(60 + q), N, pre, su=1e-2, sv=2e-3)
    ws = w.Workspace(pre, N)
    ref = _run(ws, u, v, 0.01, 1e-4, n, 0)
    assert ref[0] == _lib.OK
    for rep in range(3):
        got = _run(ws, u, v, 0.01, 1e-4, n, chunk)
        assert got[0] == _lib.OK and got[1] == ref[1] and got[4] == ref[4]
        assert np.array_equal(got[2], ref[2]) and np.array_equal(got[3], ref[3]), rep
        w.edg_rhs_(ws, u, v, 0.3, N, 1.0 / N, pre)   # plain work of the same workspace in between
    # the graph of this chunk length is kept: a longer run replays it more often
    got = _run(ws, u, v, 0.01, 1e-4, 2 * n + 1, chunk)
    ref2 = _run(ws, u, v, 0.01, 1e-4, 2 * n + 1, 0)
    assert np.array_equal(got[2], ref2[2]) and np.array_equal(got[3], ref2[3]) and got[4] == ref2[4]


def test_graph_failure_inside_a_chunk():
    """A boundary weight that fails at stage 4 of step 5 (inside the second
    graph chunk of 3 steps): the same status, side and key as the plain run."""
    from edg_westervelt_amd.batch import gaussian_profile
    from tests.test_gpu_source import _bottom_beta
    q, N = 3, 6
    pre, _ = precompute(q, N, mixed=True, mode="initial_value", coefficients=_bottom_beta(1.0),
                        boundary=dict(bot="absorbing_2beta", top="wall", lft="wall", rgt="wall"),
                        initial_profile=gaussian_profile(0.5, 0.0, 0.12, 0.05))
    dt = 1.0 / (N * 10 * q * q)
    u0 = np.zeros((N, N, pre.qu1, pre.qu1), order="F")
    v0 = np.zeros((N, N, pre.qv1, pre.qv1), order="F")
    ws = w.Workspace(pre, N)
    say("case graph failure inside a chunk")
    ref = _run(ws, u0, v0, 0.0, dt, 20, 0)
    assert ref[0] == _lib.ERR_BOUNDARY_TRACE
    for rep in range(2):
        got = _run(ws, u0, v0, 0.0, dt, 20, 3)
        assert got[0] == ref[0] and got[1] == ref[1]
