"""wedg_rhs_noalloc, the right-hand side that allocates nothing (its scratch is
a static device array), against wedg_rhs_device on a workspace: the same
kernels on the same inputs, so the outputs must agree bit for bit; and both
against the oracle to 1e-12 relative. Two calls in a row, the second on a
smaller mesh, so the second call runs on scratch the first one left behind.
The failure record reaches the caller's device wedg_error_info."""
import ctypes as C
import sys

import numpy as np
import pytest

import edg_westervelt_amd as w
from edg_westervelt_amd import _lib, precompute
from oracle import batched as ob
from tests.cases import MIXED, quadrant_coefficients, rand_state
from tests.conftest import rel

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


class _Hip:
    """Device buffers of the caller, from the HIP runtime libwedg.so uses
    (torch brings its own runtime, which cannot share the process), copied
    through pinned host memory."""

    def __init__(self):
        self.rt = C.CDLL("libamdhip64.so.7")
        self.bufs = []
        self.pin = C.c_void_p()
        self.pin_n = 1 << 22
        assert self.rt.hipHostMalloc(C.byref(self.pin), C.c_size_t(self.pin_n), 0) == 0

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.rt.hipMalloc(C.byref(p), C.c_size_t(max(nbytes, 8))) == 0
        self.bufs.append(p)
        return p

    def put(self, a):
        b = np.asfortranarray(a, dtype=np.float64).ravel(order="F").tobytes()
        assert len(b) <= self.pin_n
        p = self.alloc(len(b))
        C.memmove(self.pin, b, len(b))
        assert self.rt.hipMemcpy(p, self.pin, C.c_size_t(len(b)), 1) == 0
        return p

    def get(self, p, nbytes):
        assert nbytes <= self.pin_n
        assert self.rt.hipDeviceSynchronize() == 0
        assert self.rt.hipMemcpy(self.pin, p, C.c_size_t(nbytes), 2) == 0
        return C.string_at(self.pin, nbytes)

    def get_f64(self, p, shape):
        n = int(np.prod(shape)) * 8
        return np.frombuffer(self.get(p, n), dtype=np.float64).reshape(shape, order="F")

    def close(self):
        for p in self.bufs:
            self.rt.hipFree(p)
        self.rt.hipHostFree(self.pin)
        self.bufs = []


@pytest.fixture
def hip():
    h = _Hip()
    yield h
    h.close()


def _run_both(hip, pre, N, u, v, t, cheb):
    prob = _lib.Problem(pre, N)
    ws = _lib.DeviceWorkspace(prob, cheb)
    pol = _lib.CHEB_GLOBAL if cheb == "global" else _lib.CHEB_ELEMENT
    du, dv = hip.put(u), hip.put(v)
    wu, wv, nu, nv = (hip.alloc(u.size * 8), hip.alloc(v.size * 8), hip.alloc(u.size * 8), hip.alloc(v.size * 8))
    info = hip.alloc(C.sizeof(_lib.WedgErrorInfo))
    st = _lib.lib.wedg_rhs_device(ws.h, du, dv, float(t), wu, wv)
    assert st == _lib.OK
    st_ws = _lib.lib.wedg_sync(ws.h)
    ws_info = ws.last_error()
    st = _lib.lib.wedg_rhs_noalloc(prob.h, pol, du, dv, float(t), nu, nv, info, None)
    assert st == _lib.OK
    rec = _lib.WedgErrorInfo.from_buffer_copy(hip.get(info, C.sizeof(_lib.WedgErrorInfo)))
    return (st_ws, ws_info, hip.get_f64(wu, u.shape), hip.get_f64(wv, v.shape), rec,
            hip.get_f64(nu, u.shape), hip.get_f64(nv, v.shape))


@pytest.mark.parametrize("cheb", ["global", "element"])
def test_noalloc_matches_workspace_entry_point(hip, cheb):
    cases = [
        ("q=5 N=16 quadrants one-sided", 5, 16, False, "source",
         quadrant_coefficients(0.4, 0.6, beta=(2.0, 2.5, 3.0, 1.5))),
        ("q=3 N=6 mixed", 3, 6, True, "mms", None),   # the second call is smaller
    ]
    for k, (name, q, N, mixed, mode, coef) in enumerate(cases):
        say(f"case noalloc {cheb} call {k + 1}: {name}")
        kw = {} if coef is None else dict(coefficients=coef)
        pre, _ = precompute(q, N, mixed=mixed, mode=mode, boundary=MIXED, sigma=0.3, sigma_dirichlet=0.6,
                            kappa=0.7, **kw)
        u, v = rand_state(np.random.default_rng(40 + k), N, pre, su=1e-2, sv=2e-3)
        t = 0.021
        st_ws, ws_info, wu, wv, info, nu, nv = _run_both(hip, pre, N, u, v, t, cheb)
        assert st_ws == _lib.OK and info.status == _lib.OK
        assert np.array_equal(wu, nu) and np.array_equal(wv, nv)
        assert info.cheb_steps == ws_info.cheb_steps > 0
        a = ob.rhs(ob.OraclePre(pre, N), u, v, t, cheb_mode=0 if cheb == "global" else 1)
        assert rel(a[0], nu) < 1e-12 and rel(a[1], nv) < 1e-12


def test_noalloc_reports_the_failure_record(hip):
    """A state with 1 - 2 beta v_h < 0 on one element: the caller's record holds
    the same failure (status and element) as the workspace entry point's; the
    call after it starts from a clean record."""
    q, N = 3, 5
    pre, _ = precompute(q, N, mixed=True, mode="initial_value",
                        coefficients=w.constant_coefficients(c=1.0, alpha=1e-3, beta=2.0))
    u, v = rand_state(np.random.default_rng(7), N, pre, su=1e-2, sv=1e-4)
    v[2, 3, :, :] = 0.0
    v[2, 3, 0, 0] = 2.0   # v_h = 1 on element (3, 4): 1 - 2 beta v_h = -3
    say("case noalloc failure record")
    st_ws, ws_info, _, _, info, _, _ = _run_both(hip, pre, N, u, v, 0.0, "global")
    assert st_ws == _lib.ERR_MASS and info.status == _lib.ERR_MASS
    assert info.element == ws_info.element == 2 + N * 3 + 1
    say("case noalloc after a failure")
    v[2, 3, 0, 0] = 0.0
    st_ws, _, wu, wv, info, nu, nv = _run_both(hip, pre, N, u, v, 0.0, "global")
    assert st_ws == _lib.OK and info.status == _lib.OK
    assert np.array_equal(wu, nu) and np.array_equal(wv, nv)


def test_noalloc_rejects_arguments(hip):
    pre, _ = precompute(2, 3, mixed=False, mode="mms")
    prob = _lib.Problem(pre, 3)
    p = hip.alloc(4096)
    assert _lib.lib.wedg_rhs_noalloc(prob.h, 7, p, p, 0.0, p, p, None, None) == _lib.ERR_ARGUMENT
    assert _lib.lib.wedg_rhs_noalloc(prob.h, 0, p, None, 0.0, p, p, None, None) == _lib.ERR_ARGUMENT
