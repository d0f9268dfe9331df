"""Host side of the batched mode (no GPU): the uneven workload generator and
the piecewise-constant coefficients it feeds to precompute."""
import math

import numpy as np
import pytest

from edg_westervelt_amd import batch as B
from edg_westervelt_amd.precompute import precompute


def test_uneven_instances_are_reproducible_and_in_range():
    a = B.uneven_instances(seed=0, count=64)
    b = B.uneven_instances(seed=0, count=64)
    assert len(a) == 64
    for x, y in zip(a, b):
        assert (x.N, x.q, x.mixed, x.flux, x.boundary, x.sigma, x.kappa) == (y.N, y.q, y.mixed, y.flux, y.boundary,
                                                                               y.sigma, y.kappa)
        assert np.array_equal(x.c, y.c) and np.array_equal(x.beta, y.beta)
    assert B.uneven_instances(seed=1, count=4)[0].c[0, 0] != a[0].c[0, 0]
    for x in a:
        assert x.N in (16, 24, 32, 48, 64) and 2 <= x.q <= 6 and x.flux in B.FLUXES
        assert ((0.8 <= x.c) & (x.c <= 1.2)).all() and ((1e-6 <= x.alpha) & (x.alpha <= 1e-3)).all()
        assert ((0.0 <= x.beta) & (x.beta <= 8.0)).all()
        assert 0.3 <= x.x0 <= 0.7 and 0.3 <= x.y0 <= 0.7 and 0.03 <= x.s <= 0.08
        assert -1 <= x.sigma <= 1 and x.sigma_dirichlet == abs(x.sigma) and 0 <= x.kappa <= 1
        assert set(x.boundary.values()) <= set(B.KINDS)
    # every (q_u, pairing) group of the workload occurs
    assert len({(x.q, x.mixed) for x in a}) == 10


def test_uneven_amplitudes():
    """The "bound" rule: 1 - 2 beta |v| >= 0.5 for |v| up to 4 c_max amp /
    (s sqrt(e)). The calibrated amplitudes of seed 0 (the default): the stored
    d = 2 beta_max max|v_h| of the calibration's confirming run lies in
    [0.48, 0.5], above the bound rule's amplitudes (whose margin of 4 the
    calibration spends); the rest of each instance is unchanged."""
    bound = B.uneven_instances(seed=0, count=64, amplitude="bound")
    for x in bound:
        assert x.amp_rule == "bound"
        vmax = 4 * x.c.max() * x.amp / (x.s * math.sqrt(math.e))
        assert 1 - 2 * x.beta.max() * vmax >= 0.5 - 1e-15 and x.amp <= 0.05
    cal = B.uneven_instances(seed=0, count=64)
    with open(B.AMPLITUDE_FILE) as f:
        import json
        tab = json.load(f)["runs"][0]
    assert (tab["seed"], tab["T"], len(tab["amp"])) == (0, 0.25, 64)
    for x, y, d in zip(cal, bound, tab["d"]):
        assert x.amp_rule == "calibrated" and 0.48 <= d <= 0.5
        assert y.amp < x.amp <= 0.05
        assert (x.N, x.q, x.s, x.x0) == (y.N, y.q, y.s, y.x0) and np.array_equal(x.beta, y.beta)
    assert B.uneven_instances(seed=1, count=2)[0].amp_rule == "bound"   # no calibration stored
    with pytest.raises(ValueError):
        B.uneven_instances(seed=1, count=2, amplitude="calibrated")
    with pytest.raises(ValueError):
        B.uneven_instances(seed=0, count=65, amplitude="calibrated")
    with pytest.raises(ValueError):
        B.uneven_instances(seed=0, count=2, amplitude="margin")


def test_piecewise_constant_blocks():
    vals = np.arange(16.0).reshape(4, 4)
    f = B.piecewise_constant(vals)
    assert f(0.1, 0.9) == vals[0, 3]
    assert f(1.0, 0.0) == vals[3, 0]         # the last block is closed on the right
    X, Y = np.meshgrid(np.linspace(0.01, 0.99, 7), np.linspace(0.01, 0.99, 5), indexing="ij")
    got = f(X, Y)
    assert got.shape == X.shape
    assert np.array_equal(got, vals[np.floor(4 * X).astype(int), np.floor(4 * Y).astype(int)])


@pytest.mark.parametrize("i", [0, 12])
def test_instance_tables(i):
    """Block coefficients are constant per element on a mesh-aligned grid: one
    class per block, jumps across block faces (one-sided faces)."""
    x = B.uneven_instances(seed=0, count=16)[i]
    pre, geo = precompute(x.q, x.N, mixed=x.mixed, flux=x.flux, sigma=x.sigma, boundary=x.boundary,
                          sigma_dirichlet=x.sigma_dirichlet, kappa=x.kappa,
                          coefficients=B.block_coefficients(x.c, x.alpha, x.beta), mode="initial_value",
                          initial_profile=B.gaussian_profile(x.x0, x.y0, x.s, x.amp))
    assert pre.one_sided
    nb = x.N // 4
    # Cw[iy, e, ix] = c^2 w_iy w_ix: element (j, k) lies in block (k // nb, j // nb)
    Cw = np.asarray(pre.Cw).reshape((pre.qu1, x.N, x.N, pre.qu1), order="F")
    w = pre.w
    for j in (0, nb - 1, nb, x.N - 1):
        for k in (0, 2 * nb, x.N - 1):
            c2 = x.c[k // nb, j // nb] ** 2
            assert np.allclose(Cw[:, j, k, :], c2 * np.outer(w, w), rtol=1e-14, atol=0)


def test_solve_many_argument_checks():
    """solve_many and run_source_many refuse what they do not batch before any
    device work (solve's own messages where solve has one)."""
    with pytest.raises(ValueError, match="source"):
        B.solve_many([dict(q=2, N=4, T=0.01, mode="source")])
    with pytest.raises(ValueError, match="return_state"):
        B.solve_many([dict(q=2, N=4, T=0.01, mode="initial_value")])
    with pytest.raises(ValueError, match="unknown mode"):
        B.solve_many([dict(q=2, N=4, T=0.01, mode="adjoint")])
    with pytest.raises(ValueError, match="source"):
        B.run_source_many([dict(q=2, N=4, T=0.01, mode="mms")])
    assert B.solve_many([]) == [] and B.run_source_many([]) == []


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus N > 1 exits with a message before any device work."""
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr and not r.stdout
