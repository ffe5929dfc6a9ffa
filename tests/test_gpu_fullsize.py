"""The benchmark workloads at their real sizes: 8^4, 24^4, 32^3x64, 48^3x96.

8^4 runs end to end against the oracle: 200 Philox Metropolis sweeps
from a cold start compared link for link, then the flow to t = 1.0 row by row.
Where the oracle is too slow for a whole run (24^4 flows for 399 steps), the
first rows are compared exactly and the rest through size-independent
properties: the plaquette rises, links stay in SU(3), t^2<E> is smooth, the
mixed path stays within its stated error of the FP64 path.
"""
import numpy as np
import pytest

from latticeyangmills_amd import LatticeContext
from tests import oracle_lib as orc

pytestmark = pytest.mark.gpu


def rel(a, b):
    return abs(a - b) / abs(b)


def unitarity(links):
    M = orc.from18(links[:, ::97])
    return float(np.abs(M @ np.conj(np.swapaxes(M, -1, -2)) - np.eye(3)).max())


def test_8x8x8x8_pipeline_against_oracle():
    """8^4, beta 6.0, cold, 200 sweeps (10 hits, eps 0.24, seed 1234), flow eps 0.01 to t 1.0 every 10."""
    dims = (8, 8, 8, 8)
    cold = orc.cold_start(dims)
    want, acc_want = orc.metropolis(cold, dims, 6.0, 0.24, 10, 200, 0, 1, 1234)
    with LatticeContext(dims) as ctx:
        acc = ctx.metropolis(6.0, 0.24, 10, 200, 0, "exph", 1234)
        got = ctx.download()
    assert acc == acc_want
    assert np.abs(got - want).max() < 1e-10
    plaq = orc.observables(want, dims)[0]
    assert 0.55 < plaq < 0.65  # beta = 6.0 equilibrium is near 0.594
    series_o, _ = orc.flow(want, dims, 0.01, 1.0, 10, 0)
    with LatticeContext(dims) as ctx:
        ctx.upload(want)
        series_g = ctx.flow(0.01, 1.0, 10)
    assert series_g.shape == series_o.shape == (100, 4)
    scale = orc.observables(want, dims)[3]
    for r in range(0, 100, 10):
        assert rel(series_g[r, 1], series_o[r, 1]) < 1e-12, r
        assert rel(series_g[r, 3], series_o[r, 3]) < 1e-12, r
        assert abs(series_g[r, 2] - series_o[r, 2]) < 1e-12 * scale, r


def test_24x24x24x24_first_rows_against_oracle():
    """24^4, beta 6.0, 500 sweeps from seed 7: the GPU config flowed by both paths."""
    dims = (24, 24, 24, 24)
    with LatticeContext(dims) as ctx:
        ctx.metropolis(6.0, 0.24, 10, 500, 0, "exph", 7)
        links = ctx.download()
        series_g = ctx.flow(0.01, 0.03, 1)  # rows t = 0, 0.01, 0.02
    series_o, _ = orc.flow(links, dims, 0.01, 0.03, 1, 0)
    scale = orc.observables(links, dims)[3]
    # statistical pin: the beta = 6.0 Wilson plaquette is 0.5937 in equilibrium
    assert abs(series_o[0, 1] - 0.5937) < 0.002
    for r in range(3):
        assert rel(series_g[r, 1], series_o[r, 1]) < 1e-12, r
        assert rel(series_g[r, 3], series_o[r, 3]) < 1e-12, r
        assert abs(series_g[r, 2] - series_o[r, 2]) < 1e-12 * scale, r


def test_24x24x24x24_whole_flow_properties():
    dims = (24, 24, 24, 24)
    V = orc.volume(dims)
    with LatticeContext(dims) as ctx:
        ctx.metropolis(6.0, 0.24, 10, 500, 0, "exph", 7)
        series = ctx.flow(0.01, 4.0, 10)
        links = ctx.download()
    m = series[::10]
    assert series.shape == (400, 4) and abs(series[-1, 0] - 3.99) < 1e-12
    assert np.all(np.diff(m[:, 1]) > 0)  # plaquette rises along the flow
    t2e = m[:, 0] ** 2 * (-m[:, 3] / (64 * V))
    assert np.all(np.isfinite(t2e)) and np.all(np.diff(t2e[1:]) > 0)
    assert unitarity(links) < 1e-12


def test_32x32x32x64_mixed_within_stated_error():
    """32^3x64, beta 6.2 (Haar start, 1000 sweeps, seed 62, as the bench): mixed
    flow vs FP64 flow of the same config to t = 5.98.  Bounds: <E> and the
    plaquette 1e-6 relative at every measured row, t0/a^2 1e-6 relative,
    Q 1e-4 absolute."""
    from latticeyangmills_amd import scale
    dims = (32, 32, 32, 64)
    V = orc.volume(dims)
    with LatticeContext(dims) as ctx:
        ctx.set_hot(62)
        ctx.metropolis(6.2, 0.24, 10, 1000, 0, "exph", 62)
        links = ctx.download()
        s64 = ctx.flow(0.02, 6.0, 10)
    with LatticeContext(dims, "mixed") as ctx:
        ctx.upload(links)
        smx = ctx.flow(0.02, 6.0, 10)
        u = unitarity(ctx.download())
    rows = range(0, 300, 10)
    assert max(rel(smx[r, 3], s64[r, 3]) for r in rows) < 1e-6
    assert max(abs(smx[r, 2] - s64[r, 2]) for r in rows) < 1e-4
    assert max(rel(smx[r, 1], s64[r, 1]) for r in rows) < 1e-6
    t0m, t064 = scale.t0(smx, V), scale.t0(s64, V)
    assert t064 is not None and 4.0 < t064 < 6.0  # reached inside the flow (5.09 on this config)
    assert rel(t0m, t064) < 1e-6
    w0m, w064 = scale.w0(smx, V), scale.w0(s64, V)
    assert (w0m is None) == (w064 is None) and (w064 is None or rel(w0m, w064) < 1e-6)
    assert u < 1e-6  # FP32 storage of reunitarised links


def test_48x48x48x96_production_volume():
    """48^3x96, beta 6.4: the largest single-GPU volume, mixed flow properties."""
    dims = (48, 48, 48, 96)
    with LatticeContext(dims, "fp64") as ctx:
        ctx.metropolis(6.4, 0.24, 10, 5, 0, "exph", 64)
        links = ctx.download()
    with LatticeContext(dims, "mixed") as ctx:
        ctx.upload(links)
        series = ctx.flow(0.02, 0.5, 5)
        u = unitarity(ctx.download())
    m = series[::5]
    assert np.all(np.diff(m[:, 1]) > 0)
    assert u < 1e-6
