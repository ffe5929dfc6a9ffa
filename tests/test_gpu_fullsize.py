"""The benchmark workloads at their real sizes: 8^4, 24^4, 32^3x64, 48^3x96.

8^4 runs end to end against the oracle: 200 Philox Metropolis sweeps from a
cold start with a configuration saved every 50 sweeps, each compared link for
link, then each flowed to t = 1.0 and compared row by row.  Where the oracle is
too slow for a whole run (24^4 flows for 399 steps), the first rows are
compared, and along the trajectory the GPU's links at t = 1, 2, 3, 3.98 are
checked against the oracle (observables and one step from them); the rest
through size-independent properties: the plaquette rises, links stay in SU(3),
t^2<E> rises.  32^3x64 and 48^3x96 state the mixed path's error against the
FP64 path over the whole flow.
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


CFG1_DIMS = (8, 8, 8, 8)


@pytest.fixture(scope="module")
def cfg1_saves():
    """8^4, beta 6.0, cold start, 200 sweeps (10 hits, eps 0.24, seed 1234) with
    a configuration saved every 50 sweeps, on the GPU and by the oracle."""
    dims = CFG1_DIMS
    gpu, ora = [], []
    links = orc.cold_start(dims)
    with LatticeContext(dims) as ctx:
        for k in range(4):
            acc = ctx.metropolis(6.0, 0.24, 10, 50, 50 * k, "exph", 1234)
            gpu.append((ctx.download(), acc))
            links, acc_o = orc.metropolis(links, dims, 6.0, 0.24, 10, 50, 50 * k, 1, 1234)
            ora.append((links, acc_o))
    return gpu, ora


def test_8x8x8x8_saved_configs_against_oracle(cfg1_saves):
    """Every saved configuration link for link, accepted hits equal."""
    gpu, ora = cfg1_saves
    for k in range(4):
        assert gpu[k][1] == ora[k][1], k
        assert np.abs(gpu[k][0] - ora[k][0]).max() < 1e-10, k
    plaq = orc.observables(ora[3][0], CFG1_DIMS)[0]
    assert 0.55 < plaq < 0.65  # beta = 6.0 equilibrium is near 0.594


@pytest.mark.parametrize("save", [0, 1, 2, 3])
def test_8x8x8x8_flow_of_saved_config_against_oracle(cfg1_saves, save):
    """Each saved configuration flowed with eps 0.01 to t = 1.0, measured every
    10 steps, row by row against the oracle."""
    dims = CFG1_DIMS
    links = cfg1_saves[1][save][0]
    series_o, _ = orc.flow(links, dims, 0.01, 1.0, 10, 0)
    with LatticeContext(dims) as ctx:
        ctx.upload(links)
        series_g = ctx.flow(0.01, 1.0, 10)
    assert series_g.shape == series_o.shape == (100, 4)
    scale = orc.observables(links, dims)[3]
    for r in range(0, 100, 10):
        assert rel(series_g[r, 1], series_o[r, 1]) < 1e-12, r
        assert rel(series_g[r, 3], series_o[r, 3]) < 1e-12, r
        assert abs(series_g[r, 2] - series_o[r, 2]) < 1e-12 * scale, r


CFG2_DIMS = (24, 24, 24, 24)


@pytest.fixture(scope="module")
def cfg2():
    """24^4, beta 6.0, 500 Philox sweeps from seed 7 (cold start) on the GPU,
    and the FP64 GPU flow of that config: eps 0.01 to t = 3.99, every 10."""
    with LatticeContext(CFG2_DIMS) as ctx:
        ctx.metropolis(6.0, 0.24, 10, 500, 0, "exph", 7)
        links = ctx.download()
        series = ctx.flow(0.01, 4.0, 10)
    return links, series


def test_24x24x24x24_first_rows_against_oracle(cfg2):
    """Rows t = 0, 0.01, 0.02 of the flow against the oracle's."""
    links, _ = cfg2
    dims = CFG2_DIMS
    with LatticeContext(dims) as ctx:
        ctx.upload(links)
        series_g = ctx.flow(0.01, 0.03, 1)
    series_o, _ = orc.flow(links, dims, 0.01, 0.03, 1, 0)
    scale = orc.observables(links, dims)[3]
    # statistical pin: the beta = 6.0 Wilson plaquette is 0.5937 in equilibrium
    assert abs(series_o[0, 1] - 0.5937) < 0.002
    for r in range(3):
        assert rel(series_g[r, 1], series_o[r, 1]) < 1e-12, r
        assert rel(series_g[r, 3], series_o[r, 3]) < 1e-12, r
        assert abs(series_g[r, 2] - series_o[r, 2]) < 1e-12 * scale, r


@pytest.mark.parametrize("nsteps", [100, 200, 300, 398])
def test_24x24x24x24_trajectory_checkpoints_against_oracle(cfg2, nsteps):
    """Along the whole 399-step trajectory: the GPU's FP64 links after nsteps
    steps (t = 1.00, 2.00, 3.00, 3.98).  Their observables against the
    oracle's on the same links, and one more GPU step against one oracle step
    from them, at 1e-12; where the row is measured, the flow series holds the
    same observables bit for bit."""
    links, series = cfg2
    dims = CFG2_DIMS
    with LatticeContext(dims) as ctx:
        ctx.upload(links)
        ctx.flow_steps(0.01, nsteps)
        at = ctx.download()
        got = ctx.observables()
        ctx.flow_steps(0.01, 1)
        stepped = ctx.download()
    want = orc.observables(at, dims)
    assert rel(got[0], want[0]) < 1e-12
    assert rel(got[1], want[1]) < 1e-12
    assert abs(got[2] - want[2]) < 1e-12 * want[3]
    if nsteps % 10 == 0:
        assert series[nsteps, 1] == got[0] and series[nsteps, 3] == got[1] and series[nsteps, 2] == got[2]
    assert np.abs(stepped - orc.flow_step(at, dims, 0.01)).max() < 1e-12


def test_24x24x24x24_whole_flow_properties(cfg2):
    links0, series = cfg2
    dims = CFG2_DIMS
    V = orc.volume(dims)
    with LatticeContext(dims) as ctx:
        ctx.upload(links0)
        ctx.flow_steps(0.01, 399)
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
    assert abs(s64[0, 1] - 0.6136) < 5e-4  # the published equilibrium plaquette at beta 6.2
    assert max(rel(smx[r, 3], s64[r, 3]) for r in rows) < 1e-6
    assert max(abs(smx[r, 2] - s64[r, 2]) for r in rows) < 1e-4
    assert max(rel(smx[r, 1], s64[r, 1]) for r in rows) < 1e-6
    t0m, t064 = scale.t0(smx, V), scale.t0(s64, V)
    assert t064 is not None and 4.0 < t064 < 6.0  # reached inside the flow (5.09 on this config)
    assert rel(t0m, t064) < 1e-6
    w0m, w064 = scale.w0(smx, V), scale.w0(s64, V)
    assert (w0m is None) == (w064 is None) and (w064 is None or rel(w0m, w064) < 1e-6)
    assert u < 1e-6  # FP32 storage of reunitarised links


def test_48x48x48x96_mixed_within_stated_error():
    """48^3x96, beta 6.4: 200 Philox sweeps (cold start, seed 64), then the
    mixed and the FP64 flow of that config, eps 0.02 to t = 7.98, measured
    every 5 steps, the mixed flow with the links in three-row and in two-row
    form.  Bounds at every measured row: <E> and the plaquette 1e-6
    relative, |dQ| <= 1e-4; t0/a^2 and w0/a 1e-6 relative where the flow
    reaches them; links of the mixed path within 1e-6 of SU(3)."""
    from latticeyangmills_amd import scale
    dims = (48, 48, 48, 96)
    V = orc.volume(dims)
    with LatticeContext(dims, "fp64") as ctx:
        ctx.metropolis(6.4, 0.24, 10, 200, 0, "exph", 64)
        links = ctx.download()
        s64 = ctx.flow(0.02, 8.0, 5)
    results = {}
    with LatticeContext(dims, "mixed") as ctx:
        for form in ("three-row", "two-row"):
            if form == "two-row":  # the links as their first two rows while the flow runs
                ctx.set_schedule("column", 1024, True, two_row=True)
            ctx.upload(links)
            results[form] = (ctx.flow(0.02, 8.0, 5), ctx.check_links())
    del links
    assert s64.shape == (400, 4)
    rows = range(0, 400, 5)
    for form, (smx, (dev, bad)) in results.items():
        assert smx.shape == (400, 4)
        e_err = max(rel(smx[r, 3], s64[r, 3]) for r in rows)
        q_err = max(abs(smx[r, 2] - s64[r, 2]) for r in rows)
        p_err = max(rel(smx[r, 1], s64[r, 1]) for r in rows)
        print(f"48^3x96 mixed ({form}) vs FP64: max rel dE {e_err:.2e}, max |dQ| {q_err:.2e}, "
              f"max rel dplaq {p_err:.2e}")
        assert e_err < 1e-6 and q_err < 1e-4 and p_err < 1e-6, form
        t0m, t064 = scale.t0(smx, V), scale.t0(s64, V)
        assert (t0m is None) == (t064 is None) and (t064 is None or rel(t0m, t064) < 1e-6), form
        w0m, w064 = scale.w0(smx, V), scale.w0(s64, V)
        assert (w0m is None) == (w064 is None) and (w064 is None or rel(w0m, w064) < 1e-6), form
        assert bad == 0 and dev < 1e-6, form
    m = s64[::5]
    assert np.all(np.diff(m[:, 1]) > 0)  # the plaquette rises along the flow


@pytest.mark.parametrize("candidate,eps,start,therm", [("exph", 0.24, "hot", 300), ("su2", 0.12, "cold", 1000)])
def test_metropolis_equilibrium_plaquette(candidate, eps, start, therm):
    """Production Metropolis agrees with the Wilson action statistically: on
    16^4 at beta = 6.0, after `therm` sweeps of 10 hits, the plaquette
    averaged over 150 more sweeps (one call per sweep, as MCUpdate) is the
    published equilibrium value 0.5937 to 1e-3, for both candidates.  The
    R*S*T candidate (three SU(2) rotations by eps each, at half the step;
    18 % acceptance) drifts up from a Haar start over thousands of sweeps, so
    it starts cold and comes down; the acceptance stays between 0.05 and 0.9."""
    dims = (16, 16, 16, 16)
    V = orc.volume(dims)
    with LatticeContext(dims) as ctx:
        if start == "hot":
            ctx.set_hot(5)
        acc = ctx.metropolis(6.0, eps, 10, therm, 0, candidate, 99) / (therm * 4 * V * 10)
        assert 0.05 < acc < 0.9
        plaq = []
        for sw in range(150):
            ctx.metropolis(6.0, eps, 10, 1, therm + sw, candidate, 99)
            plaq.append(ctx.observables()[0])
    mean = float(np.mean(plaq))
    print(f"beta 6.0, {candidate} eps {eps}: acceptance {acc:.3f}, <P> = {mean:.6f} (spread {np.std(plaq):.1e})")
    assert abs(mean - 0.5937) < 1e-3


def test_64x64x64x64_fp64_equals_tiled_32x32x32x32():
    """A lattice beyond 2 GiB per direction plane in FP64 (64^4: 16.8 M sites,
    2.4 GiB per plane): a 32^4 configuration tiled 2 x 2 x 2 x 2 is periodic
    with the smaller period, so one flow step on 64^4 (whole x-rows in LDS)
    gives in every copy the 32^4 step (slab staging) bit for bit, and the
    observables are those of 32^4 (plaquette) and 16 times them (energy, Q)."""
    small_dims, big_dims = (32, 32, 32, 32), (64, 64, 64, 64)
    with LatticeContext(small_dims) as ctx:
        ctx.set_hot(81)
        ctx.metropolis(6.0, 0.24, 10, 3, 0, "exph", 81)
        start = ctx.download()
        obs_small = ctx.observables()
        ctx.flow_steps(0.02, 1)
        step_small = ctx.download()
    shape5 = (4, 32, 32, 32, 32, 18)  # mu, t, z, y, x, element
    big = np.tile(start.reshape(shape5), (1, 2, 2, 2, 2, 1)).reshape(4, -1, 18)
    del start
    with LatticeContext(big_dims) as ctx:
        ctx.upload(big)
        obs_big = ctx.observables()
        ctx.flow_steps(0.02, 1)
        big = ctx.download()
    out = big.reshape(4, 64, 64, 64, 64, 18)
    want = step_small.reshape(shape5)
    for t0 in (0, 32):
        for z0 in (0, 32):
            for y0 in (0, 32):
                for x0 in (0, 32):
                    np.testing.assert_array_equal(out[:, t0:t0 + 32, z0:z0 + 32, y0:y0 + 32, x0:x0 + 32], want)
    assert abs(obs_big[0] / obs_small[0] - 1) < 1e-13
    assert abs(obs_big[1] / (16 * obs_small[1]) - 1) < 1e-12
    assert abs(obs_big[2] - 16 * obs_small[2]) < 1e-12 * obs_big[3]
