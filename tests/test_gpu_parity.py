"""Parity of the HIP path (through the C ABI of include/lym.h) against the CPU
oracle on identical inputs.

Tolerances: FP64 path 1e-12 relative on plaquette and energy at every flow
time; Q to 1e-12 of its cancellation scale (sum of |terms|), because Q sums
terms that cancel heavily.  Metropolis in test mode consumes the same
pre-drawn uniforms as the oracle and must reproduce it link by link.
"""
import numpy as np
import pytest

from latticeyangmills_amd import LatticeContext
from tests import oracle_lib as orc

pytestmark = pytest.mark.gpu

DIMS = [(4, 4, 4, 4), (4, 6, 8, 2), (6, 4, 4, 8)]


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def assert_obs_close(got, want, tol=1e-12):
    plaq, energy, topc, scale = want
    assert rel(got[0], plaq) < tol, (got, want)
    assert rel(got[1], energy) < tol, (got, want)
    assert abs(got[2] - topc) <= tol * max(scale, 1e-300) + 1e-300, (got, want)


@pytest.fixture(scope="module")
def therm8():
    dims = (8, 8, 8, 8)
    return dims, orc.thermalised(dims, beta=6.0, sweeps=8, seed=11)


@pytest.mark.parametrize("dims", DIMS)
def test_upload_download_roundtrip(dims):
    links = orc.hot_start(dims, 3)
    with LatticeContext(dims) as ctx:
        ctx.upload(links)
        assert np.array_equal(ctx.download(), links)


@pytest.mark.parametrize("dims", DIMS)
def test_cold_start(dims):
    with LatticeContext(dims) as ctx:
        plaq, energy, topc, scale = ctx.observables()
        assert plaq == pytest.approx(1.0, abs=1e-15)
        assert abs(energy) < 1e-20 and abs(topc) < 1e-20
        assert np.array_equal(ctx.download(), orc.cold_start(dims))


@pytest.mark.parametrize("dims", DIMS)
def test_hot_start_matches_oracle(dims):
    want = orc.hot_start(dims, 1234)
    with LatticeContext(dims) as ctx:
        ctx.set_hot(1234)
        got = ctx.download()
    assert np.max(np.abs(got - want)) < 1e-12


@pytest.mark.parametrize("dims", DIMS)
def test_observables_hot(dims):
    links = orc.hot_start(dims, 5)
    want = orc.observables(links, dims, 0)
    with LatticeContext(dims) as ctx:
        ctx.upload(links)
        got = ctx.observables()
    assert_obs_close(got, want)
    assert rel(got[3], want[3]) < 1e-12


def test_observables_thermalised(therm8):
    dims, links = therm8
    want = orc.observables(links, dims, 0)
    with LatticeContext(dims) as ctx:
        ctx.upload(links)
        got = ctx.observables()
    assert_obs_close(got, want)


@pytest.mark.parametrize("dims", DIMS)
@pytest.mark.parametrize("mu", [0, 1, 2, 3])
def test_staples_and_derivative(dims, mu):
    links = orc.hot_start(dims, 17)
    with LatticeContext(dims) as ctx:
        ctx.upload(links)
        om, z = ctx.staples(mu), ctx.derivative(mu)
    om_ref, z_ref = orc.staples(links, dims, mu), orc.derivative(links, dims, mu)
    assert np.max(np.abs(om - om_ref)) < 1e-12 * np.max(np.abs(om_ref))
    assert np.max(np.abs(z - z_ref)) < 1e-12 * np.max(np.abs(z_ref))


@pytest.mark.parametrize("dims,eps", [((4, 4, 4, 4), 0.02), ((4, 6, 8, 2), 0.02), ((4, 4, 4, 4), 2.0)])
def test_one_flow_step(dims, eps):
    """On this hot start c1 = Tr Z^2 / 2 has median 3.8, so with eps = 2 even
    stage 1 (Q = eps Z / 4, median c1 0.95) puts most links above the series
    bound c1 = 0.5: the exponential takes its Morningstar-Peardon branch."""
    links = orc.hot_start(dims, 23)
    want = orc.flow_step(links, dims, eps)
    with LatticeContext(dims) as ctx:
        ctx.upload(links)
        ctx.flow_steps(eps, 1)
        got = ctx.download()
    assert np.max(np.abs(got - want)) < 1e-12


def test_flow_series_thermalised(therm8):
    """WilsonFlow::applyWilsonFlow on 8^4, eps 0.01 to t = 0.2, every row."""
    dims, links = therm8
    want, want_links = orc.flow(links, dims, 0.01, 0.2, 1, 0)
    with LatticeContext(dims) as ctx:
        ctx.upload(links)
        got = ctx.flow(0.01, 0.2, 1)
        got_links = ctx.download()
    assert got.shape == want.shape == (20, 4)
    assert np.array_equal(got[:, 0], want[:, 0])
    scale = orc.observables(links, dims)[3]
    for r in range(want.shape[0]):
        assert rel(got[r, 1], want[r, 1]) < 1e-12, r
        assert rel(got[r, 3], want[r, 3]) < 1e-12, r
        assert abs(got[r, 2] - want[r, 2]) < 1e-12 * scale, r
    assert np.max(np.abs(got_links - want_links)) < 1e-12


def test_flow_measure_every(therm8):
    dims, links = therm8
    with LatticeContext(dims) as ctx:
        ctx.upload(links)
        every = ctx.flow(0.02, 0.3, 5)
        ctx.upload(links)
        all_rows = ctx.flow(0.02, 0.3, 1)
    assert every.shape == all_rows.shape == (15, 4)
    for r in range(15):
        if r % 5 == 0:
            assert np.array_equal(every[r], all_rows[r])
        else:
            assert np.all(np.isnan(every[r, 1:]))


@pytest.mark.parametrize("candidate,eps", [("su2", 0.24), ("exph", 0.24), ("exph", 0.3)])
@pytest.mark.parametrize("dims", [(4, 4, 4, 4), (2, 4, 6, 4)])
def test_metropolis_predrawn_per_link(dims, candidate, eps):
    """Test mode: identical pre-drawn uniforms give identical links.  exp(i eps H)
    candidates take the series-only exponential for 8 eps^2 <= 0.5 (eps 0.24)
    and the general one above (eps 0.3)."""
    V = orc.volume(dims)
    nhit, nsweeps = 5, 2
    nu = 13 if candidate == "su2" else 9
    rng = np.random.default_rng(99)
    u = rng.random(nsweeps * 4 * V * nhit * nu)
    links = orc.hot_start(dims, 8)
    cid = 0 if candidate == "su2" else 1
    want, acc_want = orc.metropolis(links, dims, 6.0, eps, nhit, nsweeps, 0, cid, 0, u)
    with LatticeContext(dims) as ctx:
        ctx.upload(links)
        acc = ctx.metropolis(6.0, eps, nhit, nsweeps, 0, candidate, 0, u)
        got = ctx.download()
    assert acc == acc_want
    assert np.max(np.abs(got - want)) < 1e-12


@pytest.mark.parametrize("candidate,eps", [("su2", 0.24), ("exph", 0.24), ("exph", 0.3)])
def test_metropolis_philox_per_link(candidate, eps):
    """Production mode: the oracle draws the same Philox stream."""
    dims = (4, 4, 4, 4)
    links = orc.hot_start(dims, 4)
    cid = 0 if candidate == "su2" else 1
    want, acc_want = orc.metropolis(links, dims, 6.0, eps, 10, 3, 7, cid, 2024)
    with LatticeContext(dims) as ctx:
        ctx.upload(links)
        acc = ctx.metropolis(6.0, eps, 10, 3, 7, candidate, 2024)
        got = ctx.download()
    assert acc == acc_want
    assert np.max(np.abs(got - want)) < 1e-11


def test_metropolis_philox_per_link_column_order():
    """A lattice whose (z, y) planes hold more than one Metropolis column
    (32 x 32 x 32 x 4: four columns of 4096 sites per parity, each marched
    along t by one XCD): the block order leaves every link update as the
    oracle's."""
    dims = (32, 32, 32, 4)
    links = orc.hot_start(dims, 6)
    want, acc_want = orc.metropolis(links, dims, 6.0, 0.24, 2, 1, 0, 1, 77)
    with LatticeContext(dims) as ctx:
        ctx.upload(links)
        acc = ctx.metropolis(6.0, 0.24, 2, 1, 0, "exph", 77)
        got = ctx.download()
    assert acc == acc_want
    assert np.max(np.abs(got - want)) < 1e-11


def test_mixed_precision_against_fp64(therm8):
    """FP32 links/staples/clover, FP64 exp and reductions: error stated vs FP64."""
    dims, links = therm8
    want, _ = orc.flow(links, dims, 0.02, 0.5, 5, 1)
    with LatticeContext(dims, "mixed") as ctx:
        ctx.upload(links)
        got = ctx.flow(0.02, 0.5, 5)
    for r in range(0, want.shape[0], 5):
        assert rel(got[r, 1], want[r, 1]) < 1e-6, r
        assert rel(got[r, 3], want[r, 3]) < 1e-6, r
        assert abs(got[r, 2] - want[r, 2]) < 1e-4, r


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("kind,target,lds", [("linear", 0, False), ("contig", 0, False), ("column", 64, False),
                                             ("column", 300, False), ("linear", 0, True), ("column", 300, True)])
def test_schedule_does_not_change_results(therm8, precision, kind, target, lds):
    dims, links = therm8
    with LatticeContext(dims, precision) as ctx:
        ctx.set_schedule("linear", 0)
        ctx.upload(links)
        ref = ctx.flow(0.02, 0.2, 3)
        ref_links = ctx.download()
        ctx.set_schedule(kind, target, lds)
        ctx.upload(links)
        got = ctx.flow(0.02, 0.2, 3)
        # Each link's arithmetic depends on its own neighbourhood only (the
        # exponential's series length is chosen per lane): the links are the
        # same bit for bit under every schedule and staging mode.
        np.testing.assert_array_equal(ctx.download(), ref_links)
        # observables: the same site terms summed in another tree order
        m = ~np.isnan(ref[:, 1])
        assert np.allclose(got[m][:, [0, 1, 3]], ref[m][:, [0, 1, 3]], rtol=1e-14, atol=0)
        assert np.allclose(got[m][:, 2], ref[m][:, 2], rtol=0, atol=1e-13)


def test_autotune_picks_a_schedule(therm8, tmp_path):
    dims, links = therm8
    with LatticeContext(dims, "fp64") as ctx:
        ctx.upload(links)
        s = ctx.autotune(str(tmp_path / "tune.txt"))
        assert "schedule=" in s and "measured" in s
        assert "cached" in ctx.autotune(str(tmp_path / "tune.txt"))
        assert np.array_equal(ctx.download(), links)  # tuning leaves the field alone


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_autotune_between_flow_calls(precision):
    """The tuner times flow steps on the context's own field and puts it back:
    tuned between two flow calls (the mixed field then in two-row form), the
    links and the series go on exactly as without tuning."""
    dims = (8, 8, 4, 8)
    links = orc.hot_start(dims, 71)
    with LatticeContext(dims, precision) as a, LatticeContext(dims, precision) as b:
        for ctx in (a, b):
            ctx.upload(links)
            ctx.flow(0.02, 0.06, 1)
        mid = b.download()
        assert "measured" in a.autotune(None)
        np.testing.assert_array_equal(a.download(), mid)
        got, want = a.flow(0.02, 0.1, 2), b.flow(0.02, 0.1, 2)
        np.testing.assert_array_equal(a.download(), b.download())
        m = ~np.isnan(want[:, 1])
        assert np.allclose(got[m][:, [0, 1, 3]], want[m][:, [0, 1, 3]], rtol=1e-14, atol=0)
        assert np.allclose(got[m][:, 2], want[m][:, 2], rtol=0, atol=1e-13)


@pytest.mark.parametrize("dims", [(6, 4, 4, 8), (12, 4, 2, 4)])
def test_lds_variant_without_whole_rows(dims):
    """Lx not dividing 64: only own-site links come from LDS; results unchanged."""
    links = orc.hot_start(dims, 31)
    want = orc.flow_step(links, dims, 0.02)
    with LatticeContext(dims) as ctx:
        ctx.set_schedule("contig", 0, True)
        ctx.upload(links)
        ctx.flow_steps(0.02, 1)
        assert np.max(np.abs(ctx.download() - want)) < 1e-12


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-13), ("mixed", 1e-6)])
def test_check_links_reports_health(precision, tol):
    """NaN and non-unitary links are reported (count, max deviation), not fatal.
    A link scaled by 1.1 deviates by 0.331 in FP64 (det 1.331); an FP32
    field keeps the scaled rows 0, 1 and derives row 2 from them (scaled by
    1.21), so there it deviates by 0.4641 (det 1.4641)."""
    dims = (4, 4, 4, 4)
    links = orc.hot_start(dims, 21)
    with LatticeContext(dims, precision) as ctx:
        ctx.upload(links)
        dev, bad = ctx.check_links()
        assert dev < tol and bad == 0
        broken = links.copy()
        broken[2, 17, 5] = np.nan      # one link with a NaN
        broken[1, 3, :] *= 1.1         # one link scaled: U U^+ = 1.21, det = 1.331
        ctx.upload(broken)
        dev, bad = ctx.check_links()
        assert bad == 1
        assert abs(dev - (0.331 if precision == "fp64" else 0.4641)) < 1e-5
        ctx.set_cold()
        assert ctx.check_links() == (0.0, 0)


@pytest.mark.parametrize("dims", [(32, 4, 2, 2), (16, 8, 2, 4), (8, 8, 2, 2), (4, 16, 2, 2), (2, 32, 2, 2),
                                  (16, 4, 4, 2)])
@pytest.mark.parametrize("kind", ["linear", "contig"])
def test_slab_staging(dims, kind):
    """Runs of whole x-rows with a y halo staged in LDS (Lx | 64, Lx <= 32,
    Ly a multiple of 64 / Lx), including Ly = 64 / Lx where both halo rows are
    the run's own rows: one FP64 step equals the oracle, and the mixed path
    stays within FP32 rounding of it."""
    links = orc.hot_start(dims, 41)
    want = orc.flow_step(links, dims, 0.02)
    with LatticeContext(dims) as ctx:
        ctx.set_schedule(kind, 0, True)
        ctx.upload(links)
        ctx.flow_steps(0.02, 1)
        assert np.max(np.abs(ctx.download() - want)) < 1e-12
    with LatticeContext(dims, "mixed") as ctx:
        ctx.set_schedule(kind, 0, True)
        ctx.upload(links)
        ctx.flow_steps(0.02, 1)
        assert np.max(np.abs(ctx.download() - want)) < 2e-6


@pytest.mark.parametrize("dims", [(48, 4, 2, 2), (48, 8, 2, 2), (48, 12, 2, 4), (80, 4, 2, 2), (112, 8, 2, 2)])
@pytest.mark.parametrize("kind", ["linear", "column"])
def test_tile_staging(dims, kind):
    """Lattices with Lx a multiple of 16 that does not divide 64 take 16 x 4
    tiles of one (z, t) plane with a halo ring in x and y (Ly = 4 makes both
    halo rows the tile's own rows): one FP64 step equals the oracle, and the
    mixed path stays within FP32 rounding of it."""
    links = orc.hot_start(dims, 43)
    want = orc.flow_step(links, dims, 0.02)
    with LatticeContext(dims) as ctx:
        ctx.set_schedule(kind, 1024 if kind == "column" else 0, True)
        ctx.upload(links)
        ctx.flow_steps(0.02, 1)
        assert np.max(np.abs(ctx.download() - want)) < 1e-12
    with LatticeContext(dims, "mixed") as ctx:
        ctx.set_schedule(kind, 1024 if kind == "column" else 0, True)
        ctx.upload(links)
        ctx.flow_steps(0.02, 1)
        assert np.max(np.abs(ctx.download() - want)) < 2e-6


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_flow_graph_replays_match_eager(precision):
    """lym_flow replayed from its captured graph equals the eager launches bit
    for bit, on every replay: three replays with eager work and host
    synchronisation between them; an odd step count alternates the start
    buffer, so both cached graphs are captured and replayed; a schedule
    change drops them and a new capture follows."""
    dims = (4, 4, 4, 8)
    links = orc.hot_start(dims, 47)
    eps, tau, every = 0.02, 0.14, 2  # 7 rows, 6 steps
    with LatticeContext(dims, precision) as a, LatticeContext(dims, precision) as b:
        b.set_graph(True)
        a.upload(links)
        b.upload(links)
        for rep in range(4):
            if rep == 2:
                a.set_schedule("column", 1024, True)
                b.set_schedule("column", 1024, True)
            want = a.flow(eps, tau, every)
            got = b.flow(eps, tau, every)
            assert np.array_equal(np.isnan(want), np.isnan(got))
            assert np.array_equal(want[~np.isnan(want)], got[~np.isnan(got)]), rep
            np.testing.assert_array_equal(b.download(), a.download())
            a.flow_steps(eps, 1)  # eager work between replays: odd, so the start buffer alternates
            b.flow_steps(eps, 1)
            b.synchronize()
            a.synchronize()
        b.set_graph(False)
        np.testing.assert_array_equal(b.flow(eps, tau, every)[:, 1:3], a.flow(eps, tau, every)[:, 1:3])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_even_odd_order_is_invisible(precision):
    """lym_metropolis leaves the links in even-odd order inside the context
    until another call reads them.  One-sweep calls interleaved with every
    reading entry point equal a context that runs the same sweeps in one call
    and passes the links through host memory before each reading call."""
    dims = (4, 4, 4, 6)
    links = orc.hot_start(dims, 53)
    with LatticeContext(dims, precision) as a, LatticeContext(dims, precision) as b:
        a.upload(links)
        b.upload(links)
        a.metropolis(6.0, 0.24, 3, 1, 0, "exph", 9)
        a.metropolis(6.0, 0.24, 3, 1, 1, "exph", 9)  # already in even-odd order
        b.metropolis(6.0, 0.24, 3, 2, 0, "exph", 9)
        b.upload(b.download())
        assert np.array_equal(a.observables(), b.observables())
        a.metropolis(6.0, 0.24, 3, 1, 2, "exph", 9)
        b.metropolis(6.0, 0.24, 3, 1, 2, "exph", 9)
        b.upload(b.download())
        np.testing.assert_array_equal(a.staples(1), b.staples(1))
        a.metropolis(6.0, 0.24, 3, 1, 3, "exph", 9)
        b.metropolis(6.0, 0.24, 3, 1, 3, "exph", 9)
        b.upload(b.download())
        np.testing.assert_array_equal(a.derivative(2), b.derivative(2))
        a.metropolis(6.0, 0.24, 3, 1, 4, "exph", 9)
        b.metropolis(6.0, 0.24, 3, 1, 4, "exph", 9)
        assert a.check_links() == b.check_links()  # order-free: a max and a count
        b.upload(b.download())
        a.flow_steps(0.02, 3)
        b.flow_steps(0.02, 3)
        a.metropolis(6.0, 0.24, 3, 1, 5, "exph", 9)
        b.metropolis(6.0, 0.24, 3, 1, 5, "exph", 9)
        b.upload(b.download())
        sa, sb = a.flow(0.02, 0.1, 1), b.flow(0.02, 0.1, 1)
        np.testing.assert_array_equal(sa, sb)
        a.metropolis(6.0, 0.24, 3, 1, 6, "exph", 9)
        b.metropolis(6.0, 0.24, 3, 1, 6, "exph", 9)
        np.testing.assert_array_equal(a.download(), b.download())
        a.metropolis(6.0, 0.24, 3, 1, 7, "exph", 9)
        a.set_hot(5)  # overwrites every link: no reorder owed
        b.set_hot(5)
        np.testing.assert_array_equal(a.download(), b.download())


def test_upload_download_round_trip_large():
    """Downloads of more than one 64 MB bounce chunk, copied out by several
    host threads (32 x 32 x 32 x 16: 75.5 MB per direction, a full chunk and a
    partial one), return the uploaded links bit for bit, into fresh arrays
    and twice in a row."""
    dims = (32, 32, 32, 16)
    V = int(np.prod(dims))
    rng = np.random.default_rng(11)
    links = rng.standard_normal((4, V, 18))
    with LatticeContext(dims) as ctx:
        ctx.upload(links)
        np.testing.assert_array_equal(ctx.download(), links)
        np.testing.assert_array_equal(ctx.download(), links)


def test_staples_large_lattice_against_small_path():
    """lym_staples / lym_derivative on a lattice whose output passes through
    two bounce chunks (32 x 32 x 32 x 16, 75.5 MB) equal the oracle's at
    every site."""
    dims = (32, 32, 32, 16)
    links = orc.hot_start(dims, 19)
    with LatticeContext(dims) as ctx:
        ctx.upload(links)
        om = ctx.staples(2)
        z = ctx.derivative(3)
    want_om = orc.staples(links, dims, 2)
    want_z = orc.derivative(links, dims, 3)
    assert np.max(np.abs(om - want_om)) < 1e-12
    assert np.max(np.abs(z - want_z)) < 1e-12


@pytest.mark.parametrize("dims", [(8, 8, 2, 2), (32, 4, 2, 2), (48, 4, 2, 2), (6, 4, 4, 8), (16, 8, 2, 4)])
@pytest.mark.parametrize("kind,lds", [("linear", False), ("linear", True), ("column", True)])
def test_two_row_form_flow(dims, kind, lds):
    """Mixed path with the links in two-row form while the flow runs (row 2
    rebuilt by unitarity where read; slab, tile, own-link and no staging):
    one step within FP32 rounding of the oracle, and a 5-step series and the
    final links equal to the three-row mixed path's bit for bit (FP32 links
    hold the row 2 that the two-row form rebuilds); the field comes back in
    three-row form."""
    links = orc.hot_start(dims, 61)
    want = orc.flow_step(links, dims, 0.02)
    target = 256 if kind == "column" else 0
    with LatticeContext(dims, "mixed") as ctx:
        ctx.set_schedule(kind, target, lds, two_row=True)
        ctx.upload(links)
        ctx.flow_steps(0.02, 1)
        got = ctx.download()
        assert np.max(np.abs(got - want)) < 2e-6
        ctx.upload(links)
        s2 = ctx.flow(0.02, 0.12, 1)
        dev, bad = ctx.check_links()
        assert bad == 0 and dev < 1e-6
        l2 = ctx.download()
    with LatticeContext(dims, "mixed") as ctx:
        ctx.set_schedule(kind, target, lds)
        ctx.upload(links)
        s3 = ctx.flow(0.02, 0.12, 1)
        l3 = ctx.download()
    np.testing.assert_array_equal(s2, s3)
    np.testing.assert_array_equal(l2, l3)


def test_two_row_form_graph_replays_and_tuning():
    """The two-row flow replayed from its graph equals its eager launches bit
    for bit (the form conversions are part of the graph); the tuner, which
    times two-row candidates on a scratch copy, leaves the field as it was."""
    dims = (16, 8, 4, 4)
    links = orc.hot_start(dims, 62)
    with LatticeContext(dims, "mixed") as a, LatticeContext(dims, "mixed") as b:
        for ctx in (a, b):
            ctx.set_schedule("column", 512, True, two_row=True)
        b.set_graph(True)
        for rep in range(3):
            a.upload(links)
            b.upload(links)
            want, got = a.flow(0.02, 0.14, 2), b.flow(0.02, 0.14, 2)
            assert np.array_equal(np.isnan(want), np.isnan(got))
            assert np.array_equal(want[~np.isnan(want)], got[~np.isnan(got)]), rep
            np.testing.assert_array_equal(b.download(), a.download())
        a.upload(links)
        summary = a.autotune(None)
        assert "measured" in summary
        b.set_graph(False)
        b.upload(links)
        np.testing.assert_array_equal(a.download(), b.download())


def test_two_row_form_between_calls_is_invisible():
    """The mixed flow leaves the links in two-row form between calls; every
    entry point that reads them gives what it gives on the same links passed
    through host memory (three-row form), and the reference's loop of one
    step and one measurement per call gives lym_flow's series bit for bit."""
    dims = (8, 4, 4, 8)
    links = orc.hot_start(dims, 57)
    eps = 0.02
    readers = [
        lambda c: np.array(c.observables()),
        lambda c: c.staples(1),
        lambda c: c.derivative(2),
        lambda c: np.array(c.check_links()),
        lambda c: c.download(),
        lambda c: (c.metropolis(6.0, 0.24, 4, 1, 0, "exph", 5), c.download())[1],
    ]
    with LatticeContext(dims, "mixed") as a, LatticeContext(dims, "mixed") as b:
        for c in (a, b):
            c.set_schedule("column", 256, True, two_row=True)
            c.upload(links)
        for i, read in enumerate(readers):
            a.flow_steps(eps, 2)
            b.flow_steps(eps, 2)
            b.upload(b.download())  # the same links, now in three-row form
            np.testing.assert_array_equal(read(a), read(b), err_msg=str(i))
        a.upload(links)
        series = a.flow(eps, 0.105, 1)
        a.upload(links)
        loop = [a.observables()]
        for _ in range(4):
            a.flow_steps(eps, 1)
            loop.append(a.observables())
        np.testing.assert_array_equal(series[:, 1], [o[0] for o in loop])
        np.testing.assert_array_equal(series[:, 3], [o[1] for o in loop])
        np.testing.assert_array_equal(series[:, 2], [o[2] for o in loop])
