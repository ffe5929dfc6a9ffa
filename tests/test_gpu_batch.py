"""Batched ensemble path (include/lym_batch.h): each lattice of an uneven
batch must give what the single-lattice path gives on it, whatever the launch
schedule, with and without the HIP graph; graph replays are checked against
eager runs with eager work between them."""
import numpy as np
import pytest

from latticeyangmills_amd import LatticeContext
from latticeyangmills_amd.batch import BatchContext
from tests import oracle_lib as orc

pytestmark = pytest.mark.gpu

DIMS = [(4, 4, 4, 4), (6, 4, 4, 2), (8, 4, 4, 4), (4, 4, 6, 6)]


@pytest.fixture(scope="module")
def inputs():
    return [orc.hot_start(d, 100 + i) for i, d in enumerate(DIMS)]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_upload_download_and_observables(inputs, precision):
    with BatchContext(DIMS, precision) as b:
        for c, L in enumerate(inputs):
            b.upload(c, L)
        obs = b.observables()
        for c, L in enumerate(inputs):
            got = b.download(c)
            if precision == "fp64":
                assert np.array_equal(got, L)
            want = orc.observables(L, DIMS[c])
            tol = 1e-12 if precision == "fp64" else 1e-6
            assert abs(obs[c, 0] / want[0] - 1) < tol
            assert abs(obs[c, 1] / want[1] - 1) < tol
            assert abs(obs[c, 2] - want[2]) < tol * want[3]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("grid,kind,target,lds", [(0, "linear", 0, False), (0, "contig", 0, False),
                                                   (0, "column", 64, False), (8, "column", 128, False),
                                                   (3, "linear", 0, False), (0, "column", 64, True),
                                                   (16, "contig", 0, True)])
def test_flow_matches_single_lattice_path(inputs, precision, grid, kind, target, lds):
    eps, tau, every = 0.02, 0.2, 3
    singles = []
    for c, L in enumerate(inputs):
        with LatticeContext(DIMS[c], precision) as ctx:
            ctx.set_schedule("linear", 0)
            ctx.upload(L)
            s = ctx.flow(eps, tau, every)
            singles.append((s, ctx.download()))
    with BatchContext(DIMS, precision) as b:
        b.set_schedule(grid, kind, target, lds)
        for c, L in enumerate(inputs):
            b.upload(c, L)
        series = b.flow(eps, tau, every, use_graph=False)
        for c in range(len(DIMS)):
            s, links = singles[c]
            # The exp series length is chosen per wave (9 or 18 terms, both
            # truncating below 2e-18), so another grouping of sites into waves
            # may round a few links differently: rounding-level, not bitwise.
            # Mixed links are stored in FP32, so such a difference is one FP32 ulp.
            fp64 = precision == "fp64"
            assert np.abs(b.download(c) - links).max() < (1e-13 if fp64 else 2e-6)
            m = ~np.isnan(s[:, 1])
            assert np.array_equal(np.isnan(series[c, :, 1]), ~m)
            # observables differ by that rounding and by summation order (Q cancels: absolute bound)
            assert np.allclose(series[c][m][:, [0, 1, 3]], s[m][:, [0, 1, 3]], rtol=1e-13 if fp64 else 1e-7, atol=0)
            assert np.allclose(series[c][m][:, 2], s[m][:, 2], rtol=0, atol=1e-11 if fp64 else 1e-5)


def test_graph_replays_match_eager(inputs):
    eps, tau, every = 0.02, 0.16, 2  # 7 steps: odd, so the start buffer alternates
    with BatchContext(DIMS, "mixed") as b:
        for c, L in enumerate(inputs):
            b.upload(c, L)
        eager = b.flow(eps, tau, every, use_graph=False)
        eager_links = [b.download(c) for c in range(len(DIMS))]
        for rep in range(4):
            for c, L in enumerate(inputs):
                b.upload(c, L)
            b.observables()  # eager work between replays
            got = b.flow(eps, tau, every, use_graph=True)
            assert np.array_equal(np.nan_to_num(got, nan=7.0), np.nan_to_num(eager, nan=7.0)), rep
            for c in range(len(DIMS)):
                assert np.array_equal(b.download(c), eager_links[c]), (rep, c)


@pytest.mark.parametrize("eps", [0.24, 0.3])
def test_metropolis_matches_single_lattice_path(inputs, eps):
    seeds = [11, 22, 33, 44]
    with BatchContext(DIMS, "fp64") as b:
        for c, L in enumerate(inputs):
            b.upload(c, L)
        acc = b.metropolis(6.0, eps, 4, 2, seeds, first_sweep=3)
        for c, L in enumerate(inputs):
            with LatticeContext(DIMS[c], "fp64") as ctx:
                ctx.upload(L)
                a1 = ctx.metropolis(6.0, eps, 4, 2, 3, "exph", seeds[c])
                assert a1 == acc[c]
                assert np.array_equal(ctx.download(), b.download(c))


def test_hot_start_matches_single(inputs):
    seeds = [5, 6, 7, 8]
    with BatchContext(DIMS, "fp64") as b:
        b.set_hot(seeds)
        for c in range(len(DIMS)):
            with LatticeContext(DIMS[c], "fp64") as ctx:
                ctx.set_hot(seeds[c])
                assert np.array_equal(ctx.download(), b.download(c))


def test_autotune_keeps_results(inputs, tmp_path):
    cache = str(tmp_path / "tune.txt")
    with BatchContext(DIMS, "mixed") as b:
        for c, L in enumerate(inputs):
            b.upload(c, L)
        ref = b.flow(0.02, 0.1, 1, use_graph=False)
        summary = b.autotune(cache)
        assert "grid=" in summary and "measured" in summary
        for c, L in enumerate(inputs):
            b.upload(c, L)
        got = b.flow(0.02, 0.1, 1, use_graph=True)
        # the tuned schedule may group sites into waves differently (FP32 ulp level)
        assert np.allclose(got[..., [0, 1, 3]], ref[..., [0, 1, 3]], rtol=1e-7, atol=0)
        assert np.allclose(got[..., 2], ref[..., 2], rtol=0, atol=1e-5)
        assert "cached" in b.autotune(cache)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("grid", [0, 8])
@pytest.mark.parametrize("dims", [[(8, 8, 2, 2), (16, 4, 2, 2), (32, 2, 2, 4)],
                                  [(8, 8, 2, 2), (6, 4, 4, 2), (16, 4, 2, 2), (12, 2, 2, 2)]])
def test_slab_staging_batch(precision, grid, dims):
    """All lattices fit the slab staging (whole x-rows, y halo in LDS), or only
    some do and each workgroup picks the staging of its lattice."""
    inputs = [orc.hot_start(d, 60 + i) for i, d in enumerate(dims)]
    singles = []
    for d, L in zip(dims, inputs):
        with LatticeContext(d, precision) as ctx:
            ctx.set_schedule("linear", 0)
            ctx.upload(L)
            singles.append((ctx.flow(0.02, 0.1, 1), ctx.download()))
    with BatchContext(dims, precision) as b:
        b.set_schedule(grid, "contig", 0, True)
        for c, L in enumerate(inputs):
            b.upload(c, L)
        series = b.flow(0.02, 0.1, 1, use_graph=True)
        fp64 = precision == "fp64"
        for c in range(len(dims)):
            s, links = singles[c]
            assert np.abs(b.download(c) - links).max() < (1e-13 if fp64 else 2e-6)
            assert np.allclose(series[c][:, [0, 1, 3]], s[:, [0, 1, 3]], rtol=1e-13 if fp64 else 1e-7, atol=0)


def test_batch_even_odd_order_is_invisible(inputs):
    """The batch keeps its links in even-odd order after lym_batch_metropolis:
    one-sweep calls equal one two-sweep call, an upload of one lattice leaves
    the others' updated links in place, and observables and downloads read
    the lexicographic links."""
    seeds = list(range(31, 31 + len(DIMS)))
    with BatchContext(DIMS, "fp64") as a, BatchContext(DIMS, "fp64") as b:
        for c, L in enumerate(inputs):
            a.upload(c, L)
            b.upload(c, L)
        a.metropolis(6.0, 0.24, 3, 1, seeds, first_sweep=0)
        a.metropolis(6.0, 0.24, 3, 1, seeds, first_sweep=1)
        b.metropolis(6.0, 0.24, 3, 2, seeds, first_sweep=0)
        want = [b.download(c) for c in range(len(DIMS))]
        a.upload(0, inputs[0])  # after sweeps: lattice 0 replaced, the rest kept
        assert np.array_equal(a.download(0), inputs[0])
        for c in range(1, len(DIMS)):
            np.testing.assert_array_equal(a.download(c), want[c])
        b.upload(0, inputs[0])
        a.metropolis(6.0, 0.24, 3, 1, seeds, first_sweep=2)
        b.metropolis(6.0, 0.24, 3, 1, seeds, first_sweep=2)
        for c in range(len(DIMS)):
            b.upload(c, b.download(c))
        np.testing.assert_array_equal(a.observables(), b.observables())
