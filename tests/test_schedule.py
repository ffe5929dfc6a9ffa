"""Host logic of the XCD-aware work schedules (latticeyangmills_amd/csrc/lym_sched.hpp):
every schedule visits each 64-site run exactly once, balances the XCDs, and
COLUMN orders an XCD's runs column by column, time slice by time slice."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = r'''
#include <cstdio>
#include <cstdlib>
#include "lym_sched.hpp"
int main(int argc, char **argv) {
    int kind = atoi(argv[1]), target = atoi(argv[2]), n = (argc - 3) / 4;
    std::vector<std::array<int, 4>> dims; std::vector<long> run0; long r = 0;
    for (int c = 0; c < n; c++) {
        std::array<int, 4> d{{atoi(argv[3 + 4*c]), atoi(argv[4 + 4*c]), atoi(argv[5 + 4*c]), atoi(argv[6 + 4*c])}};
        dims.push_back(d); run0.push_back(r); r += (long(d[0]) * d[1] * d[2] * d[3] + 63) / 64;
    }
    lym::HostSchedule h = lym::build_schedule(dims, run0, kind, target);
    printf("%d\n", h.per_xcd);
    for (int v : h.order) printf("%d\n", v);
}
'''


@pytest.fixture(scope="module")
def tool(tmp_path_factory):
    d = tmp_path_factory.mktemp("sched")
    (d / "s.cpp").write_text(SRC)
    exe = d / "s"
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "latticeyangmills_amd", "csrc"),
                    str(d / "s.cpp"), "-o", str(exe)], check=True)
    return exe


def table(tool, kind, target, dims_list):
    out = subprocess.run([str(tool), str(kind), str(target)] + [str(x) for d in dims_list for x in d],
                         capture_output=True, text=True, check=True).stdout.split()
    per = int(out[0])
    return per, np.array([int(x) for x in out[1:]]).reshape(8, per) if per else np.zeros((8, 0), int)


@pytest.mark.parametrize("kind,target", [(1, 0), (2, 1024), (2, 2048), (2, 4096), (2, 8192)])
@pytest.mark.parametrize("dims_list", [[(32, 32, 32, 64)], [(24, 24, 24, 24)], [(48, 48, 48, 96)],
                                       [(8, 8, 8, 8)] * 3 + [(12, 12, 12, 12), (16, 16, 16, 32)], [(6, 4, 4, 2)]])
def test_schedule_is_balanced_permutation(tool, kind, target, dims_list):
    per, t = table(tool, kind, target, dims_list)
    nruns = sum((int(np.prod(d)) + 63) // 64 for d in dims_list)
    runs = t[t >= 0]
    assert runs.size == nruns and np.array_equal(np.sort(runs), np.arange(nruns))
    counts = (t >= 0).sum(axis=1)
    assert counts.max() - counts.min() <= 1


def test_column_order_marches_time_slices(tool):
    dims = (32, 32, 32, 64)
    per, t = table(tool, 2, 2048, [dims])
    first = t[0][t[0] >= 0]
    # a 2048-site column: 64 rows of 32 sites = 8 x 8 (y, z) rows; runs of 2 rows
    tcoord = (first * 64) // (32 * 32 * 32)
    # within one column (32 runs per time slice, 64 slices) t is non-decreasing
    steps = tcoord[: (2048 // 64) * 64]
    assert np.all(np.diff(steps) >= 0)
    assert steps[0] == 0 and steps[-1] == 63
