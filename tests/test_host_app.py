"""The C++ host classes (GpuWilsonFlow, GpuGaugeFieldFactory, GpuSuperObs,
GpuPureGauge) through the lym_app front-end, and the reference's .bin
configuration format (src/InputOutput/inputconf.cpp:114-155)."""
import os
import subprocess

import numpy as np
import pytest

from tests import oracle_lib as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
APP = os.path.join(ROOT, "latticeyangmills_amd", "lym_app")


def read_bin(path, dims):
    """.bin file: per site (x fastest) the four links, 18 doubles each -> (4, V, 18)"""
    V = int(np.prod(dims))
    raw = np.fromfile(path, dtype=np.float64)
    assert raw.size == V * 4 * 18
    return np.ascontiguousarray(raw.reshape(V, 4, 18).transpose(1, 0, 2))


def run(*args):
    r = subprocess.run([APP, *map(str, args)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return r.stdout


def test_io_selftest(tmp_path):
    assert "io-selftest ok" in run("io-selftest", tmp_path)
    a = read_bin(tmp_path / "selftest.bin", (8, 8, 8, 20))  # 10240 sites: a full I/O block and a partial one
    assert a.shape == (4, 10240, 18) and np.all(np.abs(a) <= 1)


@pytest.mark.gpu
def test_generate_then_flow(tmp_path):
    from latticeyangmills_amd import LatticeContext
    dims = (4, 4, 4, 8)
    confs, out = tmp_path / "confs", tmp_path / "out"
    run("generate", "--lattice", *dims, "--beta", 6.0, "--eps", 0.24, "--therm", 5, "--steps", 4, "--confs", 2,
        "--start", "Hot", "--seed", 3, "--output", confs)
    files = sorted(os.listdir(confs))
    assert files == ["conf0000.bin", "conf0001.bin"]
    # conf0000 = hot start + 5 thermalisation + 2 sampling sweeps (NCorr = steps/confs)
    with LatticeContext(dims, "fp64") as ctx:
        ctx.set_hot(3)
        ctx.metropolis(6.0, 0.24, 10, 7, 0, "exph", 3)
        assert np.array_equal(ctx.download(), read_bin(confs / "conf0000.bin", dims))
    run("flow", "--lattice", *dims, "--eps", 0.02, "--tau", 0.2, "--input", confs, "--output", out)
    for k, f in enumerate(files):
        got = np.loadtxt(out / "flow" / f"conf{k:04d}.dat")
        want, _ = orc.flow(read_bin(confs / f, dims), dims, 0.02, 0.2, 1, 0)
        scale = orc.observables(read_bin(confs / f, dims), dims)[3]
        assert got.shape == want.shape == (10, 4)
        assert np.allclose(got[:, 0], want[:, 0], rtol=0, atol=1e-15)
        assert np.all(np.abs(got[:, 1] / want[:, 1] - 1) < 1e-12)
        assert np.all(np.abs(got[:, 3] / want[:, 3] - 1) < 1e-12)
        assert np.all(np.abs(got[:, 2] - want[:, 2]) < 1e-10 * scale)


def run_rc(*args):
    return subprocess.run([APP, *map(str, args)], capture_output=True, text=True, timeout=120)


def test_json_front_end_rejects_bad_input(tmp_path):
    bad = tmp_path / "bad.json"
    bad.write_text('{"App": {"type": "FLOW", "params": {"epsilon": 0.01, "tauFinal": 1.0}}, "LatticeSize": {"x": 4,')
    r = run_rc("run", bad)
    assert r.returncode == 1 and "JSON:" in r.stderr
    sub = tmp_path / "sub.json"
    sub.write_text('{"App": {"type": "FLOW", "params": {"epsilon": 0.01, "tauFinal": 1.0}},'
                   '"LatticeSize": {"x": 4, "y": 4, "z": 4, "t": 8}, "SubLatticeSize": {"x": 4, "y": 4, "z": 4, "t": 4},'
                   '"Action": {"type": "puregauge", "params": {"beta": 6.0}}, "Observables": ["superobs"],'
                   '"InputDir": ".", "OutputDir": "."}')
    r = run_rc("run", sub)
    assert r.returncode == 1 and "SubLatticeSize" in r.stderr


@pytest.mark.gpu
def test_json_front_end_pgfg_then_flow(tmp_path):
    """The reference's own input format (jsonapp.cpp:49-89): PGFG writes configs, FLOW flows them."""
    import json
    confs, out = tmp_path / "confs", tmp_path / "out"
    common = {"LatticeSize": {"x": 4, "y": 4, "z": 4, "t": 8}, "SubLatticeSize": {"x": 4, "y": 4, "z": 4, "t": 8},
              "Action": {"type": "puregauge", "params": {"beta": 6.0}}, "Observables": ["superobs"]}
    pgfg = dict(common, App={"type": "PGFG", "params": {"MCSteps": 4, "thermSteps": 5, "startType": "Hot",
                                                        "NConf": 2, "SU3eps": 0.24}}, OutputDir=str(confs), seed=3)
    (tmp_path / "pgfg.json").write_text(json.dumps(pgfg))
    assert "generated 2" in run("run", tmp_path / "pgfg.json")
    flow = dict(common, App={"type": "FLOW", "params": {"epsilon": 0.02, "tauFinal": 0.2}},
                InputDir=str(confs), OutputDir=str(out))
    (tmp_path / "flow.json").write_text(json.dumps(flow))
    assert "flowed 2" in run("run", tmp_path / "flow.json")
    dims = (4, 4, 4, 8)
    got = np.loadtxt(out / "flow" / "conf0001.dat")
    want, _ = orc.flow(read_bin(confs / "conf0001.bin", dims), dims, 0.02, 0.2, 1, 0)
    assert np.all(np.abs(got[:, 1] / want[:, 1] - 1) < 1e-12)
    assert np.all(np.abs(got[:, 3] / want[:, 3] - 1) < 1e-12)


def test_scales_of_a_flow_series(tmp_path):
    """lym_app scales (host C++ Observables/flowscale.h, Utils/latticeunits.h)
    agrees with the closed forms and with latticeyangmills_amd.scale."""
    import math
    from latticeyangmills_amd import scale
    V, c = 4096, 0.06
    rows = [[t, 0.9, 0.0, -64.0 * V * (c / t if t > 0 else 0.0)] for t in np.arange(0, 8.0, 0.02)]
    f = tmp_path / "conf0000.dat"
    f.write_text("".join("%.17g %.17g %.17g %.17g\n" % tuple(r) for r in rows))
    out = dict(line.split(" = ") for line in run("scales", f, V, 6.2).splitlines())
    assert abs(float(out["t0/a^2"]) - 0.3 / c) < 1e-9
    assert abs(float(out["w0/a"]) - math.sqrt(0.3 / c)) < 1e-9
    assert abs(float(out["a"].split()[0]) - scale.lattice_spacing_fm(6.2)) < 1e-12
    assert abs(float(out["t0"].split()[0]) - 0.3 / c * scale.lattice_spacing_fm(6.2) ** 2) < 1e-12
