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


def read_series(path):
    """flow/confNNNN.dat: the reference's header line, then rows [t, plaq, topc, energy]"""
    with open(path) as f:
        assert f.readline() == "Flow Time\tPlaq.\tTopc.\tEnerg.\t\n"
    return np.loadtxt(path, skiprows=1)


def test_io_selftest(tmp_path):
    assert "io-selftest ok" in run("io-selftest", tmp_path)
    a = read_bin(tmp_path / "selftest.bin", (8, 8, 8, 20))  # 10240 sites: a full I/O block and a partial one
    assert a.shape == (4, 10240, 18)
    # the writer fills link (mu, site s, element k) with mu + 0.25 k + 1e-5 s:
    # checked here independently of the reader, on both sides of the 8192-site block
    mu = np.arange(4)[:, None, None]
    s_ = np.arange(10240)[None, :, None]
    k = np.arange(18)[None, None, :]
    want = (mu + 0.25 * k) + 1e-5 * s_
    np.testing.assert_array_equal(a, want)


@pytest.mark.gpu
def test_generate_then_flow(tmp_path):
    from latticeyangmills_amd import LatticeContext
    dims = (4, 4, 4, 8)
    confs, out = tmp_path / "confs", tmp_path / "out"
    run("generate", "--lattice", *dims, "--beta", 6.0, "--eps", 0.24, "--therm", 5, "--steps", 4, "--confs", 2,
        "--start", "Hot", "--seed", 3, "--output", confs)
    files = sorted(f for f in os.listdir(confs) if f.endswith(".bin"))
    assert files == ["conf0000.bin", "conf0001.bin"]
    assert os.path.isfile(confs / "observables.dat")
    # conf0000 = hot start + 5 thermalisation + 2 sampling sweeps (NCorr = steps/confs)
    with LatticeContext(dims, "fp64") as ctx:
        ctx.set_hot(3)
        ctx.metropolis(6.0, 0.24, 10, 7, 0, "exph", 3)
        assert np.array_equal(ctx.download(), read_bin(confs / "conf0000.bin", dims))
    run("flow", "--lattice", *dims, "--eps", 0.02, "--tau", 0.2, "--input", confs, "--output", out)
    for k, f in enumerate(files):
        got = read_series(out / "flow" / f"conf{k:04d}.dat")
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
    got = read_series(out / "flow" / "conf0001.dat")
    want, _ = orc.flow(read_bin(confs / "conf0001.bin", dims), dims, 0.02, 0.2, 1, 0)
    assert np.all(np.abs(got[:, 1] / want[:, 1] - 1) < 1e-12)
    assert np.all(np.abs(got[:, 3] / want[:, 3] - 1) < 1e-12)


@pytest.mark.gpu
def test_pgfg_reports_and_writes_observables(tmp_path):
    """PGFG as the reference's GaugeFieldFactory (gaugefieldfactory.cpp:95-128,
    outputobs.cpp:30-55): every NCorr sweeps of the thermalisation a report;
    while sampling, observables.dat gets a row (step, the observable's value,
    here the plaquette) for each saved configuration, measured on the links
    that are saved; the hits and candidates in use are printed first.  The
    reference-compatible mode takes 30 hits of R*S*T candidates."""
    import json
    confs = tmp_path / "confs"
    pgfg = {"LatticeSize": {"x": 4, "y": 4, "z": 4, "t": 8}, "SubLatticeSize": {"x": 4, "y": 4, "z": 4, "t": 8},
            "Action": {"type": "puregauge", "params": {"beta": 6.0}}, "Observables": ["superobs"],
            "App": {"type": "PGFG", "params": {"MCSteps": 6, "thermSteps": 4, "startType": "Hot", "NConf": 3,
                                               "SU3eps": 0.24}},
            "OutputDir": str(confs), "seed": 5, "referenceMetropolis": True}
    (tmp_path / "pgfg.json").write_text(json.dumps(pgfg))
    out = run("run", tmp_path / "pgfg.json")
    assert "Metropolis: 30 hits per link, R*S*T" in out and "the reference's hits and candidates" in out
    assert out.count("Thermalization Step") == 2  # steps 2 and 4 (NCorr = 6 / 3)
    assert out.count("Configuration Number") == 3
    lines = open(confs / "observables.dat").read().splitlines()
    assert lines[0] == "Step\tSuperObs\t"
    rows = [line.split("\t") for line in lines[1:]]
    assert [int(r[0]) for r in rows] == [2, 4, 6]
    dims = (4, 4, 4, 8)
    for k, r in enumerate(rows):
        plaq = orc.observables(read_bin(confs / f"conf{k:04d}.bin", dims), dims)[0]
        assert abs(float(r[1]) / plaq - 1) < 1e-12, k
    # the same sweeps through the C ABI: hot start, 4 + 6 sweeps, 30 hits, R*S*T
    from latticeyangmills_amd import LatticeContext
    with LatticeContext(dims) as ctx:
        ctx.set_hot(5)
        ctx.metropolis(6.0, 0.24, 30, 10, 0, "su2", 5)
        np.testing.assert_array_equal(ctx.download(), read_bin(confs / "conf0002.bin", dims))


@pytest.mark.gpu
def test_flow_batched_equals_one_by_one(tmp_path):
    """GpuWilsonFlow flows its input files in one lym_batch_flow call (the
    default) or one after another (--no-batch): the same flow/confNNNN.dat
    files, FP64, at 1e-12 (the links agree bit for bit; the observables are
    the same site terms summed in another order)."""
    dims = (4, 4, 4, 8)
    confs = tmp_path / "confs"
    run("generate", "--lattice", *dims, "--beta", 6.0, "--eps", 0.24, "--therm", 3, "--steps", 6, "--confs", 3,
        "--start", "Hot", "--seed", 9, "--output", confs)
    run("flow", "--lattice", *dims, "--eps", 0.02, "--tau", 0.3, "--input", confs, "--output", tmp_path / "b")
    run("flow", "--lattice", *dims, "--eps", 0.02, "--tau", 0.3, "--input", confs, "--output", tmp_path / "s",
        "--no-batch")
    for k in range(3):
        b = read_series(tmp_path / "b" / "flow" / f"conf{k:04d}.dat")
        s = read_series(tmp_path / "s" / "flow" / f"conf{k:04d}.dat")
        assert b.shape == s.shape == (15, 4)
        np.testing.assert_array_equal(b[:, 0], s[:, 0])
        assert np.all(np.abs(b[:, [1, 3]] / s[:, [1, 3]] - 1) < 1e-12), k
        assert np.all(np.abs(b[:, 2] - s[:, 2]) < 1e-12), k


@pytest.mark.gpu
@pytest.mark.parametrize("batch", [True, False])
def test_flow_observables_hold_last_measured_row(tmp_path, batch):
    """With measureEvery = 4 and 10 rows, row 9 is not measured (NaN in the
    file); the observables report row 8, the last measured one."""
    dims = (4, 4, 4, 8)
    confs = tmp_path / "confs"
    run("generate", "--lattice", *dims, "--beta", 6.0, "--eps", 0.24, "--therm", 2, "--steps", 4, "--confs", 2,
        "--start", "Hot", "--seed", 4, "--output", confs)
    args = ["flow", "--lattice", *dims, "--eps", 0.02, "--tau", 0.2, "--measure-every", 4, "--input", confs,
            "--output", tmp_path / "o"] + ([] if batch else ["--no-batch"])
    out = run(*args)
    last = read_series(tmp_path / "o" / "flow" / "conf0001.dat")
    assert last.shape == (10, 4) and np.isnan(last[9, 1]) and np.isfinite(last[8, 1])
    line = [x for x in out.splitlines() if x.startswith("last measured row")][0].split()
    plaq, topc, energy = float(line[4]), float(line[6]), float(line[8])
    assert (plaq, topc, energy) == (last[8, 1], last[8, 2], last[8, 3])


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


def write_bin(path, links):
    """(4, V, 18) links -> .bin file (per site, x fastest, the four links)"""
    np.ascontiguousarray(links.transpose(1, 0, 2)).astype(np.float64).tofile(path)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_flow_resumes_bit_identically_from_a_checkpoint(tmp_path, precision):
    """A flow stopped after 8 steps, its links written as a .bin checkpoint,
    and resumed from that file by the drop-in App (lym_app flow, which reads
    it with InputConf's format) continues bit for bit: the resumed series'
    observables equal rows 8..20 of the uninterrupted flow.  In mixed
    precision the checkpoint holds the FP32 links, whose row 2 the upload
    derives again exactly as the device had it."""
    from latticeyangmills_amd import LatticeContext
    dims, eps = (4, 4, 4, 8), 0.02
    links = orc.thermalised(dims, beta=6.0, sweeps=4, seed=13)
    with LatticeContext(dims, precision) as ctx:
        ctx.upload(links)
        whole = ctx.flow(eps, 21.5 * eps, 1)  # 21 rows, t = 0 .. 20 eps
        ctx.upload(links)
        ctx.flow_steps(eps, 8)
        ckpt = ctx.download()
    (tmp_path / "ckpt").mkdir()
    write_bin(tmp_path / "ckpt" / "conf0000.bin", ckpt)
    run("flow", "--lattice", *dims, "--eps", eps, "--tau", 13.5 * eps, "--input", tmp_path / "ckpt", "--output",
        tmp_path / "out", "--precision", precision, "--no-batch", "--no-autotune")
    resumed = read_series(tmp_path / "out" / "flow" / "conf0000.dat")
    assert resumed.shape == (13, 4)
    np.testing.assert_array_equal(resumed[:, 1:], whole[8:21, 1:])
