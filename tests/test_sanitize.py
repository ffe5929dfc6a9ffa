"""AddressSanitizer + UndefinedBehaviorSanitizer build of the CPU code: the
oracle (oracle/lym_oracle.c) and the host layer that needs no GPU
(SU(3) algebra and exp, Lattice maps, configuration and series I/O, the JSON
reader, scale setting), driven through their entry points on small lattices,
malformed input included.  Host code only: no device code is built here."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "latticeyangmills_amd", "host")
SAN = ["-fsanitize=address,undefined", "-fno-omit-frame-pointer", "-fno-sanitize-recover=all", "-g", "-O1"]

DRIVER = r'''
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>
#include "InputOutput/confio.h"
#include "InputOutput/jsonlite.h"
#include "Math/field.h"
#include "Math/lattice.h"
#include "Observables/flowscale.h"
#include "Utils/latticeunits.h"
#include "lym_oracle.h"

static int fails = 0;
#define CHECK(c) do { if (!(c)) { fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); fails++; } } while (0)

static void oracle_paths()
{
    const int dims[2][4] = {{2, 4, 4, 4}, {4, 4, 4, 4}};
    for (const auto &d : dims) {
        const int V = d[0] * d[1] * d[2] * d[3];
        std::vector<double> links(size_t(4) * V * 18), w(size_t(V) * 18);
        or_hot_start(links.data(), d, 7);
        double obs[4];
        or_observables(links.data(), d, 0, obs);
        or_observables(links.data(), d, 1, obs);
        for (int mu = 0; mu < 4; mu++) {
            or_staples(links.data(), d, mu, w.data());
            or_derivative(links.data(), d, mu, w.data());
        }
        CHECK(or_flow_step(links.data(), d, 0.02) == 0);
        CHECK(or_flow_step(links.data(), d, 2.0) == 0);  // exp above the series range
        std::vector<double> series(size_t(or_flow_rows(0.1, 0.02)) * 4);
        CHECK(or_flow(links.data(), d, 0.02, 0.1, 3, 0, series.data()) == or_flow_rows(0.1, 0.02));
        int64_t acc = 0;
        for (int cand = 0; cand < 2; cand++) {
            CHECK(or_metropolis(links.data(), d, 6.0, 0.24, 3, 2, 0, cand, 11, nullptr, &acc) == 0);
            std::vector<double> u(size_t(2) * 4 * V * 3 * or_uniforms_per_hit(cand), 0.37);
            CHECK(or_metropolis(links.data(), d, 6.0, 0.3, 3, 2, 5, cand, 11, u.data(), &acc) == 0);
        }
    }
    double q[18] = {0}, r[18];
    q[0] = 0.3; q[8] = -0.1; q[16] = -0.2; q[2] = 0.05; q[3] = 0.01; q[6] = 0.05; q[7] = -0.01;
    or_su3_exp(q, r);
    uint32_t c[4] = {1, 2, 3, 4}, k[2] = {5, 6}, o[4];
    or_philox4x32_10(c, k, o);
}

static void host_paths(const char *dir)
{
    SU3 q(0.0);
    q.mat[0] = 0.4; q.mat[8] = -0.1; q.mat[16] = -0.3; q.mat[2] = 0.2; q.mat[6] = 0.2;
    SU3 e = exp(q), z(0.0);
    CHECK(std::abs(getMultSumTrace(e, ~e) - 3.0) < 1e-13);
    CHECK(exp(z).realTrace() == 3.0);
    std::array<int, 4> size{{8, 8, 8, 20}};
    GluonField a(size), b(size);
    for (int mu = 0; mu < 4; mu++) {
        setToIdentity(a[mu]);
        a[mu] *= 0.5 + mu;
    }
    Lattice<SU3> s = a[1] + a[2] - a[0];
    s = adj(s) * a[3];
    Lattice<double> rt = realTrace(exp(a[0] * 0.0));
    CHECK(rt.at(77) == 3.0);
    CHECK(sum(a[0]).realTrace() == 1.5 * a[0].sites);
    const std::string f = std::string(dir) + "/s.bin";
    LatticeIO::OutputConf::writeConf(a, f.c_str());
    LatticeIO::InputConf::readConf(b, f.c_str());
    CHECK(b[3].at(10239).mat[0] == 3.5);
    std::array<int, 4> big{{8, 8, 8, 21}};
    GluonField c(big);
    bool threw = false;
    try { LatticeIO::InputConf::readConf(c, f.c_str()); } catch (const std::runtime_error &) { threw = true; }
    CHECK(threw);  // a short file is an error, not a read past the end
    std::vector<std::array<double, 4>> rows;
    for (int i = 0; i < 300; i++) {
        const double t = 0.02 * i;
        rows.push_back({{t, 0.9, 0.0, i % 3 ? -64.0 * 4096 * 0.06 / (t + 1e-9) : NAN}});
    }
    LatticeIO::OutputObs::writeFlowObs(std::string(dir) + "/flow.dat", rows);
    const FlowScales sc = flowScales(rows, 4096);
    CHECK(std::isfinite(sc.t0));
    LatticeUnits::initialize(6.2, size);
    CHECK(LatticeUnits::latticeSpacing > 0.06 && LatticeUnits::latticeSpacing < 0.075);
    using LatticeIO::JsonValue;
    const JsonValue j = JsonValue::parse(
        "{\"App\": {\"type\": \"FLOW\", \"params\": {\"epsilon\": 0.01, \"tauFinal\": 1.0}}, \"LatticeSize\": "
        "{\"x\": 4, \"y\": 4, \"z\": 4, \"t\": 8}, \"Observables\": [\"superobs\"], \"flag\": true, \"n\": null, "
        "\"s\": \"a\\\"b\\\\c\"}");
    CHECK(j["LatticeSize"]["t"].integer() == 8 && j["flag"].b && j["s"].string() == "a\"b\\c");
    const char *bad[] = {"{", "{\"a\": }", "[1, 2", "{\"a\": tru}", "\"abc", "{\"a\" 1}", "", "{\"a\": 1}}", "[1e999999]"};
    for (const char *t : bad) {
        try { JsonValue::parse(t); } catch (const std::runtime_error &) { }
    }
    threw = false;
    try { (void)j["missing"]; } catch (const std::runtime_error &) { threw = true; }
    CHECK(threw);
}

int main(int argc, char **argv)
{
    oracle_paths();
    host_paths(argc > 1 ? argv[1] : ".");
    if (fails) return 1;
    printf("sanitize ok\n");
    return 0;
}
'''


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("gcc") is None, reason="needs gcc and g++")
def test_oracle_and_host_code_under_asan_ubsan(tmp_path):
    (tmp_path / "driver.cpp").write_text(DRIVER)
    objs = []
    inc = ["-I", os.path.join(ROOT, "oracle"), "-I", os.path.join(ROOT, "include"), "-I", os.path.join(HOST, "include")]
    o = tmp_path / "oracle.o"
    subprocess.run(["gcc", "-std=gnu11", "-ffp-contract=off", *SAN, "-c", os.path.join(ROOT, "oracle", "lym_oracle.c"),
                    "-o", str(o)], check=True)
    objs.append(str(o))
    for src in ["math.cpp", "confio.cpp", "jsonlite.cpp", "scale.cpp"]:
        o = tmp_path / (src + ".o")
        subprocess.run(["g++", "-std=c++17", *SAN, *inc, "-c", os.path.join(HOST, "src", src), "-o", str(o)],
                       check=True)
        objs.append(str(o))
    exe = tmp_path / "san"
    subprocess.run(["g++", "-std=c++17", *SAN, *inc, str(tmp_path / "driver.cpp"), *objs, "-o", str(exe), "-lm"],
                   check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1:abort_on_error=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([str(exe), str(tmp_path)], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0 and "sanitize ok" in r.stdout, r.stdout[-2000:] + r.stderr[-4000:]
