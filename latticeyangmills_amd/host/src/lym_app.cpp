// Command-line front-end over the GPU App classes, in place of the reference's
// JSON-driven main (src/MainPrograms/LatticeQCDMain.cpp:5-64) for its FLOW and
// PGFG apps:
//   lym_app flow --lattice Lx Ly Lz Lt --eps E --tau T --input DIR --output DIR
//                [--precision fp64|mixed] [--measure-every K]
//   lym_app generate --lattice Lx Ly Lz Lt --beta B --eps E --therm N --steps M
//                --confs K [--start Cold|Hot] [--seed S] [--nhit H] --output DIR
//   lym_app run INPUT.json     (the reference's own input files: App FLOW or
//                              PGFG, LatticeSize, Action puregauge, Observables,
//                              InputDir, OutputDir; optional build keys
//                              "precision", "measureEvery", "batchFlow" (FLOW:
//                              false flows the inputs one by one), "autotune"
//                              (FLOW: false keeps the default schedule),
//                              "tuneCache" (FLOW: file of tuned schedules), "seed",
//                              "nHit", "candidate" ("exph" | "su2") and
//                              "referenceMetropolis" (PGFG: true takes the
//                              reference's 30 hits of R*S*T candidates))
//   flow also takes --no-batch, --no-autotune, --tune-cache FILE; generate
//   --candidate exph|su2
//   lym_app io-selftest DIR   (host only: writes and reads back a .bin file)
//   lym_app scales FILE VOLUME [BETA]
//                             (host only: t0/a^2 and w0/a of a flow series file
//                              flow/confNNNN.dat, and in fm when BETA is given)
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "Actions/gpupuregauge.h"
#include "Apps/gpugaugefieldfactory.h"
#include "Apps/gpuwilsonflow.h"
#include "InputOutput/confio.h"
#include "InputOutput/jsonlite.h"
#include "Observables/flowscale.h"
#include "Observables/gpusuperobs.h"
#include "Utils/latticeunits.h"

namespace {
struct Args {
    std::array<int, 4> lat{{0, 0, 0, 0}};
    double eps = 0.01, tau = 1.0, beta = 6.0;
    int therm = 0, steps = 0, confs = 1, nhit = 10, every = 1, precision = LYM_PREC_FP64, candidate = LYM_CAND_EXPH;
    bool batch = true, autotune = true;
    unsigned long long seed = 0;
    std::string input = ".", output = ".", start = "Cold", tune_cache;
};

int usage()
{
    fprintf(stderr, "usage: lym_app run INPUT.json | flow ... | generate ... | scales FILE VOLUME [BETA] | io-selftest DIR\n"
                    "       (options are listed at the top of lym_app.cpp)\n");
    return 2;
}

bool parse(int argc, char **argv, Args &a)
{
    for (int i = 2; i < argc; i++) {
        std::string k = argv[i];
        auto next = [&](void) -> const char * { return (i + 1 < argc) ? argv[++i] : nullptr; };
        const char *v = nullptr;
        if (k == "--lattice") {
            for (int d = 0; d < 4; d++) {
                if (!(v = next())) return false;
                a.lat[d] = atoi(v);
            }
            continue;
        }
        if (k == "--no-batch") {
            a.batch = false;
            continue;
        }
        if (k == "--no-autotune") {
            a.autotune = false;
            continue;
        }
        if (!(v = next())) return false;
        if (k == "--eps") a.eps = atof(v);
        else if (k == "--candidate") a.candidate = (std::string(v) == "su2") ? LYM_CAND_SU2 : LYM_CAND_EXPH;
        else if (k == "--tau") a.tau = atof(v);
        else if (k == "--beta") a.beta = atof(v);
        else if (k == "--therm") a.therm = atoi(v);
        else if (k == "--steps") a.steps = atoi(v);
        else if (k == "--confs") a.confs = atoi(v);
        else if (k == "--nhit") a.nhit = atoi(v);
        else if (k == "--measure-every") a.every = atoi(v);
        else if (k == "--seed") a.seed = strtoull(v, nullptr, 10);
        else if (k == "--input") a.input = v;
        else if (k == "--output") a.output = v;
        else if (k == "--tune-cache") a.tune_cache = v;
        else if (k == "--start") a.start = v;
        else if (k == "--precision") a.precision = (std::string(v) == "mixed") ? LYM_PREC_MIXED : LYM_PREC_FP64;
        else return false;
    }
    return true;
}
}  // namespace

// JSON input of the reference's front-end (src/InputOutput/JsonInput/JsonGeneric/jsonapp.cpp:49-89,
// jsonlattice.cpp:11-73, JsonAction/jsonaction.cpp:60-89, JsonObservable/jsonobservable.cpp:41-78,
// jsondirectories.cpp:15-55).  SubLatticeSize must equal LatticeSize (one GPU holds the whole volume).
int run_json(const char *path)
{
    using LatticeIO::JsonValue;
    const JsonValue j = JsonValue::parseFile(path);
    const JsonValue &ls = j["LatticeSize"];
    const std::array<int, 4> lat{{ls["x"].integer(), ls["y"].integer(), ls["z"].integer(), ls["t"].integer()}};
    if (j.has("SubLatticeSize")) {
        const JsonValue &ss = j["SubLatticeSize"];
        const std::array<int, 4> sub{{ss["x"].integer(), ss["y"].integer(), ss["z"].integer(), ss["t"].integer()}};
        if (sub != lat) throw std::runtime_error("SubLatticeSize must equal LatticeSize: one GPU holds the whole lattice");
    }
    if (j["Action"]["type"].string() != "puregauge") throw std::runtime_error("only the puregauge action exists");
    const double beta = j["Action"]["params"]["beta"].number();
    const int precision = (j.has("precision") && j["precision"].string() == "mixed") ? LYM_PREC_MIXED : LYM_PREC_FP64;
    const std::string out = j.has("OutputDir") ? j["OutputDir"].string() : ".";
    const std::string type = j["App"]["type"].string();
    const JsonValue &p = j["App"]["params"];
    bool superobs = false;
    if (j.has("Observables"))
        for (auto &o : j["Observables"].arr) superobs = superobs || o.string() == "superobs";
    GpuSuperObs obs(precision);
    GpuPureGauge act(beta, precision);
    if (type == "FLOW") {
        GpuWilsonFlow app(p["tauFinal"].number(), p["epsilon"].number(), precision,
                          j.has("measureEvery") ? j["measureEvery"].integer() : 1);
        if (j.has("batchFlow")) app.setBatchFlow(j["batchFlow"].b);
        app.setAutotune(!j.has("autotune") || j["autotune"].b, j.has("tuneCache") ? j["tuneCache"].string() : "");
        app.setAction(&act);
        if (superobs) app.addObservable(&obs);
        app.createLattice(lat);
        app.setInputDir(j["InputDir"].string());
        app.setOutputDir(out);
        app.execute();
        printf("flowed %zu configuration(s) to t=%g\n", app.inputList().size(), p["tauFinal"].number());
        return 0;
    }
    if (type == "PGFG") {
        const int mc = p["MCSteps"].integer();
        const int nconf = p.has("NConf") ? p["NConf"].integer() : mc / p["NCorr"].integer();
        const bool refMetro = j.has("referenceMetropolis") && j["referenceMetropolis"].b;
        int nhit = refMetro ? 30 : 10, cand = refMetro ? LYM_CAND_SU2 : LYM_CAND_EXPH;
        if (j.has("nHit")) nhit = j["nHit"].integer();
        if (j.has("candidate")) cand = j["candidate"].string() == "su2" ? LYM_CAND_SU2 : LYM_CAND_EXPH;
        GpuGaugeFieldFactory app(mc, p["thermSteps"].integer(), nconf, p["SU3eps"].number(), p["startType"].string(),
                                 beta, j.has("seed") ? (unsigned long long)j["seed"].number() : 0ull, nhit, cand);
        app.setAction(&act);
        if (superobs) app.addObservable(&obs);
        app.createLattice(lat);
        app.setOutputDir(out);
        app.execute();
        printf("generated %d configuration(s), acceptance %.4f\n", nconf, app.acceptance());
        return 0;
    }
    throw std::runtime_error("unsupported App type " + type + " (FLOW and PGFG are built)");
}

int main(int argc, char **argv)
{
    if (argc < 2) return usage();
    const std::string cmd = argv[1];
    try {
        if (cmd == "run") {
            if (argc < 3) return usage();
            return run_json(argv[2]);
        }
        if (cmd == "scales") {
            if (argc < 4) return usage();
            std::vector<std::array<double, 4>> rows;
            FILE *f = fopen(argv[2], "r");
            if (!f) throw std::runtime_error(std::string("cannot read ") + argv[2]);
            std::array<double, 4> r;
            char line[512];
            while (fgets(line, sizeof line, f))  // rows of numbers; the header line is skipped
                if (sscanf(line, "%lf %lf %lf %lf", &r[0], &r[1], &r[2], &r[3]) == 4) rows.push_back(r);
            fclose(f);
            const FlowScales sc = flowScales(rows, atol(argv[3]));
            printf("t0/a^2 = %.12g\nw0/a = %.12g\n", sc.t0, sc.w0);
            if (argc > 4) {
                const double a = LatticeUnits::calculateLatticeSpacing(atof(argv[4]));
                printf("a = %.12g fm\nt0 = %.12g fm^2\nw0 = %.12g fm\n", a, sc.t0 * a * a, sc.w0 * a);
            }
            return 0;
        }
        if (cmd == "io-selftest") {
            if (argc < 3) return usage();
            const std::array<int, 4> dims{{8, 8, 8, 20}};  // 10240 sites: more than one I/O block
            GluonField a(dims), b(dims);
            // a closed form of (mu, site, element), so a reader of the file can
            // check every value, on both sides of an I/O block boundary
            for (int mu = 0; mu < 4; mu++)
                for (int s = 0; s < a[mu].sites; s++)
                    for (int k = 0; k < 18; k++) a[mu].at(s).mat[size_t(k)] = mu + 0.25 * k + 1e-5 * s;
            const std::string f = std::string(argv[2]) + "/selftest.bin";
            LatticeIO::OutputConf::writeConf(a, f.c_str());
            LatticeIO::InputConf::readConf(b, f.c_str());
            for (int mu = 0; mu < 4; mu++)
                for (int s = 0; s < a[mu].sites; s++)
                    if (a[mu].at(s).mat != b[mu].at(s).mat) {
                        fprintf(stderr, "mismatch mu=%d s=%d\n", mu, s);
                        return 1;
                    }
            printf("io-selftest ok %s\n", f.c_str());
            return 0;
        }
        Args a;
        if (!parse(argc, argv, a) || a.lat[0] < 2) return usage();
        if (cmd == "flow") {
            GpuWilsonFlow app(a.tau, a.eps, a.precision, a.every);
            app.setBatchFlow(a.batch);
            app.setAutotune(a.autotune, a.tune_cache);
            GpuSuperObs obs(a.precision);
            GpuPureGauge act(a.beta, a.precision);
            app.setAction(&act);
            app.addObservable(&obs);
            app.createLattice(a.lat);
            app.setInputDir(a.input);
            app.setOutputDir(a.output);
            app.execute();
            printf("flowed %zu configuration(s) to t=%g\n", app.inputList().size(), a.tau);
            printf("last measured row: plaq %.17g topc %.17g energy %.17g\n", obs.plaq, obs.topc, obs.energy);
            return 0;
        }
        if (cmd == "generate") {
            GpuGaugeFieldFactory app(a.steps, a.therm, a.confs, a.eps, a.start, a.beta, a.seed, a.nhit, a.candidate);
            GpuSuperObs obs(LYM_PREC_FP64);
            GpuPureGauge act(a.beta);
            app.setAction(&act);
            app.addObservable(&obs);
            app.createLattice(a.lat);
            app.setOutputDir(a.output);
            app.execute();
            printf("generated %d configuration(s), acceptance %.4f, last plaquette %.12f\n", a.confs,
                   app.acceptance(), obs.plaq);
            return 0;
        }
    } catch (const std::exception &e) {
        fprintf(stderr, "lym_app: %s\n", e.what());
        return 1;
    }
    return usage();
}
