// GPU Action / Observable / App classes: each member replaces the reference's
// host loop of the same name by a call through include/lym.h.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <filesystem>
#include <stdexcept>
#include <string>

#include "Actions/gpupuregauge.h"
#include "Apps/gpugaugefieldfactory.h"
#include "Apps/gpuwilsonflow.h"
#include "InputOutput/confio.h"
#include "Observables/flowscale.h"
#include "Observables/gpusuperobs.h"
#include "Utils/latticeunits.h"
#include "lym_batch.h"

static_assert(sizeof(SU3) == 18 * sizeof(double), "SU3 must be 18 packed doubles for the C ABI");

// ---------------------------------------------------------------- action ---

void GpuPureGauge::computeStaples(int mu)
{
    DeviceField &d = device();
    d.toDevice();
    m_omega.allocate(m_lat->m_size);
    d.check(lym_staples(d.ctx(), mu, m_omega.lattice.data()->mat.data()), "lym_staples");
    m_omegaMu = mu;
}

Lattice<SU3> GpuPureGauge::computeDerivative(int mu)
{
    DeviceField &d = device();
    d.toDevice();
    Lattice<SU3> z(m_lat->m_size);
    d.check(lym_derivative(d.ctx(), mu, z.lattice.data()->mat.data()), "lym_derivative");
    return z;
}

double GpuPureGauge::compute(int x, int y, int z, int t, int mu, SU3 &newLink)
{
    if (m_omegaMu != mu) computeStaples(mu);
    device().toHost();
    const SU3 diff = newLink - (*m_lat)[mu].at(x, y, z, t);
    return -m_beta / 3.0 * getMultSumTrace(diff, m_omega.at(x, y, z, t));
}

// ----------------------------------------------------------- observables ---

void GpuSuperObs::compute()
{
    DeviceField &d = DeviceField::of(m_field, m_precision);
    d.toDevice();
    double out[4];
    d.check(lym_observables(d.ctx(), out), "lym_observables");
    plaq = out[0];
    energy = out[1];
    topc = out[2];
    topcScale = out[3];
    m_value = plaq;
}

// ------------------------------------------------------------ Wilson flow ---

GpuWilsonFlow::GpuWilsonFlow(double tauFinal, double epsilon, int precision, int measureEvery)
    : m_tauFinal(tauFinal), m_epsilon(epsilon), m_precision(precision), m_measureEvery(measureEvery)
{
}

void GpuWilsonFlow::createLattice(std::array<int, 4> latticeSize)
{
    App::createLattice(latticeSize);
    DeviceField::of(m_lat, m_precision);
}

void GpuWilsonFlow::execute()
{
    initialize();
    flowConfigurations();
}

void GpuWilsonFlow::initialize()
{
    for (auto *o : m_obs) o->initObservable(m_lat);
    if (m_act) m_act->initAction(m_lat);
    m_obsValues.assign(m_obs.size(), 0.0);  // src/Apps/wilsonflow.cpp:76
    m_inputConfList.clear();
    LatticeIO::InputConf::getInputList(m_inputDir, m_inputConfList);
}

void GpuWilsonFlow::computeObservables()
{
    for (auto *o : m_obs) o->compute();
}

void GpuWilsonFlow::flowConfigurations()
{
    if (m_batchFlow && m_inputConfList.size() > 1) {
        // device bytes per configuration: two link buffers, the RK accumulator,
        // the transfer staging of one direction
        const double V = double(m_size[0]) * m_size[1] * m_size[2] * m_size[3];
        const double real = m_precision == LYM_PREC_FP64 ? 8.0 : 4.0;
        const double per = V * (4 * (2 * 18 + 8) * real + 18 * 8.0);
        const size_t chunk = size_t(std::max(1.0, std::floor(m_batchBytes / per)));
        for (size_t first = 0; first < m_inputConfList.size(); first += chunk)
            flowBatch(first, std::min(m_inputConfList.size(), first + chunk));
        return;
    }
    int confNum = 0;
    for (auto &conf : m_inputConfList) {
        LatticeIO::InputConf::readConf(*m_lat, conf.c_str());
        DeviceField::of(m_lat, m_precision).hostModified();
        applyWilsonFlow(confNum++);
    }
}

// The input files [first, last) flowed together: WilsonFlow::flowConfigurations
// (src/Apps/wilsonflow.cpp:96-109) as one lym_batch_flow call, each
// configuration's series and output as applyWilsonFlow gives them.  The host
// field holds the last configuration's flowed links afterwards, as after the
// one-by-one loop.
void GpuWilsonFlow::flowBatch(size_t first, size_t last)
{
    const int n = int(last - first);
    std::vector<int> dims;
    for (int c = 0; c < n; c++) dims.insert(dims.end(), m_size.begin(), m_size.end());
    lym_batch *b = nullptr;
    int rc = lym_batch_create(&b, n, dims.data(), m_precision, 0);
    if (rc != LYM_OK) throw std::runtime_error("lym_batch_create failed with code " + std::to_string(rc));
    auto check = [&](int r, const char *what) {
        if (r == LYM_OK) return;
        const std::string m = std::string(what) + ": " + lym_batch_last_error(b);
        lym_batch_destroy(b);
        throw std::runtime_error(m);
    };
    for (int c = 0; c < n; c++) {
        LatticeIO::InputConf::readConf(*m_lat, m_inputConfList[first + size_t(c)].c_str());
        const double *p[4];
        for (int mu = 0; mu < 4; mu++) p[mu] = (*m_lat)[mu].lattice.data()->mat.data();
        check(lym_batch_upload_links(b, c, p), "lym_batch_upload_links");
    }
    if (m_autotune) {
        char summary[256] = {0};
        check(lym_batch_autotune(b, m_tuneCache.empty() ? nullptr : m_tuneCache.c_str(), summary, int(sizeof summary)),
              "lym_batch_autotune");
        if (first == 0) printf("batch flow schedule: %s\n", summary);
    }
    const int rows = std::max(1, int(m_tauFinal / m_epsilon));
    std::vector<std::array<double, 4>> all(size_t(n) * size_t(rows));
    int got = 0;
    check(lym_batch_flow(b, m_epsilon, m_tauFinal, m_measureEvery, 1, all[0].data(), rows, &got), "lym_batch_flow");
    double *p[4];
    for (int mu = 0; mu < 4; mu++) p[mu] = (*m_lat)[mu].lattice.data()->mat.data();
    check(lym_batch_download_links(b, n - 1, p), "lym_batch_download_links");
    lym_batch_destroy(b);
    DeviceField::of(m_lat, m_precision).hostModified();
    for (int c = 0; c < n; c++) {
        const auto it = all.begin() + std::ptrdiff_t(size_t(c) * size_t(rows));
        finishConfiguration(int(first) + c, std::vector<std::array<double, 4>>(it, it + got));
    }
}

void GpuWilsonFlow::tune(DeviceField &d)
{
    if (!m_autotune) return;
    const std::string s = d.autotune(m_tuneCache);
    if (!s.empty()) printf("flow schedule: %s\n", s.c_str());
}

std::vector<std::array<double, 4>> GpuWilsonFlow::applyWilsonFlow(int confNum)
{
    DeviceField &d = DeviceField::of(m_lat, m_precision);
    d.toDevice();
    tune(d);
    const int rows = int(m_tauFinal / m_epsilon);
    std::vector<std::array<double, 4>> series(size_t(std::max(rows, 1)));
    int got = 0;
    d.check(lym_flow(d.ctx(), m_epsilon, m_tauFinal, m_measureEvery, series[0].data(), int(series.size()), &got),
            "lym_flow");
    d.deviceModified();
    series.resize(size_t(got));
    // Report a field that left SU(3) or holds NaN (the reference would carry
    // on silently); the run continues.
    const DeviceField::Health h = d.health();
    const double allowed = m_precision == LYM_PREC_FP64 ? 1e-10 : 1e-5;
    if (h.nonFinite > 0 || h.maxDeviation > allowed)
        fprintf(stderr, "warning: conf %d after the flow: %lld links non-finite, max deviation from SU(3) %.3g\n",
                confNum, (long long)h.nonFinite, h.maxDeviation);
    finishConfiguration(confNum, series);
    return series;
}

void GpuWilsonFlow::finishConfiguration(int confNum, const std::vector<std::array<double, 4>> &series)
{
    // gradient-flow scales of this configuration, in lattice and physical units
    const FlowScales sc = flowScales(series, long(LatticeUnits::latticeSites));
    if (LatticeUnits::beta > 0)
        printf("conf %d: t0/a^2 = %.8g, w0/a = %.8g (a = %.5g fm at beta %g: t0 = %.5g fm^2, w0 = %.5g fm)\n", confNum,
               sc.t0, sc.w0, LatticeUnits::latticeSpacing, LatticeUnits::beta,
               sc.t0 * LatticeUnits::latticeSpacing * LatticeUnits::latticeSpacing, sc.w0 * LatticeUnits::latticeSpacing);
    else
        printf("conf %d: t0/a^2 = %.8g, w0/a = %.8g\n", confNum, sc.t0, sc.w0);
    // the observables take the values of the last measured row (rows between
    // measurements hold NaN when measureEvery > 1)
    for (auto r = series.rbegin(); r != series.rend(); ++r) {
        if (!std::isfinite((*r)[1])) continue;
        for (auto *o : m_obs) o->setResults((*r)[1], (*r)[2], (*r)[3]);
        break;
    }
    m_obsValues.resize(m_obs.size());
    for (size_t i = 0; i < m_obs.size(); i++) m_obsValues[i] = m_obs[i]->value();
    char name[64];
    snprintf(name, sizeof name, "conf%04d.dat", confNum);
    std::filesystem::create_directories(std::filesystem::path(m_outputDir) / "flow");
    LatticeIO::OutputObs::writeFlowObs((std::filesystem::path(m_outputDir) / "flow" / name).string(), series);
}

void GpuWilsonFlow::flowStep()
{
    DeviceField &d = DeviceField::of(m_lat, m_precision);
    d.toDevice();
    tune(d);
    d.check(lym_flow_steps(d.ctx(), m_epsilon, 1), "lym_flow_steps");
    d.deviceModified();
}

// ------------------------------------------------------------- Metropolis ---

GpuGaugeFieldFactory::GpuGaugeFieldFactory(int MCSteps, int thermSteps, int NConf, double epsilon,
                                           std::string startType, double beta, uint64_t seed, int nhit, int candidate)
    : m_MCSteps(MCSteps), m_thermSteps(thermSteps), m_confs(NConf),
      m_correlationSteps(NConf > 0 ? std::max(1, MCSteps / NConf) : std::max(1, MCSteps)), m_epsilon(epsilon),
      m_beta(beta), m_startType(std::move(startType)), m_seed(seed), m_nhit(nhit), m_candidate(candidate)
{
}

void GpuGaugeFieldFactory::execute()
{
    initialize();
    thermalize();
    sampleConf();
}

void GpuGaugeFieldFactory::initialize()
{
    DeviceField &d = DeviceField::of(m_lat, LYM_PREC_FP64);
    if (m_startType == "Hot")
        d.check(lym_set_hot(d.ctx(), m_seed), "lym_set_hot");
    else
        d.check(lym_set_cold(d.ctx()), "lym_set_cold");
    d.deviceModified();
    for (auto *o : m_obs) o->initObservable(m_lat);
    if (m_act) m_act->initAction(m_lat);
    const bool ref = m_nhit == 30 && m_candidate == LYM_CAND_SU2;
    printf("Metropolis: %d hits per link, %s candidates, eps %g, checkerboard sweeps, Philox seed %llu%s\n", m_nhit,
           m_candidate == LYM_CAND_SU2 ? "R*S*T SU(2)-product" : "exp(i eps H)", m_epsilon,
           (unsigned long long)m_seed,
           ref ? " (the reference's hits and candidates)"
               : " (the reference's updateLink takes 30 hits of R*S*T candidates: a different chain)");
}

void GpuGaugeFieldFactory::MCUpdate()
{
    DeviceField &d = DeviceField::of(m_lat, LYM_PREC_FP64);
    d.toDevice();
    int64_t acc = 0;
    d.check(lym_metropolis(d.ctx(), m_beta, m_epsilon, m_nhit, 1, m_sweep++, m_candidate, m_seed, nullptr, &acc),
            "lym_metropolis");
    d.deviceModified();
    m_accepted += acc;
    m_updates += int64_t(4) * m_size[0] * m_size[1] * m_size[2] * m_size[3] * m_nhit;
}

void GpuGaugeFieldFactory::computeObservables()
{
    for (auto *o : m_obs) o->compute();
}

// the terminal report of OutputTerm::printThermStep / printGenStep (src/InputOutput/outputterm.cpp:29-51)
static void printStep(const char *what, int n, std::vector<Observable *> &obs, double acceptance)
{
    printf("%s %4i\n", what, n);
    for (auto *o : obs) printf("\t%s\t = %.12g \n", o->getName(), o->value());
    printf("\tAccept. Ratio = %lf\n", acceptance);
}

void GpuGaugeFieldFactory::thermalize()
{
    for (int step = 1; step <= m_thermSteps; step++) {
        MCUpdate();
        if (step % m_correlationSteps == 0) {  // src/Apps/gaugefieldfactory.cpp:98-106
            computeObservables();
            printStep("Thermalization Step", step, m_obs, acceptance());
        }
    }
}

void GpuGaugeFieldFactory::sampleConf()
{
    int confNum = 0;
    std::filesystem::create_directories(m_outputDir);
    LatticeIO::OutputObs::initialize(m_outputDir, m_obs);
    for (int step = 1; step <= m_MCSteps; step++) {
        MCUpdate();
        if (step % m_correlationSteps == 0) {  // src/Apps/gaugefieldfactory.cpp:112-128
            computeObservables();
            LatticeIO::OutputObs::writeObs(m_outputDir, m_obs, step);
            printStep("Configuration Number", confNum, m_obs, acceptance());
            DeviceField::of(m_lat, LYM_PREC_FP64).toHost();
            char name[64];
            snprintf(name, sizeof name, "conf%04d.bin", confNum++);
            LatticeIO::OutputConf::writeConf(*m_lat, (std::filesystem::path(m_outputDir) / name).string().c_str());
        }
    }
}
