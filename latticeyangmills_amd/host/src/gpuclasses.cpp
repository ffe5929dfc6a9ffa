// GPU Action / Observable / App classes: each member replaces the reference's
// host loop of the same name by a call through include/lym.h.
#include <cmath>
#include <cstdio>
#include <filesystem>

#include "Actions/gpupuregauge.h"
#include "Apps/gpugaugefieldfactory.h"
#include "Apps/gpuwilsonflow.h"
#include "InputOutput/confio.h"
#include "Observables/flowscale.h"
#include "Observables/gpusuperobs.h"
#include "Utils/latticeunits.h"

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
    m_inputConfList.clear();
    LatticeIO::InputConf::getInputList(m_inputDir, m_inputConfList);
}

void GpuWilsonFlow::computeObservables()
{
    for (auto *o : m_obs) o->compute();
}

void GpuWilsonFlow::flowConfigurations()
{
    int confNum = 0;
    for (auto &conf : m_inputConfList) {
        LatticeIO::InputConf::readConf(*m_lat, conf.c_str());
        DeviceField::of(m_lat, m_precision).hostModified();
        applyWilsonFlow(confNum++);
    }
}

std::vector<std::array<double, 4>> GpuWilsonFlow::applyWilsonFlow(int confNum)
{
    DeviceField &d = DeviceField::of(m_lat, m_precision);
    d.toDevice();
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
    // gradient-flow scales of this configuration, in lattice and physical units
    const FlowScales sc = flowScales(series, long(LatticeUnits::latticeSites));
    if (LatticeUnits::beta > 0)
        printf("conf %d: t0/a^2 = %.8g, w0/a = %.8g (a = %.5g fm at beta %g: t0 = %.5g fm^2, w0 = %.5g fm)\n", confNum,
               sc.t0, sc.w0, LatticeUnits::latticeSpacing, LatticeUnits::beta,
               sc.t0 * LatticeUnits::latticeSpacing * LatticeUnits::latticeSpacing, sc.w0 * LatticeUnits::latticeSpacing);
    else
        printf("conf %d: t0/a^2 = %.8g, w0/a = %.8g\n", confNum, sc.t0, sc.w0);
    if (!series.empty()) {
        // the observables hold the values at the last measured row
        for (auto *o : m_obs) {
            o->plaq = series.back()[1];
            o->topc = series.back()[2];
            o->energy = series.back()[3];
        }
    }
    char name[64];
    snprintf(name, sizeof name, "conf%04d.dat", confNum);
    std::filesystem::create_directories(std::filesystem::path(m_outputDir) / "flow");
    LatticeIO::OutputObs::writeFlowObs((std::filesystem::path(m_outputDir) / "flow" / name).string(), series);
    return series;
}

void GpuWilsonFlow::flowStep()
{
    DeviceField &d = DeviceField::of(m_lat, m_precision);
    d.toDevice();
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

void GpuGaugeFieldFactory::thermalize()
{
    for (int step = 1; step <= m_thermSteps; step++) MCUpdate();
}

void GpuGaugeFieldFactory::sampleConf()
{
    int confNum = 0;
    std::filesystem::create_directories(m_outputDir);
    for (int step = 1; step <= m_MCSteps; step++) {
        MCUpdate();
        if (step % m_correlationSteps == 0) {
            computeObservables();
            DeviceField::of(m_lat, LYM_PREC_FP64).toHost();
            char name[64];
            snprintf(name, sizeof name, "conf%04d.bin", confNum++);
            LatticeIO::OutputConf::writeConf(*m_lat, (std::filesystem::path(m_outputDir) / name).string().c_str());
        }
    }
}
