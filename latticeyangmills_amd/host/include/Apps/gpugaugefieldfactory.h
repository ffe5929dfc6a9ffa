// Metropolis ensemble generation on the GPU: the drop-in for GaugeFieldFactory
// (include/Apps/gaugefieldfactory.h, src/Apps/gaugefieldfactory.cpp), with a
// checkerboard sweep and Philox random numbers.  Like the reference it prints
// the observables and the acceptance every NCorr sweeps of the
// thermalisation and of the sampling, and writes observables.dat (a "Step"
// column and one column per observable's value()) and conf%04d.bin into the
// output directory while sampling.  The reference's updateLink takes 30 hits
// of R S T candidates (src/Apps/gaugefieldfactory.cpp:210-213): nhit = 30 with
// LYM_CAND_SU2 reproduces that choice; the defaults (10 hits, exp(i eps H))
// follow the benchmark configurations.
#pragma once
#include <cstdint>
#include <string>

#include "Apps/app.h"
#include "Math/devicefield.h"

class GpuGaugeFieldFactory : public App {
public:
    GpuGaugeFieldFactory(int MCSteps, int thermSteps, int NConf, double epsilon, std::string startType,
                         double beta, uint64_t seed, int nhit = 10, int candidate = LYM_CAND_EXPH);
    void execute() override;  // initialize, thermalize, sampleConf
    void initialize();
    void thermalize();
    void sampleConf();
    void MCUpdate();  // one checkerboard sweep of every link
    void setOutputDir(const std::string &d) { m_outputDir = d; }
    double acceptance() const { return m_updates ? double(m_accepted) / double(m_updates) : 0.0; }
    int hits() const { return m_nhit; }
    int candidate() const { return m_candidate; }

private:
    void computeObservables();
    int m_MCSteps, m_thermSteps, m_confs, m_correlationSteps;
    double m_epsilon, m_beta;
    std::string m_startType, m_outputDir = ".";
    uint64_t m_seed;
    int m_nhit, m_candidate, m_sweep = 0;
    int64_t m_accepted = 0, m_updates = 0;
};
