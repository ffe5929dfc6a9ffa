// Wilson flow App on the GPU: the drop-in for WilsonFlow
// (include/Apps/wilsonflow.h:53-84, src/Apps/wilsonflow.cpp).
#pragma once
#include <string>
#include <vector>

#include "Apps/app.h"
#include "Math/devicefield.h"

class GpuWilsonFlow : public App {
public:
    GpuWilsonFlow(double tauFinal, double epsilon, int precision = LYM_PREC_FP64, int measureEvery = 1);
    void createLattice(std::array<int, 4> latticeSize) override;
    void execute() override;  // initialize + flowConfigurations
    void initialize();
    void flowConfigurations();
    /// WilsonFlow::applyWilsonFlow (wilsonflow.cpp:115-136) on the current field;
    /// returns rows x [t, plaq, topc, energy] and writes flow/conf%04d.dat
    std::vector<std::array<double, 4>> applyWilsonFlow(int confNum);
    /// one RK3 step (WilsonFlow::flowStep, wilsonflow.cpp:141-172)
    void flowStep();

    void setInputDir(const std::string &d) { m_inputDir = d; }
    void setOutputDir(const std::string &d) { m_outputDir = d; }
    std::vector<std::string> &inputList() { return m_inputConfList; }

    // getters of the reference (include/Apps/wilsonflow.h:58-60); the values
    // are those of the observables at the last measured row of the last flow
    std::array<int, 4> &getSize() { return m_size; }
    std::vector<double> &getObsValues() { return m_obsValues; }
    std::vector<Observable *> &getObs() { return m_obs; }

    /// on (default): flowConfigurations flows the input files in batches, all
    /// of a batch in one lym_batch_flow call (include/lym_batch.h), up to
    /// batchBytes of device memory per batch; off: one after another through
    /// the single-lattice context, as the reference's loop
    void setBatchFlow(bool on, double batchBytes = 32.0 * (1ull << 30))
    {
        m_batchFlow = on;
        m_batchBytes = batchBytes;
    }

    /// on (default): before its first flow the App picks the flow schedule for
    /// the lattice shape by timing the candidates (lym_autotune /
    /// lym_batch_autotune; no result changes), keeping the choice in
    /// cachePath when one is given; off: the library's default schedule
    void setAutotune(bool on, const std::string &cachePath = "")
    {
        m_autotune = on;
        m_tuneCache = cachePath;
    }

private:
    void computeObservables();
    void tune(DeviceField &d);
    void flowBatch(size_t first, size_t last);
    /// output of one flowed configuration: scales, observables, flow/conf%04d.dat
    void finishConfiguration(int confNum, const std::vector<std::array<double, 4>> &series);
    double m_tauFinal, m_epsilon;
    int m_precision, m_measureEvery;
    bool m_batchFlow = true, m_autotune = true;
    std::string m_tuneCache;
    double m_batchBytes = 32.0 * (1ull << 30);
    std::string m_inputDir = ".", m_outputDir = ".";
    std::vector<std::string> m_inputConfList;
    std::vector<double> m_obsValues;
};
