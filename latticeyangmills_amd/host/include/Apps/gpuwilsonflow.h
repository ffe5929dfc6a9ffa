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

private:
    void computeObservables();
    double m_tauFinal, m_epsilon;
    int m_precision, m_measureEvery;
    std::string m_inputDir = ".", m_outputDir = ".";
    std::vector<std::string> m_inputConfList;
};
