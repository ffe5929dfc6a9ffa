// Binary gauge configurations in the reference's file format
// (InputConf::readConf, src/InputOutput/inputconf.cpp:114-155): per site in
// lexicographic order (x fastest), the four links mu = 0..3, each 18 native
// doubles; 576 bytes per site.  The reference's writer is stubbed
// (src/InputOutput/outputconf.cpp:102); OutputConf::writeConf writes the same layout.
#pragma once
#include <string>
#include <vector>

#include "Math/field.h"
#include "Observables/observable.h"

namespace LatticeIO {
struct InputConf {
    static void readConf(GluonField &lattice, const char *inputFile);
    /// *.bin files of a directory, sorted (InputConf::getInputList, inputconf.cpp:76-82)
    static void getInputList(const std::string &dir, std::vector<std::string> &list);
};
struct OutputConf {
    static void writeConf(GluonField &lattice, const char *outputFile);
};
struct OutputObs {
    /// the reference's header line "Flow Time\tPlaq.\tTopc.\tEnerg.\t", then rows of
    /// [t, plaq, topc, energy], tab-separated, printed with %.17g where the
    /// reference uses %lf (src/InputOutput/outputobs.cpp:64-84)
    static void writeFlowObs(const std::string &path, const std::vector<std::array<double, 4>> &rows);
    /// observables.dat of the ensemble generation (outputobs.cpp:30-55): a
    /// header "Step\t<name>\t..." and one row per measurement, the step and
    /// each observable's value() (%.17g, the reference %lf)
    static void initialize(const std::string &outputDir, std::vector<Observable *> &obsList);
    static void writeObs(const std::string &outputDir, std::vector<Observable *> &obsList, int MCSteps);
};
}  // namespace LatticeIO
