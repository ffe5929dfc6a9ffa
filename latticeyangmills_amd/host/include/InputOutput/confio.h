// Binary gauge configurations in the reference's file format
// (InputConf::readConf, src/InputOutput/inputconf.cpp:114-155): per site in
// lexicographic order (x fastest), the four links mu = 0..3, each 18 native
// doubles; 576 bytes per site.  The reference's writer is stubbed
// (src/InputOutput/outputconf.cpp:102); OutputConf::writeConf writes the same layout.
#pragma once
#include <string>
#include <vector>

#include "Math/field.h"

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
    /// rows of [t, plaq, topc, energy] printed with %.17g (the reference uses %lf,
    /// src/InputOutput/outputobs.cpp:64-84)
    static void writeFlowObs(const std::string &path, const std::vector<std::array<double, 4>> &rows);
};
}  // namespace LatticeIO
