// Configuration and series files (see InputOutput/confio.h).
#include "InputOutput/confio.h"

#include <algorithm>
#include <cstdio>
#include <filesystem>
#include <stdexcept>

namespace LatticeIO {

void InputConf::readConf(GluonField &lattice, const char *inputFile)
{
    FILE *f = fopen(inputFile, "rb");
    if (!f) throw std::runtime_error(std::string("cannot open ") + inputFile);
    const int V = lattice[0].sites;
    std::vector<double> site(4 * 18);
    for (int s = 0; s < V; s++) {
        if (fread(site.data(), sizeof(double), site.size(), f) != site.size()) {
            fclose(f);
            throw std::runtime_error(std::string("short configuration file ") + inputFile);
        }
        for (int mu = 0; mu < 4; mu++) std::copy(&site[18 * mu], &site[18 * mu] + 18, lattice[mu].at(s).mat.begin());
    }
    fclose(f);
}

void InputConf::getInputList(const std::string &dir, std::vector<std::string> &list)
{
    for (auto &e : std::filesystem::directory_iterator(dir))
        if (e.path().extension() == ".bin") list.push_back(e.path().string());
    std::sort(list.begin(), list.end());
}

void OutputConf::writeConf(GluonField &lattice, const char *outputFile)
{
    FILE *f = fopen(outputFile, "wb");
    if (!f) throw std::runtime_error(std::string("cannot create ") + outputFile);
    const int V = lattice[0].sites;
    std::vector<double> site(4 * 18);
    for (int s = 0; s < V; s++) {
        for (int mu = 0; mu < 4; mu++) std::copy(lattice[mu].at(s).mat.begin(), lattice[mu].at(s).mat.end(), &site[18 * mu]);
        if (fwrite(site.data(), sizeof(double), site.size(), f) != site.size()) {
            fclose(f);
            throw std::runtime_error(std::string("write failed: ") + outputFile);
        }
    }
    fclose(f);
}

void OutputObs::writeFlowObs(const std::string &path, const std::vector<std::array<double, 4>> &rows)
{
    FILE *f = fopen(path.c_str(), "w");
    if (!f) throw std::runtime_error("cannot create " + path);
    for (auto &r : rows) fprintf(f, "%.17g %.17g %.17g %.17g\n", r[0], r[1], r[2], r[3]);
    fclose(f);
}

}  // namespace LatticeIO
