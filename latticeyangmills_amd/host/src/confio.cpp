// Configuration and series files (see InputOutput/confio.h).
#include "InputOutput/confio.h"

#include <algorithm>
#include <cstdio>
#include <filesystem>
#include <stdexcept>

namespace LatticeIO {

// Sites per fread / fwrite: the file holds, for each site in lexicographic
// order, the 4 x 18 doubles of its links (direction-major), so blocks of
// whole sites keep the layout while moving 4.7 MB per call instead of 576 B.
constexpr int kIoSites = 8192;

void InputConf::readConf(GluonField &lattice, const char *inputFile)
{
    FILE *f = fopen(inputFile, "rb");
    if (!f) throw std::runtime_error(std::string("cannot open ") + inputFile);
    const int V = lattice[0].sites;
    std::vector<double> block(size_t(kIoSites) * 4 * 18);  // whole sites, 72 doubles each
    for (int s0 = 0; s0 < V; s0 += kIoSites) {
        const int n = std::min(kIoSites, V - s0);
        if (fread(block.data(), sizeof(double) * 4 * 18, size_t(n), f) != size_t(n)) {
            fclose(f);
            throw std::runtime_error(std::string("short configuration file ") + inputFile);
        }
        for (int i = 0; i < n; i++)
            for (int mu = 0; mu < 4; mu++) {
                const double *src = &block[(size_t(i) * 4 + mu) * 18];
                std::copy(src, src + 18, lattice[mu].at(s0 + i).mat.begin());
            }
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
    std::vector<double> block(size_t(kIoSites) * 4 * 18);
    for (int s0 = 0; s0 < V; s0 += kIoSites) {
        const int n = std::min(kIoSites, V - s0);
        for (int i = 0; i < n; i++)
            for (int mu = 0; mu < 4; mu++)
                std::copy(lattice[mu].at(s0 + i).mat.begin(), lattice[mu].at(s0 + i).mat.end(),
                          &block[(size_t(i) * 4 + mu) * 18]);
        if (fwrite(block.data(), sizeof(double) * 4 * 18, size_t(n), f) != size_t(n)) {
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
    fprintf(f, "Flow Time\tPlaq.\tTopc.\tEnerg.\t\n");
    for (auto &r : rows) fprintf(f, "%.17g\t%.17g\t%.17g\t%.17g\t\n", r[0], r[1], r[2], r[3]);
    fclose(f);
}

static std::string obsFile(const std::string &dir) { return (std::filesystem::path(dir) / "observables.dat").string(); }

void OutputObs::initialize(const std::string &outputDir, std::vector<Observable *> &obsList)
{
    FILE *f = fopen(obsFile(outputDir).c_str(), "w");
    if (!f) throw std::runtime_error("cannot create " + obsFile(outputDir));
    fprintf(f, "Step\t");
    for (auto *o : obsList) fprintf(f, "%s\t", o->getName());
    fprintf(f, "\n");
    fclose(f);
}

void OutputObs::writeObs(const std::string &outputDir, std::vector<Observable *> &obsList, int MCSteps)
{
    FILE *f = fopen(obsFile(outputDir).c_str(), "a");
    if (!f) throw std::runtime_error("cannot append to " + obsFile(outputDir));
    fprintf(f, "%i\t", MCSteps);
    for (auto *o : obsList) fprintf(f, "%.17g\t", o->value());
    fprintf(f, "\n");
    fclose(f);
}

}  // namespace LatticeIO
