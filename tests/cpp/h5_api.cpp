// Writes a Dataset's attributes and fields through the C++ I/O interface (include/sphexa_amd_io.hpp)
// the way the reference's checkpoint output does (ParticlesData::loadOrStoreAttributes, then
// writeField per output field, main/src/io/ifile_io_hdf5.cpp), then reads them back with
// H5PartReader into a second Dataset and checks every value.
//
//   h5_api <path>        (host fields; prints "ok" and returns 0 on success)
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "sphexa_amd.hpp"
#include "sphexa_amd_io.hpp"

using namespace sphexa_amd;

int main(int argc, char** argv)
{
    if (argc != 2) { return 1; }
    const std::string path = argv[1];
    const size_t      n    = 1000;
    std::vector<double> x(n), temp(n);
    std::vector<float>  h(n);
    std::vector<unsigned> nc(n);
    for (size_t i = 0; i < n; ++i)
    {
        x[i]    = std::sin(double(i));
        temp[i] = 100.0 + i;
        h[i]    = 0.01f + 1e-5f * float(i);
        nc[i]   = unsigned(i % 150);
    }

    DeviceDataset<float> d;
    d.iteration = 42, d.numParticlesGlobal = n, d.ttot = 0.125, d.minDt = 1e-4, d.minDt_m1 = 2e-4, d.g = 0;
    d.ng0 = 80, d.ngmax = 120, d.gamma = 1.4, d.eosChoice = SPHX_EOS_POLYTROPIC, d.soundSpeedConst = 2.5f;
    d.alphamax = 1.5, d.sincIndex = 5.0;
    try
    {
        for (int step = 0; step < 2; ++step)
        {
            H5PartWriter w;
            w.addStep(0, n, path);
            double kcour = 0.3;
            w.fileAttribute("Kcour_file", &kcour, 1);
            loadOrStoreAttributes(&w, d);
            w.writeField("x", x.data());
            w.writeField("temp", temp.data());
            w.writeField("h", h.data());
            w.writeField("nc", nc.data());
            w.closeStep();
            d.iteration++;
        }

        DeviceDataset<float> e;
        H5PartReader         r;
        r.setStep(path, -1);
        loadOrStoreAttributes(&r, e);
        bool ok = r.globalNumParticles() == n && e.iteration == 43 && e.numParticlesGlobal == n &&
                  e.ttot == 0.125 && e.minDt == 1e-4 && e.minDt_m1 == 2e-4 && e.ng0 == 80 && e.ngmax == 120 &&
                  e.gamma == 1.4 && e.eosChoice == SPHX_EOS_POLYTROPIC && e.soundSpeedConst == 2.5f &&
                  e.alphamax == 1.5 && e.sincIndex == 5.0 && r.stepAttributeSize("iteration") == 1;
        std::vector<double>   x2(n), temp2(n);
        std::vector<float>    h2(n);
        std::vector<unsigned> nc2(n);
        r.readField("x", x2.data());
        r.readField("temp", temp2.data());
        r.readField("h", h2.data());
        r.readField("nc", nc2.data());
        ok = ok && x2 == x && temp2 == temp && h2 == h && nc2 == nc;
        double kcour = 0;
        r.fileAttribute("Kcour_file", &kcour, 1);
        ok = ok && kcour == 0.3 && r.fileAttributes().size() == 1;
        // a missing attribute is std::out_of_range, as in the reference's reader
        bool threw = false;
        try
        {
            double v;
            r.stepAttribute("no_such_attribute", &v, 1);
        }
        catch (const std::out_of_range&)
        {
            threw = true;
        }
        ok = ok && threw;
        std::printf(ok ? "ok\n" : "mismatch\n");
        return ok ? 0 : 4;
    }
    catch (const std::exception& ex)
    {
        std::fprintf(stderr, "%s\n", ex.what());
        return 3;
    }
}
