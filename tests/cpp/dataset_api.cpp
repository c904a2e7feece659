// Drives one std or VE hydro pass through the C++ interface (include/sphexa_amd.hpp) the way the
// reference propagators call the hydro functions (main/src/propagator/std_hydro.hpp:124-168,
// main/src/propagator/ve_hydro.hpp:130-204), reading the initial conditions from a binary file
// and writing the SFC-ordered results for comparison with the CPU oracle.
//
//   dataset_api <in.bin> <out.bin> std|ve|std32|ve32 [cpu|gpu|bdt] [steps]
//
// "std32"/"ve32" run the mixed form: a DeviceDataset<float> whose HydroType fields are float and
// whose x y z u temp du stay double, the reference's default types (sph/include/sph/types.hpp:40-45:
// CoordinateType double, HydroType float).
// "cpu" (the default) calls findNeighborsSfc and then computeDensity / computeXMass, as the
// reference's CPU build does; "gpu" calls the fused findNeighborsDensity / findNeighborsXMass,
// which is how the reference's GPU build runs the search (inside computeDensity / computeXMass).
// "bdt" (VE only) runs the "cpu" pass and then the block time-step calls of HydroVeBdtProp
// (main/src/propagator/ve_hydro_bdt.hpp:258-291, 337-360) over consecutive groups of 64:
// groupDivvTimestep and groupAccTimestep from groupDt = 1e10, storeRung (rung 1), computePositions
// with the rungs (dt 2e-4, dt_m1 per rung 1e-4, 2e-4, 4e-4, 8e-4) and driftPositions (forward 3e-4,
// back 1e-4), and writes out.bin: the groups' dt (floats widened to double), n uint8 rungs, then n
// doubles each of x y z vx vy vz temp.
// With `steps` the driver runs that many full std steps, forces then the integrate phase of
// HydroProp (main/src/propagator/std_hydro.hpp:170-181: computeTimestep, computePositions,
// updateSmoothingLength) from minDt = minDt_m1 = 1e-4, and writes instead
// out.bin: n doubles each of x y z vx vy vz h u, then minDt, minDt_m1, ttot, and the 10 conserved
// quantities of computeConservedQuantities.
//
// in.bin : uint64 n, uint32 periodic, double lo, double hi, then n doubles each of
//          x y z h m u vx vy vz alpha (u is the temperature for the VE pass)
// out.bin: n doubles each of x, then the pass outputs (std: rho p c c11..c33 ax ay az du;
//          VE: xm kx gradh prho c c11..c33 divv curlv alpha ax ay az du), n uint32 nc, double minDt;
//          float fields are widened to double on the host
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "sphexa_amd.hpp"

using namespace sphexa_amd;

#define HIPCHECK(x)                                                                                                    \
    do                                                                                                                 \
    {                                                                                                                  \
        if ((x) != hipSuccess)                                                                                         \
        {                                                                                                              \
            std::fprintf(stderr, "HIP error %s at %d\n", #x, __LINE__);                                                 \
            return 2;                                                                                                  \
        }                                                                                                              \
    } while (0)

namespace
{

//! device buffers of both element types, freed at exit
struct Owned
{
    std::vector<void*> ptrs;
    ~Owned()
    {
        for (void* p : ptrs) (void)hipFree(p);
    }
    template<class T>
    void operator()(T*& p, size_t n)
    {
        if (hipMalloc(&p, n * sizeof(T)) != hipSuccess) throw std::runtime_error("hipMalloc");
        if (hipMemset(p, 0, n * sizeof(T)) != hipSuccess) throw std::runtime_error("hipMemset");
        ptrs.push_back(p);
    }
};

template<class T>
bool upload(T* dst, const std::vector<double>& src)
{
    std::vector<T> buf(src.begin(), src.end());
    return hipMemcpy(dst, buf.data(), buf.size() * sizeof(T), hipMemcpyHostToDevice) == hipSuccess;
}

template<class T>
bool download(FILE* out, const T* src, size_t n)
{
    std::vector<T> buf(n);
    if (hipMemcpy(buf.data(), src, n * sizeof(T), hipMemcpyDeviceToHost) != hipSuccess) return false;
    std::vector<double> wide(buf.begin(), buf.end());
    return std::fwrite(wide.data(), 8, n, out) == n;
}

template<class H>
int run(const std::vector<std::vector<double>>& ic, size_t n, bool periodic, double lo, double hi, bool ve,
        bool fused, bool bdt, int steps, const char* outPath)
{
    DeviceDataset<H> d;
    if (sphx_workspace_create(&d.workspace, 0, nullptr) != SPHX_OK) { return 2; }
    if (sphx_workspace_set_kernel(d.workspace, 6.0, &d.K) != SPHX_OK) { return 2; }
    // the propagators change no field between the hydro calls of a pass, so the pair loops may
    // write the next loop's neighbour records (bitwise the same results)
    d.sphxFlags = SPHX_FLAG_CHAIN;
    d.minDt     = steps > 0 ? 1e-4 : 1e-3;
    d.minDt_m1  = d.minDt;

    Owned owned;
    auto& s = d.devData;
    try
    {
        for (double** p : {&s.x, &s.y, &s.z, &s.du}) owned(*p, n);
        for (H** p : {&s.h, &s.m, &s.vx, &s.vy, &s.vz, &s.c, &s.c11, &s.c12, &s.c13, &s.c22, &s.c23, &s.c33, &s.ax,
                      &s.ay, &s.az})
            owned(*p, n);
        if (ve)
        {
            owned(s.temp, n);
            for (H** p : {&s.prho, &s.xm, &s.kx, &s.gradh, &s.divv, &s.curlv, &s.alpha}) owned(*p, n);
        }
        else
        {
            owned(s.u, n);
            for (H** p : {&s.rho, &s.p}) owned(*p, n);
        }
        if (steps > 0 || bdt)
            for (H** p : {&s.x_m1, &s.y_m1, &s.z_m1, &s.du_m1}) owned(*p, n);
        owned(s.keys, n);
        owned(s.nc, n);
    }
    catch (const std::exception& e)
    {
        std::fprintf(stderr, "%s\n", e.what());
        return 2;
    }
    bool ok = upload(s.x, ic[0]) && upload(s.y, ic[1]) && upload(s.z, ic[2]) && upload(s.h, ic[3]) &&
              upload(s.m, ic[4]) && upload(ve ? s.temp : s.u, ic[5]) && upload(s.vx, ic[6]) &&
              upload(s.vy, ic[7]) && upload(s.vz, ic[8]) && (!ve || upload(s.alpha, ic[9]));
    if (!ok) { return 2; }

    Box       box0(lo, hi, periodic ? BoundaryType::periodic : BoundaryType::open);
    Domain    domain(0, 1, 64, 64, 0.5f, box0);
    GroupView grp{0, uint32_t(n), 0, nullptr, nullptr};
    const uint32_t numGroups = uint32_t((n + 63) / 64);
    float*         groupDt   = nullptr;
    uint8_t*       rungs     = nullptr;
    try
    {
        if (bdt)
        {
            owned(groupDt, numGroups);
            owned(rungs, n);
        }
        for (int step = 0; step < std::max(steps, 1); ++step)
        {
            // the pass alone keeps the given box (the CPU oracle's); steps refit it as Domain::sync does
            if (steps > 0) { domain.sync(n, d); }
            else { syncDomain(n, d, box0); }
            const Box& box = steps > 0 ? domain.box() : box0;
            if (!fused) { findNeighborsSfc(0, n, d, box); }
            if (ve)
            {
                if (fused) { findNeighborsXMass(grp, d, box); }
                else { computeXMass(grp, d, box); }
                computeVeDefGradh(grp, d, box);
                computeEOS(0, n, d);
                computeIadDivvCurlv(grp, d, box);
                computeAVswitches(grp, d, box);
                computeMomentumEnergy<false>(grp, nullptr, d, box);
            }
            else
            {
                if (fused) { findNeighborsDensity(grp, d, box); }
                else { computeDensity(grp, d, box); }
                computeEOS_HydroStd(0, n, d);
                computeIAD(grp, d, box);
                computeMomentumEnergySTD(grp, d, box);
            }
            if (bdt)
            {
                GroupView all{0, uint32_t(n), numGroups, nullptr, nullptr};
                std::vector<float> big(numGroups, 1e10f);
                if (hipMemcpy(groupDt, big.data(), numGroups * 4, hipMemcpyHostToDevice) != hipSuccess)
                    throw std::runtime_error("hipMemcpy");
                groupDivvTimestep(all, groupDt, d);
                groupAccTimestep(all, groupDt, d);
                storeRung(all, 1, rungs, d);
                std::array<float, maxNumRungs> dtM1{1e-4f, 2e-4f, 4e-4f, 8e-4f};
                computePositions(all, d, box, 2e-4f, dtM1, rungs);
                driftPositions(all, d, 3e-4f, 1e-4f, dtM1, rungs);
            }
            if (steps > 0)
            {
                computeTimestep(0, n, d);
                computePositions(grp, d, box, float(d.minDt), {float(d.minDt_m1)});
                updateSmoothingLength(grp, d);
            }
        }
    }
    catch (const std::exception& e)
    {
        std::fprintf(stderr, "%s\n", e.what());
        return 3;
    }

    FILE* out = std::fopen(outPath, "wb");
    if (!out) { return 1; }
    if (bdt)
    {
        ok = download(out, groupDt, numGroups);
        std::vector<uint8_t> r(n);
        ok = ok && hipMemcpy(r.data(), rungs, n, hipMemcpyDeviceToHost) == hipSuccess;
        std::fwrite(r.data(), 1, n, out);
        for (double* p : {s.x, s.y, s.z}) ok = ok && download(out, p, n);
        for (H* p : {s.vx, s.vy, s.vz}) ok = ok && download(out, p, n);
        ok = ok && download(out, s.temp, n);
    }
    else if (steps > 0)
    {
        for (double* p : {s.x, s.y, s.z}) ok = ok && download(out, p, n);
        for (H* p : {s.vx, s.vy, s.vz, s.h}) ok = ok && download(out, p, n);
        ok = ok && download(out, s.u, n);
        for (double v : {d.minDt, d.minDt_m1, d.ttot}) std::fwrite(&v, 8, 1, out);
        auto q = computeConservedQuantities(0, n, d);
        std::fwrite(q.data(), 8, q.size(), out);
    }
    else
    {
        ok = download(out, s.x, n);
        std::vector<H*> outs = ve ? std::vector<H*>{s.xm, s.kx, s.gradh, s.prho, s.c, s.c11, s.c12, s.c13, s.c22,
                                                    s.c23, s.c33, s.divv, s.curlv, s.alpha, s.ax, s.ay, s.az}
                                  : std::vector<H*>{s.rho, s.p, s.c, s.c11, s.c12, s.c13, s.c22, s.c23, s.c33, s.ax,
                                                    s.ay, s.az};
        for (H* p : outs) ok = ok && download(out, p, n);
        ok = ok && download(out, s.du, n);
        std::vector<uint32_t> nc(n);
        ok = ok && hipMemcpy(nc.data(), s.nc, n * 4, hipMemcpyDeviceToHost) == hipSuccess;
        std::fwrite(nc.data(), 4, n, out);
        std::fwrite(&d.minDtCourant, 8, 1, out);
    }
    std::fclose(out);
    sphx_workspace_destroy(d.workspace);
    return ok ? 0 : 2;
}

} // namespace

int main(int argc, char** argv)
{
    if (argc < 4 || argc > 6) { return 1; }
    std::string form  = argv[3];
    bool        mixed = form == "std32" || form == "ve32";
    if (mixed) { form.resize(form.size() - 2); }
    if (form != "std" && form != "ve") { return 1; }
    bool fused = argc >= 5 && std::string(argv[4]) == "gpu";
    bool bdt   = argc >= 5 && std::string(argv[4]) == "bdt";
    int  steps = argc == 6 ? std::atoi(argv[5]) : 0;
    if (steps > 0 && form != "std") { return 1; }
    if (bdt && (form != "ve" || steps > 0)) { return 1; }
    FILE* in = std::fopen(argv[1], "rb");
    if (!in) { return 1; }
    uint64_t n;
    uint32_t periodic;
    double   lo, hi;
    if (std::fread(&n, 8, 1, in) != 1 || std::fread(&periodic, 4, 1, in) != 1 || std::fread(&lo, 8, 1, in) != 1 ||
        std::fread(&hi, 8, 1, in) != 1)
        return 1;
    std::vector<std::vector<double>> ic(10, std::vector<double>(n));
    for (auto& v : ic)
        if (std::fread(v.data(), 8, n, in) != n) return 1;
    std::fclose(in);

    bool ve = form == "ve";
    return mixed ? run<float>(ic, n, periodic != 0, lo, hi, ve, fused, bdt, steps, argv[2])
                 : run<double>(ic, n, periodic != 0, lo, hi, ve, fused, bdt, steps, argv[2]);
}
