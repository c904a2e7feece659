/*
 * Definitions of the reference's GPU plug-in templates (sph/include/sph/sph_gpu.hpp:11-98) on top of the
 * sphexa_amd C ABI, for sphexa::ParticlesData<cstone::GpuTag> (sph/include/sph/particles_data.hpp:63-391).
 *
 * A maintainer adds this one translation unit to the reference's sph library in place of its hydro .cu
 * files (hydro_std/*_gpu.cu, hydro_ve/*_gpu.cu, positions_gpu.cu, update_h_gpu.cu, ts_groups.cu) and
 * links libsphexa_amd.so. Nothing in the reference's headers changes: findNeighborsSfc stays a no-op on
 * the GPU (sph/include/sph/find_neighbors.hpp:50), and computeXMass / computeDensity, the first pass after
 * it in both propagators (main/src/propagator/ve_hydro.hpp:147-152, std_hydro.hpp:139-143), build the
 * library's octree from the Domain's sorted keys, search with the h/nc iteration, then run the density
 * sum as a second kernel. Every later pair loop reads the stored lists.
 *
 * Compiled in this repository's CPU tests against the reference headers (tests/test_binding.py), which
 * also checks that every template this file instantiates is defined in the object.
 */
#include <stdexcept>
#include <string>
#include <tuple>

#include "sphexa_amd.h"

#include "cstone/sfc/box.hpp"
#include "cstone/traversal/groups.hpp"
#include "sph/particles_data.hpp"
#include "sph/sph_gpu.hpp"

namespace sph
{
namespace
{
using Data = sphexa::ParticlesData<cstone::GpuTag>;
using Box  = cstone::Box<double>;

//! one workspace per process (device current at first use); owns lists, records, tree and sort scratch
sphx_workspace_t* theWorkspace()
{
    static sphx_workspace_t* ws = nullptr;
    if (!ws)
    {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) { throw std::runtime_error("sphexa_amd: no current device"); }
        if (sphx_workspace_create(&ws, dev, nullptr) != SPHX_OK) { throw std::runtime_error("sphexa_amd: workspace"); }
    }
    return ws;
}

void check(sphx_workspace_t* ws, int code)
{
    if (code == SPHX_OK) { return; }
    if (code == SPHX_ERR_NOT_CONVERGED) { throw std::runtime_error("coupled nc/h-updated failed to converge"); }
    throw std::runtime_error(std::string("sphexa_amd: ") + sphx_last_error(ws));
}

//! ... with the kernel tables of the Dataset's sinc index, which the pair loops need
sphx_workspace_t* workspace(const Data& d)
{
    static double tableIndex = -1;
    auto*         ws         = theWorkspace();
    if (tableIndex != d.sincIndex)
    {
        double K = 0; // equals d.K (kernel_3D_k, 2000 Simpson intervals)
        check(ws, sphx_workspace_set_kernel(ws, d.sincIndex, &K));
        tableIndex = d.sincIndex;
    }
    return ws;
}

sphx_box_t box(const Box& b)
{
    auto bc = [](cstone::BoundaryType t) { return int32_t(t); }; // open 0, periodic 1, fixed 2
    return sphx_box_t{{b.xmin(), b.ymin(), b.zmin()},
                      {b.xmax(), b.ymax(), b.zmax()},
                      {bc(b.boundaryX()), bc(b.boundaryY()), bc(b.boundaryZ())}};
}

//! device pointers of the Dataset's fields; released fields are empty and pass NULL
sphx_fields_t fields(Data& d)
{
    auto&         s = d.devData;
    sphx_fields_t f{};
    f.x = rawPtr(s.x), f.y = rawPtr(s.y), f.z = rawPtr(s.z), f.u = rawPtr(s.u), f.temp = rawPtr(s.temp);
    f.du = rawPtr(s.du), f.h = rawPtr(s.h), f.m = rawPtr(s.m), f.vx = rawPtr(s.vx), f.vy = rawPtr(s.vy);
    f.vz = rawPtr(s.vz), f.rho = rawPtr(s.rho), f.p = rawPtr(s.p), f.c = rawPtr(s.c), f.prho = rawPtr(s.prho);
    f.tdpdTrho = rawPtr(s.tdpdTrho), f.c11 = rawPtr(s.c11), f.c12 = rawPtr(s.c12), f.c13 = rawPtr(s.c13);
    f.c22 = rawPtr(s.c22), f.c23 = rawPtr(s.c23), f.c33 = rawPtr(s.c33), f.xm = rawPtr(s.xm), f.kx = rawPtr(s.kx);
    f.gradh = rawPtr(s.gradh), f.divv = rawPtr(s.divv), f.curlv = rawPtr(s.curlv), f.alpha = rawPtr(s.alpha);
    f.ax = rawPtr(s.ax), f.ay = rawPtr(s.ay), f.az = rawPtr(s.az);
    f.dV11 = rawPtr(s.dV11), f.dV12 = rawPtr(s.dV12), f.dV13 = rawPtr(s.dV13);
    f.dV22 = rawPtr(s.dV22), f.dV23 = rawPtr(s.dV23), f.dV33 = rawPtr(s.dV33);
    f.keys = rawPtr(s.keys), f.nc = rawPtr(s.nc);
    return f;
}

sphx_params_t params(const Data& d)
{
    sphx_params_t p{};
    p.K = d.K, p.Kcour = d.Kcour, p.gamma = d.gamma, p.muiConst = d.muiConst, p.minDt = d.minDt;
    p.alphamin = d.alphamin, p.alphamax = d.alphamax, p.decay_constant = d.decay_constant;
    p.Atmin = d.Atmin, p.Atmax = d.Atmax, p.ramp = d.ramp, p.ng0 = d.ng0, p.ngmax = d.ngmax;
    p.hydro_type = sizeof(Data::HydroType) == 4 ? SPHX_F32 : SPHX_F64;
    p.eos = int32_t(d.eosChoice), p.soundSpeedConst = d.soundSpeedConst;
    p.polytropic_const = d.polytropic_const, p.polytropic_index = d.polytropic_index;
    return p;
}

//! the caller's GroupView (and per-group time steps) on the workspace while a call runs
struct Groups
{
    Groups(sphx_workspace_t* ws, const GroupView& g, float* groupDt = nullptr)
        : ws_(ws)
    {
        sphx_groups_t v{g.firstBody, g.lastBody, g.numGroups, 0, g.groupStart, g.groupEnd, groupDt};
        check(ws_, sphx_set_groups(ws_, &v));
    }
    ~Groups() { sphx_set_groups(ws_, nullptr); }
    Groups(const Groups&)            = delete;
    Groups& operator=(const Groups&) = delete;
    sphx_workspace_t* ws_;
};

//! a pair loop over the view's groups; lists and records are per particle
template<class F>
void pass(const GroupView& g, Data& d, const Box& b, F cfn)
{
    auto   ws = workspace(d);
    Groups scope(ws, g);
    auto   bx = box(b);
    auto   f  = fields(d);
    auto   p  = params(d);
    check(ws, cfn(ws, &bx, &f, &p, g.firstBody, g.lastBody));
}

/* the neighbour search of findNeighborsSfc (sph/include/sph/find_neighbors.hpp:48-56) with the h/nc
 * iteration, on the library's octree of the Domain's sorted keys (all local particles, halos included),
 * followed by the first density sum as its own kernel (the faster of the two forms; a fused kernel,
 * sphx_find_neighbors_xmass, follows the reference GPU build's xmassGpu structure) */
template<class Sum>
void searchThenSum(const GroupView& g, Data& d, const Box& b, Sum sum)
{
    if (d.ng0 > d.ngmax) { throw std::runtime_error("ng0 should be smaller than ngmax\n"); }
    auto       ws = workspace(d);
    auto       bx = box(b);
    auto       f  = fields(d);
    auto       p  = params(d);
    const auto n  = d.devData.keys.size();
    check(ws, sphx_build_tree(ws, rawPtr(d.devData.keys), n, 64, &bx));
    Groups scope(ws, g); // the spatial groups of computeGroups, one wavefront each
    check(ws, sphx_find_neighbors(ws, &bx, &f, &p, g.firstBody, g.lastBody, /*iterateH=*/1));
    check(ws, sum(ws, &bx, &f, &p, g.firstBody, g.lastBody));
}

template<class T>
constexpr int32_t hydroType()
{
    return sizeof(T) == 4 ? SPHX_F32 : SPHX_F64;
}
} // namespace

template<class Dataset>
void computeIADGpu(const GroupView& g, Dataset& d, const cstone::Box<typename Dataset::RealType>& b)
{
    pass(g, d, b, sphx_iad_std);
}

template<class Dataset>
void computeMomentumEnergyStdGpu(const GroupView& g, Dataset& d, const cstone::Box<typename Dataset::RealType>& b)
{
    auto   ws = workspace(d);
    Groups scope(ws, g);
    auto   bx = box(b);
    auto   f  = fields(d);
    auto   p  = params(d);
    double dt = 0;
    check(ws, sphx_momentum_energy_std(ws, &bx, &f, &p, g.firstBody, g.lastBody, &dt));
    d.minDtCourant = dt;
}

template void computeIADGpu(const GroupView&, Data&, const Box&);
template void computeMomentumEnergyStdGpu(const GroupView&, Data&, const Box&);

namespace cuda
{

template<class Dataset>
void computeXMass(const GroupView& g, Dataset& d, const cstone::Box<typename Dataset::RealType>& b)
{
    searchThenSum(g, d, b, sphx_xmass);
}

template<class Dataset>
void computeDensity(const GroupView& g, Dataset& d, const cstone::Box<typename Dataset::RealType>& b)
{
    searchThenSum(g, d, b, sphx_density_std);
}

template<class Dataset>
void computeVeDefGradh(const GroupView& g, Dataset& d, const cstone::Box<typename Dataset::RealType>& b)
{
    pass(g, d, b, sphx_ve_def_gradh);
}

template<class Dataset>
void computeIadDivvCurlv(const GroupView& g, Dataset& d, const cstone::Box<typename Dataset::RealType>& b)
{
    pass(g, d, b, sphx_iad_divv_curlv);
}

template<class Dataset>
void computeAVswitches(const GroupView& g, Dataset& d, const cstone::Box<typename Dataset::RealType>& b)
{
    pass(g, d, b, sphx_av_switches);
}

template<bool avClean, class Dataset>
void computeMomentumEnergy(const GroupView& g, float* groupDt, Dataset& d,
                           const cstone::Box<typename Dataset::RealType>& b)
{
    auto   ws = workspace(d);
    Groups scope(ws, g, groupDt); // the block time-step propagators pass groupDt
    auto   bx = box(b);
    auto   f  = fields(d);
    auto   p  = params(d);
    double dt = 0;
    check(ws, sphx_momentum_energy_ve(ws, &bx, &f, &p, g.firstBody, g.lastBody, avClean, &dt));
    d.minDtCourant = dt;
}

//! hydro_ve/eos_gpu.cu:63-76 on sphx_eos_ve (the C ABI takes temp and u as double and m as the HydroType)
template<class Tt, class Tm, class Th>
void computeIdealGasEOS(size_t first, size_t last, Tm mui, Tt gamma, const Tt* temp, const Tt* u, const Tm* m,
                        const Th* kx, const Th* xm, const Th* gradh, Th* prho, Th* c, Th* rho, Th* p)
{
    static_assert(std::is_same_v<Tt, double> && std::is_same_v<Tm, Th>, "types the C ABI carries");
    sphx_fields_t f{};
    f.temp = const_cast<Tt*>(temp), f.u = const_cast<Tt*>(u), f.m = const_cast<Tm*>(m);
    f.kx = const_cast<Th*>(kx), f.xm = const_cast<Th*>(xm), f.gradh = const_cast<Th*>(gradh);
    f.prho = prho, f.c = c, f.rho = rho, f.p = p;
    sphx_params_t prm{};
    prm.muiConst = double(mui), prm.gamma = double(gamma), prm.hydro_type = hydroType<Th>();
    prm.eos = SPHX_EOS_IDEAL_GAS;
    check(theWorkspace(), sphx_eos_ve(theWorkspace(), &f, &prm, uint32_t(first), uint32_t(last)));
}

//! hydro_ve/eos_gpu.cu:104-113 on sphx_eos_ve
template<class Th, class Tu>
void computeIsothermalEOS(size_t first, size_t last, Th cConst, Th* c, Th* rho, Th* p, const Th* m, const Th* kx,
                          const Th* xm, const Th* gradh, Th* prho, Tu* temp)
{
    static_assert(std::is_same_v<Tu, double>, "temperature is RealType in the C ABI");
    sphx_fields_t f{};
    f.c = c, f.rho = rho, f.p = p, f.m = const_cast<Th*>(m), f.kx = const_cast<Th*>(kx);
    f.xm = const_cast<Th*>(xm), f.gradh = const_cast<Th*>(gradh), f.prho = prho, f.temp = temp;
    sphx_params_t prm{};
    prm.hydro_type = hydroType<Th>(), prm.eos = SPHX_EOS_ISOTHERMAL, prm.soundSpeedConst = double(cConst);
    check(theWorkspace(), sphx_eos_ve(theWorkspace(), &f, &prm, uint32_t(first), uint32_t(last)));
}

//! hydro_ve/eos_gpu.cu:140-150 on sphx_eos_ve
template<class Th, class Tt>
void computePolytropicEOS(size_t first, size_t last, Tt polytropic_const, Tt polytropic_index, Th* rho, Th* p,
                          const Th* m, const Th* kx, const Th* xm, const Th* gradh, Th* prho, Tt* temp, Th* c)
{
    static_assert(std::is_same_v<Tt, double>, "temperature is RealType in the C ABI");
    sphx_fields_t f{};
    f.c = c, f.rho = rho, f.p = p, f.m = const_cast<Th*>(m), f.kx = const_cast<Th*>(kx);
    f.xm = const_cast<Th*>(xm), f.gradh = const_cast<Th*>(gradh), f.prho = prho, f.temp = temp;
    sphx_params_t prm{};
    prm.hydro_type = hydroType<Th>(), prm.eos = SPHX_EOS_POLYTROPIC;
    prm.polytropic_const = double(polytropic_const), prm.polytropic_index = double(polytropic_index);
    check(theWorkspace(), sphx_eos_ve(theWorkspace(), &f, &prm, uint32_t(first), uint32_t(last)));
}

//! hydro_std/eos_gpu.cu:57-70, :84-96, :111-123 on sphx_eos_std, with the Dataset's eosChoice set
template<class Dataset>
void stdEos(size_t first, size_t last, Dataset& d, int32_t eos)
{
    auto ws = workspace(d);
    auto f  = fields(d);
    auto p  = params(d);
    p.eos   = eos;
    check(ws, sphx_eos_std(ws, &f, &p, uint32_t(first), uint32_t(last)));
}

template<class Dataset>
void computeIdealGasEOS_HydroStd(size_t first, size_t last, Dataset& d)
{
    stdEos(first, last, d, SPHX_EOS_IDEAL_GAS);
}

template<class Dataset>
void computeIsothermalEOS_HydroStd(size_t first, size_t last, Dataset& d)
{
    stdEos(first, last, d, SPHX_EOS_ISOTHERMAL);
}

template<class Dataset>
void computePolytropicEOS_HydroStd(size_t first, size_t last, Dataset& d)
{
    stdEos(first, last, d, SPHX_EOS_POLYTROPIC);
}

// the instantiations of the replaced .cu files that the C ABI's types carry (ParticlesData uses
// SphTypes: double coordinates and temperatures, float HydroType and mass)
template void computeXMass(const GroupView&, Data&, const Box&);
template void computeDensity(const GroupView&, Data&, const Box&);
template void computeVeDefGradh(const GroupView&, Data&, const Box&);
template void computeIadDivvCurlv(const GroupView&, Data&, const Box&);
template void computeAVswitches(const GroupView&, Data&, const Box&);
template void computeMomentumEnergy<true>(const GroupView&, float*, Data&, const Box&);
template void computeMomentumEnergy<false>(const GroupView&, float*, Data&, const Box&);
template void computeIdealGasEOS(size_t, size_t, double, double, const double*, const double*, const double*,
                                 const double*, const double*, const double*, double*, double*, double*, double*);
template void computeIdealGasEOS(size_t, size_t, float, double, const double*, const double*, const float*,
                                 const float*, const float*, const float*, float*, float*, float*, float*);
template void computeIsothermalEOS(size_t, size_t, double, double*, double*, double*, const double*, const double*,
                                   const double*, const double*, double*, double*);
template void computeIsothermalEOS(size_t, size_t, float, float*, float*, float*, const float*, const float*,
                                   const float*, const float*, float*, double*);
template void computePolytropicEOS(size_t, size_t, double, double, double*, double*, const double*, const double*,
                                   const double*, const double*, double*, double*, double*);
template void computePolytropicEOS(size_t, size_t, double, double, float*, float*, const float*, const float*,
                                   const float*, const float*, float*, double*, float*);
template void computeIdealGasEOS_HydroStd(size_t, size_t, Data&);
template void computeIsothermalEOS_HydroStd(size_t, size_t, Data&);
template void computePolytropicEOS_HydroStd(size_t, size_t, Data&);

} // namespace cuda

// ---- the integrate phase (sph_gpu.hpp:68-98): double coordinates, temperatures and du; velocities,
// accelerations, h, mui and the previous-step fields in the HydroType (positions_gpu.cu:186-193)

template<class Tc, class Tv, class Ta, class Tdu, class Tm1, class Tu, class Thydro>
void computePositionsGpu(const GroupView& grp, float dt, util::array<float, Timestep::maxNumRungs> dt_m1, Tc* x,
                         Tc* y, Tc* z, Tv* vx, Tv* vy, Tv* vz, Tm1* x_m1, Tm1* y_m1, Tm1* z_m1, Ta* ax, Ta* ay, Ta* az,
                         const uint8_t* rung, Tu* temp, Tu* u, Tdu* du, Tm1* du_m1, Thydro* h, Thydro* mui, Tc gamma,
                         Tc constCv, const cstone::Box<Tc>& b)
{
    Groups        scope(theWorkspace(), grp);
    sphx_fields_t f{};
    f.x = x, f.y = y, f.z = z, f.vx = vx, f.vy = vy, f.vz = vz, f.ax = ax, f.ay = ay, f.az = az;
    f.temp = temp, f.u = u, f.du = du, f.h = h;
    sphx_history_t hs{x_m1, y_m1, z_m1, du_m1, mui};
    sphx_params_t  p{};
    p.gamma      = gamma;
    p.hydro_type = hydroType<Thydro>();
    auto bx      = box(b);
    check(theWorkspace(), sphx_compute_positions_rungs(theWorkspace(), &bx, &f, &hs, &p, grp.firstBody, grp.lastBody,
                                                       dt, dt_m1.data(), Timestep::maxNumRungs, rung, constCv));
}

template void computePositionsGpu(const GroupView&, float, util::array<float, Timestep::maxNumRungs>, double*,
                                  double*, double*, double*, double*, double*, double*, double*, double*, double*,
                                  double*, double*, const uint8_t*, double*, double*, double*, double*, double*,
                                  double*, double, double, const Box&);
template void computePositionsGpu(const GroupView&, float, util::array<float, Timestep::maxNumRungs>, double*,
                                  double*, double*, float*, float*, float*, float*, float*, float*, float*, float*,
                                  float*, const uint8_t*, double*, double*, double*, float*, float*, float*, double,
                                  double, const Box&);

template<class Tc, class Thydro, class Tm1, class Tdu>
void driftPositionsGpu(const GroupView& grp, float dt, float dt_back, util::array<float, Timestep::maxNumRungs> dt_m1,
                       Tc* x, Tc* y, Tc* z, Thydro* vx, Thydro* vy, Thydro* vz, const Tm1* x_m1, const Tm1* y_m1,
                       const Tm1* z_m1, const Thydro* ax, const Thydro* ay, const Thydro* az, const uint8_t* rung,
                       Tc* temp, Tc* u, Tdu* du, Tm1* du_m1, Thydro* mui, Tc gamma, Tc constCv)
{
    Groups        scope(theWorkspace(), grp);
    sphx_fields_t f{};
    f.x = x, f.y = y, f.z = z, f.vx = vx, f.vy = vy, f.vz = vz;
    f.ax = const_cast<Thydro*>(ax), f.ay = const_cast<Thydro*>(ay), f.az = const_cast<Thydro*>(az);
    f.temp = temp, f.u = u, f.du = du;
    sphx_history_t hs{const_cast<Tm1*>(x_m1), const_cast<Tm1*>(y_m1), const_cast<Tm1*>(z_m1), du_m1, mui};
    sphx_params_t  p{};
    p.gamma      = gamma;
    p.hydro_type = hydroType<Thydro>();
    check(theWorkspace(), sphx_drift_positions(theWorkspace(), &f, &hs, &p, grp.firstBody, grp.lastBody, dt, dt_back,
                                               dt_m1.data(), Timestep::maxNumRungs, rung, constCv));
}

template void driftPositionsGpu(const GroupView&, float, float, util::array<float, Timestep::maxNumRungs>, double*,
                                double*, double*, float*, float*, float*, const float*, const float*, const float*,
                                const float*, const float*, const float*, const uint8_t*, double*, double*, double*,
                                float*, float*, double, double);

template<class Th>
void updateSmoothingLengthGpu(const GroupView& grp, unsigned ng0, const unsigned* nc, Th* h)
{
    Groups        scope(theWorkspace(), grp);
    sphx_fields_t f{};
    f.h = h, f.nc = const_cast<unsigned*>(nc);
    sphx_params_t p{};
    p.ng0 = ng0, p.hydro_type = hydroType<Th>();
    check(theWorkspace(), sphx_update_smoothing_length(theWorkspace(), &f, &p, grp.firstBody, grp.lastBody));
}

template void updateSmoothingLengthGpu(const GroupView&, unsigned, const unsigned*, float*);
template void updateSmoothingLengthGpu(const GroupView&, unsigned, const unsigned*, double*);

template<class T>
void groupDivvTimestepGpu(float Krho, const GroupView& grp, const T* divv, float* groupDt)
{
    Groups        scope(theWorkspace(), grp);
    sphx_fields_t f{};
    f.divv = const_cast<T*>(divv);
    sphx_params_t p{};
    p.hydro_type = hydroType<T>();
    check(theWorkspace(),
          sphx_group_divv_timestep(theWorkspace(), &f, &p, grp.firstBody, grp.lastBody, Krho, groupDt));
}

template<class T>
void groupAccTimestepGpu(float etaAcc, const GroupView& grp, const T* ax, const T* ay, const T* az, float* groupDt)
{
    Groups        scope(theWorkspace(), grp);
    sphx_fields_t f{};
    f.ax = const_cast<T*>(ax), f.ay = const_cast<T*>(ay), f.az = const_cast<T*>(az);
    sphx_params_t p{};
    p.hydro_type = hydroType<T>();
    check(theWorkspace(),
          sphx_group_acc_timestep(theWorkspace(), &f, &p, grp.firstBody, grp.lastBody, etaAcc, groupDt));
}

template void groupDivvTimestepGpu(float, const GroupView&, const float*, float*);
template void groupDivvTimestepGpu(float, const GroupView&, const double*, float*);
template void groupAccTimestepGpu(float, const GroupView&, const float*, const float*, const float*, float*);
template void groupAccTimestepGpu(float, const GroupView&, const double*, const double*, const double*, float*);

void storeRungGpu(const GroupView& grp, uint8_t rung, uint8_t* particleRungs)
{
    Groups scope(theWorkspace(), grp);
    check(theWorkspace(), sphx_store_rung(theWorkspace(), grp.firstBody, grp.lastBody, rung, particleRungs));
}

unsigned nsGroupSize() { return 64; } // target groups of at most 64 particles, one wavefront each

} // namespace sph
