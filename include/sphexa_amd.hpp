/*
 * sphexa_amd C++ host interface: the reference's hydro-path function names and argument meaning
 * (sph/include/sph/find_neighbors.hpp and the hydro_std, hydro_ve drivers) on top of the C ABI in
 * sphexa_amd.h. A Dataset type provides device field pointers through `devData` and the scalar
 * parameters under their ParticlesData names (sph/include/sph/particles_data.hpp:91-146);
 * `DeviceDataset` below is a minimal ready-made one.
 *
 * Error behaviour follows the reference: a rejected argument or a HIP failure throws
 * std::runtime_error; ng0 > ngmax throws "ng0 should be smaller than ngmax"
 * (find_neighbors.hpp:52); a non-converged h/nc iteration throws "coupled nc/h-updated failed to
 * converge" as the reference GPU path does (hydro_ve/xmass_gpu.cu:128).
 */
#pragma once

#include <array>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>

#include "sphexa_amd.h"

namespace sphexa_amd
{

enum class BoundaryType : char
{
    open     = 0,
    periodic = 1,
    fixed    = 2
};

//! cstone::Box<double> (domain/include/cstone/sfc/box.hpp:95-174), the fields the path uses
struct Box
{
    double       lo[3], hi[3];
    BoundaryType boundary[3];

    Box(double xyzMin, double xyzMax, BoundaryType b = BoundaryType::open)
        : lo{xyzMin, xyzMin, xyzMin}
        , hi{xyzMax, xyzMax, xyzMax}
        , boundary{b, b, b}
    {
    }
    Box(double xmin, double xmax, double ymin, double ymax, double zmin, double zmax,
        BoundaryType bx = BoundaryType::open, BoundaryType by = BoundaryType::open,
        BoundaryType bz = BoundaryType::open)
        : lo{xmin, ymin, zmin}
        , hi{xmax, ymax, zmax}
        , boundary{bx, by, bz}
    {
    }

    explicit Box(const sphx_box_t& b)
        : lo{b.lo[0], b.lo[1], b.lo[2]}
        , hi{b.hi[0], b.hi[1], b.hi[2]}
        , boundary{BoundaryType(b.boundary[0]), BoundaryType(b.boundary[1]), BoundaryType(b.boundary[2])}
    {
    }

    sphx_box_t c() const
    {
        sphx_box_t b;
        for (int d = 0; d < 3; ++d)
        {
            b.lo[d] = lo[d], b.hi[d] = hi[d], b.boundary[d] = int32_t(boundary[d]);
        }
        return b;
    }
};

//! cstone::GroupView (domain/include/cstone/traversal/groups.hpp:28-34), device arrays of at most 64
//! particles per group; with null groupStart/groupEnd the targets [firstBody, lastBody) run as
//! consecutive groups of 64
struct GroupView
{
    uint32_t        firstBody, lastBody;
    uint32_t        numGroups;
    const uint32_t* groupStart;
    const uint32_t* groupEnd;
};

//! device field pointers, named and typed as DeviceParticlesData (sph/include/sph/particles_data_gpu.cuh)
template<class HydroType>
struct DeviceFields
{
    double *x{}, *y{}, *z{}, *u{}, *temp{}, *du{};
    HydroType *h{}, *m{}, *vx{}, *vy{}, *vz{}, *rho{}, *p{}, *c{}, *prho{}, *tdpdTrho{};
    HydroType *c11{}, *c12{}, *c13{}, *c22{}, *c23{}, *c33{}, *xm{}, *kx{}, *gradh{}, *divv{}, *curlv{}, *alpha{};
    HydroType *ax{}, *ay{}, *az{}, *dV11{}, *dV12{}, *dV13{}, *dV22{}, *dV23{}, *dV33{};
    HydroType *x_m1{}, *y_m1{}, *z_m1{}, *du_m1{}, *mui{}; // XM1Type = HydroType, mui optional
    uint64_t* keys{};
    uint32_t* nc{};
};

//! A minimal Dataset: device fields + the ParticlesData scalars + the workspace (scratch owner)
template<class HydroType>
struct DeviceDataset
{
    using RealType = double;
    using HydroT   = HydroType;

    DeviceFields<HydroType> devData;
    sphx_workspace_t*       workspace{nullptr};

    uint64_t iteration{1}, numParticlesGlobal{0}; // step attributes of the file I/O (particles_data.hpp:87-88)
    unsigned ng0{100}, ngmax{150};
    double   K{0}, Kcour{0.2}, gamma{5.0 / 3.0}, muiConst{10.0};
    double   sincIndex{6.0}; // the kernel's sinc exponent (sphx_workspace_set_kernel takes it)
    double   minDt{1e-12}, minDtCourant{INFINITY};
    double   alphamin{0.05}, alphamax{1.0}, decay_constant{0.2};
    double   Atmin{0.1}, Atmax{0.2}, ramp{1.0 / (0.2 - 0.1)};
    // equation of state (particles_data.hpp:117-127): eosChoice as sph::EosType (0 ideal gas,
    // 1 isothermal, 2 polytropic)
    int       eosChoice{SPHX_EOS_IDEAL_GAS};
    double    polytropic_index{5. / 3.}, polytropic_const{1.};
    HydroType soundSpeedConst{1.0};
    // time stepping (particles_data.hpp:96-141)
    double                  ttot{0.0}, minDt_m1{1e-12}, minDtRho{INFINITY};
    double                  g{0.0}, etaAcc{0.2}, eps{0.005}, Krho{0.06};
    constexpr static double maxDtIncrease = 1.1;
    // SPHX_FLAG_* bits passed with every call (no counterpart in ParticlesData); a Dataset
    // without this member passes 0, the reference behaviour
    int32_t sphxFlags{0};
};

//! block time-step rungs of the reference's integrator (sph/include/sph/timestep.h:19)
constexpr int maxNumRungs = 4;

namespace detail
{

inline void check(sphx_workspace_t* ws, int code)
{
    if (code == SPHX_OK) { return; }
    throw std::runtime_error(std::string("sphexa_amd: ") + sphx_last_error(ws));
}

template<class D, class = void>
struct HasSphxFlags : std::false_type
{
};
template<class D>
struct HasSphxFlags<D, std::void_t<decltype(std::declval<const D&>().sphxFlags)>> : std::true_type
{
};

template<class D, class = void>
struct HasEosChoice : std::false_type
{
};
template<class D>
struct HasEosChoice<D, std::void_t<decltype(std::declval<const D&>().eosChoice)>> : std::true_type
{
};

template<class Dataset>
sphx_params_t params(const Dataset& d)
{
    using H = typename Dataset::HydroT;
    sphx_params_t p{};
    p.K = d.K, p.Kcour = d.Kcour, p.gamma = d.gamma, p.muiConst = d.muiConst, p.minDt = d.minDt;
    p.alphamin = d.alphamin, p.alphamax = d.alphamax, p.decay_constant = d.decay_constant;
    p.Atmin = d.Atmin, p.Atmax = d.Atmax, p.ramp = d.ramp;
    p.ng0 = d.ng0, p.ngmax = d.ngmax;
    p.hydro_type = sizeof(H) == 4 ? SPHX_F32 : SPHX_F64;
    if constexpr (HasSphxFlags<Dataset>::value) { p.flags = d.sphxFlags; }
    // a Dataset without an EOS choice runs the ideal gas, the reference's default
    p.eos = SPHX_EOS_IDEAL_GAS, p.soundSpeedConst = 1.0, p.polytropic_const = 1.0, p.polytropic_index = 5. / 3.;
    if constexpr (HasEosChoice<Dataset>::value)
    {
        p.eos              = int32_t(d.eosChoice);
        p.soundSpeedConst  = double(d.soundSpeedConst);
        p.polytropic_const = double(d.polytropic_const);
        p.polytropic_index = double(d.polytropic_index);
    }
    return p;
}

/*! sets the caller's target groups on the workspace for the lifetime of the guard (sphx_set_groups),
 * with the per-group time steps the momentum loops lower; a view without group arrays leaves the
 * consecutive groups of 64 */
class GroupScope
{
public:
    GroupScope(sphx_workspace_t* ws, const GroupView& grp, float* groupDt = nullptr)
        : ws_(ws)
    {
        if (grp.groupStart == nullptr)
        {
            if (groupDt != nullptr) { throw std::runtime_error("per-group time steps need the group arrays"); }
            return;
        }
        sphx_groups_t g{grp.firstBody, grp.lastBody, grp.numGroups, 0, grp.groupStart, grp.groupEnd, groupDt};
        check(ws_, sphx_set_groups(ws_, &g));
        set_ = true;
    }
    ~GroupScope()
    {
        if (set_) { sphx_set_groups(ws_, nullptr); }
    }
    GroupScope(const GroupScope&)            = delete;
    GroupScope& operator=(const GroupScope&) = delete;

private:
    sphx_workspace_t* ws_;
    bool              set_ = false;
};

template<class Dataset>
sphx_fields_t fields(Dataset& d)
{
    auto&         s = d.devData;
    sphx_fields_t f{};
    f.x = s.x, f.y = s.y, f.z = s.z, f.u = s.u, f.temp = s.temp, f.du = s.du;
    f.h = s.h, f.m = s.m, f.vx = s.vx, f.vy = s.vy, f.vz = s.vz, f.rho = s.rho, f.p = s.p, f.c = s.c;
    f.prho = s.prho, f.tdpdTrho = s.tdpdTrho, f.c11 = s.c11, f.c12 = s.c12, f.c13 = s.c13, f.c22 = s.c22;
    f.c23 = s.c23, f.c33 = s.c33, f.xm = s.xm, f.kx = s.kx, f.gradh = s.gradh, f.divv = s.divv;
    f.curlv = s.curlv, f.alpha = s.alpha, f.ax = s.ax, f.ay = s.ay, f.az = s.az;
    f.dV11 = s.dV11, f.dV12 = s.dV12, f.dV13 = s.dV13, f.dV22 = s.dV22, f.dV23 = s.dV23, f.dV33 = s.dV33;
    f.keys = s.keys, f.nc = s.nc;
    return f;
}

template<class Dataset>
sphx_history_t history(Dataset& d)
{
    auto&          s = d.devData;
    sphx_history_t h{};
    h.x_m1 = s.x_m1, h.y_m1 = s.y_m1, h.z_m1 = s.z_m1, h.du_m1 = s.du_m1, h.mui = s.mui;
    return h;
}

template<class Dataset>
sphx_timestep_t timestep(const Dataset& d)
{
    sphx_timestep_t t{};
    t.minDt = d.minDt, t.minDt_m1 = d.minDt_m1, t.ttot = d.ttot, t.minDtCourant = d.minDtCourant;
    t.minDtRho = d.minDtRho, t.maxDtIncrease = d.maxDtIncrease, t.g = d.g, t.etaAcc = d.etaAcc, t.eps = d.eps;
    t.Krho = d.Krho;
    return t;
}

} // namespace detail

//! one-rank Domain::sync (domain/include/cstone/domain/domain.hpp:182-234) as the propagators call
//! it (main/src/propagator/std_hydro.hpp:108-122, ve_hydro.hpp:114-128): keys, sort, permute the
//! conserved fields (x y z h m u temp vx vy vz alpha and the integrator's previous-step fields),
//! then the tree: converged, or with treeUpdate the previous sync's tree updated once (the syncs
//! after a Domain's first, SPHX_FLAG_TREE_UPDATE). The dependent fields keep their order; every pass
//! rewrites them before reading.
template<class Dataset>
void syncDomain(size_t numParticles, Dataset& d, const Box& box, unsigned bucketSize = 64, bool treeUpdate = false)
{
    auto          b = box.c();
    auto&         s = d.devData;
    sphx_fields_t f{};
    f.x = s.x, f.y = s.y, f.z = s.z, f.h = s.h, f.m = s.m, f.u = s.u, f.temp = s.temp;
    f.vx = s.vx, f.vy = s.vy, f.vz = s.vz, f.alpha = s.alpha, f.keys = s.keys;
    auto p = detail::params(d);
    if (treeUpdate) { p.flags |= SPHX_FLAG_TREE_UPDATE; }
    detail::check(d.workspace, sphx_domain_sync(d.workspace, &b, &f, &p, numParticles, bucketSize));
    void* extra[5];
    int   bytes[5];
    size_t count = 0;
    for (void* e : {static_cast<void*>(s.x_m1), static_cast<void*>(s.y_m1), static_cast<void*>(s.z_m1),
                    static_cast<void*>(s.du_m1), static_cast<void*>(s.mui)})
    {
        if (e) { extra[count] = e, bytes[count++] = int(sizeof(typename Dataset::HydroT)); }
    }
    detail::check(d.workspace, sphx_domain_reorder(d.workspace, extra, bytes, count, numParticles));
}

/*! One-rank cstone::Domain (domain/include/cstone/domain/domain.hpp): owns the box and the
 *  first-call state. sync() refits the box as GlobalAssignment::assign does (assignment.hpp:80-82:
 *  the fitting box on the first call, then limitBoxShrinking; periodic dimensions fixed), then runs
 *  syncDomain, converging the tree on the first call and updating it once on every later one. All particles are local: startIndex() = 0, endIndex() = nParticles(). */
class Domain
{
public:
    Domain(int /*rank*/, int nRanks, unsigned bucketSize, unsigned /*bucketSizeFocus*/, float /*theta*/,
           const Box& box)
        : box_(box)
        , bucketSize_(bucketSize)
    {
        if (nRanks != 1) { throw std::runtime_error("this domain holds one rank's particles only"); }
    }

    template<class Dataset>
    void sync(size_t numParticles, Dataset& d)
    {
        auto& s = d.devData;
        auto  b = box_.c();
        detail::check(d.workspace, sphx_global_box(d.workspace, &b, s.x, s.y, s.z, numParticles, firstCall_));
        box_ = Box(b);
        // the first sync converges the tree, later ones update it once (domain.hpp:199-212)
        syncDomain(numParticles, d, box_, bucketSize_, !firstCall_);
        numParticles_ = numParticles;
        firstCall_    = false;
    }

    const Box& box() const { return box_; }
    size_t     startIndex() const { return 0; }
    size_t     endIndex() const { return numParticles_; }
    size_t     nParticles() const { return numParticles_; }
    size_t     nParticlesWithHalos() const { return numParticles_; }

private:
    Box      box_;
    unsigned bucketSize_;
    size_t   numParticles_ = 0;
    bool     firstCall_    = true;
};

/*! sph::computeGroups for the GPU (sph/include/sph/groups.hpp:42-73): spatial target groups of
 * [startIndex, endIndex) (computeGroupSplits, tolFactor 2) on the tree of the last sync; the view's
 * arrays live in the workspace until the next call */
template<class Dataset>
GroupView computeGroups(size_t startIndex, size_t endIndex, Dataset& d, const Box& box)
{
    auto          b = box.c();
    auto          f = detail::fields(d);
    sphx_groups_t g{};
    detail::check(d.workspace,
                  sphx_compute_groups(d.workspace, &b, &f, uint32_t(startIndex), uint32_t(endIndex), 2.0f, &g));
    return GroupView{g.firstBody, g.lastBody, g.numGroups, g.groupStart, g.groupEnd};
}

/*! sph::findNeighborsSfc (sph/include/sph/find_neighbors.hpp:48-56); with a group view, the search runs
 * one wavefront per group of the view (the reference GPU path searches per group inside computeXMass /
 * computeDensity) */
template<class Dataset>
void findNeighborsSfc(const GroupView& grp, Dataset& d, const Box& box)
{
    if (d.ng0 > d.ngmax) { throw std::runtime_error("ng0 should be smaller than ngmax\n"); }
    detail::GroupScope scope(d.workspace, grp);
    auto               b    = box.c();
    auto               f    = detail::fields(d);
    auto               p    = detail::params(d);
    int                code = sphx_find_neighbors(d.workspace, &b, &f, &p, grp.firstBody, grp.lastBody, 1);
    if (code == SPHX_ERR_NOT_CONVERGED) { throw std::runtime_error("coupled nc/h-updated failed to converge"); }
    detail::check(d.workspace, code);
}

//! sph::findNeighborsSfc (sph/include/sph/find_neighbors.hpp:48-56)
template<class Dataset>
void findNeighborsSfc(size_t startIndex, size_t endIndex, Dataset& d, const Box& box)
{
    if (d.ng0 > d.ngmax) { throw std::runtime_error("ng0 should be smaller than ngmax\n"); }
    auto b    = box.c();
    auto f    = detail::fields(d);
    auto p    = detail::params(d);
    int  code = sphx_find_neighbors(d.workspace, &b, &f, &p, uint32_t(startIndex), uint32_t(endIndex), 1);
    if (code == SPHX_ERR_NOT_CONVERGED) { throw std::runtime_error("coupled nc/h-updated failed to converge"); }
    detail::check(d.workspace, code);
}

namespace detail {
template<class Dataset>
void searchXMass(const GroupView& grp, Dataset& d, const Box& box, int out)
{
    if (d.ng0 > d.ngmax) { throw std::runtime_error("ng0 should be smaller than ngmax\n"); }
    GroupScope scope(d.workspace, grp);
    auto       b    = box.c();
    auto       f    = detail::fields(d);
    auto       p    = detail::params(d);
    int        code = sphx_find_neighbors_xmass(d.workspace, &b, &f, &p, grp.firstBody, grp.lastBody, 1, out);
    if (code == SPHX_ERR_NOT_CONVERGED) { throw std::runtime_error("coupled nc/h-updated failed to converge"); }
    detail::check(d.workspace, code);
}
} // namespace detail

//! GPU sph::computeXMass (sph/include/sph/hydro_ve/xmass_gpu.cu:57-131): neighbour search with the
//! h/nc iteration and the XMass sum in one pass; replaces findNeighborsSfc + computeXMass
template<class Dataset>
void findNeighborsXMass(const GroupView& grp, Dataset& d, const Box& box)
{
    detail::searchXMass(grp, d, box, SPHX_XMASS_XM);
}

//! GPU sph::computeDensity (sph/include/sph/hydro_ve/xmass_gpu.cu:150-165, rho = m / xm): neighbour
//! search with the h/nc iteration and the std density in one pass; replaces findNeighborsSfc +
//! computeDensity
template<class Dataset>
void findNeighborsDensity(const GroupView& grp, Dataset& d, const Box& box)
{
    detail::searchXMass(grp, d, box, SPHX_XMASS_RHO);
}

#define SPHEXA_AMD_GROUP_PASS(name, cfn)                                                                           \
    template<class Dataset>                                                                                            \
    void name(const GroupView& grp, Dataset& d, const Box& box)                                                        \
    {                                                                                                                  \
        detail::GroupScope scope(d.workspace, grp);                                                                    \
        auto               b = box.c();                                                                                \
        auto f = detail::fields(d);                                                                                    \
        auto p = detail::params(d);                                                                                    \
        detail::check(d.workspace, cfn(d.workspace, &b, &f, &p, grp.firstBody, grp.lastBody));                         \
    }

//! sph::computeDensity (hydro_std/density.hpp:56-60)
SPHEXA_AMD_GROUP_PASS(computeDensity, sphx_density_std)
//! sph::computeIAD (hydro_std/iad.hpp:75-79)
SPHEXA_AMD_GROUP_PASS(computeIAD, sphx_iad_std)
//! sph::computeXMass (hydro_ve/xmass.hpp:67-72)
SPHEXA_AMD_GROUP_PASS(computeXMass, sphx_xmass)
//! sph::computeVeDefGradh (hydro_ve/ve_def_gradh.hpp:79-84)
SPHEXA_AMD_GROUP_PASS(computeVeDefGradh, sphx_ve_def_gradh)
//! sph::computeIadDivvCurlv (hydro_ve/iad_divv_curlv.hpp:92-97)
SPHEXA_AMD_GROUP_PASS(computeIadDivvCurlv, sphx_iad_divv_curlv)
//! sph::computeAVswitches (hydro_ve/av_switches.hpp:80-85)
SPHEXA_AMD_GROUP_PASS(computeAVswitches, sphx_av_switches)

#undef SPHEXA_AMD_GROUP_PASS

//! sph::computeEOS_HydroStd (hydro_std/eos.hpp:146-152): ideal gas, isothermal or polytropic by d.eosChoice
template<class Dataset>
void computeEOS_HydroStd(size_t startIndex, size_t endIndex, Dataset& d)
{
    auto f = detail::fields(d);
    auto p = detail::params(d);
    detail::check(d.workspace, sphx_eos_std(d.workspace, &f, &p, uint32_t(startIndex), uint32_t(endIndex)));
}

//! sph::computeEOS (hydro_ve/eos.hpp:191-197): ideal gas, isothermal or polytropic by d.eosChoice
template<class Dataset>
void computeEOS(size_t startIndex, size_t endIndex, Dataset& d)
{
    auto f = detail::fields(d);
    auto p = detail::params(d);
    detail::check(d.workspace, sphx_eos_ve(d.workspace, &f, &p, uint32_t(startIndex), uint32_t(endIndex)));
}

//! sph::computeMomentumEnergySTD (hydro_std/momentum_energy.hpp:97-101); sets d.minDtCourant
template<class Dataset>
void computeMomentumEnergySTD(const GroupView& grp, Dataset& d, const Box& box)
{
    detail::GroupScope scope(d.workspace, grp);
    auto   b  = box.c();
    auto   f  = detail::fields(d);
    auto   p  = detail::params(d);
    double dt = 0;
    detail::check(d.workspace,
                  sphx_momentum_energy_std(d.workspace, &b, &f, &p, grp.firstBody, grp.lastBody, &dt));
    d.minDtCourant = dt;
}

//! sph::computeMomentumEnergy<avClean> (hydro_ve/momentum_energy.hpp:106-115); sets d.minDtCourant and,
//! when groupDt is given (the block time-step propagators), lowers groupDt[g] to group g's smallest
//! Courant step
template<bool avClean, class Dataset>
void computeMomentumEnergy(const GroupView& grp, float* groupDt, Dataset& d, const Box& box)
{
    detail::GroupScope scope(d.workspace, grp, groupDt);
    auto   b  = box.c();
    auto   f  = detail::fields(d);
    auto   p  = detail::params(d);
    double dt = 0;
    detail::check(d.workspace, sphx_momentum_energy_ve(d.workspace, &b, &f, &p, grp.firstBody, grp.lastBody,
                                                       avClean ? 1 : 0, &dt));
    d.minDtCourant = dt;
}

//! sph::rhoTimestep (sph/include/sph/ts_global.hpp:70-93): Krho / |max divv| over [first, last)
template<class Dataset>
double rhoTimestep(size_t first, size_t last, Dataset& d)
{
    if (d.devData.divv == nullptr) { throw std::runtime_error("Divv needs to be available in rhoTimestep\n"); }
    auto f  = detail::fields(d);
    auto p  = detail::params(d);
    auto ts = detail::timestep(d);
    detail::check(d.workspace, sphx_rho_timestep(d.workspace, &f, &p, uint32_t(first), uint32_t(last), &ts));
    return ts.minDtRho;
}

//! sph::computeTimestep (ts_global.hpp:95-113) on one rank: sets d.minDt, d.minDt_m1, d.ttot
template<class Dataset>
void computeTimestep(size_t first, size_t last, Dataset& d)
{
    auto f  = detail::fields(d);
    auto p  = detail::params(d);
    auto ts = detail::timestep(d);
    detail::check(d.workspace, sphx_compute_timestep(d.workspace, &f, &p, uint32_t(first), uint32_t(last), &ts));
    d.minDt = ts.minDt, d.minDt_m1 = ts.minDt_m1, d.ttot = ts.ttot;
}

//! sph::computePositions (sph/include/sph/positions.hpp:177-200), its GPU branch: positions,
//! velocities, x_m1..z_m1, and temp or u with du_m1. With a rung array (block time steps, a device
//! array of one uint8 per particle) particle i takes the previous step dt_m1[rung[i]], else dt_m1[0].
template<class Dataset>
void computePositions(const GroupView& grp, Dataset& d, const Box& box, float dt_forward,
                      std::array<float, maxNumRungs> dt_m1, const uint8_t* rung = nullptr)
{
    detail::GroupScope scope(d.workspace, grp);
    auto b = box.c();
    auto f = detail::fields(d);
    auto h = detail::history(d);
    auto p = detail::params(d);
    detail::check(d.workspace, sphx_compute_positions_rungs(d.workspace, &b, &f, &h, &p, grp.firstBody,
                                                            grp.lastBody, double(dt_forward), dt_m1.data(),
                                                            int(maxNumRungs), rung, -1.0));
}

//! sph::driftPositions (sph/include/sph/positions.hpp:151-174): the particles of grp, last advanced by
//! dt_backward within the current block time-step hierarchy, drift to dt_forward from its start
template<class Dataset>
void driftPositions(const GroupView& grp, Dataset& d, float dt_forward, float dt_backward,
                    std::array<float, maxNumRungs> dt_prevRung, const uint8_t* rung)
{
    detail::GroupScope scope(d.workspace, grp);
    auto f = detail::fields(d);
    auto h = detail::history(d);
    auto p = detail::params(d);
    detail::check(d.workspace, sphx_drift_positions(d.workspace, &f, &h, &p, grp.firstBody, grp.lastBody,
                                                    double(dt_forward), double(dt_backward), dt_prevRung.data(),
                                                    int(maxNumRungs), rung, -1.0));
}

//! sph::groupDivvTimestep (sph/include/sph/ts_rungs.hpp:49-56): groupDt[g] = min(groupDt[g], Krho / |max divv|)
template<class Dataset>
void groupDivvTimestep(const GroupView& grp, float* groupDt, Dataset& d)
{
    detail::GroupScope scope(d.workspace, grp);
    auto f = detail::fields(d);
    auto p = detail::params(d);
    detail::check(d.workspace, sphx_group_divv_timestep(d.workspace, &f, &p, grp.firstBody, grp.lastBody,
                                                        float(d.Krho), groupDt));
}

//! sph::groupAccTimestep (ts_rungs.hpp:58-67): groupDt[g] = min(groupDt[g], etaAcc sqrt(eps) / |a|max^(1/2))
template<class Dataset>
void groupAccTimestep(const GroupView& grp, float* groupDt, Dataset& d)
{
    detail::GroupScope scope(d.workspace, grp);
    auto f = detail::fields(d);
    auto p = detail::params(d);
    detail::check(d.workspace, sphx_group_acc_timestep(d.workspace, &f, &p, grp.firstBody, grp.lastBody,
                                                       float(d.etaAcc * std::sqrt(d.eps)), groupDt));
}

//! sph::storeRungGpu (sph/include/sph/ts_groups.cu:84-104): particleRungs[i] = rung over grp
template<class Dataset>
void storeRung(const GroupView& grp, uint8_t rung, uint8_t* particleRungs, Dataset& d)
{
    detail::GroupScope scope(d.workspace, grp);
    detail::check(d.workspace, sphx_store_rung(d.workspace, grp.firstBody, grp.lastBody, rung, particleRungs));
}

//! sph::updateSmoothingLength (sph/include/sph/update_h.hpp:26-35)
template<class Dataset>
void updateSmoothingLength(const GroupView& grp, Dataset& d)
{
    detail::GroupScope scope(d.workspace, grp);
    auto f = detail::fields(d);
    auto p = detail::params(d);
    detail::check(d.workspace, sphx_update_smoothing_length(d.workspace, &f, &p, grp.firstBody, grp.lastBody));
}

//! sphexa::computeConservedQuantities (main/src/observables/conserved_quantities.hpp:117-177) on
//! one rank: {eKin, eInt, egrav, linmom[3], angmom[3], sum nc}, the reference's quantities array
template<class Dataset>
std::array<double, 10> computeConservedQuantities(size_t startIndex, size_t endIndex, Dataset& d, double egrav = 0)
{
    auto             f = detail::fields(d);
    auto             p = detail::params(d);
    sphx_conserved_t c{};
    detail::check(d.workspace,
                  sphx_conserved_quantities(d.workspace, &f, &p, uint32_t(startIndex), uint32_t(endIndex), -1.0, &c));
    return {c.eKin, c.eInt, egrav, c.linmom[0], c.linmom[1], c.linmom[2], c.angmom[0], c.angmom[1], c.angmom[2],
            c.ncsum};
}

} // namespace sphexa_amd
