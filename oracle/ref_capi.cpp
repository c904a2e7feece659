/*
 * C ABI around the reference's own CPU/OpenMP hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * Compiled by oracle/Makefile (target `ref`) against the reference headers where they lie
 * (REF/domain/include, REF/sph/include) into oracle/_ref/libsphx_ref.so. It pins the oracle
 * restatement (oracle/sphx_oracle.hpp) bit for bit and serves as the "reference" CPU baseline.
 * Exposes the same entry points as oracle_capi.cpp with a `ref_` prefix.
 *
 * The only code here is glue: a non-owning Dataset stand-in with the members the reference
 * drivers read (sphexa::ParticlesData field names, particles_data.hpp:91-266) and calls into
 * the reference functions.
 */
#include <iostream> // sph/find_neighbors.hpp uses std::cout without including it
#include <array>
#include <chrono>
#include <cstring>
#include <numeric>
#include <vector>

#include "cstone/findneighbors.hpp"
#include "cstone/focus/octree_focus.hpp"
#include "cstone/focus/source_center.hpp"
#include "cstone/primitives/gather.hpp"
#include "cstone/sfc/sfc.hpp"
#include "cstone/tree/csarray.hpp"
#include "cstone/tree/octree.hpp"

#include "sph/kernels.hpp" // sph/find_neighbors.hpp calls updateH without including it
#include "sph/find_neighbors.hpp"
#include "sph/sph_kernel_tables.hpp"
#include "sph/table_lookup.hpp"
#include "sph/hydro_std/density.hpp"
#include "sph/hydro_std/eos.hpp"
#include "sph/hydro_std/iad.hpp"
#include "sph/hydro_std/momentum_energy.hpp"
#include "sph/hydro_ve/xmass.hpp"
#include "sph/hydro_ve/ve_def_gradh.hpp"
#include "sph/hydro_ve/eos.hpp"
#include "sph/hydro_ve/iad_divv_curlv.hpp"
#include "sph/hydro_ve/av_switches.hpp"
#include "sph/hydro_ve/momentum_energy.hpp"
#include "sph/positions.hpp"
#include "sph/update_h.hpp"

#include "sphexa_amd.h"

namespace refglue
{

using cstone::LocalIndex;
using KeyT = uint64_t;

//! non-owning vector-like view over caller memory
template<class V>
struct View
{
    V*     p = nullptr;
    size_t n = 0;
    V*       data() { return p; }
    const V* data() const { return p; }
    size_t   size() const { return n; }
    bool     empty() const { return n == 0; }
    V&       operator[](size_t i) { return p[i]; }
    const V& operator[](size_t i) const { return p[i]; }
};

template<class V>
void swap(View<V>& a, View<V>& b)
{
    std::swap(a.p, b.p);
    std::swap(a.n, b.n);
}

template<class V>
View<V> view(void* p, size_t n)
{
    return View<V>{static_cast<V*>(p), p ? n : 0};
}

//! the members of sphexa::ParticlesData that the reference CPU drivers use
template<class T>
struct Data
{
    using AcceleratorType = cstone::CpuTag;
    using RealType        = double;
    using HydroType       = T;
    using Tmass           = T;
    using KeyType         = KeyT;

    View<double> x, y, z, u, temp, du;
    View<T> h, m, vx, vy, vz, rho, p, c, prho, tdpdTrho, c11, c12, c13, c22, c23, c33, xm, kx, gradh, divv, curlv,
        alpha, ax, ay, az, dV11, dV12, dV13, dV22, dV23, dV33;
    View<unsigned>   nc;
    View<LocalIndex> neighbors;
    View<const T>    wh, whd;

    unsigned      ng0{100}, ngmax{150};
    double        K{0}, Kcour{0.2}, gamma{5.0 / 3.0}, minDt{1e-12}, minDtCourant{INFINITY}, sincIndex{6.0};
    T             muiConst{10.0};
    T             soundSpeedConst{1.0};
    double        polytropic_const{1.}, polytropic_index{5. / 3.};
    sph::EosType  eosChoice{sph::EosType::idealGas};
    T             alphamin{0.05}, alphamax{1.0}, decay_constant{0.2};
    T             Atmin{0.1}, Atmax{0.2}, ramp{1.0 / (0.2 - 0.1)};
};

template<class T>
Data<T> makeData(const sphx_fields_t* f, const sphx_params_t* p, size_t n, const uint32_t* nb, size_t nbLen,
                 const void* wh, const void* whd)
{
    Data<T> d;
    d.x  = view<double>(f->x, n), d.y = view<double>(f->y, n), d.z = view<double>(f->z, n);
    d.u  = view<double>(f->u, n), d.temp = view<double>(f->temp, n), d.du = view<double>(f->du, n);
    d.h  = view<T>(f->h, n), d.m = view<T>(f->m, n), d.vx = view<T>(f->vx, n), d.vy = view<T>(f->vy, n);
    d.vz = view<T>(f->vz, n), d.rho = view<T>(f->rho, n), d.p = view<T>(f->p, n), d.c = view<T>(f->c, n);
    d.prho = view<T>(f->prho, n), d.tdpdTrho = view<T>(f->tdpdTrho, n);
    d.c11 = view<T>(f->c11, n), d.c12 = view<T>(f->c12, n), d.c13 = view<T>(f->c13, n);
    d.c22 = view<T>(f->c22, n), d.c23 = view<T>(f->c23, n), d.c33 = view<T>(f->c33, n);
    d.xm = view<T>(f->xm, n), d.kx = view<T>(f->kx, n), d.gradh = view<T>(f->gradh, n);
    d.divv = view<T>(f->divv, n), d.curlv = view<T>(f->curlv, n), d.alpha = view<T>(f->alpha, n);
    d.ax = view<T>(f->ax, n), d.ay = view<T>(f->ay, n), d.az = view<T>(f->az, n);
    d.dV11 = view<T>(f->dV11, n), d.dV12 = view<T>(f->dV12, n), d.dV13 = view<T>(f->dV13, n);
    d.dV22 = view<T>(f->dV22, n), d.dV23 = view<T>(f->dV23, n), d.dV33 = view<T>(f->dV33, n);
    d.nc        = view<unsigned>(f->nc, n);
    d.neighbors = View<LocalIndex>{const_cast<LocalIndex*>(nb), nbLen};
    d.wh        = View<const T>{static_cast<const T*>(wh), size_t(sph::lt::kTableSize)};
    d.whd       = View<const T>{static_cast<const T*>(whd), size_t(sph::lt::kTableSize)};
    d.ng0 = p->ng0, d.ngmax = p->ngmax, d.K = p->K, d.Kcour = p->Kcour, d.gamma = p->gamma, d.minDt = p->minDt;
    d.muiConst = T(p->muiConst), d.alphamin = T(p->alphamin), d.alphamax = T(p->alphamax);
    d.decay_constant = T(p->decay_constant), d.Atmin = T(p->Atmin), d.Atmax = T(p->Atmax), d.ramp = T(p->ramp);
    d.eosChoice        = sph::EosType(p->eos);
    d.soundSpeedConst  = T(p->soundSpeedConst);
    d.polytropic_const = p->polytropic_const, d.polytropic_index = p->polytropic_index;
    return d;
}

cstone::Box<double> toBox(const sphx_box_t* b)
{
    auto bt = [](int v)
    {
        return v == 1 ? cstone::BoundaryType::periodic : v == 2 ? cstone::BoundaryType::fixed : cstone::BoundaryType::open;
    };
    return cstone::Box<double>(b->lo[0], b->hi[0], b->lo[1], b->hi[1], b->lo[2], b->hi[2], bt(b->boundary[0]),
                               bt(b->boundary[1]), bt(b->boundary[2]));
}

//! the sequence of domain/test/unit/neighbors/findneighbors.cpp:53-82 (one-rank octreeProperties)
struct Tree
{
    std::vector<KeyT>                           leaves;
    std::vector<unsigned>                       counts;
    cstone::OctreeData<KeyT, cstone::CpuTag>    oct;
    std::vector<LocalIndex>                     layout;
    std::vector<cstone::Vec3<double>>           centers, sizes;
    cstone::OctreeNsView<double, KeyT>          ns;
};

cstone::GroupView oneGroup(LocalIndex first, LocalIndex last)
{
    static thread_local LocalIndex s[2];
    s[0] = first, s[1] = last;
    return cstone::GroupView{first, last, 1, &s[0], &s[1]};
}

} // namespace refglue

using namespace refglue;

extern "C"
{

void ref_sfc_keys(const sphx_box_t* b, const double* x, const double* y, const double* z, uint64_t* keys, size_t n)
{
    cstone::computeSfcKeys(x, y, z, cstone::sfcKindPointer(keys), n, toBox(b));
}

void ref_hilbert_encode(const uint32_t* ix, const uint32_t* iy, const uint32_t* iz, uint64_t* keys, size_t n)
{
    for (size_t i = 0; i < n; ++i)
        keys[i] = cstone::iHilbert<KeyT>(ix[i], iy[i], iz[i]);
}

void ref_hilbert_decode(const uint64_t* keys, uint32_t* ix, uint32_t* iy, uint32_t* iz, size_t n)
{
    for (size_t i = 0; i < n; ++i)
    {
        auto t = cstone::decodeHilbert<KeyT>(keys[i]);
        ix[i] = util::get<0>(t), iy[i] = util::get<1>(t), iz[i] = util::get<2>(t);
    }
}

void ref_sort_keys(uint64_t* keys, uint32_t* order, size_t n)
{
    std::iota(order, order + n, uint32_t(0));
    cstone::sort_by_key(keys, keys + n, order);
}

void* ref_tree_build(const uint64_t* sortedKeys, size_t n, uint32_t bucket, const sphx_box_t* b)
{
    auto* t = new Tree;
    std::tie(t->leaves, t->counts) = cstone::computeOctree<KeyT>(std::span<const KeyT>(sortedKeys, n), bucket);
    t->oct.resize(cstone::nNodes(t->leaves));
    cstone::updateInternalTree<KeyT>(t->leaves, t->oct.data());
    t->layout.assign(cstone::nNodes(t->leaves) + 1, 0);
    std::inclusive_scan(t->counts.begin(), t->counts.end(), t->layout.begin() + 1);
    std::span<const KeyT> nodeKeys(t->oct.prefixes.data(), t->oct.numNodes);
    t->centers.resize(t->oct.numNodes);
    t->sizes.resize(t->oct.numNodes);
    cstone::nodeFpCenters<KeyT>(nodeKeys, t->centers.data(), t->sizes.data(), toBox(b));
    t->ns = cstone::OctreeNsView<double, KeyT>{t->oct.numLeafNodes,         t->oct.prefixes.data(),
                                               t->oct.childOffsets.data(),  t->oct.internalToLeaf.data(),
                                               t->oct.levelRange.data(),    t->leaves.data(),
                                               t->layout.data(),            t->centers.data(),
                                               t->sizes.data()};
    return t;
}

void ref_tree_sizes(const void* tp, int32_t* numLeaf, int32_t* numInternal, int32_t* numNodes)
{
    auto* t      = static_cast<const Tree*>(tp);
    *numLeaf     = t->oct.numLeafNodes;
    *numInternal = t->oct.numInternalNodes;
    *numNodes    = t->oct.numNodes;
}

void ref_tree_copy(const void* tp, uint64_t* leaves, uint32_t* counts, uint32_t* layout, uint64_t* prefixes,
                   int32_t* childOffsets, int32_t* parents, int32_t* internalToLeaf, int32_t* levelRange,
                   double* centers, double* sizes)
{
    auto*  t = static_cast<const Tree*>(tp);
    size_t N = t->oct.numNodes;
    if (leaves) std::memcpy(leaves, t->leaves.data(), t->leaves.size() * 8);
    if (counts) std::memcpy(counts, t->counts.data(), t->counts.size() * 4);
    if (layout) std::memcpy(layout, t->layout.data(), t->layout.size() * 4);
    if (prefixes) std::memcpy(prefixes, t->oct.prefixes.data(), N * 8);
    if (childOffsets) std::memcpy(childOffsets, t->oct.childOffsets.data(), N * 4);
    if (parents) std::memcpy(parents, t->oct.parents.data(), t->oct.parents.size() * 4);
    if (internalToLeaf) std::memcpy(internalToLeaf, t->oct.internalToLeaf.data(), N * 4);
    if (levelRange) std::memcpy(levelRange, t->oct.levelRange.data(), 23 * 4);
    for (size_t i = 0; i < N; ++i)
        for (int d = 0; d < 3; ++d)
        {
            if (centers) centers[3 * i + d] = t->centers[i][d];
            if (sizes) sizes[3 * i + d] = t->sizes[i][d];
        }
}

void ref_tree_free(void* tp) { delete static_cast<Tree*>(tp); }

/* The tree of a one-rank Domain::sync sequence through the reference's own single-rank focused octree
 * (cstone::FocusedOctreeSingleNode, focus/octree_focus.hpp:330-400), the focus spanning the whole key
 * range: ref_focus_update runs one update (rebalance on the previous counts, then counts of `keys`)
 * and returns whether the tree was unchanged. The first sync of a Domain repeats it until then. */
void* ref_focus_new(uint32_t bucket, float theta) { return new cstone::FocusedOctreeSingleNode<KeyT>(bucket, theta); }

int ref_focus_update(void* fp, const sphx_box_t* b, const uint64_t* sortedKeys, size_t n)
{
    auto* f = static_cast<cstone::FocusedOctreeSingleNode<KeyT>*>(fp);
    return f->update(toBox(b), std::span<const KeyT>(sortedKeys, n), KeyT(0), cstone::nodeRange<KeyT>(0),
                     std::span<const KeyT>{})
               ? 1
               : 0;
}

size_t ref_focus_num_leaves(const void* fp)
{
    return static_cast<const cstone::FocusedOctreeSingleNode<KeyT>*>(fp)->leafCounts().size();
}

void ref_focus_copy(const void* fp, uint64_t* leaves, uint32_t* counts)
{
    auto* f = static_cast<const cstone::FocusedOctreeSingleNode<KeyT>*>(fp);
    std::copy(f->treeLeaves().begin(), f->treeLeaves().end(), leaves);
    std::copy(f->leafCounts().begin(), f->leafCounts().end(), counts);
}

void ref_focus_free(void* fp) { delete static_cast<cstone::FocusedOctreeSingleNode<KeyT>*>(fp); }

size_t ref_find_neighbors(const void* tp, const sphx_box_t* b, const sphx_fields_t* f, const sphx_params_t* p,
                          uint32_t first, uint32_t last, int iterateH, uint32_t* nbrs)
{
    auto* t   = static_cast<const Tree*>(tp);
    auto  box = toBox(b);
    auto  run = [&](auto* h)
    {
        if (iterateH)
        {
            sph::findNeighborsSph(f->x, f->y, f->z, h, first, last, box, t->ns, p->ng0, p->ngmax, nbrs,
                                  f->nc + first);
        }
        else
        {
#pragma omp parallel for
            for (LocalIndex k = 0; k < last - first; ++k)
                f->nc[first + k] = 1 + cstone::findNeighbors(first + k, f->x, f->y, f->z, h, t->ns, box, p->ngmax,
                                                             nbrs + size_t(k) * p->ngmax);
        }
    };
    if (p->hydro_type == SPHX_F32) { run(static_cast<float*>(f->h)); }
    else { run(static_cast<double*>(f->h)); }
    return 0;
}

double ref_kernel_K(double sincIndex)
{
    return sph::kernel_3D_k(sph::getSphKernel(sph::SphKernelType::sinc_n, sincIndex), 2.0);
}

double ref_sphynx_K(double n) { return sph::sphynx_3D_k(n); }

void ref_tables_f64(double sincIndex, double* wh, double* whd)
{
    auto a = sph::tabulateFunction<double, sph::lt::kTableSize>(
        sph::getSphKernel(sph::SphKernelType::sinc_n, sincIndex), 0, 2);
    auto b = sph::tabulateFunction<double, sph::lt::kTableSize>(
        sph::getSphKernelDerivative(sph::SphKernelType::sinc_n, sincIndex), 0, 2);
    std::copy(a.begin(), a.end(), wh);
    std::copy(b.begin(), b.end(), whd);
}

void ref_tables_f32(double sincIndex, float* wh, float* whd)
{
    auto a = sph::tabulateFunction<float, sph::lt::kTableSize>(
        sph::getSphKernel(sph::SphKernelType::sinc_n, sincIndex), 0, 2);
    auto b = sph::tabulateFunction<float, sph::lt::kTableSize>(
        sph::getSphKernelDerivative(sph::SphKernelType::sinc_n, sincIndex), 0, 2);
    std::copy(a.begin(), a.end(), wh);
    std::copy(b.begin(), b.end(), whd);
}

double ref_lookup_f64(const double* table, double v) { return sph::lt::lookup(table, v); }

#define REF_DATA(T) auto d = makeData<T>(f, p, size_t(last), nb, size_t(last - first) * p->ngmax, wh, whd)

void ref_density_std(const sphx_box_t* b, const sphx_fields_t* f, const sphx_params_t* p, uint32_t first,
                     uint32_t last, const uint32_t* nb, const void* wh)
{
    const void* whd = wh;
    if (p->hydro_type == SPHX_F32) { REF_DATA(float); sph::computeDensityImpl(oneGroup(first, last), d, toBox(b)); }
    else { REF_DATA(double); sph::computeDensityImpl(oneGroup(first, last), d, toBox(b)); }
}

void ref_eos_std(const sphx_fields_t* f, const sphx_params_t* p, uint32_t first, uint32_t last)
{
    const uint32_t* nb  = nullptr;
    const void*     wh  = nullptr;
    const void*     whd = nullptr;
    if (p->hydro_type == SPHX_F32) { REF_DATA(float); sph::computeEOS_HydroStd(first, last, d); }
    else { REF_DATA(double); sph::computeEOS_HydroStd(first, last, d); }
}

void ref_iad_std(const sphx_box_t* b, const sphx_fields_t* f, const sphx_params_t* p, uint32_t first, uint32_t last,
                 const uint32_t* nb, const void* wh)
{
    const void* whd = wh;
    if (p->hydro_type == SPHX_F32) { REF_DATA(float); sph::computeIADImpl(first, last, d, toBox(b)); }
    else { REF_DATA(double); sph::computeIADImpl(first, last, d, toBox(b)); }
}

double ref_momentum_energy_std(const sphx_box_t* b, const sphx_fields_t* f, const sphx_params_t* p, uint32_t first,
                               uint32_t last, const uint32_t* nb, const void* wh, void* /*maxvs*/)
{
    const void* whd = wh;
    if (p->hydro_type == SPHX_F32)
    {
        REF_DATA(float);
        sph::computeMomentumEnergyStdImpl(first, last, d, toBox(b));
        return d.minDtCourant;
    }
    REF_DATA(double);
    sph::computeMomentumEnergyStdImpl(first, last, d, toBox(b));
    return d.minDtCourant;
}

void ref_xmass(const sphx_box_t* b, const sphx_fields_t* f, const sphx_params_t* p, uint32_t first, uint32_t last,
               const uint32_t* nb, const void* wh)
{
    const void* whd = wh;
    if (p->hydro_type == SPHX_F32) { REF_DATA(float); sph::computeXMassImpl(first, last, d, toBox(b)); }
    else { REF_DATA(double); sph::computeXMassImpl(first, last, d, toBox(b)); }
}

void ref_ve_def_gradh(const sphx_box_t* b, const sphx_fields_t* f, const sphx_params_t* p, uint32_t first,
                      uint32_t last, const uint32_t* nb, const void* wh, const void* whd)
{
    if (p->hydro_type == SPHX_F32) { REF_DATA(float); sph::computeVeDefGradhImpl(first, last, d, toBox(b)); }
    else { REF_DATA(double); sph::computeVeDefGradhImpl(first, last, d, toBox(b)); }
}

void ref_eos_ve(const sphx_fields_t* f, const sphx_params_t* p, uint32_t first, uint32_t last)
{
    const uint32_t* nb  = nullptr;
    const void*     wh  = nullptr;
    const void*     whd = nullptr;
    // the VE driver selects the temperature branch when u is empty (hydro_ve/eos.hpp:68)
    sphx_fields_t g = *f;
    if (g.temp != nullptr) { g.u = nullptr; }
    f = &g;
    // computeEOS's selection (hydro_ve/eos.hpp:191-197) among the CPU loops
    auto run = [&](auto& d)
    {
        if (p->eos == SPHX_EOS_ISOTHERMAL) { sph::computeIsothermalEOS_Impl(first, last, d); }
        else if (p->eos == SPHX_EOS_POLYTROPIC) { sph::computePolytropicEOS_Impl(first, last, d); }
        else { sph::computeIdealGasEOS_Impl(first, last, d); }
    };
    if (p->hydro_type == SPHX_F32) { REF_DATA(float); run(d); }
    else { REF_DATA(double); run(d); }
}

void ref_iad_divv_curlv(const sphx_box_t* b, const sphx_fields_t* f, const sphx_params_t* p, uint32_t first,
                        uint32_t last, const uint32_t* nb, const void* wh)
{
    const void* whd = wh;
    if (p->hydro_type == SPHX_F32) { REF_DATA(float); sph::computeIadDivvCurlvImpl(first, last, d, toBox(b)); }
    else { REF_DATA(double); sph::computeIadDivvCurlvImpl(first, last, d, toBox(b)); }
}

void ref_av_switches(const sphx_box_t* b, const sphx_fields_t* f, const sphx_params_t* p, uint32_t first,
                     uint32_t last, const uint32_t* nb, const void* wh)
{
    const void* whd = wh;
    if (p->hydro_type == SPHX_F32) { REF_DATA(float); sph::computeAVswitchesImpl(first, last, d, toBox(b)); }
    else { REF_DATA(double); sph::computeAVswitchesImpl(first, last, d, toBox(b)); }
}

double ref_momentum_energy_ve(const sphx_box_t* b, const sphx_fields_t* f, const sphx_params_t* p, uint32_t first,
                              uint32_t last, const uint32_t* nb, const void* wh, int avClean, void* /*maxvs*/)
{
    const void* whd = wh;
    auto        box = toBox(b);
    if (p->hydro_type == SPHX_F32)
    {
        REF_DATA(float);
        if (avClean) sph::computeMomentumEnergyImpl<true>(first, last, d, box);
        else sph::computeMomentumEnergyImpl<false>(first, last, d, box);
        return d.minDtCourant;
    }
    REF_DATA(double);
    if (avClean) sph::computeMomentumEnergyImpl<true>(first, last, d, box);
    else sph::computeMomentumEnergyImpl<false>(first, last, d, box);
    return d.minDtCourant;
}

// ---- time integration: the reference's element functions (sph/include/sph/positions.hpp) -------

double ref_energy_update(double u_old, double dt, double dt_m1, double du, double du_m1)
{
    return sph::energyUpdate(u_old, dt, dt_m1, du, du_m1);
}

void ref_position_update(double dt, double dt_m1, const double* X, const double* A, const double* dX,
                         const sphx_box_t* b, double* Xo, double* Vo, double* dXo)
{
    cstone::Vec3<double> Xn{X[0], X[1], X[2]}, An{A[0], A[1], A[2]}, dXn{dX[0], dX[1], dX[2]};
    auto [Xnp1, Vnp1, dXnp1] = sph::positionUpdate(dt, dt_m1, Xn, An, dXn, toBox(b));
    for (int d = 0; d < 3; ++d)
        Xo[d] = Xnp1[d], Vo[d] = Vnp1[d], dXo[d] = dXnp1[d];
}

int ref_fbc_check_f32(double coord, float h, double top, double bottom, int fbc)
{
    return sph::fbcCheck(coord, h, top, bottom, fbc != 0);
}

//! cstone::limitBoxShrinking (domain/include/cstone/sfc/box.hpp:398-412), default shrink limit
void ref_limit_box_shrinking(const double* fit6, const double* prev6, double* out6)
{
    cstone::Box<double> fit(fit6[0], fit6[1], fit6[2], fit6[3], fit6[4], fit6[5]);
    cstone::Box<double> prev(prev6[0], prev6[1], prev6[2], prev6[3], prev6[4], prev6[5]);
    auto                out = cstone::limitBoxShrinking(fit, prev);
    double              r[6] = {out.xmin(), out.xmax(), out.ymin(), out.ymax(), out.zmin(), out.zmax()};
    for (int k = 0; k < 6; ++k)
        out6[k] = r[k];
}

//! updateSmoothingLengthCpu (sph/include/sph/update_h.hpp:12-24)
void ref_update_smoothing_length(const sphx_fields_t* f, const sphx_params_t* p, uint32_t first, uint32_t last)
{
    if (p->hydro_type == SPHX_F32) { sph::updateSmoothingLengthCpu(first, last, p->ng0, f->nc, (float*)f->h); }
    else { sph::updateSmoothingLengthCpu(first, last, p->ng0, f->nc, (double*)f->h); }
}

} // extern "C"
