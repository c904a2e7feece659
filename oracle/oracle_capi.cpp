/*
 * C ABI of the CPU oracle (TEST INFRASTRUCTURE ONLY: loaded by tests/, the smoke check and the
 * CPU-baseline leg of bench.py through ctypes; never linked into the product library).
 *
 * Host pointers throughout. The particle view structs are the product's own
 * (include/sphexa_amd.h) so that one Python view feeds both sides with identical data.
 * Drivers follow the reference OpenMP drivers cited on each function; the neighbour list of
 * target i (first <= i < last) is row (i - first) of `nbrs` with `ngmax` columns, exactly like
 * the CPU Dataset's d.neighbors (sph/include/sph/particles_data.hpp:250-251).
 */
#include <chrono>
#include <cstring>
#include <memory>

#include "sphexa_amd.h"
#include "sphx_oracle.hpp"

using namespace orc;

namespace
{

Box toBox(const sphx_box_t* b)
{
    double lim[6] = {b->lo[0], b->hi[0], b->lo[1], b->hi[1], b->lo[2], b->hi[2]};
    int    bc[3]  = {b->boundary[0], b->boundary[1], b->boundary[2]};
    return Box(lim, bc);
}

template<class T>
T* F(void* p)
{
    return static_cast<T*>(p);
}

template<class T>
unsigned capped(const sphx_fields_t* f, LocalIndex i, unsigned ngmax)
{
    return std::min(f->nc[i] - 1, ngmax);
}

//! computeDensity (hydro_std/density.hpp:42-53): xm in rho's slot, then rho = m / xm
template<class T>
void densityStd(const Box& box, const sphx_fields_t* f, const sphx_params_t* p, LocalIndex first, LocalIndex last,
                const LocalIndex* nb, const T* wh)
{
    T* rho = F<T>(f->rho);
    T* h   = F<T>(f->h);
    T* m   = F<T>(f->m);
#pragma omp parallel for
    for (LocalIndex i = first; i < last; ++i)
    {
        unsigned nn = capped<T>(f, i, p->ngmax);
        rho[i]      = xmassLoop(i, p->K, box, nb + size_t(i - first) * p->ngmax, nn, f->x, f->y, f->z, h, m, wh);
    }
#pragma omp parallel for schedule(static)
    for (LocalIndex i = first; i < last; ++i)
        rho[i] = m[i] / rho[i];
}

//! computeXMassImpl, hydro_ve/xmass.hpp:40-65
template<class T>
void xmass(const Box& box, const sphx_fields_t* f, const sphx_params_t* p, LocalIndex first, LocalIndex last,
           const LocalIndex* nb, const T* wh)
{
    T* xm = F<T>(f->xm);
#pragma omp parallel for
    for (LocalIndex i = first; i < last; ++i)
    {
        unsigned nn = capped<T>(f, i, p->ngmax);
        xm[i] = xmassLoop(i, p->K, box, nb + size_t(i - first) * p->ngmax, nn, f->x, f->y, f->z, F<T>(f->h),
                          F<T>(f->m), wh);
    }
}

/* computeEOS_HydroStd, hydro_std/eos.hpp:146-152: the isothermal and polytropic branches
 * (computeIsothermalEOS_HydroStd_Impl :82-97, computePolytropicEOS_HydroStd_Impl :99-114). The isothermal
 * branch also stores c = c0, as the reference's GPU kernel does (hydro_std/eos_gpu.cu:73-82); its CPU
 * loop leaves c unchanged. */
template<class T>
void eosStdVariant(const sphx_fields_t* f, const sphx_params_t* prm, LocalIndex first, LocalIndex last)
{
    const T* rho = F<T>(f->rho);
    T*       p   = F<T>(f->p);
    T*       c   = F<T>(f->c);
#pragma omp parallel for schedule(static)
    for (LocalIndex i = first; i < last; ++i)
    {
        if (prm->eos == SPHX_EOS_ISOTHERMAL)
        {
            const T c0 = T(prm->soundSpeedConst);
            p[i]       = isothermalP(c0, rho[i]);
            c[i]       = c0;
        }
        else
        {
            double pi, ci;
            polytropicPC(prm->polytropic_const, prm->polytropic_index, rho[i], pi, ci);
            p[i] = pi;
            c[i] = ci;
        }
        if (f->temp) { f->temp[i] = 0; }
    }
}

//! computeIdealGasEOS_HydroStd_Impl, hydro_std/eos.hpp:55-80
template<class T>
void eosStd(const sphx_fields_t* f, const sphx_params_t* prm, LocalIndex first, LocalIndex last)
{
    if (prm->eos != SPHX_EOS_IDEAL_GAS) { return eosStdVariant<T>(f, prm, first, last); }
    T* rho = F<T>(f->rho);
    T* p   = F<T>(f->p);
    T* c   = F<T>(f->c);
#pragma omp parallel for schedule(static)
    for (LocalIndex i = first; i < last; ++i)
    {
        double pi, ci;
        if (f->u != nullptr) { idealGasFromU(f->u[i], rho[i], prm->gamma, pi, ci); }
        else
        {
            T    mui = T(prm->muiConst);
            auto cv  = idealGasCv(mui, prm->gamma);
            idealGasFromU(cv * f->temp[i], rho[i], prm->gamma, pi, ci);
        }
        p[i] = pi;
        c[i] = ci;
    }
}

//! computeIADImpl, hydro_std/iad.hpp:42-72
template<class T>
void iadStd(const Box& box, const sphx_fields_t* f, const sphx_params_t* p, LocalIndex first, LocalIndex last,
            const LocalIndex* nb, const T* wh)
{
    const T* m   = F<T>(f->m);
    const T* rho = F<T>(f->rho);
    auto     vol = [m, rho](LocalIndex j) { return m[j] / rho[j]; };
#pragma omp parallel for schedule(static)
    for (LocalIndex i = first; i < last; ++i)
    {
        unsigned nn = capped<T>(f, i, p->ngmax);
        iadLoop(i, p->K, box, nb + size_t(i - first) * p->ngmax, nn, f->x, f->y, f->z, F<T>(f->h), wh, vol,
                F<T>(f->c11), F<T>(f->c12), F<T>(f->c13), F<T>(f->c22), F<T>(f->c23), F<T>(f->c33));
    }
}

//! computeMomentumEnergyStdImpl, hydro_std/momentum_energy.hpp:42-95
template<class T>
double momentumEnergyStd(const Box& box, const sphx_fields_t* f, const sphx_params_t* p, LocalIndex first,
                         LocalIndex last, const LocalIndex* nb, const T* wh, T* maxvsOut)
{
    T minDt = INFINITY;
    const T* h = F<T>(f->h);
    const T* c = F<T>(f->c);
#pragma omp parallel for schedule(static) reduction(min : minDt)
    for (LocalIndex i = first; i < last; ++i)
    {
        T        maxvs = 0;
        unsigned nn    = capped<T>(f, i, p->ngmax);
        momentumEnergyStdLoop(i, p->K, box, nb + size_t(i - first) * p->ngmax, nn, f->x, f->y, f->z, F<T>(f->vx),
                              F<T>(f->vy), F<T>(f->vz), h, F<T>(f->m), F<T>(f->rho), F<T>(f->p), c,
                              F<T>(f->c11), F<T>(f->c12), F<T>(f->c13), F<T>(f->c22), F<T>(f->c23), F<T>(f->c33),
                              wh, F<T>(f->ax), F<T>(f->ay), F<T>(f->az), f->du, &maxvs,
                              (p->flags & SPHX_FLAG_STD_ALPHA) ? F<T>(f->alpha) : (const T*)nullptr);
        if (maxvsOut) { maxvsOut[i] = maxvs; }
        T dt_i = courantDt(maxvs, h[i], c[i], float(p->Kcour));
        minDt  = std::min(minDt, dt_i);
    }
    return minDt;
}

/* Adjoints: serial over targets (neighbour gradients are scattered), lists as above. */
template<class T>
void densityStdVjpAll(const Box& box, const sphx_fields_t* f, const sphx_params_t* p, LocalIndex first,
                      LocalIndex last, const LocalIndex* nb, const T* wh, const T* whd, const sphx_adjoint_t* in,
                      sphx_adjoint_t* out)
{
    for (LocalIndex i = first; i < last; ++i)
    {
        unsigned nn = capped<T>(f, i, p->ngmax);
        densityStdVjp(i, p->K, box, nb + size_t(i - first) * p->ngmax, nn, f->x, f->y, f->z, F<T>(f->h), F<T>(f->m),
                      wh, whd, F<T>(in->rho)[i], F<T>(out->x), F<T>(out->y), F<T>(out->z), F<T>(out->h));
    }
}

template<class T>
void iadStdVjpAll(const Box& box, const sphx_fields_t* f, const sphx_params_t* p, LocalIndex first, LocalIndex last,
                  const LocalIndex* nb, const T* wh, const T* whd, const sphx_adjoint_t* in, sphx_adjoint_t* out)
{
    const T* c[6]  = {F<T>(f->c11), F<T>(f->c12), F<T>(f->c13), F<T>(f->c22), F<T>(f->c23), F<T>(f->c33)};
    const T* gc[6] = {F<T>(in->c11), F<T>(in->c12), F<T>(in->c13), F<T>(in->c22), F<T>(in->c23), F<T>(in->c33)};
    for (LocalIndex i = first; i < last; ++i)
    {
        unsigned nn = capped<T>(f, i, p->ngmax);
        T        C6[6], G6[6];
        for (int q = 0; q < 6; ++q)
            C6[q] = c[q][i], G6[q] = gc[q][i];
        iadStdVjp(i, p->K, box, nb + size_t(i - first) * p->ngmax, nn, f->x, f->y, f->z, F<T>(f->h), F<T>(f->m),
                  F<T>(f->rho), C6, G6, wh, whd, F<T>(out->x), F<T>(out->y), F<T>(out->z), F<T>(out->h),
                  F<T>(out->rho));
    }
}

template<class T>
void momentumEnergyStdVjpAll(const Box& box, const sphx_fields_t* f, const sphx_params_t* p, LocalIndex first,
                             LocalIndex last, const LocalIndex* nb, const T* wh, const T* whd,
                             const sphx_adjoint_t* in, sphx_adjoint_t* out)
{
    const T*       C6[6] = {F<T>(f->c11), F<T>(f->c12), F<T>(f->c13), F<T>(f->c22), F<T>(f->c23), F<T>(f->c33)};
    const T*       alpha = (p->flags & SPHX_FLAG_STD_ALPHA) ? F<T>(f->alpha) : nullptr;
    MomStdGrads<T> g{F<T>(out->x),   F<T>(out->y), F<T>(out->z), F<T>(out->h),
                     F<T>(out->rho), F<T>(out->p), F<T>(out->alpha),
                     {F<T>(out->c11), F<T>(out->c12), F<T>(out->c13), F<T>(out->c22), F<T>(out->c23),
                      F<T>(out->c33)}};
    for (LocalIndex i = first; i < last; ++i)
    {
        unsigned nn    = capped<T>(f, i, p->ngmax);
        T        ga[3] = {F<T>(in->ax)[i], F<T>(in->ay)[i], F<T>(in->az)[i]};
        T        gdu   = in->du ? F<T>(in->du)[i] : T(0);
        momentumEnergyStdVjp(i, p->K, box, nb + size_t(i - first) * p->ngmax, nn, f->x, f->y, f->z, F<T>(f->vx),
                             F<T>(f->vy), F<T>(f->vz), F<T>(f->h), F<T>(f->m), F<T>(f->rho), F<T>(f->p),
                             F<T>(f->c), C6, alpha, wh, whd, ga, gdu, g);
    }
}

//! computeVeDefGradhImpl, hydro_ve/ve_def_gradh.hpp:40-77
template<class T>
void veDefGradh(const Box& box, const sphx_fields_t* f, const sphx_params_t* p, LocalIndex first, LocalIndex last,
                const LocalIndex* nb, const T* wh, const T* whd)
{
    T* kx    = F<T>(f->kx);
    T* gradh = F<T>(f->gradh);
#pragma omp parallel for
    for (LocalIndex i = first; i < last; ++i)
    {
        unsigned nn = capped<T>(f, i, p->ngmax);
        T        k, g;
        veDefGradhLoop(i, p->K, box, nb + size_t(i - first) * p->ngmax, nn, f->x, f->y, f->z, F<T>(f->h),
                       F<T>(f->m), wh, whd, F<T>(f->xm), k, g);
        kx[i]    = k;
        gradh[i] = g;
    }
}

/* computeIsothermalEOS_Impl and computePolytropicEOS_Impl, hydro_ve/eos.hpp:97-123 and :126-151, as
 * computeEOS selects them (:191-197) */
template<class T>
void eosVeVariant(const sphx_fields_t* f, const sphx_params_t* prm, LocalIndex first, LocalIndex last)
{
    const T* m     = F<T>(f->m);
    const T* kx    = F<T>(f->kx);
    const T* xm    = F<T>(f->xm);
    const T* gradh = F<T>(f->gradh);
    T*       prho  = F<T>(f->prho);
    T*       c     = F<T>(f->c);
    T*       rhoO  = F<T>(f->rho);
    T*       pO    = F<T>(f->p);
#pragma omp parallel for schedule(static)
    for (LocalIndex i = first; i < last; ++i)
    {
        auto rho = kx[i] * m[i] / xm[i];
        if (prm->eos == SPHX_EOS_ISOTHERMAL)
        {
            const T c0 = T(prm->soundSpeedConst);
            auto    pi = isothermalP(c0, rho);
            prho[i]    = pi / (kx[i] * m[i] * m[i] * gradh[i]);
            c[i]       = c0;
            if (pO) { pO[i] = pi; }
        }
        else
        {
            double pi, ci;
            polytropicPC(prm->polytropic_const, prm->polytropic_index, rho, pi, ci);
            prho[i] = pi / (kx[i] * m[i] * m[i] * gradh[i]);
            c[i]    = ci;
            if (pO) { pO[i] = pi; }
        }
        if (rhoO) { rhoO[i] = rho; }
        if (f->temp) { f->temp[i] = 0; }
    }
}

//! computeIdealGasEOS_Impl, hydro_ve/eos.hpp:53-95 (temp-based when temp != NULL, else u-based)
template<class T>
void eosVe(const sphx_fields_t* f, const sphx_params_t* prm, LocalIndex first, LocalIndex last)
{
    if (prm->eos != SPHX_EOS_IDEAL_GAS) { return eosVeVariant<T>(f, prm, first, last); }
    const T* m     = F<T>(f->m);
    const T* kx    = F<T>(f->kx);
    const T* xm    = F<T>(f->xm);
    const T* gradh = F<T>(f->gradh);
    T*       prho  = F<T>(f->prho);
    T*       c     = F<T>(f->c);
    T*       rhoO  = F<T>(f->rho);
    T*       pO    = F<T>(f->p);
#pragma omp parallel for schedule(static)
    for (LocalIndex i = first; i < last; ++i)
    {
        auto   rho = kx[i] * m[i] / xm[i];
        double pi, ci;
        if (f->temp != nullptr)
        {
            T    mui = T(prm->muiConst);
            auto cv  = idealGasCv(mui, prm->gamma);
            idealGasFromU(cv * f->temp[i], rho, prm->gamma, pi, ci);
        }
        else { idealGasFromU(f->u[i], rho, prm->gamma, pi, ci); }
        prho[i] = pi / (kx[i] * m[i] * m[i] * gradh[i]);
        c[i]    = ci;
        if (rhoO) { rhoO[i] = rho; }
        if (pO) { pO[i] = pi; }
    }
}

//! computeIadDivvCurlvImpl, hydro_ve/iad_divv_curlv.hpp:42-90
template<class T>
void iadDivvCurlv(const Box& box, const sphx_fields_t* f, const sphx_params_t* p, LocalIndex first, LocalIndex last,
                  const LocalIndex* nb, const T* wh)
{
    const T* xm      = F<T>(f->xm);
    const T* kx      = F<T>(f->kx);
    auto     vol     = [xm, kx](LocalIndex j) { return xm[j] / kx[j]; };
    bool     doGradV = f->dV11 != nullptr;
#pragma omp parallel for
    for (LocalIndex i = first; i < last; ++i)
    {
        unsigned          nn = capped<T>(f, i, p->ngmax);
        const LocalIndex* ni = nb + size_t(i - first) * p->ngmax;
        iadLoop(i, p->K, box, ni, nn, f->x, f->y, f->z, F<T>(f->h), wh, vol, F<T>(f->c11), F<T>(f->c12),
                F<T>(f->c13), F<T>(f->c22), F<T>(f->c23), F<T>(f->c33));
        divvCurlvLoop(i, p->K, box, ni, nn, f->x, f->y, f->z, F<T>(f->vx), F<T>(f->vy), F<T>(f->vz), F<T>(f->h),
                      F<T>(f->c11), F<T>(f->c12), F<T>(f->c13), F<T>(f->c22), F<T>(f->c23), F<T>(f->c33), wh, kx, xm,
                      F<T>(f->divv), F<T>(f->curlv), F<T>(f->dV11), F<T>(f->dV12), F<T>(f->dV13), F<T>(f->dV22),
                      F<T>(f->dV23), F<T>(f->dV33), doGradV);
    }
}

//! computeAVswitchesImpl, hydro_ve/av_switches.hpp:41-78
template<class T>
void avSwitches(const Box& box, const sphx_fields_t* f, const sphx_params_t* p, LocalIndex first, LocalIndex last,
                const LocalIndex* nb, const T* wh)
{
    T* alpha = F<T>(f->alpha);
#pragma omp parallel for
    for (LocalIndex i = first; i < last; ++i)
    {
        unsigned nn = capped<T>(f, i, p->ngmax);
        alpha[i] = avSwitchesLoop(i, p->K, box, nb + size_t(i - first) * p->ngmax, nn, f->x, f->y, f->z, F<T>(f->vx),
                                  F<T>(f->vy), F<T>(f->vz), F<T>(f->h), F<T>(f->c), F<T>(f->c11), F<T>(f->c12),
                                  F<T>(f->c13), F<T>(f->c22), F<T>(f->c23), F<T>(f->c33), wh, F<T>(f->kx),
                                  F<T>(f->xm), F<T>(f->divv), p->minDt, T(p->alphamin), T(p->alphamax),
                                  T(p->decay_constant), alpha[i]);
    }
}

//! computeMomentumEnergyImpl<avClean>, hydro_ve/momentum_energy.hpp:41-104
template<bool avClean, class T>
double momentumEnergyVe(const Box& box, const sphx_fields_t* f, const sphx_params_t* p, LocalIndex first,
                        LocalIndex last, const LocalIndex* nb, const T* wh, T* maxvsOut)
{
    T        minDt = INFINITY;
    const T* h     = F<T>(f->h);
    const T* c     = F<T>(f->c);
#pragma omp parallel for schedule(static) reduction(min : minDt)
    for (LocalIndex i = first; i < last; ++i)
    {
        unsigned nn    = capped<T>(f, i, p->ngmax);
        T        maxvs = 0;
        momentumEnergyVeLoop<avClean>(i, p->K, box, nb + size_t(i - first) * p->ngmax, nn, f->x, f->y, f->z,
                                      F<T>(f->vx), F<T>(f->vy), F<T>(f->vz), h, F<T>(f->m), F<T>(f->prho),
                                      F<T>(f->tdpdTrho), c, F<T>(f->c11), F<T>(f->c12), F<T>(f->c13), F<T>(f->c22),
                                      F<T>(f->c23), F<T>(f->c33), T(p->Atmin), T(p->Atmax), T(p->ramp), wh,
                                      F<T>(f->kx), F<T>(f->xm), F<T>(f->alpha), F<T>(f->dV11), F<T>(f->dV12),
                                      F<T>(f->dV13), F<T>(f->dV22), F<T>(f->dV23), F<T>(f->dV33), F<T>(f->ax),
                                      F<T>(f->ay), F<T>(f->az), f->du, &maxvs);
        if (maxvsOut) { maxvsOut[i] = maxvs; }
        T dt_i = courantDt(maxvs, h[i], c[i], float(p->Kcour));
        minDt  = std::min(minDt, dt_i);
    }
    return minDt;
}

} // namespace

extern "C"
{

/* ---- keys / sort ------------------------------------------------------------------------------ */
void orc_sfc_keys(const sphx_box_t* b, const double* x, const double* y, const double* z, uint64_t* keys, size_t n)
{
    computeSfcKeys(x, y, z, keys, n, toBox(b));
}

void orc_hilbert_encode(const uint32_t* ix, const uint32_t* iy, const uint32_t* iz, uint64_t* keys, size_t n)
{
    for (size_t i = 0; i < n; ++i)
        keys[i] = hilbertEncode(ix[i], iy[i], iz[i]);
}

void orc_hilbert_decode(const uint64_t* keys, uint32_t* ix, uint32_t* iy, uint32_t* iz, size_t n)
{
    for (size_t i = 0; i < n; ++i)
        hilbertDecode(keys[i], ix[i], iy[i], iz[i]);
}

//! keys sorted in place (stable); order receives the original index of each sorted key
void orc_sort_keys(uint64_t* keys, uint32_t* order, size_t n)
{
    for (size_t i = 0; i < n; ++i)
        order[i] = uint32_t(i);
    stableSortByKey(keys, order, n);
}

/* ---- tree ------------------------------------------------------------------------------------ */
void* orc_tree_build(const uint64_t* sortedKeys, size_t n, uint32_t bucket, const sphx_box_t* b)
{
    auto* t = new NsTree;
    buildNsTree(sortedKeys, n, bucket, toBox(b), *t);
    return t;
}

/*! spatial groups of [first, last) on the tree `tp` (groupSplits): writes the group starts followed by
 * `last` into out (capacity cap) and returns their count, numGroups + 1 (0 when cap is too small) */
size_t orc_group_splits(const void* tp, const sphx_box_t* b, const double* x, const double* y, const double* z,
                        uint32_t first, uint32_t last, float tolFactor, uint32_t* out, size_t cap)
{
    auto*                   t = static_cast<const NsTree*>(tp);
    std::vector<LocalIndex> g;
    groupSplits(first, last, x, y, z, t->leaves, t->layout, toBox(b), tolFactor, g);
    if (g.size() > cap) { return 0; }
    std::copy(g.begin(), g.end(), out);
    return g.size();
}

void orc_tree_sizes(const void* tp, int32_t* numLeaf, int32_t* numInternal, int32_t* numNodes)
{
    auto* t      = static_cast<const NsTree*>(tp);
    *numLeaf     = t->oct.numLeafNodes;
    *numInternal = t->oct.numInternalNodes;
    *numNodes    = t->oct.numNodes;
}

void orc_tree_copy(const void* tp, uint64_t* leaves, uint32_t* counts, uint32_t* layout, uint64_t* prefixes,
                   int32_t* childOffsets, int32_t* parents, int32_t* internalToLeaf, int32_t* levelRange,
                   double* centers, double* sizes)
{
    auto* t = static_cast<const NsTree*>(tp);
    auto  N = size_t(t->oct.numNodes);
    if (leaves) std::memcpy(leaves, t->leaves.data(), t->leaves.size() * 8);
    if (counts) std::memcpy(counts, t->counts.data(), t->counts.size() * 4);
    if (layout) std::memcpy(layout, t->layout.data(), t->layout.size() * 4);
    if (prefixes) std::memcpy(prefixes, t->oct.prefixes.data(), N * 8);
    if (childOffsets) std::memcpy(childOffsets, t->oct.childOffsets.data(), N * 4);
    if (parents) std::memcpy(parents, t->oct.parents.data(), t->oct.parents.size() * 4);
    if (internalToLeaf) std::memcpy(internalToLeaf, t->oct.internalToLeaf.data(), N * 4);
    if (levelRange) std::memcpy(levelRange, t->oct.levelRange.data(), (kMaxLevel + 2) * 4);
    if (centers) std::memcpy(centers, t->centers.data(), 3 * N * 8);
    if (sizes) std::memcpy(sizes, t->sizes.data(), 3 * N * 8);
}

void orc_tree_free(void* tp) { delete static_cast<NsTree*>(tp); }

/*! the tree update of a later sync (updateNsTree) on the tree `tp`, in place, with the new sorted keys;
 * returns 1 when no node changed */
int orc_tree_update(void* tp, const uint64_t* sortedKeys, size_t n, uint32_t bucket, const sphx_box_t* b)
{
    return updateNsTree(sortedKeys, n, bucket, toBox(b), *static_cast<NsTree*>(tp)) ? 1 : 0;
}

/* ---- neighbours ------------------------------------------------------------------------------ */
//! findNeighborsSfc -> findNeighborsSph (sph/include/sph/find_neighbors.hpp:11-56); nc[i], h[i]
//! for first <= i < last. Returns the number of particles whose h/nc iteration failed.
size_t orc_find_neighbors(const void* tp, const sphx_box_t* b, const sphx_fields_t* f, const sphx_params_t* p,
                          uint32_t first, uint32_t last, int iterateH, uint32_t* nbrs)
{
    auto* t   = static_cast<const NsTree*>(tp);
    Box   box = toBox(b);
    if (p->hydro_type == SPHX_F32)
        return findNeighborsSph(f->x, f->y, f->z, F<float>(f->h), first, last, box, *t, p->ng0, p->ngmax, nbrs,
                                f->nc + first, iterateH != 0);
    return findNeighborsSph(f->x, f->y, f->z, F<double>(f->h), first, last, box, *t, p->ng0, p->ngmax, nbrs,
                            f->nc + first, iterateH != 0);
}

//! distanceSq<true> (domain/include/cstone/findneighbors.hpp:33-48)
double orc_distance_sq_pbc(const sphx_box_t* b, double x1, double y1, double z1, double x2, double y2, double z2)
{
    return distSq<true>(x1, y1, z1, x2, y2, z2, toBox(b));
}

//! util::simpson of sin over [0, pi/2] with 2000 intervals (sph_kernel_tables.hpp:31-60)
double orc_simpson_sin() { return simpson(0, M_PI / 2, 2000, [](double x) { return std::sin(x); }); }

/* ---- tables ---------------------------------------------------------------------------------- */
double orc_kernel_K(double sincIndex) { return kernelNormalization(sincIndex); }
double orc_sphynx_K(double n) { return sphynxK(n); }

void orc_tables_f64(double sincIndex, double* wh, double* whd)
{
    tabulate<double>([sincIndex](double v) { return sincKernel(v, sincIndex); }, wh);
    tabulate<double>([sincIndex](double v) { return sincKernelDerivative(v, sincIndex); }, whd);
}

void orc_tables_f32(double sincIndex, float* wh, float* whd)
{
    tabulate<float>([sincIndex](double v) { return sincKernel(v, sincIndex); }, wh);
    tabulate<float>([sincIndex](double v) { return sincKernelDerivative(v, sincIndex); }, whd);
}

double orc_lookup_f64(const double* table, double v) { return lookup(table, v); }

/* ---- SPH passes (wh/whd in the hydro type) --------------------------------------------------- */
#define ORC_DISPATCH(call)                                                                                             \
    if (p->hydro_type == SPHX_F32) { call(float); }                                                                    \
    else { call(double); }

void orc_density_std(const sphx_box_t* b, const sphx_fields_t* f, const sphx_params_t* p, uint32_t first,
                     uint32_t last, const uint32_t* nb, const void* wh)
{
#define C(T) densityStd<T>(toBox(b), f, p, first, last, nb, (const T*)wh)
    ORC_DISPATCH(C)
#undef C
}

void orc_eos_std(const sphx_fields_t* f, const sphx_params_t* p, uint32_t first, uint32_t last)
{
#define C(T) eosStd<T>(f, p, first, last)
    ORC_DISPATCH(C)
#undef C
}

void orc_iad_std(const sphx_box_t* b, const sphx_fields_t* f, const sphx_params_t* p, uint32_t first, uint32_t last,
                 const uint32_t* nb, const void* wh)
{
#define C(T) iadStd<T>(toBox(b), f, p, first, last, nb, (const T*)wh)
    ORC_DISPATCH(C)
#undef C
}

double orc_momentum_energy_std(const sphx_box_t* b, const sphx_fields_t* f, const sphx_params_t* p, uint32_t first,
                               uint32_t last, const uint32_t* nb, const void* wh, void* maxvs)
{
    if (p->hydro_type == SPHX_F32)
        return momentumEnergyStd<float>(toBox(b), f, p, first, last, nb, (const float*)wh, (float*)maxvs);
    return momentumEnergyStd<double>(toBox(b), f, p, first, last, nb, (const double*)wh, (double*)maxvs);
}

void orc_density_std_vjp(const sphx_box_t* b, const sphx_fields_t* f, const sphx_params_t* p, uint32_t first,
                         uint32_t last, const uint32_t* nb, const void* wh, const void* whd, const sphx_adjoint_t* in,
                         sphx_adjoint_t* out)
{
#define C(T) densityStdVjpAll<T>(toBox(b), f, p, first, last, nb, (const T*)wh, (const T*)whd, in, out)
    ORC_DISPATCH(C)
#undef C
}

void orc_iad_std_vjp(const sphx_box_t* b, const sphx_fields_t* f, const sphx_params_t* p, uint32_t first,
                     uint32_t last, const uint32_t* nb, const void* wh, const void* whd, const sphx_adjoint_t* in,
                     sphx_adjoint_t* out)
{
#define C(T) iadStdVjpAll<T>(toBox(b), f, p, first, last, nb, (const T*)wh, (const T*)whd, in, out)
    ORC_DISPATCH(C)
#undef C
}

void orc_momentum_energy_std_vjp(const sphx_box_t* b, const sphx_fields_t* f, const sphx_params_t* p,
                                 uint32_t first, uint32_t last, const uint32_t* nb, const void* wh, const void* whd,
                                 const sphx_adjoint_t* in, sphx_adjoint_t* out)
{
#define C(T) momentumEnergyStdVjpAll<T>(toBox(b), f, p, first, last, nb, (const T*)wh, (const T*)whd, in, out)
    ORC_DISPATCH(C)
#undef C
}

void orc_xmass(const sphx_box_t* b, const sphx_fields_t* f, const sphx_params_t* p, uint32_t first, uint32_t last,
               const uint32_t* nb, const void* wh)
{
#define C(T) xmass<T>(toBox(b), f, p, first, last, nb, (const T*)wh)
    ORC_DISPATCH(C)
#undef C
}

void orc_ve_def_gradh(const sphx_box_t* b, const sphx_fields_t* f, const sphx_params_t* p, uint32_t first,
                      uint32_t last, const uint32_t* nb, const void* wh, const void* whd)
{
#define C(T) veDefGradh<T>(toBox(b), f, p, first, last, nb, (const T*)wh, (const T*)whd)
    ORC_DISPATCH(C)
#undef C
}

void orc_eos_ve(const sphx_fields_t* f, const sphx_params_t* p, uint32_t first, uint32_t last)
{
#define C(T) eosVe<T>(f, p, first, last)
    ORC_DISPATCH(C)
#undef C
}

void orc_iad_divv_curlv(const sphx_box_t* b, const sphx_fields_t* f, const sphx_params_t* p, uint32_t first,
                        uint32_t last, const uint32_t* nb, const void* wh)
{
#define C(T) iadDivvCurlv<T>(toBox(b), f, p, first, last, nb, (const T*)wh)
    ORC_DISPATCH(C)
#undef C
}

//! divV_curlVJLoop alone on the stored c_ij (the JLoop the reference's ve.cpp:122-139 exercises)
void orc_divv_curlv(const sphx_box_t* b, const sphx_fields_t* f, const sphx_params_t* p, uint32_t first,
                    uint32_t last, const uint32_t* nb, const void* whv)
{
    const double* wh = static_cast<const double*>(whv);
    Box           box = toBox(b);
    for (LocalIndex i = first; i < last; ++i)
    {
        unsigned nn = capped<double>(f, i, p->ngmax);
        divvCurlvLoop(i, p->K, box, nb + size_t(i - first) * p->ngmax, nn, f->x, f->y, f->z, F<double>(f->vx),
                      F<double>(f->vy), F<double>(f->vz), F<double>(f->h), F<double>(f->c11), F<double>(f->c12),
                      F<double>(f->c13), F<double>(f->c22), F<double>(f->c23), F<double>(f->c33), wh,
                      F<double>(f->kx), F<double>(f->xm), F<double>(f->divv), F<double>(f->curlv), F<double>(f->dV11),
                      F<double>(f->dV12), F<double>(f->dV13), F<double>(f->dV22), F<double>(f->dV23),
                      F<double>(f->dV33), f->dV11 != nullptr);
    }
}

void orc_av_switches(const sphx_box_t* b, const sphx_fields_t* f, const sphx_params_t* p, uint32_t first,
                     uint32_t last, const uint32_t* nb, const void* wh)
{
#define C(T) avSwitches<T>(toBox(b), f, p, first, last, nb, (const T*)wh)
    ORC_DISPATCH(C)
#undef C
}

double orc_momentum_energy_ve(const sphx_box_t* b, const sphx_fields_t* f, const sphx_params_t* p, uint32_t first,
                              uint32_t last, const uint32_t* nb, const void* wh, int avClean, void* maxvs)
{
    Box box = toBox(b);
    if (p->hydro_type == SPHX_F32)
    {
        return avClean ? momentumEnergyVe<true, float>(box, f, p, first, last, nb, (const float*)wh, (float*)maxvs)
                       : momentumEnergyVe<false, float>(box, f, p, first, last, nb, (const float*)wh, (float*)maxvs);
    }
    return avClean ? momentumEnergyVe<true, double>(box, f, p, first, last, nb, (const double*)wh, (double*)maxvs)
                   : momentumEnergyVe<false, double>(box, f, p, first, last, nb, (const double*)wh, (double*)maxvs);
}

// ---- time integration ------------------------------------------------------------------------

double orc_energy_update(double u_old, double dt, double dt_m1, double du, double du_m1)
{
    return energyUpdate(u_old, dt, dt_m1, du, du_m1);
}

void orc_position_update(double dt, double dt_m1, const double* X, const double* A, const double* dX,
                         const sphx_box_t* b, double* Xo, double* Vo, double* dXo)
{
    positionUpdate(dt, dt_m1, X, A, dX, toBox(b), Xo, Vo, dXo);
}

//! computePositions (positions_gpu.cu:116-168) on host arrays; hydro_type selects T
void orc_compute_positions(const sphx_box_t* b, const sphx_fields_t* f, const sphx_history_t* hs,
                           const sphx_params_t* p, uint32_t first, uint32_t last, double dt, double dt_m1)
{
    Box box = toBox(b);
    auto run = [&](auto tag) {
        using T          = decltype(tag);
        double constCv   = hs->mui ? -1.0 : double(idealGasCv(T(p->muiConst), p->gamma));
        double* u        = f->temp ? nullptr : f->u;
        computePositions<T>(box, dt, dt_m1, f->x, f->y, f->z, (T*)f->vx, (T*)f->vy, (T*)f->vz, (T*)hs->x_m1,
                            (T*)hs->y_m1, (T*)hs->z_m1, (const T*)f->ax, (const T*)f->ay, (const T*)f->az, f->temp, u,
                            f->du, (T*)hs->du_m1, (const T*)f->h, (const T*)hs->mui, p->gamma, constCv, first, last);
    };
    if (p->hydro_type == SPHX_F32) { run(float{}); }
    else { run(double{}); }
}

//! computePositions with a rung array (positions_gpu.cu:116-168): previous step dtM1Rung[rung[i]]
void orc_compute_positions_rungs(const sphx_box_t* b, const sphx_fields_t* f, const sphx_history_t* hs,
                                 const sphx_params_t* p, uint32_t first, uint32_t last, double dt,
                                 const float* dtM1Rung, const uint8_t* rung)
{
    Box box = toBox(b);
    auto run = [&](auto tag) {
        using T        = decltype(tag);
        double constCv = hs->mui ? -1.0 : double(idealGasCv(T(p->muiConst), p->gamma));
        double* u      = f->temp ? nullptr : f->u;
        computePositions<T>(box, dt, double(dtM1Rung[0]), f->x, f->y, f->z, (T*)f->vx, (T*)f->vy, (T*)f->vz,
                            (T*)hs->x_m1, (T*)hs->y_m1, (T*)hs->z_m1, (const T*)f->ax, (const T*)f->ay,
                            (const T*)f->az, f->temp, u, f->du, (T*)hs->du_m1, (const T*)f->h, (const T*)hs->mui,
                            p->gamma, constCv, first, last, dtM1Rung, rung);
    };
    if (p->hydro_type == SPHX_F32) { run(float{}); }
    else { run(double{}); }
}

//! driftPositionsGpu / driftKernel (positions_gpu.cu:17-85) on host arrays
void orc_drift_positions(const sphx_fields_t* f, const sphx_history_t* hs, const sphx_params_t* p, uint32_t first,
                         uint32_t last, double dt, double dtBack, const float* dtM1Rung, const uint8_t* rung)
{
    auto run = [&](auto tag) {
        using T        = decltype(tag);
        double constCv = hs->mui ? -1.0 : double(idealGasCv(T(p->muiConst), p->gamma));
        double* u      = f->temp ? nullptr : f->u;
        driftPositions<T>(dt, dtBack, dtM1Rung, rung, f->x, f->y, f->z, (T*)f->vx, (T*)f->vy, (T*)f->vz,
                          (const T*)hs->x_m1, (const T*)hs->y_m1, (const T*)hs->z_m1, (const T*)f->ax,
                          (const T*)f->ay, (const T*)f->az, f->temp, u, f->du, (const T*)hs->du_m1, (const T*)hs->mui,
                          p->gamma, constCv, first, last);
    };
    if (p->hydro_type == SPHX_F32) { run(float{}); }
    else { run(double{}); }
}

//! groupDivvTimestepGpu (ts_groups.cu:16-46) over explicit groups
void orc_group_divv_timestep(const sphx_fields_t* f, const sphx_params_t* p, const uint32_t* groupStart,
                             const uint32_t* groupEnd, uint32_t numGroups, float Krho, float* groupDt)
{
    if (p->hydro_type == SPHX_F32)
        groupDivvTimestep(Krho, groupStart, groupEnd, numGroups, (const float*)f->divv, groupDt);
    else
        groupDivvTimestep(Krho, groupStart, groupEnd, numGroups, (const double*)f->divv, groupDt);
}

//! groupAccTimestepGpu (ts_groups.cu:51-81) over explicit groups
void orc_group_acc_timestep(const sphx_fields_t* f, const sphx_params_t* p, const uint32_t* groupStart,
                            const uint32_t* groupEnd, uint32_t numGroups, float etaAcc, float* groupDt)
{
    if (p->hydro_type == SPHX_F32)
        groupAccTimestep(etaAcc, groupStart, groupEnd, numGroups, (const float*)f->ax, (const float*)f->ay,
                         (const float*)f->az, groupDt);
    else
        groupAccTimestep(etaAcc, groupStart, groupEnd, numGroups, (const double*)f->ax, (const double*)f->ay,
                         (const double*)f->az, groupDt);
}

//! storeRungGpu (ts_groups.cu:84-104)
void orc_store_rung(const uint32_t* groupStart, const uint32_t* groupEnd, uint32_t numGroups, uint8_t rung,
                    uint8_t* particleRungs)
{
    storeRung(groupStart, groupEnd, numGroups, rung, particleRungs);
}

void orc_update_smoothing_length(const sphx_fields_t* f, const sphx_params_t* p, uint32_t first, uint32_t last)
{
    if (p->hydro_type == SPHX_F32) { updateSmoothingLength(first, last, p->ng0, f->nc, (float*)f->h); }
    else { updateSmoothingLength(first, last, p->ng0, f->nc, (double*)f->h); }
}

//! computeTimestep (ts_global.hpp:95-113) with accelerationTimestep's maximum (:46-67)
void orc_compute_timestep(const sphx_fields_t* f, const sphx_params_t* p, uint32_t first, uint32_t last,
                          sphx_timestep_t* ts)
{
    double maxAccSq = 0;
    if (ts->g != 0.0)
    {
        maxAccSq = p->hydro_type == SPHX_F32
                       ? maxNormSquare((const float*)f->ax, (const float*)f->ay, (const float*)f->az, first, last)
                       : maxNormSquare((const double*)f->ax, (const double*)f->ay, (const double*)f->az, first, last);
    }
    computeTimestep(maxAccSq, ts->minDt, ts->minDt_m1, ts->ttot, ts->minDtCourant, ts->minDtRho, ts->maxDtIncrease,
                    ts->g, ts->etaAcc, ts->eps);
}

//! rhoTimestep (ts_global.hpp:70-93): Krho / |max divv|
void orc_rho_timestep(const sphx_fields_t* f, const sphx_params_t* p, uint32_t first, uint32_t last,
                      sphx_timestep_t* ts)
{
    double m = -INFINITY;
    for (uint32_t i = first; i < last; ++i)
    {
        double v = p->hydro_type == SPHX_F32 ? double(((const float*)f->divv)[i]) : ((const double*)f->divv)[i];
        m        = std::max(m, v);
    }
    ts->minDtRho = ts->Krho / std::abs(m);
}

//! box update of one Domain::sync (globalBox): limits in box->lo / box->hi, updated in place
void orc_global_box(sphx_box_t* box, const double* x, const double* y, const double* z, size_t n, int firstCall)
{
    double lim[6] = {box->lo[0], box->hi[0], box->lo[1], box->hi[1], box->lo[2], box->hi[2]};
    int    bc[3]  = {box->boundary[0], box->boundary[1], box->boundary[2]};
    globalBox(x, y, z, n, bc, firstCall != 0, lim);
    for (int d = 0; d < 3; ++d)
        box->lo[d] = lim[2 * d], box->hi[d] = lim[2 * d + 1];
}

//! limitBoxShrinking alone (limits6 = xmin xmax ymin ymax zmin zmax)
void orc_limit_box_shrinking(const double* fit6, const double* prev6, double* out6)
{
    limitBoxShrinking(fit6, prev6, 0.05, out6);
}

//! conserved quantities over [first, last): out = eKin, eInt, linmom[3], angmom[3], ncsum
void orc_conserved_quantities(const sphx_fields_t* f, const sphx_params_t* p, uint32_t first, uint32_t last,
                              double* out)
{
    if (p->hydro_type == SPHX_F32)
    {
        double cv = double(idealGasCv(float(p->muiConst), p->gamma));
        conservedQuantities(f->x, f->y, f->z, (const float*)f->vx, (const float*)f->vy, (const float*)f->vz,
                            (const float*)f->m, f->u, f->temp, cv, f->nc, first, last, out);
    }
    else
    {
        double cv = idealGasCv(p->muiConst, p->gamma);
        conservedQuantities(f->x, f->y, f->z, (const double*)f->vx, (const double*)f->vy, (const double*)f->vz,
                            (const double*)f->m, f->u, f->temp, cv, f->nc, first, last, out);
    }
}

} // extern "C"
