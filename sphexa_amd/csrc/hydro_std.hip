// Standard SPH pair loops: density (XMass), ideal-gas EOS, IAD, momentum-energy with the
// Courant time step (sph/include/sph/hydro_std/*, sph/include/sph/hydro_ve/xmass_kern.hpp).
// Neighbour data is read from packed per-particle records (records.hpp).
#include <cstring>

#include "records.hpp"

namespace sphx
{

namespace
{

constexpr unsigned kPackThreads = 256;

template<class T>
__global__ void __launch_bounds__(kPackThreads) packMass(DevBox b, const double* __restrict__ x,
                                                         const double* __restrict__ y, const double* __restrict__ z,
                                                         const T* __restrict__ m, RecMass<T>* __restrict__ rec,
                                                         uint32_t n)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) { return; }
    RecMass<T> r;
    r.x = coordOf<T>(b, 0, x[i]), r.y = coordOf<T>(b, 1, y[i]), r.z = coordOf<T>(b, 2, z[i]), r.m = m[i];
    rec[i] = r;
}

template<class T>
__global__ void __launch_bounds__(kPackThreads) packVol(DevBox b, const double* __restrict__ x,
                                                        const double* __restrict__ y, const double* __restrict__ z,
                                                        const T* __restrict__ m, const T* __restrict__ rho,
                                                        RecVol<T>* __restrict__ rec, uint32_t n)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) { return; }
    rec[i] = volRecord<T>(coordOf<T>(b, 0, x[i]), coordOf<T>(b, 1, y[i]), coordOf<T>(b, 2, z[i]), m[i], rho[i]);
}

struct MomStdFields
{
    const void *vx, *vy, *vz, *h, *m, *rho, *p, *c, *c11, *c12, *c13, *c22, *c23, *c33;
    void *      ax, *ay, *az;
    double*     du;
    const void* alpha; // per-particle AV coefficient (SPHX_FLAG_STD_ALPHA), else the reference's 1
};

/* The std momentum-energy record, built by the pack kernel or by the IAD epilogue under
 * SPHX_FLAG_CHAIN with this one function, so both give the same bits. C = {c11, .., c33}. */
template<class T>
__device__ __forceinline__ RecMomStd<T> momStdRecord(Coord<T> x, Coord<T> y, Coord<T> z, T vx, T vy, T vz, T h, T m,
                                                     T rho, T p, T c, const T C[6])
{
    RecMomStd<T> r;
    r.x = x, r.y = y, r.z = z;
    r.vx = vx, r.vy = vy, r.vz = vz;
    r.hInv = T(1) / h;
    r.m    = m;
    r.vol  = fastDiv(m, rho);
    r.pro  = fastDiv(p, rho); // p_j / (rho_j gradh_j), gradh_j = 1
    r.c    = c;
    r.c11 = C[0], r.c12 = C[1], r.c13 = C[2], r.c22 = C[3], r.c23 = C[4], r.c33 = C[5];
    return r;
}

template<class T>
__global__ void __launch_bounds__(kPackThreads) packMomStd(PairArgs a, MomStdFields f, RecMomStd<T>* __restrict__ rec)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.numParticles) { return; }
    const T C[6] = {ld<T>(f.c11, i), ld<T>(f.c12, i), ld<T>(f.c13, i), ld<T>(f.c22, i), ld<T>(f.c23, i), ld<T>(f.c33, i)};
    rec[i] = momStdRecord<T>(coordOf<T>(a.box, 0, a.x[i]), coordOf<T>(a.box, 1, a.y[i]), coordOf<T>(a.box, 2, a.z[i]),
                             ld<T>(f.vx, i), ld<T>(f.vy, i), ld<T>(f.vz, i), ld<T>(f.h, i), ld<T>(f.m, i),
                             ld<T>(f.rho, i), ld<T>(f.p, i), ld<T>(f.c, i), C);
}

/* xmassJLoop (sph/include/sph/hydro_ve/xmass_kern.hpp:51-79):
 *   xm_i = m_i / (K / h_i^3 * (m_i + sum_j m_j W(r_ij / h_i)))
 * toRho: computeDensity (hydro_std/density.hpp:42-53) stores rho = m / xm instead. */
template<class T, bool toRho, bool Wide>
__global__ void SPHX_LIGHT_LOOP(T)
    xmassKernel(PairArgs a, const RecMass<T>* __restrict__ rec, const T* __restrict__ h, SincKernel<T> kern,
                T* __restrict__ out, RecGradh<T>* __restrict__ next, RecVol<T>* __restrict__ nextVol)
{
    Lane t = laneTarget(a.L);
    if (!t.valid) { return; }
    const uint32_t   i     = t.i;
    const RecMass<T> ri    = rec[i];
    T                hi    = h[i];
    T                hInv  = T(1) / hi;
    T                h3Inv = hInv * hInv * hInv;
    double           r2h   = 2.0 * double(hi);
    Acc<T>           rho0(ri.m);
    withFold(recordNeedsFold<T, Wide>(a.box, ri, r2h), [&](auto fold) {
        constexpr int                U = kBatchLight<T>;
        NeighborBatch<U, RecMass<T>> nb(t, rec);
        for (unsigned k = 0; k < t.n; k += U)
        {
            nb.load(k);
            typename SincKernel<T>::FetchW w[U];
#pragma unroll
            for (int u = 0; u < U; ++u)
                w[u] = kern.fetchW(pairSeparation<decltype(fold)::value, T>(a.box, r2h, ri, nb.r[u]).d * hInv);
#pragma unroll
            for (int u = 0; u < U; ++u)
                rho0.add(nb.live(k, u, value(w[u])) * nb.r[u].m);
        }
    });
    double rhoZero = double(rho0.value()) * a.K * double(h3Inv);
    T      xmi     = T(double(ri.m) / rhoZero);
    const T outi   = toRho ? T(ri.m / xmi) : xmi;
    out[i]         = outi;
    if (!toRho && next) { next[i] = gradhRecord<T>(ri.x, ri.y, ri.z, ri.m, xmi); }
    if (toRho && nextVol) { nextVol[i] = volRecord<T>(ri.x, ri.y, ri.z, ri.m, outi); }
}

//! idealGasEOS_u (sph/include/sph/eos.hpp:35-44) or idealGasEOS with cv from mui (:47-51)
template<class T>
__global__ void __launch_bounds__(256) eosStdKernel(const double* __restrict__ u, const double* __restrict__ temp,
                                                    const T* __restrict__ rho, T* __restrict__ p, T* __restrict__ c,
                                                    double gamma, double cv, uint32_t first, uint32_t last)
{
    uint32_t i = first + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= last) { return; }
    double ui  = u ? u[i] : cv * temp[i];
    double tmp = ui * (gamma - 1.0);
    p[i]       = T(double(rho[i]) * tmp);
    c[i]       = T(sqrt(gamma * tmp));
}

/* isothermalEOS (sph/include/sph/eos.hpp:59-64) as the reference's GPU kernel applies it
 * (hydro_std/eos_gpu.cu:73-82): p = rho c0 c0 (in FP64, once rounded to the HydroType in the FP32
 * form), c = c0, temp = 0 */
template<class T>
__global__ void __launch_bounds__(256) eosStdIsothermalKernel(const T* __restrict__ rho, T* __restrict__ p,
                                                              T* __restrict__ c, double* __restrict__ temp, T c0,
                                                              uint32_t first, uint32_t last)
{
    uint32_t i = first + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= last) { return; }
    const double cc = double(c0);
    p[i]            = T(double(rho[i]) * cc * cc);
    c[i]            = c0;
    if (temp) { temp[i] = 0.0; }
}

/* polytropicEOS (sph/include/sph/eos.hpp:77-86) per particle (hydro_std/eos.hpp:99-114):
 * p = K rho^n, c = sqrt(n p / rho) in FP64, temp = 0 */
template<class T>
__global__ void __launch_bounds__(256) eosStdPolytropicKernel(const T* __restrict__ rho, T* __restrict__ p,
                                                              T* __restrict__ c, double* __restrict__ temp, double K,
                                                              double n, uint32_t first, uint32_t last)
{
    uint32_t i = first + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= last) { return; }
    const double r  = double(rho[i]);
    const double pi = K * pow(r, n);
    p[i]            = T(pi);
    c[i]            = T(sqrt(n * pi / r));
    if (temp) { temp[i] = 0.0; }
}

/* IADJLoopSTD (hydro_std/iad_kern.hpp:11-76): tau_ab = sum_j m_j/rho_j W_ij r_a r_b, inverted */
template<class T, bool Wide>
__global__ void SPHX_LIGHT_LOOP(T)
    iadStdKernel(PairArgs a, const RecVol<T>* __restrict__ rec, const T* __restrict__ h, SincKernel<T> kern,
                 T* c11, T* c12, T* c13, T* c22, T* c23, T* c33, MomStdFields nf, RecMomStd<T>* __restrict__ next)
{
    Lane t = laneTarget(a.L);
    if (!t.valid) { return; }
    const uint32_t  i     = t.i;
    const RecVol<T> ri    = rec[i];
    T               hi    = h[i];
    T               hiInv = T(1) / hi;
    double          r2h   = 2.0 * double(hi);
    T               t11 = 0, t12 = 0, t13 = 0, t22 = 0, t23 = 0, t33 = 0;
    withFold(recordNeedsFold<T, Wide>(a.box, ri, r2h), [&](auto fold) {
        constexpr int               U = kBatchLight<T>;
        NeighborBatch<U, RecVol<T>> nb(t, rec);
        for (unsigned k = 0; k < t.n; k += U)
        {
            nb.load(k);
            Sep<T>                    r[U];
            typename SincKernel<T>::FetchW w[U];
#pragma unroll
            for (int u = 0; u < U; ++u)
            {
                r[u] = pairSeparation<decltype(fold)::value, T>(a.box, r2h, ri, nb.r[u]);
                w[u] = kern.fetchW(r[u].d * hiInv);
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
            {
                T rx = r[u].x, ry = r[u].y, rz = r[u].z;
                T vol_w = nb.r[u].vol * nb.live(k, u, value(w[u]));
                t11 += rx * rx * vol_w;
                t12 += rx * ry * vol_w;
                t13 += rx * rz * vol_w;
                t22 += ry * ry * vol_w;
                t23 += ry * rz * vol_w;
                t33 += rz * rz * vol_w;
            }
        }
    });
    double cij[6];
    iadInvert(t11, t12, t13, t22, t23, t33, double(hi), a.K, cij);
    const T C[6] = {T(cij[0]), T(cij[1]), T(cij[2]), T(cij[3]), T(cij[4]), T(cij[5])};
    c11[i] = C[0], c12[i] = C[1], c13[i] = C[2];
    c22[i] = C[3], c23[i] = C[4], c33[i] = C[5];
    if (next)
    {
        next[i] = momStdRecord<T>(ri.x, ri.y, ri.z, ld<T>(nf.vx, i), ld<T>(nf.vy, i), ld<T>(nf.vz, i), hi,
                                  ld<T>(nf.m, i), ld<T>(nf.rho, i), ld<T>(nf.p, i), ld<T>(nf.c, i), C);
    }
}

/* momentumAndEnergyJLoop, std (hydro_std/momentum_energy_kern.hpp:11-134) + tsKCourant.
 * PerAlpha: AV coefficients alpha_i, alpha_j from f.alpha instead of the constant 1. */
template<class T, bool PerAlpha, bool Wide>
__global__ void SPHX_MOM_LOOP(T)
    momentumStdKernel(PairArgs a, MomStdFields f, const RecMomStd<T>* __restrict__ rec, SincKernel<T> kern,
                      float Kcour, unsigned long long* minDt)
{
    Lane   t      = laneTarget(a.L);
    double dtLane = INFINITY;
    if (t.valid)
    {
        const uint32_t     i      = t.i;
        const RecMomStd<T> ri     = rec[i];
        T                  hi     = ld<T>(f.h, i);
        T                  roi    = ld<T>(f.rho, i);
        T                  hiInv  = ri.hInv;
        T                  hiInv3 = hiInv * hiInv * hiInv;
        T                  pro_i  = fastDiv(ld<T>(f.p, i), roi * roi); // p_i / (gradh_i rho_i^2), gradh_i = 1
        double             r2h    = 2.0 * double(hi);
        const T            alpha_i = PerAlpha ? ld<T>(f.alpha, i) : T(1);

        T      maxvs = 0;
        T      momx = 0, momy = 0, momz = 0;
        Acc<T>       energy;
        withFold(recordNeedsFold<T, Wide>(a.box, ri, r2h), [&](auto fold) {
            constexpr int                  U = kBatchMom<T>;
            NeighborBatch<U, RecMomStd<T>> nb(t, rec);
            for (unsigned k = 0; k < t.n; k += U)
            {
                nb.load(k);
                Sep<T>                    r[U];
                typename SincKernel<T>::FetchW wi[U], wj[U];
#pragma unroll
                for (int u = 0; u < U; ++u)
                {
                    r[u]  = pairSeparation<decltype(fold)::value, T>(a.box, r2h, ri, nb.r[u]);
                    wi[u] = kern.fetchW(r[u].d * hiInv);
                    wj[u] = kern.fetchW(r[u].d * nb.r[u].hInv);
                }
#pragma unroll
                for (int u = 0; u < U; ++u)
                {
                    const RecMomStd<T>& rj = nb.r[u];
                    T rx = r[u].x, ry = r[u].y, rz = r[u].z, dist = r[u].d;
                    T vx_ij = ri.vx - rj.vx, vy_ij = ri.vy - rj.vy, vz_ij = ri.vz - rj.vz;
                    T rv    = rx * vx_ij + ry * vy_ij + rz * vz_ij;
                    T Wi    = hiInv3 * nb.live(k, u, value(wi[u]));
                    T Wj    = rj.hInv * rj.hInv * rj.hInv * nb.live(k, u, value(wj[u]));

                    T tA1_i = ri.c11 * rx + ri.c12 * ry + ri.c13 * rz;
                    T tA2_i = ri.c12 * rx + ri.c22 * ry + ri.c23 * rz;
                    T tA3_i = ri.c13 * rx + ri.c23 * ry + ri.c33 * rz;
                    T tA1_j = rj.c11 * rx + rj.c12 * ry + rj.c13 * rz;
                    T tA2_j = rj.c12 * rx + rj.c22 * ry + rj.c23 * rz;
                    T tA3_j = rj.c13 * rx + rj.c23 * ry + rj.c33 * rz;

                    T wij  = fastDiv(rv, dist);
                    T alpha_j = PerAlpha ? ld<T>(f.alpha, nb.j[u]) : T(1);
                    T visc    = T(0.5) * artificialViscosity(alpha_i, alpha_j, ri.c, rj.c, wij);
                    T vsig = ri.c + rj.c - T(3) * wij;
                    maxvs  = nb.in(k, u) && vsig > maxvs ? vsig : maxvs;

                    T mj_roj_Wj = rj.vol * Wj;
                    T mj_pro_i  = rj.m * pro_i;
                    {
                        T aa = Wi * (mj_pro_i + visc * ri.vol);
                        T bb = mj_roj_Wj * (rj.pro + visc);
                        momx += aa * tA1_i + bb * tA1_j;
                        momy += aa * tA2_i + bb * tA2_j;
                        momz += aa * tA3_i + bb * tA3_j;
                    }
                    {
                        T aa = Wi * (T(2) * mj_pro_i + visc * ri.vol);
                        T bb = visc * mj_roj_Wj;
                        energy.add(vx_ij * (aa * tA1_i + bb * tA1_j) + vy_ij * (aa * tA2_i + bb * tA2_j) +
                                   vz_ij * (aa * tA3_i + bb * tA3_j));
                    }
                }
            }
        });
        f.du[i] = -a.K * 0.5 * double(energy.value());
        st<T>(f.ax, i, T(a.K * double(momx)));
        st<T>(f.ay, i, T(a.K * double(momy)));
        st<T>(f.az, i, T(a.K * double(momz)));
        dtLane = courantDt(maxvs, hi, ri.c, Kcour);
    }
    reduceMinDt<T>(minDt, dtLane, a.L.groups);
}

template<class T>
void launchDensity(sphx_workspace* ws, const PairArgs& a, const sphx_fields_t* f, const sphx_params_t* p, bool toRho,
                   void* out)
{
    unsigned blocks = pairBlocks(a.L);
    if (blocks == 0) { return; }
    RecMass<T>* rec = recordBuffer<RecMass<T>>(ws, a.numParticles, 0);
    packMass<T><<<ceilDiv(a.numParticles, kPackThreads), kPackThreads, 0, ws->stream>>>(a.box, a.x, a.y, a.z,
                                                                                         (const T*)f->m, rec,
                                                                                         a.numParticles);
    SPHX_CHECK_LAUNCH();
    // XMass leaves the VeDefGradh records (position, m, xm), density the IAD ones (position, m/rho)
    RecGradh<T>* next = !toRho && chainable(a, p) ? recordBuffer<RecGradh<T>>(ws, a.numParticles, 1) : nullptr;
    RecVol<T>*   nextVol = toRho && chainable(a, p) ? recordBuffer<RecVol<T>>(ws, a.numParticles, 1) : nullptr;
    {
        KernelScope timer(ws, toRho ? "density_std" : "xmass");
        launchWide<T>(ws->nbWide, [&](auto wide) {
            constexpr bool W = decltype(wide)::value;
            if (toRho)
                xmassKernel<T, true, W><<<blocks, kPairWaves * kWave, isMixed<T> ? occupancyCap(SPHX_CAP_XMASS) : 0, ws->stream>>>(
                    a, rec, (const T*)f->h, sincKernel<T>(ws), (T*)out, nullptr, nextVol);
            else
                xmassKernel<T, false, W><<<blocks, kPairWaves * kWave, isMixed<T> ? occupancyCap(SPHX_CAP_XMASS) : 0, ws->stream>>>(
                    a, rec, (const T*)f->h, sincKernel<T>(ws), (T*)out, next, nullptr);
        });
        SPHX_CHECK_LAUNCH();
    }
    if (next) { leaveChain(ws, kChainGradh, 1, a, f, p); }
    if (nextVol) { leaveChain(ws, kChainVol, 1, a, f, p); }
}

template<class T>
void launchIadStd(sphx_workspace* ws, const PairArgs& a, const sphx_fields_t* f, const sphx_params_t* p)
{
    unsigned blocks = pairBlocks(a.L);
    int      in     = takeChain(ws, kChainVol, a, f, p);
    if (blocks == 0) { return; }
    RecVol<T>* rec = recordBuffer<RecVol<T>>(ws, a.numParticles, in < 0 ? 0 : in);
    if (in < 0)
    {
        in = 0;
        packVol<T><<<ceilDiv(a.numParticles, kPackThreads), kPackThreads, 0, ws->stream>>>(
            a.box, a.x, a.y, a.z, (const T*)f->m, (const T*)f->rho, rec, a.numParticles);
        SPHX_CHECK_LAUNCH();
    }
    // IAD leaves the momentum-energy records when chaining (it needs v, p and c besides its own fields)
    MomStdFields  nf{f->vx, f->vy, f->vz, f->h, f->m, f->rho, f->p, f->c};
    RecMomStd<T>* next = chainable(a, p) && f->vx && f->vy && f->vz && f->p && f->c
                             ? recordBuffer<RecMomStd<T>>(ws, a.numParticles, 1 - in)
                             : nullptr;
    {
        KernelScope timer(ws, "iad_std");
        launchWide<T>(ws->nbWide, [&](auto wide) {
            iadStdKernel<T, decltype(wide)::value><<<blocks, kPairWaves * kWave, isMixed<T> ? occupancyCap(SPHX_CAP_IAD) : 0, ws->stream>>>(
                a, rec, (const T*)f->h, sincKernel<T>(ws), (T*)f->c11, (T*)f->c12, (T*)f->c13, (T*)f->c22,
                (T*)f->c23, (T*)f->c33, nf, next);
        });
        SPHX_CHECK_LAUNCH();
    }
    if (next) { leaveChain(ws, kChainMomStd, 1 - in, a, f, p); }
}

template<class T>
void launchMomentumStd(sphx_workspace* ws, const PairArgs& a, const MomStdFields& mf, const sphx_fields_t* f,
                       const sphx_params_t* p, float Kcour, unsigned long long* dt, bool perAlpha)
{
    unsigned blocks = pairBlocks(a.L);
    int      in     = takeChain(ws, kChainMomStd, a, f, p);
    if (blocks == 0) { return; }
    RecMomStd<T>* rec = recordBuffer<RecMomStd<T>>(ws, a.numParticles, in < 0 ? 0 : in);
    if (in < 0)
    {
        packMomStd<T><<<ceilDiv(a.numParticles, kPackThreads), kPackThreads, 0, ws->stream>>>(a, mf, rec);
        SPHX_CHECK_LAUNCH();
    }
    KernelScope timer(ws, "momentum_std");
    launchWide<T>(ws->nbWide, [&](auto wide) {
        constexpr bool W = decltype(wide)::value;
        if (perAlpha)
            momentumStdKernel<T, true, W><<<blocks, kPairWaves * kWave, isMixed<T> ? occupancyCap(SPHX_CAP_MOM) : 0, ws->stream>>>(a, mf, rec, sincKernel<T>(ws),
                                                                                        Kcour, dt);
        else
            momentumStdKernel<T, false, W><<<blocks, kPairWaves * kWave, isMixed<T> ? occupancyCap(SPHX_CAP_MOM) : 0, ws->stream>>>(a, mf, rec,
                                                                                         sincKernel<T>(ws), Kcour, dt);
    });
    SPHX_CHECK_LAUNCH();
}

} // namespace

void densityStd(sphx_workspace* ws, const DevBox& box, const sphx_fields_t* f, const sphx_params_t* p, uint32_t first,
                uint32_t last)
{
    PairArgs a = makePairArgs(ws, box, f, p, first, last);
    SPHX_REQUIRE(f->h && f->m && f->rho, SPHX_ERR_ARG, "density needs h, m and rho");
    if (p->hydro_type == SPHX_F32) launchDensity<float>(ws, a, f, p, true, f->rho);
    else launchDensity<double>(ws, a, f, p, true, f->rho);
}

void xmassVe(sphx_workspace* ws, const DevBox& box, const sphx_fields_t* f, const sphx_params_t* p, uint32_t first,
             uint32_t last)
{
    PairArgs a = makePairArgs(ws, box, f, p, first, last);
    SPHX_REQUIRE(f->h && f->m && f->xm, SPHX_ERR_ARG, "xmass needs h, m and xm");
    if (p->hydro_type == SPHX_F32) launchDensity<float>(ws, a, f, p, false, f->xm);
    else launchDensity<double>(ws, a, f, p, false, f->xm);
}

template<class T>
void launchEosStdVariant(sphx_workspace* ws, const sphx_fields_t* f, const sphx_params_t* p, uint32_t first,
                         uint32_t last)
{
    const unsigned blocks = ceilDiv(last - first, 256);
    if (p->eos == SPHX_EOS_ISOTHERMAL)
        eosStdIsothermalKernel<T><<<blocks, 256, 0, ws->stream>>>((const T*)f->rho, (T*)f->p, (T*)f->c, f->temp,
                                                                  T(p->soundSpeedConst), first, last);
    else
        eosStdPolytropicKernel<T><<<blocks, 256, 0, ws->stream>>>((const T*)f->rho, (T*)f->p, (T*)f->c, f->temp,
                                                                  p->polytropic_const, p->polytropic_index, first,
                                                                  last);
    SPHX_CHECK_LAUNCH();
}

void eosStd(sphx_workspace* ws, const sphx_fields_t* f, const sphx_params_t* p, uint32_t first, uint32_t last)
{
    SPHX_REQUIRE(p->eos >= SPHX_EOS_IDEAL_GAS && p->eos <= SPHX_EOS_POLYTROPIC, SPHX_ERR_ARG, "unknown EOS choice");
    SPHX_REQUIRE(first <= last, SPHX_ERR_ARG, "invalid range");
    if (p->eos != SPHX_EOS_IDEAL_GAS)
    {
        SPHX_REQUIRE(f->rho && f->p && f->c, SPHX_ERR_ARG, "EOS needs rho, p and c");
        if (last == first) { return; }
        if (p->hydro_type == SPHX_F32) launchEosStdVariant<float>(ws, f, p, first, last);
        else launchEosStdVariant<double>(ws, f, p, first, last);
        return;
    }
    SPHX_REQUIRE(f->rho && f->p && f->c && (f->u || f->temp), SPHX_ERR_ARG, "EOS needs rho, p, c and u or temp");
    if (last == first) { return; }
    // idealGasCv (sph/include/sph/eos.hpp:19-23), R = 8.317e7
    double cv = 8.317e7 / p->muiConst / (p->gamma - 1.0);
    if (p->hydro_type == SPHX_F32)
    {
        cv = double(float(8.317e7f / float(p->muiConst) / (p->gamma - 1.0)));
        eosStdKernel<float><<<ceilDiv(last - first, 256), 256, 0, ws->stream>>>(
            f->u, f->temp, (const float*)f->rho, (float*)f->p, (float*)f->c, p->gamma, cv, first, last);
    }
    else
    {
        eosStdKernel<double><<<ceilDiv(last - first, 256), 256, 0, ws->stream>>>(
            f->u, f->temp, (const double*)f->rho, (double*)f->p, (double*)f->c, p->gamma, cv, first, last);
    }
    SPHX_CHECK_LAUNCH();
}

void iadStd(sphx_workspace* ws, const DevBox& box, const sphx_fields_t* f, const sphx_params_t* p, uint32_t first,
            uint32_t last)
{
    PairArgs a = makePairArgs(ws, box, f, p, first, last);
    SPHX_REQUIRE(f->h && f->m && f->rho && f->c11 && f->c12 && f->c13 && f->c22 && f->c23 && f->c33, SPHX_ERR_ARG,
                 "IAD needs h, m, rho and c11..c33");
    if (p->hydro_type == SPHX_F32) launchIadStd<float>(ws, a, f, p);
    else launchIadStd<double>(ws, a, f, p);
}

double momentumEnergyStd(sphx_workspace* ws, const DevBox& box, const sphx_fields_t* f, const sphx_params_t* p,
                         uint32_t first, uint32_t last)
{
    PairArgs a = makePairArgs(ws, box, f, p, first, last);
    SPHX_REQUIRE(f->vx && f->vy && f->vz && f->h && f->m && f->rho && f->p && f->c && f->c11 && f->c12 && f->c13 &&
                     f->c22 && f->c23 && f->c33 && f->ax && f->ay && f->az && f->du,
                 SPHX_ERR_ARG, "momentum-energy needs v, h, m, rho, p, c, c11..c33, ax, ay, az and du");
    ws->counters.reserve(2 * kCounterSlots);
    unsigned long long* dt = reinterpret_cast<unsigned long long*>(ws->counters.get()) + 3;
    fill64(ws, reinterpret_cast<uint64_t*>(dt), 0x7ff0000000000000ull, 1); // +inf
    const bool perAlpha = (p->flags & SPHX_FLAG_STD_ALPHA) != 0;
    SPHX_REQUIRE(!perAlpha || f->alpha, SPHX_ERR_ARG, "SPHX_FLAG_STD_ALPHA set but fields.alpha is NULL");
    MomStdFields mf{f->vx,  f->vy,  f->vz,  f->h,  f->m,  f->rho, f->p,  f->c,  f->c11, f->c12,
                    f->c13, f->c22, f->c23, f->c33, f->ax, f->ay,  f->az, f->du, f->alpha};
    if (p->hydro_type == SPHX_F32) launchMomentumStd<float>(ws, a, mf, f, p, float(p->Kcour), dt, perAlpha);
    else launchMomentumStd<double>(ws, a, mf, f, p, float(p->Kcour), dt, perAlpha);
    unsigned long long bits = 0;
    SPHX_HIP(hipMemcpyAsync(&bits, dt, 8, hipMemcpyDeviceToHost, ws->stream));
    SPHX_HIP(hipStreamSynchronize(ws->stream));
    double r;
    std::memcpy(&r, &bits, 8);
    return r;
}

} // namespace sphx
