// Standard SPH pair loops: density (XMass), ideal-gas EOS, IAD, momentum-energy with the
// Courant time step (sph/include/sph/hydro_std/*, sph/include/sph/hydro_ve/xmass_kern.hpp).
#include <cstring>

#include "pairs.hpp"

namespace sphx
{

namespace
{

struct Common
{
    const double* x;
    const double* y;
    const double* z;
    DevBox        box;
    double        K;
    Lists         L;
};

Common makeCommon(sphx_workspace* ws, const DevBox& box, const sphx_fields_t* f, const sphx_params_t* p,
                  uint32_t first, uint32_t last)
{
    SPHX_REQUIRE(ws->nbValid, SPHX_ERR_STATE, "pair loop called before find_neighbors");
    SPHX_REQUIRE(first >= ws->nbFirst && last <= ws->nbLast && first <= last, SPHX_ERR_ARG,
                 "target range outside the searched range");
    SPHX_REQUIRE(first == ws->nbFirst, SPHX_ERR_ARG, "target range must start at the searched first index");
    SPHX_REQUIRE(p->ngmax == ws->nbNgmax, SPHX_ERR_ARG, "ngmax differs from the one used by find_neighbors");
    SPHX_REQUIRE(ws->tablesValid, SPHX_ERR_STATE, "kernel tables not set (sphx_workspace_set_kernel)");
    SPHX_REQUIRE(f->x && f->y && f->z && f->nc, SPHX_ERR_ARG, "x, y, z and nc are required");
    Common c;
    c.x = f->x, c.y = f->y, c.z = f->z, c.box = box, c.K = p->K;
    c.L = Lists{ws->nbrs.get(), f->nc, first, last, p->ngmax};
    return c;
}

template<class T>
const T* whTable(sphx_workspace* ws)
{
    if constexpr (isMixed<T>) { return ws->wh32.get(); }
    else { return ws->wh64.get(); }
}

/* xmassJLoop (sph/include/sph/hydro_ve/xmass_kern.hpp:51-79):
 *   xm_i = m_i / (K / h_i^3 * (m_i + sum_j m_j W(r_ij / h_i)))
 * toRho: computeDensity (hydro_std/density.hpp:42-53) stores rho = m / xm instead. */
template<class T, bool toRho>
__global__ void __launch_bounds__(kPairWaves * kWave)
    xmassKernel(Common a, const T* __restrict__ h, const T* __restrict__ m, const T* __restrict__ wh, T* __restrict__ out)
{
    Lane t = laneTarget(a.L);
    if (!t.valid) { return; }
    const uint32_t i  = t.i;
    double         xi = a.x[i], yi = a.y[i], zi = a.z[i];
    T              hi = h[i], mi = m[i];
    T              hInv  = T(1) / hi;
    T              h3Inv = hInv * hInv * hInv;
    double         r2h   = 2.0 * double(hi);
    Acc<T>         rho0(mi);
    for (unsigned k = 0; k < t.n; ++k)
    {
        uint32_t j = t.list[size_t(k) * kGroup];
        T        rx, ry, rz;
        separation(a.box, r2h, xi, yi, zi, a.x[j], a.y[j], a.z[j], rx, ry, rz);
        T dist = fastSqrt(rx * rx + ry * ry + rz * rz);
        T w    = lookup(wh, dist * hInv);
        rho0.add(w * m[j]);
    }
    double rhoZero = double(rho0.value()) * a.K * double(h3Inv);
    T      xmi     = T(double(mi) / rhoZero);
    out[i]         = toRho ? T(mi / xmi) : xmi;
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

/* IADJLoopSTD (hydro_std/iad_kern.hpp:11-76): tau_ab = sum_j m_j/rho_j W_ij r_a r_b, inverted */
template<class T>
__global__ void __launch_bounds__(kPairWaves * kWave)
    iadStdKernel(Common a, const T* __restrict__ h, const T* __restrict__ m, const T* __restrict__ rho,
                 const T* __restrict__ wh, T* c11, T* c12, T* c13, T* c22, T* c23, T* c33)
{
    Lane t = laneTarget(a.L);
    if (!t.valid) { return; }
    const uint32_t i  = t.i;
    double         xi = a.x[i], yi = a.y[i], zi = a.z[i];
    T              hi    = h[i];
    T              hiInv = T(1) / hi;
    double         r2h   = 2.0 * double(hi);
    T              t11 = 0, t12 = 0, t13 = 0, t22 = 0, t23 = 0, t33 = 0;
    for (unsigned k = 0; k < t.n; ++k)
    {
        uint32_t j = t.list[size_t(k) * kGroup];
        T        rx, ry, rz;
        separation(a.box, r2h, xi, yi, zi, a.x[j], a.y[j], a.z[j], rx, ry, rz);
        T dist  = fastSqrt(rx * rx + ry * ry + rz * rz);
        T w     = lookup(wh, dist * hiInv);
        T vol_w = fastDiv(m[j], rho[j]) * w;
        t11 += rx * rx * vol_w;
        t12 += rx * ry * vol_w;
        t13 += rx * rz * vol_w;
        t22 += ry * ry * vol_w;
        t23 += ry * rz * vol_w;
        t33 += rz * rz * vol_w;
    }
    double cij[6];
    iadInvert(t11, t12, t13, t22, t23, t33, double(hi), a.K, cij);
    c11[i] = T(cij[0]), c12[i] = T(cij[1]), c13[i] = T(cij[2]);
    c22[i] = T(cij[3]), c23[i] = T(cij[4]), c33[i] = T(cij[5]);
}

struct MomStdFields
{
    const void *vx, *vy, *vz, *h, *m, *rho, *p, *c, *c11, *c12, *c13, *c22, *c23, *c33;
    void *      ax, *ay, *az;
    double*     du;
};

/* momentumAndEnergyJLoop, std (hydro_std/momentum_energy_kern.hpp:11-134) + tsKCourant */
template<class T>
__global__ void __launch_bounds__(kPairWaves * kWave)
    momentumStdKernel(Common a, MomStdFields f, const T* __restrict__ wh, float Kcour, unsigned long long* minDt)
{
    Lane t = laneTarget(a.L);
    double dtLane = INFINITY;
    if (t.valid)
    {
        const uint32_t i = t.i;
        const T* __restrict__ vx  = (const T*)f.vx;
        const T* __restrict__ vy  = (const T*)f.vy;
        const T* __restrict__ vz  = (const T*)f.vz;
        const T* __restrict__ h   = (const T*)f.h;
        const T* __restrict__ m   = (const T*)f.m;
        const T* __restrict__ rho = (const T*)f.rho;
        const T* __restrict__ p   = (const T*)f.p;
        const T* __restrict__ c   = (const T*)f.c;
        const T* __restrict__ c11 = (const T*)f.c11;
        const T* __restrict__ c12 = (const T*)f.c12;
        const T* __restrict__ c13 = (const T*)f.c13;
        const T* __restrict__ c22 = (const T*)f.c22;
        const T* __restrict__ c23 = (const T*)f.c23;
        const T* __restrict__ c33 = (const T*)f.c33;

        double xi = a.x[i], yi = a.y[i], zi = a.z[i];
        T vxi = vx[i], vyi = vy[i], vzi = vz[i];
        T hi = h[i], roi = rho[i], pri = p[i], ci = c[i];
        T mi_roi = fastDiv(m[i], roi);
        T hiInv  = T(1) / hi;
        T hiInv3 = hiInv * hiInv * hiInv;
        T c11i = c11[i], c12i = c12[i], c13i = c13[i], c22i = c22[i], c23i = c23[i], c33i = c33[i];
        T pro_i = fastDiv(pri, roi * roi); // p_i / (gradh_i rho_i^2), gradh_i = 1
        double r2h = 2.0 * double(hi);

        T      maxvs = 0;
        T      momx = 0, momy = 0, momz = 0;
        Acc<T> energy;
        for (unsigned k = 0; k < t.n; ++k)
        {
            uint32_t j = t.list[size_t(k) * kGroup];
            T        rx, ry, rz;
            separation(a.box, r2h, xi, yi, zi, a.x[j], a.y[j], a.z[j], rx, ry, rz);
            T r2    = rx * rx + ry * ry + rz * rz;
            T dist  = fastSqrt(r2);
            T vx_ij = vxi - vx[j], vy_ij = vyi - vy[j], vz_ij = vzi - vz[j];
            T hjInv = T(1) / h[j];
            T v1    = dist * hiInv;
            T v2    = dist * hjInv;
            T rv    = rx * vx_ij + ry * vy_ij + rz * vz_ij;
            T Wi    = hiInv3 * lookup(wh, v1);
            T Wj    = hjInv * hjInv * hjInv * lookup(wh, v2);

            T tA1_i = c11i * rx + c12i * ry + c13i * rz;
            T tA2_i = c12i * rx + c22i * ry + c23i * rz;
            T tA3_i = c13i * rx + c23i * ry + c33i * rz;
            T c11j = c11[j], c12j = c12[j], c13j = c13[j], c22j = c22[j], c23j = c23[j], c33j = c33[j];
            T tA1_j = c11j * rx + c12j * ry + c13j * rz;
            T tA2_j = c12j * rx + c22j * ry + c23j * rz;
            T tA3_j = c13j * rx + c23j * ry + c33j * rz;

            T roj  = rho[j];
            T cj   = c[j];
            T wij  = fastDiv(rv, dist);
            T visc = T(0.5) * artificialViscosity(T(1), T(1), ci, cj, wij);
            T vsig = ci + cj - T(3) * wij;
            maxvs  = vsig > maxvs ? vsig : maxvs;

            T mj        = m[j];
            T mj_roj_Wj = fastDiv(mj, roj) * Wj;
            T mj_pro_i  = mj * pro_i;
            {
                T aa = Wi * (mj_pro_i + visc * mi_roi);
                T bb = mj_roj_Wj * (fastDiv(p[j], roj) + visc);
                momx += aa * tA1_i + bb * tA1_j;
                momy += aa * tA2_i + bb * tA2_j;
                momz += aa * tA3_i + bb * tA3_j;
            }
            {
                T aa = Wi * (T(2) * mj_pro_i + visc * mi_roi);
                T bb = visc * mj_roj_Wj;
                energy.add(vx_ij * (aa * tA1_i + bb * tA1_j) + vy_ij * (aa * tA2_i + bb * tA2_j) +
                           vz_ij * (aa * tA3_i + bb * tA3_j));
            }
        }
        f.du[i] = -a.K * 0.5 * double(energy.value());
        st<T>(f.ax, i, T(a.K * double(momx)));
        st<T>(f.ay, i, T(a.K * double(momy)));
        st<T>(f.az, i, T(a.K * double(momz)));
        dtLane = courantDt(maxvs, hi, ci, Kcour);
    }
    reduceMinDt<T>(minDt, dtLane);
}

template<class T>
void launchDensity(sphx_workspace* ws, const Common& a, const sphx_fields_t* f, bool toRho, void* out)
{
    unsigned blocks = pairBlocks(a.L.first, a.L.last);
    if (blocks == 0) { return; }
    if (toRho)
        xmassKernel<T, true><<<blocks, kPairWaves * kWave, 0, ws->stream>>>(a, (const T*)f->h, (const T*)f->m,
                                                                            whTable<T>(ws), (T*)out);
    else
        xmassKernel<T, false><<<blocks, kPairWaves * kWave, 0, ws->stream>>>(a, (const T*)f->h, (const T*)f->m,
                                                                             whTable<T>(ws), (T*)out);
    SPHX_CHECK_LAUNCH();
}

} // namespace

void densityStd(sphx_workspace* ws, const DevBox& box, const sphx_fields_t* f, const sphx_params_t* p, uint32_t first,
                uint32_t last)
{
    Common a = makeCommon(ws, box, f, p, first, last);
    SPHX_REQUIRE(f->h && f->m && f->rho, SPHX_ERR_ARG, "density needs h, m and rho");
    if (p->hydro_type == SPHX_F32) launchDensity<float>(ws, a, f, true, f->rho);
    else launchDensity<double>(ws, a, f, true, f->rho);
}

void xmassVe(sphx_workspace* ws, const DevBox& box, const sphx_fields_t* f, const sphx_params_t* p, uint32_t first,
             uint32_t last)
{
    Common a = makeCommon(ws, box, f, p, first, last);
    SPHX_REQUIRE(f->h && f->m && f->xm, SPHX_ERR_ARG, "xmass needs h, m and xm");
    if (p->hydro_type == SPHX_F32) launchDensity<float>(ws, a, f, false, f->xm);
    else launchDensity<double>(ws, a, f, false, f->xm);
}

void eosStd(sphx_workspace* ws, const sphx_fields_t* f, const sphx_params_t* p, uint32_t first, uint32_t last)
{
    SPHX_REQUIRE(f->rho && f->p && f->c && (f->u || f->temp), SPHX_ERR_ARG, "EOS needs rho, p, c and u or temp");
    SPHX_REQUIRE(first <= last, SPHX_ERR_ARG, "invalid range");
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
    Common a = makeCommon(ws, box, f, p, first, last);
    SPHX_REQUIRE(f->h && f->m && f->rho && f->c11 && f->c12 && f->c13 && f->c22 && f->c23 && f->c33, SPHX_ERR_ARG,
                 "IAD needs h, m, rho and c11..c33");
    unsigned blocks = pairBlocks(first, last);
    if (blocks == 0) { return; }
#define LAUNCH(T)                                                                                                      \
    iadStdKernel<T><<<blocks, kPairWaves * kWave, 0, ws->stream>>>(                                                    \
        a, (const T*)f->h, (const T*)f->m, (const T*)f->rho, whTable<T>(ws), (T*)f->c11, (T*)f->c12, (T*)f->c13,      \
        (T*)f->c22, (T*)f->c23, (T*)f->c33)
    if (p->hydro_type == SPHX_F32) LAUNCH(float);
    else LAUNCH(double);
#undef LAUNCH
    SPHX_CHECK_LAUNCH();
}

double momentumEnergyStd(sphx_workspace* ws, const DevBox& box, const sphx_fields_t* f, const sphx_params_t* p,
                         uint32_t first, uint32_t last)
{
    Common a = makeCommon(ws, box, f, p, first, last);
    SPHX_REQUIRE(f->vx && f->vy && f->vz && f->h && f->m && f->rho && f->p && f->c && f->c11 && f->c12 && f->c13 &&
                     f->c22 && f->c23 && f->c33 && f->ax && f->ay && f->az && f->du,
                 SPHX_ERR_ARG, "momentum-energy needs v, h, m, rho, p, c, c11..c33, ax, ay, az and du");
    ws->counters.reserve(8);
    unsigned long long* dt = reinterpret_cast<unsigned long long*>(ws->counters.get()) + 3;
    fill64(ws, reinterpret_cast<uint64_t*>(dt), 0x7ff0000000000000ull, 1); // +inf
    MomStdFields mf{f->vx, f->vy, f->vz, f->h, f->m, f->rho, f->p, f->c, f->c11, f->c12, f->c13, f->c22, f->c23, f->c33,
                    f->ax, f->ay, f->az, f->du};
    unsigned blocks = pairBlocks(first, last);
    if (blocks > 0)
    {
        if (p->hydro_type == SPHX_F32)
            momentumStdKernel<float><<<blocks, kPairWaves * kWave, 0, ws->stream>>>(a, mf, ws->wh32.get(),
                                                                                  float(p->Kcour), dt);
        else
            momentumStdKernel<double><<<blocks, kPairWaves * kWave, 0, ws->stream>>>(a, mf, ws->wh64.get(),
                                                                                   float(p->Kcour), dt);
        SPHX_CHECK_LAUNCH();
    }
    unsigned long long bits = 0;
    SPHX_HIP(hipMemcpyAsync(&bits, dt, 8, hipMemcpyDeviceToHost, ws->stream));
    SPHX_HIP(hipStreamSynchronize(ws->stream));
    double r;
    std::memcpy(&r, &bits, 8);
    return r;
}

} // namespace sphx
