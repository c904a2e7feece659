// Generalised-volume-element (VE) SPH pair loops (sph/include/sph/hydro_ve/*):
// XMass (shared with std density, hydro_std.hip), VeDefGradh, EOS, IAD + div/curl v,
// AV switches, momentum-energy with optional AV cleaning.
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
    SPHX_REQUIRE(first == ws->nbFirst && last <= ws->nbLast && first <= last, SPHX_ERR_ARG,
                 "target range must start at the searched first index and stay inside the searched range");
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

template<class T>
const T* whdTable(sphx_workspace* ws)
{
    if constexpr (isMixed<T>) { return ws->whd32.get(); }
    else { return ws->whd64.get(); }
}

/* veDefGradhJLoop (hydro_ve/ve_def_gradh_kern.hpp:44-90): kx and the grad-h term */
template<class T>
__global__ void __launch_bounds__(kPairWaves * kWave)
    veDefGradhKernel(Common a, const T* __restrict__ h, const T* __restrict__ m, const T* __restrict__ xm,
                     const T* __restrict__ wh, const T* __restrict__ whd, T* __restrict__ kx, T* __restrict__ gradh)
{
    Lane t = laneTarget(a.L);
    if (!t.valid) { return; }
    const uint32_t i  = t.i;
    double         xi = a.x[i], yi = a.y[i], zi = a.z[i];
    T              hi = h[i], mi = m[i], xmi = xm[i];
    T              hInv = T(1) / hi;
    double         r2h  = 2.0 * double(hi);
    Acc<T>         kxs(xmi), wom(-T(3) * xmi), wr0(-T(3) * mi);
    for (unsigned k = 0; k < t.n; ++k)
    {
        uint32_t j = t.list[size_t(k) * kGroup];
        T        rx, ry, rz;
        separation(a.box, r2h, xi, yi, zi, a.x[j], a.y[j], a.z[j], rx, ry, rz);
        T dist  = fastSqrt(rx * rx + ry * ry + rz * rz);
        T vloc  = dist * hInv;
        T w     = lookup(wh, vloc);
        T dw    = lookup(whd, vloc);
        T dterh = -(T(3) * w + vloc * dw);
        T xmj   = xm[j];
        kxs.add(w * xmj);
        wom.add(dterh * xmj);
        wr0.add(dterh * m[j]);
    }
    // per-target epilogue in FP64, in the reference's order
    double dhInv  = double(hInv);
    double h3Inv  = dhInv * dhInv * dhInv;
    double kxi    = double(kxs.value()) * (a.K * h3Inv);
    double whom   = double(wom.value()) * (a.K * h3Inv * dhInv);
    double wrho0  = double(wr0.value()) * (a.K * h3Inv * dhInv);
    if constexpr (!isMixed<T>)
    {
        // keep the reference's T-typed rounding of hInv^3
        T th3 = hInv * hInv * hInv;
        kxi   = double(kxs.value()) * (a.K * double(th3));
        whom  = double(wom.value()) * (a.K * double(th3) * dhInv);
        wrho0 = double(wr0.value()) * (a.K * double(th3) * dhInv);
        h3Inv = double(th3);
    }
    whom          = whom * double(mi) / double(xmi) + (kxi - a.K * double(xmi) * h3Inv) * wrho0;
    double rhoi   = kxi * double(mi) / double(xmi);
    double dhdrho = -double(hi) / (rhoi * 3.0);
    kx[i]         = T(kxi);
    gradh[i]      = T(1.0 - dhdrho * whom);
}

/* VE ideal-gas EOS (hydro_ve/eos.hpp:53-95) */
template<class T>
__global__ void __launch_bounds__(256) eosVeKernel(const double* __restrict__ u, const double* __restrict__ temp,
                                                   const T* __restrict__ m, const T* __restrict__ kx,
                                                   const T* __restrict__ xm, const T* __restrict__ gradh,
                                                   T* __restrict__ prho, T* __restrict__ c, T* rhoOut, T* pOut,
                                                   double gamma, double cv, uint32_t first, uint32_t last)
{
    uint32_t i = first + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= last) { return; }
    T      rho = kx[i] * m[i] / xm[i];
    double ui  = temp ? cv * temp[i] : u[i];
    double tmp = ui * (gamma - 1.0);
    double pi  = double(rho) * tmp;
    prho[i]    = T(pi / double(kx[i] * m[i] * m[i] * gradh[i]));
    c[i]       = T(sqrt(gamma * tmp));
    if (rhoOut) { rhoOut[i] = rho; }
    if (pOut) { pOut[i] = T(pi); }
}

struct IadVeFields
{
    const void *h, *xm, *kx, *vx, *vy, *vz;
    void *      c11, *c12, *c13, *c22, *c23, *c33, *divv, *curlv, *dV11, *dV12, *dV13, *dV22, *dV23, *dV33;
};

/* IADJLoop (hydro_ve/iad_kern.hpp:44-109) + divV_curlVJLoop (hydro_ve/divv_curlv_kern.hpp:44-123)
 * in one neighbour pass. The velocity-gradient terms are linear in the target's IAD matrix C_i:
 *   dV_a[b] = sum_j (v_ja - v_ia) xm_j (-(C_i r_ij)_b W_ij) = -sum_c C_i[b][c] M[a][c],
 *   M[a][c] = sum_j (v_ja - v_ia) xm_j W_ij r_ij,c,
 * so the 9 moments M are accumulated alongside tau and combined with C_i after the inversion. */
template<class T>
__global__ void __launch_bounds__(kPairWaves * kWave)
    iadDivvCurlvKernel(Common a, IadVeFields f, const T* __restrict__ wh)
{
    Lane t = laneTarget(a.L);
    if (!t.valid) { return; }
    const T* __restrict__ h  = (const T*)f.h;
    const T* __restrict__ xm = (const T*)f.xm;
    const T* __restrict__ kx = (const T*)f.kx;
    const T* __restrict__ vx = (const T*)f.vx;
    const T* __restrict__ vy = (const T*)f.vy;
    const T* __restrict__ vz = (const T*)f.vz;

    const uint32_t i  = t.i;
    double         xi = a.x[i], yi = a.y[i], zi = a.z[i];
    T              hi = h[i], vxi = vx[i], vyi = vy[i], vzi = vz[i];
    T              hiInv = T(1) / hi;
    double         r2h   = 2.0 * double(hi);
    T              t11 = 0, t12 = 0, t13 = 0, t22 = 0, t23 = 0, t33 = 0;
    T              M[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (unsigned k = 0; k < t.n; ++k)
    {
        uint32_t j = t.list[size_t(k) * kGroup];
        T        rx, ry, rz;
        separation(a.box, r2h, xi, yi, zi, a.x[j], a.y[j], a.z[j], rx, ry, rz);
        T dist  = fastSqrt(rx * rx + ry * ry + rz * rz);
        T w     = lookup(wh, dist * hiInv);
        T xmj   = xm[j];
        T vol_w = fastDiv(xmj, kx[j]) * w;
        t11 += rx * rx * vol_w;
        t12 += rx * ry * vol_w;
        t13 += rx * rz * vol_w;
        t22 += ry * ry * vol_w;
        t23 += ry * rz * vol_w;
        t33 += rz * rz * vol_w;
        T xw  = xmj * w;
        T dvx = (vx[j] - vxi) * xw, dvy = (vy[j] - vyi) * xw, dvz = (vz[j] - vzi) * xw;
        M[0][0] += dvx * rx, M[0][1] += dvx * ry, M[0][2] += dvx * rz;
        M[1][0] += dvy * rx, M[1][1] += dvy * ry, M[1][2] += dvy * rz;
        M[2][0] += dvz * rx, M[2][1] += dvz * ry, M[2][2] += dvz * rz;
    }
    double cij[6];
    iadInvert(t11, t12, t13, t22, t23, t33, double(hi), a.K, cij);
    T c[6];
    for (int q = 0; q < 6; ++q)
        c[q] = T(cij[q]);
    st<T>(f.c11, i, c[0]), st<T>(f.c12, i, c[1]), st<T>(f.c13, i, c[2]);
    st<T>(f.c22, i, c[3]), st<T>(f.c23, i, c[4]), st<T>(f.c33, i, c[5]);

    // rows of the symmetric C_i
    T C[3][3] = {{c[0], c[1], c[2]}, {c[1], c[3], c[4]}, {c[2], c[4], c[5]}};
    T g[3][3]; // g[a][b] = d v_a / d x_b before normalisation
    for (int aa = 0; aa < 3; ++aa)
        for (int b = 0; b < 3; ++b)
            g[aa][b] = -(C[b][0] * M[aa][0] + C[b][1] * M[aa][1] + C[b][2] * M[aa][2]);

    T hiInv3   = hiInv * hiInv * hiInv;
    T norm_kxi = T(a.K * double(hiInv3) / double(kx[i]));
    st<T>(f.divv, i, norm_kxi * (g[0][0] + g[1][1] + g[2][2]));
    if (f.curlv)
    {
        T cx = g[2][1] - g[1][2], cy = g[0][2] - g[2][0], cz = g[1][0] - g[0][1];
        st<T>(f.curlv, i, norm_kxi * fastSqrt(cx * cx + (cy * cy + cz * cz)));
    }
    if (f.dV11)
    {
        st<T>(f.dV11, i, norm_kxi * g[0][0]);
        st<T>(f.dV12, i, norm_kxi * (g[0][1] + g[1][0]));
        st<T>(f.dV13, i, norm_kxi * (g[0][2] + g[2][0]));
        st<T>(f.dV22, i, norm_kxi * g[1][1]);
        st<T>(f.dV23, i, norm_kxi * (g[1][2] + g[2][1]));
        st<T>(f.dV33, i, norm_kxi * g[2][2]);
    }
}

struct AvFields
{
    const void *h, *c, *xm, *kx, *vx, *vy, *vz, *divv, *c11, *c12, *c13, *c22, *c23, *c33;
    void*       alpha;
};

/* AVswitchesJLoop (hydro_ve/av_switches_kern.hpp:44-137). grad(div v) is linear in C_i:
 * gd_b = -K/h_i^3 sum_c C_i[b][c] sum_j f_j W_ij r_ij,c with f_j = xm_j/kx_j (divv_i - divv_j). */
template<class T>
__global__ void __launch_bounds__(kPairWaves * kWave)
    avSwitchesKernel(Common a, AvFields f, const T* __restrict__ wh, double dt, T alphamin, T alphamax, T decay)
{
    Lane t = laneTarget(a.L);
    if (!t.valid) { return; }
    const T* __restrict__ h    = (const T*)f.h;
    const T* __restrict__ c    = (const T*)f.c;
    const T* __restrict__ xm   = (const T*)f.xm;
    const T* __restrict__ kx   = (const T*)f.kx;
    const T* __restrict__ vx   = (const T*)f.vx;
    const T* __restrict__ vy   = (const T*)f.vy;
    const T* __restrict__ vz   = (const T*)f.vz;
    const T* __restrict__ divv = (const T*)f.divv;

    const uint32_t i  = t.i;
    double         xi = a.x[i], yi = a.y[i], zi = a.z[i];
    T              hi = h[i], ci = c[i], vxi = vx[i], vyi = vy[i], vzi = vz[i], divv_i = divv[i];
    T              hiInv = T(1) / hi;
    double         r2h   = 2.0 * double(hi);
    T              vsig  = T(1.e-40) * ci;
    T              S[3]  = {0, 0, 0};
    for (unsigned k = 0; k < t.n; ++k)
    {
        uint32_t j = t.list[size_t(k) * kGroup];
        T        rx, ry, rz;
        separation(a.box, r2h, xi, yi, zi, a.x[j], a.y[j], a.z[j], rx, ry, rz);
        T dist = fastSqrt(rx * rx + ry * ry + rz * rz);
        T rv   = rx * (vxi - vx[j]) + ry * (vyi - vy[j]) + rz * (vzi - vz[j]);
        T vs   = rv < T(0) ? ci + c[j] - T(3) * fastDiv(rv, dist) : T(0);
        vsig   = vsig < vs ? vs : vsig;
        T w    = lookup(wh, dist * hiInv);
        T fw   = fastDiv(xm[j], kx[j]) * (divv_i - divv[j]) * w;
        S[0] += fw * rx, S[1] += fw * ry, S[2] += fw * rz;
    }
    T c11i = ld<T>(f.c11, i), c12i = ld<T>(f.c12, i), c13i = ld<T>(f.c13, i);
    T c22i = ld<T>(f.c22, i), c23i = ld<T>(f.c23, i), c33i = ld<T>(f.c33, i);
    T kh3  = T(a.K * double(hiInv * hiInv * hiInv));
    T gx   = -(c11i * S[0] + c12i * S[1] + c13i * S[2]) * kh3;
    T gy   = -(c12i * S[0] + c22i * S[1] + c23i * S[2]) * kh3;
    T gz   = -(c13i * S[0] + c23i * S[1] + c33i * S[2]) * kh3;
    T graddivv = fastSqrt(gx * gx + gy * gy + gz * gz);

    T alpha_i  = ld<T>(f.alpha, i);
    T alphaloc = 0;
    if (divv_i < T(0))
    {
        T a_const = hi * hi * graddivv;
        alphaloc  = alphamax * a_const / (a_const + hi * fabs(divv_i) + T(0.05) * ci);
    }
    if (alphaloc >= alpha_i) { alpha_i = alphaloc; }
    else
    {
        T decayT   = hi / (decay * vsig);
        T alphadot = alphaloc >= alphamin ? (alphaloc - alpha_i) / decayT : (alphamin - alpha_i) / decayT;
        alpha_i    = T(double(alpha_i) + double(alphadot) * dt);
    }
    st<T>(f.alpha, i, alpha_i);
}

struct MomVeFields
{
    const void *vx, *vy, *vz, *h, *m, *prho, *tdpdTrho, *c, *c11, *c12, *c13, *c22, *c23, *c33, *kx, *xm, *alpha;
    const void *dV11, *dV12, *dV13, *dV22, *dV23, *dV33;
    void *      ax, *ay, *az;
    double*     du;
};

//! avRvCorrection (hydro_ve/momentum_energy_kern.hpp:44-63) with the reference's symv layout
template<class T>
__device__ __forceinline__ T avRvCorrection(T rx, T ry, T rz, T eta_ab, T eta_crit, const T gi[6], const T gj[6])
{
    auto q = [rx, ry, rz](const T* g)
    {
        T s0 = g[0] * rx + g[1] * ry + g[2] * rz;
        T s1 = g[3] * ry + g[4] * rz;
        T s2 = g[5] * rz;
        return rx * s0 + (ry * s1 + rz * s2);
    };
    T dmy1 = q(gi), dmy2 = q(gj), dmy3 = T(1);
    if (eta_ab < eta_crit)
    {
        T e  = T(5) * (eta_ab - eta_crit);
        dmy3 = exp(-e * e);
    }
    T A_ab   = dmy2 != T(0) ? dmy1 / dmy2 : T(0);
    T A_abp1 = T(1) + A_ab;
    T r      = T(4) * A_ab / (A_abp1 * A_abp1);
    r        = r < T(1) ? r : T(1);
    r        = r > T(0) ? r : T(0);
    T phi_ab = T(0.5) * dmy3 * r;
    return -phi_ab * (dmy1 + dmy2);
}

/* momentumAndEnergyJLoop<avClean> (hydro_ve/momentum_energy_kern.hpp:66-222) + tsKCourant */
template<class T, bool avClean>
__global__ void __launch_bounds__(kPairWaves * kWave)
    momentumVeKernel(Common a, MomVeFields f, const T* __restrict__ wh, T Atmin, T Atmax, T ramp, float Kcour,
                     unsigned long long* minDt)
{
    Lane   t      = laneTarget(a.L);
    double dtLane = INFINITY;
    if (t.valid)
    {
        const T* __restrict__ vx    = (const T*)f.vx;
        const T* __restrict__ vy    = (const T*)f.vy;
        const T* __restrict__ vz    = (const T*)f.vz;
        const T* __restrict__ h     = (const T*)f.h;
        const T* __restrict__ m     = (const T*)f.m;
        const T* __restrict__ prho  = (const T*)f.prho;
        const T* __restrict__ c     = (const T*)f.c;
        const T* __restrict__ c11   = (const T*)f.c11;
        const T* __restrict__ c12   = (const T*)f.c12;
        const T* __restrict__ c13   = (const T*)f.c13;
        const T* __restrict__ c22   = (const T*)f.c22;
        const T* __restrict__ c23   = (const T*)f.c23;
        const T* __restrict__ c33   = (const T*)f.c33;
        const T* __restrict__ kx    = (const T*)f.kx;
        const T* __restrict__ xm    = (const T*)f.xm;
        const T* __restrict__ alpha = (const T*)f.alpha;

        const uint32_t i  = t.i;
        double         xi = a.x[i], yi = a.y[i], zi = a.z[i];
        T vxi = vx[i], vyi = vy[i], vzi = vz[i];
        T hi = h[i], mi = m[i], ci = c[i], kxi = kx[i], alpha_i = alpha[i], xmassi = xm[i];
        T rhoi   = kxi * mi / xmassi;
        T prhoi  = prho[i];
        T hiInv  = T(1) / hi;
        T hiInv3 = hiInv * hiInv * hiInv;
        T c11i = c11[i], c12i = c12[i], c13i = c13[i], c22i = c22[i], c23i = c23[i], c33i = c33[i];
        T gradV_i[6] = {0, 0, 0, 0, 0, 0};
        if constexpr (avClean)
        {
            gradV_i[0] = ld<T>(f.dV11, i), gradV_i[1] = ld<T>(f.dV12, i), gradV_i[2] = ld<T>(f.dV13, i);
            gradV_i[3] = ld<T>(f.dV22, i), gradV_i[4] = ld<T>(f.dV23, i), gradV_i[5] = ld<T>(f.dV33, i);
        }
        T      eta_crit = T(cbrt(32.0 * M_PI / 3.0 / double(t.n + 1)));
        double r2h      = 2.0 * double(hi);

        T      maxvs = 0;
        T      momx = 0, momy = 0, momz = 0;
        Acc<T> energy, avEnergy;
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
            T Wi    = hiInv3 * lookup(wh, v1);
            T Wj    = hjInv * hjInv * hjInv * lookup(wh, v2);

            T tA1_i = -(c11i * rx + c12i * ry + c13i * rz) * Wi;
            T tA2_i = -(c12i * rx + c22i * ry + c23i * rz) * Wi;
            T tA3_i = -(c13i * rx + c23i * ry + c33i * rz) * Wi;
            T c11j = c11[j], c12j = c12[j], c13j = c13[j], c22j = c22[j], c23j = c23[j], c33j = c33[j];
            T tA1_j = -(c11j * rx + c12j * ry + c13j * rz) * Wj;
            T tA2_j = -(c12j * rx + c22j * ry + c23j * rz) * Wj;
            T tA3_j = -(c13j * rx + c23j * ry + c33j * rz) * Wj;

            T mj = m[j], cj = c[j], kxj = kx[j], xmassj = xm[j];
            T rhoj = kxj * mj / xmassj;

            T rv = rx * vx_ij + ry * vy_ij + rz * vz_ij;
            if constexpr (avClean)
            {
                T gradV_j[6] = {ld<T>(f.dV11, j), ld<T>(f.dV12, j), ld<T>(f.dV13, j),
                                ld<T>(f.dV22, j), ld<T>(f.dV23, j), ld<T>(f.dV33, j)};
                rv += avRvCorrection(rx, ry, rz, v1 < v2 ? v1 : v2, eta_crit, gradV_i, gradV_j);
            }
            T wij  = fastDiv(rv, dist);
            T visc = artificialViscosity(alpha_i, alpha[j], ci, cj, wij);
            T vsig = T(0.5) * (ci + cj) - T(2) * wij;
            maxvs  = vsig > maxvs ? vsig : maxvs;

            T a_mom, b_mom;
            T Atwood = fastDiv(fabs(rhoi - rhoj), rhoi + rhoj);
            if (Atwood < Atmin)
            {
                a_mom = xmassi * xmassi;
                b_mom = xmassj * xmassj;
            }
            else if (Atwood > Atmax)
            {
                a_mom = xmassi * xmassj;
                b_mom = a_mom;
            }
            else
            {
                T sigma = ramp * (Atwood - Atmin);
                a_mom   = pow(xmassi, T(2) - sigma) * pow(xmassj, sigma);
                b_mom   = pow(xmassj, T(2) - sigma) * pow(xmassi, sigma);
            }

            T a_visc   = fastDiv(mj, rhoi) * visc;
            T b_visc   = fastDiv(mj, rhoj) * visc;
            T a_visc_x = T(0.5) * (a_visc * tA1_i + b_visc * tA1_j);
            T a_visc_y = T(0.5) * (a_visc * tA2_i + b_visc * tA2_j);
            T a_visc_z = T(0.5) * (a_visc * tA3_i + b_visc * tA3_j);
            avEnergy.add(a_visc_x * vx_ij + a_visc_y * vy_ij + a_visc_z * vz_ij);
            energy.add(mj * a_mom * (vx_ij * tA1_i + vy_ij * tA2_i + vz_ij * tA3_i));

            T mom_i = mj * prhoi * a_mom;
            T mom_j = mj * prho[j] * b_mom;
            momx += mom_i * tA1_i + mom_j * tA1_j + a_visc_x;
            momy += mom_i * tA2_i + mom_j * tA2_j + a_visc_y;
            momz += mom_i * tA3_i + mom_j * tA3_j + a_visc_z;
        }
        T avE    = avEnergy.value() > T(0) ? avEnergy.value() : T(0);
        T eCoeff = f.tdpdTrho ? ld<T>(f.tdpdTrho, i) : prhoi;
        f.du[i]  = a.K * double(eCoeff * energy.value() + T(0.5) * avE);
        st<T>(f.ax, i, T(-a.K * double(momx)));
        st<T>(f.ay, i, T(-a.K * double(momy)));
        st<T>(f.az, i, T(-a.K * double(momz)));
        dtLane = courantDt(maxvs, hi, ci, Kcour);
    }
    reduceMinDt<T>(minDt, dtLane);
}

unsigned long long* resetDt(sphx_workspace* ws)
{
    ws->counters.reserve(8);
    unsigned long long* dt  = reinterpret_cast<unsigned long long*>(ws->counters.get()) + 3;
    fill64(ws, reinterpret_cast<uint64_t*>(dt), 0x7ff0000000000000ull, 1); // +inf
    return dt;
}

double readDt(sphx_workspace* ws, unsigned long long* dt)
{
    unsigned long long bits = 0;
    SPHX_HIP(hipMemcpyAsync(&bits, dt, 8, hipMemcpyDeviceToHost, ws->stream));
    SPHX_HIP(hipStreamSynchronize(ws->stream));
    double r;
    std::memcpy(&r, &bits, 8);
    return r;
}

} // namespace

void veDefGradh(sphx_workspace* ws, const DevBox& box, const sphx_fields_t* f, const sphx_params_t* p, uint32_t first,
                uint32_t last)
{
    Common a = makeCommon(ws, box, f, p, first, last);
    SPHX_REQUIRE(f->h && f->m && f->xm && f->kx && f->gradh, SPHX_ERR_ARG, "VeDefGradh needs h, m, xm, kx, gradh");
    unsigned blocks = pairBlocks(first, last);
    if (blocks == 0) { return; }
#define LAUNCH(T)                                                                                                      \
    veDefGradhKernel<T><<<blocks, kPairWaves * kWave, 0, ws->stream>>>(                                                \
        a, (const T*)f->h, (const T*)f->m, (const T*)f->xm, whTable<T>(ws), whdTable<T>(ws), (T*)f->kx, (T*)f->gradh)
    if (p->hydro_type == SPHX_F32) LAUNCH(float);
    else LAUNCH(double);
#undef LAUNCH
    SPHX_CHECK_LAUNCH();
}

void eosVe(sphx_workspace* ws, const sphx_fields_t* f, const sphx_params_t* p, uint32_t first, uint32_t last)
{
    SPHX_REQUIRE(f->m && f->kx && f->xm && f->gradh && f->prho && f->c && (f->temp || f->u), SPHX_ERR_ARG,
                 "VE EOS needs m, kx, xm, gradh, prho, c and temp or u");
    SPHX_REQUIRE(first <= last, SPHX_ERR_ARG, "invalid range");
    if (last == first) { return; }
    double cv = 8.317e7 / p->muiConst / (p->gamma - 1.0);
    if (p->hydro_type == SPHX_F32)
    {
        cv = double(float(8.317e7f / float(p->muiConst) / (p->gamma - 1.0)));
        eosVeKernel<float><<<ceilDiv(last - first, 256), 256, 0, ws->stream>>>(
            f->u, f->temp, (const float*)f->m, (const float*)f->kx, (const float*)f->xm, (const float*)f->gradh,
            (float*)f->prho, (float*)f->c, (float*)f->rho, (float*)f->p, p->gamma, cv, first, last);
    }
    else
    {
        eosVeKernel<double><<<ceilDiv(last - first, 256), 256, 0, ws->stream>>>(
            f->u, f->temp, (const double*)f->m, (const double*)f->kx, (const double*)f->xm, (const double*)f->gradh,
            (double*)f->prho, (double*)f->c, (double*)f->rho, (double*)f->p, p->gamma, cv, first, last);
    }
    SPHX_CHECK_LAUNCH();
}

void iadDivvCurlv(sphx_workspace* ws, const DevBox& box, const sphx_fields_t* f, const sphx_params_t* p,
                  uint32_t first, uint32_t last)
{
    Common a = makeCommon(ws, box, f, p, first, last);
    SPHX_REQUIRE(f->h && f->xm && f->kx && f->vx && f->vy && f->vz && f->c11 && f->c12 && f->c13 && f->c22 &&
                     f->c23 && f->c33 && f->divv,
                 SPHX_ERR_ARG, "IadDivvCurlv needs h, xm, kx, v, c11..c33 and divv");
    bool gradV = f->dV11 != nullptr;
    SPHX_REQUIRE(!gradV || (f->dV12 && f->dV13 && f->dV22 && f->dV23 && f->dV33), SPHX_ERR_ARG,
                 "dV11..dV33 must be all set or all NULL");
    IadVeFields vf{f->h,   f->xm,   f->kx,    f->vx,   f->vy,   f->vz,   f->c11,  f->c12,  f->c13,  f->c22,
                   f->c23, f->c33, f->divv, f->curlv, f->dV11, f->dV12, f->dV13, f->dV22, f->dV23, f->dV33};
    unsigned    blocks = pairBlocks(first, last);
    if (blocks == 0) { return; }
    if (p->hydro_type == SPHX_F32)
        iadDivvCurlvKernel<float><<<blocks, kPairWaves * kWave, 0, ws->stream>>>(a, vf, ws->wh32.get());
    else
        iadDivvCurlvKernel<double><<<blocks, kPairWaves * kWave, 0, ws->stream>>>(a, vf, ws->wh64.get());
    SPHX_CHECK_LAUNCH();
}

void avSwitches(sphx_workspace* ws, const DevBox& box, const sphx_fields_t* f, const sphx_params_t* p, uint32_t first,
                uint32_t last)
{
    Common a = makeCommon(ws, box, f, p, first, last);
    SPHX_REQUIRE(f->h && f->c && f->xm && f->kx && f->vx && f->vy && f->vz && f->divv && f->c11 && f->c12 && f->c13 &&
                     f->c22 && f->c23 && f->c33 && f->alpha,
                 SPHX_ERR_ARG, "AVswitches needs h, c, xm, kx, v, divv, c11..c33 and alpha");
    AvFields af{f->h, f->c, f->xm, f->kx, f->vx, f->vy, f->vz, f->divv, f->c11, f->c12, f->c13, f->c22, f->c23, f->c33,
                f->alpha};
    unsigned blocks = pairBlocks(first, last);
    if (blocks == 0) { return; }
    if (p->hydro_type == SPHX_F32)
        avSwitchesKernel<float><<<blocks, kPairWaves * kWave, 0, ws->stream>>>(
            a, af, ws->wh32.get(), p->minDt, float(p->alphamin), float(p->alphamax), float(p->decay_constant));
    else
        avSwitchesKernel<double><<<blocks, kPairWaves * kWave, 0, ws->stream>>>(
            a, af, ws->wh64.get(), p->minDt, p->alphamin, p->alphamax, p->decay_constant);
    SPHX_CHECK_LAUNCH();
}

double momentumEnergyVe(sphx_workspace* ws, const DevBox& box, const sphx_fields_t* f, const sphx_params_t* p,
                        uint32_t first, uint32_t last, bool avClean)
{
    Common a = makeCommon(ws, box, f, p, first, last);
    SPHX_REQUIRE(f->vx && f->vy && f->vz && f->h && f->m && f->prho && f->c && f->c11 && f->c12 && f->c13 && f->c22 &&
                     f->c23 && f->c33 && f->kx && f->xm && f->alpha && f->ax && f->ay && f->az && f->du,
                 SPHX_ERR_ARG, "VE momentum-energy needs v, h, m, prho, c, c11..c33, kx, xm, alpha, a and du");
    SPHX_REQUIRE(!avClean || (f->dV11 && f->dV12 && f->dV13 && f->dV22 && f->dV23 && f->dV33), SPHX_ERR_ARG,
                 "AV cleaning needs dV11..dV33");
    MomVeFields mf{f->vx,  f->vy,   f->vz,   f->h,    f->m,    f->prho, f->tdpdTrho, f->c,  f->c11, f->c12,
                   f->c13, f->c22,  f->c23,  f->c33,  f->kx,   f->xm,   f->alpha,    f->dV11, f->dV12, f->dV13,
                   f->dV22, f->dV23, f->dV33, f->ax,   f->ay,   f->az,   f->du};
    unsigned long long* dt     = resetDt(ws);
    unsigned            blocks = pairBlocks(first, last);
    if (blocks > 0)
    {
#define LAUNCH(T, C)                                                                                                   \
    momentumVeKernel<T, C><<<blocks, kPairWaves * kWave, 0, ws->stream>>>(                                             \
        a, mf, whTable<T>(ws), T(p->Atmin), T(p->Atmax), T(p->ramp), float(p->Kcour), dt)
        if (p->hydro_type == SPHX_F32)
        {
            if (avClean) LAUNCH(float, true);
            else LAUNCH(float, false);
        }
        else
        {
            if (avClean) LAUNCH(double, true);
            else LAUNCH(double, false);
        }
#undef LAUNCH
        SPHX_CHECK_LAUNCH();
    }
    return readDt(ws, dt);
}

} // namespace sphx
