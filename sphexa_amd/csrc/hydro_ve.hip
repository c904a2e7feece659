// Generalised-volume-element (VE) SPH pair loops (sph/include/sph/hydro_ve/*):
// XMass (shared with std density, hydro_std.hip), VeDefGradh, EOS, IAD + div/curl v,
// AV switches, momentum-energy with optional AV cleaning. Neighbour data is read from packed
// per-particle records (records.hpp).

#include "records.hpp"

namespace sphx
{

namespace
{

constexpr unsigned kPackThreads = 256;

template<class T>
__global__ void __launch_bounds__(kPackThreads) packGradh(PairArgs a, const T* __restrict__ m,
                                                          const T* __restrict__ xm, RecGradh<T>* __restrict__ rec)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.numParticles) { return; }
    rec[i] = gradhRecord<T>(coordOf<T>(a.box, 0, a.x[i]), coordOf<T>(a.box, 1, a.y[i]), coordOf<T>(a.box, 2, a.z[i]),
                            m[i], xm[i]);
}

/* veDefGradhJLoop (hydro_ve/ve_def_gradh_kern.hpp:44-90): kx and the grad-h term */
template<class T, bool Wide>
__global__ void SPHX_LIGHT_LOOP(T)
    veDefGradhKernel(PairArgs a, const RecGradh<T>* __restrict__ rec, const T* __restrict__ h,
                     SincKernel<T> kern, T* __restrict__ kx, T* __restrict__ gradh, RecIadVe<T>* __restrict__ next,
                     const T* __restrict__ vx, const T* __restrict__ vy, const T* __restrict__ vz)
{
    Lane t = laneTarget(a.L);
    if (!t.valid) { return; }
    const uint32_t    i   = t.i;
    const RecGradh<T> ri  = rec[i];
    T                 hi  = h[i], mi = ri.m, xmi = ri.xm;
    T                 hInv = T(1) / hi;
    double            r2h  = 2.0 * double(hi);
    Acc<T>            kxs(xmi), wom(-T(3) * xmi), wr0(-T(3) * mi);
    withFold(recordNeedsFold<T, Wide>(a.box, ri, r2h), [&](auto fold) {
        constexpr int                 U = kBatchLight<T>;
        NeighborBatch<U, RecGradh<T>> nb(t, rec);
        for (unsigned k = 0; k < t.n; k += U)
        {
            nb.load(k);
            typename SincKernel<T>::FetchWD w[U];
#pragma unroll
            for (int u = 0; u < U; ++u)
                w[u] = kern.fetchWD(pairSeparation<decltype(fold)::value, T>(a.box, r2h, ri, nb.r[u]).d * hInv);
#pragma unroll
            for (int u = 0; u < U; ++u)
            {
                T wv, dw;
                values(w[u], wv, dw);
                wv = nb.live(k, u, wv), dw = nb.live(k, u, dw);
                T dterh = -(T(3) * wv + w[u].v * dw);
                kxs.add(wv * nb.r[u].xm);
                wom.add(dterh * nb.r[u].xm);
                wr0.add(dterh * nb.r[u].m);
            }
        }
    });
    // per-target epilogue in FP64 following the reference's order; h^-3 is rounded in T as there
    T      th3   = hInv * hInv * hInv;
    double dhInv = double(hInv);
    double h3Inv = double(th3);
    double kxi   = double(kxs.value()) * (a.K * h3Inv);
    double whom  = double(wom.value()) * (a.K * h3Inv * dhInv);
    double wrho0 = double(wr0.value()) * (a.K * h3Inv * dhInv);
    whom          = whom * double(mi) / double(xmi) + (kxi - a.K * double(xmi) * h3Inv) * wrho0;
    double rhoi   = kxi * double(mi) / double(xmi);
    double dhdrho = -double(hi) / (rhoi * 3.0);
    const T kxT   = T(kxi);
    kx[i]         = kxT;
    gradh[i]      = T(1.0 - dhdrho * whom);
    if (next) { next[i] = iadVeRecord<T>(ri.x, ri.y, ri.z, xmi, kxT, vx[i], vy[i], vz[i]); }
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

/* VE isothermal EOS (hydro_ve/eos.hpp:97-123, hydro_ve/eos_gpu.cu:89-103): rho = kx m / xm in the
 * HydroType, p = rho c0^2, prho = p / (kx m^2 gradh) (in FP64, rounded once in the FP32 form), c = c0,
 * rho and p stored when given, temp = 0 */
template<class T>
__global__ void __launch_bounds__(256) eosVeIsothermalKernel(const T* __restrict__ m, const T* __restrict__ kx,
                                                             const T* __restrict__ xm, const T* __restrict__ gradh,
                                                             T* __restrict__ prho, T* __restrict__ c, T* rhoOut,
                                                             T* pOut, double* temp, T c0, uint32_t first,
                                                             uint32_t last)
{
    uint32_t i = first + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= last) { return; }
    const T      rho = kx[i] * m[i] / xm[i];
    const double cc  = double(c0);
    const double pi  = double(rho) * cc * cc;
    prho[i]          = T(pi / double(kx[i] * m[i] * m[i] * gradh[i]));
    c[i]             = c0;
    if (rhoOut) { rhoOut[i] = rho; }
    if (pOut) { pOut[i] = T(pi); }
    if (temp) { temp[i] = 0.0; }
}

/* VE polytropic EOS (computePolytropicEOS_Impl, hydro_ve/eos.hpp:125-151): rho = kx m / xm in the
 * HydroType, p = K rho^n and c = sqrt(n p / rho) in FP64 (polytropicEOS, sph/include/sph/eos.hpp:77-86),
 * prho = p / (kx m^2 gradh), rho and p stored when given, temp = 0. c is stored as the reference's CPU
 * path does; its GPU kernel leaves c unchanged (hydro_ve/eos_gpu.cu:134). */
template<class T>
__global__ void __launch_bounds__(256) eosVePolytropicKernel(const T* __restrict__ m, const T* __restrict__ kx,
                                                             const T* __restrict__ xm, const T* __restrict__ gradh,
                                                             T* __restrict__ prho, T* __restrict__ c, T* rhoOut,
                                                             T* pOut, double* temp, double K, double n,
                                                             uint32_t first, uint32_t last)
{
    uint32_t i = first + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= last) { return; }
    const T      rho = kx[i] * m[i] / xm[i];
    const double r   = double(rho);
    const double pi  = K * pow(r, n);
    prho[i]          = T(pi / double(kx[i] * m[i] * m[i] * gradh[i]));
    c[i]             = T(sqrt(n * pi / r));
    if (rhoOut) { rhoOut[i] = rho; }
    if (pOut) { pOut[i] = T(pi); }
    if (temp) { temp[i] = 0.0; }
}

/* Record builders shared by the pack kernels and the epilogues that write the next pass's records
 * (SPHX_FLAG_CHAIN), so that both give the same bits. */
template<class T>
__device__ __forceinline__ RecAv<T> avRecord(Coord<T> x, Coord<T> y, Coord<T> z, T xm, T kx, T c, T divv, T vx, T vy,
                                             T vz)
{
    RecAv<T> r;
    r.x = x, r.y = y, r.z = z;
    r.vol  = fastDiv(xm, kx);
    r.c    = c;
    r.divv = divv;
    r.vx = vx, r.vy = vy, r.vz = vz;
    return r;
}

//! C = {c11, c12, c13, c22, c23, c33}
template<class T>
__device__ __forceinline__ RecMomVe<T> momVeRecord(Coord<T> x, Coord<T> y, Coord<T> z, T vx, T vy, T vz, T h, T m,
                                                   T c, T xm, T kx, T alpha, T prho, const T C[6])
{
    RecMomVe<T> r;
    r.x = x, r.y = y, r.z = z;
    r.vx = vx, r.vy = vy, r.vz = vz;
    r.hInv  = T(1) / h;
    r.m     = m;
    r.c     = c;
    r.xm    = xm;
    r.rho   = kx * m / xm;
    r.vol   = fastDiv(m, r.rho);
    r.alpha = alpha;
    r.prho  = prho;
    r.c11 = C[0], r.c12 = C[1], r.c13 = C[2], r.c22 = C[3], r.c23 = C[4], r.c33 = C[5];
    return r;
}

struct IadVeFields
{
    const void *h, *xm, *kx, *vx, *vy, *vz;
    void *      c11, *c12, *c13, *c22, *c23, *c33, *divv, *curlv, *dV11, *dV12, *dV13, *dV22, *dV23, *dV33;
};

template<class T>
__global__ void __launch_bounds__(kPackThreads) packIadVe(PairArgs a, IadVeFields f, RecIadVe<T>* __restrict__ rec)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.numParticles) { return; }
    rec[i] = iadVeRecord<T>(coordOf<T>(a.box, 0, a.x[i]), coordOf<T>(a.box, 1, a.y[i]), coordOf<T>(a.box, 2, a.z[i]),
                            ld<T>(f.xm, i), ld<T>(f.kx, i), ld<T>(f.vx, i), ld<T>(f.vy, i), ld<T>(f.vz, i));
}

/* IADJLoop (hydro_ve/iad_kern.hpp:44-109) + divV_curlVJLoop (hydro_ve/divv_curlv_kern.hpp:44-123)
 * in one neighbour pass. The velocity-gradient terms are linear in the target's IAD matrix C_i:
 *   dV_a[b] = sum_j (v_ja - v_ia) xm_j (-(C_i r_ij)_b W_ij) = -sum_c C_i[b][c] M[a][c],
 *   M[a][c] = sum_j (v_ja - v_ia) xm_j W_ij r_ij,c,
 * so the 9 moments M are accumulated alongside tau and combined with C_i after the inversion. */
template<class T, bool Wide>
__global__ void SPHX_LIGHT_LOOP(T)
    iadDivvCurlvKernel(PairArgs a, IadVeFields f, const RecIadVe<T>* __restrict__ rec, SincKernel<T> kern,
                       RecAv<T>* __restrict__ next, const T* __restrict__ cNext)
{
    Lane t = laneTarget(a.L);
    if (!t.valid) { return; }
    const uint32_t    i     = t.i;
    const RecIadVe<T> ri    = rec[i];
    T                 hi    = ld<T>(f.h, i);
    T                 hiInv = T(1) / hi;
    double            r2h   = 2.0 * double(hi);
    T                 t11 = 0, t12 = 0, t13 = 0, t22 = 0, t23 = 0, t33 = 0;
    T                 M[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    withFold(recordNeedsFold<T, Wide>(a.box, ri, r2h), [&](auto fold) {
        constexpr int                 U = kBatchLight<T>;
        NeighborBatch<U, RecIadVe<T>> nb(t, rec);
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
                const RecIadVe<T>& rj = nb.r[u];
                T rx = r[u].x, ry = r[u].y, rz = r[u].z;
                T wv    = nb.live(k, u, value(w[u]));
                T vol_w = rj.vol * wv;
                t11 += rx * rx * vol_w;
                t12 += rx * ry * vol_w;
                t13 += rx * rz * vol_w;
                t22 += ry * ry * vol_w;
                t23 += ry * rz * vol_w;
                t33 += rz * rz * vol_w;
                T xw  = rj.xm * wv;
                T dvx = (rj.vx - ri.vx) * xw, dvy = (rj.vy - ri.vy) * xw, dvz = (rj.vz - ri.vz) * xw;
                M[0][0] += dvx * rx, M[0][1] += dvx * ry, M[0][2] += dvx * rz;
                M[1][0] += dvy * rx, M[1][1] += dvy * ry, M[1][2] += dvy * rz;
                M[2][0] += dvz * rx, M[2][1] += dvz * ry, M[2][2] += dvz * rz;
            }
        }
    });
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
    T kxi      = ld<T>(f.kx, i);
    T norm_kxi = T(a.K * double(hiInv3) / double(kxi));
    T divv_i   = norm_kxi * (g[0][0] + g[1][1] + g[2][2]);
    st<T>(f.divv, i, divv_i);
    if (next) { next[i] = avRecord<T>(ri.x, ri.y, ri.z, ri.xm, kxi, cNext[i], divv_i, ri.vx, ri.vy, ri.vz); }
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
    const void *m, *prho; // read only to write the momentum-energy records (next != NULL)
};

template<class T>
__global__ void __launch_bounds__(kPackThreads) packAv(PairArgs a, AvFields f, RecAv<T>* __restrict__ rec)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.numParticles) { return; }
    rec[i] = avRecord<T>(coordOf<T>(a.box, 0, a.x[i]), coordOf<T>(a.box, 1, a.y[i]), coordOf<T>(a.box, 2, a.z[i]),
                         ld<T>(f.xm, i), ld<T>(f.kx, i), ld<T>(f.c, i), ld<T>(f.divv, i), ld<T>(f.vx, i),
                         ld<T>(f.vy, i), ld<T>(f.vz, i));
}

/* AVswitchesJLoop (hydro_ve/av_switches_kern.hpp:44-137). grad(div v) is linear in C_i:
 * gd_b = -K/h_i^3 sum_c C_i[b][c] sum_j f_j W_ij r_ij,c with f_j = xm_j/kx_j (divv_i - divv_j). */
template<class T, bool Wide>
__global__ void SPHX_LIGHT_LOOP(T)
    avSwitchesKernel(PairArgs a, AvFields f, const RecAv<T>* __restrict__ rec, SincKernel<T> kern, double dt,
                     T alphamin, T alphamax, T decay, RecMomVe<T>* __restrict__ next)
{
    Lane t = laneTarget(a.L);
    if (!t.valid) { return; }
    const uint32_t i     = t.i;
    const RecAv<T> ri    = rec[i];
    T              hi    = ld<T>(f.h, i);
    T              hiInv = T(1) / hi;
    double         r2h   = 2.0 * double(hi);
    T              vsig  = T(1.e-40) * ri.c;
    T              S[3]  = {0, 0, 0};
    withFold(recordNeedsFold<T, Wide>(a.box, ri, r2h), [&](auto fold) {
        constexpr int              U = kBatchLight<T>;
        NeighborBatch<U, RecAv<T>> nb(t, rec);
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
                const RecAv<T>& rj = nb.r[u];
                T rx = r[u].x, ry = r[u].y, rz = r[u].z;
                T rv = rx * (ri.vx - rj.vx) + ry * (ri.vy - rj.vy) + rz * (ri.vz - rj.vz);
                T vs = nb.in(k, u) && rv < T(0) ? ri.c + rj.c - T(3) * fastDiv(rv, r[u].d) : T(0);
                vsig = vsig < vs ? vs : vsig;
                T fw = rj.vol * (ri.divv - rj.divv) * nb.live(k, u, value(w[u]));
                S[0] += fw * rx, S[1] += fw * ry, S[2] += fw * rz;
            }
        }
    });
    T c11i = ld<T>(f.c11, i), c12i = ld<T>(f.c12, i), c13i = ld<T>(f.c13, i);
    T c22i = ld<T>(f.c22, i), c23i = ld<T>(f.c23, i), c33i = ld<T>(f.c33, i);
    T kh3  = T(a.K * double(hiInv * hiInv * hiInv));
    T gx   = -(c11i * S[0] + c12i * S[1] + c13i * S[2]) * kh3;
    T gy   = -(c12i * S[0] + c22i * S[1] + c23i * S[2]) * kh3;
    T gz   = -(c13i * S[0] + c23i * S[1] + c33i * S[2]) * kh3;
    T graddivv = fastSqrt(gx * gx + gy * gy + gz * gz);

    T divv_i   = ri.divv;
    T alpha_i  = ld<T>(f.alpha, i);
    T alphaloc = 0;
    if (divv_i < T(0))
    {
        T a_const = hi * hi * graddivv;
        alphaloc  = alphamax * a_const / (a_const + hi * fabs(divv_i) + T(0.05) * ri.c);
    }
    if (alphaloc >= alpha_i) { alpha_i = alphaloc; }
    else
    {
        T decayT   = hi / (decay * vsig);
        T alphadot = alphaloc >= alphamin ? (alphaloc - alpha_i) / decayT : (alphamin - alpha_i) / decayT;
        alpha_i    = T(double(alpha_i) + double(alphadot) * dt);
    }
    st<T>(f.alpha, i, alpha_i);
    if (next)
    {
        const T C[6] = {c11i, c12i, c13i, c22i, c23i, c33i};
        next[i] = momVeRecord<T>(ri.x, ri.y, ri.z, ri.vx, ri.vy, ri.vz, hi, ld<T>(f.m, i), ri.c, ld<T>(f.xm, i),
                                 ld<T>(f.kx, i), alpha_i, ld<T>(f.prho, i), C);
    }
}

struct MomVeFields
{
    const void *vx, *vy, *vz, *h, *m, *prho, *tdpdTrho, *c, *c11, *c12, *c13, *c22, *c23, *c33, *kx, *xm, *alpha;
    const void *dV11, *dV12, *dV13, *dV22, *dV23, *dV33;
    void *      ax, *ay, *az;
    double*     du;
};

template<class T>
__global__ void __launch_bounds__(kPackThreads) packMomVe(PairArgs a, MomVeFields f, RecMomVe<T>* __restrict__ rec)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.numParticles) { return; }
    const T C[6] = {ld<T>(f.c11, i), ld<T>(f.c12, i), ld<T>(f.c13, i), ld<T>(f.c22, i), ld<T>(f.c23, i), ld<T>(f.c33, i)};
    rec[i] = momVeRecord<T>(coordOf<T>(a.box, 0, a.x[i]), coordOf<T>(a.box, 1, a.y[i]), coordOf<T>(a.box, 2, a.z[i]),
                            ld<T>(f.vx, i), ld<T>(f.vy, i), ld<T>(f.vz, i), ld<T>(f.h, i), ld<T>(f.m, i),
                            ld<T>(f.c, i), ld<T>(f.xm, i), ld<T>(f.kx, i), ld<T>(f.alpha, i), ld<T>(f.prho, i), C);
}

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

//! log2 / exp2 of the Atwood-ramp weights: the hardware transcendentals in the FP32 form
template<class T>
__device__ __forceinline__ T ramLog2(T v)
{
    if constexpr (isMixed<T>) { return __log2f(v); }
    else { return log2(v); }
}

template<class T>
__device__ __forceinline__ T ramExp2(T v)
{
    if constexpr (isMixed<T>) { return exp2f(v); }
    else { return exp2(v); }
}

/* momentumAndEnergyJLoop<avClean> (hydro_ve/momentum_energy_kern.hpp:66-222) + tsKCourant */
template<class T, bool avClean, bool Wide>
__global__ void SPHX_MOM_LOOP(T)
    momentumVeKernel(PairArgs a, MomVeFields f, const RecMomVe<T>* __restrict__ rec, SincKernel<T> kern, T Atmin,
                     T Atmax, T ramp, float Kcour, unsigned long long* minDt)
{
    Lane   t      = laneTarget(a.L);
    double dtLane = INFINITY;
    if (t.valid)
    {
        const uint32_t    i      = t.i;
        const RecMomVe<T> ri     = rec[i];
        T                 hi     = ld<T>(f.h, i);
        T                 hiInv  = ri.hInv;
        T                 hiInv3 = hiInv * hiInv * hiInv;
        T                 gradV_i[6] = {0, 0, 0, 0, 0, 0};
        if constexpr (avClean)
        {
            gradV_i[0] = ld<T>(f.dV11, i), gradV_i[1] = ld<T>(f.dV12, i), gradV_i[2] = ld<T>(f.dV13, i);
            gradV_i[3] = ld<T>(f.dV22, i), gradV_i[4] = ld<T>(f.dV23, i), gradV_i[5] = ld<T>(f.dV33, i);
        }
        T      eta_crit = T(cbrt(32.0 * M_PI / 3.0 / double(t.n + 1)));
        double r2h      = 2.0 * double(hi);

        T      maxvs = 0;
        T      momx = 0, momy = 0, momz = 0;
        Acc<T>       energy, avEnergy;
        withFold(recordNeedsFold<T, Wide>(a.box, ri, r2h), [&](auto fold) {
            constexpr int                 U = kBatchMom<T>;
            NeighborBatch<U, RecMomVe<T>> nb(t, rec);
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
                const uint32_t     j  = nb.j[u];
                const RecMomVe<T>& rj = nb.r[u];
                T rx = r[u].x, ry = r[u].y, rz = r[u].z, dist = r[u].d;
                T vx_ij = ri.vx - rj.vx, vy_ij = ri.vy - rj.vy, vz_ij = ri.vz - rj.vz;
                T v1    = wi[u].v;
                T v2    = wj[u].v;
                T Wi    = hiInv3 * nb.live(k, u, value(wi[u]));
                T Wj    = rj.hInv * rj.hInv * rj.hInv * nb.live(k, u, value(wj[u]));

                T tA1_i = -(ri.c11 * rx + ri.c12 * ry + ri.c13 * rz) * Wi;
                T tA2_i = -(ri.c12 * rx + ri.c22 * ry + ri.c23 * rz) * Wi;
                T tA3_i = -(ri.c13 * rx + ri.c23 * ry + ri.c33 * rz) * Wi;
                T tA1_j = -(rj.c11 * rx + rj.c12 * ry + rj.c13 * rz) * Wj;
                T tA2_j = -(rj.c12 * rx + rj.c22 * ry + rj.c23 * rz) * Wj;
                T tA3_j = -(rj.c13 * rx + rj.c23 * ry + rj.c33 * rz) * Wj;

                T rv = rx * vx_ij + ry * vy_ij + rz * vz_ij;
                if constexpr (avClean)
                {
                    T gradV_j[6] = {ld<T>(f.dV11, j), ld<T>(f.dV12, j), ld<T>(f.dV13, j),
                                    ld<T>(f.dV22, j), ld<T>(f.dV23, j), ld<T>(f.dV33, j)};
                    rv += avRvCorrection(rx, ry, rz, v1 < v2 ? v1 : v2, eta_crit, gradV_i, gradV_j);
                }
                T wij  = fastDiv(rv, dist);
                T visc = artificialViscosity(ri.alpha, rj.alpha, ri.c, rj.c, wij);
                T vsig = T(0.5) * (ri.c + rj.c) - T(2) * wij;
                maxvs  = nb.in(k, u) && vsig > maxvs ? vsig : maxvs;

                T a_mom, b_mom;
                T dRho   = ri.rho - rj.rho;
                T Atwood = fastDiv(dRho < T(0) ? -dRho : dRho, ri.rho + rj.rho);
                if (Atwood < Atmin)
                {
                    a_mom = ri.xm * ri.xm;
                    b_mom = rj.xm * rj.xm;
                }
                else if (Atwood > Atmax)
                {
                    a_mom = ri.xm * rj.xm;
                    b_mom = a_mom;
                }
                else
                {
                    // xm_i^(2-s) xm_j^s = xm_i^2 2^(s log2(xm_j/xm_i)), and symmetrically for b:
                    // one logarithm and two exponentials instead of four powers
                    T sigma = ramp * (Atwood - Atmin);
                    T lr    = ramLog2(fastDiv(rj.xm, ri.xm));
                    a_mom   = ri.xm * ri.xm * ramExp2(sigma * lr);
                    b_mom   = rj.xm * rj.xm * ramExp2(-sigma * lr);
                }

                T a_visc   = fastDiv(rj.m, ri.rho) * visc;
                T b_visc   = rj.vol * visc;
                T a_visc_x = T(0.5) * (a_visc * tA1_i + b_visc * tA1_j);
                T a_visc_y = T(0.5) * (a_visc * tA2_i + b_visc * tA2_j);
                T a_visc_z = T(0.5) * (a_visc * tA3_i + b_visc * tA3_j);
                avEnergy.add(a_visc_x * vx_ij + a_visc_y * vy_ij + a_visc_z * vz_ij);
                energy.add(rj.m * a_mom * (vx_ij * tA1_i + vy_ij * tA2_i + vz_ij * tA3_i));

                T mom_i = rj.m * ri.prho * a_mom;
                T mom_j = rj.m * rj.prho * b_mom;
                momx += mom_i * tA1_i + mom_j * tA1_j + a_visc_x;
                momy += mom_i * tA2_i + mom_j * tA2_j + a_visc_y;
                momz += mom_i * tA3_i + mom_j * tA3_j + a_visc_z;
              }
            }
        });
        T avE    = avEnergy.value() > T(0) ? avEnergy.value() : T(0);
        T eCoeff = f.tdpdTrho ? ld<T>(f.tdpdTrho, i) : ri.prho;
        f.du[i]  = a.K * double(eCoeff * energy.value() + T(0.5) * avE);
        st<T>(f.ax, i, T(-a.K * double(momx)));
        st<T>(f.ay, i, T(-a.K * double(momy)));
        st<T>(f.az, i, T(-a.K * double(momz)));
        dtLane = courantDt(maxvs, hi, ri.c, Kcour);
    }
    reduceMinDt<T>(minDt, dtLane, a.L.groups);
}

unsigned long long* resetDt(sphx_workspace* ws)
{
    ws->counters.reserve(2 * kCounterSlots);
    unsigned long long* dt = reinterpret_cast<unsigned long long*>(ws->counters.get()) + 3;
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

template<class T>
void launchGradh(sphx_workspace* ws, const PairArgs& a, const sphx_fields_t* f, const sphx_params_t* p)
{
    unsigned blocks = pairBlocks(a.L);
    int      in     = takeChain(ws, kChainGradh, a, f, p);
    if (blocks == 0) { return; }
    RecGradh<T>* rec = recordBuffer<RecGradh<T>>(ws, a.numParticles, in < 0 ? 0 : in);
    if (in < 0)
    {
        packGradh<T><<<ceilDiv(a.numParticles, kPackThreads), kPackThreads, 0, ws->stream>>>(a, (const T*)f->m,
                                                                                              (const T*)f->xm, rec);
        SPHX_CHECK_LAUNCH();
    }
    // VeDefGradh leaves the IAD records (position, xm, xm/kx, v) when chaining
    const int    out  = in < 0 ? 1 : 1 - in;
    RecIadVe<T>* next = chainable(a, p) && f->vx && f->vy && f->vz
                            ? recordBuffer<RecIadVe<T>>(ws, a.numParticles, out)
                            : nullptr;
    {
        KernelScope timer(ws, "ve_def_gradh");
        launchWide<T>(ws->nbWide, [&](auto wide) {
            veDefGradhKernel<T, decltype(wide)::value><<<blocks, kPairWaves * kWave, isMixed<T> ? occupancyCap(SPHX_CAP_GRADH) : 0, ws->stream>>>(
                a, rec, (const T*)f->h, sincKernel<T>(ws), (T*)f->kx, (T*)f->gradh, next, (const T*)f->vx,
                (const T*)f->vy, (const T*)f->vz);
        });
        SPHX_CHECK_LAUNCH();
    }
    if (next) { leaveChain(ws, kChainIadVe, out, a, f, p); }
}

template<class T>
void launchIadVe(sphx_workspace* ws, const PairArgs& a, const IadVeFields& vf, const sphx_fields_t* f,
                 const sphx_params_t* p)
{
    unsigned blocks = pairBlocks(a.L);
    int      in     = takeChain(ws, kChainIadVe, a, f, p);
    if (blocks == 0) { return; }
    RecIadVe<T>* rec = recordBuffer<RecIadVe<T>>(ws, a.numParticles, in < 0 ? 0 : in);
    if (in < 0)
    {
        in = 0;
        packIadVe<T><<<ceilDiv(a.numParticles, kPackThreads), kPackThreads, 0, ws->stream>>>(a, vf, rec);
        SPHX_CHECK_LAUNCH();
    }
    RecAv<T>* next = chainable(a, p) && f->c ? recordBuffer<RecAv<T>>(ws, a.numParticles, 1 - in) : nullptr;
    {
        KernelScope timer(ws, "iad_divv_curlv");
        launchWide<T>(ws->nbWide, [&](auto wide) {
            iadDivvCurlvKernel<T, decltype(wide)::value><<<blocks, kPairWaves * kWave, isMixed<T> ? occupancyCap(SPHX_CAP_IAD) : 0, ws->stream>>>(
                a, vf, rec, sincKernel<T>(ws), next, (const T*)f->c);
        });
        SPHX_CHECK_LAUNCH();
    }
    if (next) { leaveChain(ws, kChainAv, 1 - in, a, f, p); }
}

template<class T>
void launchAv(sphx_workspace* ws, const PairArgs& a, const AvFields& af, const sphx_fields_t* f,
              const sphx_params_t* p)
{
    unsigned blocks = pairBlocks(a.L);
    int      in     = takeChain(ws, kChainAv, a, f, p);
    if (blocks == 0) { return; }
    RecAv<T>* rec = recordBuffer<RecAv<T>>(ws, a.numParticles, in < 0 ? 0 : in);
    if (in < 0)
    {
        in = 0;
        packAv<T><<<ceilDiv(a.numParticles, kPackThreads), kPackThreads, 0, ws->stream>>>(a, af, rec);
        SPHX_CHECK_LAUNCH();
    }
    RecMomVe<T>* next =
        chainable(a, p) && f->m && f->prho ? recordBuffer<RecMomVe<T>>(ws, a.numParticles, 1 - in) : nullptr;
    {
        KernelScope timer(ws, "av_switches");
        launchWide<T>(ws->nbWide, [&](auto wide) {
            avSwitchesKernel<T, decltype(wide)::value><<<blocks, kPairWaves * kWave, isMixed<T> ? occupancyCap(SPHX_CAP_AV) : 0, ws->stream>>>(
                a, af, rec, sincKernel<T>(ws), p->minDt, T(p->alphamin), T(p->alphamax), T(p->decay_constant), next);
        });
        SPHX_CHECK_LAUNCH();
    }
    if (next) { leaveChain(ws, kChainMomVe, 1 - in, a, f, p); }
}

template<class T>
void launchMomVe(sphx_workspace* ws, const PairArgs& a, const MomVeFields& mf, const sphx_fields_t* f,
                 const sphx_params_t* p, bool avClean, unsigned long long* dt)
{
    unsigned blocks = pairBlocks(a.L);
    int      in     = takeChain(ws, kChainMomVe, a, f, p);
    if (blocks == 0) { return; }
    RecMomVe<T>* rec = recordBuffer<RecMomVe<T>>(ws, a.numParticles, in < 0 ? 0 : in);
    if (in < 0)
    {
        packMomVe<T><<<ceilDiv(a.numParticles, kPackThreads), kPackThreads, 0, ws->stream>>>(a, mf, rec);
        SPHX_CHECK_LAUNCH();
    }
    KernelScope timer(ws, "momentum_ve");
    launchWide<T>(ws->nbWide, [&](auto wide) {
        constexpr bool W = decltype(wide)::value;
        if (avClean)
            momentumVeKernel<T, true, W><<<blocks, kPairWaves * kWave, isMixed<T> ? occupancyCap(SPHX_CAP_MOM) : 0, ws->stream>>>(
                a, mf, rec, sincKernel<T>(ws), T(p->Atmin), T(p->Atmax), T(p->ramp), float(p->Kcour), dt);
        else
            momentumVeKernel<T, false, W><<<blocks, kPairWaves * kWave, isMixed<T> ? occupancyCap(SPHX_CAP_MOM) : 0, ws->stream>>>(
                a, mf, rec, sincKernel<T>(ws), T(p->Atmin), T(p->Atmax), T(p->ramp), float(p->Kcour), dt);
    });
    SPHX_CHECK_LAUNCH();
}

} // namespace

void veDefGradh(sphx_workspace* ws, const DevBox& box, const sphx_fields_t* f, const sphx_params_t* p, uint32_t first,
                uint32_t last)
{
    PairArgs a = makePairArgs(ws, box, f, p, first, last);
    SPHX_REQUIRE(f->h && f->m && f->xm && f->kx && f->gradh, SPHX_ERR_ARG, "VeDefGradh needs h, m, xm, kx, gradh");
    if (p->hydro_type == SPHX_F32) launchGradh<float>(ws, a, f, p);
    else launchGradh<double>(ws, a, f, p);
}

template<class T>
void launchEosVeVariant(sphx_workspace* ws, const sphx_fields_t* f, const sphx_params_t* p, uint32_t first,
                        uint32_t last)
{
    const unsigned blocks = ceilDiv(last - first, 256);
    if (p->eos == SPHX_EOS_ISOTHERMAL)
        eosVeIsothermalKernel<T><<<blocks, 256, 0, ws->stream>>>(
            (const T*)f->m, (const T*)f->kx, (const T*)f->xm, (const T*)f->gradh, (T*)f->prho, (T*)f->c, (T*)f->rho,
            (T*)f->p, f->temp, T(p->soundSpeedConst), first, last);
    else
        eosVePolytropicKernel<T><<<blocks, 256, 0, ws->stream>>>(
            (const T*)f->m, (const T*)f->kx, (const T*)f->xm, (const T*)f->gradh, (T*)f->prho, (T*)f->c, (T*)f->rho,
            (T*)f->p, f->temp, p->polytropic_const, p->polytropic_index, first, last);
    SPHX_CHECK_LAUNCH();
}

void eosVe(sphx_workspace* ws, const sphx_fields_t* f, const sphx_params_t* p, uint32_t first, uint32_t last)
{
    SPHX_REQUIRE(p->eos >= SPHX_EOS_IDEAL_GAS && p->eos <= SPHX_EOS_POLYTROPIC, SPHX_ERR_ARG, "unknown EOS choice");
    SPHX_REQUIRE(first <= last, SPHX_ERR_ARG, "invalid range");
    if (p->eos != SPHX_EOS_IDEAL_GAS)
    {
        SPHX_REQUIRE(f->m && f->kx && f->xm && f->gradh && f->prho && f->c, SPHX_ERR_ARG,
                     "VE EOS needs m, kx, xm, gradh, prho and c");
        if (last == first) { return; }
        if (p->hydro_type == SPHX_F32) launchEosVeVariant<float>(ws, f, p, first, last);
        else launchEosVeVariant<double>(ws, f, p, first, last);
        return;
    }
    SPHX_REQUIRE(f->m && f->kx && f->xm && f->gradh && f->prho && f->c && (f->temp || f->u), SPHX_ERR_ARG,
                 "VE EOS needs m, kx, xm, gradh, prho, c and temp or u");
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
    PairArgs a = makePairArgs(ws, box, f, p, first, last);
    SPHX_REQUIRE(f->h && f->xm && f->kx && f->vx && f->vy && f->vz && f->c11 && f->c12 && f->c13 && f->c22 &&
                     f->c23 && f->c33 && f->divv,
                 SPHX_ERR_ARG, "IadDivvCurlv needs h, xm, kx, v, c11..c33 and divv");
    bool gradV = f->dV11 != nullptr;
    SPHX_REQUIRE(!gradV || (f->dV12 && f->dV13 && f->dV22 && f->dV23 && f->dV33), SPHX_ERR_ARG,
                 "dV11..dV33 must be all set or all NULL");
    IadVeFields vf{f->h,   f->xm,   f->kx,    f->vx,   f->vy,   f->vz,   f->c11,  f->c12,  f->c13,  f->c22,
                   f->c23, f->c33, f->divv, f->curlv, f->dV11, f->dV12, f->dV13, f->dV22, f->dV23, f->dV33};
    if (p->hydro_type == SPHX_F32) launchIadVe<float>(ws, a, vf, f, p);
    else launchIadVe<double>(ws, a, vf, f, p);
}

void avSwitches(sphx_workspace* ws, const DevBox& box, const sphx_fields_t* f, const sphx_params_t* p, uint32_t first,
                uint32_t last)
{
    PairArgs a = makePairArgs(ws, box, f, p, first, last);
    SPHX_REQUIRE(f->h && f->c && f->xm && f->kx && f->vx && f->vy && f->vz && f->divv && f->c11 && f->c12 && f->c13 &&
                     f->c22 && f->c23 && f->c33 && f->alpha,
                 SPHX_ERR_ARG, "AVswitches needs h, c, xm, kx, v, divv, c11..c33 and alpha");
    AvFields af{f->h,   f->c,   f->xm,  f->kx,  f->vx,  f->vy,  f->vz,    f->divv, f->c11,
                f->c12, f->c13, f->c22, f->c23, f->c33, f->alpha, f->m,  f->prho};
    if (p->hydro_type == SPHX_F32) launchAv<float>(ws, a, af, f, p);
    else launchAv<double>(ws, a, af, f, p);
}

double momentumEnergyVe(sphx_workspace* ws, const DevBox& box, const sphx_fields_t* f, const sphx_params_t* p,
                        uint32_t first, uint32_t last, bool avClean)
{
    PairArgs a = makePairArgs(ws, box, f, p, first, last);
    SPHX_REQUIRE(f->vx && f->vy && f->vz && f->h && f->m && f->prho && f->c && f->c11 && f->c12 && f->c13 && f->c22 &&
                     f->c23 && f->c33 && f->kx && f->xm && f->alpha && f->ax && f->ay && f->az && f->du,
                 SPHX_ERR_ARG, "VE momentum-energy needs v, h, m, prho, c, c11..c33, kx, xm, alpha, a and du");
    SPHX_REQUIRE(!avClean || (f->dV11 && f->dV12 && f->dV13 && f->dV22 && f->dV23 && f->dV33), SPHX_ERR_ARG,
                 "AV cleaning needs dV11..dV33");
    MomVeFields mf{f->vx,  f->vy,   f->vz,   f->h,    f->m,    f->prho, f->tdpdTrho, f->c,  f->c11, f->c12,
                   f->c13, f->c22,  f->c23,  f->c33,  f->kx,   f->xm,   f->alpha,    f->dV11, f->dV12, f->dV13,
                   f->dV22, f->dV23, f->dV33, f->ax,   f->ay,   f->az,   f->du};
    unsigned long long* dt = resetDt(ws);
    if (p->hydro_type == SPHX_F32) launchMomVe<float>(ws, a, mf, f, p, avClean, dt);
    else launchMomVe<double>(ws, a, mf, f, p, avClean, dt);
    return readDt(ws, dt);
}

} // namespace sphx
