// Adjoint (vector-Jacobian product) kernels of the std density, IAD and momentum-energy passes.
//
// No counterpart in the reference. Each kernel re-walks the neighbour lists of the last search
// (held fixed: they are the checkpoint of the forward), recomputes the pair terms of its forward
// loop and accumulates the reverse-mode derivatives. A pair term of target i and neighbour j has
// derivatives for both: the target's are summed in registers and added to its gradient once; the
// neighbour's are written to the pair's record in a per-pair buffer (no atomics), and a second
// kernel sums, for every particle j, the records of the pairs that have j as neighbour, in the
// order of their pair index (AdjointPairs: target i's pairs are p = pairStart[i] + k, and a sort of
// (j, p) keys gives each j its pairs in increasing p). Results are therefore the same bit for bit
// from run to run. Kernel derivatives use the tabulated dW/dv interpolated like W (FP64) or are
// evaluated like W (FP32). Launch shape and list layout as the forward loops (pairs.hpp).
// Derivations, with r = x_i - x_j after the fold, d = |r|, v = d / h_i:
//
//  density   rho_i = K h_i^-3 (m_i + sum_j m_j W(v))
//            d/dr = K h_i^-4 m_j W'(v) r/d,   d/dh_i = -K h_i^-4 (3 rho0_i + sum_j m_j v W'(v))
//  IAD       C_i = h_i^3/K tau_i^-1, tau_i = sum_j V_j W(v) r r^T, V_j = m_j/rho_j; with G the
//            matrix cotangent of C_i (a stored off-diagonal entry stands for two matrix entries):
//            tauBar = -(K/h_i^3) C_i G C_i, dL/dh_i += 3/h_i tr(G C_i); per pair q = r^T tauBar r:
//            d/dr = V_j (W'(v) q r/(h_i d) + 2 W(v) tauBar r), d/dh_i = -V_j W'(v) q d/h_i^2,
//            d/drho_j = -m_j/rho_j^2 W(v) q
//  momentum  reverse mode through the pair term of hydro_std/momentum_energy_kern.hpp:11-134 with
//            a per-particle AV coefficient (the reference's is the constant 1):
//            a_i = K sum_j (aa C_i r + bb C_j r), du_i = -K/2 sum_j v_ij.(ae C_i r + be C_j r)
#include <cstring> // rocprim uses memset without including it
#include <rocprim/rocprim.hpp>

#include "pairs.hpp"

namespace sphx
{

namespace
{

//! the target's own gradient: one lane per target, so a plain read-modify-write
template<class T>
__device__ __forceinline__ void addOwn(void* p, uint32_t i, T v)
{
    static_cast<T*>(p)[i] += v;
}

/*! values per pair record: NQ rounded up to whole 16-byte words (FP32) or 32-byte sectors (FP64); records
 * move as 16-byte vector loads and stores. Measured at 2M particles: whole sectors help the FP64 records
 * (momentum 17.7 -> 16.5 ms), while the FP32 ones run faster at their 16-byte size (density 4.9 against
 * 6.0 ms padded). */
template<class T, int NQ>
constexpr int kPaddedNq = int((NQ * sizeof(T) + (sizeof(T) == 8 ? 31 : 15)) / (sizeof(T) == 8 ? 32 : 16) *
                              (sizeof(T) == 8 ? 32 : 16) / sizeof(T));

/*! where the pair records of a call go: the record of pair p (NQP values of T, 16-byte aligned) at
 * buf + recordOf[p] * NQP, its place among the pairs sorted by neighbour, so that the records a
 * neighbour sums are one contiguous row */
template<class T>
struct PairOut
{
    T*              buf;
    const uint32_t* pairStart; // per searched target, from the searched range's first
    const uint32_t* recordOf;
    uint32_t        nbFirst;

    //! the first pair index of target i
    __device__ __forceinline__ uint32_t firstPair(uint32_t i) const { return pairStart[i - nbFirst]; }

    template<int NQP>
    __device__ __forceinline__ void store(uint32_t e, const T (&v)[NQP]) const
    {
        static_assert(NQP * sizeof(T) % 16 == 0, "whole 16-byte words");
        uint4* dst = reinterpret_cast<uint4*>(buf + size_t(e) * NQP);
#pragma unroll
        for (int c = 0; c < int(NQP * sizeof(T) / 16); ++c)
        {
            uint4 w;
            __builtin_memcpy(&w, reinterpret_cast<const char*>(v) + 16 * c, 16);
            dst[c] = w;
        }
    }
};

template<class T>
__device__ __forceinline__ T pow4Inv(T hInv)
{
    T h2 = hInv * hInv;
    return h2 * h2;
}

struct DensityVjpArgs
{
    const void *h, *m, *gRho;
    void *      gx, *gy, *gz, *gh;
};

constexpr int kDensityNq = 3; // neighbour-side record: gx, gy, gz

template<class T>
__global__ void __launch_bounds__(kPairWaves * kWave)
    densityStdVjpKernel(PairArgs a, DensityVjpArgs f, SincKernel<T> kern, PairOut<T> out)
{
    Lane t = laneTarget(a.L);
    if (!t.valid) { return; }
    const uint32_t i    = t.i;
    const T        hi   = ld<T>(f.h, i);
    const T        hInv = T(1) / hi;
    const double   r2h  = 2.0 * double(hi);
    const double   xi = a.x[i], yi = a.y[i], zi = a.z[i];
    const T        coef = ld<T>(f.gRho, i) * T(a.K) * pow4Inv(hInv);
    T              rho0 = ld<T>(f.m, i), sumVdW = 0, gxi = 0, gyi = 0, gzi = 0;
    const uint32_t ps = out.firstPair(i);
    withFold(waveNeedsFold(a.box, xi, yi, zi, r2h), [&](auto fold) {
        // U neighbours at a time: their list entries, record places and fields load before any is used
        constexpr int U = 4;
        for (unsigned k0 = 0; k0 < t.n; k0 += U)
        {
            uint32_t j[U], e[U];
#pragma unroll
            for (int u = 0; u < U; ++u)
            {
                const unsigned kk = min(k0 + u, t.n - 1);
                j[u]              = t.list[laneSlot(kk)];
                e[u]              = out.recordOf[ps + kk];
            }
            double xj[U], yj[U], zj[U];
            T      mj[U];
#pragma unroll
            for (int u = 0; u < U; ++u)
                xj[u] = a.x[j[u]], yj[u] = a.y[j[u]], zj[u] = a.z[j[u]], mj[u] = ld<T>(f.m, j[u]);
#pragma unroll
            for (int u = 0; u < U; ++u)
            {
                if (k0 + u >= t.n) { break; }
                T rx, ry, rz;
                separation<decltype(fold)::value>(a.box, r2h, xi, yi, zi, xj[u], yj[u], zj[u], rx, ry, rz);
                T d = sqrt(rx * rx + ry * ry + rz * rz);
                T v = d * hInv;
                T w, dw;
                values(kern.fetchWD(v), w, dw);
                rho0 += mj[u] * w;
                sumVdW += mj[u] * v * dw;
                T c = d > T(0) ? coef * mj[u] * dw / d : T(0);
                gxi += c * rx, gyi += c * ry, gzi += c * rz;
                T rec[kPaddedNq<T, kDensityNq>] = {};
                rec[0] = -c * rx, rec[1] = -c * ry, rec[2] = -c * rz;
                out.store(e[u], rec);
            }
        }
    });
    addOwn<T>(f.gx, i, gxi), addOwn<T>(f.gy, i, gyi), addOwn<T>(f.gz, i, gzi);
    addOwn<T>(f.gh, i, -coef * (T(3) * rho0 + sumVdW));
}

struct IadVjpArgs
{
    const void *h, *m, *rho, *c[6], *g[6];
    void *      gx, *gy, *gz, *gh, *grho;
};

constexpr int kIadNq = 4; // neighbour-side record: gx, gy, gz, grho

template<class T>
__global__ void __launch_bounds__(kPairWaves * kWave)
    iadStdVjpKernel(PairArgs a, IadVjpArgs f, SincKernel<T> kern, PairOut<T> out)
{
    Lane t = laneTarget(a.L);
    if (!t.valid) { return; }
    const uint32_t i    = t.i;
    const T        hi   = ld<T>(f.h, i);
    const T        hInv = T(1) / hi;
    const double   r2h  = 2.0 * double(hi);
    const double   xi = a.x[i], yi = a.y[i], zi = a.z[i];

    T C6[6], G6[6];
    for (int q = 0; q < 6; ++q)
        C6[q] = ld<T>(f.c[q], i), G6[q] = ld<T>(f.g[q], i);
    const T C[3][3] = {{C6[0], C6[1], C6[2]}, {C6[1], C6[3], C6[4]}, {C6[2], C6[4], C6[5]}};
    const T G[3][3] = {{G6[0], G6[1] / 2, G6[2] / 2}, {G6[1] / 2, G6[3], G6[4] / 2}, {G6[2] / 2, G6[4] / 2, G6[5]}};
    T       CG[3][3], B[3][3];
    for (int p = 0; p < 3; ++p)
        for (int q = 0; q < 3; ++q)
            CG[p][q] = C[p][0] * G[0][q] + C[p][1] * G[1][q] + C[p][2] * G[2][q];
    const T s = -T(a.K) * hInv * hInv * hInv;
    for (int p = 0; p < 3; ++p)
        for (int q = 0; q < 3; ++q)
            B[p][q] = s * (CG[p][0] * C[0][q] + CG[p][1] * C[1][q] + CG[p][2] * C[2][q]);
    T ghi = T(3) * hInv *
            (G6[0] * C6[0] + G6[1] * C6[1] + G6[2] * C6[2] + G6[3] * C6[3] + G6[4] * C6[4] + G6[5] * C6[5]);
    T gxi = 0, gyi = 0, gzi = 0;

    const uint32_t ps = out.firstPair(i);
    withFold(waveNeedsFold(a.box, xi, yi, zi, r2h), [&](auto fold) {
        // U neighbours at a time: their list entries, record places and fields load before any is used
        constexpr int U = 4;
        for (unsigned k0 = 0; k0 < t.n; k0 += U)
        {
            uint32_t j[U], e[U];
#pragma unroll
            for (int u = 0; u < U; ++u)
            {
                const unsigned kk = min(k0 + u, t.n - 1);
                j[u]              = t.list[laneSlot(kk)];
                e[u]              = out.recordOf[ps + kk];
            }
            double xj[U], yj[U], zj[U];
            T      mjs[U], rhojs[U];
#pragma unroll
            for (int u = 0; u < U; ++u)
            {
                xj[u] = a.x[j[u]], yj[u] = a.y[j[u]], zj[u] = a.z[j[u]];
                mjs[u] = ld<T>(f.m, j[u]), rhojs[u] = ld<T>(f.rho, j[u]);
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
            {
                if (k0 + u >= t.n) { break; }
                T r[3];
                separation<decltype(fold)::value>(a.box, r2h, xi, yi, zi, xj[u], yj[u], zj[u], r[0], r[1], r[2]);
                T d = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
                T v = d * hInv;
                T w, dw;
                values(kern.fetchWD(v), w, dw);
                T mj   = mjs[u];
                T rhoj = rhojs[u];
                T Vj   = mj / rhoj;
                T Br[3];
                for (int p = 0; p < 3; ++p)
                    Br[p] = B[p][0] * r[0] + B[p][1] * r[1] + B[p][2] * r[2];
                T q      = r[0] * Br[0] + r[1] * Br[1] + r[2] * Br[2];
                T radial = d > T(0) ? Vj * dw * q * hInv / d : T(0);
                T dr[3];
                for (int p = 0; p < 3; ++p)
                    dr[p] = radial * r[p] + T(2) * Vj * w * Br[p];
                gxi += dr[0], gyi += dr[1], gzi += dr[2];
                ghi += -Vj * dw * q * d * hInv * hInv;
                T rec[kPaddedNq<T, kIadNq>] = {};
                rec[0] = -dr[0], rec[1] = -dr[1], rec[2] = -dr[2], rec[3] = -mj / (rhoj * rhoj) * w * q;
                out.store(e[u], rec);
            }
        }
    });
    addOwn<T>(f.gx, i, gxi), addOwn<T>(f.gy, i, gyi), addOwn<T>(f.gz, i, gzi);
    addOwn<T>(f.gh, i, ghi);
}

struct MomVjpArgs
{
    const void *vx, *vy, *vz, *h, *m, *rho, *p, *c, *alpha, *C[6];
    const void *gax, *gay, *gaz, *gdu;
    void *      gx, *gy, *gz, *gh, *grho, *gp, *galpha, *gC[6];
};

template<class T>
__device__ __forceinline__ void loadC(const MomVjpArgs& f, uint32_t k, T M[3][3])
{
    T c11 = ld<T>(f.C[0], k), c12 = ld<T>(f.C[1], k), c13 = ld<T>(f.C[2], k);
    T c22 = ld<T>(f.C[3], k), c23 = ld<T>(f.C[4], k), c33 = ld<T>(f.C[5], k);
    M[0][0] = c11, M[0][1] = c12, M[0][2] = c13;
    M[1][0] = c12, M[1][1] = c22, M[1][2] = c23;
    M[2][0] = c13, M[2][1] = c23, M[2][2] = c33;
}

//! stored-entry cotangent of a symmetric matrix used as M r, from the matrix cotangent u r^T
template<class T>
__device__ __forceinline__ void symOuter(const T u[3], const T r[3], T out[6])
{
    out[0] = u[0] * r[0];
    out[1] = u[0] * r[1] + u[1] * r[0];
    out[2] = u[0] * r[2] + u[2] * r[0];
    out[3] = u[1] * r[1];
    out[4] = u[1] * r[2] + u[2] * r[1];
    out[5] = u[2] * r[2];
}

// neighbour-side record: gx, gy, gz, gh, grho, gp, galpha, gC[6]
constexpr int kMomNq = 13;

template<class T>
__global__ void __launch_bounds__(kPairWaves * kWave)
    momentumStdVjpKernel(PairArgs a, MomVjpArgs f, SincKernel<T> kern, PairOut<T> out)
{
    Lane t = laneTarget(a.L);
    if (!t.valid) { return; }
    const uint32_t i      = t.i;
    const T        hi     = ld<T>(f.h, i);
    const T        hiInv  = T(1) / hi;
    const T        hiInv3 = hiInv * hiInv * hiInv;
    const double   r2h    = 2.0 * double(hi);
    const double   xi = a.x[i], yi = a.y[i], zi = a.z[i];
    const T        vxi = ld<T>(f.vx, i), vyi = ld<T>(f.vy, i), vzi = ld<T>(f.vz, i);
    const T        roi = ld<T>(f.rho, i), pri = ld<T>(f.p, i), ci = ld<T>(f.c, i), mi = ld<T>(f.m, i);
    const T        vol_i = mi / roi, pro_i = pri / (roi * roi);
    const T        alpha_i = f.alpha ? ld<T>(f.alpha, i) : T(1);
    T              Ci[3][3];
    loadC<T>(f, i, Ci);
    const T Ga[3] = {T(a.K) * ld<T>(f.gax, i), T(a.K) * ld<T>(f.gay, i), T(a.K) * ld<T>(f.gaz, i)};
    const T ge    = f.gdu ? -T(a.K) * T(0.5) * ld<T>(f.gdu, i) : T(0);

    T gx_i = 0, gy_i = 0, gz_i = 0, gh_i = 0, grho_i = 0, gp_i = 0, galpha_i = 0;
    T gC_i[6] = {0, 0, 0, 0, 0, 0};

    const uint32_t ps = out.firstPair(i);
    withFold(waveNeedsFold(a.box, xi, yi, zi, r2h), [&](auto fold) {
        for (unsigned k = 0; k < t.n; ++k)
        {
            const uint32_t j = t.list[laneSlot(k)];
            const uint32_t e = out.recordOf[ps + k];
            T              r[3];
            separation<decltype(fold)::value>(a.box, r2h, xi, yi, zi, a.x[j], a.y[j], a.z[j], r[0], r[1], r[2]);
            T d      = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
            T vij[3] = {vxi - ld<T>(f.vx, j), vyi - ld<T>(f.vy, j), vzi - ld<T>(f.vz, j)};
            T hjInv  = T(1) / ld<T>(f.h, j);
            T hjInv3 = hjInv * hjInv * hjInv;
            T v1 = d * hiInv, v2 = d * hjInv;
            T W1, dW1, W2, dW2;
            values(kern.fetchWD(v1), W1, dW1);
            values(kern.fetchWD(v2), W2, dW2);
            T Wi = hiInv3 * W1, Wj = hjInv3 * W2;
            T Cj[3][3];
            loadC<T>(f, j, Cj);
            T tAi[3], tAj[3];
            for (int p = 0; p < 3; ++p)
            {
                tAi[p] = Ci[p][0] * r[0] + Ci[p][1] * r[1] + Ci[p][2] * r[2];
                tAj[p] = Cj[p][0] * r[0] + Cj[p][1] * r[1] + Cj[p][2] * r[2];
            }
            T rv      = r[0] * vij[0] + r[1] * vij[1] + r[2] * vij[2];
            T w       = d > T(0) ? rv / d : T(0);
            T alpha_j = f.alpha ? ld<T>(f.alpha, j) : T(1);
            T cj      = ld<T>(f.c, j);
            T visc = 0, dvisc_dw = 0, dvisc_dalpha = 0;
            if (w < T(0))
            {
                T a4         = (alpha_i + alpha_j) / T(4) * (ci + cj);
                visc         = T(0.5) * (-(a4 - T(2) * w) * w);
                dvisc_dw     = T(0.5) * (-a4 + T(4) * w);
                dvisc_dalpha = T(0.5) * (-(ci + cj) / T(4) * w);
            }
            T mj = ld<T>(f.m, j), roj = ld<T>(f.rho, j), pj = ld<T>(f.p, j);
            T vol_j = mj / roj, prj = pj / roj;

            T aa = Wi * (mj * pro_i + visc * vol_i);
            T bb = vol_j * Wj * (prj + visc);
            T ae = Wi * (T(2) * mj * pro_i + visc * vol_i);
            T be = visc * vol_j * Wj;

            T aaBar = Ga[0] * tAi[0] + Ga[1] * tAi[1] + Ga[2] * tAi[2];
            T bbBar = Ga[0] * tAj[0] + Ga[1] * tAj[1] + Ga[2] * tAj[2];
            T aeBar = ge * (vij[0] * tAi[0] + vij[1] * tAi[1] + vij[2] * tAi[2]);
            T beBar = ge * (vij[0] * tAj[0] + vij[1] * tAj[1] + vij[2] * tAj[2]);

            T tAiBar[3], tAjBar[3], rBar[3];
            for (int p = 0; p < 3; ++p)
            {
                tAiBar[p] = aa * Ga[p] + ge * ae * vij[p];
                tAjBar[p] = bb * Ga[p] + ge * be * vij[p];
            }
            for (int p = 0; p < 3; ++p)
                rBar[p] = Ci[0][p] * tAiBar[0] + Ci[1][p] * tAiBar[1] + Ci[2][p] * tAiBar[2] + Cj[0][p] * tAjBar[0] +
                          Cj[1][p] * tAjBar[1] + Cj[2][p] * tAjBar[2];
            T rec[kPaddedNq<T, kMomNq>] = {};
            T ci6[6], cj6[6];
            symOuter(tAiBar, r, ci6);
            symOuter(tAjBar, r, cj6);
            for (int q = 0; q < 6; ++q)
            {
                gC_i[q] += ci6[q];
                rec[7 + q] = cj6[q];
            }

            T WiBar    = aaBar * (mj * pro_i + visc * vol_i) + aeBar * (T(2) * mj * pro_i + visc * vol_i);
            T WjBar    = bbBar * vol_j * (prj + visc) + beBar * visc * vol_j;
            T proBar_i = (aaBar + T(2) * aeBar) * Wi * mj;
            T viscBar  = (aaBar + aeBar) * Wi * vol_i + (bbBar + beBar) * vol_j * Wj;
            T volBar_i = (aaBar + aeBar) * Wi * visc;
            T volBar_j = bbBar * Wj * (prj + visc) + beBar * visc * Wj;
            T prjBar   = bbBar * vol_j * Wj;

            T wBar = viscBar * dvisc_dw;
            galpha_i += viscBar * dvisc_dalpha;
            rec[6] = viscBar * dvisc_dalpha;
            if (d > T(0))
            {
                T dBar = WiBar * hiInv3 * hiInv * dW1 + WjBar * hjInv3 * hjInv * dW2;
                for (int p = 0; p < 3; ++p)
                    rBar[p] += dBar * r[p] / d + wBar * (vij[p] / d - w * r[p] / (d * d));
            }
            gh_i += -WiBar * hiInv3 * hiInv * (T(3) * W1 + v1 * dW1);
            rec[3] = -WjBar * hjInv3 * hjInv * (T(3) * W2 + v2 * dW2);

            gp_i += proBar_i / (roi * roi);
            grho_i += -T(2) * proBar_i * pri / (roi * roi * roi) - volBar_i * mi / (roi * roi);
            rec[5] = prjBar / roj;
            rec[4] = -volBar_j * mj / (roj * roj) - prjBar * pj / (roj * roj);

            gx_i += rBar[0], gy_i += rBar[1], gz_i += rBar[2];
            rec[0] = -rBar[0], rec[1] = -rBar[1], rec[2] = -rBar[2];
            out.store(e, rec);
        }
    });
    addOwn<T>(f.gx, i, gx_i), addOwn<T>(f.gy, i, gy_i), addOwn<T>(f.gz, i, gz_i), addOwn<T>(f.gh, i, gh_i);
    addOwn<T>(f.grho, i, grho_i), addOwn<T>(f.gp, i, gp_i), addOwn<T>(f.galpha, i, galpha_i);
    for (int q = 0; q < 6; ++q)
        addOwn<T>(f.gC[q], i, gC_i[q]);
}

/* ---- the pair index of the adjoints --------------------------------------------------------------- */

//! stored pairs of each searched target (the last entry 0, so the exclusive scan ends with the total)
__global__ void pairCountKernel(const uint32_t* __restrict__ nc, uint32_t nbFirst, uint32_t nt, uint32_t ngmax,
                                uint32_t* __restrict__ cnt)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nt) { cnt[t] = min(nc[nbFirst + t] - 1u, ngmax); }
    else if (t == nt) { cnt[t] = 0; }
}

//! key (j << pbits) | p of every stored pair p = pairStart[t] + k, j the neighbour in slot k of target t
__global__ void pairKeysKernel(const uint32_t* __restrict__ nbrs, const uint32_t* __restrict__ pairStart, uint32_t nt,
                               uint32_t ngmax4, int pbits, uint64_t* __restrict__ keys)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) { return; }
    const uint32_t p0 = pairStart[t], n = pairStart[t + 1] - p0;
    for (uint32_t k = 0; k < n; ++k)
    {
        const uint32_t j = nbrs[listOffset(t / kGroup, ngmax4, k, t % kGroup)];
        keys[p0 + k]     = (uint64_t(j) << pbits) | uint64_t(p0 + k);
    }
}

//! rowStart[j] = first sorted key of neighbour j (lower bound of j << pbits), j = 0 .. n
__global__ void rowStartKernel(const uint64_t* __restrict__ sorted, uint64_t pairs, int pbits, uint32_t n,
                               uint32_t* __restrict__ rowStart)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j > n) { return; }
    const uint64_t key = uint64_t(j) << pbits;
    uint64_t       lo = 0, hi = pairs;
    while (lo < hi)
    {
        const uint64_t mid = (lo + hi) / 2;
        if (sorted[mid] < key) { lo = mid + 1; }
        else { hi = mid; }
    }
    rowStart[j] = uint32_t(lo);
}

//! recordOf[p] = e for the e-th sorted key, pair p
__global__ void recordOfKernel(const uint64_t* __restrict__ sorted, uint64_t pairs, uint64_t pmask,
                               uint32_t* __restrict__ recordOf)
{
    const uint64_t e = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e < pairs) { recordOf[sorted[e] & pmask] = uint32_t(e); }
}

template<class T, int NQ>
struct PullOut
{
    T* g[NQ];
};

/*! g_q[j] += sum of field q of the records of the pairs that have j as neighbour (the contiguous row
 * rowStart[j] .. rowStart[j + 1] of the records), in increasing pair index, counting only the pairs of
 * targets [tFirst, tLast) of the call (pair indices [pairStart[tFirst], pairStart[tLast]); with allPairs
 * every pair belongs to the call and the keys are not read) */
template<class T, int NQ>
__global__ void __launch_bounds__(256) pullKernel(const uint64_t* __restrict__ sorted, const uint32_t* __restrict__ rowStart,
                                                  const uint32_t* __restrict__ pairStart, uint32_t tFirst, uint32_t tLast,
                                                  uint64_t pmask, bool allPairs, const T* __restrict__ buf,
                                                  PullOut<T, NQ> out, uint32_t n)
{
    constexpr int NQP = kPaddedNq<T, NQ>;
    constexpr int NV  = int(NQP * sizeof(T) / 16); // 16-byte words per record
    constexpr int U   = 4;                         // records in flight per thread
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) { return; }
    const uint64_t pLo = pairStart[tFirst], pHi = pairStart[tLast];
    T              acc[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q)
        acc[q] = T(0);
    const uint32_t e0 = rowStart[j], e1 = rowStart[j + 1];
    for (uint32_t e = e0; e < e1; e += U)
    {
        bool use[U];
#pragma unroll
        for (int u = 0; u < U; ++u)
        {
            use[u] = e + u < e1;
            if (!allPairs && use[u])
            {
                const uint64_t p = sorted[e + u] & pmask;
                use[u]           = p >= pLo && p < pHi;
            }
        }
        uint4 w[U][NV];
#pragma unroll
        for (int u = 0; u < U; ++u)
        {
            const uint4* rec = reinterpret_cast<const uint4*>(buf + size_t(use[u] ? e + u : e0) * NQP);
#pragma unroll
            for (int c = 0; c < NV; ++c)
                w[u][c] = rec[c];
        }
        // summed in pair order; a slot past the row or outside the call's targets adds nothing
#pragma unroll
        for (int u = 0; u < U; ++u)
        {
            T v[NQP];
            __builtin_memcpy(v, w[u], sizeof(v));
#pragma unroll
            for (int q = 0; q < NQ; ++q)
                acc[q] = use[u] ? acc[q] + v[q] : acc[q];
        }
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q)
        out.g[q][j] += acc[q];
}

int bitsFor(uint64_t v)
{
    int b = 1;
    while (b < 64 && (v >> b) != 0) { ++b; }
    return b;
}

//! the pair index of the last search (built once per search, on the first adjoint call after it)
void buildAdjointPairs(sphx_workspace* ws)
{
    auto& A = ws->adj;
    if (A.serial == ws->searchSerial) { return; }
    const uint32_t nt = ws->nbLast - ws->nbFirst;
    const uint32_t n  = uint32_t(ws->tree.numParticles);
    SPHX_REQUIRE(uint64_t(nt) * ws->nbNgmax < (uint64_t(1) << 32), SPHX_ERR_ARG,
                 "adjoints: more than 2^32 neighbour pairs");
    A.pairStart.reserve(size_t(nt) + 1);
    // the counts, then the scan's temporaries, in the workspace scratch
    size_t tmpBytes = 0;
    SPHX_HIP(rocprim::exclusive_scan(nullptr, tmpBytes, static_cast<const uint32_t*>(nullptr), A.pairStart.get(), 0u,
                                     size_t(nt) + 1, rocprim::plus<uint32_t>(), ws->stream));
    const size_t cntBytes = ((size_t(nt) + 1) * sizeof(uint32_t) + 255) & ~size_t(255);
    char*        base     = static_cast<char*>(scratchBytes(ws, cntBytes + tmpBytes));
    uint32_t*    cnt      = reinterpret_cast<uint32_t*>(base);
    pairCountKernel<<<ceilDiv(size_t(nt) + 1, 256), 256, 0, ws->stream>>>(ws->nbNc, ws->nbFirst, nt, ws->nbNgmax, cnt);
    SPHX_CHECK_LAUNCH();
    SPHX_HIP(rocprim::exclusive_scan(base + cntBytes, tmpBytes, cnt, A.pairStart.get(), 0u, size_t(nt) + 1,
                                     rocprim::plus<uint32_t>(), ws->stream));
    uint32_t pairs = 0;
    SPHX_HIP(hipMemcpyAsync(&pairs, A.pairStart.get() + nt, sizeof(pairs), hipMemcpyDeviceToHost, ws->stream));
    SPHX_HIP(hipStreamSynchronize(ws->stream));
    const int pbits = bitsFor(pairs), jbits = bitsFor(n);
    A.keys.reserve(std::max<uint64_t>(pairs, 1));
    A.sorted.reserve(std::max<uint64_t>(pairs, 1));
    A.rowStart.reserve(size_t(n) + 1);
    if (pairs > 0)
    {
        pairKeysKernel<<<ceilDiv(nt, 256), 256, 0, ws->stream>>>(ws->nbrs.get(), A.pairStart.get(), nt,
                                                                  paddedNgmax(ws->nbNgmax), pbits, A.keys.get());
        SPHX_CHECK_LAUNCH();
        size_t sortBytes = 0;
        SPHX_HIP(rocprim::radix_sort_keys(nullptr, sortBytes, A.keys.get(), A.sorted.get(), size_t(pairs), 0,
                                          pbits + jbits, ws->stream));
        void* tmp = scratchBytes(ws, sortBytes);
        SPHX_HIP(rocprim::radix_sort_keys(tmp, sortBytes, A.keys.get(), A.sorted.get(), size_t(pairs), 0,
                                          pbits + jbits, ws->stream));
    }
    rowStartKernel<<<ceilDiv(size_t(n) + 1, 256), 256, 0, ws->stream>>>(A.sorted.get(), pairs, pbits, n,
                                                                        A.rowStart.get());
    SPHX_CHECK_LAUNCH();
    A.recordOf.reserve(std::max<uint64_t>(pairs, 1));
    if (pairs > 0)
    {
        recordOfKernel<<<ceilDiv(pairs, 256), 256, 0, ws->stream>>>(A.sorted.get(), pairs,
                                                                    (uint64_t(1) << pbits) - 1, A.recordOf.get());
        SPHX_CHECK_LAUNCH();
    }
    A.pairs  = pairs;
    A.pbits  = pbits;
    A.serial = ws->searchSerial;
}

//! the pair records of a call with nq values of T per pair
template<class T>
PairOut<T> pairOut(sphx_workspace* ws, int nq)
{
    auto& A = ws->adj;
    A.pairBuf.reserve(std::max<uint64_t>(A.pairs, 1) * nq * sizeof(T) + 16);
    return PairOut<T>{reinterpret_cast<T*>(A.pairBuf.get()), A.pairStart.get(), A.recordOf.get(), ws->nbFirst};
}

//! add every particle's neighbour-side sums of the call over targets [first, last) to g[0..NQ)
template<class T, int NQ>
void pull(sphx_workspace* ws, const PairOut<T>& po, uint32_t first, uint32_t last, void* const (&g)[NQ])
{
    const auto&    A = ws->adj;
    const uint32_t n = uint32_t(ws->tree.numParticles);
    PullOut<T, NQ> out;
    for (int q = 0; q < NQ; ++q)
        out.g[q] = static_cast<T*>(g[q]);
    const uint64_t pmask = (A.pbits >= 64) ? ~uint64_t(0) : ((uint64_t(1) << A.pbits) - 1);
    if (n == 0) { return; }
    const bool allPairs = first == ws->nbFirst && last == ws->nbLast;
    pullKernel<T, NQ><<<ceilDiv(n, 256), 256, 0, ws->stream>>>(A.sorted.get(), A.rowStart.get(), A.pairStart.get(),
                                                                first - ws->nbFirst, last - ws->nbFirst, pmask,
                                                                allPairs, po.buf, out, n);
    SPHX_CHECK_LAUNCH();
}

//! the adjoints take consecutive groups of 64 (no group view), the layout of their pair index
void requireConsecutive(const PairArgs& a)
{
    SPHX_REQUIRE(a.L.groups.start == nullptr, SPHX_ERR_ARG,
                 "the adjoints run over consecutive groups; clear the group view (sphx_set_groups with NULL) first");
}

void requireAll(std::initializer_list<const void*> ptrs, const char* what)
{
    for (const void* p : ptrs)
        SPHX_REQUIRE(p != nullptr, SPHX_ERR_ARG, what);
}

} // namespace

void densityStdVjp(sphx_workspace* ws, const DevBox& box, const sphx_fields_t* f, const sphx_params_t* p,
                   uint32_t first, uint32_t last, const sphx_adjoint_t* in, sphx_adjoint_t* out)
{
    PairArgs a = makePairArgs(ws, box, f, p, first, last);
    requireAll({f->h, f->m, in->rho, out->x, out->y, out->z, out->h},
               "density adjoint needs fields h, m, cotangent rho and gradients x, y, z, h");
    requireConsecutive(a);
    DensityVjpArgs d{f->h, f->m, in->rho, out->x, out->y, out->z, out->h};
    unsigned       blocks = pairBlocks(a.L);
    if (blocks == 0) { return; }
    buildAdjointPairs(ws);
    void* const g[kDensityNq] = {out->x, out->y, out->z};
    auto        run           = [&](auto zero) {
        using T      = decltype(zero);
        PairOut<T> o = pairOut<T>(ws, kPaddedNq<T, kDensityNq>);
        densityStdVjpKernel<T><<<blocks, kPairWaves * kWave, 0, ws->stream>>>(a, d, sincKernel<T>(ws), o);
        SPHX_CHECK_LAUNCH();
        pull<T, kDensityNq>(ws, o, first, last, g);
    };
    if (p->hydro_type == SPHX_F32) { run(0.0f); }
    else { run(0.0); }
}

void iadStdVjp(sphx_workspace* ws, const DevBox& box, const sphx_fields_t* f, const sphx_params_t* p, uint32_t first,
               uint32_t last, const sphx_adjoint_t* in, sphx_adjoint_t* out)
{
    PairArgs a = makePairArgs(ws, box, f, p, first, last);
    requireAll({f->h, f->m, f->rho, f->c11, f->c12, f->c13, f->c22, f->c23, f->c33},
               "IAD adjoint needs fields h, m, rho and c11..c33");
    requireAll({in->c11, in->c12, in->c13, in->c22, in->c23, in->c33, out->x, out->y, out->z, out->h, out->rho},
               "IAD adjoint needs cotangents c11..c33 and gradients x, y, z, h, rho");
    IadVjpArgs d{f->h,
                 f->m,
                 f->rho,
                 {f->c11, f->c12, f->c13, f->c22, f->c23, f->c33},
                 {in->c11, in->c12, in->c13, in->c22, in->c23, in->c33},
                 out->x,
                 out->y,
                 out->z,
                 out->h,
                 out->rho};
    requireConsecutive(a);
    unsigned blocks = pairBlocks(a.L);
    if (blocks == 0) { return; }
    buildAdjointPairs(ws);
    void* const g[kIadNq] = {out->x, out->y, out->z, out->rho};
    auto        run       = [&](auto zero) {
        using T      = decltype(zero);
        PairOut<T> o = pairOut<T>(ws, kPaddedNq<T, kIadNq>);
        iadStdVjpKernel<T><<<blocks, kPairWaves * kWave, 0, ws->stream>>>(a, d, sincKernel<T>(ws), o);
        SPHX_CHECK_LAUNCH();
        pull<T, kIadNq>(ws, o, first, last, g);
    };
    if (p->hydro_type == SPHX_F32) { run(0.0f); }
    else { run(0.0); }
}

void momentumEnergyStdVjp(sphx_workspace* ws, const DevBox& box, const sphx_fields_t* f, const sphx_params_t* p,
                          uint32_t first, uint32_t last, const sphx_adjoint_t* in, sphx_adjoint_t* out)
{
    PairArgs a = makePairArgs(ws, box, f, p, first, last);
    requireAll({f->vx, f->vy, f->vz, f->h, f->m, f->rho, f->p, f->c, f->c11, f->c12, f->c13, f->c22, f->c23, f->c33},
               "momentum-energy adjoint needs the forward's fields");
    const bool perAlpha = (p->flags & SPHX_FLAG_STD_ALPHA) != 0;
    SPHX_REQUIRE(!perAlpha || f->alpha, SPHX_ERR_ARG, "SPHX_FLAG_STD_ALPHA set but fields.alpha is NULL");
    requireAll({in->ax, in->ay, in->az, out->x, out->y, out->z, out->h, out->rho, out->p, out->alpha, out->c11,
                out->c12, out->c13, out->c22, out->c23, out->c33},
               "momentum-energy adjoint needs cotangents ax, ay, az and gradients x, y, z, h, rho, p, alpha, "
               "c11..c33");
    MomVjpArgs d{f->vx,
                 f->vy,
                 f->vz,
                 f->h,
                 f->m,
                 f->rho,
                 f->p,
                 f->c,
                 perAlpha ? f->alpha : nullptr,
                 {f->c11, f->c12, f->c13, f->c22, f->c23, f->c33},
                 in->ax,
                 in->ay,
                 in->az,
                 in->du,
                 out->x,
                 out->y,
                 out->z,
                 out->h,
                 out->rho,
                 out->p,
                 out->alpha,
                 {out->c11, out->c12, out->c13, out->c22, out->c23, out->c33}};
    requireConsecutive(a);
    unsigned blocks = pairBlocks(a.L);
    if (blocks == 0) { return; }
    buildAdjointPairs(ws);
    void* const g[kMomNq] = {out->x,   out->y,   out->z,   out->h,   out->rho, out->p,  out->alpha,
                             out->c11, out->c12, out->c13, out->c22, out->c23, out->c33};
    auto        run       = [&](auto zero) {
        using T      = decltype(zero);
        PairOut<T> o = pairOut<T>(ws, kPaddedNq<T, kMomNq>);
        momentumStdVjpKernel<T><<<blocks, kPairWaves * kWave, 0, ws->stream>>>(a, d, sincKernel<T>(ws), o);
        SPHX_CHECK_LAUNCH();
        pull<T, kMomNq>(ws, o, first, last, g);
    };
    if (p->hydro_type == SPHX_F32) { run(0.0f); }
    else { run(0.0); }
}

} // namespace sphx
