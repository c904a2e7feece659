// Adjoint (vector-Jacobian product) kernels of the std density, IAD and momentum-energy passes.
//
// No counterpart in the reference. Each kernel re-walks the neighbour lists of the last search
// (held fixed: they are the checkpoint of the forward), recomputes the pair terms of its forward
// loop and accumulates the reverse-mode derivatives: the target's own in registers, the
// neighbour's with an atomic add (a pair gradient scatters to j, and neighbour relations are not
// symmetric). Kernel derivatives use the tabulated dW/dv interpolated like W. Launch shape and
// list layout as the forward loops (pairs.hpp). Derivations, with r = x_i - x_j after the fold,
// d = |r|, v = d / h_i:
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
#include "pairs.hpp"

namespace sphx
{

namespace
{

template<class T>
__device__ __forceinline__ void scatter(void* p, uint32_t j, T v)
{
    atomicAdd(static_cast<T*>(p) + j, v);
}

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

template<class T>
__global__ void __launch_bounds__(kPairWaves * kWave)
    densityStdVjpKernel(PairArgs a, DensityVjpArgs f, SincKernel<T> kern)
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
    withFold(waveNeedsFold(a.box, xi, yi, zi, r2h), [&](auto fold) {
        for (unsigned k = 0; k < t.n; ++k)
        {
            const uint32_t j = t.list[laneSlot(k)];
            T              rx, ry, rz;
            separation<decltype(fold)::value>(a.box, r2h, xi, yi, zi, a.x[j], a.y[j], a.z[j], rx, ry, rz);
            T d  = sqrt(rx * rx + ry * ry + rz * rz);
            T v  = d * hInv;
            T w  = kern.W(v);
            T dw = kern.dW(v);
            T mj = ld<T>(f.m, j);
            rho0 += mj * w;
            sumVdW += mj * v * dw;
            if (d > T(0))
            {
                T c = coef * mj * dw / d;
                gxi += c * rx, gyi += c * ry, gzi += c * rz;
                scatter<T>(f.gx, j, -c * rx), scatter<T>(f.gy, j, -c * ry), scatter<T>(f.gz, j, -c * rz);
            }
        }
    });
    scatter<T>(f.gx, i, gxi), scatter<T>(f.gy, i, gyi), scatter<T>(f.gz, i, gzi);
    scatter<T>(f.gh, i, -coef * (T(3) * rho0 + sumVdW));
}

struct IadVjpArgs
{
    const void *h, *m, *rho, *c[6], *g[6];
    void *      gx, *gy, *gz, *gh, *grho;
};

template<class T>
__global__ void __launch_bounds__(kPairWaves * kWave)
    iadStdVjpKernel(PairArgs a, IadVjpArgs f, SincKernel<T> kern)
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

    withFold(waveNeedsFold(a.box, xi, yi, zi, r2h), [&](auto fold) {
        for (unsigned k = 0; k < t.n; ++k)
        {
            const uint32_t j = t.list[laneSlot(k)];
            T              r[3];
            separation<decltype(fold)::value>(a.box, r2h, xi, yi, zi, a.x[j], a.y[j], a.z[j], r[0], r[1], r[2]);
            T d    = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
            T v    = d * hInv;
            T w    = kern.W(v);
            T dw   = kern.dW(v);
            T mj   = ld<T>(f.m, j);
            T rhoj = ld<T>(f.rho, j);
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
            scatter<T>(f.gx, j, -dr[0]), scatter<T>(f.gy, j, -dr[1]), scatter<T>(f.gz, j, -dr[2]);
            ghi += -Vj * dw * q * d * hInv * hInv;
            scatter<T>(f.grho, j, -mj / (rhoj * rhoj) * w * q);
        }
    });
    scatter<T>(f.gx, i, gxi), scatter<T>(f.gy, i, gyi), scatter<T>(f.gz, i, gzi);
    scatter<T>(f.gh, i, ghi);
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

template<class T>
__global__ void __launch_bounds__(kPairWaves * kWave)
    momentumStdVjpKernel(PairArgs a, MomVjpArgs f, SincKernel<T> kern)
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

    withFold(waveNeedsFold(a.box, xi, yi, zi, r2h), [&](auto fold) {
        for (unsigned k = 0; k < t.n; ++k)
        {
            const uint32_t j = t.list[laneSlot(k)];
            T              r[3];
            separation<decltype(fold)::value>(a.box, r2h, xi, yi, zi, a.x[j], a.y[j], a.z[j], r[0], r[1], r[2]);
            T d      = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
            T vij[3] = {vxi - ld<T>(f.vx, j), vyi - ld<T>(f.vy, j), vzi - ld<T>(f.vz, j)};
            T hjInv  = T(1) / ld<T>(f.h, j);
            T hjInv3 = hjInv * hjInv * hjInv;
            T v1 = d * hiInv, v2 = d * hjInv;
            T W1 = kern.W(v1), dW1 = kern.dW(v1), W2 = kern.W(v2), dW2 = kern.dW(v2);
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
            T ci6[6], cj6[6];
            symOuter(tAiBar, r, ci6);
            symOuter(tAjBar, r, cj6);
            for (int q = 0; q < 6; ++q)
            {
                gC_i[q] += ci6[q];
                scatter<T>(f.gC[q], j, cj6[q]);
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
            scatter<T>(f.galpha, j, viscBar * dvisc_dalpha);
            if (d > T(0))
            {
                T dBar = WiBar * hiInv3 * hiInv * dW1 + WjBar * hjInv3 * hjInv * dW2;
                for (int p = 0; p < 3; ++p)
                    rBar[p] += dBar * r[p] / d + wBar * (vij[p] / d - w * r[p] / (d * d));
            }
            gh_i += -WiBar * hiInv3 * hiInv * (T(3) * W1 + v1 * dW1);
            scatter<T>(f.gh, j, -WjBar * hjInv3 * hjInv * (T(3) * W2 + v2 * dW2));

            gp_i += proBar_i / (roi * roi);
            grho_i += -T(2) * proBar_i * pri / (roi * roi * roi) - volBar_i * mi / (roi * roi);
            scatter<T>(f.gp, j, prjBar / roj);
            scatter<T>(f.grho, j, -volBar_j * mj / (roj * roj) - prjBar * pj / (roj * roj));

            gx_i += rBar[0], gy_i += rBar[1], gz_i += rBar[2];
            scatter<T>(f.gx, j, -rBar[0]), scatter<T>(f.gy, j, -rBar[1]), scatter<T>(f.gz, j, -rBar[2]);
        }
    });
    scatter<T>(f.gx, i, gx_i), scatter<T>(f.gy, i, gy_i), scatter<T>(f.gz, i, gz_i), scatter<T>(f.gh, i, gh_i);
    scatter<T>(f.grho, i, grho_i), scatter<T>(f.gp, i, gp_i), scatter<T>(f.galpha, i, galpha_i);
    for (int q = 0; q < 6; ++q)
        scatter<T>(f.gC[q], i, gC_i[q]);
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
    DensityVjpArgs d{f->h, f->m, in->rho, out->x, out->y, out->z, out->h};
    unsigned       blocks = pairBlocks(a.L);
    if (blocks == 0) { return; }
    if (p->hydro_type == SPHX_F32)
        densityStdVjpKernel<float><<<blocks, kPairWaves * kWave, 0, ws->stream>>>(a, d, sincKernel<float>(ws));
    else
        densityStdVjpKernel<double><<<blocks, kPairWaves * kWave, 0, ws->stream>>>(a, d, sincKernel<double>(ws));
    SPHX_CHECK_LAUNCH();
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
    unsigned blocks = pairBlocks(a.L);
    if (blocks == 0) { return; }
    if (p->hydro_type == SPHX_F32)
        iadStdVjpKernel<float><<<blocks, kPairWaves * kWave, 0, ws->stream>>>(a, d, sincKernel<float>(ws));
    else
        iadStdVjpKernel<double><<<blocks, kPairWaves * kWave, 0, ws->stream>>>(a, d, sincKernel<double>(ws));
    SPHX_CHECK_LAUNCH();
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
    unsigned blocks = pairBlocks(a.L);
    if (blocks == 0) { return; }
    if (p->hydro_type == SPHX_F32)
        momentumStdVjpKernel<float><<<blocks, kPairWaves * kWave, 0, ws->stream>>>(a, d, sincKernel<float>(ws));
    else
        momentumStdVjpKernel<double><<<blocks, kPairWaves * kWave, 0, ws->stream>>>(a, d, sincKernel<double>(ws));
    SPHX_CHECK_LAUNCH();
}

} // namespace sphx
