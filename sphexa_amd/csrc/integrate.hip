// Time integration of one rank's particles: the global time step (device reductions of |a|^2 and
// div v), the Press second-order position update with the Adams-Bashforth energy update, and the
// smoothing-length update. Elementwise and HBM-bound: one thread per particle, coalesced SoA
// accesses, no LDS.
//
// Types follow the reference's GPU instantiations (sph/include/sph/positions_gpu.cu:186-193):
// positions, temperature / internal energy and du are FP64; velocities, accelerations, h, mui and
// the previous-step fields x_m1, y_m1, z_m1, du_m1 are the HydroType T (XM1Type = HydroType in
// both SphTypes and the plain FP64 layout). The update itself runs in FP64 with contraction off,
// in the reference's operation order, so the FP64 form reproduces the reference bit for bit.
#include <algorithm>
#include <cstring>

#include "pairs.hpp"

namespace sphx
{

namespace
{

constexpr unsigned kIntThreads = 256;

//! idealGasCv (sph/include/sph/eos.hpp:18-23): R / mui / (gamma - 1) with R and the result in T1
template<class T1, class T2>
__host__ __device__ __forceinline__ T1 idealGasCv(T1 mui, T2 gamma)
{
    constexpr T1 R = 8.317e7;
    return R / mui / (gamma - T1(1));
}

//! energyUpdate (sph/include/sph/positions.hpp:54-61), Adams-Bashforth 2nd order
__device__ __forceinline__ double energyUpdate(double u_old, double dt, double dt_m1, double du, double du_m1)
{
#pragma clang fp contract(off)
    double u_new = u_old + du * dt + 0.5 * (du - du_m1) / dt_m1 * fabs(dt) * dt;
    if (u_new < 0.) { u_new = u_old * exp(u_new * dt / u_old); }
    return u_new;
}

//! fbcCheck (positions.hpp:46-51)
template<class Th>
__device__ __forceinline__ bool fbcCheck(double coord, Th h, double top, double bottom, bool fbc)
{
#pragma clang fp contract(off)
    return fbc && (fabs(top - coord) < double(Th(2) * h) || fabs(bottom - coord) < double(Th(2) * h));
}

struct PosArgs
{
    DevBox      box;
    int         fbc[3];
    double      dt, dt_m1, gamma, constCv;
    Groups      groups; // the call's targets
    double *    x, *y, *z, *temp, *u;
    const double* du;
    void *      vx, *vy, *vz, *x_m1, *y_m1, *z_m1, *du_m1;
    const void *ax, *ay, *az, *h, *mui;
};

/* computePositionsKernel (positions_gpu.cu:116-168) with positionUpdate (positions.hpp:63-85)
 * and putInBox (domain/include/cstone/sfc/box.hpp:193-214) */
template<class T>
__global__ void __launch_bounds__(kIntThreads) positionsKernel(PosArgs a)
{
#pragma clang fp contract(off)
    uint32_t i;
    if (!groupedIndex(a.groups, i)) { return; }
    T* vx = static_cast<T*>(a.vx);
    T* vy = static_cast<T*>(a.vy);
    T* vz = static_cast<T*>(a.vz);

    const bool anyFBC = a.fbc[0] || a.fbc[1] || a.fbc[2];
    if (anyFBC && vx[i] == T(0) && vy[i] == T(0) && vz[i] == T(0))
    {
        const T h = ld<T>(a.h, i);
        if (fbcCheck(a.x[i], h, a.box.hi[0], a.box.lo[0], a.fbc[0]) ||
            fbcCheck(a.y[i], h, a.box.hi[1], a.box.lo[1], a.fbc[1]) ||
            fbcCheck(a.z[i], h, a.box.hi[2], a.box.lo[2], a.fbc[2]))
        {
            return;
        }
    }

    const double dt = a.dt, dt_m1 = a.dt_m1;
    const double X[3]  = {a.x[i], a.y[i], a.z[i]};
    const double An[3] = {double(ld<T>(a.ax, i)), double(ld<T>(a.ay, i)), double(ld<T>(a.az, i))};
    const double dX[3] = {double(ld<T>(a.x_m1, i)), double(ld<T>(a.y_m1, i)), double(ld<T>(a.z_m1, i))};
    T*           m1[3] = {static_cast<T*>(a.x_m1), static_cast<T*>(a.y_m1), static_cast<T*>(a.z_m1)};
    double*      xo[3] = {a.x, a.y, a.z};
    T*           vo[3] = {vx, vy, vz};
    for (int d = 0; d < 3; ++d)
    {
        double Vnmhalf = dX[d] * (1.0 / dt_m1);
        double Vn      = Vnmhalf + 0.5 * dt_m1 * An[d];
        double Vnp1    = Vn + An[d] * dt;
        double dXnp1   = (Vn + 0.5 * An[d] * fabs(dt)) * dt;
        double Xnp1    = X[d] + dXnp1;
        if (a.box.pbc[d] && Xnp1 > a.box.hi[d]) { Xnp1 -= a.box.len[d]; }
        else if (a.box.pbc[d] && Xnp1 < a.box.lo[d]) { Xnp1 += a.box.len[d]; }
        xo[d][i] = Xnp1;
        m1[d][i] = T(dXnp1);
        vo[d][i] = T(Vnp1);
    }

    T* du_m1 = static_cast<T*>(a.du_m1);
    if (a.temp != nullptr)
    {
        const T cv    = (a.constCv < 0) ? idealGasCv(ld<T>(a.mui, i), a.gamma) : T(a.constCv);
        double  u_old = a.temp[i] * double(cv);
        a.temp[i]     = energyUpdate(u_old, dt, dt_m1, a.du[i], double(du_m1[i])) / double(cv);
        du_m1[i]      = T(a.du[i]);
    }
    else if (a.u != nullptr)
    {
        a.u[i]   = energyUpdate(a.u[i], dt, dt_m1, a.du[i], double(du_m1[i]));
        du_m1[i] = T(a.du[i]);
    }
}

//! updateSmoothingLengthGpuKernel (sph/include/sph/update_h_gpu.cu:40-50): h = updateH(ng0, nc, h)
template<class T>
__global__ void __launch_bounds__(kIntThreads) updateHKernel(Groups G, unsigned ng0, const uint32_t* __restrict__ nc,
                                                           T* __restrict__ h)
{
#pragma clang fp contract(off)
    uint32_t i;
    if (!groupedIndex(G, i)) { return; }
    constexpr T c0 = 1023.0;
    constexpr T ex = 1.0 / 10.0;
    h[i]           = h[i] * T(0.5) * pow(T(1) + c0 * ng0 / T(nc[i]), ex);
}

//! order-preserving map of a double onto uint64 (a > b iff key(a) > key(b)), and its inverse
__device__ __forceinline__ unsigned long long orderedKey(double v)
{
    unsigned long long b = (unsigned long long)__double_as_longlong(v);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__host__ __device__ __forceinline__ double fromOrderedKey(unsigned long long k)
{
    unsigned long long b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    double             v;
    memcpy(&v, &b, sizeof(v));
    return v;
}

/* Maximum over [first, last) of |a|^2 = (ax^2 + ay^2) + az^2 in T (the reference's GPU reduction,
 * domain/include/cstone/primitives/primitives_gpu.cu:179-196), or of divv (MinMaxGpu, :162-171):
 * a grid-stride loop per thread, wave then block maxima through LDS, and one atomic maximum per
 * block on the ordered key (a few hundred atomics on the one address, not one per wave). */
constexpr unsigned kMaxBlocks = 1024;

template<class T, bool Norm>
__global__ void __launch_bounds__(kIntThreads) maxKernel(uint32_t first, uint32_t last, const T* __restrict__ x,
                                                       const T* __restrict__ y, const T* __restrict__ z,
                                                       unsigned long long* out)
{
#pragma clang fp contract(off)
    __shared__ double waveMaxima[kIntThreads / kWave];
    double            v = -INFINITY;
    for (uint32_t i = first + blockIdx.x * blockDim.x + threadIdx.x; i < last; i += gridDim.x * blockDim.x)
    {
        double e;
        if constexpr (Norm) { e = double(x[i] * x[i] + y[i] * y[i] + z[i] * z[i]); }
        else { e = double(x[i]); }
        v = e > v ? e : v;
    }
    v = waveMax(v);
    if (laneId() == 0) { waveMaxima[threadIdx.x / kWave] = v; }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        for (unsigned w = 1; w < kIntThreads / kWave; ++w)
            v = waveMaxima[w] > v ? waveMaxima[w] : v;
        atomicMax(out, orderedKey(v));
    }
}

template<class T>
double deviceMax(sphx_workspace* ws, uint32_t first, uint32_t last, const void* x, const void* y, const void* z,
                 bool norm)
{
    ws->counters.reserve(2 * kCounterSlots);
    unsigned long long* slot = reinterpret_cast<unsigned long long*>(ws->counters.get()) + 4;
    fill64(ws, reinterpret_cast<uint64_t*>(slot), 0, 1); // below the key of every double
    const unsigned blocks = std::min(ceilDiv(last - first, kIntThreads), kMaxBlocks);
    if (norm)
        maxKernel<T, true><<<blocks, kIntThreads, 0, ws->stream>>>(first, last, static_cast<const T*>(x),
                                                                   static_cast<const T*>(y),
                                                                   static_cast<const T*>(z), slot);
    else
        maxKernel<T, false><<<blocks, kIntThreads, 0, ws->stream>>>(first, last, static_cast<const T*>(x), nullptr,
                                                                    nullptr, slot);
    SPHX_CHECK_LAUNCH();
    unsigned long long key = 0;
    SPHX_HIP(hipMemcpyAsync(&key, slot, sizeof(key), hipMemcpyDeviceToHost, ws->stream));
    SPHX_HIP(hipStreamSynchronize(ws->stream));
    return fromOrderedKey(key);
}

template<class T>
void launchPositions(sphx_workspace* ws, PosArgs& a, const sphx_fields_t* f, const sphx_history_t* hist,
                     const sphx_params_t* p, double constCv)
{
    // constCv: -1 selects the per-particle idealGasCv(mui, gamma); without mui the caller's value,
    // or the reference host side's idealGasCv(muiConst, gamma) in the mass type (positions.hpp:183)
    if (hist->mui) { a.constCv = -1.0; }
    else { a.constCv = constCv >= 0 ? constCv : double(idealGasCv(T(p->muiConst), p->gamma)); }
    KernelScope timer(ws, "positions");
    positionsKernel<T><<<ceilDiv(groupedThreads(a.groups), kIntThreads), kIntThreads, 0, ws->stream>>>(a);
    SPHX_CHECK_LAUNCH();
}

} // namespace

void rhoTimestep(sphx_workspace* ws, const sphx_fields_t* f, const sphx_params_t* p, uint32_t first, uint32_t last,
                 sphx_timestep_t* ts)
{
    SPHX_REQUIRE(f->divv, SPHX_ERR_ARG, "Divv needs to be available in rhoTimestep");
    SPHX_REQUIRE(first < last, SPHX_ERR_ARG, "invalid range");
    double maxDivv = p->hydro_type == SPHX_F32 ? deviceMax<float>(ws, first, last, f->divv, nullptr, nullptr, false)
                                               : deviceMax<double>(ws, first, last, f->divv, nullptr, nullptr, false);
    ts->minDtRho = ts->Krho / fabs(maxDivv);
}

void computeTimestep(sphx_workspace* ws, const sphx_fields_t* f, const sphx_params_t* p, uint32_t first,
                     uint32_t last, sphx_timestep_t* ts)
{
    double minDtAcc = INFINITY;
    if (ts->g != 0.0)
    {
        SPHX_REQUIRE(f->ax && f->ay && f->az, SPHX_ERR_ARG, "the acceleration limit needs ax, ay, az");
        SPHX_REQUIRE(first < last, SPHX_ERR_ARG, "invalid range");
        double maxAccSq = p->hydro_type == SPHX_F32 ? deviceMax<float>(ws, first, last, f->ax, f->ay, f->az, true)
                                                    : deviceMax<double>(ws, first, last, f->ax, f->ay, f->az, true);
        minDtAcc = ts->etaAcc * sqrt(ts->eps / sqrt(maxAccSq));
    }
    double dt    = std::min({minDtAcc, ts->minDtCourant, ts->minDtRho, ts->maxDtIncrease * ts->minDt});
    ts->ttot     = ts->ttot + dt;
    ts->minDt_m1 = ts->minDt;
    ts->minDt    = dt;
}

void computePositions(sphx_workspace* ws, const sphx_box_t* box, const sphx_fields_t* f, const sphx_history_t* hist,
                      const sphx_params_t* p, uint32_t first, uint32_t last, double dt, double dt_m1, double constCv)
{
    SPHX_REQUIRE(f->x && f->y && f->z && f->vx && f->vy && f->vz && f->ax && f->ay && f->az, SPHX_ERR_ARG,
                 "positions need x, y, z, vx, vy, vz, ax, ay, az");
    SPHX_REQUIRE(hist->x_m1 && hist->y_m1 && hist->z_m1, SPHX_ERR_ARG, "positions need x_m1, y_m1, z_m1");
    SPHX_REQUIRE(!(f->temp || f->u) || (f->du && hist->du_m1), SPHX_ERR_ARG, "the energy update needs du and du_m1");
    SPHX_REQUIRE(first <= last, SPHX_ERR_ARG, "invalid range");
    if (first == last) { return; }
    PosArgs a;
    a.box = makeDevBox(box);
    bool anyFbc = false;
    for (int d = 0; d < 3; ++d)
    {
        a.fbc[d] = box->boundary[d] == 2;
        anyFbc   = anyFbc || a.fbc[d];
    }
    SPHX_REQUIRE(!anyFbc || f->h, SPHX_ERR_ARG, "fixed boundaries need h");
    a.dt = dt, a.dt_m1 = dt_m1, a.gamma = p->gamma, a.constCv = -1;
    a.groups = groupsFor(ws, first, last);
    a.x = f->x, a.y = f->y, a.z = f->z, a.temp = f->temp, a.u = f->temp ? nullptr : f->u, a.du = f->du;
    a.vx = f->vx, a.vy = f->vy, a.vz = f->vz, a.ax = f->ax, a.ay = f->ay, a.az = f->az, a.h = f->h;
    a.x_m1 = hist->x_m1, a.y_m1 = hist->y_m1, a.z_m1 = hist->z_m1, a.du_m1 = hist->du_m1, a.mui = hist->mui;
    if (p->hydro_type == SPHX_F32) { launchPositions<float>(ws, a, f, hist, p, constCv); }
    else { launchPositions<double>(ws, a, f, hist, p, constCv); }
}

void updateSmoothingLength(sphx_workspace* ws, const sphx_fields_t* f, const sphx_params_t* p, uint32_t first,
                           uint32_t last)
{
    SPHX_REQUIRE(f->h && f->nc, SPHX_ERR_ARG, "updateSmoothingLength needs h and nc");
    SPHX_REQUIRE(first <= last, SPHX_ERR_ARG, "invalid range");
    if (first == last) { return; }
    const Groups   G      = groupsFor(ws, first, last);
    const unsigned blocks = ceilDiv(groupedThreads(G), kIntThreads);
    KernelScope    timer(ws, "update_h");
    if (p->hydro_type == SPHX_F32)
        updateHKernel<float><<<blocks, kIntThreads, 0, ws->stream>>>(G, p->ng0, f->nc, (float*)f->h);
    else
        updateHKernel<double><<<blocks, kIntThreads, 0, ws->stream>>>(G, p->ng0, f->nc, (double*)f->h);
    SPHX_CHECK_LAUNCH();
}

} // namespace sphx
