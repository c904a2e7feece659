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

constexpr int kMaxRungs = SPHX_MAX_RUNGS;

//! a particle's rung, clamped to the rungs the arguments hold (a larger one is the caller's error)
__device__ __forceinline__ unsigned rungOf(const uint8_t* rung, uint32_t i)
{
    return min(unsigned(rung[i]), unsigned(kMaxRungs - 1));
}

struct PosArgs
{
    DevBox      box;
    int         fbc[3];
    double      dt, dt_m1, gamma, constCv;
    double      dtBack;                // drift: the step taken since the hierarchy's start
    float       dtM1Rung[kMaxRungs];   // block time steps: previous step per rung ...
    const uint8_t* rung;               // ... and each particle's rung (NULL: dt_m1)
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

    const double dt = a.dt, dt_m1 = a.rung ? double(a.dtM1Rung[rungOf(a.rung, i)]) : a.dt_m1;
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

/* driftKernel (positions_gpu.cu:17-60): a particle of a block time-step hierarchy, advanced by dtBack
 * from (Xn, An, dXn) at its last force evaluation, is taken back to Xn with positionUpdate(-dtBack)
 * and advanced by dt from there, both without the periodic wrap (the reference drifts in an open
 * box so that the octree stays valid); velocities follow, temp or u is recovered and advanced with
 * energyUpdate. x_m1, y_m1, z_m1 and du_m1 are read only. */
template<class T>
__global__ void __launch_bounds__(kIntThreads) driftKernel(PosArgs a)
{
#pragma clang fp contract(off)
    uint32_t i;
    if (!groupedIndex(a.groups, i)) { return; }
    const double dt = a.dt, dtBack = a.dtBack;
    const double dt_m1 = double(a.dtM1Rung[a.rung ? rungOf(a.rung, i) : 0u]);
    const double X[3]  = {a.x[i], a.y[i], a.z[i]};
    const double An[3] = {double(ld<T>(a.ax, i)), double(ld<T>(a.ay, i)), double(ld<T>(a.az, i))};
    const double dX[3] = {double(ld<T>(a.x_m1, i)), double(ld<T>(a.y_m1, i)), double(ld<T>(a.z_m1, i))};
    double*      xo[3] = {a.x, a.y, a.z};
    T*           vo[3] = {static_cast<T*>(a.vx), static_cast<T*>(a.vy), static_cast<T*>(a.vz)};
    for (int d = 0; d < 3; ++d)
    {
        // back: positionUpdate(-dtBack, ...) keeps Xn
        double Vnmhalf = dX[d] * (1.0 / dt_m1);
        double Vn      = Vnmhalf + 0.5 * dt_m1 * An[d];
        double dXback  = (Vn + 0.5 * An[d] * fabs(-dtBack)) * -dtBack;
        double Xn      = X[d] + dXback;
        // forward from Xn with the same (An, dXn)
        double Vnp1  = Vn + An[d] * dt;
        double dXnp1 = (Vn + 0.5 * An[d] * fabs(dt)) * dt;
        xo[d][i]     = Xn + dXnp1;
        vo[d][i]     = T(Vnp1);
    }
    const T* du_m1 = static_cast<const T*>(a.du_m1);
    if (a.temp != nullptr)
    {
        const T cv      = (a.constCv < 0) ? idealGasCv(ld<T>(a.mui, i), a.gamma) : T(a.constCv);
        double  u_recov = energyUpdate(a.temp[i] * double(cv), -dtBack, dt_m1, a.du[i], double(du_m1[i]));
        a.temp[i]       = energyUpdate(u_recov, dt, dt_m1, a.du[i], double(du_m1[i])) / double(cv);
    }
    else if (a.u != nullptr)
    {
        double u_recov = energyUpdate(a.u[i], -dtBack, dt_m1, a.du[i], double(du_m1[i]));
        a.u[i]         = energyUpdate(u_recov, dt, dt_m1, a.du[i], double(du_m1[i]));
    }
}

/* groupDivvKernel / groupAccKernel (ts_groups.cu:16-36, 51-71), one wavefront per group instead of
 * one thread: the group's largest divv (signed), or largest |a|^2 = ax^2 + (ay^2 + az^2) in T
 * (norm2, the right fold of util/array.hpp:236-240), as a wave maximum; rounding that maximum to
 * float once equals the reference's running float maximum (rounding is monotone). Then
 * groupDt[g] = min(groupDt[g], Krho / |max|), or min(groupDt[g], etaAcc / sqrt(sqrt(max))), in
 * float with correctly rounded divide and root. Groups of any size: lanes stride over [start, end). */
template<class T, bool Acc>
__global__ void __launch_bounds__(kIntThreads) groupDtKernel(Groups G, float coeff, const T* __restrict__ a,
                                                           const T* __restrict__ b, const T* __restrict__ c,
                                                           float* __restrict__ groupDt)
{
#pragma clang fp contract(off)
    const uint32_t g = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    if (g >= G.num) { return; }
    uint32_t s, e;
    if (G.start == nullptr)
    {
        s = G.first + g * uint32_t(kGroup);
        e = min(s + uint32_t(kGroup), G.last);
    }
    else { s = G.start[g], e = G.end[g]; }
    double m = Acc ? 0.0 : -INFINITY;
    for (uint32_t i = s + laneId(); i < e; i += kWave)
    {
        double v;
        if constexpr (Acc) { v = double(T(a[i] * a[i] + (b[i] * b[i] + c[i] * c[i]))); }
        else { v = double(a[i]); }
        m = v > m ? v : m;
    }
    m = waveMax(m);
    if (laneId() == 0)
    {
        const float mf = float(m);
        const float dt = Acc ? coeff / sqrtf(sqrtf(mf)) : coeff / fabsf(mf);
        groupDt[g]     = fminf(groupDt[g], dt);
    }
}

//! storeRungKernel (ts_groups.cu:84-93): particleRungs[i] = rung over the call's targets
__global__ void __launch_bounds__(kIntThreads) storeRungKernel(Groups G, uint8_t rung, uint8_t* particleRungs)
{
    uint32_t i;
    if (!groupedIndex(G, i)) { return; }
    particleRungs[i] = rung;
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
                     const sphx_params_t* p, double constCv, bool drift)
{
    // constCv: -1 selects the per-particle idealGasCv(mui, gamma); without mui the caller's value,
    // or the reference host side's idealGasCv(muiConst, gamma) in the mass type (positions.hpp:183)
    if (hist->mui) { a.constCv = -1.0; }
    else { a.constCv = constCv >= 0 ? constCv : double(idealGasCv(T(p->muiConst), p->gamma)); }
    const unsigned blocks = ceilDiv(groupedThreads(a.groups), kIntThreads);
    if (drift)
    {
        KernelScope timer(ws, "drift");
        driftKernel<T><<<blocks, kIntThreads, 0, ws->stream>>>(a);
    }
    else
    {
        KernelScope timer(ws, "positions");
        positionsKernel<T><<<blocks, kIntThreads, 0, ws->stream>>>(a);
    }
    SPHX_CHECK_LAUNCH();
}

//! the rung arguments of the block time-step entry points into `a`
void setRungs(PosArgs& a, const float* dtM1Rung, int numRungs, const uint8_t* rung)
{
    SPHX_REQUIRE(dtM1Rung != nullptr, SPHX_ERR_ARG, "the previous steps per rung are missing");
    SPHX_REQUIRE(numRungs >= 1 && numRungs <= kMaxRungs, SPHX_ERR_ARG,
                 "numRungs must lie in [1, " + std::to_string(kMaxRungs) + "]");
    for (int r = 0; r < kMaxRungs; ++r)
        a.dtM1Rung[r] = r < numRungs ? dtM1Rung[r] : 0.0f;
    a.rung  = rung;
    a.dt_m1 = dtM1Rung[0];
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

namespace
{

PosArgs positionArgs(sphx_workspace* ws, const sphx_fields_t* f, const sphx_history_t* hist, uint32_t first,
                     uint32_t last)
{
    SPHX_REQUIRE(f->x && f->y && f->z && f->vx && f->vy && f->vz && f->ax && f->ay && f->az, SPHX_ERR_ARG,
                 "positions need x, y, z, vx, vy, vz, ax, ay, az");
    SPHX_REQUIRE(hist->x_m1 && hist->y_m1 && hist->z_m1, SPHX_ERR_ARG, "positions need x_m1, y_m1, z_m1");
    SPHX_REQUIRE(!(f->temp || f->u) || (f->du && hist->du_m1), SPHX_ERR_ARG, "the energy update needs du and du_m1");
    SPHX_REQUIRE(first <= last, SPHX_ERR_ARG, "invalid range");
    PosArgs a{};
    a.groups = groupsFor(ws, first, last);
    a.x = f->x, a.y = f->y, a.z = f->z, a.temp = f->temp, a.u = f->temp ? nullptr : f->u, a.du = f->du;
    a.vx = f->vx, a.vy = f->vy, a.vz = f->vz, a.ax = f->ax, a.ay = f->ay, a.az = f->az, a.h = f->h;
    a.x_m1 = hist->x_m1, a.y_m1 = hist->y_m1, a.z_m1 = hist->z_m1, a.du_m1 = hist->du_m1, a.mui = hist->mui;
    a.rung = nullptr;
    return a;
}

} // namespace

void computePositions(sphx_workspace* ws, const sphx_box_t* box, const sphx_fields_t* f, const sphx_history_t* hist,
                      const sphx_params_t* p, uint32_t first, uint32_t last, double dt, double dt_m1, double constCv,
                      const float* dtM1Rung, int numRungs, const uint8_t* rung)
{
    PosArgs a = positionArgs(ws, f, hist, first, last);
    if (dtM1Rung != nullptr) { setRungs(a, dtM1Rung, numRungs, rung); }
    if (first == last) { return; }
    a.box = makeDevBox(box);
    bool anyFbc = false;
    for (int d = 0; d < 3; ++d)
    {
        a.fbc[d] = box->boundary[d] == 2;
        anyFbc   = anyFbc || a.fbc[d];
    }
    SPHX_REQUIRE(!anyFbc || f->h, SPHX_ERR_ARG, "fixed boundaries need h");
    a.dt = dt, a.gamma = p->gamma, a.constCv = -1;
    if (dtM1Rung == nullptr) { a.dt_m1 = dt_m1; }
    if (p->hydro_type == SPHX_F32) { launchPositions<float>(ws, a, f, hist, p, constCv, false); }
    else { launchPositions<double>(ws, a, f, hist, p, constCv, false); }
}

void driftPositions(sphx_workspace* ws, const sphx_fields_t* f, const sphx_history_t* hist, const sphx_params_t* p,
                    uint32_t first, uint32_t last, double dt, double dtBack, const float* dtM1Rung, int numRungs,
                    const uint8_t* rung, double constCv)
{
    PosArgs a = positionArgs(ws, f, hist, first, last);
    setRungs(a, dtM1Rung, numRungs, rung);
    if (first == last) { return; }
    a.dt = dt, a.dtBack = dtBack, a.gamma = p->gamma, a.constCv = -1;
    if (p->hydro_type == SPHX_F32) { launchPositions<float>(ws, a, f, hist, p, constCv, true); }
    else { launchPositions<double>(ws, a, f, hist, p, constCv, true); }
}

void groupTimestep(sphx_workspace* ws, const sphx_fields_t* f, const sphx_params_t* p, uint32_t first, uint32_t last,
                   bool acc, float coeff, float* groupDt)
{
    SPHX_REQUIRE(groupDt != nullptr, SPHX_ERR_ARG, "groupDt is missing");
    SPHX_REQUIRE(acc ? (f->ax && f->ay && f->az) : f->divv != nullptr, SPHX_ERR_ARG,
                 acc ? "the acceleration step needs ax, ay, az" : "the divv step needs divv");
    SPHX_REQUIRE(first <= last, SPHX_ERR_ARG, "invalid range");
    const Groups G = groupsFor(ws, first, last);
    if (G.num == 0) { return; }
    const unsigned blocks = ceilDiv(size_t(G.num) * kWave, kIntThreads);
    KernelScope    timer(ws, "group_dt");
    if (p->hydro_type == SPHX_F32)
    {
        if (acc)
            groupDtKernel<float, true><<<blocks, kIntThreads, 0, ws->stream>>>(
                G, coeff, (const float*)f->ax, (const float*)f->ay, (const float*)f->az, groupDt);
        else
            groupDtKernel<float, false><<<blocks, kIntThreads, 0, ws->stream>>>(G, coeff, (const float*)f->divv,
                                                                               nullptr, nullptr, groupDt);
    }
    else
    {
        if (acc)
            groupDtKernel<double, true><<<blocks, kIntThreads, 0, ws->stream>>>(
                G, coeff, (const double*)f->ax, (const double*)f->ay, (const double*)f->az, groupDt);
        else
            groupDtKernel<double, false><<<blocks, kIntThreads, 0, ws->stream>>>(G, coeff, (const double*)f->divv,
                                                                                nullptr, nullptr, groupDt);
    }
    SPHX_CHECK_LAUNCH();
}

void storeRung(sphx_workspace* ws, uint32_t first, uint32_t last, uint8_t rung, uint8_t* particleRungs)
{
    SPHX_REQUIRE(particleRungs != nullptr, SPHX_ERR_ARG, "particleRungs is missing");
    SPHX_REQUIRE(first <= last, SPHX_ERR_ARG, "invalid range");
    if (first == last) { return; }
    const Groups G = groupsFor(ws, first, last);
    storeRungKernel<<<ceilDiv(groupedThreads(G), kIntThreads), kIntThreads, 0, ws->stream>>>(G, rung, particleRungs);
    SPHX_CHECK_LAUNCH();
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
