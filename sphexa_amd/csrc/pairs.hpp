// Helpers shared by the SPH pair-loop kernels (std and VE).
//
// Launch shape: one wavefront per group of 64 consecutive targets (the neighbour-search group),
// 4 waves per block, blocks renumbered per XCD (xcdBlock) so that neighbouring groups share an
// L2. Lane l of wave g owns target first + 64 g + l and walks its own list slots
// k < min(nc-1, ngmax) in the quad layout of common.hpp (listOffset).
//
// Two arithmetic forms share each kernel template:
//  * T = double ("plain" FP64): the reference JLoop formulas in FP64.
//  * T = float ("mixed"): HydroType fields in float (the reference's SphTypes layout); pair
//    separations are formed in FP64 from the FP64 positions, folded, then rounded to float;
//    kernel lookups, gradients and pair terms in FP32; density and energy sums with
//    compensated (Kahan) summation; the IAD 3x3 inversion in FP64.
#pragma once

#include <type_traits>

#include "common.hpp"

namespace sphx
{

#ifndef SPHX_PAIR_WAVES_PER_BLOCK
#define SPHX_PAIR_WAVES_PER_BLOCK 4
#endif
constexpr int kPairWaves = SPHX_PAIR_WAVES_PER_BLOCK;

template<class T>
constexpr bool isMixed = std::is_same_v<T, float>;

// neighbours per batch (NeighborBatch) of the light loops (XMass, grad-h, IAD, AV switches) and of
// the momentum-energy loops, per form; larger batches hide more latency and cost registers
#ifndef SPHX_BATCH_LIGHT_F32
#define SPHX_BATCH_LIGHT_F32 4
#endif
#ifndef SPHX_BATCH_LIGHT_F64
#define SPHX_BATCH_LIGHT_F64 8
#endif
#ifndef SPHX_BATCH_MOM_F32
#define SPHX_BATCH_MOM_F32 2
#endif
#ifndef SPHX_BATCH_MOM_F64
#define SPHX_BATCH_MOM_F64 2
#endif

// waves per SIMD the register allocator must leave room for (amdgpu_waves_per_eu) in the light
// loops and in momentum-energy, per form; 1 leaves the choice to the compiler
#ifndef SPHX_WAVES_LIGHT_F32
#define SPHX_WAVES_LIGHT_F32 1
#endif
#ifndef SPHX_WAVES_LIGHT_F64
#define SPHX_WAVES_LIGHT_F64 1
#endif
#ifndef SPHX_WAVES_MOM_F32
#define SPHX_WAVES_MOM_F32 1
#endif
#ifndef SPHX_WAVES_MOM_F64
#define SPHX_WAVES_MOM_F64 1
#endif
template<class T>
constexpr int kWavesLight = isMixed<T> ? SPHX_WAVES_LIGHT_F32 : SPHX_WAVES_LIGHT_F64;
template<class T>
constexpr int kWavesMom = isMixed<T> ? SPHX_WAVES_MOM_F32 : SPHX_WAVES_MOM_F64;

//! launch attributes of the pair-loop kernels (T: the HydroType template parameter)
#define SPHX_LIGHT_LOOP(T) __launch_bounds__(kPairWaves * kWave) __attribute__((amdgpu_waves_per_eu(kWavesLight<T>)))
#define SPHX_MOM_LOOP(T) __launch_bounds__(kPairWaves * kWave) __attribute__((amdgpu_waves_per_eu(kWavesMom<T>)))
template<class T>
constexpr int kBatchLight = isMixed<T> ? SPHX_BATCH_LIGHT_F32 : SPHX_BATCH_LIGHT_F64;
template<class T>
constexpr int kBatchMom = isMixed<T> ? SPHX_BATCH_MOM_F32 : SPHX_BATCH_MOM_F64;

struct Lists
{
    const uint32_t* nbrs;
    const uint32_t* nc;
    uint32_t        first, last, ngmax, ngmax4; // [first, last): the searched targets
    Groups          groups;                     // the call's target groups, one wavefront each
};

struct Lane
{
    uint32_t        i;     // target index
    bool            valid; // lane holds a target
    unsigned        n;     // neighbours to visit
    const uint32_t* list;  // this lane's slot 0; slot k at list[laneSlot(k)]
};

//! the wave's group index among the call's target groups
__device__ __forceinline__ uint32_t waveGroup()
{
    return __builtin_amdgcn_readfirstlane(xcdBlock(blockIdx.x, gridDim.x) * kPairWaves + threadIdx.x / kWave);
}

/*! lane l of the wave of group g takes the group's l-th target; its list sits at block (i - first) / 64,
 * lane (i - first) % 64 of the searched range, whatever the grouping. Targets outside the searched
 * range [first, last) are not taken. */
__device__ __forceinline__ Lane laneTarget(const Lists& L)
{
    Lane           t;
    const unsigned lane = threadIdx.x % kWave;
    const uint32_t g    = waveGroup();
    uint32_t       i0 = 0, i1 = 0;
    bool           oversized;
    if (g < L.groups.num) { groupRange(L.groups, g, i0, i1, oversized); }
    t.i            = i0 + lane;
    t.valid        = t.i < i1 && t.i >= L.first && t.i < L.last;
    t.n            = t.valid ? min(L.nc[t.i] - 1, L.ngmax) : 0u;
    const uint32_t r = t.valid ? t.i - L.first : 0u;
    t.list         = L.nbrs + listOffset(r / kGroup, L.ngmax4, 0, r % kGroup);
    return t;
}

inline unsigned pairBlocks(const Lists& L) { return ceilDiv(L.groups.num, kPairWaves); }

/*! Dynamic LDS per block that caps a pair loop at `waves` waves per SIMD (0: no cap): the CU's 160 KB
 * then holds at most 4 `waves` / kPairWaves blocks. A gather-bound loop can run faster with fewer waves in
 * flight than its register use allows, since more waves evict each other's records from L1. Caps per
 * loop family of the FP32 form, from a sweep of 2-6 at 8M particles: AV switches 3 (2.98 -> 2.75 ms,
 * 4 waves by registers otherwise), XMass 6 (1.27 -> 1.22 ms, 8 otherwise); the others ran as fast or
 * faster uncapped. */
inline size_t occupancyCap(int waves)
{
    // blocks per CU = 4 SIMDs x waves / (waves per block)
    return waves > 0 ? size_t(160 * 1024) * kPairWaves / (4 * size_t(waves)) : 0;
}
#ifndef SPHX_CAP_XMASS
#define SPHX_CAP_XMASS 6
#endif
#ifndef SPHX_CAP_GRADH
#define SPHX_CAP_GRADH 0
#endif
#ifndef SPHX_CAP_IAD
#define SPHX_CAP_IAD 0
#endif
#ifndef SPHX_CAP_AV
#define SPHX_CAP_AV 3
#endif
#ifndef SPHX_CAP_MOM
#define SPHX_CAP_MOM 0
#endif

//! lt::lookup (sph/include/sph/table_lookup.hpp:10-26): linear interpolation of a 20000-point
//! table on [0,2], zero beyond the support. The table is stored per interval (both end values
//! adjacent), so one lookup is one vector gather instead of two scalar ones. Interval 19999, one
//! past the last, holds zeros: arguments at or beyond the support clamp onto it and interpolate
//! to +0, the reference's value there, without a branch around the load. Arguments are >= 0.
template<class T>
__device__ __forceinline__ T interpolate(T t0, T t1, T v, int idx)
{
    constexpr T dx         = T(2.0) / (kTableSize - 1);
    constexpr T invDx      = T(1) / dx;
    T           derivative = (t1 - t0) * invDx;
    return t0 + derivative * (v - T(idx) * dx);
}

template<class T>
__device__ __forceinline__ int tableInterval(T v)
{
    constexpr int numIntervals = kTableSize - 1;
    constexpr T   dx           = T(2.0) / numIntervals;
    constexpr T   invDx        = T(1) / dx;
    T             vc           = v < T(4) ? v : T(4); // keeps the conversion in range
    return min(int(vc * invDx), numIntervals);
}

template<class T>
__device__ __forceinline__ T lookup(const TabPair<T>* __restrict__ table, T v)
{
    int              idx = tableInterval(v);
    const TabPair<T> e   = table[idx];
    return interpolate(e.t0, e.t1, v, idx);
}

//! W(v) and dW/dv(v) from one interval record
template<class T>
__device__ __forceinline__ void lookupBoth(const TabQuad<T>* __restrict__ table, T v, T& w, T& dw)
{
    int              idx = tableInterval(v);
    const TabQuad<T> e   = table[idx];
    w                    = interpolate(e.w0, e.w1, v, idx);
    dw                   = interpolate(e.d0, e.d1, v, idx);
}

//! fast reciprocal / square root for the FP32 form, IEEE operations for FP64
template<class T>
__device__ __forceinline__ T fastDiv(T a, T b)
{
    if constexpr (isMixed<T>) { return a * __builtin_amdgcn_rcpf(b); }
    else { return a / b; }
}

template<class T>
__device__ __forceinline__ T fastSqrt(T a)
{
    if constexpr (isMixed<T>) { return __builtin_amdgcn_sqrtf(a); }
    else { return sqrt(a); }
}

/*! Batched walk over one lane's neighbour list.
 *
 * A pair loop visits its neighbours U at a time (U = 1, 2 or a multiple of 4): load(k) makes the records of
 * slots k..k+U-1 current (their list entries were requested one batch earlier with aligned vector
 * loads from the lane's quads) and requests the list entries of the following batch.
 * The loop then forms the U separations and issues the U table fetches before it consumes any of
 * them, so one batch costs two memory round trips (records, tables) instead of three per
 * neighbour. Slots beyond the lane's count take the batch's first valid entry and their pair
 * terms are zeroed (live). Slot 0 of a valid target always holds a valid particle index (the
 * search writes the target itself there when it has no neighbours), so every index used for a
 * record load names a particle. Valid targets only. */
template<int U, class R>
struct NeighborBatch
{
    static_assert(U == 1 || U == 2 || U % 4 == 0, "batch of 1, 2 or a multiple of 4 slots");

    const uint32_t* list;
    const R*        rec;
    unsigned        n, lastSlot;
    uint32_t        next[U]; // list entries of the next batch
    uint32_t        j[U];    // particle indices of the current batch
    R               r[U];    // records of the current batch

    __device__ __forceinline__ NeighborBatch(const Lane& t, const R* __restrict__ records)
        : list(t.list), rec(records), n(t.n), lastSlot(t.n ? t.n - 1 : 0)
    {
        slots(0, next);
#pragma unroll
        for (int u = 0; u < U; ++u)
        {
            // an opaque definition: otherwise the compiler merges this load with the look-ahead
            // load of the loop into one load at the top of each batch, undoing the look-ahead
            asm volatile("" : "+v"(next[u]));
        }
    }

    //! list entries of slots k..k+U-1 (clamped as described above)
    __device__ __forceinline__ void slots(unsigned k, uint32_t out[U]) const
    {
        const unsigned  kq = min(k, lastSlot) & ~unsigned(U - 1);
        const uint32_t* p  = list + laneSlot(kq);
        uint32_t        v[U];
        if constexpr (U % 4 == 0)
        {
#pragma unroll
            for (int c = 0; c < U; c += 4)
            {
                const uint4 q = *reinterpret_cast<const uint4*>(list + laneSlot(kq + c));
                v[c] = q.x, v[c + 1] = q.y, v[c + 2] = q.z, v[c + 3] = q.w;
            }
        }
        else if constexpr (U == 2)
        {
            const uint2 q = *reinterpret_cast<const uint2*>(p);
            v[0] = q.x, v[1] = q.y;
        }
        else { v[0] = p[0]; }
#pragma unroll
        for (int u = 0; u < U; ++u)
            out[u] = kq + u <= lastSlot ? v[u] : v[0];
    }

    __device__ __forceinline__ void load(unsigned k)
    {
#pragma unroll
        for (int u = 0; u < U; ++u)
        {
            j[u] = next[u];
            r[u] = rec[j[u]];
        }
        slots(k + U, next);
    }

    //! whether member u of the batch starting at slot k is a real neighbour
    __device__ __forceinline__ bool in(unsigned k, int u) const { return k + u < n; }

    //! a kernel weight of member u, zeroed for a clamped slot. Every pair term of the loops is
    //! proportional to a kernel weight, so a clamped slot contributes exactly zero without a branch
    //! (a branch would let the compiler sink that member's loads behind it).
    template<class T>
    __device__ __forceinline__ T live(unsigned k, int u, T w) const
    {
        return in(k, u) ? w : T(0);
    }
};

//! table entry of argument v, fetched ahead of its interpolation
template<class T, class E>
struct TableFetch
{
    E   e;
    T   v;
    int idx;
};

template<class T, class E>
__device__ __forceinline__ TableFetch<T, E> fetch(const E* __restrict__ table, T v)
{
    TableFetch<T, E> f;
    f.v   = v;
    f.idx = tableInterval(v);
    f.e   = table[f.idx];
    return f;
}

template<class T>
__device__ __forceinline__ T value(const TableFetch<T, TabPair<T>>& f)
{
    return interpolate(f.e.t0, f.e.t1, f.v, f.idx);
}

template<class T>
__device__ __forceinline__ void values(const TableFetch<T, TabQuad<T>>& f, T& w, T& dw)
{
    w  = interpolate(f.e.w0, f.e.w1, f.v, f.idx);
    dw = interpolate(f.e.d0, f.e.d1, f.v, f.idx);
}

/*! The SPH kernel W(v) = sinc(pi v / 2)^n and dW/dv on [0, 2), zero from v = 2 on.
 *
 * The FP64 form interpolates the reference's 20000-point tables (lookup), so its values are the
 * reference's. The FP32 form evaluates the function itself. A table lookup is a 64-lane gather
 * that touches about one cache line per lane, and the pair loops are bound by that gather
 * traffic, while the ALUs are mostly idle. The evaluation runs in FP64 arithmetic, which on this
 * GPU issues at the FP32 rate (FP32 ops are not packed here), and is rounded once to FP32: its
 * error is the final rounding, below that of the FP32 table (values rounded to FP32, then
 * interpolated in FP32), and its difference from the FP64 table's interpolant is the
 * interpolation error, about 6e-9 of W(0) for n = 6.
 *
 * With x = pi v / 2: sin x from its Taylor polynomial on [0, pi/2] (to t^13; the remainder is
 * below 7e-10), reached for v > 1 through sin x = sin y, cos x = -cos y, y = (2 - v) pi/2 formed
 * from the exact 2 - v, so sinc keeps its relative accuracy up to the edge of the support;
 * sinc = sin x / x and d sinc/dx = (cos x - sinc) / x (1 and 0 at x = 0), with 1/x from the
 * hardware reciprocal and one Newton step; W = sinc^n by repeated multiplication for an integer
 * n (the reference's sinc indices), else through FP32 log2 / exp2; dW/dv = n sinc^(n-1)
 * d sinc/dx pi/2. */
struct DirectW
{
    float w, v;
};
struct DirectWD
{
    float w, dw, v;
};

__device__ __forceinline__ float value(const DirectW& f) { return f.w; }

__device__ __forceinline__ void values(const DirectWD& f, float& w, float& dw)
{
    w  = f.w;
    dw = f.dw;
}

//! x^k for a small integer k >= 0 (wave-uniform loop), or 2^(k log2 x) when k is not an integer
__device__ __forceinline__ double sincPower(double x, float k, int ki)
{
    if (ki == 5)
    {
        // the default sinc index 6 (sinc^5 here): the products of the loop below, unrolled
        const double x2 = x * x;
        return x * (x2 * x2);
    }
    if (ki < 0) { return double(__builtin_amdgcn_exp2f(k * __builtin_amdgcn_logf(float(x)))); }
    double r = 1.0, b = x;
    for (int e = ki; e > 0; e >>= 1)
    {
        if (e & 1) { r *= b; }
        b *= b;
    }
    return r;
}

template<bool Derivative>
__device__ __forceinline__ void sincDirect(float vf, float n, int ni, float& w, float& dw)
{
    constexpr double halfPi = 1.57079632679489661923;
    const double     v      = vf;
    const bool       upper  = v > 1.0;
    const double     x      = v * halfPi;
    const double     y      = upper ? (2.0 - v) * halfPi : x; // in [0, pi/2] on the support
    const double     y2     = y * y;
    // sin y = y (1 - y^2/3! + ... - y^12/13!)
    double sy = -1.0 / 6227020800.0;
    sy        = sy * y2 + 1.0 / 39916800.0;
    sy        = sy * y2 - 1.0 / 362880.0;
    sy        = sy * y2 + 1.0 / 5040.0;
    sy        = sy * y2 - 1.0 / 120.0;
    sy        = sy * y2 + 1.0 / 6.0;
    sy        = y - y * y2 * sy;
    double xInv = double(__builtin_amdgcn_rcpf(float(x)));
    xInv        = xInv + xInv * (1.0 - x * xInv);
    const bool   zero   = vf == 0.0f;
    const double sinc   = zero ? 1.0 : sy * xInv;
    const bool   inside = vf < 2.0f;
    const double pm1    = sincPower(sinc, n - 1.0f, ni - 1); // sinc^(n-1)
    w                   = inside ? float(pm1 * sinc) : 0.0f;
    if constexpr (Derivative)
    {
        // cos y = 1 - y^2/2! + ... + y^14/14!
        double cy = -1.0 / 87178291200.0;
        cy        = cy * y2 + 1.0 / 479001600.0;
        cy        = cy * y2 - 1.0 / 3628800.0;
        cy        = cy * y2 + 1.0 / 40320.0;
        cy        = cy * y2 - 1.0 / 720.0;
        cy        = cy * y2 + 1.0 / 24.0;
        cy        = cy * y2 - 0.5;
        cy        = 1.0 + y2 * cy;
        const double cx = upper ? -cy : cy;
        const double ds = zero ? 0.0 : (cx - sinc) * xInv;
        dw              = inside ? float(double(n) * pm1 * ds * halfPi) : 0.0f;
    }
}

template<class T>
struct SincKernel;

//! FP64: the reference's tables (W, dW/dv, and both per interval)
template<>
struct SincKernel<double>
{
    const TabPair<double>* w;
    const TabPair<double>* dw;
    const TabQuad<double>* wq;
    using FetchW  = TableFetch<double, TabPair<double>>;
    using FetchWD = TableFetch<double, TabQuad<double>>;

    __device__ __forceinline__ FetchW  fetchW(double v) const { return fetch(w, v); }
    __device__ __forceinline__ FetchWD fetchWD(double v) const { return fetch(wq, v); }
    __device__ __forceinline__ double  W(double v) const { return lookup(w, v); }
    __device__ __forceinline__ double  dW(double v) const { return lookup(dw, v); }
};

//! FP32: evaluated directly (sincDirect)
template<>
struct SincKernel<float>
{
    float n;
    int   ni; // n when it is an integer, else -1
    using FetchW  = DirectW;
    using FetchWD = DirectWD;

    __device__ __forceinline__ FetchW fetchW(float v) const
    {
        FetchW f;
        float  unused;
        f.v = v;
        sincDirect<false>(v, n, ni, f.w, unused);
        return f;
    }
    __device__ __forceinline__ FetchWD fetchWD(float v) const
    {
        FetchWD f;
        f.v = v;
        sincDirect<true>(v, n, ni, f.w, f.dw);
        return f;
    }
    __device__ __forceinline__ float W(float v) const { return fetchW(v).w; }
    __device__ __forceinline__ float dW(float v) const { return fetchWD(v).dw; }
};

/*! Whether a wave must apply the legacy fold (sfc/box.hpp:218-238). The fold changes a separation
 * only when |x_i - x_j| exceeds the cut-off 2h_i. For a listed neighbour (minimum-image distance
 * below 2h_i) that needs the raw difference to wrap across a periodic face, which puts the target
 * within 2h_i of that face. Waves without such a target run the loop without the fold. */
__device__ __forceinline__ bool waveNeedsFold(const DevBox& b, double x, double y, double z, double r2h)
{
    const double p[3] = {x, y, z};
    const double r    = r2h * (1.0 + 1e-6);
    bool         near = false;
    for (int d = 0; d < 3; ++d)
        near |= b.pbc[d] && (p[d] - b.lo[d] <= r || b.hi[d] - p[d] <= r);
    return ballot(near) != 0;
}

/*! separation x_i - x_j in the pair type, formed in FP64 (exact for T = double, then rounded for
 * float). With Fold, the legacy minimum image with cut-off r (sfc/box.hpp:218-238): per periodic
 * dimension subtract L if d > r, add L if d < -r. Written with selects, d - 0 == d exactly, so the
 * result is that of the reference's branches. */
template<bool Fold, class T>
__device__ __forceinline__ void separation(const DevBox& b, double r, double xi, double yi, double zi, double xj,
                                           double yj, double zj, T& rx, T& ry, T& rz)
{
    double d[3] = {xi - xj, yi - yj, zi - zj};
    if constexpr (Fold)
    {
        for (int q = 0; q < 3; ++q)
        {
            bool   up   = b.pbc[q] && d[q] > r;
            bool   down = b.pbc[q] && d[q] < -r;
            double s    = up ? b.len[q] : (down ? -b.len[q] : 0.0);
            d[q]        = d[q] - s;
        }
    }
    rx = T(d[0]), ry = T(d[1]), rz = T(d[2]);
}

//! separation of a pair and its length
template<class T>
struct Sep
{
    T x, y, z, d;
};

/*! Position of a particle in the neighbour records: FP64 in the plain form; in the FP32 form
 * 32-bit fixed point over a period P = L / 2^s of the box length L, u = rint((x - lo) 2^32 / P)
 * mod 2^32 per dimension. The shift s is set after each search from its largest h: the largest
 * s <= 16 with P >= 8 h_max (fixedPointShift), so a listed pair, closer than 2h_max in each
 * dimension, is closer than P/4. A record position is 12 bytes instead of 24; coordinates are
 * quantized to P / 2^32, about 2e-9 h_max whatever the box size (an open box much larger than the
 * particles' extent, or a wide range of h, keeps the resolution of the smoothing lengths, as FP32
 * offsets from a nearby cell would). The separation is converted to FP32 and scaled by P / 2^32
 * (about one ulp in all). Which integer difference stands for x_i - x_j:
 *  - while every target has 2h < L/2 in every dimension, the wrapping difference u_i - u_j read as
 *    a signed 32-bit integer. A listed pair is closer than 2h < P/2 in each dimension, so this is
 *    the raw difference in an open dimension, and in a periodic one (P divides L) the minimum
 *    image, which equals the reference's fold with cut-off 2h (sfc/box.hpp:218-238) for any pair
 *    closer than 2h < L/2;
 *  - otherwise (2h >= L/2: a box a few smoothing lengths wide, as for a few hundred particles or
 *    fewer; then s = 0, P = L), the exact difference u_i - u_j in 64 bits, the raw difference of
 *    coordinates that lie in [lo, hi] (the box update of the domain sync keeps open dimensions
 *    there), folded in a periodic dimension with cut-off 2h as the reference does. */
template<class T>
using Coord = std::conditional_t<isMixed<T>, uint32_t, double>;

__device__ __forceinline__ uint32_t fixedCoord(const DevBox& b, int d, double x)
{
    return uint32_t(uint64_t(int64_t(rint((x - b.lo[d]) * b.fixScale[d]))));
}

template<class T>
__device__ __forceinline__ Coord<T> coordOf(const DevBox& b, int d, double x)
{
    if constexpr (isMixed<T>) { return fixedCoord(b, d, x); }
    else { return x; }
}

/*! exact fixed-point separation of one dimension for boxes a few smoothing lengths wide: the 64-bit
 * difference, then the legacy fold with cut-off r in a periodic dimension (see Coord above) */
__device__ __forceinline__ float wideFixedSeparation(const DevBox& b, int q, double r, uint32_t ui, uint32_t uj)
{
    double d = double(int64_t(ui) - int64_t(uj)); // exact: |d| < 2^33
    if (b.pbc[q])
    {
        double dx = d * b.fixStepD[q];
        d         = dx > r ? d - 4294967296.0 : (dx < -r ? d + 4294967296.0 : d);
    }
    return float(d) * b.fixStep[q];
}

template<bool Fold, class T, class Ri, class Rj>
__device__ __forceinline__ Sep<T> pairSeparation(const DevBox& b, double r2h, const Ri& ri, const Rj& rj)
{
    Sep<T> s;
    if constexpr (isMixed<T> && Fold)
    {
        s.x = wideFixedSeparation(b, 0, r2h, ri.x, rj.x);
        s.y = wideFixedSeparation(b, 1, r2h, ri.y, rj.y);
        s.z = wideFixedSeparation(b, 2, r2h, ri.z, rj.z);
    }
    else if constexpr (isMixed<T>)
    {
        s.x = float(int32_t(ri.x - rj.x)) * b.fixStep[0];
        s.y = float(int32_t(ri.y - rj.y)) * b.fixStep[1];
        s.z = float(int32_t(ri.z - rj.z)) * b.fixStep[2];
    }
    else { separation<Fold>(b, r2h, ri.x, ri.y, ri.z, rj.x, rj.y, rj.z, s.x, s.y, s.z); }
    s.d = fastSqrt(s.x * s.x + s.y * s.y + s.z * s.z);
    return s;
}

/*! whether the wave of target record ri must fold its FP64 separations; in the FP32 form, Wide: the
 * last search met a target whose support 2h reaches half a box length, where the wrapping 32-bit
 * difference stops standing for the reference's separation (see Coord above), so the loop takes the
 * exact 64-bit difference. Wide is a kernel template argument, chosen on the host (launchWide), so
 * that the common loop carries none of the registers of the exact path. */
template<class T, bool Wide, class R>
__device__ __forceinline__ bool recordNeedsFold(const DevBox& b, const R& ri, double r2h)
{
    if constexpr (isMixed<T>) { return Wide; }
    else { return waveNeedsFold(b, ri.x, ri.y, ri.z, r2h); }
}

//! host: launch(std::bool_constant<Wide>) with Wide = `wide` in the FP32 form, false in the FP64 form
template<class T, class F>
inline void launchWide(bool wide, F&& launch)
{
    if constexpr (isMixed<T>)
    {
        if (wide)
        {
            launch(std::true_type{});
            return;
        }
    }
    launch(std::false_type{});
}

//! run `loop` with the fold compiled in or out, chosen once per wave
template<class F>
__device__ __forceinline__ void withFold(bool fold, F&& loop)
{
    if (fold) { loop(std::true_type{}); }
    else { loop(std::false_type{}); }
}

//! compensated accumulator (Kahan) for the FP32 form, plain sum for FP64
template<class T>
struct Acc
{
    T s, c;
    __device__ __forceinline__ explicit Acc(T init = T(0)) : s(init), c(0) {}
    __device__ __forceinline__ void add(T v)
    {
        if constexpr (isMixed<T>)
        {
            T y  = v - c;
            T t2 = s + y;
            c    = (t2 - s) - y;
            s    = t2;
        }
        else { s += v; }
    }
    //! add(v) where p holds, the state untouched elsewhere (branch-free)
    __device__ __forceinline__ void addIf(bool p, T v)
    {
        if constexpr (isMixed<T>)
        {
            T y  = v - c;
            T t2 = s + y;
            T c2 = (t2 - s) - y;
            s    = p ? t2 : s;
            c    = p ? c2 : c;
        }
        else { s = p ? s + v : s; }
    }
    __device__ __forceinline__ T value() const { return s; }
};

//! positive-double minimum through the integer order of their bit patterns
__device__ __forceinline__ void atomicMinPositive(unsigned long long* dst, double v)
{
    atomicMin(dst, (unsigned long long)__double_as_longlong(v));
}

template<class T>
__device__ __forceinline__ T ld(const void* p, size_t i)
{
    return static_cast<const T*>(p)[i];
}

template<class T>
__device__ __forceinline__ void st(void* p, size_t i, T v)
{
    static_cast<T*>(p)[i] = v;
}

//! IAD: scale tau by 2^-mean(ilogb) and invert through the adjugate (hydro_std/iad_kern.hpp:52-76),
//! computed in FP64 for both forms; returns c11, c12, c13, c22, c23, c33
__device__ __forceinline__ void iadInvert(double t11, double t12, double t13, double t22, double t23, double t33,
                                          double hi, double K, double out[6])
{
    auto getExp   = [](double v) { return v == 0.0 ? 0 : ilogb(v); };
    int  expSum   = getExp(t11) + getExp(t12) + getExp(t13) + getExp(t22) + getExp(t23) + getExp(t33);
    double norm   = ldexp(1.0, -expSum / 6);
    t11 *= norm, t12 *= norm, t13 *= norm, t22 *= norm, t23 *= norm, t33 *= norm;
    double det    = t11 * t22 * t33 + 2.0 * t12 * t23 * t13 - t11 * t23 * t23 - t22 * t13 * t13 - t33 * t12 * t12;
    double factor = norm * (hi * hi * hi) / (det * K);
    out[0]        = (t22 * t33 - t23 * t23) * factor;
    out[1]        = (t13 * t23 - t33 * t12) * factor;
    out[2]        = (t12 * t23 - t22 * t13) * factor;
    out[3]        = (t11 * t33 - t13 * t13) * factor;
    out[4]        = (t13 * t12 - t11 * t23) * factor;
    out[5]        = (t11 * t22 - t12 * t12) * factor;
}

//! tsKCourant (sph/include/sph/kernels.hpp:11-16); Kcour is passed as float like the reference
template<class T>
__device__ __forceinline__ double courantDt(T maxvs, T h, T c, float Kcour)
{
    T v = maxvs > T(0) ? maxvs : c;
    return double(T(Kcour * h / v));
}

//! artificial_viscosity (kernels.hpp:71-84), beta = 2
template<class T>
__device__ __forceinline__ T artificialViscosity(T alpha_i, T alpha_j, T c_i, T c_j, T w_ij)
{
    constexpr T beta = 2.0;
    T           visc = 0.0;
    if (w_ij < T(0))
    {
        T vsig = (alpha_i + alpha_j) * T(0.25) * (c_i + c_j) - beta * w_ij;
        visc   = -vsig * w_ij;
    }
    return visc;
}

//! arguments shared by all pair kernels
struct PairArgs
{
    const double* x;
    const double* y;
    const double* z;
    DevBox        box;
    double        K;
    Lists         L;
    uint32_t      numParticles; // all particles that can appear as neighbours
};

inline PairArgs makePairArgs(sphx_workspace* ws, const DevBox& box, const sphx_fields_t* f, const sphx_params_t* p,
                             uint32_t first, uint32_t last)
{
    SPHX_REQUIRE(ws->nbValid, SPHX_ERR_STATE, "pair loop called before find_neighbors");
    SPHX_REQUIRE(ws->nbFirst <= first && first <= last && last <= ws->nbLast, SPHX_ERR_ARG,
                 "target range must stay inside the searched range");
    SPHX_REQUIRE(p->ngmax == ws->nbNgmax, SPHX_ERR_ARG, "ngmax differs from the one used by find_neighbors");
    SPHX_REQUIRE(ws->tablesValid, SPHX_ERR_STATE, "kernel tables not set (sphx_workspace_set_kernel)");
    SPHX_REQUIRE(f->x && f->y && f->z && f->nc, SPHX_ERR_ARG, "x, y, z and nc are required");
    PairArgs c;
    c.x = f->x, c.y = f->y, c.z = f->z, c.box = box, c.K = p->K;
    c.L            = Lists{ws->nbrs.get(), f->nc, ws->nbFirst, ws->nbLast, p->ngmax, paddedNgmax(p->ngmax),
                               groupsFor(ws, first, last)};
    c.numParticles = uint32_t(ws->tree.numParticles);
    return c;
}

//! the kernel function of form T for the device (the workspace's tables, or its sinc index)
template<class T>
inline SincKernel<T> sincKernel(sphx_workspace* ws)
{
    SincKernel<T> k;
    if constexpr (isMixed<T>)
    {
        k.n  = float(ws->sincIndex);
        k.ni = ws->sincIndex == double(int(ws->sincIndex)) && ws->sincIndex < 64 ? int(ws->sincIndex) : -1;
    }
    else
    {
        k.w  = ws->wh64.get();
        k.dw = ws->whd64.get();
        k.wq = ws->whq64.get();
    }
    return k;
}

/*! wave minimum then one atomic per wave; with per-group steps (block time steps), also
 * groupDt[g] = min(groupDt[g], the group's minimum) as a float (hydro_ve/momentum_energy_gpu.cu:96-104) */
template<class T>
__device__ __forceinline__ void reduceMinDt(unsigned long long* dst, double dt, const Groups& G)
{
    double m = waveMin(dt);
    if (threadIdx.x % kWave == 0)
    {
        atomicMinPositive(dst, m);
        const uint32_t g = waveGroup();
        if (G.dt != nullptr && g < G.num) { G.dt[g] = fminf(G.dt[g], float(m)); }
    }
}

} // namespace sphx
