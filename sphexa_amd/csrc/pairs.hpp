// Helpers shared by the SPH pair-loop kernels (std and VE).
//
// Launch shape: one wavefront per group of 64 consecutive targets (the neighbour-search group),
// 4 waves per block. Lane l of wave g owns target first + 64 g + l and walks its own list
// nbrs[(g*ngmax + k)*64 + l], k < min(nc-1, ngmax): one coalesced 256-byte load per slot.
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

constexpr int kPairWaves = 4;

template<class T>
constexpr bool isMixed = std::is_same_v<T, float>;

struct Lists
{
    const uint32_t* nbrs;
    const uint32_t* nc;
    uint32_t        first, last, ngmax;
};

struct Lane
{
    uint32_t        i;     // target index
    bool            valid; // lane holds a target
    unsigned        n;     // neighbours to visit
    const uint32_t* list;  // this lane's first slot; slot k at list[k * 64]
};

__device__ __forceinline__ Lane laneTarget(const Lists& L)
{
    Lane           t;
    const unsigned lane = threadIdx.x % kWave;
    const uint32_t g    = blockIdx.x * kPairWaves + threadIdx.x / kWave;
    t.i                 = L.first + g * kGroup + lane;
    t.valid             = t.i < L.last;
    t.n                 = t.valid ? min(L.nc[t.i] - 1, L.ngmax) : 0u;
    t.list              = L.nbrs + size_t(g) * L.ngmax * kGroup + lane;
    return t;
}

inline unsigned pairBlocks(uint32_t first, uint32_t last)
{
    return ceilDiv(ceilDiv(last - first, kGroup), kPairWaves);
}

//! lt::lookup (sph/include/sph/table_lookup.hpp:10-26): linear interpolation of a 20000-point
//! table on [0,2], zero beyond the support
template<class T>
__device__ __forceinline__ T lookup(const T* __restrict__ table, T v)
{
    constexpr int numIntervals = kTableSize - 1;
    constexpr T   dx           = T(2.0) / numIntervals;
    constexpr T   invDx        = T(1) / dx;
    int           idx          = int(v * invDx);
    bool          inside       = idx < numIntervals;
    int           k            = inside ? max(idx, 0) : 0;
    T             t0 = table[k], t1 = table[k + 1];
    T             derivative = (t1 - t0) * invDx;
    return inside ? t0 + derivative * (v - T(idx) * dx) : T(0);
}

//! legacy minimum image of a separation with cut-off r (sfc/box.hpp:218-238)
template<class T>
__device__ __forceinline__ void foldLegacy(const DevBox& b, T r, T& xx, T& yy, T& zz)
{
    if (b.pbc[0]) { xx = xx > r ? xx - T(b.len[0]) : (xx < -r ? xx + T(b.len[0]) : xx); }
    if (b.pbc[1]) { yy = yy > r ? yy - T(b.len[1]) : (yy < -r ? yy + T(b.len[1]) : yy); }
    if (b.pbc[2]) { zz = zz > r ? zz - T(b.len[2]) : (zz < -r ? zz + T(b.len[2]) : zz); }
}

//! separation x_i - x_j in the pair type: folded in FP64, then rounded (exact for T = double)
template<class T>
__device__ __forceinline__ void separation(const DevBox& b, double hi2, double xi, double yi, double zi, double xj,
                                           double yj, double zj, T& rx, T& ry, T& rz)
{
    double dx = xi - xj, dy = yi - yj, dz = zi - zj;
    foldLegacy(b, hi2, dx, dy, dz);
    rx = T(dx), ry = T(dy), rz = T(dz);
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
    SPHX_REQUIRE(first == ws->nbFirst && last <= ws->nbLast && first <= last, SPHX_ERR_ARG,
                 "target range must start at the searched first index and stay inside the searched range");
    SPHX_REQUIRE(p->ngmax == ws->nbNgmax, SPHX_ERR_ARG, "ngmax differs from the one used by find_neighbors");
    SPHX_REQUIRE(ws->tablesValid, SPHX_ERR_STATE, "kernel tables not set (sphx_workspace_set_kernel)");
    SPHX_REQUIRE(f->x && f->y && f->z && f->nc, SPHX_ERR_ARG, "x, y, z and nc are required");
    PairArgs c;
    c.x = f->x, c.y = f->y, c.z = f->z, c.box = box, c.K = p->K;
    c.L            = Lists{ws->nbrs.get(), f->nc, first, last, p->ngmax};
    c.numParticles = uint32_t(ws->tree.numParticles);
    return c;
}

template<class T>
inline const T* whTable(sphx_workspace* ws)
{
    if constexpr (isMixed<T>) { return ws->wh32.get(); }
    else { return ws->wh64.get(); }
}

template<class T>
inline const T* whdTable(sphx_workspace* ws)
{
    if constexpr (isMixed<T>) { return ws->whd32.get(); }
    else { return ws->whd64.get(); }
}

//! wave minimum then one atomic per wave
template<class T>
__device__ __forceinline__ void reduceMinDt(unsigned long long* dst, double dt)
{
    double m = waveMin(dt);
    if (threadIdx.x % kWave == 0) { atomicMinPositive(dst, m); }
}

} // namespace sphx
