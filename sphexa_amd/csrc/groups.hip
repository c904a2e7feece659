// Spatial target groups: the reference GPU path's computeSpatialGroups / computeGroupSplits
// (sph/include/sph/groups.hpp:42-55, domain/include/cstone/traversal/groups_gpu.cu:34-101,
// groups_gpu.cuh:37-76,149-216) for groups of 64, one wavefront per fixed group.
#include <cstring> // rocprim uses memset without including it
#include <rocprim/rocprim.hpp>

#include "common.hpp"

namespace sphx
{

namespace
{

/* One wavefront per fixed group of 64 consecutive particles (lanes past `last` repeat particle
 * last - 1). distCrit = tolFactor * cbrt(v) with v the smallest unit-box volume of the lanes' leaves:
 * for a leaf of level L that volume is 2^-3L and its cube root 2^-L, formed exactly; distCrit is float
 * (h's type on the reference's path) and squared in double. Lane l flags a split when the distance
 * from its particle to lane l+1's, in unit-box coordinates, exceeds distCrit. */
__global__ void __launch_bounds__(256) groupSplitsKernel(uint32_t first, uint32_t last, const double* __restrict__ x,
                                                         const double* __restrict__ y, const double* __restrict__ z,
                                                         const Key* __restrict__ leaves,
                                                         const uint32_t* __restrict__ layout, int numLeaves,
                                                         double il0, double il1, double il2, float tolFactor,
                                                         uint64_t* __restrict__ splitMask, uint32_t* __restrict__ count,
                                                         uint32_t numFixed)
{
#pragma clang fp contract(off)
    const uint32_t g    = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const unsigned lane = laneId();
    if (g >= numFixed) { return; }
    const uint32_t b = min(first + g * uint32_t(kWave) + lane, last - 1);

    // the leaf holding b: the last leaf whose first particle is <= b
    int lo = 0, hi = numLeaves; // layout[lo] <= b < layout[hi]
    while (hi - lo > 1)
    {
        const int mid = (lo + hi) / 2;
        if (layout[mid] <= b) { lo = mid; }
        else { hi = mid; }
    }
    const Key      range = leaves[lo + 1] - leaves[lo];
    const unsigned level = (unsigned(__clzll((long long)(range - 1))) - 1u) / 3u;
    const unsigned deep  = waveMax(level);

    const float  distCrit = ldexpf(1.0f, -int(deep)) * tolFactor;
    const double dc2      = double(distCrit) * double(distCrit);

    const double px = x[b] * il0, py = y[b] * il1, pz = z[b] * il2;
    const double nx = __shfl_down(px, 1, kWave), ny = __shfl_down(py, 1, kWave), nz = __shfl_down(pz, 1, kWave);
    const double dx = nx - px, dy = ny - py, dz = nz - pz;
    const double d2 = dx * dx + (dy * dy + dz * dz);
    const bool   split = lane + 1 < unsigned(kWave) && d2 > dc2;
    const uint64_t mask = ballot(split);
    if (lane == 0)
    {
        splitMask[g] = mask;
        count[g]     = 1u + uint32_t(__popcll(mask));
    }
}

//! group starts of fixed group g at offset[g]: its first particle, then one past each split lane
__global__ void __launch_bounds__(256) makeGroupsKernel(uint32_t first, const uint64_t* __restrict__ splitMask,
                                                        const uint32_t* __restrict__ offset,
                                                        uint32_t* __restrict__ groups, uint32_t numFixed)
{
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= numFixed) { return; }
    const uint32_t start = first + g * uint32_t(kWave);
    uint32_t       pos   = offset[g];
    uint64_t       m     = splitMask[g];
    groups[pos++]        = start;
    while (m)
    {
        groups[pos++] = start + uint32_t(__builtin_ctzll(m)) + 1u;
        m &= m - 1;
    }
}

__global__ void setLastKernel(uint32_t* p, uint32_t v) { *p = v; }

} // namespace

void computeGroups(sphx_workspace* ws, const DevBox& box, const double* x, const double* y, const double* z,
                   uint32_t first, uint32_t last, float tolFactor, sphx_groups_t* out)
{
    SPHX_REQUIRE(ws->tree.valid, SPHX_ERR_STATE, "compute_groups called before the tree was built");
    SPHX_REQUIRE(x && y && z, SPHX_ERR_ARG, "compute_groups needs x, y and z");
    SPHX_REQUIRE(first <= last && last <= ws->tree.numParticles, SPHX_ERR_ARG, "invalid range");
    SPHX_REQUIRE(tolFactor > 0, SPHX_ERR_ARG, "tolFactor must be positive");
    *out           = sphx_groups_t{};
    out->firstBody = first, out->lastBody = last;
    const uint32_t numFixed = ceilDiv(last - first, uint32_t(kWave));
    if (numFixed == 0)
    {
        ws->groupBuf.reserve(1);
        setLastKernel<<<1, 1, 0, ws->stream>>>(ws->groupBuf.get(), last);
        SPHX_CHECK_LAUNCH();
        out->groupStart = out->groupEnd = ws->groupBuf.get();
        return;
    }
    ws->groupSplit.reserve(numFixed);
    ws->groupCount.reserve(2 * size_t(numFixed) + 2);
    uint32_t* count  = ws->groupCount.get();
    uint32_t* offset = count + numFixed + 1;
    const unsigned threads = 256;
    groupSplitsKernel<<<ceilDiv(size_t(numFixed) * kWave, threads), threads, 0, ws->stream>>>(
        first, last, x, y, z, ws->tree.leaves.get(), ws->tree.layout.get(), ws->tree.numLeafNodes, box.ilen[0],
        box.ilen[1], box.ilen[2], tolFactor, ws->groupSplit.get(), count, numFixed);
    SPHX_CHECK_LAUNCH();
    setLastKernel<<<1, 1, 0, ws->stream>>>(count + numFixed, 0u);
    SPHX_CHECK_LAUNCH();
    size_t bytes = 0;
    SPHX_HIP(rocprim::exclusive_scan(nullptr, bytes, count, offset, 0u, size_t(numFixed) + 1, rocprim::plus<uint32_t>(),
                                     ws->stream));
    void* tmp = scratchBytes(ws, bytes);
    SPHX_HIP(rocprim::exclusive_scan(tmp, bytes, count, offset, 0u, size_t(numFixed) + 1, rocprim::plus<uint32_t>(),
                                     ws->stream));
    uint32_t numGroups = 0;
    SPHX_HIP(hipMemcpyAsync(&numGroups, offset + numFixed, sizeof(numGroups), hipMemcpyDeviceToHost, ws->stream));
    SPHX_HIP(hipStreamSynchronize(ws->stream));
    ws->groupBuf.reserve(size_t(numGroups) + 1);
    makeGroupsKernel<<<ceilDiv(numFixed, threads), threads, 0, ws->stream>>>(first, ws->groupSplit.get(), offset,
                                                                            ws->groupBuf.get(), numFixed);
    SPHX_CHECK_LAUNCH();
    setLastKernel<<<1, 1, 0, ws->stream>>>(ws->groupBuf.get() + numGroups, last);
    SPHX_CHECK_LAUNCH();
    out->numGroups  = numGroups;
    out->groupStart = ws->groupBuf.get();
    out->groupEnd   = ws->groupBuf.get() + 1;
}

} // namespace sphx
