// Hilbert keys, stable key sort and field permutation (the one-rank Domain::sync front end).
#include <algorithm>
#include <vector>
#include <cstring> // rocprim uses memset without including it
#include <rocprim/rocprim.hpp>

#include "common.hpp"

namespace sphx
{

/* Hilbert key of one particle. Integer coordinates follow sfc3D (domain/include/cstone/sfc/
 * sfc.hpp:142-178): i = int(floor(x*m) - xmin*m), clamped to 2^21-1, with m = 2^21/L, evaluated
 * without contraction so the keys equal the CPU's. The curve is walked from the top level: the
 * Morton octant o = 4x+2y+z of each level maps to the Hilbert digit o ^ (o >> 1) ^ (o >> 2)
 * (prefix parity: 0,1,3,2,7,6,4,5), after which the lower coordinate bits are reflected and
 * permuted into the child's frame. */
__device__ __forceinline__ Key hilbertKeyOf(double x, double y, double z, const DevBox& b)
{
#pragma clang fp contract(off)
    const double cube = double(1u << kMaxLevel);
    double       mx = cube * b.ilen[0], my = cube * b.ilen[1], mz = cube * b.ilen[2];
    int          ix = int(floor(x * mx) - b.lo[0] * mx);
    int          iy = int(floor(y * my) - b.lo[1] * my);
    int          iz = int(floor(z * mz) - b.lo[2] * mz);
    const int    mc = (1 << kMaxLevel) - 1;
    unsigned     px = unsigned(min(ix, mc)), py = unsigned(min(iy, mc)), pz = unsigned(min(iz, mc));

    Key key = 0;
#pragma unroll
    for (int level = int(kMaxLevel) - 1; level >= 0; --level)
    {
        unsigned bx = (px >> level) & 1u, by = (py >> level) & 1u, bz = (pz >> level) & 1u;
        unsigned o  = (bx << 2) | (by << 1) | bz;
        key         = (key << 3) | Key(o ^ (o >> 1) ^ (o >> 2));

        unsigned nby = by ^ 1u, nbz = bz ^ 1u;
        px ^= 0u - (bx & (nby | bz));
        py ^= 0u - ((bx & (by | bz)) | (by & nbz));
        pz ^= 0u - ((bx & nby & nbz) | (by & nbz));
        unsigned tx = px, ty = py, tz = pz;
        // bz: rotate (x,y,z) <- (y,z,x); else !by: exchange x and z
        px = bz ? ty : (nby ? tz : tx);
        py = bz ? tz : ty;
        pz = bz ? tx : (nby ? tx : tz);
    }
    return key;
}

//! keys of n particles; with KeepRemoved, keys already equal to removeKey = 2^63 stay
template<bool KeepRemoved>
__global__ void __launch_bounds__(256) keysKernel(DevBox box, const double* __restrict__ x,
                                                  const double* __restrict__ y, const double* __restrict__ z,
                                                  Key* __restrict__ keys, size_t n)
{
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) { return; }
    constexpr Key removeKey = Key(1) << 63;
    if (!KeepRemoved || keys[i] != removeKey) { keys[i] = hilbertKeyOf(x[i], y[i], z[i], box); }
}

void computeKeys(sphx_workspace* ws, const DevBox& box, const double* x, const double* y, const double* z, Key* keys,
                 size_t n, bool keepRemoved)
{
    if (n == 0) { return; }
    if (keepRemoved) { keysKernel<true><<<ceilDiv(n, 256), 256, 0, ws->stream>>>(box, x, y, z, keys, n); }
    else { keysKernel<false><<<ceilDiv(n, 256), 256, 0, ws->stream>>>(box, x, y, z, keys, n); }
    SPHX_CHECK_LAUNCH();
}

__global__ void fill64Kernel(uint64_t* p, uint64_t v, size_t n)
{
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) { p[i] = v; }
}

//! device fill (a kernel rather than hipMemsetAsync, so callers may capture it in a graph)
void fill64(sphx_workspace* ws, uint64_t* p, uint64_t v, size_t n)
{
    if (n == 0) { return; }
    fill64Kernel<<<ceilDiv(n, 256), 256, 0, ws->stream>>>(p, v, n);
    SPHX_CHECK_LAUNCH();
}

//! Stable LSD radix sort of (key, original index) over all 64 key bits (removeKey = 2^63 sorts
//! last), i.e. the order of std::stable_sort on the keys.
void sortKeys(sphx_workspace* ws, Key* keys, uint32_t* order, size_t n)
{
    if (n == 0) { return; }
    ws->keyTmp.reserve(n);
    // the original indices come from a counting iterator: no index array is written or read
    const rocprim::counting_iterator<uint32_t> iota(0);

    size_t tmpBytes = 0;
    SPHX_HIP(rocprim::radix_sort_pairs(nullptr, tmpBytes, keys, ws->keyTmp.get(), iota, order, n, 0, 64, ws->stream));
    void* tmp = scratchBytes(ws, tmpBytes);
    SPHX_HIP(rocprim::radix_sort_pairs(tmp, tmpBytes, keys, ws->keyTmp.get(), iota, order, n, 0, 64, ws->stream));
    SPHX_HIP(hipMemcpyAsync(keys, ws->keyTmp.get(), n * sizeof(Key), hipMemcpyDeviceToDevice, ws->stream));
}

} // namespace sphx

/* Fixed scratch of the allocation-free sort (sortKeysNoAlloc): the sorted-key buffer, then the
 * radix sort's temporaries. Every GPU the library runs on holds its own copy for the life of the
 * process. In host code `scratch` names a host-side stand-in, not device memory: its device
 * address comes from hipGetSymbolAddress, once per device (noAllocScratch). */
static __device__ __align__(256) char scratch[sphx::kNoAllocScratchBytes];

namespace sphx
{

namespace
{
std::mutex g_scratchMutex;
void*      g_scratchPtr[kMaxDevices] = {}; // device address of `scratch` per device, once looked up
} // namespace

//! the device address of `scratch` on `device` (the current device), looked up on first use
void* noAllocScratch(int device)
{
    SPHX_REQUIRE(device >= 0 && device < kMaxDevices, SPHX_ERR_ARG, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratchMutex);
    if (g_scratchPtr[device] == nullptr)
    {
        void* ptr = nullptr;
        SPHX_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        SPHX_REQUIRE(ptr != nullptr, SPHX_ERR_HIP, "no device address for the sort scratch");
        g_scratchPtr[device] = ptr;
    }
    return g_scratchPtr[device];
}

/* sortKeys with the fixed scratch instead of workspace buffers: no device allocation. The sorted keys
 * go to the scratch's first n*8 bytes (rounded up to 256), the radix sort's temporaries after them,
 * and the keys are copied back, as sortKeys does. n is limited to kNoAllocMaxKeys. */
void sortKeysNoAlloc(int device, hipStream_t stream, Key* keys, uint32_t* order, size_t n)
{
    SPHX_REQUIRE(n <= kNoAllocMaxKeys, SPHX_ERR_ARG,
                 "sort without allocation: " + std::to_string(n) + " keys exceed the maximum of " +
                     std::to_string(kNoAllocMaxKeys));
    SPHX_REQUIRE(n == 0 || (keys && order), SPHX_ERR_ARG, "null keys or order");
    char* base = static_cast<char*>(noAllocScratch(device));
    if (n == 0) { return; }
    const rocprim::counting_iterator<uint32_t> iota(0);
    Key*         keyTmp   = reinterpret_cast<Key*>(base);
    const size_t keyBytes = (n * sizeof(Key) + 255) & ~size_t(255);
    size_t       tmpBytes = 0;
    SPHX_HIP(rocprim::radix_sort_pairs(nullptr, tmpBytes, keys, keyTmp, iota, order, n, 0, 64, stream));
    SPHX_REQUIRE(keyBytes + tmpBytes <= kNoAllocScratchBytes, SPHX_ERR_ARG,
                 "sort without allocation: " + std::to_string(n) + " keys need " +
                     std::to_string(keyBytes + tmpBytes) + " bytes of scratch, more than " +
                     std::to_string(kNoAllocScratchBytes));
    SPHX_HIP(rocprim::radix_sort_pairs(base + keyBytes, tmpBytes, keys, keyTmp, iota, order, n, 0, 64, stream));
    SPHX_HIP(hipMemcpyAsync(keys, keyTmp, n * sizeof(Key), hipMemcpyDeviceToDevice, stream));
}

template<class E>
__global__ void __launch_bounds__(256) gatherKernel(const uint32_t* __restrict__ order, const E* __restrict__ in,
                                                    E* __restrict__ out, size_t n)
{
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) { out[i] = in[order[i]]; }
}

void gather(sphx_workspace* ws, const uint32_t* order, const void* in, void* out, size_t n, int elemBytes)
{
    if (n == 0) { return; }
    if (elemBytes == 8)
        gatherKernel<uint64_t><<<ceilDiv(n, 256), 256, 0, ws->stream>>>(order, (const uint64_t*)in, (uint64_t*)out, n);
    else if (elemBytes == 4)
        gatherKernel<uint32_t><<<ceilDiv(n, 256), 256, 0, ws->stream>>>(order, (const uint32_t*)in, (uint32_t*)out, n);
    else
        throw Error{SPHX_ERR_ARG, "gather: element size must be 4 or 8 bytes"};
    SPHX_CHECK_LAUNCH();
}

//! up to kGatherBatch fields of one element size, permuted or copied in one launch
constexpr int kGatherBatch = 16;

template<class E>
struct FieldBatch
{
    const E* in[kGatherBatch];
    E*       out[kGatherBatch];
    int      count;
};

//! out_f[i] = in_f[order[i]] for every field f of the batch: the permutation is read once
template<class E>
__global__ void __launch_bounds__(256) gatherBatchKernel(const uint32_t* __restrict__ order, FieldBatch<E> b, size_t n)
{
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) { return; }
    const uint32_t o = order[i];
    for (int f = 0; f < b.count; ++f)
        b.out[f][i] = b.in[f][o];
}

template<class E>
__global__ void __launch_bounds__(256) copyBatchKernel(FieldBatch<E> b, size_t n)
{
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) { return; }
    for (int f = 0; f < b.count; ++f)
        b.out[f][i] = b.in[f][i];
}

template<class E>
void permuteBatches(sphx_workspace* ws, const uint32_t* order, const std::vector<void*>& fields, size_t n)
{
    for (size_t f0 = 0; f0 < fields.size(); f0 += kGatherBatch)
    {
        const int     count = int(std::min(fields.size() - f0, size_t(kGatherBatch)));
        FieldBatch<E> g{}, c{};
        g.count = c.count = count;
        E* tmp           = reinterpret_cast<E*>(ws->fieldTmp.get());
        for (int f = 0; f < count; ++f)
        {
            g.in[f]  = static_cast<const E*>(fields[f0 + f]);
            g.out[f] = tmp + size_t(f) * n;
            c.in[f]  = tmp + size_t(f) * n;
            c.out[f] = static_cast<E*>(fields[f0 + f]);
        }
        gatherBatchKernel<E><<<ceilDiv(n, 256), 256, 0, ws->stream>>>(order, g, n);
        SPHX_CHECK_LAUNCH();
        copyBatchKernel<E><<<ceilDiv(n, 256), 256, 0, ws->stream>>>(c, n);
        SPHX_CHECK_LAUNCH();
    }
}

void permuteFields(sphx_workspace* ws, const uint32_t* order, const std::vector<std::pair<void*, int>>& fields,
                   size_t n)
{
    if (n == 0) { return; }
    // an array named twice is permuted once; arrays that overlap otherwise have no defined result
    std::vector<std::pair<void*, int>> sorted(fields);
    std::sort(sorted.begin(), sorted.end());
    std::vector<void*> f8, f4;
    for (size_t k = 0; k < sorted.size(); ++k)
    {
        auto [ptr, eb] = sorted[k];
        if (eb != 8 && eb != 4) { throw Error{SPHX_ERR_ARG, "permute: element size must be 4 or 8 bytes"}; }
        if (k > 0)
        {
            auto [prev, peb] = sorted[k - 1];
            if (ptr == prev && eb == peb) { continue; }
            if (static_cast<char*>(prev) + n * size_t(peb) > static_cast<char*>(ptr))
            {
                throw Error{SPHX_ERR_ARG, "permute: two of the arrays overlap"};
            }
        }
        (eb == 8 ? f8 : f4).push_back(ptr);
    }
    // scratch for the widest batch: up to kGatherBatch fields of one element size
    const size_t w8 = std::min(f8.size(), size_t(kGatherBatch)), w4 = std::min(f4.size(), size_t(kGatherBatch));
    ws->fieldTmp.reserve(n * std::max(8 * w8, 4 * w4)); // bytes
    permuteBatches<uint64_t>(ws, order, f8, n);
    permuteBatches<uint32_t>(ws, order, f4, n);
}

namespace
{
constexpr unsigned kExtentThreads = 256;
constexpr unsigned kExtentBlocks  = 512;

/*! per block: min x, max x, min y, max y, min z, max z over a grid-stride range, each as
 *  std::min / std::max from +-INFINITY would accumulate it (min and max are exact, so the order of
 *  the reduction does not change the result) */
__global__ void __launch_bounds__(kExtentThreads) extentKernel(const double* __restrict__ x,
                                                              const double* __restrict__ y,
                                                              const double* __restrict__ z, size_t n,
                                                              double* partial)
{
    double        lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    const double* c[3]  = {x, y, z};
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
    {
        for (int d = 0; d < 3; ++d)
        {
            const double v = c[d][i];
            lo[d]          = v < lo[d] ? v : lo[d];
            hi[d]          = v > hi[d] ? v : hi[d];
        }
    }
    __shared__ double part[kExtentThreads / kWave][6];
    for (int d = 0; d < 3; ++d)
    {
        lo[d] = waveMin(lo[d]);
        hi[d] = waveMax(hi[d]);
    }
    if (laneId() == 0)
    {
        for (int d = 0; d < 3; ++d)
            part[threadIdx.x / kWave][2 * d] = lo[d], part[threadIdx.x / kWave][2 * d + 1] = hi[d];
    }
    __syncthreads();
    if (threadIdx.x < 6)
    {
        const bool isMax = threadIdx.x % 2;
        double     r     = part[0][threadIdx.x];
        for (unsigned w = 1; w < kExtentThreads / kWave; ++w)
        {
            const double v = part[w][threadIdx.x];
            r              = isMax ? (v > r ? v : r) : (v < r ? v : r);
        }
        partial[size_t(blockIdx.x) * 6 + threadIdx.x] = r;
    }
}

//! one block of 64: the partials of `blocks` blocks into out[6]
__global__ void extentFinalKernel(const double* __restrict__ partial, unsigned blocks, double* out)
{
    if (threadIdx.x >= 6) { return; }
    const bool isMax = threadIdx.x % 2;
    double     r     = isMax ? -INFINITY : INFINITY;
    for (unsigned b = 0; b < blocks; ++b)
    {
        const double v = partial[size_t(b) * 6 + threadIdx.x];
        r              = isMax ? (v > r ? v : r) : (v < r ? v : r);
    }
    out[threadIdx.x] = r;
}
} // namespace

void coordinateExtent(sphx_workspace* ws, const double* x, const double* y, const double* z, size_t n, double ext[6])
{
    const unsigned blocks = std::max(1u, std::min(kExtentBlocks, ceilDiv(n, kExtentThreads)));
    double*        buf    = static_cast<double*>(scratchBytes(ws, (size_t(blocks) + 1) * 6 * sizeof(double)));
    extentKernel<<<blocks, kExtentThreads, 0, ws->stream>>>(x, y, z, n, buf + 6);
    SPHX_CHECK_LAUNCH();
    extentFinalKernel<<<1, 64, 0, ws->stream>>>(buf + 6, blocks, buf);
    SPHX_CHECK_LAUNCH();
    SPHX_HIP(hipMemcpyAsync(ext, buf, 6 * sizeof(double), hipMemcpyDeviceToHost, ws->stream));
    SPHX_HIP(hipStreamSynchronize(ws->stream));
}

} // namespace sphx
