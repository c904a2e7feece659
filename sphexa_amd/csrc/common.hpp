// Shared device helpers, workspace state and launch utilities of the gfx950 hydro path.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "sphexa_amd.h"

namespace sphx
{

using Key           = uint64_t;
using LocalIndex    = uint32_t;
using TreeNodeIndex = int32_t;

constexpr unsigned kMaxLevel  = 21; // 64-bit Hilbert keys, 21 bits per dimension
constexpr int      kTableSize = 20000;
constexpr int      kWave      = 64; // gfx950 wavefront
constexpr int      kGroup     = 64; // targets per neighbour-search group (one wavefront)

/* Neighbour-list layout. Group g of 64 consecutive targets owns ngmax4 = ngmax rounded up to 4
 * slots per lane; slot k of lane l is at
 *     (g * ngmax4 + (k & ~3)) * 64 + 4 l + (k & 3),
 * so each lane's slots come in 16-byte quads and quad q of all 64 lanes is one contiguous
 * kilobyte: the search stores a lane's entries a quad at a time, and a pair loop reads four
 * neighbour indices of all its lanes with one coalesced 16-byte-per-lane load. */
__host__ __device__ inline uint32_t paddedNgmax(uint32_t ngmax) { return (ngmax + 3u) & ~3u; }

__host__ __device__ inline size_t listOffset(uint32_t g, uint32_t ngmax4, uint32_t k, uint32_t lane)
{
    return (size_t(g) * ngmax4 + (k & ~3u)) * kGroup + 4u * lane + (k & 3u);
}

//! slot k relative to a lane's slot 0
__host__ __device__ inline size_t laneSlot(uint32_t k) { return size_t(k & ~3u) * kGroup + (k & 3u); }

/* ---------------------------------------------------------------------------------------------
 * error handling: every entry point returns an sphx_status and records a message
 * -------------------------------------------------------------------------------------------*/
struct Error
{
    int         code;
    std::string msg;
};

#define SPHX_HIP(call)                                                                                                 \
    do                                                                                                                 \
    {                                                                                                                  \
        hipError_t e_ = (call);                                                                                        \
        if (e_ != hipSuccess)                                                                                          \
            throw ::sphx::Error{SPHX_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)};                      \
    } while (0)

#define SPHX_CHECK_LAUNCH() SPHX_HIP(hipGetLastError())

#define SPHX_REQUIRE(cond, code, text)                                                                                 \
    do                                                                                                                 \
    {                                                                                                                  \
        if (!(cond)) throw ::sphx::Error{(code), (text)};                                                              \
    } while (0)

//! growable device buffer owned by the workspace
template<class T>
struct DevBuf
{
    T*     ptr = nullptr;
    size_t cap = 0;

    T* get() const { return ptr; }
    void reserve(size_t n)
    {
        if (n <= cap) { return; }
        release();
        size_t bytes = (n == 0 ? 1 : n) * sizeof(T);
        if (hipMalloc(&ptr, bytes) != hipSuccess)
        {
            ptr = nullptr;
            cap = 0;
            (void)hipGetLastError();
            throw Error{SPHX_ERR_ALLOC, "hipMalloc of " + std::to_string(bytes) + " bytes failed"};
        }
        cap = n;
    }
    void release()
    {
        if (ptr) { (void)hipFree(ptr); }
        ptr = nullptr;
        cap = 0;
    }
};

//! per-build temporaries of the top-down tree construction, kept for reuse
struct TreeBuildScratch
{
    DevBuf<Key>      nodeKey, leafKey, leafKeySorted;
    DevBuf<uint32_t> lo, hi, flag, rank, leafFlag, leafRank, leafNode, leafNodeSorted;
    DevBuf<uint8_t>  level;
    DevBuf<int32_t>  levelStart;
    // the one-step update of a later sync: node counts, node and leaf ops, the new leaves
    DevBuf<uint32_t> nodeCnt, changed;
    DevBuf<int32_t>  nodeOp, leafOp;
    DevBuf<Key>      nextLeaves;

    void releaseAll()
    {
        nodeKey.release(), leafKey.release(), leafKeySorted.release(), lo.release(), hi.release();
        flag.release(), rank.release(), leafFlag.release(), leafRank.release(), leafNode.release();
        leafNodeSorted.release(), level.release(), levelStart.release();
        nodeCnt.release(), changed.release(), nodeOp.release(), leafOp.release(), nextLeaves.release();
    }
};

//! one interpolation interval of a kernel table, values at both ends, fetched as one vector load
template<class T>
struct alignas(2 * sizeof(T)) TabPair
{
    T t0, t1;
};

//! W and dW/dv over one interval, for loops that evaluate both at the same argument
template<class T>
struct alignas(4 * sizeof(T)) TabQuad
{
    T w0, w1, d0, d1;
};

//! 16-byte storage unit of the neighbour-record buffer
struct alignas(16) Chunk16
{
    uint64_t lo, hi;
};

struct Tree
{
    int32_t              numLeafNodes = 0, numInternalNodes = 0, numNodes = 0;
    size_t               numParticles = 0;
    DevBuf<Key>          prefixes, leaves;
    DevBuf<TreeNodeIndex> childOffsets, parents, internalToLeaf, levelRange;
    DevBuf<uint32_t>     counts, layout;
    DevBuf<double>       centers, sizes;
    // the neighbour search's visiting order (singleTraversal, traversal/traversal.hpp:52-93):
    // leaf children in octant order, then internal children in reverse octant order, depth first.
    // rankLo: position of a node's first leaf in that order; childPush: an internal node's
    // children from last to first visited, 3 bits per octant; subtreeLeaves: scratch
    DevBuf<int32_t>      rankLo, subtreeLeaves;
    DevBuf<uint32_t>     childPush;
    bool                 valid = false;
    TreeBuildScratch     build;

    void releaseAll()
    {
        prefixes.release(), leaves.release(), childOffsets.release(), parents.release();
        internalToLeaf.release(), levelRange.release(), counts.release(), layout.release();
        centers.release(), sizes.release(), rankLo.release(), subtreeLeaves.release(), childPush.release();
        build.releaseAll();
        valid = false;
    }
};

/*! target groups of a call over [first, last) (cstone::GroupView): consecutive groups of 64 from
 * `first` when start is null, else [start[g], min(end[g], start[g] + 64)) for g < num */
struct Groups
{
    const uint32_t* start = nullptr;
    const uint32_t* end   = nullptr;
    float*          dt    = nullptr; // optional per-group Courant step (min-updated)
    uint32_t        num   = 0;
    uint32_t        first = 0, last = 0;
};

} // namespace sphx

struct sphx_workspace
{
    int         device = 0;
    hipStream_t stream = nullptr;
    std::string lastError;

    // kernel tables: W and dW/dv on [0,2], 20000 points, in both precisions, stored per interval
    double                              K = 0, sincIndex = 0;
    // the reference's W and dW/dv tables for the FP64 form (the FP32 form evaluates W directly)
    sphx::DevBuf<sphx::TabPair<double>> wh64, whd64;
    sphx::DevBuf<sphx::TabQuad<double>> whq64;
    bool                   tablesValid = false;

    sphx::Tree tree;

    // neighbour lists of the last search: group-interleaved [group][k][lane]
    sphx::DevBuf<uint32_t> nbrs;
    uint32_t               nbFirst = 0, nbLast = 0, nbNgmax = 0;
    const uint32_t*        nbNc = nullptr; // nc array the lists were built with
    bool                   nbValid = false;
    bool                   nbWide  = false; // some target's 2h reaches half a box length (FP32 loops: exact separations)
    double                 nbHMax  = 0; // largest h of the last search's targets (the FP32 loops' fixed-point period)

    // target groups set by sphx_set_groups, and the spatial groups of sphx_compute_groups
    bool                   groupsSet = false;
    sphx_groups_t          groupView{};
    sphx::DevBuf<uint32_t> groupBuf;   // group starts then last
    sphx::DevBuf<uint32_t> groupCount; // per fixed group: sub-groups, then their exclusive scan
    sphx::DevBuf<uint64_t> groupSplit; // per fixed group: split mask
    uint64_t               statSumPairs = 0;
    uint32_t               statMaxCount = 0;
    uint64_t               statTests    = 0; // candidate distance tests of the last search

    // scratch
    sphx::DevBuf<char>     scratch;      // sort / scan temporaries
    sphx::DevBuf<uint64_t> keyTmp;
    sphx::DevBuf<char>     fieldTmp;     // permutation staging
    sphx::DevBuf<uint32_t> order;        // sort permutation of the domain sync
    size_t                 syncN = 0;    // particles of that sync (0: none yet)
    bool                   syncTree   = false; // the workspace tree is that sync's (its counts are of its keys)
    uint32_t               syncBucket = 0;     // its bucket size
    sphx::DevBuf<sphx::Chunk16> records; // packed neighbour records of the current pair loop
    sphx::DevBuf<sphx::Chunk16> records2; // records a pair loop's epilogue writes for the next loop

    // record chaining (SPHX_FLAG_CHAIN): every entry point that may write particle fields or the
    // neighbour lists increments callSerial; a pair loop that wrote the next loop's records in its
    // epilogue leaves this stamp, and the next call uses those records only if it is the very
    // next such call, on the same particles, fields, box and form
    uint64_t callSerial = 0;
    struct RecordChain
    {
        int           kind   = 0; // record type written (0: none)
        uint64_t      serial = 0; // callSerial of the producing call
        uint32_t      n      = 0;
        int           form   = -1;
        int           buf    = 0; // record buffer holding them
        sphx_fields_t fields{};
        double        lo[3]{}, hi[3]{};
        int           pbc[3]{};
    } chain;
    // the last EOS call (its callSerial, fields and kind: 1 std, 2 VE). The std EOS writes p and
    // c, the VE EOS p, c, rho and p/rho: the IAD records that density / VeDefGradh leave survive
    // an EOS of their own formulation
    uint64_t      eosSerial = 0;
    sphx_fields_t eosFields{};
    int           eosKind = 0;
    sphx::DevBuf<uint32_t> counters;     // small device counters / flags: kCounterSlots 64-bit slots
    sphx::DevBuf<uint32_t> hostVisible;  // (unused placeholder for future pinned results)
    sphx::DevBuf<double>   reduce64;     // per-block partial reductions

    // the adjoints' pair index of the last search (built by the first adjoint call after it): target i's
    // stored pairs are p = pairStart[i - nbFirst] + k; sorted holds (j << pbits | p) in increasing order, so
    // the pairs that have particle j as neighbour are sorted[rowStart[j] .. rowStart[j + 1]); pairBuf holds
    // the neighbour-side contributions of each pair at its place in sorted (recordOf[p]), so each
    // neighbour sums a contiguous row
    uint64_t searchSerial = 0; // incremented by every search
    struct AdjointPairs
    {
        uint64_t               serial = 0; // searchSerial the index was built for (0: none)
        uint64_t               pairs  = 0;
        int                    pbits  = 0;
        sphx::DevBuf<uint32_t> pairStart, rowStart, recordOf; // recordOf[p]: the pair's place in sorted
        sphx::DevBuf<uint64_t> keys, sorted;
        sphx::DevBuf<char>     pairBuf;
    } adj;

    // optional per-kernel timing: HIP events around the main kernel of each pass
    struct KernelEvents
    {
        hipEvent_t start = nullptr, stop = nullptr;
        bool       recorded = false;
    };
    bool                                profile = false;
    std::map<std::string, KernelEvents> kernelEvents;
};

namespace sphx
{
//! records HIP events on the workspace stream around the enclosed launch when profiling is on
class KernelScope
{
public:
    KernelScope(sphx_workspace* ws, const char* name)
        : ws_(ws)
    {
        if (!ws->profile) { return; }
        ev_ = &ws->kernelEvents[name];
        if (!ev_->start)
        {
            SPHX_HIP(hipEventCreate(&ev_->start));
            SPHX_HIP(hipEventCreate(&ev_->stop));
        }
        SPHX_HIP(hipEventRecord(ev_->start, ws->stream));
    }
    ~KernelScope()
    {
        if (ev_ && hipEventRecord(ev_->stop, ws_->stream) == hipSuccess) { ev_->recorded = true; }
    }
    KernelScope(const KernelScope&)            = delete;
    KernelScope& operator=(const KernelScope&) = delete;

private:
    sphx_workspace*                ws_;
    sphx_workspace::KernelEvents* ev_ = nullptr;
};
} // namespace sphx

namespace sphx
{

//! scratch of at least `bytes`, 256-byte aligned
inline void* scratchBytes(sphx_workspace* ws, size_t bytes)
{
    ws->scratch.reserve(bytes);
    return ws->scratch.get();
}

inline unsigned ceilDiv(size_t a, unsigned b) { return unsigned((a + b - 1) / b); }

//! the target groups of a call over [first, last): the view of sphx_set_groups when it covers exactly
//! that range, else consecutive groups of 64 from first
inline Groups groupsFor(const sphx_workspace* ws, uint32_t first, uint32_t last)
{
    Groups g;
    g.first = first, g.last = last;
    if (ws->groupsSet && ws->groupView.firstBody == first && ws->groupView.lastBody == last)
    {
        g.start = ws->groupView.groupStart, g.end = ws->groupView.groupEnd, g.dt = ws->groupView.groupDt;
        g.num   = ws->groupView.numGroups;
    }
    else { g.num = ceilDiv(last - first, uint32_t(kGroup)); }
    return g;
}

/*! the particle range [i0, i1) of group g of a call (device): clipped to the call's range and to 64
 * particles; sets `oversized` when the caller's group holds more than 64 */
__device__ __forceinline__ void groupRange(const Groups& G, uint32_t g, uint32_t& i0, uint32_t& i1, bool& oversized);

//! threads an elementwise kernel over the call's targets launches: one per target, or 64 per group
inline size_t groupedThreads(const Groups& G) { return G.start ? size_t(G.num) * kGroup : size_t(G.last - G.first); }

//! particle of this thread of an elementwise kernel over the call's targets (false: none)
__device__ __forceinline__ bool groupedIndex(const Groups& G, uint32_t& i)
{
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (G.start == nullptr)
    {
        i = G.first + tid;
        return i < G.last;
    }
    const uint32_t g = tid / uint32_t(kGroup);
    if (g >= G.num) { return false; }
    uint32_t i0, i1;
    bool     oversized;
    groupRange(G, g, i0, i1, oversized);
    i = i0 + tid % uint32_t(kGroup);
    return i < i1;
}

__device__ __forceinline__ void groupRange(const Groups& G, uint32_t g, uint32_t& i0, uint32_t& i1, bool& oversized)
{
    if (G.start == nullptr)
    {
        i0        = G.first + g * uint32_t(kGroup);
        i1        = min(i0 + uint32_t(kGroup), G.last);
        oversized = false;
        return;
    }
    const uint32_t s = G.start[g], e = G.end[g];
    oversized        = e > s + uint32_t(kGroup);
    i0               = max(s, G.first);
    i1               = min(min(e, s + uint32_t(kGroup)), G.last);
    if (i1 < i0) { i1 = i0; }
}

/* 64-bit slots of the workspace counters: the search's statistics (stored pairs, maximum count,
 * non-converged targets) in 0-2, the Courant dt of the momentum loops in 3, the device maxima of the
 * time step in 4, the search's wide-support count, candidate tests and oversized groups in 5-7, the
 * largest h of the searched targets (FP64 bits) in 8 */
constexpr int kCounterSlots  = 9;
constexpr int kStatWide      = 5;
constexpr int kStatTests     = 6;
constexpr int kStatBadGroups = 7;
constexpr int kStatHMax      = 8;

constexpr unsigned kNumXcd = 8; // MI355X: 8 accelerator complex dies, one L2 each

//! The hardware deals the blocks of a launch round-robin over the XCDs (block b runs on XCD b % 8).
//! Renumber them so that each XCD works through one contiguous run of logical blocks: groups that are
//! adjacent in SFC order, and so share most of their neighbours, then run under the same L2.
//! A bijection of [0, nb) for every nb.
__host__ __device__ inline uint32_t xcdBlock(uint32_t b, uint32_t nb)
{
    uint32_t q = nb / kNumXcd, r = nb % kNumXcd;
    uint32_t xcd = b % kNumXcd, idx = b / kNumXcd;
    return xcd < r ? xcd * (q + 1) + idx : r * (q + 1) + (xcd - r) * q + idx;
}

/* ---------------------------------------------------------------------------------------------
 * bit helpers (placeholder-bit keys, cf. domain/include/cstone/sfc/common.hpp)
 * -------------------------------------------------------------------------------------------*/
__host__ __device__ inline Key nodeRange(unsigned level) { return Key(1) << (3u * (kMaxLevel - level)); }

__host__ __device__ inline Key encodePlaceholder(Key key, int prefixLength)
{
    return (Key(1) << prefixLength) | (key >> (3 * int(kMaxLevel) - prefixLength));
}

//! length in bits of the prefix of a placeholder-bit key (decodePrefixLength, sfc/common.hpp)
__host__ __device__ inline unsigned decodePrefixLength(Key code) { return 63u - unsigned(__builtin_clzll(code)); }

//! first SFC key of the node of a placeholder-bit key (decodePlaceholderBit, sfc/common.hpp)
__host__ __device__ inline Key decodePlaceholder(Key code)
{
    const unsigned len = decodePrefixLength(code);
    return (code ^ (Key(1) << len)) << (3u * kMaxLevel - len);
}

//! octree level of a node spanning `range` keys (treeLevel, sfc/common.hpp)
__host__ __device__ inline unsigned treeLevel(Key range) { return (unsigned(__builtin_clzll(range - 1)) - 1u) / 3u; }

//! Hilbert key -> integer coordinates (same curve as domain/include/cstone/sfc/hilbert.hpp:130-172)
__host__ __device__ inline void hilbertDecode(Key key, unsigned& ox, unsigned& oy, unsigned& oz)
{
    unsigned px = 0, py = 0, pz = 0;
    for (unsigned level = 0; level < kMaxLevel; ++level)
    {
        unsigned d  = unsigned(key >> (3 * level)) & 7u;
        unsigned bx = d >> 2, by = (d >> 1) & 1u, bz = d & 1u;
        if (by ^ bz)
        {
            unsigned t = px;
            px = pz, pz = py, py = t;
        }
        else if (bx == by && by == bz)
        {
            unsigned t = px;
            px = pz, pz = t;
        }
        unsigned mask = (1u << level) - 1;
        px ^= mask & (0u - (bx & (by | bz)));
        py ^= mask & (0u - ((bx & ((by ^ 1u) | (bz ^ 1u))) | ((bx ^ 1u) & by & bz)));
        pz ^= mask & (0u - ((bx & (by ^ 1u) & (bz ^ 1u)) | (by & bz)));
        px |= bx << level;
        py |= (bx ^ by) << level;
        pz |= (by ^ bz) << level;
    }
    ox = px, oy = py, oz = pz;
}

/* ---------------------------------------------------------------------------------------------
 * wave-level primitives (wave64)
 * -------------------------------------------------------------------------------------------*/
__device__ inline unsigned laneId() { return __lane_id(); }

__device__ inline uint64_t ballot(bool p) { return __ballot(p); }

//! number of set bits of `mask` below this lane
__device__ inline unsigned laneRank(uint64_t mask)
{
    return __builtin_amdgcn_mbcnt_hi(unsigned(mask >> 32), __builtin_amdgcn_mbcnt_lo(unsigned(mask), 0u));
}

template<class T>
__device__ inline T waveMin(T v)
{
    for (int o = 32; o > 0; o >>= 1)
    {
        T w = __shfl_xor(v, o, 64);
        v   = w < v ? w : v;
    }
    return v;
}

template<class T>
__device__ inline T waveMax(T v)
{
    for (int o = 32; o > 0; o >>= 1)
    {
        T w = __shfl_xor(v, o, 64);
        v   = w > v ? w : v;
    }
    return v;
}

template<class T>
__device__ inline T waveSum(T v)
{
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o, 64);
    return v;
}

//! read a value held by lane `src` (uniform result)
template<class T>
__device__ inline T readLane(T v, int src)
{
    return __shfl(v, src, 64);
}

/* ---------------------------------------------------------------------------------------------
 * box in device-friendly form
 * -------------------------------------------------------------------------------------------*/
struct DevBox
{
    double lo[3], hi[3], len[3], ilen[3];
    int    pbc[3];
    int    anyPbc;
    // 32-bit fixed-point coordinates of the FP32 pair loops over the period P = len / 2^s (s the
    // fixed-point shift of the last search, 0 by default): u = rint((x - lo) 2^32 / P) mod 2^32
    double fixScale[3]; // 2^32 / P
    double fixStepD[3]; // P / 2^32
    float  fixStep[3];  // P / 2^32, rounded to float
};

inline DevBox makeDevBox(const sphx_box_t* b)
{
    DevBox d;
    d.anyPbc = 0;
    for (int k = 0; k < 3; ++k)
    {
        d.lo[k]   = b->lo[k];
        d.hi[k]   = b->hi[k];
        d.len[k]  = b->hi[k] - b->lo[k];
        d.ilen[k] = 1.0 / (b->hi[k] - b->lo[k]);
        d.pbc[k]  = b->boundary[k] == 1;
        d.anyPbc |= d.pbc[k];
        d.fixScale[k] = 4294967296.0 / d.len[k];
        d.fixStepD[k] = d.len[k] / 4294967296.0;
        d.fixStep[k]  = float(d.fixStepD[k]);
    }
    return d;
}

//! the box of the pair loops; in the FP32 form its fixed point runs over the period len / 2^shift[k]
//! per dimension (fixedPointShift)
inline DevBox makePairBox(const sphx_box_t* b, const int shift[3])
{
    DevBox d = makeDevBox(b);
    for (int k = 0; k < 3; ++k)
    {
        const double p = std::ldexp(d.len[k], -shift[k]);
        d.fixScale[k]  = 4294967296.0 / p;
        d.fixStepD[k]  = p / 4294967296.0;
        d.fixStep[k]   = float(d.fixStepD[k]);
    }
    return d;
}

//! host: the fixed-point shift per dimension for a search whose largest h is hMax (see Coord)
inline void fixedPointShift(const sphx_box_t* b, double hMax, bool wide, int shift[3])
{
    for (int d = 0; d < 3; ++d)
    {
        const double len = b->hi[d] - b->lo[d];
        int          s   = 0;
        while (!wide && s < 16 && std::ldexp(len, -(s + 1)) >= 8.0 * hMax) { ++s; }
        shift[d] = s;
    }
}

//! host: the box of the FP32 pair loops over the last search's lists (FP64 form: the plain box)
inline DevBox pairBox(const sphx_workspace* ws, const sphx_box_t* b)
{
    int shift[3];
    fixedPointShift(b, ws->nbHMax, ws->nbWide, shift);
    return makePairBox(b, shift);
}

/* ---------------------------------------------------------------------------------------------
 * host launchers implemented in the .hip translation units
 * -------------------------------------------------------------------------------------------*/
void fill64(sphx_workspace* ws, uint64_t* p, uint64_t v, size_t n);
//! Hilbert keys; keepRemoved: keys equal to removeKey (2^63) on entry stay
void computeKeys(sphx_workspace* ws, const DevBox& box, const double* x, const double* y, const double* z, Key* keys,
                 size_t n, bool keepRemoved);
void sortKeys(sphx_workspace* ws, Key* keys, uint32_t* order, size_t n);
//! the allocation-free sort: fixed device scratch of kNoAllocScratchBytes per GPU, at most kNoAllocMaxKeys keys
constexpr size_t kNoAllocScratchBytes = size_t(32) << 20;
constexpr size_t kNoAllocMaxKeys      = size_t(1) << 20;
static_assert(kNoAllocMaxKeys == SPHX_SORT_NOALLOC_MAX_KEYS, "the C header states the same limit");
constexpr int    kMaxDevices          = 64;
void sortKeysNoAlloc(int device, hipStream_t stream, Key* keys, uint32_t* order, size_t n);
//! spatial target groups of [first, last) into the workspace (computeGroupSplits)
void computeGroups(sphx_workspace* ws, const DevBox& box, const double* x, const double* y, const double* z,
                   uint32_t first, uint32_t last, float tolFactor, sphx_groups_t* out);
void gather(sphx_workspace* ws, const uint32_t* order, const void* in, void* out, size_t n, int elemBytes);
//! permute every field in place by `order` (4- or 8-byte elements), batched per element size
//! min x, max x, min y, max y, min z, max z of n coordinates (+-INFINITY for n = 0), on the host
void coordinateExtent(sphx_workspace* ws, const double* x, const double* y, const double* z, size_t n, double ext[6]);
void permuteFields(sphx_workspace* ws, const uint32_t* order, const std::vector<std::pair<void*, int>>& fields,
                   size_t n);
void buildTree(sphx_workspace* ws, const Key* sortedKeys, size_t n, unsigned bucket, const DevBox& box);
//! one rebalance of the workspace tree on its counts, then counts of the new sorted keys (the tree update
//! of a later Domain::sync); returns whether no node changed
bool updateTree(sphx_workspace* ws, const Key* sortedKeys, size_t n, unsigned bucket, const DevBox& box);

//! what findNeighbors computes besides the lists: nothing, or the fused XMass sum as xm or as rho
constexpr int kSearchOnly = 0, kSearchXm = 1, kSearchRho = 2;

int findNeighbors(sphx_workspace* ws, const DevBox& box, const sphx_fields_t* f, const sphx_params_t* p,
                  uint32_t first, uint32_t last, bool iterateH, int xmassOut = kSearchOnly);
void copyNeighbors(sphx_workspace* ws, uint32_t* out, uint32_t first, uint32_t last);

void densityStd(sphx_workspace* ws, const DevBox& box, const sphx_fields_t* f, const sphx_params_t* p,
                uint32_t first, uint32_t last);
void eosStd(sphx_workspace* ws, const sphx_fields_t* f, const sphx_params_t* p, uint32_t first, uint32_t last);
void iadStd(sphx_workspace* ws, const DevBox& box, const sphx_fields_t* f, const sphx_params_t* p, uint32_t first,
            uint32_t last);
double momentumEnergyStd(sphx_workspace* ws, const DevBox& box, const sphx_fields_t* f, const sphx_params_t* p,
                         uint32_t first, uint32_t last);

void densityStdVjp(sphx_workspace* ws, const DevBox& box, const sphx_fields_t* f, const sphx_params_t* p,
                   uint32_t first, uint32_t last, const sphx_adjoint_t* in, sphx_adjoint_t* out);
void iadStdVjp(sphx_workspace* ws, const DevBox& box, const sphx_fields_t* f, const sphx_params_t* p, uint32_t first,
               uint32_t last, const sphx_adjoint_t* in, sphx_adjoint_t* out);
void momentumEnergyStdVjp(sphx_workspace* ws, const DevBox& box, const sphx_fields_t* f, const sphx_params_t* p,
                          uint32_t first, uint32_t last, const sphx_adjoint_t* in, sphx_adjoint_t* out);

void xmassVe(sphx_workspace* ws, const DevBox& box, const sphx_fields_t* f, const sphx_params_t* p, uint32_t first,
             uint32_t last);
void veDefGradh(sphx_workspace* ws, const DevBox& box, const sphx_fields_t* f, const sphx_params_t* p,
                uint32_t first, uint32_t last);
void eosVe(sphx_workspace* ws, const sphx_fields_t* f, const sphx_params_t* p, uint32_t first, uint32_t last);
void rhoTimestep(sphx_workspace* ws, const sphx_fields_t* f, const sphx_params_t* p, uint32_t first, uint32_t last,
                 sphx_timestep_t* ts);
void computeTimestep(sphx_workspace* ws, const sphx_fields_t* f, const sphx_params_t* p, uint32_t first,
                     uint32_t last, sphx_timestep_t* ts);
void computePositions(sphx_workspace* ws, const sphx_box_t* box, const sphx_fields_t* f, const sphx_history_t* hist,
                      const sphx_params_t* p, uint32_t first, uint32_t last, double dt, double dt_m1, double constCv,
                      const float* dtM1Rung = nullptr, int numRungs = 0, const uint8_t* rung = nullptr);
//! block time steps: the drift of a hierarchy's particles, the per-group steps, the rung store
void driftPositions(sphx_workspace* ws, const sphx_fields_t* f, const sphx_history_t* hist, const sphx_params_t* p,
                    uint32_t first, uint32_t last, double dt, double dtBack, const float* dtM1Rung, int numRungs,
                    const uint8_t* rung, double constCv);
void groupTimestep(sphx_workspace* ws, const sphx_fields_t* f, const sphx_params_t* p, uint32_t first, uint32_t last,
                   bool acc, float coeff, float* groupDt);
void storeRung(sphx_workspace* ws, uint32_t first, uint32_t last, uint8_t rung, uint8_t* particleRungs);
void updateSmoothingLength(sphx_workspace* ws, const sphx_fields_t* f, const sphx_params_t* p, uint32_t first,
                           uint32_t last);
void conservedQuantities(sphx_workspace* ws, const sphx_fields_t* f, const sphx_params_t* p, uint32_t first,
                         uint32_t last, double cv, sphx_conserved_t* out);
void iadDivvCurlv(sphx_workspace* ws, const DevBox& box, const sphx_fields_t* f, const sphx_params_t* p,
                  uint32_t first, uint32_t last);
void avSwitches(sphx_workspace* ws, const DevBox& box, const sphx_fields_t* f, const sphx_params_t* p,
                uint32_t first, uint32_t last);
double momentumEnergyVe(sphx_workspace* ws, const DevBox& box, const sphx_fields_t* f, const sphx_params_t* p,
                        uint32_t first, uint32_t last, bool avClean);

} // namespace sphx
