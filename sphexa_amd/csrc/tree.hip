// Cornerstone leaf array + linked octree + layout + geometric node centres, built top-down.
//
// The reference converges a leaf array by repeated count/split/merge sweeps (computeOctree,
// domain/include/cstone/tree/csarray.hpp:409-440) and then derives the internal tree by sorting
// placeholder-bit prefixes (buildOctreeGpu, tree/octree_gpu.cu:137-157). Its fixed point is
// unique: a node is internal iff it holds more than `bucket` particles and lies above the
// deepest level. Descending level by level from the root and splitting exactly those nodes
// produces that same tree, and because children are appended in key order behind their
// level-ordered parents, the node array comes out directly in the reference's sorted
// prefix order (level-major, SFC-minor). One host read-back per level sizes the next one.
#include <cstring> // rocprim uses memset without including it
#include <rocprim/rocprim.hpp>

#include "common.hpp"

namespace sphx
{

namespace
{

__device__ __forceinline__ uint32_t lowerBound(const Key* keys, uint32_t lo, uint32_t hi, Key v)
{
    while (lo < hi)
    {
        uint32_t mid = lo + (hi - lo) / 2;
        if (keys[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

__global__ void initRoot(Key* nodeKey, uint32_t* lo, uint32_t* hi, uint8_t* level, uint32_t n)
{
    nodeKey[0] = 0;
    lo[0]      = 0;
    hi[0]      = n;
    level[0]   = 0;
}

__global__ void classifyKernel(const uint32_t* lo, const uint32_t* hi, uint32_t* flag, uint32_t begin, uint32_t end,
                               unsigned bucket, unsigned level)
{
    uint32_t i = begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= end) { return; }
    flag[i] = (hi[i] - lo[i] > bucket) && (level < kMaxLevel);
}

//! each internal node of the level writes its 8 children behind the level; 8 consecutive threads
//! per node, thread c finding the first particle of child c (the searches run side by side
//! instead of as one chain of dependent loads)
__global__ void expandKernel(const Key* __restrict__ keys, Key* nodeKey, uint32_t* lo, uint32_t* hi, uint8_t* nodeLevel,
                             const uint32_t* flag, const uint32_t* inclRank, uint32_t internalBefore,
                             uint32_t begin, uint32_t end, unsigned level)
{
    const size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const size_t i = begin + t / 8;
    const int    c = int(t % 8);
    if (i >= end || !flag[i]) { return; } // uniform over the node's 8 threads
    uint32_t r     = inclRank[i] - 1 - internalBefore; // rank among this level's internal nodes
    uint32_t slot  = end + 8 * r + uint32_t(c);
    Key      k     = nodeKey[i];
    Key      range = nodeRange(level + 1);
    uint32_t a = lo[i], b = hi[i];
    uint32_t first = (c == 0) ? a : lowerBound(keys, a, b, k + Key(c) * range);
    uint32_t next  = __shfl_down(first, 1, 8); // first particle of child c + 1
    nodeKey[slot]   = k + Key(c) * range;
    lo[slot]        = first;
    hi[slot]        = (c == 7) ? b : next;
    nodeLevel[slot] = uint8_t(level + 1);
}

//! prefixes, child links, parents, centres; leaves are collected for the key sort
__global__ void linkKernel(const Key* nodeKey, const uint8_t* nodeLevel, const uint32_t* flag, const uint32_t* inclRank,
                           const int32_t* levelStart, uint32_t numNodes, DevBox box, Key* prefixes,
                           TreeNodeIndex* childOffsets, TreeNodeIndex* parents, TreeNodeIndex* internalToLeaf,
                           double* centers, double* sizes, Key* leafKey, uint32_t* leafNode, const uint32_t* leafRank)
{
#pragma clang fp contract(off)
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= numNodes) { return; }
    Key      k = nodeKey[i];
    unsigned l = nodeLevel[i];
    prefixes[i] = encodePlaceholder(k, 3 * l);

    if (flag[i])
    {
        // ranks of internal nodes are global (inclusive scan over all nodes); the level's
        // first internal rank is the count of internal nodes on shallower levels
        uint32_t rankInLevel = inclRank[i] - 1 - (levelStart[l] == 0 ? 0 : inclRank[levelStart[l] - 1]);
        TreeNodeIndex child  = levelStart[l + 1] + 8 * TreeNodeIndex(rankInLevel);
        childOffsets[i]      = child;
        parents[(child - 1) / 8] = TreeNodeIndex(i);
        internalToLeaf[i]    = -1;
    }
    else
    {
        childOffsets[i]   = 0;
        uint32_t s        = leafRank[i];
        leafKey[s]        = k;
        leafNode[s]       = i;
    }

    // geometric centre / half size: hilbertIBox + centerAndSize (sfc/hilbert.hpp:259-274,
    // sfc/box.hpp:318-334)
    unsigned ix, iy, iz;
    hilbertDecode(k, ix, iy, iz);
    unsigned     cube = (1u << kMaxLevel) >> l;
    unsigned     mask = ~(cube - 1);
    unsigned     ic[3] = {ix & mask, iy & mask, iz & mask};
    const double uL    = 1.0 / double(1u << kMaxLevel);
    for (int d = 0; d < 3; ++d)
    {
        int    imin = int(ic[d]), imax = int(ic[d] + cube);
        double half = 0.5 * uL * box.len[d];
        centers[3 * size_t(i) + d] = box.lo[d] + double(imax + imin) * half;
        sizes[3 * size_t(i) + d]   = double(imax - imin) * half;
    }
}

__global__ void leafFlagKernel(const uint32_t* flag, uint32_t* leafFlag, uint32_t numNodes)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < numNodes) { leafFlag[i] = 1u - flag[i]; }
}

__global__ void exclusiveFromInclusive(const uint32_t* incl, const uint32_t* v, uint32_t* excl, uint32_t n)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) { excl[i] = incl[i] - v[i]; }
}

__global__ void leafFinalize(const Key* sortedLeafKey, const uint32_t* sortedLeafNode, const uint32_t* lo,
                             const uint32_t* hi, Key* leaves, uint32_t* counts, uint32_t* layout,
                             TreeNodeIndex* internalToLeaf, uint32_t L, uint32_t n)
{
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s > L) { return; }
    if (s == L)
    {
        leaves[L] = nodeRange(0);
        layout[L] = n;
        return;
    }
    uint32_t node        = sortedLeafNode[s];
    leaves[s]            = sortedLeafKey[s];
    counts[s]            = hi[node] - lo[node];
    layout[s]            = lo[node];
    internalToLeaf[node] = TreeNodeIndex(s);
}

//! leaves in the subtree of each node of one level (the deeper levels are done)
__global__ void subtreeLeavesKernel(const TreeNodeIndex* childOffsets, int32_t* leaves, int32_t first, int32_t last)
{
    const int32_t i = first + int32_t(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= last) { return; }
    const TreeNodeIndex c = childOffsets[i];
    int32_t             s = 1;
    if (c != 0)
    {
        s = 0;
        for (int o = 0; o < 8; ++o)
            s += leaves[c + o];
    }
    leaves[i] = s;
}

/*! Visiting order of the search for the children of each internal node of one level (the
 *  shallower levels are done): singleTraversal (traversal/traversal.hpp:52-93) handles a node's
 *  leaf children in octant order as it meets them, pushes its internal children in octant order
 *  and pops the last pushed first. rankLo of a child = rankLo of the parent plus the leaves visited
 *  before it; childPush lists the children from last to first visited. */
__global__ void visitOrderKernel(const TreeNodeIndex* childOffsets, const int32_t* leaves, int32_t* rankLo,
                                 uint32_t* childPush, int32_t first, int32_t last)
{
    const int32_t i = first + int32_t(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= last) { return; }
    const TreeNodeIndex c = childOffsets[i];
    if (c == 0) { return; }
    int32_t  base = rankLo[i];
    bool     isLeaf[8];
    for (int o = 0; o < 8; ++o)
        isLeaf[o] = childOffsets[c + o] == 0;
    uint32_t push = 0;
    int      slot = 7; // childPush slot 7 = the first visited
    for (int o = 0; o < 8; ++o)
    {
        if (!isLeaf[o]) { continue; }
        rankLo[c + o] = base++;
        push |= uint32_t(o) << (3 * slot--);
    }
    for (int o = 7; o >= 0; --o)
    {
        if (isLeaf[o]) { continue; }
        rankLo[c + o] = base;
        base += leaves[c + o];
        push |= uint32_t(o) << (3 * slot--);
    }
    childPush[i] = push;
}

__global__ void zeroRankKernel(int32_t* rankLo) { rankLo[0] = 0; }

template<class T>
void growCopy(sphx_workspace* ws, DevBuf<T>& buf, size_t used, size_t need)
{
    if (need <= buf.cap) { return; }
    DevBuf<T> nb;
    nb.reserve(std::max(need, 2 * buf.cap));
    if (used) SPHX_HIP(hipMemcpyAsync(nb.ptr, buf.ptr, used * sizeof(T), hipMemcpyDeviceToDevice, ws->stream));
    SPHX_HIP(hipStreamSynchronize(ws->stream));
    buf.release();
    buf = nb;
}

} // namespace

namespace
{

/* The linked tree whose nodes are exactly the octree nodes holding more than `bucket` of the sorted
 * keys `keys` (and their children): the converged cornerstone tree of particle keys with the bucket
 * size, or, with bucket 1 and the leaf starts of a cornerstone array as keys, the octree of that leaf
 * array (a node is internal iff more than one leaf starts inside it). Leaf counts and layout come out
 * in units of keys. */
void buildLinked(sphx_workspace* ws, const Key* keys, size_t nParticles, unsigned bucket, const DevBox& box)
{
    hipStream_t  st = ws->stream;
    uint32_t     n  = uint32_t(nParticles);
    auto&        s   = ws->tree.build;
    size_t       cap = std::max<size_t>(4096, 16 * (nParticles / (bucket + 1) + 1));
    s.nodeKey.reserve(cap), s.lo.reserve(cap), s.hi.reserve(cap), s.flag.reserve(cap), s.rank.reserve(cap);
    s.level.reserve(cap);

    initRoot<<<1, 1, 0, st>>>(s.nodeKey.get(), s.lo.get(), s.hi.get(), s.level.get(), n);
    SPHX_CHECK_LAUNCH();

    std::vector<int32_t> levelStart(kMaxLevel + 3, 0);
    uint32_t             begin = 0, end = 1;
    unsigned             level = 0;
    uint32_t             internalBefore = 0;
    size_t               scanBytes      = 0;
    while (true)
    {
        levelStart[level] = int32_t(begin);
        uint32_t cnt      = end - begin;
        classifyKernel<<<ceilDiv(cnt, 256), 256, 0, st>>>(s.lo.get(), s.hi.get(), s.flag.get(), begin, end, bucket,
                                                           level);
        SPHX_CHECK_LAUNCH();
        // global inclusive scan of flags up to `end` (so ranks are tree-global)
        SPHX_HIP(rocprim::inclusive_scan(nullptr, scanBytes, s.flag.get(), s.rank.get(), end, rocprim::plus<uint32_t>(),
                                         st));
        void* tmp = scratchBytes(ws, scanBytes);
        SPHX_HIP(rocprim::inclusive_scan(tmp, scanBytes, s.flag.get(), s.rank.get(), end, rocprim::plus<uint32_t>(),
                                         st));
        uint32_t totalInternal = 0;
        SPHX_HIP(hipMemcpyAsync(&totalInternal, s.rank.get() + end - 1, 4, hipMemcpyDeviceToHost, st));
        SPHX_HIP(hipStreamSynchronize(st));
        uint32_t numInternalHere = totalInternal - internalBefore;
        if (numInternalHere == 0) { break; }

        size_t need = size_t(end) + 8 * size_t(numInternalHere);
        growCopy(ws, s.nodeKey, end, need), growCopy(ws, s.lo, end, need), growCopy(ws, s.hi, end, need);
        growCopy(ws, s.flag, end, need), growCopy(ws, s.rank, end, need), growCopy(ws, s.level, end, need);

        expandKernel<<<ceilDiv(size_t(cnt) * 8, 256), 256, 0, st>>>(keys, s.nodeKey.get(), s.lo.get(), s.hi.get(),
                                                                   s.level.get(), s.flag.get(), s.rank.get(),
                                                                   internalBefore, begin, end, level);
        SPHX_CHECK_LAUNCH();
        internalBefore = totalInternal;
        begin          = end;
        end            = uint32_t(need);
        level++;
    }
    for (unsigned l = level + 1; l <= kMaxLevel + 2; ++l)
        levelStart[l] = int32_t(end);

    uint32_t numNodes    = end;
    uint32_t numInternal = internalBefore;
    uint32_t L           = numNodes - numInternal;

    Tree& t = ws->tree;
    t.prefixes.reserve(numNodes), t.childOffsets.reserve(numNodes + 1), t.internalToLeaf.reserve(numNodes);
    t.parents.reserve(std::max<uint32_t>(1, (numNodes - 1) / 8)), t.levelRange.reserve(kMaxLevel + 2);
    t.centers.reserve(3 * size_t(numNodes)), t.sizes.reserve(3 * size_t(numNodes));
    t.leaves.reserve(L + 1), t.counts.reserve(L), t.layout.reserve(L + 1);

    DevBuf<int32_t>& dLevelStart = s.levelStart;
    dLevelStart.reserve(kMaxLevel + 3);
    SPHX_HIP(hipMemcpyAsync(dLevelStart.get(), levelStart.data(), (kMaxLevel + 3) * 4, hipMemcpyHostToDevice, st));

    // leaf ranks in level order, then sort leaves by key
    s.leafFlag.reserve(numNodes), s.leafRank.reserve(numNodes);
    s.leafKey.reserve(L), s.leafNode.reserve(L), s.leafKeySorted.reserve(L), s.leafNodeSorted.reserve(L);
    leafFlagKernel<<<ceilDiv(numNodes, 256), 256, 0, st>>>(s.flag.get(), s.leafFlag.get(), numNodes);
    SPHX_CHECK_LAUNCH();
    SPHX_HIP(rocprim::inclusive_scan(nullptr, scanBytes, s.leafFlag.get(), s.leafRank.get(), numNodes,
                                     rocprim::plus<uint32_t>(), st));
    SPHX_HIP(rocprim::inclusive_scan(scratchBytes(ws, scanBytes), scanBytes, s.leafFlag.get(), s.leafRank.get(),
                                     numNodes, rocprim::plus<uint32_t>(), st));
    exclusiveFromInclusive<<<ceilDiv(numNodes, 256), 256, 0, st>>>(s.leafRank.get(), s.leafFlag.get(),
                                                                    s.leafRank.get(), numNodes);
    SPHX_CHECK_LAUNCH();

    linkKernel<<<ceilDiv(numNodes, 256), 256, 0, st>>>(
        s.nodeKey.get(), s.level.get(), s.flag.get(), s.rank.get(), dLevelStart.get(), numNodes, box, t.prefixes.get(),
        t.childOffsets.get(), t.parents.get(), t.internalToLeaf.get(), t.centers.get(), t.sizes.get(), s.leafKey.get(),
        s.leafNode.get(), s.leafRank.get());
    SPHX_CHECK_LAUNCH();

    // leaf keys lie below 2^63 and are multiples of nodeRange(deepest level): only the bits
    // between decide their order
    const unsigned lowBit    = 3u * (kMaxLevel - std::min(level, kMaxLevel));
    size_t         sortBytes = 0;
    SPHX_HIP(rocprim::radix_sort_pairs(nullptr, sortBytes, s.leafKey.get(), s.leafKeySorted.get(), s.leafNode.get(),
                                       s.leafNodeSorted.get(), L, lowBit, 63, st));
    SPHX_HIP(rocprim::radix_sort_pairs(scratchBytes(ws, sortBytes), sortBytes, s.leafKey.get(), s.leafKeySorted.get(),
                                       s.leafNode.get(), s.leafNodeSorted.get(), L, lowBit, 63, st));
    leafFinalize<<<ceilDiv(L + 1, 256), 256, 0, st>>>(s.leafKeySorted.get(), s.leafNodeSorted.get(), s.lo.get(),
                                                      s.hi.get(), t.leaves.get(), t.counts.get(), t.layout.get(),
                                                      t.internalToLeaf.get(), L, n);
    SPHX_CHECK_LAUNCH();

    std::vector<int32_t> levelRange(kMaxLevel + 2);
    for (unsigned l = 0; l <= kMaxLevel; ++l)
        levelRange[l] = levelStart[l];
    levelRange[kMaxLevel + 1] = int32_t(numNodes);
    SPHX_HIP(hipMemcpyAsync(t.levelRange.get(), levelRange.data(), (kMaxLevel + 2) * 4, hipMemcpyHostToDevice, st));

    // the search's visiting order: subtree leaf counts bottom-up, then ranks top-down
    t.rankLo.reserve(numNodes), t.subtreeLeaves.reserve(numNodes), t.childPush.reserve(numNodes);
    for (int l = int(level); l >= 0; --l)
    {
        const int32_t a = levelRange[l], b = levelRange[l + 1];
        if (b > a)
        {
            subtreeLeavesKernel<<<ceilDiv(size_t(b - a), 256), 256, 0, st>>>(t.childOffsets.get(),
                                                                          t.subtreeLeaves.get(), a, b);
            SPHX_CHECK_LAUNCH();
        }
    }
    zeroRankKernel<<<1, 1, 0, st>>>(t.rankLo.get());
    SPHX_CHECK_LAUNCH();
    for (int l = 0; l < int(level); ++l)
    {
        const int32_t a = levelRange[l], b = levelRange[l + 1];
        if (b > a)
        {
            visitOrderKernel<<<ceilDiv(size_t(b - a), 256), 256, 0, st>>>(
                t.childOffsets.get(), t.subtreeLeaves.get(), t.rankLo.get(), t.childPush.get(), a, b);
            SPHX_CHECK_LAUNCH();
        }
    }
    SPHX_HIP(hipStreamSynchronize(st));

    t.numLeafNodes     = int32_t(L);
    t.numInternalNodes = int32_t(numInternal);
    t.numNodes         = int32_t(numNodes);
    t.numParticles     = nParticles;
    t.valid            = true;
}

//! particle count of every node of the tree: a leaf's from its count, an internal node's from the layout
//! of the leaves it spans
__global__ void nodeCountKernel(const Key* __restrict__ prefixes, const TreeNodeIndex* __restrict__ childOffsets,
                                const TreeNodeIndex* __restrict__ internalToLeaf, const Key* __restrict__ leaves,
                                const uint32_t* __restrict__ layout, const uint32_t* __restrict__ counts, uint32_t L,
                                uint32_t N, uint32_t* __restrict__ nodeCnt)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) { return; }
    if (childOffsets[i] == 0)
    {
        nodeCnt[i] = counts[internalToLeaf[i]];
        return;
    }
    const Key      p     = prefixes[i];
    const Key      start = decodePlaceholder(p);
    const Key      end   = start + nodeRange(decodePrefixLength(p) / 3u);
    const uint32_t a     = lowerBound(leaves, 0, L, start);
    const uint32_t b     = lowerBound(leaves, a, L + 1, end);
    nodeCnt[i]           = layout[b] - layout[a];
}

/* mergeCountAndMacOp (focus/rebalance.hpp:31-68) on one rank, the focus spanning every key: node i > 0
 * merges (0) when its parent holds at most `bucket` particles; a leaf above the deepest level holding
 * more splits in 8; else 1 */
__global__ void nodeOpKernel(const Key* __restrict__ prefixes, const TreeNodeIndex* __restrict__ childOffsets,
                             const TreeNodeIndex* __restrict__ parents, const uint32_t* __restrict__ nodeCnt,
                             uint32_t N, unsigned bucket, int32_t* __restrict__ op)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) { return; }
    if (i > 0 && nodeCnt[parents[(i - 1) / 8]] <= bucket)
    {
        op[i] = 0;
        return;
    }
    const unsigned level = decodePrefixLength(prefixes[i]) / 3u;
    op[i]                = (childOffsets[i] == 0 && level < kMaxLevel && nodeCnt[i] > bucket) ? 8 : 1;
}

/* protectAncestors / nzAncestorOp (focus/rebalance.hpp:88-107,219-233) for every node, reading the ops
 * before protection: a node marked 0 takes the op of its closest ancestor not marked 0 when it is that
 * ancestor's first descendant, else stays 0. An ancestor marked 0 that protection would turn to 1 is the
 * first descendant of its own closest unmarked ancestor, so walking past it gives the same answer as the
 * reference's in-place loop. Leaves store their final op by leaf index; any node left != 1 marks a change. */
__global__ void protectKernel(const Key* __restrict__ prefixes, const TreeNodeIndex* __restrict__ parents,
                              const TreeNodeIndex* __restrict__ internalToLeaf, const int32_t* __restrict__ op,
                              uint32_t N, int32_t* __restrict__ leafOp, uint32_t* __restrict__ changed)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) { return; }
    int32_t o = op[i];
    if (o == 0) // never the root
    {
        TreeNodeIndex a = TreeNodeIndex(i);
        while (op[a] == 0) { a = parents[(a - 1) / 8]; }
        o = decodePlaceholder(prefixes[i]) == decodePlaceholder(prefixes[a]) ? op[a] : 0;
    }
    if (o != 1) { *changed = 1u; }
    const TreeNodeIndex leaf = internalToLeaf[i];
    if (leaf >= 0) { leafOp[leaf] = o; }
}

//! rebalanceTree / processNode (tree/csarray.hpp:340-389) with ops 0, 1 and 8 at offsets `offs`
__global__ void rebalanceKernel(const Key* __restrict__ leaves, const int32_t* __restrict__ offs, uint32_t L,
                                Key* __restrict__ next)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s > L) { return; }
    if (s == L)
    {
        next[offs[L]] = leaves[L];
        return;
    }
    const Key      node = leaves[s];
    const int32_t  k    = offs[s + 1] - offs[s];
    if (k == 1) { next[offs[s]] = node; }
    else if (k == 8)
    {
        const Key r = nodeRange(treeLevel(leaves[s + 1] - node) + 1u);
        for (int c = 0; c < 8; ++c)
            next[offs[s] + c] = node + Key(c) * r;
    }
}

//! first particle and count of every leaf of the linked leaf array (computeNodeCounts: binary searches)
__global__ void leafLayoutKernel(const Key* __restrict__ leaves, uint32_t L, const Key* __restrict__ keys, uint32_t n,
                                 uint32_t* __restrict__ layout, uint32_t* __restrict__ counts)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s > L) { return; }
    const uint32_t lo = s == L ? n : lowerBound(keys, 0, n, leaves[s]);
    layout[s]         = lo;
    if (s < L) { counts[s] = (s + 1 == L ? n : lowerBound(keys, lo, n, leaves[s + 1])) - lo; }
}

__global__ void setLeafOpEnd(int32_t* leafOp, uint32_t L, uint32_t* changed)
{
    leafOp[L] = 0;
    *changed  = 0u;
}

} // namespace

void buildTree(sphx_workspace* ws, const Key* keys, size_t nParticles, unsigned bucket, const DevBox& box)
{
    SPHX_REQUIRE(nParticles < (size_t(1) << 32), SPHX_ERR_ARG, "particle count must fit 32-bit indices");
    SPHX_REQUIRE(bucket >= 1, SPHX_ERR_ARG, "bucket size must be positive");
    buildLinked(ws, keys, nParticles, bucket, box);
}

bool updateTree(sphx_workspace* ws, const Key* keys, size_t nParticles, unsigned bucket, const DevBox& box)
{
    SPHX_REQUIRE(nParticles < (size_t(1) << 32), SPHX_ERR_ARG, "particle count must fit 32-bit indices");
    SPHX_REQUIRE(bucket >= 1, SPHX_ERR_ARG, "bucket size must be positive");
    Tree& t = ws->tree;
    SPHX_REQUIRE(t.valid, SPHX_ERR_STATE, "no tree to update");
    hipStream_t    st = ws->stream;
    auto&          s  = t.build;
    const uint32_t N = uint32_t(t.numNodes), L = uint32_t(t.numLeafNodes);
    s.nodeCnt.reserve(N), s.nodeOp.reserve(N), s.leafOp.reserve(size_t(L) + 1), s.changed.reserve(1);
    setLeafOpEnd<<<1, 1, 0, st>>>(s.leafOp.get(), L, s.changed.get());
    SPHX_CHECK_LAUNCH();
    nodeCountKernel<<<ceilDiv(N, 256), 256, 0, st>>>(t.prefixes.get(), t.childOffsets.get(), t.internalToLeaf.get(),
                                                      t.leaves.get(), t.layout.get(), t.counts.get(), L, N,
                                                      s.nodeCnt.get());
    SPHX_CHECK_LAUNCH();
    nodeOpKernel<<<ceilDiv(N, 256), 256, 0, st>>>(t.prefixes.get(), t.childOffsets.get(), t.parents.get(),
                                                   s.nodeCnt.get(), N, bucket, s.nodeOp.get());
    SPHX_CHECK_LAUNCH();
    protectKernel<<<ceilDiv(N, 256), 256, 0, st>>>(t.prefixes.get(), t.parents.get(), t.internalToLeaf.get(),
                                                    s.nodeOp.get(), N, s.leafOp.get(), s.changed.get());
    SPHX_CHECK_LAUNCH();
    size_t bytes = 0;
    SPHX_HIP(rocprim::exclusive_scan(nullptr, bytes, s.leafOp.get(), s.leafOp.get(), 0, size_t(L) + 1,
                                     rocprim::plus<int32_t>(), st));
    SPHX_HIP(rocprim::exclusive_scan(scratchBytes(ws, bytes), bytes, s.leafOp.get(), s.leafOp.get(), 0,
                                     size_t(L) + 1, rocprim::plus<int32_t>(), st));
    int32_t  newL    = 0;
    uint32_t changed = 0;
    SPHX_HIP(hipMemcpyAsync(&newL, s.leafOp.get() + L, 4, hipMemcpyDeviceToHost, st));
    SPHX_HIP(hipMemcpyAsync(&changed, s.changed.get(), 4, hipMemcpyDeviceToHost, st));
    SPHX_HIP(hipStreamSynchronize(st));
    SPHX_REQUIRE(newL >= 1, SPHX_ERR_STATE, "tree update left no leaf");
    s.nextLeaves.reserve(size_t(newL) + 1);
    rebalanceKernel<<<ceilDiv(size_t(L) + 1, 256), 256, 0, st>>>(t.leaves.get(), s.leafOp.get(), L,
                                                                 s.nextLeaves.get());
    SPHX_CHECK_LAUNCH();

    // link the new leaf array (its leaf starts as the keys, one per leaf), then count the particles
    buildLinked(ws, s.nextLeaves.get(), size_t(newL), 1, box);
    SPHX_REQUIRE(t.numLeafNodes == newL, SPHX_ERR_STATE, "linking the updated leaves changed their number");
    const uint32_t n = uint32_t(nParticles);
    leafLayoutKernel<<<ceilDiv(size_t(newL) + 1, 256), 256, 0, st>>>(t.leaves.get(), uint32_t(newL), keys, n,
                                                                     t.layout.get(), t.counts.get());
    SPHX_CHECK_LAUNCH();
    SPHX_HIP(hipStreamSynchronize(st));
    t.numParticles = nParticles;
    return changed == 0;
}

} // namespace sphx
