/*
 * CPU restatement of SPH-EXA's hydro hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * This header is the oracle: the checker the HIP product path is compared
 * against. Only tests, the smoke check and the CPU-baseline leg of the bench
 * may load it; the product library never links or calls it.
 *
 * Every function restates one reference function (file:line cited on each,
 * paths relative to the reference repository root) and keeps the reference's
 * operation order, operand types and promotions, so that -- built with
 * -ffp-contract=off like the reference wrapper in oracle/_ref -- its results are
 * bitwise identical to the reference headers instantiated with the same types.
 *
 * Types: Tc = coordinate type (always double here), T = HydroType (double for the
 * plain FP64 form, float for the reference's SphTypes layout), Tm = mass type.
 */
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <limits>
#include <numeric>
#include <set>
#include <vector>

namespace orc
{

using LocalIndex    = uint32_t;
using TreeNodeIndex = int;
using Key           = uint64_t;

constexpr unsigned kMaxLevel = 21; // 64-bit keys: 21 bits per dimension
constexpr unsigned kUnused   = 1;  // leading unused bit in a 64-bit key
constexpr int      kTableSize = 20000;

/* ------------------------------------------------------------------------------------------------
 * Box  (domain/include/cstone/sfc/box.hpp:95-174)
 * ----------------------------------------------------------------------------------------------*/
struct Box
{
    double lo[3], hi[3];
    double len[3], ilen[3];
    int    pbc[3]; // 1 = periodic
    int    fbc[3]; // 2 = fixed

    Box() = default;
    Box(const double* limits6, const int* bc3)
    {
        for (int d = 0; d < 3; ++d)
        {
            lo[d]   = limits6[2 * d];
            hi[d]   = limits6[2 * d + 1];
            len[d]  = hi[d] - lo[d];
            ilen[d] = 1.0 / (hi[d] - lo[d]);
            pbc[d]  = bc3[d] == 1;
            fbc[d]  = bc3[d] == 2;
        }
    }
    bool anyPbc() const { return pbc[0] || pbc[1] || pbc[2]; }
};

/* ------------------------------------------------------------------------------------------------
 * Bit helpers (domain/include/cstone/sfc/common.hpp:109-213, 252-255, 271-275)
 * ----------------------------------------------------------------------------------------------*/
inline int clz64(Key x) { return x == 0 ? 64 : __builtin_clzll(x); }

inline Key nodeRange(unsigned level) { return Key(1) << (3u * (kMaxLevel - level)); }

//! level of a node whose key range is a power of 8 (common.hpp:186-191)
inline unsigned treeLevel(Key range) { return (clz64(range - 1) - kUnused) / 3; }

//! common.hpp:170-173
inline int commonPrefix(Key a, Key b) { return clz64(a ^ b) - int(kUnused); }

//! placeholder-bit (Warren-Salmon) encoding, common.hpp:203-210
inline Key encodePlaceholder(Key key, int prefixLength)
{
    int nShifts = 3 * int(kMaxLevel) - prefixLength;
    return (Key(1) << prefixLength) | (key >> nShifts);
}

//! common.hpp:220-224
inline unsigned decodePrefixLength(Key code) { return 8 * sizeof(Key) - 1 - clz64(code); }

//! common.hpp:234-241
inline Key decodePlaceholder(Key code)
{
    int prefixLength = decodePrefixLength(code);
    Key stripped     = code ^ (Key(1) << prefixLength);
    return stripped << (3 * kMaxLevel - prefixLength);
}

inline unsigned octalDigit(Key code, unsigned position) { return (code >> (3u * (kMaxLevel - position))) & 7u; }

/* ------------------------------------------------------------------------------------------------
 * Hilbert curve (domain/include/cstone/sfc/hilbert.hpp:44-93 encode, :130-172 decode)
 *
 * Per level, the Morton octant (x<<2|y<<1|z) is mapped to the Hilbert digit with the
 * table {0,1,3,2,7,6,4,5}; the remaining lower bits of the coordinates are then
 * reflected and permuted so that the next level is traversed in the child's frame.
 * ----------------------------------------------------------------------------------------------*/
inline Key hilbertEncode(unsigned px, unsigned py, unsigned pz)
{
    static const unsigned digitOf[8] = {0, 1, 3, 2, 7, 6, 4, 5};
    Key key = 0;
    for (int level = int(kMaxLevel) - 1; level >= 0; --level)
    {
        unsigned bx = (px >> level) & 1u, by = (py >> level) & 1u, bz = (pz >> level) & 1u;
        key = (key << 3) + digitOf[(bx << 2) | (by << 1) | bz];

        // reflections, applied to all bits (bits above `level` are never read again)
        unsigned rx = bx & ((!by) | bz);
        unsigned ry = (bx & (by | bz)) | (by & (!bz));
        unsigned rz = (bx & (!by) & (!bz)) | (by & (!bz));
        px ^= 0u - rx;
        py ^= 0u - ry;
        pz ^= 0u - rz;

        if (bz) // rotate (x,y,z) <- (y,z,x)
        {
            unsigned t = px;
            px = py, py = pz, pz = t;
        }
        else if (!by) // exchange x and z
        {
            std::swap(px, pz);
        }
    }
    return key;
}

inline void hilbertDecode(Key key, unsigned& ox, unsigned& oy, unsigned& oz)
{
    unsigned px = 0, py = 0, pz = 0;
    for (unsigned level = 0; level < kMaxLevel; ++level)
    {
        unsigned digit = (key >> (3 * level)) & 7u;
        unsigned bx = digit >> 2u, by = (digit >> 1u) & 1u, bz = digit & 1u;

        if (by ^ bz) // inverse rotation: (x,y,z) <- (z,x,y)
        {
            unsigned t = px;
            px = pz, pz = py, py = t;
        }
        else if ((!bx & !by & !bz) || (bx & by & bz)) { std::swap(px, pz); }

        unsigned mask = (1u << level) - 1;
        px ^= mask & (0u - (bx & (by | bz)));
        py ^= mask & (0u - ((bx & ((!by) | (!bz))) | ((!bx) & by & bz)));
        pz ^= mask & (0u - ((bx & (!by) & (!bz)) | (by & bz)));

        px |= (bx << level);
        py |= ((bx ^ by) << level);
        pz |= ((by ^ bz) << level);
    }
    ox = px, oy = py, oz = pz;
}

//! sfc3D, domain/include/cstone/sfc/sfc.hpp:142-178 (cube length 2^21, clamp to 2^21-1)
inline Key sfcKey(double x, double y, double z, const Box& box)
{
    constexpr int mcoord = (1 << kMaxLevel) - 1;
    const unsigned cube  = 1u << kMaxLevel;
    double mx = cube * box.ilen[0], my = cube * box.ilen[1], mz = cube * box.ilen[2];

    int ix = std::floor(x * mx) - box.lo[0] * mx;
    int iy = std::floor(y * my) - box.lo[1] * my;
    int iz = std::floor(z * mz) - box.lo[2] * mz;
    ix = std::min(ix, mcoord);
    iy = std::min(iy, mcoord);
    iz = std::min(iz, mcoord);
    return hilbertEncode(ix, iy, iz);
}

constexpr Key kRemoveKey = Key(1) << 63; // definitions.h removeKey

//! computeSfcKeys, sfc/sfc.hpp:268-275 (keys equal to removeKey are left untouched)
inline void computeSfcKeys(const double* x, const double* y, const double* z, Key* keys, size_t n, const Box& box)
{
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; ++i)
    {
        if (keys[i] != kRemoveKey) { keys[i] = sfcKey(x[i], y[i], z[i], box); }
    }
}

//! sort_by_key with std::stable_sort, primitives/gather.hpp:43-68
inline void stableSortByKey(Key* keys, LocalIndex* values, size_t n)
{
    std::vector<size_t> order(n);
    std::iota(order.begin(), order.end(), size_t(0));
    std::stable_sort(order.begin(), order.end(), [keys](size_t a, size_t b) { return keys[a] < keys[b]; });
    std::vector<Key>        k2(n);
    std::vector<LocalIndex> v2(n);
    for (size_t i = 0; i < n; ++i)
    {
        k2[i] = keys[order[i]];
        v2[i] = values[order[i]];
    }
    std::copy(k2.begin(), k2.end(), keys);
    std::copy(v2.begin(), v2.end(), values);
}

/* ------------------------------------------------------------------------------------------------
 * Cornerstone leaf array (domain/include/cstone/tree/csarray.hpp:182-440)
 * ----------------------------------------------------------------------------------------------*/
inline unsigned nodeCount(Key a, Key b, const Key* keys, size_t n)
{
    auto lo = std::lower_bound(keys, keys + n, a);
    auto hi = std::lower_bound(keys, keys + n, b);
    return unsigned(hi - lo);
}

//! calculateNodeOp, csarray.hpp:247-290: 0 merge, 1 keep, 8/64/512/4096 split
inline int nodeOp(const std::vector<Key>& tree, TreeNodeIndex i, const std::vector<unsigned>& counts, unsigned bucket)
{
    Key      node  = tree[i];
    unsigned level = treeLevel(tree[i + 1] - node);
    int      sib   = -1;
    if (level > 0)
    {
        sib = octalDigit(node, level);
        bool complete = tree[i - sib + 8] == tree[i - sib] + nodeRange(level - 1);
        if (!complete) { sib = -1; }
    }
    if (sib > 0)
    {
        size_t parentCount = 0;
        for (int s = 0; s < 8; ++s)
            parentCount += counts[i - sib + s];
        if (parentCount <= size_t(bucket)) { return 0; }
    }
    size_t c = counts[i];
    if (c > size_t(bucket) * 512 && level + 3 < kMaxLevel) { return 4096; }
    if (c > size_t(bucket) * 64 && level + 2 < kMaxLevel) { return 512; }
    if (c > size_t(bucket) * 8 && level + 1 < kMaxLevel) { return 64; }
    if (c > size_t(bucket) && level < kMaxLevel) { return 8; }
    return 1;
}

//! computeOctree (csarray.hpp:409-440): iterate count / rebalance until no node changes
inline void computeOctree(const Key* keys, size_t n, unsigned bucket, std::vector<Key>& tree, std::vector<unsigned>& counts)
{
    tree   = {0, nodeRange(0)};
    counts = {unsigned(n)};
    while (true)
    {
        TreeNodeIndex         numNodes = TreeNodeIndex(tree.size()) - 1;
        std::vector<int>      ops(numNodes + 1, 0);
        bool                  converged = true;
        for (TreeNodeIndex i = 0; i < numNodes; ++i)
        {
            ops[i] = nodeOp(tree, i, counts, bucket);
            if (ops[i] != 1) { converged = false; }
        }
        // rebalanceTree / processNode, csarray.hpp:340-389
        std::exclusive_scan(ops.begin(), ops.end(), ops.begin(), 0);
        std::vector<Key> next(ops[numNodes] + 1);
        for (TreeNodeIndex i = 0; i < numNodes; ++i)
        {
            Key      node  = tree[i];
            unsigned level = treeLevel(tree[i + 1] - node);
            int      op    = ops[i + 1] - ops[i];
            if (op == 1) { next[ops[i]] = node; }
            else if (op >= 8)
            {
                unsigned down = (op == 8) ? 1 : (op == 64) ? 2 : (op == 512) ? 3 : 4;
                for (int s = 0; s < op; ++s)
                    next[ops[i] + s] = node + Key(s) * nodeRange(level + down);
            }
        }
        next.back() = tree.back();
        tree.swap(next);
        counts.resize(tree.size() - 1);
        for (size_t i = 0; i + 1 < tree.size(); ++i)
            counts[i] = nodeCount(tree[i], tree[i + 1], keys, n);
        if (converged) { break; }
    }
}

/* ------------------------------------------------------------------------------------------------
 * Linked octree (domain/include/cstone/tree/octree.hpp:79-197)
 *
 * Same output as buildOctreeCpu: all nodes (internal + leaves) in placeholder-bit
 * prefix order (level-major, SFC-minor), childOffsets (0 = leaf), parents per sibling
 * group, levelRange[kMaxLevel+2] and internalToLeaf (= cornerstone leaf index for
 * leaves). Internal nodes are found here as the set of proper ancestors of all
 * leaves instead of the reference's binary-radix-tree index mapping.
 * ----------------------------------------------------------------------------------------------*/
struct Octree
{
    TreeNodeIndex              numLeafNodes{0}, numInternalNodes{0}, numNodes{0};
    std::vector<Key>           prefixes;
    std::vector<TreeNodeIndex> childOffsets, parents, levelRange, internalToLeaf;
};

inline void buildOctree(const std::vector<Key>& leaves, Octree& o)
{
    TreeNodeIndex L = TreeNodeIndex(leaves.size()) - 1;
    o.numLeafNodes     = L;
    o.numInternalNodes = (L - 1) / 7;
    o.numNodes         = L + o.numInternalNodes;

    std::vector<std::pair<Key, TreeNodeIndex>> nodes; // (prefix, leaf index or -1)
    std::set<Key>                              internal;
    for (TreeNodeIndex t = 0; t < L; ++t)
    {
        Key      k     = leaves[t];
        unsigned level = treeLevel(leaves[t + 1] - k);
        nodes.push_back({encodePlaceholder(k, 3 * level), t});
        for (int l = int(level) - 1; l >= 0; --l)
        {
            Key anc = encodePlaceholder(k & ~(nodeRange(l) - 1), 3 * l);
            if (!internal.insert(anc).second) { break; }
        }
    }
    for (Key p : internal)
        nodes.push_back({p, -1});
    std::sort(nodes.begin(), nodes.end());

    TreeNodeIndex N = TreeNodeIndex(nodes.size());
    o.prefixes.resize(N);
    o.internalToLeaf.resize(N);
    for (TreeNodeIndex i = 0; i < N; ++i)
    {
        o.prefixes[i]       = nodes[i].first;
        o.internalToLeaf[i] = nodes[i].second; // internal nodes: negative
    }
    o.levelRange.resize(kMaxLevel + 2);
    for (unsigned l = 0; l <= kMaxLevel; ++l)
        o.levelRange[l] =
            TreeNodeIndex(std::lower_bound(o.prefixes.begin(), o.prefixes.end(), encodePlaceholder(0, 3 * l)) -
                          o.prefixes.begin());
    o.levelRange[kMaxLevel + 1] = N;

    o.childOffsets.assign(N + 1, 0);
    o.parents.assign(std::max(1, (N - 1) / 8), 0);
    for (TreeNodeIndex i = 0; i < N; ++i)
    {
        if (o.internalToLeaf[i] >= 0) { continue; }
        Key      prefix = o.prefixes[i];
        unsigned plen   = decodePrefixLength(prefix);
        unsigned level  = plen / 3;
        Key      child  = encodePlaceholder(decodePlaceholder(prefix), plen + 3);
        auto     b      = o.prefixes.begin() + o.levelRange[level + 1];
        auto     e      = o.prefixes.begin() + o.levelRange[level + 2];
        auto     it     = std::lower_bound(b, e, child);
        if (it != e && *it == child)
        {
            TreeNodeIndex c   = TreeNodeIndex(it - o.prefixes.begin());
            o.childOffsets[i] = c;
            o.parents[(c - 1) / 8] = i;
        }
    }
}

//! nodeFpCenters, focus/source_center.hpp:130-141 with hilbertIBox (hilbert.hpp:259-274) and
//! centerAndSize (sfc/box.hpp:318-334)
inline void nodeCenters(const std::vector<Key>& prefixes, const Box& box, std::vector<double>& centers,
                        std::vector<double>& sizes)
{
    size_t n = prefixes.size();
    centers.resize(3 * n);
    sizes.resize(3 * n);
    const double uL = 1.0 / double(1 << kMaxLevel);
    for (size_t i = 0; i < n; ++i)
    {
        Key      start = decodePlaceholder(prefixes[i]);
        unsigned level = decodePrefixLength(prefixes[i]) / 3;
        unsigned cube  = (1u << kMaxLevel) >> level;
        unsigned mask  = ~(cube - 1);
        unsigned ic[3];
        hilbertDecode(start, ic[0], ic[1], ic[2]);
        for (int d = 0; d < 3; ++d)
        {
            int    imin = int(ic[d] & mask), imax = imin + int(cube);
            double half = 0.5 * uL * box.len[d];
            centers[3 * i + d] = box.lo[d] + (imax + imin) * half;
            sizes[3 * i + d]   = (imax - imin) * half;
        }
    }
}

//! OctreeNsView, tree/octree.hpp:280-300 (+ layout from leaf counts, focus centres)
struct NsTree
{
    std::vector<Key>        leaves;
    std::vector<unsigned>   counts;
    std::vector<LocalIndex> layout;
    Octree                  oct;
    std::vector<double>     centers, sizes;
};

//! link, layout and centres of the leaves and counts already in t
inline void finishNsTree(const Box& box, NsTree& t)
{
    buildOctree(t.leaves, t.oct);
    t.layout.assign(t.counts.size() + 1, 0);
    std::inclusive_scan(t.counts.begin(), t.counts.end(), t.layout.begin() + 1);
    nodeCenters(t.oct.prefixes, box, t.centers, t.sizes);
}

inline void buildNsTree(const Key* sortedKeys, size_t n, unsigned bucket, const Box& box, NsTree& t)
{
    computeOctree(sortedKeys, n, bucket, t.leaves, t.counts);
    finishNsTree(box, t);
}

/* ------------------------------------------------------------------------------------------------
 * Tree update of the syncs after the first (Domain::sync, domain/include/cstone/domain/domain.hpp:
 * 199-212): the first sync converges the tree; every later one applies ONE rebalance of the previous
 * leaves, decided on the previous counts, then counts the new keys in the new leaves
 * (focusTree_.updateTree + updateCounts). On one rank the focus covers the whole key range, which is
 * FocusedOctreeSingleNode::update (focus/octree_focus.hpp:348-385) around CombinedUpdate::updateFocus
 * (:65-130):
 *  - node counts of every node: leaf counts scattered to the leaves, summed upwards (upsweep);
 *  - mergeCountAndMacOp (focus/rebalance.hpp:31-68): node i > 0 merges (0) when its parent holds at
 *    most `bucket` particles; a leaf above the deepest level holding more splits in 8; else 1. Every
 *    node lies in the focus and every sibling group in its fringe, so the MAC never decides;
 *  - protectAncestors / nzAncestorOp (rebalance.hpp:88-107,219-233): a node marked 0 takes the op of its
 *    closest ancestor not marked 0 when it is that ancestor's first descendant, else stays 0, so a
 *    merged subtree becomes one leaf (in index order, every ancestor is decided before its descendants,
 *    which gives the same result as the reference's parallel in-place loop);
 *  - the mandatory keys 0 and 2^63 (enforceKeys) are always leaf bounds: nothing to enforce;
 *  - the leaves' ops to rebalanceTree / processNode (tree/csarray.hpp:340-389); counts of the new
 *    leaves from the keys (computeNodeCounts).
 * Leaves may hold more than `bucket` particles afterwards, until later syncs split them. Returns
 * whether no node changed (the updateFocus result).
 * ----------------------------------------------------------------------------------------------*/
inline bool updateNsTree(const Key* sortedKeys, size_t n, unsigned bucket, const Box& box, NsTree& t)
{
    const Octree&         o = t.oct;
    const TreeNodeIndex   N = o.numNodes;
    std::vector<unsigned> cnt(N, 0);
    for (TreeNodeIndex i = 0; i < N; ++i)
        if (o.internalToLeaf[i] >= 0) { cnt[i] = t.counts[o.internalToLeaf[i]]; }
    // level-major order: every child follows its parent, so a reverse sweep sums children first
    for (TreeNodeIndex i = N - 1; i >= 0; --i)
    {
        if (o.childOffsets[i] == 0) { continue; }
        unsigned s = 0;
        for (int c = 0; c < 8; ++c)
            s += cnt[o.childOffsets[i] + c];
        cnt[i] = s;
    }
    std::vector<int> op(N, 1);
    for (TreeNodeIndex i = 0; i < N; ++i)
    {
        unsigned level = decodePrefixLength(o.prefixes[i]) / 3;
        if (i > 0 && cnt[o.parents[(i - 1) / 8]] <= bucket)
        {
            op[i] = 0;
            continue;
        }
        bool leaf = o.childOffsets[i] == 0;
        op[i]     = (leaf && level < kMaxLevel && cnt[i] > bucket) ? 8 : 1;
    }
    bool converged = true;
    for (TreeNodeIndex i = 0; i < N; ++i)
    {
        if (op[i] == 0)
        {
            TreeNodeIndex a = i;
            while (op[a] == 0) { a = o.parents[(a - 1) / 8]; }
            op[i] = decodePlaceholder(o.prefixes[i]) == decodePlaceholder(o.prefixes[a]) ? op[a] : 0;
        }
        if (op[i] != 1) { converged = false; }
    }
    const TreeNodeIndex L = o.numLeafNodes;
    std::vector<int>    ops(L + 1, 0);
    for (TreeNodeIndex i = 0; i < N; ++i)
        if (o.internalToLeaf[i] >= 0) { ops[o.internalToLeaf[i]] = op[i]; }
    std::exclusive_scan(ops.begin(), ops.end(), ops.begin(), 0);
    std::vector<Key> next(ops[L] + 1);
    for (TreeNodeIndex i = 0; i < L; ++i)
    {
        Key      node  = t.leaves[i];
        unsigned level = treeLevel(t.leaves[i + 1] - node);
        int      k     = ops[i + 1] - ops[i];
        if (k == 1) { next[ops[i]] = node; }
        else if (k == 8)
        {
            for (int s = 0; s < 8; ++s)
                next[ops[i] + s] = node + Key(s) * nodeRange(level + 1);
        }
    }
    next.back() = t.leaves.back();
    t.leaves.swap(next);
    t.counts.resize(t.leaves.size() - 1);
    for (size_t i = 0; i + 1 < t.leaves.size(); ++i)
        t.counts[i] = nodeCount(t.leaves[i], t.leaves[i + 1], sortedKeys, n);
    finishNsTree(box, t);
    return converged;
}

/* ------------------------------------------------------------------------------------------------
 * Spatial target groups: computeGroupSplits with groupSize = nsGroupSize() = 64, one 64-wide
 * wavefront per fixed group (N = 1), as the reference builds its GroupView on the GPU
 * (sph/include/sph/groups.hpp:42-55, traversal/groups_gpu.cu:34-101, groups_gpu.cuh:37-76,149-216).
 * ----------------------------------------------------------------------------------------------*/

/*! Group starts of [first, last) followed by `last`. Each fixed group of 64 consecutive particles
 * (lanes past last repeat particle last - 1) is split between consecutive particles whose
 * distance in unit-box coordinates (x / Lx, ...) exceeds distCrit = tolFactor * cbrt(v), v the
 * smallest unit-box volume of the leaves holding the group's particles. That volume is 8 (s/2)^3
 * for a leaf of side s = 2^-level, so its cube root is exactly 2^-level, formed here directly;
 * distCrit is float (h's type on the reference's GPU path) and squared in double. */
inline void groupSplits(LocalIndex first, LocalIndex last, const double* x, const double* y, const double* z,
                        const std::vector<Key>& leaves, const std::vector<LocalIndex>& layout, const Box& box,
                        float tolFactor, std::vector<LocalIndex>& groups)
{
    groups.clear();
    const TreeNodeIndex numLeaves = TreeNodeIndex(leaves.size()) - 1;
    const double        il[3]     = {1.0 / box.len[0], 1.0 / box.len[1], 1.0 / box.len[2]};
    for (LocalIndex g0 = first; g0 < last; g0 += 64)
    {
        auto body = [&](unsigned l) { return std::min(LocalIndex(g0 + l), LocalIndex(last - 1)); };
        unsigned maxLevel = 0;
        for (unsigned l = 0; l < 64; ++l)
        {
            LocalIndex    b    = body(l);
            TreeNodeIndex leaf = TreeNodeIndex(std::upper_bound(layout.begin(), layout.begin() + numLeaves, b) -
                                               layout.begin()) - 1;
            maxLevel           = std::max(maxLevel, treeLevel(leaves[leaf + 1] - leaves[leaf]));
        }
        const float  distCrit = std::ldexp(1.0f, -int(maxLevel)) * tolFactor;
        const double dc2      = double(distCrit) * double(distCrit);
        groups.push_back(g0);
        for (unsigned l = 0; l + 1 < 64; ++l)
        {
            LocalIndex b = body(l), c = body(l + 1);
            double     d[3] = {x[c] * il[0] - x[b] * il[0], y[c] * il[1] - y[b] * il[1], z[c] * il[2] - z[b] * il[2]};
            double     d2   = d[0] * d[0] + (d[1] * d[1] + d[2] * d[2]);
            if (d2 > dc2) { groups.push_back(g0 + l + 1); }
        }
    }
    groups.push_back(last);
}

/* ------------------------------------------------------------------------------------------------
 * Neighbour search (domain/include/cstone/findneighbors.hpp:33-170,
 * traversal/traversal.hpp:52-93, traversal/boxoverlap.hpp:184-215)
 * ----------------------------------------------------------------------------------------------*/
inline double pbcFold(double d, const Box& box, int dim)
{
    return d - double(box.pbc[dim]) * box.len[dim] * std::rint(d * box.ilen[dim]);
}

//! distanceSq<Pbc>(x1.., x2.., box): (x1-x2) folded, then dx*dx + dy*dy + dz*dz
template<bool Pbc>
inline double distSq(double x1, double y1, double z1, double x2, double y2, double z2, const Box& box)
{
    double dx = x1 - x2, dy = y1 - y2, dz = z1 - z2;
    if (Pbc)
    {
        dx = pbcFold(dx, box, 0);
        dy = pbcFold(dy, box, 1);
        dz = pbcFold(dz, box, 2);
    }
    return dx * dx + dy * dy + dz * dz;
}

//! norm2(minDistance(X, center, size[, box])): right-fold dot product (util/array.hpp:237-240)
template<bool Pbc>
inline double cellDistSq(const double* X, const double* c, const double* s, const Box& box)
{
    double d[3];
    for (int k = 0; k < 3; ++k)
    {
        double v = c[k] - X[k];
        if (Pbc) { v = pbcFold(v, box, k); }
        v = std::abs(v) - s[k];
        v += std::abs(v);
        d[k] = v * 0.5;
    }
    return d[0] * d[0] + (d[1] * d[1] + d[2] * d[2]);
}

template<bool Pbc, class Th>
inline unsigned searchOne(LocalIndex i, const double* x, const double* y, const double* z, Th hi, const NsTree& t,
                          const Box& box, unsigned ngmax, LocalIndex* out)
{
    const double X[3]         = {x[i], y[i], z[i]};
    auto         radiusSq     = Th(4.0) * hi * hi;
    auto         cellRadiusSq = radiusSq * 1.0f * 1.0f; // searchExtFactor = 1 (octree.hpp:298)
    unsigned     count        = 0;

    auto overlaps = [&](TreeNodeIndex idx)
    { return cellDistSq<Pbc>(X, &t.centers[3 * idx], &t.sizes[3 * idx], box) < cellRadiusSq; };

    auto scanLeaf = [&](TreeNodeIndex idx)
    {
        TreeNodeIndex leaf = t.oct.internalToLeaf[idx];
        for (LocalIndex j = t.layout[leaf]; j < t.layout[leaf + 1]; ++j)
        {
            if (j == i) { continue; }
            if (distSq<Pbc>(x[j], y[j], z[j], X[0], X[1], X[2], box) < radiusSq)
            {
                if (count < ngmax) { out[count] = j; }
                count++;
            }
        }
    };

    // depth-first walk: children 0..7 in order, leaves handled immediately, internal
    // children pushed on a LIFO stack (traversal.hpp:52-93)
    const TreeNodeIndex* child = t.oct.childOffsets.data();
    if (!overlaps(0)) { return 0; }
    if (child[0] == 0)
    {
        scanLeaf(0);
        return count;
    }
    TreeNodeIndex stack[128];
    stack[0]          = 0;
    TreeNodeIndex top = 1, node = 0;
    do
    {
        for (int oct = 0; oct < 8; ++oct)
        {
            TreeNodeIndex c = child[node] + oct;
            if (overlaps(c))
            {
                if (child[c] == 0) { scanLeaf(c); }
                else { stack[top++] = c; }
            }
        }
        node = stack[--top];
    } while (node != 0);
    return count;
}

//! cstone::findNeighbors(i, ...), findneighbors.hpp:78-147 (count excludes self)
template<class Th>
inline unsigned findNeighbors(LocalIndex i, const double* x, const double* y, const double* z, const Th* h,
                              const NsTree& t, const Box& box, unsigned ngmax, LocalIndex* out)
{
    Th     hi   = h[i];
    double ext  = double(2) * hi;
    bool   in   = true;
    double X[3] = {x[i], y[i], z[i]};
    for (int d = 0; d < 3; ++d)
        in = in && (X[d] - ext >= box.lo[d]) && (X[d] + ext <= box.hi[d]);
    bool usePbc = box.anyPbc() && !in;
    return usePbc ? searchOne<true>(i, x, y, z, hi, t, box, ngmax, out)
                  : searchOne<false>(i, x, y, z, hi, t, box, ngmax, out);
}

//! updateH, sph/include/sph/kernels.hpp:27-32
template<class T>
inline T updateH(unsigned ng0, unsigned nc, T h)
{
    constexpr T c0 = 1023.0;
    constexpr T ex = 1.0 / 10.0;
    return h * T(0.5) * std::pow(T(1) + c0 * ng0 / T(nc), ex);
}

//! findNeighborsSph, sph/include/sph/find_neighbors.hpp:11-44. nc includes self.
//! Returns the number of particles whose h/nc iteration did not converge.
template<class Th>
inline size_t findNeighborsSph(const double* x, const double* y, const double* z, Th* h, LocalIndex first,
                               LocalIndex last, const Box& box, const NsTree& t, unsigned ng0, unsigned ngmax,
                               LocalIndex* neighbors, unsigned* nc, bool iterateH)
{
    unsigned ngmin    = ng0 / 4;
    size_t   numFails = 0;
#pragma omp parallel for reduction(+ : numFails) schedule(dynamic, 256)
    for (LocalIndex k = 0; k < last - first; ++k)
    {
        LocalIndex id = k + first;
        unsigned   n  = 1 + findNeighbors(id, x, y, z, h, t, box, ngmax, neighbors + size_t(k) * ngmax);
        int        it = 0;
        while (iterateH && (ngmin > n || (n - 1) > ngmax) && it++ < 10)
        {
            h[id] = updateH(ng0, n, h[id]);
            n     = 1 + findNeighbors(id, x, y, z, h, t, box, ngmax, neighbors + size_t(k) * ngmax);
        }
        numFails += (it >= 10);
        nc[k] = n;
    }
    return numFails;
}

/* ------------------------------------------------------------------------------------------------
 * SPH kernel tables (sph/include/sph/kernels.hpp:36-61, sph_kernel_tables.hpp:31-101,
 * table_lookup.hpp:10-26, particles_data.hpp:380-387)
 * ----------------------------------------------------------------------------------------------*/
inline double sincHalfPi(double v)
{
    if (v == 0.0) { return 1.0; }
    const double Pv = M_PI_2 * v;
    return std::sin(Pv) / Pv;
}

inline double sincHalfPiDerivative(double v)
{
    if (v == 0.0) return 0.0;
    constexpr double piHalf = M_PI_2;
    const double     Pv     = piHalf * v;
    const double     s      = std::sin(Pv) / (Pv);
    return s * piHalf * ((std::cos(Pv) / std::sin(Pv)) - 1.0 / Pv);
}

inline double sincKernel(double v, double n) { return std::pow(sincHalfPi(v), n); }
inline double sincKernelDerivative(double v, double n)
{
    return n * std::pow(sincHalfPi(v), n - 1) * sincHalfPiDerivative(v);
}

//! composite Simpson rule with sorted partial sums, sph_kernel_tables.hpp:31-60
template<class F>
inline double simpson(double a, double b, uint64_t n, F&& f)
{
    uint64_t            numOdd  = n / 2;
    uint64_t            numEven = (numOdd >= 1) ? numOdd - 1 : 0;
    double              h       = (b - a) / double(n);
    std::vector<double> odd(numOdd), even(numEven);
    for (uint64_t i = 0; i < numOdd; ++i)
        odd[i] = f(a + double(2 * (i + 1) - 1) * h);
    for (uint64_t i = 0; i < numEven; ++i)
        even[i] = f(a + double(2 * (i + 1)) * h);
    std::sort(odd.begin(), odd.end());
    std::sort(even.begin(), even.end());
    return h / 3.0 *
           (f(a) + f(b) + 4.0 * std::accumulate(odd.begin(), odd.end(), 0.0) +
            2.0 * std::accumulate(even.begin(), even.end(), 0.0));
}

//! K = 1 / int_0^2 4 pi v^2 W(v) dv with 2000 Simpson intervals (sph_kernel_tables.hpp:78-83)
inline double kernelNormalization(double sincIndex)
{
    auto vol = [sincIndex](double v) { return 4.0 * M_PI * v * v * sincKernel(v, sincIndex); };
    return 1.0 / simpson(0, 2.0, 2000, vol);
}

//! SPHYNX interpolation constant used by the reference's known-answer tests (sph_kernel_tables.hpp:65-76)
inline double sphynxK(double n)
{
    double b0 = 2.7012593e-2, b1 = 2.0410827e-2, b2 = 3.7451957e-3, b3 = 4.7013839e-2;
    return b0 + b1 * std::sqrt(n) + b2 * n + b3 * std::sqrt(n * n * n);
}

//! tabulateFunction<T, 20000>(f, 0, 2): abscissa computed in T (sph_kernel_tables.hpp:88-101)
template<class T, class F>
inline void tabulate(F&& f, T* table)
{
    const T dx = (2.0 - 0.0) / (kTableSize - 1);
    for (size_t i = 0; i < size_t(kTableSize); ++i)
    {
        T v      = 0.0 + i * dx;
        table[i] = f(v);
    }
}

//! lt::lookup, table_lookup.hpp:10-26
template<class T>
inline T lookup(const T* table, T v)
{
    constexpr int numIntervals = kTableSize - 1;
    constexpr T   support      = 2.0;
    constexpr T   dx           = support / numIntervals;
    constexpr T   invDx        = T(1) / dx;

    int idx        = v * invDx;
    T   derivative = (idx >= numIntervals) ? 0.0 : (table[idx + 1] - table[idx]) * invDx;
    return (idx >= numIntervals) ? 0.0 : table[idx] + derivative * (v - T(idx) * dx);
}

/* ------------------------------------------------------------------------------------------------
 * Periodic images and kernels helpers (sfc/box.hpp:218-250, sph/include/sph/kernels.hpp:11-99)
 * ----------------------------------------------------------------------------------------------*/
//! legacy minimum image: shift by one box length if |d| > r (box.hpp:218-238)
template<class T>
inline void foldLegacy(const Box& box, T r, T& xx, T& yy, T& zz)
{
    if (box.pbc[0] && xx > r) xx -= box.len[0];
    else if (box.pbc[0] && xx < -r) xx += box.len[0];
    if (box.pbc[1] && yy > r) yy -= box.len[1];
    else if (box.pbc[1] && yy < -r) yy += box.len[1];
    if (box.pbc[2] && zz > r) zz -= box.len[2];
    else if (box.pbc[2] && zz < -r) zz += box.len[2];
}

//! distancePBC, box.hpp:241-250: pair distance in T after a legacy fold with r = 2h
template<class T>
inline T pairDistance(const Box& box, T hi, double x1, double y1, double z1, double x2, double y2, double z2)
{
    T xx = x1 - x2, yy = y1 - y2, zz = z1 - z2;
    foldLegacy(box, T(2) * hi, xx, yy, zz);
    return std::sqrt(xx * xx + yy * yy + zz * zz);
}

//! tsKCourant, kernels.hpp:11-16
template<class T1, class T2, class T3>
inline auto courantDt(T1 maxvsignal, T2 h, T3 c, float Kcour)
{
    using T = std::common_type_t<T1, T2, T3>;
    T v     = maxvsignal > T(0) ? maxvsignal : c;
    return T(Kcour * h / v);
}

//! artificial_viscosity, kernels.hpp:71-84 (beta = 2)
template<class T>
inline T avisc(T alpha_i, T alpha_j, T c_i, T c_j, T w_ij)
{
    constexpr T beta = 2.0;
    T           visc = 0.0;
    if (w_ij < 0.0)
    {
        T vsig = (alpha_i + alpha_j) / 4.0 * (c_i + c_j) - beta * w_ij;
        visc   = -vsig * w_ij;
    }
    return visc;
}

/* ------------------------------------------------------------------------------------------------
 * Per-target pair loops ("JLoops") with neighbour lists
 * ----------------------------------------------------------------------------------------------*/

//! xmassJLoop, sph/include/sph/hydro_ve/xmass_kern.hpp:51-79
template<class T, class Tm>
inline T xmassLoop(LocalIndex i, double K, const Box& box, const LocalIndex* nb, unsigned nn, const double* x,
                   const double* y, const double* z, const T* h, const Tm* m, const T* wh)
{
    auto xi = x[i], yi = y[i], zi = z[i];
    auto hi = h[i];
    auto mi = m[i];
    T    hInv  = 1.0 / hi;
    T    h3Inv = hInv * hInv * hInv;
    T    rho0i = mi;
    for (unsigned k = 0; k < nn; ++k)
    {
        LocalIndex j    = nb[k];
        T          dist = pairDistance(box, hi, xi, yi, zi, x[j], y[j], z[j]);
        T          vloc = dist * hInv;
        T          w    = lookup(wh, vloc);
        rho0i += w * m[j];
    }
    auto rhoZero = rho0i * K * h3Inv; // double (K is the coordinate type)
    T    xmassi  = mi / rhoZero;      // veDefinition
    return xmassi;
}

//! IADJLoopSTD (hydro_std/iad_kern.hpp:11-76) and IADJLoop (hydro_ve/iad_kern.hpp:44-109):
//! tau accumulation with volume vol(j), 2^-mean(ilogb) scaling, adjugate inverse.
template<class T, class Vol>
inline void iadLoop(LocalIndex i, double K, const Box& box, const LocalIndex* nb, unsigned nn, const double* x,
                    const double* y, const double* z, const T* h, const T* wh, Vol&& vol, T* c11, T* c12, T* c13,
                    T* c22, T* c23, T* c33)
{
    T    tau11 = 0.0, tau12 = 0.0, tau13 = 0.0, tau22 = 0.0, tau23 = 0.0, tau33 = 0.0;
    auto xi = x[i], yi = y[i], zi = z[i];
    auto hi    = h[i];
    auto hiInv = T(1) / hi;
    for (unsigned k = 0; k < nn; ++k)
    {
        LocalIndex j  = nb[k];
        T          rx = (xi - x[j]), ry = (yi - y[j]), rz = (zi - z[j]);
        foldLegacy(box, T(2) * hi, rx, ry, rz);
        T dist  = std::sqrt(rx * rx + ry * ry + rz * rz);
        T vloc  = dist * hiInv;
        T w     = lookup(wh, vloc);
        T vol_w = vol(j) * w;
        tau11 += rx * rx * vol_w;
        tau12 += rx * ry * vol_w;
        tau13 += rx * rz * vol_w;
        tau22 += ry * ry * vol_w;
        tau23 += ry * rz * vol_w;
        tau33 += rz * rz * vol_w;
    }
    auto getExp = [](T val) { return (val == T(0) ? 0 : std::ilogb(val)); };
    int  expSum = getExp(tau11) + getExp(tau12) + getExp(tau13) + getExp(tau22) + getExp(tau23) + getExp(tau33);
    T    norm   = std::ldexp(T(1), -expSum / 6);
    tau11 *= norm;
    tau12 *= norm;
    tau13 *= norm;
    tau22 *= norm;
    tau23 *= norm;
    tau33 *= norm;
    T det = tau11 * tau22 * tau33 + T(2) * tau12 * tau23 * tau13 - tau11 * tau23 * tau23 - tau22 * tau13 * tau13 -
            tau33 * tau12 * tau12;
    T factor = norm * (hi * hi * hi) / (det * K);
    c11[i]   = (tau22 * tau33 - tau23 * tau23) * factor;
    c12[i]   = (tau13 * tau23 - tau33 * tau12) * factor;
    c13[i]   = (tau12 * tau23 - tau22 * tau13) * factor;
    c22[i]   = (tau11 * tau33 - tau13 * tau13) * factor;
    c23[i]   = (tau13 * tau12 - tau11 * tau23) * factor;
    c33[i]   = (tau11 * tau22 - tau12 * tau12) * factor;
}

//! momentumAndEnergyJLoop (std), hydro_std/momentum_energy_kern.hpp:11-134. `alpha` (may be null)
//! replaces the loop's constant AV coefficient 1 by a per-particle one; null is the reference loop.
template<class T, class Tm>
inline void momentumEnergyStdLoop(LocalIndex i, double K, const Box& box, const LocalIndex* nb, unsigned nn,
                                  const double* x, const double* y, const double* z, const T* vx, const T* vy,
                                  const T* vz, const T* h, const Tm* m, const T* rho, const T* p, const T* c,
                                  const T* c11, const T* c12, const T* c13, const T* c22, const T* c23, const T* c33,
                                  const T* wh, T* ax, T* ay, T* az, double* du, T* maxvsignal,
                                  const T* alpha = nullptr)
{
    constexpr T gradh_i = 1.0, gradh_j = 1.0;
    auto        xi = x[i], yi = y[i], zi = z[i];
    auto        vxi = vx[i], vyi = vy[i], vzi = vz[i];
    auto        hi = h[i], roi = rho[i], pri = p[i], ci = c[i];
    auto        mi_roi = m[i] / rho[i];
    T           hiInv  = T(1) / hi;
    T           hiInv3 = hiInv * hiInv * hiInv;
    T           maxvs  = 0.0;
    T           momx = 0.0, momy = 0.0, momz = 0.0, energy = 0.0;
    auto        c11i = c11[i], c12i = c12[i], c13i = c13[i], c22i = c22[i], c23i = c23[i], c33i = c33[i];

    for (unsigned k = 0; k < nn; ++k)
    {
        LocalIndex j  = nb[k];
        T          rx = xi - x[j], ry = yi - y[j], rz = zi - z[j];
        foldLegacy(box, T(2) * hi, rx, ry, rz);
        T r2    = rx * rx + ry * ry + rz * rz;
        T dist  = std::sqrt(r2);
        T vx_ij = vxi - vx[j], vy_ij = vyi - vy[j], vz_ij = vzi - vz[j];
        T hj    = h[j];
        T hjInv = T(1) / hj;
        T v1    = dist * hiInv;
        T v2    = dist * hjInv;
        T rv    = rx * vx_ij + ry * vy_ij + rz * vz_ij;
        T hjInv3 = hjInv * hjInv * hjInv;
        T Wi     = hiInv3 * lookup(wh, v1);
        T Wj     = hjInv3 * lookup(wh, v2);

        T termA1_i = c11i * rx + c12i * ry + c13i * rz;
        T termA2_i = c12i * rx + c22i * ry + c23i * rz;
        T termA3_i = c13i * rx + c23i * ry + c33i * rz;
        auto c11j = c11[j], c12j = c12[j], c13j = c13[j], c22j = c22[j], c23j = c23[j], c33j = c33[j];
        T termA1_j = c11j * rx + c12j * ry + c13j * rz;
        T termA2_j = c12j * rx + c22j * ry + c23j * rz;
        T termA3_j = c13j * rx + c23j * ry + c33j * rz;

        auto        roj   = rho[j];
        auto        cj    = c[j];
        T           wij   = rv / dist;
        T           alpha_i = alpha ? alpha[i] : T(1), alpha_j = alpha ? alpha[j] : T(1);
        T           visc  = T(0.5) * avisc(alpha_i, alpha_j, ci, cj, wij);
        T           vsig  = ci + cj - T(3) * wij;
        maxvs             = (vsig > maxvs) ? vsig : maxvs;

        auto mj        = m[j];
        auto mj_roj_Wj = mj / roj * Wj;
        T    mj_pro_i  = mj * pri / (gradh_i * roi * roi);
        {
            T a = Wi * (mj_pro_i + visc * mi_roi);
            T b = mj_roj_Wj * (p[j] / (roj * gradh_j) + visc);
            momx += a * termA1_i + b * termA1_j;
            momy += a * termA2_i + b * termA2_j;
            momz += a * termA3_i + b * termA3_j;
        }
        {
            T a = Wi * (T(2) * mj_pro_i + visc * mi_roi);
            T b = visc * mj_roj_Wj;
            energy += vx_ij * (a * termA1_i + b * termA1_j) + vy_ij * (a * termA2_i + b * termA2_j) +
                      vz_ij * (a * termA3_i + b * termA3_j);
        }
    }
    du[i]       = -K * double(0.5) * energy;
    ax[i]       = K * momx;
    ay[i]       = K * momy;
    az[i]       = K * momz;
    *maxvsignal = maxvs;
}

/* ------------------------------------------------------------------------------------------------
 * Adjoints of the std density, IAD and momentum-energy loops (no reference counterpart).
 *
 * Vector-Jacobian products of the functions above with the neighbour lists held fixed: given the
 * cotangent of a loop's outputs for target i, they add dL/d(input) for i and for its neighbours j.
 * The kernel derivative dW/dv is the tabulated whd interpolated like W (the derivative of the
 * smooth kernel, which central differences of the interpolated forward reproduce to O(dx^2)).
 * The periodic fold is a constant shift, so d r_ij / d x_i = -d r_ij / d x_j = identity.
 * Derivations (r = x_i - x_j after the fold, d = |r|, v = d/h_i):
 *   density  rho_i = K h_i^-3 (m_i + sum_j m_j W(v))
 *            d/dr = K h_i^-4 m_j W'(v) r/d,  d/dh_i = -K h_i^-4 (3 rho0_i + sum_j m_j v W'(v))
 *   IAD      C_i = h_i^3/K tau_i^-1,  tau_i = sum_j V_j W(v) r r^T,  V_j = m_j/rho_j
 *            with G the matrix cotangent of C_i (off-diagonal stored entries count twice):
 *            tauBar = -(K/h_i^3) C_i G C_i,  dL/dh_i (direct) = 3/h_i tr(G C_i),
 *            per pair q = r^T tauBar r:  d/dr = V_j (W'(v) q r/(h_i d) + 2 W(v) tauBar r),
 *            d/dh_i = -V_j W'(v) q d/h_i^2,  d/drho_j = -m_j/rho_j^2 W(v) q
 *   momentum-energy: reverse mode through the loop body of momentumEnergyStdLoop, term by term.
 * ----------------------------------------------------------------------------------------------*/

template<class T>
inline void foldedSeparation(const Box& box, T hi, double xi, double yi, double zi, double xj, double yj, double zj,
                             T& rx, T& ry, T& rz)
{
    rx = xi - xj, ry = yi - yj, rz = zi - zj;
    foldLegacy(box, T(2) * hi, rx, ry, rz);
}

//! adjoint of the std density rho_i = m_i / xmassLoop(i): adds into gx, gy, gz (i and j) and gh (i)
template<class T, class Tm>
inline void densityStdVjp(LocalIndex i, double K, const Box& box, const LocalIndex* nb, unsigned nn,
                          const double* x, const double* y, const double* z, const T* h, const Tm* m, const T* wh,
                          const T* whd, T gRho, T* gx, T* gy, T* gz, T* gh)
{
    T hi = h[i], hInv = T(1) / hi;
    T coef = gRho * T(K) * hInv * hInv * hInv * hInv;
    T rho0 = m[i], sumVdW = 0, gxi = 0, gyi = 0, gzi = 0;
    for (unsigned k = 0; k < nn; ++k)
    {
        LocalIndex j = nb[k];
        T          rx, ry, rz;
        foldedSeparation(box, hi, x[i], y[i], z[i], x[j], y[j], z[j], rx, ry, rz);
        T d  = std::sqrt(rx * rx + ry * ry + rz * rz);
        T v  = d * hInv;
        T w  = lookup(wh, v);
        T dw = lookup(whd, v);
        rho0 += m[j] * w;
        sumVdW += m[j] * v * dw;
        if (d > T(0))
        {
            T f = coef * m[j] * dw / d;
            gxi += f * rx, gyi += f * ry, gzi += f * rz;
            gx[j] -= f * rx, gy[j] -= f * ry, gz[j] -= f * rz;
        }
    }
    gx[i] += gxi, gy[i] += gyi, gz[i] += gzi;
    gh[i] += -coef * (T(3) * rho0 + sumVdW);
}

//! adjoint of IADJLoopSTD: cotangent gC (c11..c33 of target i); adds into gx, gy, gz, gh, gRho
template<class T, class Tm>
inline void iadStdVjp(LocalIndex i, double K, const Box& box, const LocalIndex* nb, unsigned nn, const double* x,
                      const double* y, const double* z, const T* h, const Tm* m, const T* rho, const T C6[6],
                      const T G6[6], const T* wh, const T* whd, T* gx, T* gy, T* gz, T* gh, T* gRho)
{
    T hi = h[i], hInv = T(1) / hi;
    // symmetric matrices from the 6 stored entries (c11 c12 c13 c22 c23 c33)
    T C[3][3] = {{C6[0], C6[1], C6[2]}, {C6[1], C6[3], C6[4]}, {C6[2], C6[4], C6[5]}};
    T G[3][3] = {{G6[0], G6[1] / 2, G6[2] / 2}, {G6[1] / 2, G6[3], G6[4] / 2}, {G6[2] / 2, G6[4] / 2, G6[5]}};
    T CG[3][3], B[3][3];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
            CG[a][b] = C[a][0] * G[0][b] + C[a][1] * G[1][b] + C[a][2] * G[2][b];
    T s = -T(K) * hInv * hInv * hInv;
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
            B[a][b] = s * (CG[a][0] * C[0][b] + CG[a][1] * C[1][b] + CG[a][2] * C[2][b]);
    T trGC = G6[0] * C6[0] + G6[1] * C6[1] + G6[2] * C6[2] + G6[3] * C6[3] + G6[4] * C6[4] + G6[5] * C6[5];
    T ghi  = T(3) * hInv * trGC;
    T gxi = 0, gyi = 0, gzi = 0;
    for (unsigned k = 0; k < nn; ++k)
    {
        LocalIndex j = nb[k];
        T          r[3];
        foldedSeparation(box, hi, x[i], y[i], z[i], x[j], y[j], z[j], r[0], r[1], r[2]);
        T d  = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        T v  = d * hInv;
        T w  = lookup(wh, v);
        T dw = lookup(whd, v);
        T Vj = m[j] / rho[j];
        T Br[3];
        for (int a = 0; a < 3; ++a)
            Br[a] = B[a][0] * r[0] + B[a][1] * r[1] + B[a][2] * r[2];
        T q = r[0] * Br[0] + r[1] * Br[1] + r[2] * Br[2];
        T radial = d > T(0) ? Vj * dw * q * hInv / d : T(0);
        T dr[3];
        for (int a = 0; a < 3; ++a)
            dr[a] = radial * r[a] + T(2) * Vj * w * Br[a];
        gxi += dr[0], gyi += dr[1], gzi += dr[2];
        gx[j] -= dr[0], gy[j] -= dr[1], gz[j] -= dr[2];
        ghi += -Vj * dw * q * d * hInv * hInv;
        gRho[j] += -m[j] / (rho[j] * rho[j]) * w * q;
    }
    gx[i] += gxi, gy[i] += gyi, gz[i] += gzi;
    gh[i] += ghi;
}

//! gradient accumulators of the momentum-energy adjoint
template<class T>
struct MomStdGrads
{
    T *x, *y, *z, *h, *rho, *p, *alpha, *c[6];
};

//! adjoint of momentumEnergyStdLoop with per-particle alpha (null = 1): cotangents ga (ax, ay, az)
//! and gdu of target i
template<class T, class Tm>
inline void momentumEnergyStdVjp(LocalIndex i, double K, const Box& box, const LocalIndex* nb, unsigned nn,
                                 const double* x, const double* y, const double* z, const T* vx, const T* vy,
                                 const T* vz, const T* h, const Tm* m, const T* rho, const T* p, const T* c,
                                 const T* const* C6, const T* alpha, const T* wh, const T* whd, const T ga[3], T gdu,
                                 MomStdGrads<T> g)
{
    auto load6 = [C6](LocalIndex k, T out[3][3])
    {
        T c11 = C6[0][k], c12 = C6[1][k], c13 = C6[2][k], c22 = C6[3][k], c23 = C6[4][k], c33 = C6[5][k];
        T M[3][3] = {{c11, c12, c13}, {c12, c22, c23}, {c13, c23, c33}};
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                out[a][b] = M[a][b];
    };
    // scatter a matrix cotangent of C (outer product u r^T, symmetric use) into the 6 stored entries
    auto addC = [](T* const* gc, LocalIndex k, const T u[3], const T r[3], T sign)
    {
        gc[0][k] += sign * u[0] * r[0];
        gc[1][k] += sign * (u[0] * r[1] + u[1] * r[0]);
        gc[2][k] += sign * (u[0] * r[2] + u[2] * r[0]);
        gc[3][k] += sign * u[1] * r[1];
        gc[4][k] += sign * (u[1] * r[2] + u[2] * r[1]);
        gc[5][k] += sign * u[2] * r[2];
    };

    T hi = h[i], hiInv = T(1) / hi, hiInv3 = hiInv * hiInv * hiInv;
    T roi = rho[i], pri = p[i], ci = c[i], mi = m[i];
    T vol_i = mi / roi, pro_i = pri / (roi * roi);
    T alpha_i = alpha ? alpha[i] : T(1);
    T Ci[3][3];
    load6(i, Ci);
    T Ga[3] = {T(K) * ga[0], T(K) * ga[1], T(K) * ga[2]};
    T ge    = -T(K) * T(0.5) * gdu;

    T gx_i = 0, gy_i = 0, gz_i = 0, gh_i = 0, grho_i = 0, gp_i = 0, galpha_i = 0;
    T gC_i[6] = {0, 0, 0, 0, 0, 0};
    T* gCi[6] = {&gC_i[0], &gC_i[1], &gC_i[2], &gC_i[3], &gC_i[4], &gC_i[5]};

    for (unsigned k = 0; k < nn; ++k)
    {
        LocalIndex j = nb[k];
        T          r[3];
        foldedSeparation(box, hi, x[i], y[i], z[i], x[j], y[j], z[j], r[0], r[1], r[2]);
        T d     = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        T vij[3] = {vx[i] - vx[j], vy[i] - vy[j], vz[i] - vz[j]};
        T hj = h[j], hjInv = T(1) / hj, hjInv3 = hjInv * hjInv * hjInv;
        T v1 = d * hiInv, v2 = d * hjInv;
        T W1 = lookup(wh, v1), dW1 = lookup(whd, v1), W2 = lookup(wh, v2), dW2 = lookup(whd, v2);
        T Wi = hiInv3 * W1, Wj = hjInv3 * W2;
        T Cj[3][3];
        load6(j, Cj);
        T tAi[3], tAj[3];
        for (int a = 0; a < 3; ++a)
        {
            tAi[a] = Ci[a][0] * r[0] + Ci[a][1] * r[1] + Ci[a][2] * r[2];
            tAj[a] = Cj[a][0] * r[0] + Cj[a][1] * r[1] + Cj[a][2] * r[2];
        }
        T rv      = r[0] * vij[0] + r[1] * vij[1] + r[2] * vij[2];
        T w       = d > T(0) ? rv / d : T(0);
        T alpha_j = alpha ? alpha[j] : T(1);
        T cj      = c[j];
        T visc = 0, dvisc_dw = 0, dvisc_dalpha = 0;
        if (w < T(0))
        {
            T a4      = (alpha_i + alpha_j) / T(4) * (ci + cj);
            visc      = T(0.5) * (-(a4 - T(2) * w) * w);
            dvisc_dw  = T(0.5) * (-a4 + T(4) * w);
            dvisc_dalpha = T(0.5) * (-(ci + cj) / T(4) * w);
        }
        T mj = m[j], roj = rho[j], vol_j = mj / roj, prj = p[j] / roj;

        T aa = Wi * (mj * pro_i + visc * vol_i);
        T bb = vol_j * Wj * (prj + visc);
        T ae = Wi * (T(2) * mj * pro_i + visc * vol_i);
        T be = visc * vol_j * Wj;

        T s_i = Ga[0] * tAi[0] + Ga[1] * tAi[1] + Ga[2] * tAi[2];
        T s_j = Ga[0] * tAj[0] + Ga[1] * tAj[1] + Ga[2] * tAj[2];
        T e_i = vij[0] * tAi[0] + vij[1] * tAi[1] + vij[2] * tAi[2];
        T e_j = vij[0] * tAj[0] + vij[1] * tAj[1] + vij[2] * tAj[2];

        T aaBar = s_i, bbBar = s_j, aeBar = ge * e_i, beBar = ge * e_j;
        T tAiBar[3], tAjBar[3];
        for (int a = 0; a < 3; ++a)
        {
            tAiBar[a] = aa * Ga[a] + ge * ae * vij[a];
            tAjBar[a] = bb * Ga[a] + ge * be * vij[a];
        }
        T rBar[3];
        for (int a = 0; a < 3; ++a)
            rBar[a] = Ci[0][a] * tAiBar[0] + Ci[1][a] * tAiBar[1] + Ci[2][a] * tAiBar[2] + Cj[0][a] * tAjBar[0] +
                      Cj[1][a] * tAjBar[1] + Cj[2][a] * tAjBar[2];
        addC(gCi, 0, tAiBar, r, T(1));
        addC(g.c, j, tAjBar, r, T(1));

        T WiBar = aaBar * (mj * pro_i + visc * vol_i) + aeBar * (T(2) * mj * pro_i + visc * vol_i);
        T WjBar = bbBar * vol_j * (prj + visc) + beBar * visc * vol_j;
        T proBar_i = (aaBar + T(2) * aeBar) * Wi * mj;
        T viscBar  = (aaBar + aeBar) * Wi * vol_i + (bbBar + beBar) * vol_j * Wj;
        T volBar_i = (aaBar + aeBar) * Wi * visc;
        T volBar_j = bbBar * Wj * (prj + visc) + beBar * visc * Wj;
        T prjBar   = bbBar * vol_j * Wj;

        T wBar = viscBar * dvisc_dw;
        galpha_i += viscBar * dvisc_dalpha;
        g.alpha[j] += viscBar * dvisc_dalpha;
        if (d > T(0))
        {
            T dBar = WiBar * hiInv3 * hiInv * dW1 + WjBar * hjInv3 * hjInv * dW2;
            for (int a = 0; a < 3; ++a)
                rBar[a] += dBar * r[a] / d + wBar * (vij[a] / d - w * r[a] / (d * d));
        }
        gh_i += -WiBar * hiInv3 * hiInv * (T(3) * W1 + v1 * dW1);
        g.h[j] += -WjBar * hjInv3 * hjInv * (T(3) * W2 + v2 * dW2);

        gp_i += proBar_i / (roi * roi);
        grho_i += -T(2) * proBar_i * pri / (roi * roi * roi) - volBar_i * mi / (roi * roi);
        g.p[j] += prjBar / roj;
        g.rho[j] += -volBar_j * mj / (roj * roj) - prjBar * p[j] / (roj * roj);

        gx_i += rBar[0], gy_i += rBar[1], gz_i += rBar[2];
        g.x[j] -= rBar[0], g.y[j] -= rBar[1], g.z[j] -= rBar[2];
    }
    g.x[i] += gx_i, g.y[i] += gy_i, g.z[i] += gz_i, g.h[i] += gh_i;
    g.rho[i] += grho_i, g.p[i] += gp_i, g.alpha[i] += galpha_i;
    for (int q = 0; q < 6; ++q)
        g.c[q][i] += gC_i[q];
}

//! veDefGradhJLoop, hydro_ve/ve_def_gradh_kern.hpp:44-90 -> (kx, gradh)
template<class T, class Tm>
inline void veDefGradhLoop(LocalIndex i, double K, const Box& box, const LocalIndex* nb, unsigned nn, const double* x,
                           const double* y, const double* z, const T* h, const Tm* m, const T* wh, const T* whd,
                           const T* xm, T& kxOut, T& gradhOut)
{
    auto xi = x[i], yi = y[i], zi = z[i];
    auto hi = h[i];
    auto mi = m[i];
    auto xmassi = xm[i];
    auto hInv   = T(1) / hi;
    auto h3Inv  = hInv * hInv * hInv;
    auto kxi      = xmassi;
    auto whomegai = -T(3) * xmassi;
    auto wrho0i   = -T(3) * mi;
    for (unsigned k = 0; k < nn; ++k)
    {
        LocalIndex j     = nb[k];
        T          dist  = pairDistance(box, hi, xi, yi, zi, x[j], y[j], z[j]);
        T          vloc  = dist * hInv;
        T          w     = lookup(wh, vloc);
        T          dw    = lookup(whd, vloc);
        T          dterh = -(T(3) * w + vloc * dw);
        T          xmj   = xm[j];
        kxi += w * xmj;
        whomegai += dterh * xmj;
        wrho0i += dterh * m[j];
    }
    kxi *= K * h3Inv;
    whomegai *= K * h3Inv * hInv;
    wrho0i *= K * h3Inv * hInv;
    whomegai = whomegai * mi / xmassi + (kxi - K * xmassi * h3Inv) * wrho0i;
    T rhoi   = kxi * mi / xmassi;
    T dhdrho = -hi / (rhoi * T(3));
    gradhOut = T(1) - dhdrho * whomegai;
    kxOut    = kxi;
}

//! divV_curlVJLoop, hydro_ve/divv_curlv_kern.hpp:44-123
template<class T>
inline void divvCurlvLoop(LocalIndex i, double K, const Box& box, const LocalIndex* nb, unsigned nn, const double* x,
                          const double* y, const double* z, const T* vx, const T* vy, const T* vz, const T* h,
                          const T* c11, const T* c12, const T* c13, const T* c22, const T* c23, const T* c33,
                          const T* wh, const T* kx, const T* xm, T* divv, T* curlv, T* dV11, T* dV12, T* dV13, T* dV22,
                          T* dV23, T* dV33, bool doGradV)
{
    auto xi = x[i], yi = y[i], zi = z[i];
    auto vxi = vx[i], vyi = vy[i], vzi = vz[i];
    auto hi = h[i], kxi = kx[i];
    auto hiInv  = T(1) / hi;
    auto hiInv3 = hiInv * hiInv * hiInv;
    T    gX[3] = {0., 0., 0.}, gY[3] = {0., 0., 0.}, gZ[3] = {0., 0., 0.};
    auto c11i = c11[i], c12i = c12[i], c13i = c13[i], c22i = c22[i], c23i = c23[i], c33i = c33[i];

    for (unsigned k = 0; k < nn; ++k)
    {
        LocalIndex j  = nb[k];
        T          rx = xi - x[j], ry = yi - y[j], rz = zi - z[j];
        foldLegacy(box, T(2) * hi, rx, ry, rz);
        T r2    = rx * rx + ry * ry + rz * rz;
        T dist  = std::sqrt(r2);
        T vx_ji = vx[j] - vxi, vy_ji = vy[j] - vyi, vz_ji = vz[j] - vzi;
        T v1    = dist * hiInv;
        T Wi    = lookup(wh, v1);
        T tA[3];
        tA[0]   = -(c11i * rx + c12i * ry + c13i * rz) * Wi;
        tA[1]   = -(c12i * rx + c22i * ry + c23i * rz) * Wi;
        tA[2]   = -(c13i * rx + c23i * ry + c33i * rz) * Wi;
        T xmj   = xm[j];
        T sx = vx_ji * xmj, sy = vy_ji * xmj, sz = vz_ji * xmj;
        for (int d = 0; d < 3; ++d)
        {
            gX[d] += sx * tA[d];
            gY[d] += sy * tA[d];
            gZ[d] += sz * tA[d];
        }
    }
    T norm_kxi = K * hiInv3 / kxi;
    divv[i]    = norm_kxi * (gX[0] + gY[1] + gZ[2]);
    if (curlv != nullptr)
    {
        T cv[3] = {gZ[1] - gY[2], gX[2] - gZ[0], gY[0] - gX[1]};
        curlv[i] = norm_kxi * std::sqrt(cv[0] * cv[0] + (cv[1] * cv[1] + cv[2] * cv[2]));
    }
    if (doGradV)
    {
        dV11[i] = norm_kxi * gX[0];
        dV12[i] = norm_kxi * (gX[1] + gY[0]);
        dV13[i] = norm_kxi * (gX[2] + gZ[0]);
        dV22[i] = norm_kxi * gY[1];
        dV23[i] = norm_kxi * (gY[2] + gZ[1]);
        dV33[i] = norm_kxi * gZ[2];
    }
}

//! AVswitchesJLoop, hydro_ve/av_switches_kern.hpp:44-137 -> new alpha_i
template<class T>
inline T avSwitchesLoop(LocalIndex i, double K, const Box& box, const LocalIndex* nb, unsigned nn, const double* x,
                        const double* y, const double* z, const T* vx, const T* vy, const T* vz, const T* h,
                        const T* c, const T* c11, const T* c12, const T* c13, const T* c22, const T* c23,
                        const T* c33, const T* wh, const T* kx, const T* xm, const T* divv, const double dt,
                        const T alphamin, const T alphamax, const T decay_constant, T alpha_i)
{
    auto xi = x[i], yi = y[i], zi = z[i];
    auto vxi = vx[i], vyi = vy[i], vzi = vz[i];
    auto hi = h[i], ci = c[i];
    auto c11i = c11[i], c12i = c12[i], c13i = c13[i], c22i = c22[i], c23i = c23[i], c33i = c33[i];
    T    vsig_i = T(1.e-40) * ci;
    auto hiInv  = T(1) / hi;
    auto hiInv3 = hiInv * hiInv * hiInv;
    auto divv_i = divv[i];
    T    gdx = 0.0, gdy = 0.0, gdz = 0.0;

    for (unsigned k = 0; k < nn; ++k)
    {
        LocalIndex j  = nb[k];
        T          rx = xi - x[j], ry = yi - y[j], rz = zi - z[j];
        foldLegacy(box, T(2) * hi, rx, ry, rz);
        T r2    = rx * rx + ry * ry + rz * rz;
        T dist  = std::sqrt(r2);
        T vx_ij = vxi - vx[j], vy_ij = vyi - vy[j], vz_ij = vzi - vz[j];
        T rv    = rx * vx_ij + ry * vy_ij + rz * vz_ij;
        T vs_ij = 0.0;
        if (rv < T(0)) { vs_ij = ci + c[j] - T(3) * rv / dist; }
        vsig_i = std::max(vsig_i, vs_ij);
        T v1   = dist * hiInv;
        T Wi   = K * hiInv3 * lookup(wh, v1);
        T tA1  = -(c11i * rx + c12i * ry + c13i * rz) * Wi;
        T tA2  = -(c12i * rx + c22i * ry + c23i * rz) * Wi;
        T tA3  = -(c13i * rx + c23i * ry + c33i * rz) * Wi;
        T volj = xm[j] / kx[j];
        T fac  = volj * (divv_i - divv[j]);
        gdx += fac * tA1;
        gdy += fac * tA2;
        gdz += fac * tA3;
    }
    T graddivv = std::sqrt(gdx * gdx + gdy * gdy + gdz * gdz);
    T alphaloc = 0.0;
    if (divv_i < T(0))
    {
        T a_const = hi * hi * graddivv;
        alphaloc  = alphamax * a_const / (a_const + hi * std::abs(divv_i) + T(0.05) * ci);
    }
    if (alphaloc >= alpha_i) { alpha_i = alphaloc; }
    else
    {
        T decay    = hi / (decay_constant * vsig_i);
        T alphadot = 0.0;
        if (alphaloc >= alphamin) { alphadot = (alphaloc - alpha_i) / decay; }
        else { alphadot = (alphamin - alpha_i) / decay; }
        alpha_i += alphadot * dt;
    }
    return alpha_i;
}

//! symv with the reference's upper-triangular row layout (kernels.hpp:87-96) and the AV cleaning
//! correction avRvCorrection (hydro_ve/momentum_energy_kern.hpp:44-63)
template<class T>
inline T avRvCorrection(const T R[3], T eta_ab, T eta_crit, const T gi[6], const T gj[6])
{
    auto symvDot = [&R](const T* g)
    {
        T s0 = g[0] * R[0] + g[1] * R[1] + g[2] * R[2];
        T s1 = g[3] * R[1] + g[4] * R[2];
        T s2 = g[5] * R[2];
        return R[0] * s0 + (R[1] * s1 + R[2] * s2);
    };
    T dmy1 = symvDot(gi);
    T dmy2 = symvDot(gj);
    T dmy3 = T(1);
    if (eta_ab < eta_crit)
    {
        T etaDiff = T(5) * (eta_ab - eta_crit);
        dmy3      = std::exp(-etaDiff * etaDiff);
    }
    T A_ab   = (dmy2 != T(0)) ? dmy1 / dmy2 : T(0);
    T A_abp1 = T(1) + A_ab;
    T phi_ab = T(0.5) * dmy3 * std::max(T(0), std::min(T(1), T(4) * A_ab / (A_abp1 * A_abp1)));
    return -phi_ab * (dmy1 + dmy2);
}

//! momentumAndEnergyJLoop<avClean> (VE), hydro_ve/momentum_energy_kern.hpp:66-222
template<bool avClean, class T, class Tm>
inline void momentumEnergyVeLoop(LocalIndex i, double K, const Box& box, const LocalIndex* nb, unsigned nn,
                                 const double* x, const double* y, const double* z, const T* vx, const T* vy,
                                 const T* vz, const T* h, const Tm* m, const T* prho, const T* tdpdTrho, const T* c,
                                 const T* c11, const T* c12, const T* c13, const T* c22, const T* c23, const T* c33,
                                 const T Atmin, const T Atmax, const T ramp, const T* wh, const T* kx, const T* xm,
                                 const T* alpha, const T* dV11, const T* dV12, const T* dV13, const T* dV22,
                                 const T* dV23, const T* dV33, T* ax, T* ay, T* az, double* du, T* maxvsignal)
{
    auto xi = x[i], yi = y[i], zi = z[i];
    auto vxi = vx[i], vyi = vy[i], vzi = vz[i];
    auto hi = h[i];
    auto mi = m[i];
    auto ci = c[i];
    auto kxi = kx[i];
    auto alpha_i = alpha[i];
    auto xmassi  = xm[i];
    auto rhoi    = kxi * mi / xmassi;
    auto prhoi   = prho[i];
    T    hiInv   = T(1) / hi;
    T    hiInv3  = hiInv * hiInv * hiInv;
    T    maxvs   = 0.0;
    T    momx = 0.0, momy = 0.0, momz = 0.0, energy = 0.0;
    T    avEnergy = 0.0;
    auto c11i = c11[i], c12i = c12[i], c13i = c13[i], c22i = c22[i], c23i = c23[i], c33i = c33[i];
    T    gradV_i[6] = {0, 0, 0, 0, 0, 0};
    if (avClean) { gradV_i[0] = dV11[i], gradV_i[1] = dV12[i], gradV_i[2] = dV13[i], gradV_i[3] = dV22[i],
                   gradV_i[4] = dV23[i], gradV_i[5] = dV33[i]; }
    T eta_crit = std::cbrt(T(32) * M_PI / T(3) / T(nn + 1));

    for (unsigned k = 0; k < nn; ++k)
    {
        LocalIndex j  = nb[k];
        T          rx = xi - x[j], ry = yi - y[j], rz = zi - z[j];
        auto       vxj = vx[j], vyj = vy[j], vzj = vz[j];
        foldLegacy(box, T(2) * hi, rx, ry, rz);
        T r2     = rx * rx + ry * ry + rz * rz;
        T dist   = std::sqrt(r2);
        T vx_ij  = vxi - vxj, vy_ij = vyi - vyj, vz_ij = vzi - vzj;
        T hj     = h[j];
        T hjInv  = T(1) / hj;
        T v1     = dist * hiInv;
        T v2     = dist * hjInv;
        T hjInv3 = hjInv * hjInv * hjInv;
        T Wi     = hiInv3 * lookup(wh, v1);
        T Wj     = hjInv3 * lookup(wh, v2);

        T tA1_i = -(c11i * rx + c12i * ry + c13i * rz) * Wi;
        T tA2_i = -(c12i * rx + c22i * ry + c23i * rz) * Wi;
        T tA3_i = -(c13i * rx + c23i * ry + c33i * rz) * Wi;
        auto c11j = c11[j], c12j = c12[j], c13j = c13[j], c22j = c22[j], c23j = c23[j], c33j = c33[j];
        T tA1_j = -(c11j * rx + c12j * ry + c13j * rz) * Wj;
        T tA2_j = -(c12j * rx + c22j * ry + c23j * rz) * Wj;
        T tA3_j = -(c13j * rx + c23j * ry + c33j * rz) * Wj;

        auto mj  = m[j];
        auto cj  = c[j];
        auto kxj = kx[j];
        auto xmassj = xm[j];
        auto rhoj   = kxj * mj / xmassj;

        T rv = rx * vx_ij + ry * vy_ij + rz * vz_ij;
        if (avClean)
        {
            T R[3]      = {rx, ry, rz};
            T gradV_j[6] = {dV11[j], dV12[j], dV13[j], dV22[j], dV23[j], dV33[j]};
            rv += avRvCorrection(R, std::min(v1, v2), eta_crit, gradV_i, gradV_j);
        }
        T wij  = rv / dist;
        T visc = avisc(alpha_i, alpha[j], ci, cj, wij);
        T vsig = T(0.5) * (ci + cj) - T(2) * wij;
        maxvs  = (vsig > maxvs) ? vsig : maxvs;

        T a_mom, b_mom;
        T Atwood = (std::abs(rhoi - rhoj)) / (rhoi + rhoj);
        if (Atwood < Atmin)
        {
            a_mom = xmassi * xmassi;
            b_mom = xmassj * xmassj;
        }
        else if (Atwood > Atmax)
        {
            a_mom = xmassi * xmassj;
            b_mom = a_mom;
        }
        else
        {
            T sigma_ij = ramp * (Atwood - Atmin);
            // unqualified pow in the reference resolves to the C library's double pow
            a_mom = ::pow(xmassi, T(2) - sigma_ij) * ::pow(xmassj, sigma_ij);
            b_mom = ::pow(xmassj, T(2) - sigma_ij) * ::pow(xmassi, sigma_ij);
        }

        auto a_visc   = mj / rhoi * visc;
        auto b_visc   = mj / rhoj * visc;
        T    a_visc_x = T(0.5) * (a_visc * tA1_i + b_visc * tA1_j);
        T    a_visc_y = T(0.5) * (a_visc * tA2_i + b_visc * tA2_j);
        T    a_visc_z = T(0.5) * (a_visc * tA3_i + b_visc * tA3_j);
        avEnergy += a_visc_x * vx_ij + a_visc_y * vy_ij + a_visc_z * vz_ij;

        energy += mj * a_mom * (vx_ij * tA1_i + vy_ij * tA2_i + vz_ij * tA3_i);

        auto mom_i = mj * prhoi * a_mom;
        auto mom_j = mj * prho[j] * b_mom;
        momx += mom_i * tA1_i + mom_j * tA1_j + a_visc_x;
        momy += mom_i * tA2_i + mom_j * tA2_j + a_visc_y;
        momz += mom_i * tA3_i + mom_j * tA3_j + a_visc_z;
    }
    avEnergy = std::max(T(0), avEnergy);
    T eCoeff = (tdpdTrho == nullptr) ? prhoi : tdpdTrho[i];
    du[i]    = K * (eCoeff * energy + T(0.5) * avEnergy);
    ax[i]    = -K * momx;
    ay[i]    = -K * momy;
    az[i]    = -K * momz;
    *maxvsignal = maxvs;
}

/* ------------------------------------------------------------------------------------------------
 * Equations of state (sph/include/sph/eos.hpp:19-86)
 * ----------------------------------------------------------------------------------------------*/
template<class T1, class T2>
inline T1 idealGasCv(T1 mui, T2 gamma)
{
    constexpr T1 R = 8.317e7;
    return R / mui / (gamma - T1(1));
}

template<class T1, class T2, class T3>
inline void idealGasFromU(T1 u, T2 rho, T3 gamma, std::common_type_t<T1, T2, T3>& p,
                          std::common_type_t<T1, T2, T3>& c)
{
    using Tc = std::common_type_t<T1, T2, T3>;
    Tc tmp   = u * (gamma - Tc(1));
    p        = rho * tmp;
    c        = std::sqrt(gamma * tmp);
}

//! isothermalEOS, eos.hpp:58-64: p = rho c c in the common type
template<class T1, class T2>
inline std::common_type_t<T1, T2> isothermalP(T1 c, T2 rho)
{
    using Tc = std::common_type_t<T1, T2>;
    Tc p     = rho * c * c;
    return p;
}

//! polytropicEOS, eos.hpp:77-86: p = K rho^n, c = sqrt(n p / rho) in the common type
template<class T1, class T2, class T3>
inline void polytropicPC(T1 K, T2 n, T3 rho, std::common_type_t<T1, T2, T3>& p, std::common_type_t<T1, T2, T3>& c)
{
    p = K * std::pow(rho, n);
    c = std::sqrt(n * p / rho);
}

/* ------------------------------------------------------------------------------------------------
 * Time integration (sph/include/sph/positions.hpp, positions_gpu.cu, update_h.hpp, ts_global.hpp).
 * Follows the reference's GPU kernel and its instantiated types (positions_gpu.cu:116-193):
 * positions, temperature / internal energy, du in double; velocities, accelerations, h, mui and
 * the previous-step fields x_m1, y_m1, z_m1, du_m1 in the HydroType T.
 * ----------------------------------------------------------------------------------------------*/

//! energyUpdate, positions.hpp:54-61 (Adams-Bashforth 2nd order)
template<class TU>
inline TU energyUpdate(TU u_old, double dt, double dt_m1, double du, double du_m1)
{
    TU u_new = u_old + du * dt + 0.5 * (du - du_m1) / dt_m1 * std::abs(dt) * dt;
    if (u_new < 0.) { u_new = u_old * std::exp(u_new * dt / u_old); }
    return u_new;
}

//! putInBox, domain/include/cstone/sfc/box.hpp:193-214
inline void putInBox(double X[3], const Box& box)
{
    for (int d = 0; d < 3; ++d)
    {
        if (box.pbc[d] && X[d] > box.hi[d]) { X[d] -= box.len[d]; }
        else if (box.pbc[d] && X[d] < box.lo[d]) { X[d] += box.len[d]; }
    }
}

/*! makeGlobalBox (sfc/box_mpi.hpp:69-100) on one rank. limits6 = xmin xmax ymin ymax zmin zmax.
 * A periodic dimension keeps the previous limits; any other takes the min and max of the
 * coordinates, accumulated with std::min / std::max from +INFINITY / -INFINITY (MinMax,
 * box_mpi.hpp:36-52). */
inline void fittingBox(const double* x, const double* y, const double* z, size_t n, const int bc[3],
                       const double prev[6], double out[6])
{
    const double* c[3] = {x, y, z};
    for (int d = 0; d < 3; ++d)
    {
        if (bc[d] == 1)
        {
            out[2 * d] = prev[2 * d], out[2 * d + 1] = prev[2 * d + 1];
            continue;
        }
        double lo = INFINITY, hi = -INFINITY;
        for (size_t i = 0; i < n; ++i)
        {
            lo = std::min(lo, c[d][i]);
            hi = std::max(hi, c[d][i]);
        }
        out[2 * d] = lo, out[2 * d + 1] = hi;
    }
}

/*! limitBoxShrinking (sfc/box.hpp:398-412): each side moves inwards by at most shrinkLimit times
 * the previous edge length (the length as the Box constructor forms it, xmax - xmin). */
inline void limitBoxShrinking(const double fit[6], const double prev[6], double shrinkLimit, double out[6])
{
    for (int d = 0; d < 3; ++d)
    {
        double len   = prev[2 * d + 1] - prev[2 * d];
        double step  = shrinkLimit * len;
        double lower = prev[2 * d] + step;
        double upper = prev[2 * d + 1] - step;
        out[2 * d]     = std::min(fit[2 * d], lower);
        out[2 * d + 1] = std::max(fit[2 * d + 1], upper);
    }
}

/*! The box update of GlobalAssignment::assign (domain/assignment.hpp:80-82): the fitting box on the
 * first call, afterwards the fitting box with limited shrinking against the previous one. */
inline void globalBox(const double* x, const double* y, const double* z, size_t n, const int bc[3], bool firstCall,
                      double limits6[6])
{
    double fit[6];
    fittingBox(x, y, z, n, bc, limits6, fit);
    if (firstCall) { std::copy(fit, fit + 6, limits6); }
    else
    {
        double prev[6];
        std::copy(limits6, limits6 + 6, prev);
        limitBoxShrinking(fit, prev, 0.05, limits6);
    }
}

//! positionUpdate, positions.hpp:63-85 (Press 2nd order): (Xn, An, dXn) -> (Xn+1, Vn+1, dXn+1)
inline void positionUpdate(double dt, double dt_m1, const double Xn[3], const double An[3], const double dXn[3],
                           const Box& box, double Xnp1[3], double Vnp1[3], double dXnp1[3])
{
    for (int d = 0; d < 3; ++d)
    {
        double Vnmhalf = dXn[d] * (1.0 / dt_m1);
        double Vn      = Vnmhalf + 0.5 * dt_m1 * An[d];
        Vnp1[d]        = Vn + An[d] * dt;
        dXnp1[d]       = (Vn + 0.5 * An[d] * std::abs(dt)) * dt;
        Xnp1[d]        = Xn[d] + dXnp1[d];
    }
    putInBox(Xnp1, box);
}

//! fbcCheck, positions.hpp:46-51
template<class Th>
inline bool fbcCheck(double coord, Th h, double top, double bottom, bool fbc)
{
    return fbc && (std::abs(top - coord) < Th(2) * h || std::abs(bottom - coord) < Th(2) * h);
}

/*! computePositionsKernel, positions_gpu.cu:116-168, over [first, last). constCv < 0 selects the
 * per-particle idealGasCv(mui[i], gamma). temp (when non-null) takes precedence over u. With a rung
 * array (block time steps), particle i takes the previous step dtM1Rung[rung[i]] (:140). */
template<class T>
inline void computePositions(const Box& box, double dt, double dt_m1_all, double* x, double* y, double* z, T* vx,
                             T* vy, T* vz, T* x_m1, T* y_m1, T* z_m1, const T* ax, const T* ay, const T* az,
                             double* temp, double* u, const double* du, T* du_m1, const T* h, const T* mui,
                             double gamma, double constCv, size_t first, size_t last, const float* dtM1Rung = nullptr,
                             const uint8_t* rung = nullptr)
{
    bool anyFBC = box.fbc[0] || box.fbc[1] || box.fbc[2];
    for (size_t i = first; i < last; ++i)
    {
        if (anyFBC && vx[i] == T(0) && vy[i] == T(0) && vz[i] == T(0))
        {
            if (fbcCheck(x[i], h[i], box.hi[0], box.lo[0], box.fbc[0]) ||
                fbcCheck(y[i], h[i], box.hi[1], box.lo[1], box.fbc[1]) ||
                fbcCheck(z[i], h[i], box.hi[2], box.lo[2], box.fbc[2]))
            {
                continue;
            }
        }
        const double dt_m1 = rung ? double(dtM1Rung[rung[i]]) : dt_m1_all;
        double A[3]  = {double(ax[i]), double(ay[i]), double(az[i])};
        double X[3]  = {x[i], y[i], z[i]};
        double dX[3] = {double(x_m1[i]), double(y_m1[i]), double(z_m1[i])};
        double Xn1[3], V[3], dXn1[3];
        positionUpdate(dt, dt_m1, X, A, dX, box, Xn1, V, dXn1);
        x[i] = Xn1[0], y[i] = Xn1[1], z[i] = Xn1[2];
        x_m1[i] = T(dXn1[0]), y_m1[i] = T(dXn1[1]), z_m1[i] = T(dXn1[2]);
        vx[i] = T(V[0]), vy[i] = T(V[1]), vz[i] = T(V[2]);

        if (temp != nullptr)
        {
            T      cv    = (constCv < 0) ? idealGasCv(mui[i], gamma) : constCv;
            double u_old = temp[i] * cv;
            temp[i]      = energyUpdate(u_old, dt, dt_m1, du[i], du_m1[i]) / cv;
            du_m1[i]     = du[i];
        }
        else if (u != nullptr)
        {
            u[i]     = energyUpdate(u[i], dt, dt_m1, du[i], du_m1[i]);
            du_m1[i] = du[i];
        }
    }
}

/*! driftKernel, positions_gpu.cu:17-60, over [first, last): a particle of a block time-step
 * hierarchy, last advanced by dtBack from (Xn, An, dXn), is taken back to Xn (positionUpdate with
 * -dtBack) and advanced by dt from there, both without the periodic wrap (the reference drifts in
 * an open unit box, so that the octree stays valid); velocities follow, and temp or u is recovered
 * and advanced the same way with energyUpdate. x_m1, y_m1, z_m1 and du_m1 are read only. The
 * previous step is dtM1Rung[rung[i]], or dtM1Rung[0] without a rung array. */
template<class T>
inline void driftPositions(double dt, double dtBack, const float* dtM1Rung, const uint8_t* rung, double* x, double* y,
                           double* z, T* vx, T* vy, T* vz, const T* x_m1, const T* y_m1, const T* z_m1, const T* ax,
                           const T* ay, const T* az, double* temp, double* u, const double* du, const T* du_m1,
                           const T* mui, double gamma, double constCv, size_t first, size_t last)
{
    Box noPbc{};
    for (int d = 0; d < 3; ++d)
    {
        noPbc.lo[d] = 0, noPbc.hi[d] = 1, noPbc.len[d] = 1, noPbc.ilen[d] = 1, noPbc.pbc[d] = 0, noPbc.fbc[d] = 0;
    }
    for (size_t i = first; i < last; ++i)
    {
        const double dt_m1 = double(rung ? dtM1Rung[rung[i]] : dtM1Rung[0]);
        double A[3]  = {double(ax[i]), double(ay[i]), double(az[i])};
        double Xb[3] = {x[i], y[i], z[i]};
        double dX[3] = {double(x_m1[i]), double(y_m1[i]), double(z_m1[i])};
        double Xn[3], V[3], dXn1[3];
        positionUpdate(-dtBack, dt_m1, Xb, A, dX, noPbc, Xn, V, dXn1);
        double Xn1[3];
        positionUpdate(dt, dt_m1, Xn, A, dX, noPbc, Xn1, V, dXn1);
        x[i] = Xn1[0], y[i] = Xn1[1], z[i] = Xn1[2];
        vx[i] = T(V[0]), vy[i] = T(V[1]), vz[i] = T(V[2]);

        if (temp != nullptr)
        {
            T      cv      = (constCv < 0) ? idealGasCv(mui[i], gamma) : constCv;
            double u_recov = energyUpdate(double(temp[i] * cv), -dtBack, dt_m1, du[i], double(du_m1[i]));
            temp[i]        = energyUpdate(u_recov, dt, dt_m1, du[i], double(du_m1[i])) / cv;
        }
        else if (u != nullptr)
        {
            double u_recov = energyUpdate(u[i], -dtBack, dt_m1, du[i], double(du_m1[i]));
            u[i]           = energyUpdate(u_recov, dt, dt_m1, du[i], double(du_m1[i]));
        }
    }
}

/*! groupDivvKernel, ts_groups.cu:16-36: per group the largest divv (signed, in float), then
 * groupDt[g] = min(groupDt[g], Krho / |max divv|) in float */
template<class T>
inline void groupDivvTimestep(float Krho, const uint32_t* groupStart, const uint32_t* groupEnd, size_t numGroups,
                              const T* divv, float* groupDt)
{
    for (size_t g = 0; g < numGroups; ++g)
    {
        float localMax = -INFINITY;
        for (uint32_t i = groupStart[g]; i < groupEnd[g]; ++i)
            localMax = float(std::max<T>(T(localMax), divv[i]));
        groupDt[g] = std::min(groupDt[g], Krho / std::abs(localMax));
    }
}

/*! groupAccKernel, ts_groups.cu:51-71: per group the largest |a|^2 (norm2, util/array.hpp:236-240
 * and 316-320: the right fold ax^2 + (ay^2 + az^2) in T, as float), then
 * groupDt[g] = min(groupDt[g], etaAcc / sqrt(sqrt(max |a|^2))) in float */
template<class T>
inline void groupAccTimestep(float etaAcc, const uint32_t* groupStart, const uint32_t* groupEnd, size_t numGroups,
                             const T* ax, const T* ay, const T* az, float* groupDt)
{
    for (size_t g = 0; g < numGroups; ++g)
    {
        float maxAcc = 0;
        for (uint32_t i = groupStart[g]; i < groupEnd[g]; ++i)
            maxAcc = float(std::max<T>(T(maxAcc), T(ax[i] * ax[i] + (ay[i] * ay[i] + az[i] * az[i]))));
        groupDt[g] = std::min(groupDt[g], etaAcc / std::sqrt(std::sqrt(maxAcc)));
    }
}

//! storeRungKernel, ts_groups.cu:84-95: particleRungs[i] = rung over every group
inline void storeRung(const uint32_t* groupStart, const uint32_t* groupEnd, size_t numGroups, uint8_t rung,
                      uint8_t* particleRungs)
{
    for (size_t g = 0; g < numGroups; ++g)
        for (uint32_t i = groupStart[g]; i < groupEnd[g]; ++i)
            particleRungs[i] = rung;
}

//! updateSmoothingLengthCpu, update_h.hpp:12-24
template<class T>
inline void updateSmoothingLength(size_t first, size_t last, unsigned ng0, const uint32_t* nc, T* h)
{
    for (size_t i = first; i < last; ++i)
        h[i] = updateH(ng0, nc[i], h[i]);
}

//! maxNormSquareGpu's element (primitives_gpu.cu:179-185): x*x + y*y + z*z in T
template<class T>
inline double maxNormSquare(const T* x, const T* y, const T* z, size_t first, size_t last)
{
    T m = 0;
    for (size_t i = first; i < last; ++i)
        m = std::max(m, T(x[i] * x[i] + y[i] * y[i] + z[i] * z[i]));
    return double(m);
}

//! computeTimestep, ts_global.hpp:95-113, on one rank; dtAcc from accelerationTimestep (:46-67)
inline void computeTimestep(double maxAccSq, double& minDt, double& minDt_m1, double& ttot, double minDtCourant,
                            double minDtRho, double maxDtIncrease, double g, double etaAcc, double eps)
{
    double minDtAcc = (g != 0.0) ? etaAcc * std::sqrt(eps / std::sqrt(maxAccSq)) : INFINITY;
    double dt       = std::min({minDtAcc, minDtCourant, minDtRho, maxDtIncrease * minDt});
    ttot += dt;
    minDt_m1 = minDt;
    minDt    = dt;
}

/* ------------------------------------------------------------------------------------------------
 * Conserved quantities (main/src/observables/conserved_quantities.hpp:50-107), with the GPU path's
 * shared cv for temperature (conserved_gpu.cu:71-95): eKin = 1/2 sum m |v|^2, eInt, linmom, angmom.
 * ----------------------------------------------------------------------------------------------*/
template<class T>
inline void conservedQuantities(const double* x, const double* y, const double* z, const T* vx, const T* vy,
                                const T* vz, const T* m, const double* u, const double* temp, double cv,
                                const uint32_t* nc, size_t first, size_t last, double out[9])
{
    double eKin = 0, eInt = 0, lin[3] = {0, 0, 0}, ang[3] = {0, 0, 0}, ncsum = 0;
    for (size_t i = first; i < last; ++i)
    {
        double X[3] = {x[i], y[i], z[i]};
        double V[3] = {double(vx[i]), double(vy[i]), double(vz[i])};
        double mi   = double(m[i]);
        eKin += mi * (V[0] * V[0] + (V[1] * V[1] + V[2] * V[2]));
        for (int d = 0; d < 3; ++d)
            lin[d] += mi * V[d];
        ang[0] += mi * (X[1] * V[2] - X[2] * V[1]);
        ang[1] += mi * (X[2] * V[0] - X[0] * V[2]);
        ang[2] += mi * (X[0] * V[1] - X[1] * V[0]);
        if (temp) { eInt += mi * temp[i]; }
        else if (u) { eInt += mi * u[i]; }
        if (nc) { ncsum += double(nc[i]); }
    }
    out[0] = 0.5 * eKin;
    out[1] = temp ? cv * eInt : eInt;
    for (int d = 0; d < 3; ++d)
        out[2 + d] = lin[d], out[5 + d] = ang[d];
    out[8] = ncsum;
}

} // namespace orc
