// Neighbour search on the cornerstone octree: one wavefront per group of 64 SFC-consecutive
// targets, with the coupled smoothing-length / neighbour-count iteration.
//
// Per group the wave walks the octree once with the group's bounding box grown by the largest
// search radius 2*h_max (exact periodic box-box distance, so no cell holding a neighbour of any
// target is skipped), collecting candidate leaves in LDS; internal nodes live on an LDS stack
// popped 64 at a time, one node per lane, compacted with ballot/mbcnt. The stack is kept in the
// visiting order of the reference's per-target traversal (leaf children in octant order, then
// internal children in reverse octant order, depth first), and found leaves are swept in that
// order, so every list holds its neighbours in the reference's order. Candidate leaves that no
// target's sphere reaches are dropped (one leaf per lane stages its box, moved to the periodic
// image the group sees, in LDS; the wave reads each back as a broadcast, one ballot per leaf).
// The particles of the surviving leaves within reach of the group box are compacted into full
// tiles of 64 in a double-buffered LDS area (coalesced loads of the next window overlap the
// sweep), and the wave tests each tile reading every candidate with one broadcast LDS access,
// every lane against its own target with the reference predicate
// (domain/include/cstone/findneighbors.hpp:33-60,93,116-140): d2 = dx*dx + dy*dy + dz*dz with
// dx = x_j - x_i, periodic images folded by one box length, unfused FP64 arithmetic, keep if
// d2 < Th(4)*h_i*h_i and j != i. The lists, counts and, past ngmax, the kept first ngmax hits
// are therefore bit-identical to the CPU search; they are stored in the group layout of
// common.hpp (listOffset): a lane
// shifts its entries through a register quad and stores 16 bytes at a time, and the pair loops read
// four slots of all 64 lanes with one coalesced load.
#include <cstring>

#include "pairs.hpp"

namespace sphx
{

namespace
{

// one wave per block: a group's search time varies with its candidates, and a block's LDS is held
// until its slowest wave ends, so single-wave blocks free their share as soon as their group is done
// (8M turbulence 5.29 -> 5.10 ms, 4M Evrard 3.59 -> 3.44 ms against 4-wave blocks)
#ifndef SPHX_SEARCH_WAVES_PER_BLOCK
#define SPHX_SEARCH_WAVES_PER_BLOCK 1
#endif
constexpr int kWavesPerBlock = SPHX_SEARCH_WAVES_PER_BLOCK;
constexpr int kStackCap      = 952;  // LDS node stack per wave (bound argued at searchGroupInto)
constexpr int kLeafCap       = 64;   // leaves per sweep of candidate leaves (one per lane)
constexpr int kBufCap        = 128;  // LDS buffer of found leaves awaiting their turn in visiting order
constexpr int kPerPairImage  = 64;   // image code of a leaf swept with the per-pair fold
constexpr int kNoShiftImage  = 13;   // image code of shift (0, 0, 0)

struct SearchArgs
{
    const double*        x;
    const double*        y;
    const double*        z;
    const TreeNodeIndex* childOffsets;
    const TreeNodeIndex* internalToLeaf;
    const int32_t*       rankLo;    // visiting order of the reference traversal (first leaf of a node)
    const uint32_t*      childPush; // an internal node's children from last to first visited
    const uint32_t*      layout;
    const double*        centers;
    const double*        sizes;
    DevBox               box;
    double               slack; // absolute enlargement of search radii (rounding margin)
    uint32_t             first, last, ng0, ngmax, ngmax4;
    Groups               groups; // target groups, one wavefront each
    int                  iterateH;
    uint32_t*            nbrs;
    uint32_t*            nc;
    unsigned long long*  stats; // [0] stored pairs, [1] max count, [2] non-converged particles, [kStatWide] wide
                                // support, [kStatTests] candidate distance tests (swept candidates x 64 lanes)
    double               minLen; // shortest box length
    // fused XMass: masses, kernel table (TabPair of h's type), normalisation, output (xm or rho)
    const void*          m;
    const void*          wh;    // FP64 form: the W table
    float                sincN; // FP32 form: the sinc exponent (W evaluated directly) ...
    int                  sincInt; // ... and that exponent when it is an integer, else -1
    double               K;
    void*                xmOut;
    int                  toRho;
};

//! per-wave LDS: traversal stack, candidate leaves, the kept leaves' particle ranges, and the
//! double-buffered candidate tile (with the candidates' masses for the fused XMass only)
template<bool Mass>
struct WaveLds
{
    int32_t  stack[kStackCap];      // pending nodes, the first in visiting order on top
    int32_t  bufNode[kBufCap];      // found leaves ...
    int32_t  bufRank[kBufCap];      // ... and their visiting-order ranks
    uint32_t leafLo[kLeafCap];      // kept leaves: first particle
    uint32_t leafPre[kLeafCap + 1]; // kept leaves: exclusive prefix of particle counts, then the total
    int32_t  leafImg[kLeafCap];     // kept leaves: periodic image code (imageShift)
    double   tx[2][kWave], ty[2][kWave], tz[2][kWave];
    uint32_t tj[2][kWave]; // particle index of each staged candidate
    double   tm[2][Mass ? kWave : 1]; // its mass (fused XMass only)
};

// sixteen waves of the search share a CU's 160 KB of LDS (the fused search + XMass, with its masses,
// holds three waves per SIMD by its registers)
static_assert(sizeof(WaveLds<false>) <= 160 * 1024 / 16, "search LDS must leave room for 16 waves per CU");
static_assert(sizeof(WaveLds<true>) <= 160 * 1024 / 12, "fused search LDS must leave room for 12 waves per CU");

//! per-target state of a lane
struct Target
{
    double   x, y, z;
    double   r2;     // exact predicate radius^2 (formed in h's type)
    double   reach2; // (2h + slack)^2, conservative radius for cell tests
    uint32_t i;
    bool     active;
};

//! A lane's appending writer into its neighbour list: entries collect in a register quad that is
//! stored (16 bytes) when full and, for the incomplete last quad, by finish(). Slots past the
//! stored count inside that quad receive don't-care values; readers never use them.
struct ListWriter
{
    uint32_t* base; // the lane's slot 0
    uint32_t  q0 = 0, q1 = 0, q2 = 0, q3 = 0; // the last four entries, oldest first
    uint32_t  swept = 0; // candidates swept by the wave (each tested by all 64 lanes)

    //! entry k = j: shifted in behind the previous three, stored when it completes a quad
    __device__ __forceinline__ void put(unsigned k, uint32_t j)
    {
        q0 = q1, q1 = q2, q2 = q3, q3 = j;
        if ((k & 3u) == 3u) { store(k, q0, q1, q2, q3); }
    }
    __device__ __forceinline__ void store(unsigned k, uint32_t a, uint32_t b, uint32_t c, uint32_t d)
    {
        *reinterpret_cast<uint4*>(base + laneSlot(k & ~3u)) = make_uint4(a, b, c, d);
    }
    //! flush after `stored` entries: the r = stored % 4 last ones (q3 the newest) open the quad;
    //! its remaining slots (never read) repeat the newest entry
    __device__ __forceinline__ void finish(unsigned stored)
    {
        const unsigned r = stored & 3u;
        if (r == 1) { store(stored - 1, q3, q3, q3, q3); }
        else if (r == 2) { store(stored - 1, q2, q3, q3, q3); }
        else if (r == 3) { store(stored - 1, q1, q2, q3, q3); }
    }
};

//! what the search does with each stored hit besides listing it: nothing ...
struct ListOnly
{
    static constexpr bool kStageMass = false;
    static constexpr int  kBatch     = 1;
    __device__ __forceinline__ void stage(int, bool, double, double, double, double) {}
    __device__ __forceinline__ void flush() {}
};

//! ... or the XMass sum of xmassJLoop (hydro_ve/xmass_kern.hpp:51-79): m_i + sum_j m_j W(r_ij / h_i)
//! over the stored neighbours in list order, r_ij from the same separation as the pair loop
//! (minimum image, then rounded to h's type), compensated for the FP32 form. Hits are staged
//! kBatch at a time: their kernel values are fetched together, then accumulated in list order,
//! slots without a stored hit leaving the sum untouched (bitwise the one-at-a-time sum).
template<class Th>
struct XMassSum
{
    static constexpr bool kStageMass = true;
    static constexpr int  kBatch     = isMixed<Th> ? 1 : 2; // FP32 evaluates W, no gather to overlap
    SincKernel<Th>        kern;
    Th                    hInv;
    Acc<Th>               rho0;
    typename SincKernel<Th>::FetchW w[kBatch];
    Th                          mj[kBatch];
    bool                        on[kBatch];

    __device__ __forceinline__ void stage(int u, bool stored, double dx, double dy, double dz, double m)
    {
        Th rx = Th(dx), ry = Th(dy), rz = Th(dz);
        Th dist = fastSqrt(rx * rx + ry * ry + rz * rz);
        on[u]   = stored;
        mj[u]   = Th(m);
        w[u]    = kern.fetchW(stored ? dist * hInv : Th(4)); // 4: past the support, W = 0
    }
    __device__ __forceinline__ void flush()
    {
#pragma unroll
        for (int u = 0; u < kBatch; ++u)
        {
            rho0.addIf(on[u], value(w[u]) * mj[u]);
        }
    }
};

__device__ __forceinline__ uint32_t uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ double uniformD(double v)
{
    uint64_t b = __double_as_longlong(v);
    uint32_t lo = __builtin_amdgcn_readfirstlane(uint32_t(b)), hi = __builtin_amdgcn_readfirstlane(uint32_t(b >> 32));
    return __longlong_as_double((long long)((uint64_t(hi) << 32) | lo));
}

enum class Fold
{
    none,    // all pairs of the tile use the direct image
    shifted, // one periodic image (shift multiple of the box length) serves all pairs
    perPair  // box too small for a single image: rint fold per pair, as the reference
};

//! Test the `n` staged candidates against the lane's target. Hits are recorded as bits of a
//! per-lane mask over the tile's candidates and appended afterwards in candidate order, so the
//! list keeps the traversal order while the test loop stays free of the divergent append.
template<Fold mode, class Sink>
__device__ __forceinline__ void sweepTile(const SearchArgs& a, const Target& t, const double* tx, const double* ty,
                                          const double* tz, const uint32_t* tj, const double* tm, uint32_t n,
                                          double sx, double sy, double sz, unsigned& count, ListWriter& out,
                                          Sink& sink)
{
#pragma clang fp contract(off)
    const DevBox&  b     = a.box;
    const unsigned ngmax = a.ngmax;
    auto test32 = [&](uint32_t c0, uint32_t cn) {
        uint32_t m = 0;
#pragma unroll 4
        for (uint32_t c = 0; c < cn; ++c)
        {
            double dx = tx[c0 + c] - t.x, dy = ty[c0 + c] - t.y, dz = tz[c0 + c] - t.z;
            if constexpr (mode == Fold::shifted) { dx = dx - sx, dy = dy - sy, dz = dz - sz; }
            if constexpr (mode == Fold::perPair)
            {
                dx = dx - double(b.pbc[0]) * b.len[0] * rint(dx * b.ilen[0]);
                dy = dy - double(b.pbc[1]) * b.len[1] * rint(dy * b.ilen[1]);
                dz = dz - double(b.pbc[2]) * b.len[2] * rint(dz * b.ilen[2]);
            }
            double d2 = dx * dx + dy * dy + dz * dz;
            m |= d2 < t.r2 ? (1u << c) : 0u;
        }
        return m;
    };
    // a tile holds at most 64 candidates: one 64-bit hit mask per lane, drained in one pass (the
    // drain runs as many rounds as the busiest lane has hits, and one mask over the whole tile
    // needs fewer rounds than two over its halves)
    for (uint32_t c0 = 0; c0 < n; c0 += 64)
    {
        const uint32_t cn = uniform(min(n - c0, 64u));
        uint64_t       m  = test32(c0, min(cn, 32u));
        if (cn > 32u) { m |= uint64_t(test32(c0 + 32u, cn - 32u)) << 32; }
        if (!t.active) { m = 0; }
        while (m)
        {
#pragma unroll
            for (int u = 0; u < Sink::kBatch; ++u)
            {
                const bool     have  = m != 0;
                const uint32_t c     = c0 + (have ? uint32_t(__builtin_ctzll(m)) : 0u);
                const uint32_t j     = tj[c];
                const bool     keep  = have && j != t.i;
                const bool     store = keep && count < ngmax;
                m &= m - 1;
                if (store) { out.put(count, j); }
                count += keep ? 1u : 0u;
                if constexpr (Sink::kStageMass)
                {
                    double dx = tx[c] - t.x, dy = ty[c] - t.y, dz = tz[c] - t.z;
                    if constexpr (mode == Fold::shifted) { dx = dx - sx, dy = dy - sy, dz = dz - sz; }
                    if constexpr (mode == Fold::perPair)
                    {
                        dx = dx - double(b.pbc[0]) * b.len[0] * rint(dx * b.ilen[0]);
                        dy = dy - double(b.pbc[1]) * b.len[1] * rint(dy * b.ilen[1]);
                        dz = dz - double(b.pbc[2]) * b.len[2] * rint(dz * b.ilen[2]);
                    }
                    sink.stage(u, store, dx, dy, dz, tm[c]);
                }
            }
            sink.flush();
        }
    }
}

__device__ __forceinline__ double readlaneD(double v, int q)
{
    uint64_t b  = __double_as_longlong(v);
    uint32_t lo = __builtin_amdgcn_readlane(uint32_t(b), q), hi = __builtin_amdgcn_readlane(uint32_t(b >> 32), q);
    return __longlong_as_double((long long)((uint64_t(hi) << 32) | lo));
}

//! inclusive prefix sum over the wave
__device__ __forceinline__ uint32_t waveInclusiveSum(uint32_t v)
{
    const unsigned lane = laneId();
    for (int o = 1; o < kWave; o <<= 1)
    {
        uint32_t w = __shfl_up(v, o, kWave);
        v += lane >= unsigned(o) ? w : 0u;
    }
    return v;
}

/* Periodic image of a candidate leaf relative to the group centre: the shift by whole box lengths
 * k_d = rint((c_d - g_d) / L_d), k_d in {-1, 0, 1}, coded (k_x+1) + 3 (k_y+1) + 9 (k_z+1), which
 * serves all pairs of the leaf; kPerPairImage when the group, the leaf and the reach together
 * exceed a quarter box in a periodic dimension (then the per-pair fold of the reference). */
__device__ __forceinline__ int imageCode(const DevBox& b, const double c[3], const double s[3], const double gc[3],
                                         const double ge[3], double reach)
{
    int  code    = 0;
    bool perPair = false;
    for (int d = 2; d >= 0; --d)
    {
        int k = 0;
        if (b.pbc[d])
        {
            k       = int(rint((c[d] - gc[d]) * b.ilen[d]));
            perPair = perPair || (ge[d] + s[d] + reach > 0.25 * b.len[d]);
        }
        code = 3 * code + (k + 1);
    }
    return perPair ? kPerPairImage : code;
}

__device__ __forceinline__ double imageShift(const DevBox& b, int code, int d)
{
    const int k = (d == 0 ? code : d == 1 ? code / 3 : code / 9) % 3 - 1;
    return double(k) * b.len[d];
}

//! sweep the first n candidates of tile buffer `buf`, all of image `code`
template<class Sink, class Lds>
__device__ __forceinline__ void sweepBuffer(const SearchArgs& a, const Target& t, Lds& lds, int buf, uint32_t n,
                                            int code, unsigned& count, ListWriter& out, Sink& sink)
{
    __builtin_amdgcn_wave_barrier(); // a wave's LDS accesses complete in order
    out.swept += n;
    const double* tx = lds.tx[buf];
    const double* ty = lds.ty[buf];
    const double* tz = lds.tz[buf];
    if (code == kPerPairImage)
        sweepTile<Fold::perPair>(a, t, tx, ty, tz, lds.tj[buf], lds.tm[buf], n, 0, 0, 0, count, out, sink);
    else if (code != kNoShiftImage)
        sweepTile<Fold::shifted>(a, t, tx, ty, tz, lds.tj[buf], lds.tm[buf], n, imageShift(a.box, code, 0),
                                 imageShift(a.box, code, 1), imageShift(a.box, code, 2), count, out, sink);
    else
        sweepTile<Fold::none>(a, t, tx, ty, tz, lds.tj[buf], lds.tm[buf], n, 0, 0, 0, count, out, sink);
    __builtin_amdgcn_wave_barrier();
}

/* Sweep the buffered candidate leaves (nLeaf <= 64, traversal order).
 *
 * 1. Lane q takes leaf q: its box, particle range and periodic image load in parallel. The leaf
 *    is kept if some active target's sphere reaches its box (the box is broadcast to the wave
 *    lane by lane). Kept leaves get their first particle, image and the exclusive prefix of
 *    their particle counts in LDS, in traversal order.
 * 2. Their particles are read as one stream, 64 per window (a window never spans two images).
 *    Each lane loads one particle and keeps it if it lies within reach of the group box under
 *    its leaf's image. Kept particles are appended, in stream order, to the tile being filled.
 *    A full tile of 64 (or the last one, or one whose image ends) is swept, while the next
 *    window's loads are in flight. The candidate order is therefore the traversal order, as in
 *    the reference. */
template<class Sink, class Lds>
__device__ __forceinline__ void flushLeaves(const SearchArgs& a, const Target& t, Lds& lds, const int32_t* leaves, int nLeaf,
                            const double gc[3],
                            const double ge[3], double reach, unsigned& count, ListWriter& out, Sink& sink)
{
#pragma clang fp contract(off)
    const DevBox&  b    = a.box;
    const unsigned lane = laneId();

    // 1. leaf boxes, ranges and images in parallel; keep test with the box broadcast per leaf
    double   c[3] = {0, 0, 0}, sz[3] = {0, 0, 0};
    uint32_t lo = 0, cnt = 0;
    int      code = kNoShiftImage;
    if (int(lane) < nLeaf)
    {
        const int32_t node = leaves[lane];
        for (int d = 0; d < 3; ++d)
        {
            c[d]  = a.centers[3 * size_t(node) + d];
            sz[d] = a.sizes[3 * size_t(node) + d];
        }
        const uint32_t leaf = uint32_t(a.internalToLeaf[node]);
        lo                  = a.layout[leaf];
        cnt                 = a.layout[leaf + 1] - lo;
        code                = imageCode(b, c, sz, gc, ge, reach);
    }
    uint64_t keepMask = 0;
    {
        const double p[3] = {t.x, t.y, t.z};
        if (!ballot(code == kPerPairImage))
        {
            // every leaf has one image for the whole group: its box, centre moved to that image,
            // goes to the idle tile buffers and is read back as an LDS broadcast per leaf
            if (int(lane) < nLeaf)
            {
                lds.tx[0][lane] = c[0] - imageShift(b, code, 0), lds.tx[1][lane] = sz[0];
                lds.ty[0][lane] = c[1] - imageShift(b, code, 1), lds.ty[1][lane] = sz[1];
                lds.tz[0][lane] = c[2] - imageShift(b, code, 2), lds.tz[1][lane] = sz[2];
            }
            __builtin_amdgcn_wave_barrier();
            for (int q = 0; q < nLeaf; ++q)
            {
                const double cq[3] = {lds.tx[0][q], lds.ty[0][q], lds.tz[0][q]};
                const double sq[3] = {lds.tx[1][q], lds.ty[1][q], lds.tz[1][q]};
                double       d2    = 0.0;
                for (int d = 0; d < 3; ++d)
                {
                    double v = fabs(cq[d] - p[d]) - sq[d];
                    v        = v > 0.0 ? v : 0.0;
                    d2 += v * v;
                }
                keepMask |= ballot(t.active && d2 < t.reach2) ? (uint64_t(1) << q) : uint64_t(0);
            }
            __builtin_amdgcn_wave_barrier();
        }
        else
        {
            for (int q = 0; q < nLeaf; ++q)
            {
                double d2 = 0.0;
                for (int d = 0; d < 3; ++d)
                {
                    double v = readlaneD(c[d], q) - p[d];
                    if (b.pbc[d]) { v = v - b.len[d] * rint(v * b.ilen[d]); }
                    v = fabs(v) - readlaneD(sz[d], q);
                    v = v > 0.0 ? v : 0.0;
                    d2 += v * v;
                }
                keepMask |= ballot(t.active && d2 < t.reach2) ? (uint64_t(1) << q) : uint64_t(0);
            }
        }
    }
    const bool     kept   = (keepMask >> lane) & 1u;
    const int      nKept  = __popcll(keepMask);
    const uint32_t incl   = waveInclusiveSum(kept ? cnt : 0u);
    if (kept)
    {
        const unsigned r = laneRank(keepMask);
        lds.leafLo[r]    = lo;
        lds.leafImg[r]   = code;
        lds.leafPre[r]   = incl - cnt;
    }
    const uint32_t total = uniform(__shfl(incl, kWave - 1, kWave));
    if (lane == 0) { lds.leafPre[nKept] = total; }
    __builtin_amdgcn_wave_barrier();
    if (total == 0) { return; }

    // 2. the kept leaves' particles as one stream of windows, compacted into full tiles
    const double reach2 = reach * reach;
    uint32_t     S      = 0;  // stream position of the current window
    int          qs     = 0;  // kept leaf holding S
    uint32_t     runEnd = 0;  // stream end of the run of leaves with the current window's image
    int          runImg = kNoShiftImage;
    // plan the window at S: its end, and this lane's particle (valid, index)
    auto plan = [&](uint32_t& E, bool& valid, uint32_t& idx) {
        while (qs + 1 < nKept && lds.leafPre[qs + 1] <= S) { qs++; }
        if (S >= runEnd)
        {
            runImg = lds.leafImg[qs];
            int q  = qs + 1;
            while (q < nKept && lds.leafImg[q] == runImg) { q++; }
            runEnd = lds.leafPre[q];
        }
        E                = min(S + uint32_t(kWave), runEnd);
        const uint32_t p = S + lane;
        valid            = p < E;
        idx              = 0;
        if (valid)
        {
            int q = qs;
            while (q + 1 < nKept && lds.leafPre[q + 1] <= p) { q++; }
            idx = lds.leafLo[q] + (p - lds.leafPre[q]);
        }
    };
    double   px = 0, py = 0, pz = 0, pm = 0;
    auto     load = [&](bool valid, uint32_t idx) {
        if (valid)
        {
            px = a.x[idx], py = a.y[idx], pz = a.z[idx];
            if constexpr (Sink::kStageMass) { pm = double(static_cast<const decltype(sink.hInv)*>(a.m)[idx]); }
        }
    };

    uint32_t E;
    bool     valid;
    uint32_t idx;
    plan(E, valid, idx);
    load(valid, idx);
    int      cur     = 0;             // tile buffer being filled
    uint32_t acc     = 0;             // candidates in it
    int      tileImg = runImg;        // their image
    while (true)
    {
        const int      img = runImg;  // image of the loaded window
        const uint32_t end = E;
        if (img != tileImg && acc > 0) // a tile holds one image
        {
            sweepBuffer(a, t, lds, cur, acc, tileImg, count, out, sink);
            acc = 0, cur ^= 1;
        }
        tileImg = img;
        // keep the window's particles within reach of the group box under its image
        bool keep = valid;
        if (keep && img != kPerPairImage)
        {
            const double q[3] = {px - imageShift(b, img, 0), py - imageShift(b, img, 1), pz - imageShift(b, img, 2)};
            double       d2   = 0.0;
            for (int d = 0; d < 3; ++d)
            {
                double v = fabs(q[d] - gc[d]) - ge[d];
                v        = v > 0.0 ? v : 0.0;
                d2 += v * v;
            }
            keep = d2 < reach2;
        }
        const uint64_t kmask = ballot(keep);
        if (keep)
        {
            const uint32_t slot = acc + laneRank(kmask);
            const int      bb   = slot < uint32_t(kWave) ? cur : cur ^ 1;
            const uint32_t sl   = slot < uint32_t(kWave) ? slot : slot - kWave;
            lds.tx[bb][sl] = px, lds.ty[bb][sl] = py, lds.tz[bb][sl] = pz;
            lds.tj[bb][sl] = idx;
            if constexpr (Sink::kStageMass) { lds.tm[bb][sl] = pm; }
        }
        acc += uint32_t(__popcll(kmask));
        // next window's loads go out before the sweep
        S                = end;
        const bool more  = S < total;
        if (more)
        {
            plan(E, valid, idx);
            load(valid, idx);
        }
        if (acc >= uint32_t(kWave))
        {
            sweepBuffer(a, t, lds, cur, kWave, tileImg, count, out, sink);
            acc -= kWave, cur ^= 1;
        }
        if (!more) { break; }
    }
    if (acc > 0) { sweepBuffer(a, t, lds, cur, acc, tileImg, count, out, sink); }
}

//! Sort the first n entries of the found-leaf buffer by rank (bitonic network over the next
//! power of two, the padding ranked last; ranks of distinct leaves are distinct)
template<class Lds>
__device__ __forceinline__ void sortFoundLeaves(Lds& lds, int n)
{
    const unsigned lane = laneId();
    // leaves found in one pop arrive in order: often there is nothing to sort
    bool inversion = false;
    for (int q = int(lane); q + 1 < n; q += kWave)
        inversion = inversion || lds.bufRank[q] > lds.bufRank[q + 1];
    if (!ballot(inversion)) { return; }
    int size = 2;
    while (size < n) { size <<= 1; }
    for (int q = n + int(lane); q < size; q += kWave)
    {
        lds.bufRank[q] = INT32_MAX;
        lds.bufNode[q] = -1;
    }
    __builtin_amdgcn_wave_barrier();
    for (int k = 2; k <= size; k <<= 1)
    {
        for (int j = k >> 1; j > 0; j >>= 1)
        {
            for (int p = int(lane); p < (size >> 1); p += kWave)
            {
                const int  i  = 2 * j * (p / j) + (p % j);
                const int  l  = i + j;
                const bool up = (i & k) == 0;
                const int  ri = lds.bufRank[i], rl = lds.bufRank[l];
                if ((ri > rl) == up)
                {
                    const int ni = lds.bufNode[i], nl = lds.bufNode[l];
                    lds.bufRank[i] = rl, lds.bufRank[l] = ri;
                    lds.bufNode[i] = nl, lds.bufNode[l] = ni;
                }
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
}

/* Collect all leaves whose (periodic) distance to the active targets' bounding box is below the
 * largest active search radius and sweep them in the reference's visiting order
 * (singleTraversal, traversal/traversal.hpp:52-93: at each node its leaf children in octant
 * order, then its internal children in reverse octant order, depth first), which is the order
 * of each target's list. Returns each lane's neighbour count.
 *
 * Every node carries the rank of its first leaf in that order (rankLo), and the leaves of a
 * subtree have consecutive ranks. The stack holds pending nodes sorted by rank, the first on top:
 * a pop takes the up to 64 first pending nodes (lane 0 the first), whose subtrees all precede the
 * rest of the stack, and pushes the children of the overlapping internal ones back on top in the
 * same order (childPush). Found leaves wait in an LDS buffer with their ranks; once no pending
 * node can hold an earlier leaf (rank below the top node's), they are final and are swept in rank
 * order, 64 at a time.
 *
 * Bounds: nodes are popped 64 at a time only while top + 7*64 <= kStackCap - 160 and the buffer
 * can take 64 more leaves with 8 to spare, else one at a time. A single pop takes the first
 * pending node and replaces it by at most 8 children, the next pop taking the first of them,
 * i.e. it proceeds depth first and adds at most 7 entries per tree level below the popped node
 * (<= 7*20 = 140 < 160) on top of the stack it started from. A single pop adds at most one leaf,
 * the first in order of all, and the buffer is swept whenever it reaches its spare slots. */
template<class Sink, class Lds>
__device__ __forceinline__ unsigned searchGroupInto(const SearchArgs& a, const Target& t, double radius, Lds& lds, ListWriter& out,
                                    Sink& sink)
{
#pragma clang fp contract(off)
    const unsigned lane = laneId();
    const DevBox&  b    = a.box;

    double gmin[3], gmax[3], gc[3], ge[3];
    double p[3] = {t.x, t.y, t.z};
    for (int d = 0; d < 3; ++d)
    {
        gmin[d] = waveMin(t.active ? p[d] : INFINITY);
        gmax[d] = waveMax(t.active ? p[d] : -INFINITY);
        gc[d]   = 0.5 * (gmin[d] + gmax[d]);
        ge[d]   = 0.5 * (gmax[d] - gmin[d]);
    }
    double reach  = waveMax(t.active ? radius : 0.0) + a.slack;
    double reach2 = reach * reach;

    unsigned count = 0;
    int      top   = 1;
    int      nBuf  = 0;
    if (lane == 0) { lds.stack[0] = 0; }
    __builtin_amdgcn_wave_barrier();

    // sweep the buffered leaves ranked below `limit`, in rank order, and keep the others
    auto sweepFound = [&](int32_t limit) {
        sortFoundLeaves(lds, nBuf);
        int final = 0;
        for (int q = 0; q < nBuf; q += kWave)
        {
            const int i = q + int(lane);
            final += __popcll(ballot(i < nBuf && lds.bufRank[i] < limit));
        }
        for (int done = 0; done < final; done += kLeafCap)
        {
            flushLeaves(a, t, lds, lds.bufNode + done, min(kLeafCap, final - done), gc, ge, reach, count, out, sink);
        }
        // move the rest to the front
        const int rest = nBuf - final;
        int32_t   n0 = 0, r0 = 0, n1 = 0, r1 = 0;
        if (int(lane) < rest) { n0 = lds.bufNode[final + lane], r0 = lds.bufRank[final + lane]; }
        if (int(lane) + kWave < rest) { n1 = lds.bufNode[final + kWave + lane], r1 = lds.bufRank[final + kWave + lane]; }
        __builtin_amdgcn_wave_barrier();
        if (int(lane) < rest) { lds.bufNode[lane] = n0, lds.bufRank[lane] = r0; }
        if (int(lane) + kWave < rest) { lds.bufNode[kWave + lane] = n1, lds.bufRank[kWave + lane] = r1; }
        __builtin_amdgcn_wave_barrier();
        nBuf = rest;
    };

    while (top > 0)
    {
        const bool batch = top + 7 * kWave <= kStackCap - 160 && nBuf + kWave <= kBufCap - 8;
        const int  k     = batch ? min(kWave, top) : 1;
        int        node  = -1;
        if (int(lane) < k) { node = lds.stack[top - 1 - int(lane)]; }
        top -= k;
        __builtin_amdgcn_wave_barrier();

        bool     ov = false, isLeaf = false;
        int      child = 0;
        int32_t  rank  = 0;
        uint32_t push  = 0;
        if (node >= 0)
        {
            // the node's rank and child order load with its box (childPush is read for leaves too,
            // unused there)
            rank = a.rankLo[node];
            push = a.childPush[node];
            double d2 = 0.0;
            for (int d = 0; d < 3; ++d)
            {
                double v = a.centers[3 * size_t(node) + d] - gc[d];
                if (b.pbc[d]) { v = v - b.len[d] * rint(v * b.ilen[d]); }
                v = fabs(v) - a.sizes[3 * size_t(node) + d] - ge[d];
                v = v > 0.0 ? v : 0.0;
                d2 += v * v;
            }
            ov     = d2 < reach2;
            child  = a.childOffsets[node];
            isLeaf = child == 0;
        }
        const uint64_t mLeaf = ballot(ov && isLeaf);
        const uint64_t mInt  = ballot(ov && !isLeaf);
        if (ov && isLeaf)
        {
            const int pos  = nBuf + int(laneRank(mLeaf));
            lds.bufNode[pos] = node;
            lds.bufRank[pos] = rank;
        }
        if (ov && !isLeaf)
        {
            // the children of later lanes go deeper: lane 0's, the first in order, end on top
            const int later = __popcll(lane + 1 < kWave ? mInt >> (lane + 1) : uint64_t(0));
            const int pos   = top + 8 * later;
            for (int c = 0; c < 8; ++c)
                lds.stack[pos + c] = child + int((push >> (3 * c)) & 7u);
        }
        nBuf += __popcll(mLeaf);
        top += 8 * __popcll(mInt);
        __builtin_amdgcn_wave_barrier();

        // leaves ranked below the first pending node are final
        if (nBuf >= kLeafCap || nBuf > kBufCap - 8 - kWave)
        {
            // the first pending node: the first visited child of the first pushed node, whose rank
            // is its parent's, or else the node left below the popped ones
            int32_t limit = INT32_MAX;
            if (mInt) { limit = __builtin_amdgcn_readlane(rank, __builtin_ctzll(mInt)); }
            else if (top > 0) { limit = a.rankLo[lds.stack[top - 1]]; }
            int           final = 0;
            for (int q = 0; q < nBuf; q += kWave)
            {
                const int i = q + int(lane);
                final += __popcll(ballot(i < nBuf && lds.bufRank[i] < limit));
            }
            if (final >= kLeafCap || (nBuf > kBufCap - 8 && final > 0)) { sweepFound(limit); }
        }
    }
    sweepFound(INT32_MAX);
    return count;
}

//! search of the active lanes' targets into their lists (slot 0 at myList); returns the counts
template<class Sink, class Lds>
__device__ __forceinline__ unsigned searchGroup(const SearchArgs& a, const Target& t, double radius, Lds& lds, uint32_t* myList,
                                Sink& sink, uint32_t& swept)
{
    ListWriter out{myList};
    unsigned   count = searchGroupInto(a, t, radius, lds, out, sink);
    if (t.active) { out.finish(min(count, a.ngmax)); }
    swept += out.swept;
    return count;
}

//! updateH (sph/include/sph/kernels.hpp:27-32)
template<class Th>
__device__ __forceinline__ Th updateH(unsigned ng0, unsigned nc, Th h)
{
#pragma clang fp contract(off)
    constexpr Th c0 = 1023.0;
    constexpr Th ex = 1.0 / 10.0;
    return h * Th(0.5) * pow(Th(1) + c0 * ng0 / Th(nc), ex);
}

template<class Th>
__device__ __forceinline__ void setRadius(Target& t, Th hi, double slack)
{
#pragma clang fp contract(off)
    Th r2      = Th(4.0) * hi * hi; // radius in h's type, as the reference
    t.r2       = double(r2);
    double rr  = 2.0 * double(hi) + slack;
    t.reach2   = rr * rr;
}

/* One wave per group: search, h/nc iteration, and with XMass the fused density sum of the final
 * search, stored as xm = m / rho0 (xmassJLoop) or rho = m / xm (computeDensity,
 * hydro_std/density.hpp:42-53), with the per-target epilogue of the XMass kernel. */
// waves per SIMD the search kernel's register allocation must leave room for, per form (the
// fused XMass variant is left to the compiler)
#ifndef SPHX_WAVES_SEARCH_F32
#define SPHX_WAVES_SEARCH_F32 4
#endif
#ifndef SPHX_WAVES_SEARCH_F64
#define SPHX_WAVES_SEARCH_F64 4
#endif

template<class Th, bool XMass>
__global__ void __launch_bounds__(kWavesPerBlock * kWave)
__attribute__((amdgpu_waves_per_eu(XMass ? 1 : isMixed<Th> ? SPHX_WAVES_SEARCH_F32 : SPHX_WAVES_SEARCH_F64)))
searchKernel(SearchArgs a, Th* __restrict__ h)
{
    __shared__ WaveLds<XMass> ldsMem[kWavesPerBlock];

    const unsigned wid  = threadIdx.x / kWave;
    const unsigned lane = laneId();
    const uint32_t g    = uniform(xcdBlock(blockIdx.x, gridDim.x) * kWavesPerBlock + wid);
    if (g >= a.groups.num) { return; }
    uint32_t base, end;
    bool     oversized;
    groupRange(a.groups, g, base, end, oversized);
    base = uniform(base), end = uniform(end);
    if (oversized && lane == 0) { atomicAdd(&a.stats[kStatBadGroups], 1ull); }
    WaveLds<XMass>& lds = ldsMem[wid];

    Target t;
    t.i      = base + lane;
    t.active = t.i < end;
    Th hi    = t.active ? h[t.i] : Th(0);
    t.x      = t.active ? a.x[t.i] : 0.0;
    t.y      = t.active ? a.y[t.i] : 0.0;
    t.z      = t.active ? a.z[t.i] : 0.0;
    setRadius(t, hi, a.slack);
    // lists are indexed by particle: target i at group (i - first) / 64, lane (i - first) % 64
    const uint32_t r      = t.active ? t.i - a.first : 0u;
    uint32_t*      myList = a.nbrs + listOffset(r / kGroup, a.ngmax4, 0, r % kGroup);

    using Sink = std::conditional_t<XMass, XMassSum<Th>, ListOnly>;
    Sink      sink;
    const Th  mi = XMass && t.active ? static_cast<const Th*>(a.m)[t.i] : Th(0);
    auto      resetSink = [&]()
    {
        if constexpr (XMass)
        {
            if constexpr (isMixed<Th>) { sink.kern.n = a.sincN, sink.kern.ni = a.sincInt; }
            else { sink.kern.w = static_cast<const TabPair<double>*>(a.wh); }
            sink.hInv = Th(1) / hi;
            sink.rho0 = Acc<Th>(mi);
        }
    };
    resetSink();

    bool     valid = t.active;
    uint32_t swept = 0;
    unsigned count = searchGroup(a, t, double(2) * double(hi), lds, myList, sink, swept);
    unsigned ncSph = 1 + count;

    int      updates = 0;
    unsigned ngmin   = a.ng0 / 4;
    if (a.iterateH)
    {
        while (true)
        {
            bool repeat = valid && (ngmin > ncSph || (ncSph - 1) > a.ngmax) && updates < 10;
            if (!ballot(repeat)) { break; }
            if (repeat)
            {
                updates++;
                hi = updateH(a.ng0, ncSph, hi);
                setRadius(t, hi, a.slack);
                resetSink();
            }
            t.active       = repeat;
            unsigned again = searchGroup(a, t, double(2) * double(hi), lds, myList, sink, swept);
            if (repeat) { ncSph = 1 + again; }
        }
    }
    if (valid)
    {
        if (updates > 0) { h[t.i] = hi; }
        a.nc[t.i] = ncSph;
        // slot 0 always names a particle: the pair loops read one slot ahead without a branch
        if (ncSph == 1) { myList[0] = t.i; }
        if constexpr (XMass)
        {
            // epilogue of xmassKernel: FP64 normalisation with h^-3 rounded in h's type
            const Th hInv    = Th(1) / hi;
            const Th h3Inv   = hInv * hInv * hInv;
            double   rhoZero = double(sink.rho0.value()) * a.K * double(h3Inv);
            Th       xmi     = Th(double(mi) / rhoZero);
            static_cast<Th*>(a.xmOut)[t.i] = a.toRho ? Th(mi / xmi) : xmi;
        }
    }

    unsigned stored = valid ? min(ncSph - 1, a.ngmax) : 0;
    unsigned sum    = waveSum(stored);
    unsigned mx     = waveMax(valid ? ncSph : 0u);
    unsigned fails  = __popcll(ballot(valid && updates >= 10));
    // support 2h reaching half a box length: the FP32 pair loops then take exact separations
    unsigned wide = __popcll(ballot(valid && 4.0 * double(hi) * (1.0 + 1e-6) >= a.minLen));
    // largest h: sets the fixed-point period of the FP32 pair loops (positive doubles order as
    // their bit patterns)
    double hMax = waveMax(valid ? double(hi) : 0.0);
    if (lane == 0)
    {
        atomicMax(&a.stats[kStatHMax], (unsigned long long)__double_as_longlong(hMax));
        atomicAdd(&a.stats[0], (unsigned long long)sum);
        atomicMax(&a.stats[1], (unsigned long long)mx);
        if (fails) { atomicAdd(&a.stats[2], (unsigned long long)fails); }
        if (wide) { atomicAdd(&a.stats[kStatWide], (unsigned long long)wide); }
        atomicAdd(&a.stats[kStatTests], (unsigned long long)swept * kWave);
    }
}

//! row i - first of `out` = stored list of target i; groups count from the searched range's first target
__global__ void copyListsKernel(const uint32_t* __restrict__ nbrs, const uint32_t* __restrict__ nc,
                                uint32_t* __restrict__ out, uint32_t searchFirst, uint32_t first, uint32_t last,
                                uint32_t ngmax)
{
    uint32_t i = first + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= last) { return; }
    uint32_t r = i - searchFirst, g = r / kGroup, lane = r % kGroup;
    uint32_t n = min(nc[i] - 1, ngmax);
    for (uint32_t k = 0; k < n; ++k)
        out[size_t(i - first) * ngmax + k] = nbrs[listOffset(g, paddedNgmax(ngmax), k, lane)];
}

} // namespace

int findNeighbors(sphx_workspace* ws, const DevBox& box, const sphx_fields_t* f, const sphx_params_t* p,
                  uint32_t first, uint32_t last, bool iterateH, int xmassOut)
{
    SPHX_REQUIRE(ws->tree.valid, SPHX_ERR_STATE, "find_neighbors called before the tree was built");
    SPHX_REQUIRE(f->x && f->y && f->z && f->h && f->nc, SPHX_ERR_ARG, "find_neighbors needs x, y, z, h and nc");
    SPHX_REQUIRE(first <= last && last <= ws->tree.numParticles, SPHX_ERR_ARG, "invalid target range");
    SPHX_REQUIRE(p->ng0 <= p->ngmax, SPHX_ERR_ARG, "ng0 should be smaller than ngmax");
    SPHX_REQUIRE(p->ngmax > 0, SPHX_ERR_ARG, "ngmax must be positive");

    const Tree&  t      = ws->tree;
    const Groups grp    = groupsFor(ws, first, last);
    const uint32_t groups = grp.num;
    // list blocks of 64 consecutive particles from first, whatever the target groups
    ws->nbrs.reserve(std::max<size_t>(1, size_t(ceilDiv(last - first, kGroup)) * paddedNgmax(p->ngmax) * kGroup));
    ws->counters.reserve(2 * kCounterSlots);
    fill64(ws, reinterpret_cast<uint64_t*>(ws->counters.get()), 0, 3);
    fill64(ws, reinterpret_cast<uint64_t*>(ws->counters.get()) + kStatWide, 0, kCounterSlots - kStatWide);

    double scale = 0;
    for (int d = 0; d < 3; ++d)
        scale = std::max(scale, std::max(std::abs(box.lo[d]), std::abs(box.hi[d])));

    SearchArgs a;
    a.x = f->x, a.y = f->y, a.z = f->z;
    a.childOffsets   = t.childOffsets.get();
    a.internalToLeaf = t.internalToLeaf.get();
    a.rankLo         = t.rankLo.get();
    a.childPush      = t.childPush.get();
    a.layout         = t.layout.get();
    a.centers        = t.centers.get();
    a.sizes          = t.sizes.get();
    a.box            = box;
    a.slack          = 1e-12 * scale;
    a.first = first, a.last = last, a.ng0 = p->ng0, a.ngmax = p->ngmax, a.ngmax4 = paddedNgmax(p->ngmax);
    a.groups = grp;
    a.iterateH = iterateH ? 1 : 0;
    a.nbrs     = ws->nbrs.get();
    a.nc       = f->nc;
    a.stats    = reinterpret_cast<unsigned long long*>(ws->counters.get());
    a.minLen   = std::min(box.len[0], std::min(box.len[1], box.len[2]));
    a.m = nullptr, a.wh = nullptr, a.sincN = 0, a.sincInt = -1, a.K = 0, a.xmOut = nullptr, a.toRho = 0;
    const bool f32 = p->hydro_type == SPHX_F32;
    if (xmassOut != kSearchOnly)
    {
        SPHX_REQUIRE(ws->tablesValid, SPHX_ERR_STATE, "kernel tables not set (sphx_workspace_set_kernel)");
        void* outField = xmassOut == kSearchRho ? f->rho : f->xm;
        SPHX_REQUIRE(f->m && outField, SPHX_ERR_ARG, "the fused XMass needs m and its output field (xm or rho)");
        a.m     = f->m;
        a.wh    = ws->wh64.get();
        a.sincN   = sincKernel<float>(ws).n;
        a.sincInt = sincKernel<float>(ws).ni;
        a.K     = p->K;
        a.xmOut = outField;
        a.toRho = xmassOut == kSearchRho;
    }

    if (groups > 0)
    {
        unsigned    blocks = ceilDiv(groups, kWavesPerBlock);
        KernelScope timer(ws, "search");
        const dim3  grid(blocks), block(kWavesPerBlock * kWave);
        if (xmassOut == kSearchOnly)
        {
            if (f32) searchKernel<float, false><<<grid, block, 0, ws->stream>>>(a, static_cast<float*>(f->h));
            else searchKernel<double, false><<<grid, block, 0, ws->stream>>>(a, static_cast<double*>(f->h));
        }
        else
        {
            if (f32) searchKernel<float, true><<<grid, block, 0, ws->stream>>>(a, static_cast<float*>(f->h));
            else searchKernel<double, true><<<grid, block, 0, ws->stream>>>(a, static_cast<double*>(f->h));
        }
        SPHX_CHECK_LAUNCH();
    }
    unsigned long long stats[kCounterSlots];
    SPHX_HIP(hipMemcpyAsync(stats, ws->counters.get(), sizeof(stats), hipMemcpyDeviceToHost, ws->stream));
    SPHX_HIP(hipStreamSynchronize(ws->stream));

    ws->nbFirst      = first;
    ws->nbLast       = last;
    ws->nbNgmax      = p->ngmax;
    ws->nbNc         = f->nc;
    ws->nbValid      = true;
    ++ws->searchSerial;
    ws->statSumPairs = stats[0];
    ws->statMaxCount = uint32_t(stats[1]);
    ws->nbWide       = stats[kStatWide] > 0;
    ws->statTests    = stats[kStatTests];
    std::memcpy(&ws->nbHMax, &stats[kStatHMax], sizeof(double));
    if (stats[kStatBadGroups] > 0)
    {
        ws->nbValid = false;
        throw Error{SPHX_ERR_ARG, std::to_string(stats[kStatBadGroups]) + " target groups hold more than 64 particles"};
    }
    return stats[2] > 0 ? SPHX_ERR_NOT_CONVERGED : SPHX_OK;
}

void copyNeighbors(sphx_workspace* ws, uint32_t* out, uint32_t first, uint32_t last)
{
    SPHX_REQUIRE(ws->nbValid, SPHX_ERR_STATE, "no neighbour lists: call find_neighbors first");
    SPHX_REQUIRE(ws->nbFirst <= first && first <= last && last <= ws->nbLast, SPHX_ERR_ARG,
                 "copy range must lie inside the searched range");
    SPHX_REQUIRE(out != nullptr, SPHX_ERR_ARG, "null output");
    if (last == first) { return; }
    size_t             len = size_t(last - first) * ws->nbNgmax;
    hipPointerAttribute_t attr{};
    bool               onDevice = hipPointerGetAttributes(&attr, out) == hipSuccess && attr.type == hipMemoryTypeDevice;
    (void)hipGetLastError();
    uint32_t* dst = out;
    DevBuf<uint32_t> staging;
    if (!onDevice)
    {
        staging.reserve(len);
        dst = staging.get();
    }
    copyListsKernel<<<ceilDiv(last - first, 256), 256, 0, ws->stream>>>(ws->nbrs.get(), ws->nbNc, dst, ws->nbFirst, first,
                                                                          last, ws->nbNgmax);
    SPHX_CHECK_LAUNCH();
    if (!onDevice)
    {
        SPHX_HIP(hipMemcpyAsync(out, dst, len * sizeof(uint32_t), hipMemcpyDeviceToHost, ws->stream));
        SPHX_HIP(hipStreamSynchronize(ws->stream));
        staging.release();
    }
}

} // namespace sphx
