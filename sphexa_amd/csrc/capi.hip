// extern "C" entry points of include/sphexa_amd.h: argument checks, error translation,
// kernel-table generation and the one-rank domain sync.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "common.hpp"

namespace sphx
{

namespace
{

//! the last error of a call without a workspace on this thread (sphx_last_error(NULL)): a workspace
//! entry point called with NULL, or an entry point that takes a device instead
thread_local std::string g_lastErrorNoWs;

//! runs an entry point's body with error translation; `writes` = it may change particle fields
template<class F>
int guarded(sphx_workspace* ws, F&& body, bool writes = true)
{
    if (ws == nullptr)
    {
        g_lastErrorNoWs = "null workspace";
        return SPHX_ERR_ARG;
    }
    if (writes) { ++ws->callSerial; }
    try
    {
        SPHX_HIP(hipSetDevice(ws->device));
        int r = body();
        if (r == SPHX_OK) { ws->lastError.clear(); }
        return r;
    }
    catch (const Error& e)
    {
        ws->lastError = e.msg;
        return e.code;
    }
    catch (const std::exception& e)
    {
        ws->lastError = e.what();
        return SPHX_ERR_ARG;
    }
}

//! guarded() for the entry points that take a device instead of a workspace
template<class F>
int guardedNoWorkspace(int device, F&& body)
{
    try
    {
        SPHX_REQUIRE(device >= 0, SPHX_ERR_ARG, "negative device index");
        SPHX_HIP(hipSetDevice(device));
        int r = body();
        if (r == SPHX_OK) { g_lastErrorNoWs.clear(); }
        return r;
    }
    catch (const Error& e)
    {
        g_lastErrorNoWs = e.msg;
        return e.code;
    }
    catch (const std::exception& e)
    {
        g_lastErrorNoWs = e.what();
        return SPHX_ERR_ARG;
    }
}

//! sinc(pi/2 v)^n and its v-derivative (sph/include/sph/kernels.hpp:36-61, sph_kernel_tables.hpp:105-109)
double sincPow(double v, double n)
{
    double s = v == 0.0 ? 1.0 : std::sin(M_PI_2 * v) / (M_PI_2 * v);
    return std::pow(s, n);
}

double sincPowDerivative(double v, double n)
{
    if (v == 0.0) { return 0.0; }
    const double pv = M_PI_2 * v;
    const double s  = std::sin(pv) / pv;
    const double ds = s * M_PI_2 * ((std::cos(pv) / std::sin(pv)) - 1.0 / pv);
    return n * std::pow(s, n - 1) * ds;
}

//! 1 / int_0^2 4 pi v^2 W(v) dv, composite Simpson with 2000 intervals and ascending partial sums
//! (sph_kernel_tables.hpp:31-60,78-83)
double normalization(double n)
{
    const uint64_t      intervals = 2000;
    const double        a = 0.0, b = 2.0, h = (b - a) / double(intervals);
    auto                f = [n](double v) { return 4.0 * M_PI * v * v * sincPow(v, n); };
    std::vector<double> odd(intervals / 2), even(intervals / 2 - 1);
    for (uint64_t i = 0; i < odd.size(); ++i)
        odd[i] = f(a + double(2 * i + 1) * h);
    for (uint64_t i = 0; i < even.size(); ++i)
        even[i] = f(a + double(2 * i + 2) * h);
    std::sort(odd.begin(), odd.end());
    std::sort(even.begin(), even.end());
    double so = std::accumulate(odd.begin(), odd.end(), 0.0), se = std::accumulate(even.begin(), even.end(), 0.0);
    return 1.0 / (h / 3.0 * (f(a) + f(b) + 4.0 * so + 2.0 * se));
}

//! tabulateFunction<T, 20000>(f, 0, 2): abscissae i * dx formed in T (sph_kernel_tables.hpp:88-101)
template<class T, class F>
std::vector<T> tabulate(F&& f)
{
    std::vector<T> t(kTableSize);
    const T        dx = T(2.0 / (kTableSize - 1));
    for (int i = 0; i < kTableSize; ++i)
    {
        T v  = T(0.0 + double(T(i * dx)));
        t[i] = T(f(double(v)));
    }
    return t;
}

//! interval k holds the table values at both of its ends; interval kTableSize-1, past the
//! support, holds zeros (the value lookups return there)
template<class T>
std::vector<TabPair<T>> intervals(const std::vector<T>& t)
{
    std::vector<TabPair<T>> p(kTableSize, TabPair<T>{T(0), T(0)});
    for (int k = 0; k + 1 < kTableSize; ++k) { p[k] = TabPair<T>{t[k], t[k + 1]}; }
    return p;
}

template<class T>
std::vector<TabQuad<T>> intervals(const std::vector<T>& w, const std::vector<T>& d)
{
    std::vector<TabQuad<T>> p(kTableSize, TabQuad<T>{T(0), T(0), T(0), T(0)});
    for (int k = 0; k + 1 < kTableSize; ++k) { p[k] = TabQuad<T>{w[k], w[k + 1], d[k], d[k + 1]}; }
    return p;
}

template<class E>
void upload(sphx_workspace* ws, DevBuf<E>& buf, const std::vector<E>& host)
{
    buf.reserve(host.size());
    SPHX_HIP(hipMemcpyAsync(buf.get(), host.data(), host.size() * sizeof(E), hipMemcpyHostToDevice, ws->stream));
}

void requireTables(const sphx_workspace* ws)
{
    SPHX_REQUIRE(ws->tablesValid, SPHX_ERR_STATE, "kernel tables not set (sphx_workspace_set_kernel)");
}

size_t hydroBytes(const sphx_params_t* p) { return p->hydro_type == SPHX_F32 ? 4 : 8; }

} // namespace
} // namespace sphx

using namespace sphx;

extern "C"
{

int sphx_workspace_create(sphx_workspace_t** out, int device, void* stream)
{
    if (out == nullptr) { return SPHX_ERR_ARG; }
    *out = nullptr;
    if (hipSetDevice(device) != hipSuccess)
    {
        (void)hipGetLastError();
        return SPHX_ERR_HIP;
    }
    auto* ws   = new sphx_workspace;
    ws->device = device;
    ws->stream = static_cast<hipStream_t>(stream);
    *out       = ws;
    return SPHX_OK;
}

int sphx_workspace_destroy(sphx_workspace_t* ws)
{
    if (ws == nullptr) { return SPHX_OK; }
    (void)hipSetDevice(ws->device);
    (void)hipStreamSynchronize(ws->stream);
    ws->wh64.release(), ws->whd64.release(), ws->whq64.release();
    for (auto& kv : ws->kernelEvents)
    {
        if (kv.second.start) { (void)hipEventDestroy(kv.second.start); }
        if (kv.second.stop) { (void)hipEventDestroy(kv.second.stop); }
    }
    ws->tree.releaseAll();
    ws->nbrs.release(), ws->scratch.release(), ws->keyTmp.release();
    ws->fieldTmp.release(), ws->order.release(), ws->records.release(), ws->counters.release(), ws->hostVisible.release(), ws->reduce64.release();
    ws->adj.pairStart.release(), ws->adj.rowStart.release(), ws->adj.recordOf.release(), ws->adj.keys.release(), ws->adj.sorted.release();
    ws->adj.pairBuf.release(), ws->records2.release(), ws->groupBuf.release(), ws->groupCount.release();
    ws->groupSplit.release();
    delete ws;
    return SPHX_OK;
}

const char* sphx_last_error(const sphx_workspace_t* ws) { return ws ? ws->lastError.c_str() : g_lastErrorNoWs.c_str(); }

int sphx_workspace_set_profiling(sphx_workspace_t* ws, int on)
{
    return guarded(ws, [&] {
        ws->profile = on != 0;
        for (auto& kv : ws->kernelEvents)
            kv.second.recorded = false;
        return SPHX_OK;
    });
}

int sphx_kernel_ms(sphx_workspace_t* ws, const char* kernel, double* ms)
{
    return guarded(ws, [&] {
        SPHX_REQUIRE(kernel && ms, SPHX_ERR_ARG, "null argument");
        auto it = ws->kernelEvents.find(kernel);
        SPHX_REQUIRE(it != ws->kernelEvents.end() && it->second.recorded, SPHX_ERR_STATE,
                     std::string("no timed launch of ") + kernel);
        SPHX_HIP(hipEventSynchronize(it->second.stop));
        float t = 0;
        SPHX_HIP(hipEventElapsedTime(&t, it->second.start, it->second.stop));
        *ms = t;
        return SPHX_OK;
    });
}

int sphx_workspace_set_kernel(sphx_workspace_t* ws, double sincIndex, double* K)
{
    return guarded(ws, [&] {
        SPHX_REQUIRE(sincIndex > 0, SPHX_ERR_ARG, "sinc index must be positive");
        auto w64  = tabulate<double>([sincIndex](double v) { return sincPow(v, sincIndex); });
        auto wd64 = tabulate<double>([sincIndex](double v) { return sincPowDerivative(v, sincIndex); });
        upload(ws, ws->wh64, intervals(w64));
        upload(ws, ws->whd64, intervals(wd64));
        upload(ws, ws->whq64, intervals(w64, wd64));
        SPHX_HIP(hipStreamSynchronize(ws->stream));
        ws->K           = normalization(sincIndex);
        ws->sincIndex   = sincIndex;
        ws->tablesValid = true;
        if (K) { *K = ws->K; }
        return SPHX_OK;
    });
}

int sphx_compute_sfc_keys(sphx_workspace_t* ws, const sphx_box_t* box, const double* x, const double* y,
                          const double* z, uint64_t* keys, size_t n)
{
    return guarded(ws, [&] {
        SPHX_REQUIRE(box && (n == 0 || (x && y && z && keys)), SPHX_ERR_ARG, "null argument");
        computeKeys(ws, makeDevBox(box), x, y, z, keys, n, true);
        return SPHX_OK;
    });
}

int sphx_sort_keys(sphx_workspace_t* ws, uint64_t* keys, uint32_t* order, size_t n)
{
    return guarded(ws, [&] {
        SPHX_REQUIRE(n == 0 || (keys && order), SPHX_ERR_ARG, "null argument");
        SPHX_REQUIRE(n < (size_t(1) << 32), SPHX_ERR_ARG, "more than 2^32-1 particles");
        sortKeys(ws, keys, order, n);
        return SPHX_OK;
    });
}

int sphx_sort_keys_noalloc(int device, void* stream, uint64_t* keys, uint32_t* order, size_t n)
{
    return guardedNoWorkspace(device, [&] {
        sortKeysNoAlloc(device, static_cast<hipStream_t>(stream), keys, order, n);
        return SPHX_OK;
    });
}

int sphx_gather(sphx_workspace_t* ws, const uint32_t* order, const void* in, void* out, size_t n, int elemBytes)
{
    return guarded(ws, [&] {
        SPHX_REQUIRE(n == 0 || (order && in && out && in != out), SPHX_ERR_ARG, "null or aliased argument");
        gather(ws, order, in, out, n, elemBytes);
        return SPHX_OK;
    });
}

int sphx_build_tree(sphx_workspace_t* ws, const uint64_t* sortedKeys, size_t n, uint32_t bucketSize,
                    const sphx_box_t* box)
{
    return guarded(ws, [&] {
        SPHX_REQUIRE(box && (n == 0 || sortedKeys), SPHX_ERR_ARG, "null argument");
        ws->nbValid  = false;
        ws->syncTree = false;
        buildTree(ws, sortedKeys, n, bucketSize, makeDevBox(box));
        return SPHX_OK;
    });
}

int sphx_tree_view(const sphx_workspace_t* wsc, sphx_tree_view_t* out)
{
    if (wsc == nullptr || out == nullptr) { return SPHX_ERR_ARG; }
    const Tree& t = wsc->tree;
    if (!t.valid) { return SPHX_ERR_STATE; }
    out->numLeafNodes     = t.numLeafNodes;
    out->numInternalNodes = t.numInternalNodes;
    out->numNodes         = t.numNodes;
    out->reserved         = 0;
    out->prefixes         = t.prefixes.get();
    out->childOffsets     = t.childOffsets.get();
    out->parents          = t.parents.get();
    out->internalToLeaf   = t.internalToLeaf.get();
    out->levelRange       = t.levelRange.get();
    out->leaves           = t.leaves.get();
    out->counts           = t.counts.get();
    out->layout           = t.layout.get();
    out->centers          = t.centers.get();
    out->sizes            = t.sizes.get();
    return SPHX_OK;
}

int sphx_domain_sync(sphx_workspace_t* ws, const sphx_box_t* box, sphx_fields_t* f, const sphx_params_t* prm,
                     size_t n, uint32_t bucketSize)
{
    return guarded(ws, [&] {
        SPHX_REQUIRE(box && f && prm, SPHX_ERR_ARG, "null argument");
        SPHX_REQUIRE(n == 0 || (f->x && f->y && f->z), SPHX_ERR_ARG, "x, y, z are required");
        SPHX_REQUIRE(n < (size_t(1) << 32), SPHX_ERR_ARG, "more than 2^32-1 particles");
        ws->nbValid  = false;
        DevBox b     = makeDevBox(box);
        ws->keyTmp.reserve(n);
        uint64_t* keys = f->keys;
        DevBuf<uint64_t> own;
        if (keys == nullptr)
        {
            own.reserve(n);
            keys = own.get();
        }
        {
            KernelScope timer(ws, "sfc_keys");
            computeKeys(ws, b, f->x, f->y, f->z, keys, n, false); // the sync keys every particle
        }
        ws->order.reserve(n);
        DevBuf<uint32_t>& order = ws->order;
        {
            KernelScope timer(ws, "sort");
            sortKeys(ws, keys, order.get(), n);
        }

        // permute every allocated field, batched per element size (the permutation read once per batch)
        std::vector<std::pair<void*, int>> fields;
        for (double* r : {f->x, f->y, f->z, f->u, f->temp, f->du})
            if (r) fields.push_back({r, 8});
        int   hb      = int(hydroBytes(prm));
        void* hyd[]   = {f->h,   f->m,    f->vx,   f->vy,    f->vz,    f->rho,   f->p,    f->c,   f->prho, f->tdpdTrho,
                         f->c11, f->c12,  f->c13,  f->c22,   f->c23,   f->c33,   f->xm,   f->kx,  f->gradh, f->divv,
                         f->curlv, f->alpha, f->ax, f->ay,   f->az,    f->dV11,  f->dV12, f->dV13, f->dV22, f->dV23,
                         f->dV33};
        for (void* h : hyd)
            if (h) fields.push_back({h, hb});
        if (f->nc) fields.push_back({f->nc, 4});
        {
            KernelScope timer(ws, "permute");
            permuteFields(ws, order.get(), fields, n);
        }
        {
            KernelScope timer(ws, "tree");
            const bool update = (prm->flags & SPHX_FLAG_TREE_UPDATE) && ws->syncTree && ws->tree.valid &&
                                ws->syncBucket == bucketSize;
            ws->syncTree = false;
            if (update) { updateTree(ws, keys, n, bucketSize, b); }
            else { buildTree(ws, keys, n, bucketSize, b); }
        }
        SPHX_HIP(hipStreamSynchronize(ws->stream));
        own.release();
        ws->syncN      = n;
        ws->syncTree   = true;
        ws->syncBucket = bucketSize;
        return SPHX_OK;
    });
}

int sphx_global_box(sphx_workspace_t* ws, sphx_box_t* box, const double* x, const double* y, const double* z,
                    size_t n, int firstCall)
{
#pragma clang fp contract(off)
    return guarded(ws, [&] {
        SPHX_REQUIRE(box, SPHX_ERR_ARG, "null argument");
        SPHX_REQUIRE(n == 0 || (x && y && z), SPHX_ERR_ARG, "x, y, z are required");
        if (box->boundary[0] == 1 && box->boundary[1] == 1 && box->boundary[2] == 1) { return SPHX_OK; }
        double ext[6];
        coordinateExtent(ws, x, y, z, n, ext);
        for (int d = 0; d < 3; ++d)
        {
            if (box->boundary[d] == 1) { continue; } // periodic: the limits stay
            double lo = ext[2 * d], hi = ext[2 * d + 1];
            if (!firstCall)
            {
                // limitBoxShrinking with shrink limit 0.05 of the previous edge length
                const double step  = 0.05 * (box->hi[d] - box->lo[d]);
                const double lower = box->lo[d] + step;
                const double upper = box->hi[d] - step;
                lo                 = std::min(lo, lower);
                hi                 = std::max(hi, upper);
            }
            box->lo[d] = lo;
            box->hi[d] = hi;
        }
        return SPHX_OK;
    });
}

int sphx_domain_reorder(sphx_workspace_t* ws, void* const* fields, const int* elemBytes, size_t count, size_t n)
{
    return guarded(ws, [&] {
        SPHX_REQUIRE(count == 0 || (fields && elemBytes), SPHX_ERR_ARG, "null argument");
        SPHX_REQUIRE(n == ws->syncN, SPHX_ERR_STATE, "no sphx_domain_sync over the same particle count");
        std::vector<std::pair<void*, int>> batch;
        for (size_t k = 0; k < count; ++k)
        {
            SPHX_REQUIRE(elemBytes[k] == 4 || elemBytes[k] == 8, SPHX_ERR_ARG, "element size must be 4 or 8 bytes");
            SPHX_REQUIRE(fields[k] || n == 0, SPHX_ERR_ARG, "null field");
            batch.push_back({fields[k], elemBytes[k]});
        }
        permuteFields(ws, ws->order.get(), batch, n);
        return SPHX_OK;
    });
}

int sphx_find_neighbors(sphx_workspace_t* ws, const sphx_box_t* box, const sphx_fields_t* f,
                        const sphx_params_t* prm, uint32_t first, uint32_t last, int iterateH)
{
    return guarded(ws, [&] {
        SPHX_REQUIRE(box && f && prm, SPHX_ERR_ARG, "null argument");
        return findNeighbors(ws, makeDevBox(box), f, prm, first, last, iterateH != 0);
    });
}

int sphx_find_neighbors_xmass(sphx_workspace_t* ws, const sphx_box_t* box, const sphx_fields_t* f,
                              const sphx_params_t* prm, uint32_t first, uint32_t last, int iterateH, int out)
{
    return guarded(ws, [&] {
        SPHX_REQUIRE(box && f && prm, SPHX_ERR_ARG, "null argument");
        SPHX_REQUIRE(out == SPHX_XMASS_XM || out == SPHX_XMASS_RHO, SPHX_ERR_ARG,
                     "out must be SPHX_XMASS_XM or SPHX_XMASS_RHO");
        return findNeighbors(ws, makeDevBox(box), f, prm, first, last, iterateH != 0,
                             out == SPHX_XMASS_RHO ? kSearchRho : kSearchXm);
    });
}

int sphx_copy_neighbors(sphx_workspace_t* ws, uint32_t* out, uint32_t first, uint32_t last)
{
    return guarded(ws, [&] {
        copyNeighbors(ws, out, first, last);
        return SPHX_OK;
    });
}

int sphx_neighbor_stats(const sphx_workspace_t* ws, uint64_t* sumPairs, uint32_t* maxCount)
{
    if (ws == nullptr) { return SPHX_ERR_ARG; }
    if (!ws->nbValid) { return SPHX_ERR_STATE; }
    if (sumPairs) { *sumPairs = ws->statSumPairs; }
    if (maxCount) { *maxCount = ws->statMaxCount; }
    return SPHX_OK;
}

int sphx_compute_groups(sphx_workspace_t* ws, const sphx_box_t* box, const sphx_fields_t* f, uint32_t first,
                        uint32_t last, float tolFactor, sphx_groups_t* out)
{
    return guarded(ws, [&] {
        SPHX_REQUIRE(box && f && out, SPHX_ERR_ARG, "null argument");
        computeGroups(ws, makeDevBox(box), f->x, f->y, f->z, first, last, tolFactor, out);
        return SPHX_OK;
    }, false);
}

int sphx_set_groups(sphx_workspace_t* ws, const sphx_groups_t* groups)
{
    return guarded(ws, [&] {
        if (groups == nullptr)
        {
            ws->groupsSet = false;
            ws->groupView = sphx_groups_t{};
            return SPHX_OK;
        }
        SPHX_REQUIRE(groups->firstBody <= groups->lastBody, SPHX_ERR_ARG, "invalid group range");
        SPHX_REQUIRE(groups->numGroups == 0 || (groups->groupStart && groups->groupEnd), SPHX_ERR_ARG,
                     "null group arrays");
        ws->groupsSet = true;
        ws->groupView = *groups;
        return SPHX_OK;
    }, false);
}

int sphx_search_tests(const sphx_workspace_t* ws, uint64_t* tests)
{
    if (ws == nullptr || tests == nullptr) { return SPHX_ERR_ARG; }
    if (!ws->nbValid) { return SPHX_ERR_STATE; }
    *tests = ws->statTests;
    return SPHX_OK;
}

#define SPHX_PAIR_ENTRY(name, fn)                                                                                      \
    int name(sphx_workspace_t* ws, const sphx_box_t* box, const sphx_fields_t* f, const sphx_params_t* prm,            \
             uint32_t first, uint32_t last)                                                                            \
    {                                                                                                                  \
        return guarded(ws, [&] {                                                                                       \
            SPHX_REQUIRE(box && f && prm, SPHX_ERR_ARG, "null argument");                                              \
            requireTables(ws);                                                                                         \
            fn(ws, pairBox(ws, box), f, prm, first, last);                                                              \
            return SPHX_OK;                                                                                            \
        });                                                                                                            \
    }

SPHX_PAIR_ENTRY(sphx_density_std, densityStd)
SPHX_PAIR_ENTRY(sphx_iad_std, iadStd)
SPHX_PAIR_ENTRY(sphx_xmass, xmassVe)
SPHX_PAIR_ENTRY(sphx_ve_def_gradh, veDefGradh)
SPHX_PAIR_ENTRY(sphx_iad_divv_curlv, iadDivvCurlv)
SPHX_PAIR_ENTRY(sphx_av_switches, avSwitches)

int sphx_eos_std(sphx_workspace_t* ws, const sphx_fields_t* f, const sphx_params_t* prm, uint32_t first,
                 uint32_t last)
{
    return guarded(ws, [&] {
        SPHX_REQUIRE(f && prm, SPHX_ERR_ARG, "null argument");
        eosStd(ws, f, prm, first, last);
        ws->eosSerial = ws->callSerial;
        ws->eosFields = *f;
        ws->eosKind   = 1;
        return SPHX_OK;
    });
}

int sphx_eos_ve(sphx_workspace_t* ws, const sphx_fields_t* f, const sphx_params_t* prm, uint32_t first,
                uint32_t last)
{
    return guarded(ws, [&] {
        SPHX_REQUIRE(f && prm, SPHX_ERR_ARG, "null argument");
        eosVe(ws, f, prm, first, last);
        ws->eosSerial = ws->callSerial;
        ws->eosFields = *f;
        ws->eosKind   = 2;
        return SPHX_OK;
    });
}

int sphx_rho_timestep(sphx_workspace_t* ws, const sphx_fields_t* f, const sphx_params_t* prm, uint32_t first,
                      uint32_t last, sphx_timestep_t* ts)
{
    return guarded(ws, [&] {
        SPHX_REQUIRE(f && prm && ts, SPHX_ERR_ARG, "null argument");
        rhoTimestep(ws, f, prm, first, last, ts);
        return SPHX_OK;
    }, false);
}

int sphx_compute_timestep(sphx_workspace_t* ws, const sphx_fields_t* f, const sphx_params_t* prm, uint32_t first,
                          uint32_t last, sphx_timestep_t* ts)
{
    return guarded(ws, [&] {
        SPHX_REQUIRE(f && prm && ts, SPHX_ERR_ARG, "null argument");
        computeTimestep(ws, f, prm, first, last, ts);
        return SPHX_OK;
    }, false);
}

int sphx_compute_positions(sphx_workspace_t* ws, const sphx_box_t* box, const sphx_fields_t* f,
                           const sphx_history_t* hist, const sphx_params_t* prm, uint32_t first, uint32_t last,
                           double dt, double dt_m1, double constCv)
{
    return guarded(ws, [&] {
        SPHX_REQUIRE(box && f && hist && prm, SPHX_ERR_ARG, "null argument");
        computePositions(ws, box, f, hist, prm, first, last, dt, dt_m1, constCv);
        return SPHX_OK;
    });
}

int sphx_compute_positions_rungs(sphx_workspace_t* ws, const sphx_box_t* box, const sphx_fields_t* f,
                                 const sphx_history_t* hist, const sphx_params_t* prm, uint32_t first,
                                 uint32_t last, double dt, const float* dtM1, int numRungs, const uint8_t* rung,
                                 double constCv)
{
    return guarded(ws, [&] {
        SPHX_REQUIRE(box && f && hist && prm && dtM1, SPHX_ERR_ARG, "null argument");
        computePositions(ws, box, f, hist, prm, first, last, dt, dtM1[0], constCv, dtM1, numRungs, rung);
        return SPHX_OK;
    });
}

int sphx_drift_positions(sphx_workspace_t* ws, const sphx_fields_t* f, const sphx_history_t* hist,
                         const sphx_params_t* prm, uint32_t first, uint32_t last, double dt, double dtBack,
                         const float* dtM1, int numRungs, const uint8_t* rung, double constCv)
{
    return guarded(ws, [&] {
        SPHX_REQUIRE(f && hist && prm && dtM1, SPHX_ERR_ARG, "null argument");
        driftPositions(ws, f, hist, prm, first, last, dt, dtBack, dtM1, numRungs, rung, constCv);
        return SPHX_OK;
    });
}

int sphx_group_divv_timestep(sphx_workspace_t* ws, const sphx_fields_t* f, const sphx_params_t* prm,
                             uint32_t first, uint32_t last, float Krho, float* groupDt)
{
    return guarded(ws, [&] {
        SPHX_REQUIRE(f && prm, SPHX_ERR_ARG, "null argument");
        groupTimestep(ws, f, prm, first, last, false, Krho, groupDt);
        return SPHX_OK;
    });
}

int sphx_group_acc_timestep(sphx_workspace_t* ws, const sphx_fields_t* f, const sphx_params_t* prm,
                            uint32_t first, uint32_t last, float etaAcc, float* groupDt)
{
    return guarded(ws, [&] {
        SPHX_REQUIRE(f && prm, SPHX_ERR_ARG, "null argument");
        groupTimestep(ws, f, prm, first, last, true, etaAcc, groupDt);
        return SPHX_OK;
    });
}

int sphx_store_rung(sphx_workspace_t* ws, uint32_t first, uint32_t last, uint8_t rung, uint8_t* particleRungs)
{
    return guarded(ws, [&] {
        storeRung(ws, first, last, rung, particleRungs);
        return SPHX_OK;
    });
}

int sphx_update_smoothing_length(sphx_workspace_t* ws, const sphx_fields_t* f, const sphx_params_t* prm,
                                 uint32_t first, uint32_t last)
{
    return guarded(ws, [&] {
        SPHX_REQUIRE(f && prm, SPHX_ERR_ARG, "null argument");
        updateSmoothingLength(ws, f, prm, first, last);
        return SPHX_OK;
    });
}

int sphx_conserved_quantities(sphx_workspace_t* ws, const sphx_fields_t* f, const sphx_params_t* prm,
                              uint32_t first, uint32_t last, double cv, sphx_conserved_t* out)
{
    return guarded(ws, [&] {
        SPHX_REQUIRE(f && prm && out, SPHX_ERR_ARG, "null argument");
        conservedQuantities(ws, f, prm, first, last, cv, out);
        return SPHX_OK;
    }, false);
}

int sphx_momentum_energy_std(sphx_workspace_t* ws, const sphx_box_t* box, const sphx_fields_t* f,
                             const sphx_params_t* prm, uint32_t first, uint32_t last, double* minDtCourant)
{
    return guarded(ws, [&] {
        SPHX_REQUIRE(box && f && prm, SPHX_ERR_ARG, "null argument");
        requireTables(ws);
        double dt = momentumEnergyStd(ws, pairBox(ws, box), f, prm, first, last);
        if (minDtCourant) { *minDtCourant = dt; }
        return SPHX_OK;
    });
}

#define SPHX_VJP_ENTRY(name, fn)                                                                                        \
    int name(sphx_workspace_t* ws, const sphx_box_t* box, const sphx_fields_t* f, const sphx_params_t* prm,           \
             uint32_t first, uint32_t last, const sphx_adjoint_t* in, sphx_adjoint_t* out)                             \
    {                                                                                                                  \
        return guarded(ws, [&] {                                                                                       \
            SPHX_REQUIRE(box && f && prm && in && out, SPHX_ERR_ARG, "null argument");                                 \
            requireTables(ws);                                                                                         \
            fn(ws, pairBox(ws, box), f, prm, first, last, in, out);                                                     \
            return SPHX_OK;                                                                                            \
        });                                                                                                            \
    }

SPHX_VJP_ENTRY(sphx_density_std_vjp, densityStdVjp)
SPHX_VJP_ENTRY(sphx_iad_std_vjp, iadStdVjp)
SPHX_VJP_ENTRY(sphx_momentum_energy_std_vjp, momentumEnergyStdVjp)

int sphx_momentum_energy_ve(sphx_workspace_t* ws, const sphx_box_t* box, const sphx_fields_t* f,
                            const sphx_params_t* prm, uint32_t first, uint32_t last, int avClean,
                            double* minDtCourant)
{
    return guarded(ws, [&] {
        SPHX_REQUIRE(box && f && prm, SPHX_ERR_ARG, "null argument");
        requireTables(ws);
        double dt = momentumEnergyVe(ws, pairBox(ws, box), f, prm, first, last, avClean != 0);
        if (minDtCourant) { *minDtCourant = dt; }
        return SPHX_OK;
    });
}

} // extern "C"
