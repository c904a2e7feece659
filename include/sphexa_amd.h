/*
 * sphexa_amd -- C ABI of the MI355X (gfx950) SPH hydro hot path.
 *
 * One extern "C" entry point per phase of the reference's hydro step. Each one
 * replaces the GPU plug-in point of the reference (the `HaveGpu` branch of the
 * public template, dispatching to the extern templates declared in
 * sph/include/sph/sph_gpu.hpp:11-98); the reference function each entry point
 * stands in for is cited above its declaration (paths relative to the reference
 * repository root).
 *
 * Conventions
 *  - All field pointers are DEVICE pointers owned by the caller (the Dataset);
 *    kernels borrow them and write outputs in place. Scratch (tree, neighbour
 *    lists, counters) is owned by the workspace.
 *  - Fields are SoA arrays indexed by SFC-sorted particle position, laid out as in
 *    sphexa::ParticlesData (sph/include/sph/particles_data.hpp:220-266):
 *    x,y,z,u,temp,du are RealType (double); the HydroType fields are double
 *    (plain FP64 form, SPHX_F64) or float (the reference's SphTypes layout with the
 *    mixed-precision kernels, SPHX_F32), selected by sphx_params_t.hydro_type.
 *  - Work is queued on the workspace's stream (the stream passed at creation);
 *    entry points that return a scalar (minDtCourant) or a status that depends on
 *    device results synchronise that stream before returning.
 *  - Every entry point returns an sphx_status (0 = success) and never exits the
 *    process. Workspaces are independent: calls on different workspaces may run
 *    concurrently; calls on one workspace must be serialised by the caller.
 */
#ifndef SPHEXA_AMD_H
#define SPHEXA_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum sphx_status
{
    SPHX_OK                = 0,
    SPHX_ERR_HIP           = 1, /* a HIP runtime call or kernel launch failed */
    SPHX_ERR_ARG           = 2, /* invalid argument: null pointer, empty/inverted range, ng0 > ngmax */
    SPHX_ERR_NOT_CONVERGED = 3, /* coupled h / neighbour-count iteration did not converge in 10 rounds */
    SPHX_ERR_STATE         = 4, /* called out of order, e.g. pair loop before find_neighbors */
    SPHX_ERR_ALLOC         = 5  /* device allocation failed */
} sphx_status;

typedef enum sphx_dtype
{
    SPHX_F64 = 0,
    SPHX_F32 = 1
} sphx_dtype;

/* cstone::Box<double> (domain/include/cstone/sfc/box.hpp:95-174); boundary: 0 open, 1 periodic,
 * 2 fixed (BoundaryType, box.hpp:80-85: open for the search; the integrator leaves particles at rest
 * within 2h of a fixed face) */
typedef struct sphx_box
{
    double  lo[3];
    double  hi[3];
    int32_t boundary[3];
} sphx_box_t;

/* Device pointers to the ParticlesData fields used on the hot path
 * (sph/include/sph/particles_data.hpp:220-248). Unused fields may be NULL. */
typedef struct sphx_fields
{
    /* RealType (double) fields */
    double* x;
    double* y;
    double* z;
    double* u;
    double* temp;
    double* du;
    /* HydroType fields (double or float, see sphx_params_t.hydro_type); m is Tmass (same type) */
    void* h;
    void* m;
    void* vx;
    void* vy;
    void* vz;
    void* rho;
    void* p;
    void* c;
    void* prho;
    void* tdpdTrho;
    void* c11;
    void* c12;
    void* c13;
    void* c22;
    void* c23;
    void* c33;
    void* xm;
    void* kx;
    void* gradh;
    void* divv;
    void* curlv;
    void* alpha;
    void* ax;
    void* ay;
    void* az;
    void* dV11;
    void* dV12;
    void* dV13;
    void* dV22;
    void* dV23;
    void* dV33;
    /* integer fields */
    uint64_t* keys;
    uint32_t* nc;
} sphx_fields_t;

/* Dataset scalars (particles_data.hpp:91-146) */
typedef struct sphx_params
{
    double   K;              /* kernel normalisation, see sphx_workspace_set_kernel */
    double   Kcour;          /* Courant factor (passed to tsKCourant as float, like the reference) */
    double   gamma;          /* adiabatic index */
    double   muiConst;       /* mean molecular weight, for temperature-based EOS */
    double   minDt;          /* current global time step (AV switches) */
    double   alphamin, alphamax, decay_constant;
    double   Atmin, Atmax, ramp;
    uint32_t ng0, ngmax;
    int32_t  hydro_type;     /* sphx_dtype of the HydroType fields */
    int32_t  flags;          /* SPHX_FLAG_* bits, 0 = reference behaviour */
    int32_t  eos;            /* eosChoice (sph/include/sph/eos.hpp:10-15): SPHX_EOS_IDEAL_GAS (0), SPHX_EOS_ISOTHERMAL,
                                SPHX_EOS_POLYTROPIC */
    int32_t  reserved;
    double   soundSpeedConst;  /* isothermal sound speed (particles_data.hpp:127, default 1) */
    double   polytropic_const; /* polytropic constant K (particles_data.hpp:123, default 1) */
    double   polytropic_index; /* polytropic exponent (particles_data.hpp:121, default 5/3) */
} sphx_params_t;

/* sphx_params_t.eos, the reference's EosType (sph/include/sph/eos.hpp:10-15) */
#define SPHX_EOS_IDEAL_GAS 0
#define SPHX_EOS_ISOTHERMAL 1
#define SPHX_EOS_POLYTROPIC 2

/* std momentum-energy with a per-particle AV coefficient from fields.alpha instead of the
 * reference's constant alpha = 1 (hydro_std/momentum_energy_kern.hpp:98); the coefficient the
 * std adjoint differentiates against. */
#define SPHX_FLAG_STD_ALPHA 1

/* Record chaining. Set on the pair-loop calls of a pass over the full range [0, n), it promises
 * that the caller changes no particle field between consecutive calls on one workspace. Then
 * sphx_xmass writes the neighbour records of sphx_ve_def_gradh in its epilogue,
 * sphx_ve_def_gradh those of sphx_iad_divv_curlv, sphx_iad_divv_curlv those of
 * sphx_av_switches, sphx_av_switches those of sphx_momentum_energy_ve, sphx_density_std those of
 * sphx_iad_std and sphx_iad_std those of sphx_momentum_energy_std, so the next call skips its
 * packing kernel. The next call uses them
 * only
 * if it comes directly after the producing one (sphx_rho_timestep, sphx_compute_timestep and
 * sphx_conserved_quantities, which read fields only, may sit between, and one sphx_eos_ve or
 * sphx_eos_std on the same fields between sphx_ve_def_gradh and sphx_iad_divv_curlv or between
 * sphx_density_std and sphx_iad_std), on the same fields, box,
 * form and particle count; otherwise it packs as without the flag. Results are bitwise the same
 * with or without the flag. */
#define SPHX_FLAG_CHAIN 2

/* Tree update of the syncs after a Domain's first (domain/include/cstone/domain/domain.hpp:199-212).
 * Set on sphx_domain_sync, the tree of the workspace's previous sync (same bucket size) is updated
 * ONCE instead of converged from scratch: one rebalance decided on its counts (the previous keys),
 * merging sibling groups whose parent holds at most bucketSize particles and splitting in 8 the leaves
 * holding more, then the counts of the new keys, as the reference's focused octree does on one rank
 * (FocusedOctreeSingleNode::update, focus/octree_focus.hpp:348-385). Leaves may then hold more than
 * bucketSize particles until later syncs split them; the neighbour search and the pair loops take any
 * leaf size. Without a previous sync tree (first call, or after sphx_build_tree) the sync converges,
 * as the reference's first sync does. The C++ Domain sets it on every sync after its first. */
#define SPHX_FLAG_TREE_UPDATE 4

/* Previous-step fields of the time integrator (particles_data.hpp:221,233,238). x_m1, y_m1, z_m1
 * (X_n - X_n-1) and du_m1 are XM1Type, which is the HydroType here (double in the plain form,
 * float in the mixed form, as SphTypes); mui is HydroType, NULL for the constant muiConst. */
typedef struct sphx_history
{
    void* x_m1;
    void* y_m1;
    void* z_m1;
    void* du_m1;
    void* mui;
} sphx_history_t;

/* Global time-step state, the Dataset scalars of particles_data.hpp:96-141 (defaults in
 * parentheses): minDt, minDt_m1 (1e-12), ttot (0), minDtCourant and minDtRho (INFINITY),
 * maxDtIncrease (1.1), g (0: no acceleration limit), etaAcc (0.2), eps (0.005), Krho (0.06). */
typedef struct sphx_timestep
{
    double minDt, minDt_m1, ttot;
    double minDtCourant, minDtRho;
    double maxDtIncrease;
    double g, etaAcc, eps;
    double Krho;
} sphx_timestep_t;

/* Cotangents and gradients of the adjoint entry points: one HydroType array per named quantity
 * (du included). An adjoint reads the cotangents of its forward outputs from `in` and adds
 * (+=) the gradients of its differentiable inputs into `out`, so callers zero `out` first; the
 * members each entry point uses are listed with it. */
typedef struct sphx_adjoint
{
    void *x, *y, *z, *h, *rho, *p, *alpha;
    void *c11, *c12, *c13, *c22, *c23, *c33;
    void *ax, *ay, *az, *du;
} sphx_adjoint_t;

/* OctreeNsView<double, uint64_t> (domain/include/cstone/tree/octree.hpp:280-300) plus the
 * cornerstone leaf counts; all device pointers into workspace-owned memory. */
typedef struct sphx_tree_view
{
    int32_t         numLeafNodes;
    int32_t         numInternalNodes;
    int32_t         numNodes;
    int32_t         reserved;
    const uint64_t* prefixes;       /* [numNodes] placeholder-bit node keys, level-major order */
    const int32_t*  childOffsets;   /* [numNodes] first child, 0 = leaf */
    const int32_t*  parents;        /* [(numNodes-1)/8] */
    const int32_t*  internalToLeaf; /* [numNodes] leaf index for leaves, negative for internal nodes */
    const int32_t*  levelRange;     /* [23] */
    const uint64_t* leaves;         /* [numLeafNodes+1] cornerstone leaf keys */
    const uint32_t* counts;         /* [numLeafNodes] particles per leaf */
    const uint32_t* layout;         /* [numLeafNodes+1] first particle of each leaf */
    const double*   centers;        /* [3*numNodes] geometric centres (x,y,z interleaved) */
    const double*   sizes;          /* [3*numNodes] half sizes */
} sphx_tree_view_t;

/* cstone::GroupView (domain/include/cstone/traversal/groups.hpp:28-34): target groups of at most 64
 * particles, [groupStart[g], groupEnd[g]) for g < numGroups, inside [firstBody, lastBody); device
 * pointers. groupDt, optional (NULL), is the per-group time step the momentum loops lower to the
 * group's smallest Courant step (groupDt[g] = min(groupDt[g], ...), momentum_energy_gpu.cu:96-104). */
typedef struct sphx_groups
{
    uint32_t        firstBody, lastBody;
    uint32_t        numGroups;
    uint32_t        reserved;
    const uint32_t* groupStart;
    const uint32_t* groupEnd;
    float*          groupDt;
} sphx_groups_t;

typedef struct sphx_workspace sphx_workspace_t;

/* ---- workspace ------------------------------------------------------------------------------- */

/* Create a workspace on `device`; all work is queued on `stream` (a hipStream_t, NULL = default
 * stream). Replaces the Dataset-owned scratch of the reference: devData.traversalStack grown by
 * allocateNcStacks (domain/include/cstone/traversal/find_neighbors.cuh:493-505). */
int sphx_workspace_create(sphx_workspace_t** ws, int device, void* stream);
int sphx_workspace_destroy(sphx_workspace_t* ws);

/* Tabulate W and dW/dv (20000 points on [0,2]) for sinc kernel index `sincIndex` and upload them
 * in both precisions; returns K through *K. Replaces ParticlesData::createTables
 * (sph/include/sph/particles_data.hpp:380-387) + DeviceParticlesData::uploadTables. */
int sphx_workspace_set_kernel(sphx_workspace_t* ws, double sincIndex, double* K);

/* Human-readable text for the last error recorded on this workspace; with NULL, that of the last
 * call without a workspace (sphx_sort_keys_noalloc) on the calling thread. */
const char* sphx_last_error(const sphx_workspace_t* ws);

/* Kernel timing (no counterpart in the reference; for benchmarks and profiling). When on, the
 * search and every pair-loop entry point record HIP events on the workspace stream around their
 * main kernel (not around the record packing or host reads). sphx_kernel_ms waits for and returns
 * the duration of the last such launch of `kernel`: "search", "density_std", "iad_std",
 * "momentum_std", "xmass", "ve_def_gradh", "iad_divv_curlv", "av_switches", "momentum_ve",
 * "positions", "update_h"; and the stages of sphx_domain_sync: "sfc_keys", "sort", "permute",
 * "tree" (the last with its host synchronisations).
 * Returns SPHX_ERR_STATE when that kernel has not run since profiling was switched on. */
int sphx_workspace_set_profiling(sphx_workspace_t* ws, int on);
int sphx_kernel_ms(sphx_workspace_t* ws, const char* kernel, double* ms);

/* ---- keys, sort, tree (one-rank Domain::sync) ------------------------------------------------ */

/* 64-bit Hilbert keys of n particles; keys equal to 2^63 (removeKey) are left unchanged.
 * Replaces computeSfcKeysGpu (domain/include/cstone/sfc/sfc_gpu.cu:25-43). */
int sphx_compute_sfc_keys(sphx_workspace_t* ws, const sphx_box_t* box, const double* x, const double* y,
                          const double* z, uint64_t* keys, size_t n);

/* Stable sort of keys; order[i] receives the original index of the i-th sorted key.
 * Replaces sortByKeyGpu / GpuSfcSorter::setMapFromCodes (domain/include/cstone/primitives/
 * primitives_gpu.cu:323-346, primitives/gather.cuh:50-128). */
int sphx_sort_keys(sphx_workspace_t* ws, uint64_t* keys, uint32_t* order, size_t n);

/* sphx_sort_keys without a workspace and without any device allocation. No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). Same keys and order, bit for bit, as
 * sphx_sort_keys. Runs on `stream` (a hipStream_t, NULL = default stream) of GPU `device`, which it
 * makes current. Its temporaries live in a fixed device array of the library, 32 MiB on every GPU
 * it runs on, held until the process ends: n is limited to SPHX_SORT_NOALLOC_MAX_KEYS, and larger n
 * returns SPHX_ERR_ARG with the message in sphx_last_error(NULL) (per thread). Calls share that
 * array per GPU, so calls that overlap on one GPU (other streams, other threads) race on it and
 * must be serialised by the caller. */
#define SPHX_SORT_NOALLOC_MAX_KEYS 1048576
int sphx_sort_keys_noalloc(int device, void* stream, uint64_t* keys, uint32_t* order, size_t n);

/* out[i] = in[order[i]] for elements of `elemBytes` (4 or 8). Replaces gatherGpu
 * (domain/include/cstone/primitives/primitives_gpu.h:34-40). */
int sphx_gather(sphx_workspace_t* ws, const uint32_t* order, const void* in, void* out, size_t n, int elemBytes);

/* Cornerstone leaves (bucket size `bucketSize`), linked octree, leaf layout and geometric node
 * centres from sorted keys, kept in the workspace. Replaces computeOctree / buildOctreeGpu /
 * computeNodeLayout / nodeFpCenters as combined by Domain::sync + octreeProperties on one rank
 * (domain/include/cstone/tree/csarray.hpp:430-440, tree/octree_gpu.cu:137-157,
 * focus/source_center.hpp:130-141, domain/include/cstone/domain/domain.hpp:416-428). */
int sphx_build_tree(sphx_workspace_t* ws, const uint64_t* sortedKeys, size_t n, uint32_t bucketSize,
                    const sphx_box_t* box);

/* Device view of the tree built by the last sphx_build_tree / sphx_domain_sync. */
int sphx_tree_view(const sphx_workspace_t* ws, sphx_tree_view_t* out);

/* The box update at the start of Domain::sync (GlobalAssignment::assign,
 * domain/include/cstone/domain/assignment.hpp:80-82, with makeGlobalBox, sfc/box_mpi.hpp:69-100, and
 * limitBoxShrinking, sfc/box.hpp:398-412): per non-periodic dimension, the min and max of the n
 * coordinates; with firstCall = 0 each side then moves inwards by at most 0.05 of the previous edge
 * length. Periodic dimensions keep their limits. Updates box in place (reads its limits as the
 * previous box). */
int sphx_global_box(sphx_workspace_t* ws, sphx_box_t* box, const double* x, const double* y, const double* z,
                    size_t n, int firstCall);

/* keys -> stable sort -> permute every non-NULL field of `f` -> tree (converged, or updated once with
 * SPHX_FLAG_TREE_UPDATE in prm->flags). One-rank Domain::sync
 * (domain/include/cstone/domain/domain.hpp:182-234). As there, only the conserved fields need to be
 * passed (std: x y z h m u vx vy vz, HydroProp::ConservedFields, main/src/propagator/std_hydro.hpp:73;
 * VE: x y z h m temp vx vy vz alpha, ve_hydro.hpp:72): a pass recomputes every dependent field
 * before it reads it, so leaving those NULL here saves their permutation. keys, if given, receives
 * the sorted SFC keys. */
int sphx_domain_sync(sphx_workspace_t* ws, const sphx_box_t* box, sphx_fields_t* f, const sphx_params_t* prm,
                     size_t n, uint32_t bucketSize);

/* Reorder `count` more per-particle arrays (fields[k] of elemBytes[k] = 4 or 8 bytes each) by the
 * permutation of the last sphx_domain_sync over the same n, in batches that read the permutation
 * once. For conserved fields outside sphx_fields_t (x_m1, y_m1, z_m1, du_m1, mui, ...), which
 * Domain::sync reorders together with the others (domain/include/cstone/domain/domain.hpp:182-234). */
int sphx_domain_reorder(sphx_workspace_t* ws, void* const* fields, const int* elemBytes, size_t count,
                        size_t n);

/* ---- target groups ------------------------------------------------------------------------- */

/* Spatial target groups of [first, last), as the reference's GPU path builds its GroupView
 * (computeSpatialGroups, sph/include/sph/groups.hpp:42-55, with computeGroupSplits,
 * domain/include/cstone/traversal/groups_gpu.cu:34-101): fixed groups of 64 consecutive particles, each
 * split between consecutive particles farther apart (in unit-box coordinates) than tolFactor times
 * the side of the smallest leaf holding one of the group's particles (tolFactor = 2 in the reference).
 * Needs the tree of the last sphx_build_tree / sphx_domain_sync; fields: x, y, z. The groups live in
 * the workspace until the next call; *out receives their view (groupDt NULL). */
int sphx_compute_groups(sphx_workspace_t* ws, const sphx_box_t* box, const sphx_fields_t* f, uint32_t first,
                        uint32_t last, float tolFactor, sphx_groups_t* out);

/* Target groups for the calls that follow (the GroupView argument of the reference's drivers): the
 * neighbour search, every pair loop, sphx_compute_positions and sphx_update_smoothing_length, when
 * called with [first, last) = [groups->firstBody, groups->lastBody), run one wavefront per group
 * instead of consecutive groups of 64 from first. The view is copied (its device arrays must stay
 * valid while in use); NULL restores consecutive groups. Groups of the search must partition
 * [firstBody, lastBody) and hold at most 64 particles each (the search returns SPHX_ERR_ARG for a
 * larger one); the pair loops accept any subset of the searched targets (the block time-step
 * propagators' active groups), whose neighbour lists stay indexed by particle. Record chaining is
 * not applied while a view is set. */
int sphx_set_groups(sphx_workspace_t* ws, const sphx_groups_t* groups);

/* ---- neighbour search ---------------------------------------------------------------------- */

/* Neighbour lists of targets [first,last) with the coupled h / nc iteration (at most 10 updates,
 * h <- updateH(ng0, nc, h) while nc < ng0/4 or nc-1 > ngmax); writes f->nc[i] (self included)
 * and updates f->h in place; lists stay in the workspace for the pair loops. With iterateH = 0
 * the search runs once with the given h. Returns SPHX_ERR_NOT_CONVERGED (after completing all
 * lists) if any particle did not converge. Replaces sph::findNeighborsSfc / findNeighborsSph
 * (sph/include/sph/find_neighbors.hpp:11-56) and the h-loop of xmassGpu
 * (sph/include/sph/hydro_ve/xmass_gpu.cu:57-102). */
int sphx_find_neighbors(sphx_workspace_t* ws, const sphx_box_t* box, const sphx_fields_t* f,
                        const sphx_params_t* prm, uint32_t first, uint32_t last, int iterateH);

/* sphx_find_neighbors fused with the first pair loop, as the reference GPU path runs them
 * (the h/nc iteration inside computeXMass, sph/include/sph/hydro_ve/xmass_gpu.cu:57-129): the
 * final search of each target also forms m_i + sum_j m_j W(r_ij/h_i) over its stored neighbours
 * in list order and writes xm = m_i / (K h_i^-3 (...)) (out = SPHX_XMASS_XM, xmassJLoop,
 * hydro_ve/xmass_kern.hpp:51-79) or rho = m / xm (out = SPHX_XMASS_RHO, computeDensity,
 * hydro_std/density.hpp:42-60). Equivalent to sphx_find_neighbors followed by sphx_xmass or
 * sphx_density_std, without their second pass over the lists. Needs the kernel tables. */
#define SPHX_XMASS_XM 1
#define SPHX_XMASS_RHO 2
int sphx_find_neighbors_xmass(sphx_workspace_t* ws, const sphx_box_t* box, const sphx_fields_t* f,
                              const sphx_params_t* prm, uint32_t first, uint32_t last, int iterateH, int out);

/* Copy the stored neighbour list of targets [first,last), a sub-range of the last search's range,
 * into `out` (device or host memory, row-major [last-first][ngmax], row k for target first + k,
 * entries beyond min(nc-1, ngmax) unspecified). Test/inspection
 * helper matching the CPU Dataset's d.neighbors layout (particles_data.hpp:250-251). */
int sphx_copy_neighbors(sphx_workspace_t* ws, uint32_t* out, uint32_t first, uint32_t last);

/* Pair-test statistics of the last search: total stored pairs and maximum count. */
int sphx_neighbor_stats(const sphx_workspace_t* ws, uint64_t* sumPairs, uint32_t* maxCount);

/* Candidate distance tests the last search performed (each candidate swept against all 64 targets of
 * its group, h iterations included), for work accounting; no counterpart in the reference, whose
 * neighbour driver counts the same quantity as sumP2P (domain/test/performance/neighbor_driver.cu:146). */
int sphx_search_tests(const sphx_workspace_t* ws, uint64_t* tests);

/* ---- std SPH (sph/include/sph/hydro_std) ---------------------------------------------------- */

/* rho = m / xm with xm from xmassJLoop. Replaces computeDensity (hydro_std/density.hpp:56-60). */
int sphx_density_std(sphx_workspace_t* ws, const sphx_box_t* box, const sphx_fields_t* f,
                     const sphx_params_t* prm, uint32_t first, uint32_t last);

/* std EOS selected by prm->eos, as computeEOS_HydroStd selects by eosChoice (hydro_std/eos.hpp:146-152).
 * Ideal gas (computeIdealGasEOS_HydroStd, :55-82): p, c from u (or temp when u is NULL). Isothermal
 * (:82-97): p = rho c0^2 with c0 = soundSpeedConst, c = c0 as the reference's GPU kernel stores it
 * (hydro_std/eos_gpu.cu:74-82), temp = 0 when given. Polytropic (:99-114): p = K rho^n,
 * c = sqrt(n p / rho), temp = 0 when given. */
int sphx_eos_std(sphx_workspace_t* ws, const sphx_fields_t* f, const sphx_params_t* prm, uint32_t first,
                 uint32_t last);

/* IAD tensor inverse c11..c33. Replaces computeIAD (hydro_std/iad.hpp:75-79). */
int sphx_iad_std(sphx_workspace_t* ws, const sphx_box_t* box, const sphx_fields_t* f, const sphx_params_t* prm,
                 uint32_t first, uint32_t last);

/* ax, ay, az, du and the Courant time step (min over targets) in *minDtCourant.
 * Replaces computeMomentumEnergySTD (hydro_std/momentum_energy.hpp:97-101). */
int sphx_momentum_energy_std(sphx_workspace_t* ws, const sphx_box_t* box, const sphx_fields_t* f,
                             const sphx_params_t* prm, uint32_t first, uint32_t last, double* minDtCourant);

/* ---- VE SPH (sph/include/sph/hydro_ve) ------------------------------------------------------ */

/* xm. Replaces computeXMass (hydro_ve/xmass.hpp:67-72). */
int sphx_xmass(sphx_workspace_t* ws, const sphx_box_t* box, const sphx_fields_t* f, const sphx_params_t* prm,
               uint32_t first, uint32_t last);

/* kx, gradh. Replaces computeVeDefGradh (hydro_ve/ve_def_gradh.hpp:79-84). */
int sphx_ve_def_gradh(sphx_workspace_t* ws, const sphx_box_t* box, const sphx_fields_t* f,
                      const sphx_params_t* prm, uint32_t first, uint32_t last);

/* VE EOS selected by prm->eos, as computeEOS selects by eosChoice (hydro_ve/eos.hpp:191-197); each writes
 * prho = p / (kx m^2 gradh) and c, and rho = kx m / xm and p when those are non-NULL. Ideal gas
 * (computeIdealGasEOS_Impl, :53-95): from temp (u when temp is NULL). Isothermal (:97-123): p = rho c0^2,
 * c = c0, temp = 0 when given. Polytropic (:125-151): p = K rho^n, c = sqrt(n p / rho), temp = 0 when
 * given. Replaces computeIdealGasEOS, computeIsothermalEOS and computePolytropicEOS
 * (sph/include/sph/sph_gpu.hpp:48-62). */
int sphx_eos_ve(sphx_workspace_t* ws, const sphx_fields_t* f, const sphx_params_t* prm, uint32_t first,
                uint32_t last);

/* c11..c33, divv, curlv (and dV11..dV33 when non-NULL). Replaces computeIadDivvCurlv
 * (hydro_ve/iad_divv_curlv.hpp:92-97). */
int sphx_iad_divv_curlv(sphx_workspace_t* ws, const sphx_box_t* box, const sphx_fields_t* f,
                        const sphx_params_t* prm, uint32_t first, uint32_t last);

/* alpha (in place). Replaces computeAVswitches (hydro_ve/av_switches.hpp:80-85). */
int sphx_av_switches(sphx_workspace_t* ws, const sphx_box_t* box, const sphx_fields_t* f,
                     const sphx_params_t* prm, uint32_t first, uint32_t last);

/* ax, ay, az, du, *minDtCourant; avClean selects the AV-cleaning variant.
 * Replaces computeMomentumEnergy<avClean> (hydro_ve/momentum_energy.hpp:106-115). */
int sphx_momentum_energy_ve(sphx_workspace_t* ws, const sphx_box_t* box, const sphx_fields_t* f,
                            const sphx_params_t* prm, uint32_t first, uint32_t last, int avClean,
                            double* minDtCourant);

/* ---- time integration ------------------------------------------------------------------------
 * One rank's share of the step's integrate phase (main/src/propagator/std_hydro.hpp:170-181,
 * ve_hydro.hpp:206-217): time step, positions and energy, smoothing lengths. FP64 arithmetic on
 * FP64 positions and energies in both forms, as the reference's mixed types keep them. */

/* ts->minDtRho = Krho / |max_i divv_i| over [first, last). Replaces rhoTimestep
 * (sph/include/sph/ts_global.hpp:70-93). fields: divv. */
int sphx_rho_timestep(sphx_workspace_t* ws, const sphx_fields_t* f, const sphx_params_t* prm, uint32_t first,
                      uint32_t last, sphx_timestep_t* ts);

/* New global step min(dtAcc, minDtCourant, minDtRho, maxDtIncrease * minDt), dtAcc =
 * etaAcc sqrt(eps / |a|max) when g != 0 (else infinite); then ttot += step, minDt_m1 = minDt,
 * minDt = step. Replaces computeTimestep (ts_global.hpp:95-113) on one rank (the MPI_Allreduce
 * of the minimum is the identity). fields: ax, ay, az when g != 0. */
int sphx_compute_timestep(sphx_workspace_t* ws, const sphx_fields_t* f, const sphx_params_t* prm, uint32_t first,
                          uint32_t last, sphx_timestep_t* ts);

/* Press second-order position update and Adams-Bashforth energy update over [first, last):
 * x, y, z (put back into a periodic box), x_m1..z_m1, vx, vy, vz, and temp (when fields.temp is
 * set) or u, with du_m1 = du. dt and dt_m1 are the step and the previous step (the reference
 * passes them as float). The temperature update uses cv = idealGasCv(mui_i, gamma) when hist->mui
 * is set, else constCv, or idealGasCv(prm->muiConst, gamma) formed as the reference's host does
 * when constCv < 0. Replaces computePositions / computePositionsGpu and its kernel
 * (sph/include/sph/positions.hpp:177-200, sph_gpu.hpp:78-82, positions_gpu.cu:116-168).
 * fields: x, y, z, vx, vy, vz, ax, ay, az, du, temp or u, h (fixed boundaries only). */
int sphx_compute_positions(sphx_workspace_t* ws, const sphx_box_t* box, const sphx_fields_t* f,
                           const sphx_history_t* hist, const sphx_params_t* prm, uint32_t first, uint32_t last,
                           double dt, double dt_m1, double constCv);

/* ---- block time steps (the ve-bdt propagators, main/src/propagator/ve_hydro_bdt.hpp) ---------
 * The entry points below run over the targets of [first, last): the groups of the view set with
 * sphx_set_groups for that range (the propagator's active rungs), else consecutive groups of 64.
 * Rungs index the previous steps dtM1[0 .. numRungs) (Timestep::dt_m1, sph/include/sph/timestep.h:26);
 * rung values must lie below numRungs. */
#define SPHX_MAX_RUNGS 4

/* sphx_compute_positions with a rung per particle: particle i takes the previous step
 * dtM1[rung[i]] (dtM1[0] when rung is NULL). Replaces computePositionsGpu with its rung array
 * (positions_gpu.cu:116-168, :140). */
int sphx_compute_positions_rungs(sphx_workspace_t* ws, const sphx_box_t* box, const sphx_fields_t* f,
                                 const sphx_history_t* hist, const sphx_params_t* prm, uint32_t first,
                                 uint32_t last, double dt, const float* dtM1, int numRungs, const uint8_t* rung,
                                 double constCv);

/* Drift to a time inside a block time-step hierarchy: each particle is taken back by dtBack to the
 * state of its last force evaluation (positionUpdate with -dtBack) and advanced by dt from there,
 * without the periodic wrap; vx, vy, vz follow and temp (cv as in sphx_compute_positions) or u is
 * recovered and advanced with energyUpdate. x_m1, y_m1, z_m1 and du_m1 are read, not written.
 * Replaces driftPositionsGpu (positions_gpu.cu:17-85) and driftPositions (positions.hpp:151-174).
 * fields: x, y, z, vx, vy, vz, ax, ay, az, du, temp or u. */
int sphx_drift_positions(sphx_workspace_t* ws, const sphx_fields_t* f, const sphx_history_t* hist,
                         const sphx_params_t* prm, uint32_t first, uint32_t last, double dt, double dtBack,
                         const float* dtM1, int numRungs, const uint8_t* rung, double constCv);

/* groupDt[g] = min(groupDt[g], Krho / |max divv over group g|) (the signed maximum, in float).
 * Replaces groupDivvTimestepGpu (ts_groups.cu:16-46; groupDivvTimestep passes d.Krho,
 * ts_rungs.hpp:49-56). fields: divv. */
int sphx_group_divv_timestep(sphx_workspace_t* ws, const sphx_fields_t* f, const sphx_params_t* prm,
                             uint32_t first, uint32_t last, float Krho, float* groupDt);

/* groupDt[g] = min(groupDt[g], etaAcc / sqrt(sqrt(max |a|^2 over group g))) in float, |a|^2 =
 * ax^2 + (ay^2 + az^2) in the HydroType. Replaces groupAccTimestepGpu (ts_groups.cu:51-81;
 * groupAccTimestep passes etaAcc * sqrt(eps), ts_rungs.hpp:58-67). fields: ax, ay, az. */
int sphx_group_acc_timestep(sphx_workspace_t* ws, const sphx_fields_t* f, const sphx_params_t* prm,
                            uint32_t first, uint32_t last, float etaAcc, float* groupDt);

/* particleRungs[i] = rung for every target (groups of at most 64 particles, as for the other
 * elementwise entry points). Replaces storeRungGpu (ts_groups.cu:84-104). */
int sphx_store_rung(sphx_workspace_t* ws, uint32_t first, uint32_t last, uint8_t rung, uint8_t* particleRungs);

/* h = updateH(ng0, nc, h) over [first, last). Replaces updateSmoothingLength
 * (sph/include/sph/update_h.hpp:26-35). fields: h, nc. */
int sphx_update_smoothing_length(sphx_workspace_t* ws, const sphx_fields_t* f, const sphx_params_t* prm,
                                 uint32_t first, uint32_t last);

/* ---- observables -------------------------------------------------------------------------------
 * Totals over [first, last) of one rank: eKin = 1/2 sum m |v|^2, eInt = sum m u, or
 * cv sum m temp when fields.temp is set (cv as given, or idealGasCv(muiConst, gamma) formed as the
 * reference's host does when cv < 0), linmom = sum m v,
 * angmom = sum m (x cross v), ncsum = sum nc (0 when fields.nc is NULL), all in FP64. Replaces
 * computeConservedQuantities / conservedQuantitiesGpu (main/src/observables/
 * conserved_quantities.hpp:50-177, conserved_gpu.cu:50-95) before its MPI reduction. The sums run
 * in a fixed order (per-block partials, then one block), so repeated calls agree bit for bit. */
typedef struct sphx_conserved
{
    double eKin, eInt;
    double linmom[3], angmom[3];
    double ncsum;
} sphx_conserved_t;

int sphx_conserved_quantities(sphx_workspace_t* ws, const sphx_fields_t* f, const sphx_params_t* prm,
                              uint32_t first, uint32_t last, double cv, sphx_conserved_t* out);

/* ---- adjoints (no counterpart in the reference) ---------------------------------------------
 * Vector-Jacobian products of the std density, IAD and momentum-energy passes with respect to
 * positions, smoothing lengths and a per-particle AV coefficient, for gradient-based calibration
 * and inverse problems. Each recomputes its pair terms from the neighbour lists of the last
 * sphx_find_neighbors (the lists are a checkpoint: they are held fixed, as is the h of the
 * forward). Kernel derivatives use the tabulated dW/dv interpolated like W in the FP64 form and
 * are evaluated like W in the FP32 form, as the forward loops do. A pair's derivatives for its
 * neighbour go to a per-pair buffer of the workspace and are summed per particle in a fixed order
 * (no atomics), so repeated calls give the same bits. The first adjoint call after a search builds
 * that pair index (one sort of the search's pairs); the adjoints run over consecutive groups (no
 * group view set). */

/* rho = K h^-3 (m_i + sum_j m_j W(r_ij/h_i)) (sphx_density_std).
 * in: rho. out: x, y, z, h. fields: x, y, z, h, m, nc. */
int sphx_density_std_vjp(sphx_workspace_t* ws, const sphx_box_t* box, const sphx_fields_t* f,
                         const sphx_params_t* prm, uint32_t first, uint32_t last, const sphx_adjoint_t* in,
                         sphx_adjoint_t* out);

/* c_ij = h^3/K tau^-1, tau = sum_j m_j/rho_j W(r_ij/h_i) r_ij r_ij^T (sphx_iad_std).
 * in: c11..c33. out: x, y, z, h, rho. fields: x, y, z, h, m, rho, c11..c33 (forward output), nc. */
int sphx_iad_std_vjp(sphx_workspace_t* ws, const sphx_box_t* box, const sphx_fields_t* f,
                     const sphx_params_t* prm, uint32_t first, uint32_t last, const sphx_adjoint_t* in,
                     sphx_adjoint_t* out);

/* ax, ay, az, du of sphx_momentum_energy_std, with the AV coefficient taken from fields.alpha when
 * prm->flags has SPHX_FLAG_STD_ALPHA, else 1. in: ax, ay, az, du (du may be NULL).
 * out: x, y, z, h, rho, p, alpha, c11..c33. fields: as the forward. */
int sphx_momentum_energy_std_vjp(sphx_workspace_t* ws, const sphx_box_t* box, const sphx_fields_t* f,
                                 const sphx_params_t* prm, uint32_t first, uint32_t last,
                                 const sphx_adjoint_t* in, sphx_adjoint_t* out);

#ifdef __cplusplus
}
#endif

#endif /* SPHEXA_AMD_H */
