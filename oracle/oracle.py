"""ctypes loader of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Loads oracle/liboracle.so (the restatement, prefix ``orc_``) or oracle/_ref/libsphx_ref.so
(the reference's own headers compiled by oracle/Makefile, prefix ``ref_``). Used by tests/,
__graft_entry__.smoke() and the CPU-baseline leg of bench.py only, as the checker; the
product library never imports this module.
"""
import ctypes as C
import os

import numpy as np

from sphexa_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIBS = {"orc": os.path.join(HERE, "liboracle.so"), "ref": os.path.join(HERE, "_ref", "libsphx_ref.so"),
        # the reference built with its Release flags (timed CPU baseline only), per x86-64 level
        "ref_rel_v3": os.path.join(HERE, "_ref", "libsphx_ref_rel_v3.so"),
        "ref_rel_v4": os.path.join(HERE, "_ref", "libsphx_ref_rel_v4.so")}
PREFIX = {"orc": "orc", "ref": "ref", "ref_rel_v3": "ref", "ref_rel_v4": "ref"}

_P = C.c_void_p
_SZ = C.c_size_t
_U32 = C.c_uint32
_BOXP = C.POINTER(abi.Box)
_FP = C.POINTER(abi.Fields)
_PP = C.POINTER(abi.Params)
_AP = C.POINTER(abi.Adjoint)
_HP = C.POINTER(abi.History)
_TSP = C.POINTER(abi.Timestep)

_SIGS = {
    "sfc_keys": (None, [_BOXP, _P, _P, _P, _P, _SZ]),
    "hilbert_encode": (None, [_P, _P, _P, _P, _SZ]),
    "hilbert_decode": (None, [_P, _P, _P, _P, _SZ]),
    "sort_keys": (None, [_P, _P, _SZ]),
    "tree_build": (_P, [_P, _SZ, _U32, _BOXP]),
    "tree_sizes": (None, [_P, _P, _P, _P]),
    "tree_copy": (None, [_P] * 11),
    "tree_free": (None, [_P]),
    "tree_update": (C.c_int, [_P, _P, _SZ, _U32, _BOXP]),
    "focus_new": (_P, [_U32, C.c_float]),
    "focus_update": (C.c_int, [_P, _BOXP, _P, _SZ]),
    "focus_num_leaves": (_SZ, [_P]),
    "focus_copy": (None, [_P, _P, _P]),
    "focus_free": (None, [_P]),
    "group_splits": (_SZ, [_P, _BOXP, _P, _P, _P, _U32, _U32, C.c_float, _P, _SZ]),
    "find_neighbors": (_SZ, [_P, _BOXP, _FP, _PP, _U32, _U32, C.c_int, _P]),
    "kernel_K": (C.c_double, [C.c_double]),
    "distance_sq_pbc": (C.c_double, [_BOXP] + [C.c_double] * 6),
    "simpson_sin": (C.c_double, []),
    "sphynx_K": (C.c_double, [C.c_double]),
    "tables_f64": (None, [C.c_double, _P, _P]),
    "tables_f32": (None, [C.c_double, _P, _P]),
    "lookup_f64": (C.c_double, [_P, C.c_double]),
    "density_std": (None, [_BOXP, _FP, _PP, _U32, _U32, _P, _P]),
    "eos_std": (None, [_FP, _PP, _U32, _U32]),
    "iad_std": (None, [_BOXP, _FP, _PP, _U32, _U32, _P, _P]),
    "momentum_energy_std": (C.c_double, [_BOXP, _FP, _PP, _U32, _U32, _P, _P, _P]),
    "xmass": (None, [_BOXP, _FP, _PP, _U32, _U32, _P, _P]),
    "ve_def_gradh": (None, [_BOXP, _FP, _PP, _U32, _U32, _P, _P, _P]),
    "eos_ve": (None, [_FP, _PP, _U32, _U32]),
    "iad_divv_curlv": (None, [_BOXP, _FP, _PP, _U32, _U32, _P, _P]),
    "av_switches": (None, [_BOXP, _FP, _PP, _U32, _U32, _P, _P]),
    "divv_curlv": (None, [_BOXP, _FP, _PP, _U32, _U32, _P, _P]),
    "momentum_energy_ve": (C.c_double, [_BOXP, _FP, _PP, _U32, _U32, _P, _P, C.c_int, _P]),
    "energy_update": (C.c_double, [C.c_double] * 5),
    "position_update": (None, [C.c_double, C.c_double, _P, _P, _P, _BOXP, _P, _P, _P]),
    "fbc_check_f32": (C.c_int, [C.c_double, C.c_float, C.c_double, C.c_double, C.c_int]),
    "compute_positions": (None, [_BOXP, _FP, _HP, _PP, _U32, _U32, C.c_double, C.c_double]),
    "compute_positions_rungs": (None, [_BOXP, _FP, _HP, _PP, _U32, _U32, C.c_double, _P, _P]),
    "drift_positions": (None, [_FP, _HP, _PP, _U32, _U32, C.c_double, C.c_double, _P, _P]),
    "group_divv_timestep": (None, [_FP, _PP, _P, _P, _U32, C.c_float, _P]),
    "group_acc_timestep": (None, [_FP, _PP, _P, _P, _U32, C.c_float, _P]),
    "store_rung": (None, [_P, _P, _U32, C.c_uint8, _P]),
    "update_smoothing_length": (None, [_FP, _PP, _U32, _U32]),
    "compute_timestep": (None, [_FP, _PP, _U32, _U32, _TSP]),
    "rho_timestep": (None, [_FP, _PP, _U32, _U32, _TSP]),
    "conserved_quantities": (None, [_FP, _PP, _U32, _U32, _P]),
    "global_box": (None, [_BOXP, _P, _P, _P, _SZ, C.c_int]),
    "limit_box_shrinking": (None, [_P, _P, _P]),
    "density_std_vjp": (None, [_BOXP, _FP, _PP, _U32, _U32, _P, _P, _P, _AP, _AP]),
    "iad_std_vjp": (None, [_BOXP, _FP, _PP, _U32, _U32, _P, _P, _P, _AP, _AP]),
    "momentum_energy_std_vjp": (None, [_BOXP, _FP, _PP, _U32, _U32, _P, _P, _P, _AP, _AP]),
}

_cache = {}


def available(kind="orc"):
    return os.path.exists(LIBS[kind])


def load(kind="orc"):
    """Return a namespace of the oracle's C functions (kind 'orc', 'ref', 'ref_rel_v3' or 'ref_rel_v4')."""
    if kind in _cache:
        return _cache[kind]
    path = LIBS[kind]
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is not built (run: make -C oracle{' ref' if kind == 'ref' else ''})")
    lib = C.CDLL(path)
    ns = {}
    for name, (res, args) in _SIGS.items():
        if not hasattr(lib, f"{PREFIX[kind]}_{name}"):
            continue
        fn = getattr(lib, f"{PREFIX[kind]}_{name}")
        fn.restype = res
        fn.argtypes = args
        ns[name] = fn
    ns["_lib"] = lib
    _cache[kind] = type("OracleLib", (), ns)
    return _cache[kind]


def ptr(a):
    return None if a is None else a.ctypes.data


class Tree:
    """Cornerstone leaves + linked octree + layout + node centres (numpy copies)."""

    def __init__(self, lib, sorted_keys, box, bucket=64):
        keys = np.ascontiguousarray(sorted_keys, dtype=np.uint64)
        self.lib = lib
        self.handle = lib.tree_build(ptr(keys), keys.size, bucket, C.byref(box))
        self._copy()

    def update(self, sorted_keys, box, bucket=64):
        """The tree update of a later Domain::sync (one rebalance on the previous counts, then counts of
        the new keys), in place; returns True when no node changed."""
        keys = np.ascontiguousarray(sorted_keys, dtype=np.uint64)
        conv = self.lib.tree_update(self.handle, ptr(keys), keys.size, bucket, C.byref(box))
        self._copy()
        return bool(conv)

    def _copy(self):
        lib = self.lib
        nl, ni, nn = C.c_int32(), C.c_int32(), C.c_int32()
        lib.tree_sizes(self.handle, C.byref(nl), C.byref(ni), C.byref(nn))
        self.numLeafNodes, self.numInternalNodes, self.numNodes = nl.value, ni.value, nn.value
        L, N = self.numLeafNodes, self.numNodes
        self.leaves = np.zeros(L + 1, np.uint64)
        self.counts = np.zeros(L, np.uint32)
        self.layout = np.zeros(L + 1, np.uint32)
        self.prefixes = np.zeros(N, np.uint64)
        self.childOffsets = np.zeros(N, np.int32)
        self.parents = np.zeros(max(1, (N - 1) // 8), np.int32)
        self.internalToLeaf = np.zeros(N, np.int32)
        self.levelRange = np.zeros(23, np.int32)
        self.centers = np.zeros((N, 3), np.float64)
        self.sizes = np.zeros((N, 3), np.float64)
        lib.tree_copy(self.handle, ptr(self.leaves), ptr(self.counts), ptr(self.layout), ptr(self.prefixes),
                      ptr(self.childOffsets), ptr(self.parents), ptr(self.internalToLeaf), ptr(self.levelRange),
                      ptr(self.centers), ptr(self.sizes))

    def __del__(self):
        try:
            self.lib.tree_free(self.handle)
        except Exception:
            pass


class FocusTree:
    """The reference's single-rank focused octree (oracle/_ref only): the leaves and counts of a
    sequence of one-rank Domain::sync calls. sync() converges on the first call and updates once on
    every later one."""

    def __init__(self, lib, bucket=64, theta=0.5):
        self.lib = lib
        self.handle = lib.focus_new(bucket, theta)
        self.first = True

    def sync(self, sorted_keys, box):
        keys = np.ascontiguousarray(sorted_keys, dtype=np.uint64)
        if self.first:
            while not self.lib.focus_update(self.handle, C.byref(box), ptr(keys), keys.size):
                pass
            self.first = False
        else:
            self.lib.focus_update(self.handle, C.byref(box), ptr(keys), keys.size)
        L = self.lib.focus_num_leaves(self.handle)
        leaves, counts = np.zeros(L + 1, np.uint64), np.zeros(L, np.uint32)
        self.lib.focus_copy(self.handle, ptr(leaves), ptr(counts))
        return leaves, counts

    def __del__(self):
        try:
            self.lib.focus_free(self.handle)
        except Exception:
            pass


def group_splits(lib, tree, box, x, y, z, first, last, tol_factor=2.0):
    """Spatial target groups of [first, last) (computeGroupSplits, groupSize 64): the group starts
    followed by `last`, as a uint32 array of numGroups + 1 entries."""
    cap = (last - first) + 2
    out = np.zeros(cap, np.uint32)
    k = lib.group_splits(tree.handle, C.byref(box), ptr(x), ptr(y), ptr(z), first, last, tol_factor, ptr(out), cap)
    assert k > 0
    return out[:k]


def tables(lib, sinc_index=6.0, dtype=np.float64):
    wh = np.zeros(20000, dtype)
    whd = np.zeros(20000, dtype)
    fn = lib.tables_f64 if dtype == np.float64 else lib.tables_f32
    fn(sinc_index, ptr(wh), ptr(whd))
    return wh, whd


def sfc_keys(lib, box, x, y, z):
    keys = np.zeros(x.size, np.uint64)
    lib.sfc_keys(C.byref(box), ptr(x), ptr(y), ptr(z), ptr(keys), x.size)
    return keys


def sort_keys(lib, keys):
    k = np.ascontiguousarray(keys, dtype=np.uint64).copy()
    order = np.zeros(k.size, np.uint32)
    lib.sort_keys(ptr(k), ptr(order), k.size)
    return k, order


def find_neighbors(lib, tree, box, arrays, params, first, last, iterate_h=True):
    """Run findNeighborsSph on [first,last); updates arrays['h'] and arrays['nc'] in place.
    Returns (neighbors[last-first, ngmax], number of non-converged particles)."""
    nb = np.zeros((last - first) * params.ngmax, np.uint32)
    f = abi.fields_from(arrays)
    fails = lib.find_neighbors(tree.handle, C.byref(box), C.byref(f), C.byref(params), first, last,
                               int(iterate_h), ptr(nb))
    return nb.reshape(last - first, params.ngmax), fails


def sorted_lists(nb, nc, ngmax):
    """Per-target sorted neighbour sets (only the stored min(nc-1, ngmax) entries)."""
    out = []
    for k in range(nb.shape[0]):
        n = min(int(nc[k]) - 1, ngmax)
        out.append(np.sort(nb[k, :n]))
    return out


def adjoint_from(arrays):
    """abi.Adjoint over host arrays (missing names stay NULL)."""
    a = abi.Adjoint()
    for name in abi.ADJOINT_FIELDS:
        if name in arrays and arrays[name] is not None:
            setattr(a, name, arrays[name].ctypes.data)
    return a


class CpuPass:
    """One std or VE hydro pass on host arrays through the oracle (or the reference wrapper)."""

    def __init__(self, lib, box, arrays, params, nbrs, first, last, wh, whd):
        self.lib, self.box, self.a, self.p = lib, box, arrays, params
        self.nb, self.first, self.last, self.wh, self.whd = nbrs, first, last, wh, whd

    def _f(self):
        return abi.fields_from(self.a)

    def density_std(self):
        self.lib.density_std(C.byref(self.box), C.byref(self._f()), C.byref(self.p), self.first, self.last,
                             ptr(self.nb), ptr(self.wh))

    def eos_std(self):
        self.lib.eos_std(C.byref(self._f()), C.byref(self.p), self.first, self.last)

    def iad_std(self):
        self.lib.iad_std(C.byref(self.box), C.byref(self._f()), C.byref(self.p), self.first, self.last,
                         ptr(self.nb), ptr(self.wh))

    def momentum_energy_std(self, maxvs=None):
        return self.lib.momentum_energy_std(C.byref(self.box), C.byref(self._f()), C.byref(self.p), self.first,
                                            self.last, ptr(self.nb), ptr(self.wh), ptr(maxvs))

    def xmass(self):
        self.lib.xmass(C.byref(self.box), C.byref(self._f()), C.byref(self.p), self.first, self.last, ptr(self.nb),
                       ptr(self.wh))

    def ve_def_gradh(self):
        self.lib.ve_def_gradh(C.byref(self.box), C.byref(self._f()), C.byref(self.p), self.first, self.last,
                              ptr(self.nb), ptr(self.wh), ptr(self.whd))

    def eos_ve(self):
        self.lib.eos_ve(C.byref(self._f()), C.byref(self.p), self.first, self.last)

    def iad_divv_curlv(self):
        self.lib.iad_divv_curlv(C.byref(self.box), C.byref(self._f()), C.byref(self.p), self.first, self.last,
                                ptr(self.nb), ptr(self.wh))

    def av_switches(self):
        self.lib.av_switches(C.byref(self.box), C.byref(self._f()), C.byref(self.p), self.first, self.last,
                             ptr(self.nb), ptr(self.wh))

    def momentum_energy_ve(self, av_clean=False, maxvs=None):
        return self.lib.momentum_energy_ve(C.byref(self.box), C.byref(self._f()), C.byref(self.p), self.first,
                                           self.last, ptr(self.nb), ptr(self.wh), int(av_clean), ptr(maxvs))

    def vjp(self, name, cot, grad):
        """Adjoint `name` (density_std_vjp, iad_std_vjp, momentum_energy_std_vjp): cotangent arrays
        `cot` and gradient accumulators `grad` (dicts of host arrays, added into)."""
        getattr(self.lib, name)(C.byref(self.box), C.byref(self._f()), C.byref(self.p), self.first, self.last,
                                ptr(self.nb), ptr(self.wh), ptr(self.whd), C.byref(adjoint_from(cot)),
                                C.byref(adjoint_from(grad)))

    def run_std(self):
        self.density_std()
        self.eos_std()
        self.iad_std()
        return self.momentum_energy_std()

    def run_ve(self, av_clean=False):
        self.xmass()
        self.ve_def_gradh()
        self.eos_ve()
        self.iad_divv_curlv()
        self.av_switches()
        return self.momentum_energy_ve(av_clean)
