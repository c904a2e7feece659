"""Full hydro pass on the CPU oracle (or the reference wrapper) -- TEST INFRASTRUCTURE ONLY.

keys -> stable sort -> permute -> cornerstone tree (bucket 64) -> findNeighborsSph -> the std or
VE pair loops, in the order of the reference propagators (main/src/propagator/std_hydro.hpp:124-168,
main/src/propagator/ve_hydro.hpp:130-204). Returns every array in SFC order plus the tree and the
neighbour lists, with per-phase wall times.
"""
import time

import numpy as np

from sphexa_amd import abi

from . import oracle as O

STD_OUT = ("rho", "p", "c", "c11", "c12", "c13", "c22", "c23", "c33", "ax", "ay", "az", "du")
VE_OUT = ("xm", "kx", "gradh", "prho", "c", "c11", "c12", "c13", "c22", "c23", "c33", "divv", "curlv", "alpha", "ax",
          "ay", "az", "du")
GRADV = ("dV11", "dV12", "dV13", "dV22", "dV23", "dV33")


def run(lib, arrays, box, params, form="std", av_clean=False, bucket=64, iterate_h=True, hydro=np.float64,
        prev_tree=None):
    """arrays: host IC (unsorted). form: 'std', 've' or 'search' (stop after the neighbour search).
    prev_tree: the tree of the previous sync of the same particles (a later Domain::sync): it is
    updated once (O.Tree.update) instead of converged anew."""
    times = {}
    t0 = time.perf_counter()
    x, y, z = (np.ascontiguousarray(arrays[k], np.float64) for k in "xyz")
    n = x.size
    keys = O.sfc_keys(lib, box, x, y, z)
    skeys, order = O.sort_keys(lib, keys)
    a = {}
    for k, v in arrays.items():
        if v is None:
            continue
        dt = np.float64 if k in abi.REAL_FIELDS else hydro
        a[k] = np.ascontiguousarray(np.asarray(v)[order], dtype=dt)
    if prev_tree is None:
        tree = O.Tree(lib, skeys, box, bucket)
    else:
        tree = prev_tree
        tree.update(skeys, box, bucket)
    times["domain::sync"] = time.perf_counter() - t0

    t0 = time.perf_counter()
    a["nc"] = np.zeros(n, np.uint32)
    a["h"] = np.ascontiguousarray(a["h"], hydro)
    nb, fails = O.find_neighbors(lib, tree, box, a, params, 0, n, iterate_h)
    times["FindNeighbors"] = time.perf_counter() - t0
    res = dict(keys=skeys, order=order, tree=tree, nbrs=nb, fails=fails, a=a, times=times)
    if form == "search":
        return res

    wh, whd = O.tables(lib, 6.0, hydro)
    for k in abi.HYDRO_FIELDS:
        if k not in a:
            a[k] = np.zeros(n, hydro)
    a.setdefault("du", np.zeros(n))
    a["tdpdTrho"] = None
    if form == "ve":
        a["rho"] = None
        a["p"] = None
        if not av_clean:
            for k in GRADV:
                a[k] = None
        a.setdefault("temp", None)
    else:
        a["temp"] = None
    cp = O.CpuPass(lib, box, a, params, nb, 0, n, wh, whd)

    def timed(name, fn, *args):
        t = time.perf_counter()
        r = fn(*args)
        times[name] = time.perf_counter() - t
        return r

    if form == "std":
        timed("Density", cp.density_std)
        timed("EquationOfState", cp.eos_std)
        timed("IAD", cp.iad_std)
        res["minDt"] = timed("MomentumEnergyIAD", cp.momentum_energy_std)
    else:
        timed("XMass", cp.xmass)
        timed("Normalization & Gradh", cp.ve_def_gradh)
        timed("EquationOfState", cp.eos_ve)
        timed("IadVelocityDivCurl", cp.iad_divv_curlv)
        timed("AVswitches", cp.av_switches)
        res["minDt"] = timed("MomentumAndEnergy", cp.momentum_energy_ve, av_clean)
    return res
