"""Write the golden vectors of tests/golden from the reference's own headers (oracle/_ref).

    make -C oracle ref && python tests/golden/make_golden.py

Runs the full std and VE passes of the compiled reference wrapper (keys, stable sort,
cornerstone tree, findNeighborsSph, JLoops) on small seeded inputs and stores inputs and outputs
as compressed .npz files. tests/test_oracle.py compares the oracle restatement against them
bit for bit, which pins the oracle wherever the reference checkout is absent.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from oracle import pipeline  # noqa: E402
from sphexa_amd import abi, ic  # noqa: E402
from sphexa_amd.particles import ideal_gas_cv  # noqa: E402

CASES = {
    # name: (generator, form, av_clean, hydro dtype)
    "sedov10_std_f64": ("sedov", "std", False, np.float64),
    "sedov10_std_f32": ("sedov", "std", False, np.float32),
    "turb10_ve_f64": ("turb", "ve", False, np.float64),
    "turb10_veclean_f64": ("turb", "ve", True, np.float64),
    "turb10_ve_f32": ("turb", "ve", False, np.float32),
    "evrard1k_search_f64": ("evrard", "search", False, np.float64),
}


def inputs(kind):
    if kind == "sedov":
        arr, spec = ic.sedov_lattice(10, seed=2024)
        arr["vx"] = 0.1 * np.sin(6.0 * arr["x"])
        arr["vy"] = 0.05 * np.cos(5.0 * arr["y"])
    elif kind == "turb":
        arr, spec = ic.turbulence_box(10, seed=2025)
        arr["temp"] = arr.pop("u") / ideal_gas_cv(10.0, 5.0 / 3.0)
        arr["alpha"] = np.random.default_rng(7).uniform(0.05, 1.0, arr["x"].size)
    else:
        arr, spec = ic.evrard_sphere(1000, seed=2026)
    box = abi.Box.make(spec["lo"], spec["hi"], abi.PERIODIC if spec["periodic"] else abi.OPEN)
    return arr, box


def run_case(lib, name):
    kind, form, clean, hydro = CASES[name]
    arr, box = inputs(kind)
    ht = abi.SPHX_F32 if hydro == np.float32 else abi.SPHX_F64
    params = abi.Params.defaults(lib.kernel_K(6.0), hydro_type=ht, minDt=1e-3)
    return arr, box, params, pipeline.run(lib, arr, box, params, form=form, av_clean=clean, hydro=hydro)


def outputs(res, form, clean):
    a = res["a"]
    names = {"std": pipeline.STD_OUT, "ve": pipeline.VE_OUT + (pipeline.GRADV if clean else ()),
             "search": ()}[form]
    out = {k: a[k] for k in names}
    nc = a["nc"]
    out["nc"] = nc
    out["h_final"] = a["h"]
    out["keys"] = res["keys"]
    out["order"] = res["order"]
    ngmax = res["nbrs"].shape[1]
    counts = np.minimum(nc.astype(np.int64) - 1, ngmax)
    out["nbrs_flat"] = np.concatenate([res["nbrs"][i, :counts[i]] for i in range(nc.size)]).astype(np.uint32)
    t = res["tree"]
    for f in ("leaves", "counts", "layout", "prefixes", "childOffsets", "levelRange", "centers", "sizes"):
        out["tree_" + f] = getattr(t, f)
    if "minDt" in res:
        out["minDt"] = np.array(res["minDt"])
    return out


def main():
    ref = O.load("ref")
    for name in CASES:
        arr, box, params, res = run_case(ref, name)
        kind, form, clean, hydro = CASES[name]
        data = {"in_" + k: np.asarray(v) for k, v in arr.items()}
        data.update(outputs(res, form, clean))
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **data)
        print(f"wrote {path} ({os.path.getsize(path) // 1024} KiB)")


if __name__ == "__main__":
    main()
