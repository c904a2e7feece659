"""Measure the reference's own float path on the Evrard-like sphere (VE pass) and store its error.

    make -C oracle ref && python tests/golden/make_evrard_pin.py

Runs the compiled reference wrapper (oracle/_ref, the reference's headers instantiated with
double and with float HydroType, i.e. its plain and SphTypes layouts, sph/include/sph/types.hpp:39-46)
through one VE pass on the 4,000,000-particle sphere of the GPU workload test (seed 11, u = 0.05 as
temperature, v = 0, open box [-1,1]^3, minDt = 1e-4), from the same initial h, with the h/nc
iteration. It writes, per field, the max and RMS error of the float pass against the double pass:
pointwise relative for the positive fields (rho = kx m / xm, xm, kx, prho), relative to the field's
largest magnitude otherwise. tests/test_gpu_workloads.py holds the HIP mixed form to these figures
on the same sphere: the grad-h term of an unrelaxed random sampling passes near zero, so prho and
the accelerations are ill-conditioned in any precision, and the reference's float path is the bar.
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from oracle import pipeline  # noqa: E402
from sphexa_amd import abi, ic  # noqa: E402
from sphexa_amd.particles import ideal_gas_cv  # noqa: E402

N, SEED = 4_000_000, 11
FIELDS = ("rho", "xm", "kx", "gradh", "prho", "c", "alpha", "ax", "ay", "az")
POSITIVE = ("rho", "xm", "kx", "prho", "c")
OUT = os.path.join(HERE, "evrard4m_ve_ref_float_error.json")


def field_errors(got, ref):
    out = {}
    for k in FIELDS:
        g = np.asarray(got[k], np.float64)
        r = np.asarray(ref[k], np.float64)
        d = np.abs(g - r)
        rel = d / np.abs(r) if k in POSITIVE else d / max(float(np.max(np.abs(r))), 1e-300)
        out[k] = {"max": float(np.max(rel)), "rms": float(np.sqrt(np.mean(rel * rel)))}
    return out


def main():
    lib = O.load("ref")
    arr, spec = ic.evrard_sphere(N, seed=SEED)
    arr["temp"] = arr.pop("u") / ideal_gas_cv(10.0, 5.0 / 3.0)
    box = abi.Box.make(spec["lo"], spec["hi"], abi.OPEN)
    K = lib.kernel_K(6.0)
    res = {}
    for hyd in (np.float64, np.float32):
        p = abi.Params.defaults(K, hydro_type=abi.SPHX_F32 if hyd == np.float32 else abi.SPHX_F64, minDt=1e-4)
        t = time.perf_counter()
        o = pipeline.run(lib, arr, box, p, form="ve", hydro=hyd)
        assert o["fails"] == 0
        a = o["a"]
        a["rho"] = np.asarray(a["kx"], np.float64) * np.asarray(a["m"], np.float64) / np.asarray(a["xm"], np.float64)
        res[hyd] = a
        print(f"reference VE pass, {hyd.__name__}: {time.perf_counter() - t:.1f} s", flush=True)
    err = field_errors(res[np.float32], res[np.float64])
    nc_diff = int(np.count_nonzero(res[np.float32]["nc"] != res[np.float64]["nc"]))
    doc = {
        "what": "reference float path (SphTypes) vs reference double path, one VE pass, Evrard-like sphere",
        "made_by": "tests/golden/make_evrard_pin.py, from the reference headers compiled by oracle/Makefile (ref)",
        "particles": N,
        "seed": SEED,
        "minDt": 1e-4,
        "targets_with_different_nc": nc_diff,
        "errors": err,
    }
    with open(OUT, "w") as fh:
        json.dump(doc, fh, indent=1)
    for k, v in err.items():
        print(f"{k:6s} max {v['max']:.3e} rms {v['rms']:.3e}")
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
