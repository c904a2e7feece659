#!/usr/bin/env python3
"""Throughput of the SPH hydro hot path on one MI355X.

Workload (the headline configuration): subsonic turbulence box, VE formulation, 200^3 = 8,000,000
jittered-lattice particles in the periodic unit box, rho = 1, u = 1 (temperature-based EOS),
seeded solenoidal velocity field P(k) ~ k^-4 on 1 <= |k| <= 2 at rms Mach 0.3, ng0 = 100.

One step = one full pass with all fields resident in HBM:
  keys -> stable radix sort + permutation of every field -> cornerstone tree (bucket 64) ->
  neighbour search with the h/nc iteration -> XMass -> VeDefGradh -> EOS ->
  IAD + div/curl v -> AV switches -> momentum-energy + Courant time step.
It runs in the mixed form (the headline value) and in the plain FP64 form; the JSON line reports
particle-updates/s = N * steps / T for the mixed form, the FP64 value beside it, the roofline of
the dominant kernel from HIP events on the launch stream, and the reference's own OpenMP CPU path
(oracle/_ref) timed on a bounded sample of the same workload on this host.

    python bench.py [--gpus 1] [--steps 5] [--warmup 2] [--side 200]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "SPH particle-updates/s (nbr search+density+IAD+force), mixed vs FP64, 1 GPU"

# MI355X published peaks
PEAK_FP64_VECTOR_TFLOPS = 78.6
PEAK_FP32_VECTOR_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0

# Algorithmic flops per neighbour pair of the VE momentum-energy loop
# (sph/include/sph/hydro_ve/momentum_energy_kern.hpp:121-212, avClean = false), counting
# + - * as 1, division and square root as 1, each table lookup as 7 (index, two loads'
# difference, slope, offset, fma): separation 3, fold/r2/dist 6, v_ij 3, 1/h_j 1, v1 v2 2,
# h_j^-3 2, W_i W_j 2+14, IAD projections 2*(3*5+1) = 32, rho_j 2, rv 5, w_ij 1, viscosity 9,
# signal 4, Atwood 4 + branch 2, a_mom 2, AV terms 2+2+3*4 = 16, AV energy 5, energy 7,
# momentum 3*6 = 18, accumulations 4 -> 142.
FLOPS_PER_PAIR_MOM_VE = 142
# per-target part of the same kernel (loads aside): rho_i, 1/h_i, h_i^-3, eta_crit, epilogue
FLOPS_PER_TARGET_MOM_VE = 40

# HBM bytes per launch of the momentum-energy kernel on this workload (200^3), from rocprofv3
# FETCH_SIZE and WRITE_SIZE counter runs scaled by a plain copy kernel of known size (read x 2.0,
# write x 1.0); see profiles/ve_turbulence_8M_counters.md. Counters cannot be read from inside the
# timed run, so the measured figure is carried here for the default workload only.
MEASURED_TRAFFIC_MOM_VE = {(200, "mixed"): 4.992e9, (200, "fp64"): 6.928e9}


def parse():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[1])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--side", type=int, default=200, help="lattice side (N = side^3)")
    ap.add_argument("--cpu-side", type=int, default=100, help="lattice side of the CPU-baseline sample")
    ap.add_argument("--cpu-passes", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-std", action="store_true", help="skip the std-pass side measurement")
    ap.add_argument("--forms", default="mixed,fp64")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB device copy three times (counter calibration under rocprofv3)")
    return ap.parse_args()


def calibrate_copy():
    """A plain copy kernel of known size: 2^28 float32 elements read and written (1 GiB each way)."""
    import torch

    n = 1 << 28
    a = torch.rand(n, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        torch.mul(a, 1.0, out=b)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "launches": 3}), flush=True)


def make_ic(side, seed=42):
    from sphexa_amd import ic
    from sphexa_amd.particles import ideal_gas_cv

    arr, spec = ic.turbulence_box(side, seed=seed)
    arr["temp"] = arr.pop("u") / ideal_gas_cv(10.0, 5.0 / 3.0)
    return arr, spec


# previous-step fields of the integrator, permuted by the domain sync with the others
HISTORY = ("x_m1", "y_m1", "z_m1", "du_m1")

# HBM bytes per particle of the position / energy update (temperature form, no fixed boundary):
# read x, y, z (FP64), x_m1.., ax.. (T), temp, du (FP64), du_m1 (T); write x, y, z, x_m1..,
# vx.. , temp, du_m1
POSITIONS_BYTES = {"mixed": 24 + 12 + 12 + 8 + 8 + 4 + 24 + 12 + 12 + 8 + 4,
                   "fp64": 24 + 24 + 24 + 8 + 8 + 8 + 24 + 24 + 24 + 8 + 8}

# main kernels timed with HIP events on the workspace stream (sphx_kernel_ms), and the stages of
# the domain sync
KERNELS = ("search", "xmass", "ve_def_gradh", "iad_divv_curlv", "av_switches", "momentum_ve", "sfc_keys", "sort",
           "permute", "tree")
PAIR_KERNELS = ("xmass", "ve_def_gradh", "iad_divv_curlv", "av_switches", "momentum_ve")

# outputs of the VE pass compared between the mixed and FP64 forms (rho = kx m / xm added);
# positive quantities pointwise, signed ones relative to the field's largest magnitude
COMPARED = ("h", "xm", "kx", "gradh", "prho", "c", "c11", "c12", "c13", "c22", "c23", "c33", "divv", "curlv",
            "alpha", "ax", "ay", "az", "du")
POINTWISE = ("h", "xm", "kx", "gradh", "prho", "c", "rho")


def field_errors(mixed, ref):
    """Max and RMS relative error of each field of the mixed form against the FP64 form."""
    out = {}
    for k in sorted(ref):
        d = np.abs(mixed[k] - ref[k])
        if k in POINTWISE:
            r = d / np.abs(ref[k])
            norm = "pointwise"
        else:
            r = d / max(float(np.max(np.abs(ref[k]))), 1e-300)
            norm = "field max"
        out[k] = {"max": float(f"{np.max(r):.3e}"), "rms": float(f"{np.sqrt(np.mean(r * r)):.3e}"), "norm": norm}
    return out


PHASES = ("domain::sync", "FindNeighbors+XMass", "Normalization & Gradh", "EquationOfState",
          "IadVelocityDivCurl", "AVswitches", "MomentumAndEnergy")


def run_gpu(arr, spec, mixed, steps, warmup):
    import torch

    from sphexa_amd import abi
    from sphexa_amd.lib import Workspace
    from sphexa_amd.particles import VE_FIELDS, ParticleSet, integrate, search_first_sum

    ws = Workspace(0)
    box = abi.Box.make(spec["lo"], spec["hi"], abi.PERIODIC)
    params = abi.Params.defaults(ws.K, hydro_type=abi.SPHX_F32 if mixed else abi.SPHX_F64, minDt=1e-4)
    ps = ParticleSet.from_host(arr, VE_FIELDS + HISTORY, mixed=mixed)
    n = ps.n
    f = ps.fields()
    sf = ps.sync_fields()  # the sync permutes the conserved fields only, as Domain::sync does
    cparams = abi.chained(params)  # XMass, IAD and AV write the next pass's records (bitwise the same)
    stream = torch.cuda.current_stream()

    ws.set_profiling(True)
    kernels = KERNELS
    kernel_ms = {k: [] for k in kernels}

    def one_step(ev):
        ev[0].record(stream)
        ws.domain_sync(box, sf, params, n, 64)
        ps.reorder_history(ws)
        ev[1].record(stream)
        search_first_sum(ws, box, f, cparams, n, rho=False)
        ev[2].record(stream)
        ws.pair("sphx_ve_def_gradh", box, f, cparams, 0, n)
        ev[3].record(stream)
        ws.eos("sphx_eos_ve", f, params, 0, n)
        ev[4].record(stream)
        ws.pair("sphx_iad_divv_curlv", box, f, cparams, 0, n)
        ev[5].record(stream)
        ws.pair("sphx_av_switches", box, f, cparams, 0, n)
        ev[6].record(stream)
        dt = ws.momentum_energy_ve(box, f, cparams, 0, n, False)
        ev[7].record(stream)
        return dt

    dt_courant = 0.0

    def read_kernels():
        # the search and momentum-energy entry points end with a stream synchronisation, so every
        # event of the step has completed and these reads do not stall
        for k in kernels:
            kernel_ms[k].append(ws.kernel_ms(k))

    mk = lambda: [torch.cuda.Event(enable_timing=True) for _ in range(len(PHASES) + 1)]
    for _ in range(warmup):
        one_step(mk())
    torch.cuda.synchronize()
    events = []
    t0 = time.perf_counter()
    for _ in range(steps):
        ev = mk()
        dt_courant = one_step(ev)
        read_kernels()
        events.append(ev)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    phase_ms = {p: float(np.mean([e[i].elapsed_time(e[i + 1]) for e in events])) for i, p in enumerate(PHASES)}
    step_ms_median = float(np.median([e[0].elapsed_time(e[-1]) for e in events]))
    pairs, maxnc = ws.neighbor_stats()
    out = dict(n=n, wall_s=wall, ms_per_step=1e3 * wall / steps, step_ms_median=step_ms_median, phase_ms=phase_ms,
               pairs=int(pairs),
               max_nc=int(maxnc), rate=n * steps / wall,
               kernel_ms={k: float(np.mean(v)) for k, v in kernel_ms.items()},
               fields={k: ps.host(k).astype(np.float64) for k in COMPARED},
               nc=ps.host("nc").copy())
    out["fields"]["rho"] = out["fields"]["kx"] * ps.host("m").astype(np.float64) / out["fields"]["xm"]

    # the integrate phase (HydroVeProp::integrate: rho and global time step, positions and energy,
    # smoothing lengths), timed apart from the hydro pass that the headline value measures
    ts = abi.Timestep.defaults(minDt=1e-4, minDt_m1=1e-4)
    int_ms, kms = [], {"positions": [], "update_h": []}
    for it in range(warmup + steps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ts.minDtCourant = dt_courant
        e0.record(stream)
        ws.rho_timestep(f, params, 0, n, ts)
        integrate(ws, ps, box, params, ts)
        e1.record(stream)
        torch.cuda.synchronize()
        if it >= warmup:
            int_ms.append(e0.elapsed_time(e1))
            for k in kms:
                kms[k].append(ws.kernel_ms(k))
    out["integrate"] = dict(ms=float(np.mean(int_ms)), kernel_ms={k: float(np.mean(v)) for k, v in kms.items()})
    del ps, f
    ws.close()
    torch.cuda.empty_cache()
    return out


# main kernels of the std pass (sphx_kernel_ms names)
STD_KERNELS = ("search", "density_std", "iad_std", "momentum_std", "sfc_keys", "sort", "permute", "tree")


def run_std(arr, spec, mixed, steps, warmup):
    """Side measurement of the std formulation on the same particles (u-based EOS): domain sync,
    search with the h iteration, density, EOS, IAD, momentum-energy, with record chaining. Not
    the headline value: the headline configuration is the VE pass."""
    import torch

    from sphexa_amd import abi
    from sphexa_amd.lib import Workspace
    from sphexa_amd.particles import STD_FIELDS, ParticleSet, ideal_gas_cv, std_step

    a = dict(arr)
    a["u"] = a.pop("temp") * ideal_gas_cv(10.0, 5.0 / 3.0)
    ws = Workspace(0)
    box = abi.Box.make(spec["lo"], spec["hi"], abi.PERIODIC)
    params = abi.Params.defaults(ws.K, hydro_type=abi.SPHX_F32 if mixed else abi.SPHX_F64, minDt=1e-4)
    ps = ParticleSet.from_host(a, STD_FIELDS, mixed=mixed)
    stream = torch.cuda.current_stream()
    ws.set_profiling(True)
    for _ in range(warmup):
        std_step(ws, ps, box, params)
    torch.cuda.synchronize()
    ms, kms = [], {k: [] for k in STD_KERNELS}
    for _ in range(steps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        std_step(ws, ps, box, params)
        e1.record(stream)
        torch.cuda.synchronize()
        ms.append(e0.elapsed_time(e1))
        for k in kms:
            kms[k].append(ws.kernel_ms(k))
    t = float(np.mean(ms))
    out = dict(ms_per_step=round(t, 4), rate=round(ps.n / (t * 1e-3), 1),
               kernel_ms={k: round(float(np.mean(v)), 4) for k, v in kms.items()})
    del ps, ws
    torch.cuda.empty_cache()
    return out


def cpu_baseline(side, passes):
    from oracle import oracle as O
    from oracle import pipeline
    from sphexa_amd import abi

    kind = "reference" if O.available("ref") else "port"
    lib = O.load("ref" if kind == "reference" else "orc")
    arr, spec = make_ic(side, seed=7)
    box = abi.Box.make(spec["lo"], spec["hi"], abi.PERIODIC)
    params = abi.Params.defaults(lib.kernel_K(6.0), minDt=1e-4)
    times = []
    for _ in range(passes):
        t0 = time.perf_counter()
        pipeline.run(lib, arr, box, params, form="ve")
        times.append(time.perf_counter() - t0)
    n = side ** 3
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return dict(value=n / float(np.median(times)), unit="particle-updates/s", cores=threads, kind=kind,
                sample=f"VE FP64 pass (keys, sort, tree, findNeighbors+h-iteration, XMass, VeDefGradh, EOS, "
                       f"IAD+divv/curlv, AVswitches, momentum-energy) of the same turbulence workload at "
                       f"{side}^3 = {n} particles, median of {passes} passes, OpenMP threads = {threads}")


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: --gpus {args.gpus} is not supported (multi-GPU sharding is a "
              f"design note, not part of this build)", file=sys.stderr)
        sys.exit(2)
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if world > 1:
        print("bench.py is single-GPU; launch it without torch.distributed", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy()
        return

    import torch

    arr, spec = make_ic(args.side)
    forms = [f.strip() for f in args.forms.split(",") if f.strip()]
    res = {}
    for form in forms:
        res[form] = run_gpu(arr, spec, form == "mixed", args.steps, args.warmup)
        print(f"# {form}: {res[form]['ms_per_step']:.3f} ms/step, phases (ms) "
              + ", ".join(f"{k}={v:.3f}" for k, v in res[form]["phase_ms"].items()), file=sys.stderr, flush=True)

    std = {}
    if not args.no_std:
        for form in forms:
            std[form] = run_std(arr, spec, form == "mixed", args.steps, args.warmup)
            print(f"# std {form}: {std[form]['ms_per_step']:.3f} ms/step, kernels (ms) "
                  + ", ".join(f"{k}={v:.3f}" for k, v in std[form]["kernel_ms"].items()), file=sys.stderr,
                  flush=True)

    head = res.get("mixed") or res[forms[0]]
    mixed = "mixed" in res
    mom_ms = head["kernel_ms"]["momentum_ve"]
    flops = FLOPS_PER_PAIR_MOM_VE * head["pairs"] + FLOPS_PER_TARGET_MOM_VE * head["n"]
    achieved = flops / (mom_ms * 1e-3) / 1e12
    peak = PEAK_FP32_VECTOR_TFLOPS if mixed else PEAK_FP64_VECTOR_TFLOPS
    traffic = MEASURED_TRAFFIC_MOM_VE.get((args.side, "mixed" if mixed else "fp64"))
    # compute-bound roofline: MI355X's FP32 / FP64 matrix peaks (157.3 / 78.6 TFLOP/s) equal its
    # non-packed vector peaks, and this gather-reduce runs on the vector ALUs ("pipe")
    roofline = dict(bound="mfma", pipe="valu", kernel="momentumVeKernel", achieved=round(achieved, 3), peak=peak,
                    unit="TFLOP/s", frac=round(achieved / peak, 4), traffic=traffic,
                    traffic_gbs=round(traffic / (mom_ms * 1e-3) / 1e9, 1) if traffic else None,
                    hbm_peak_gbs=PEAK_HBM_GBS,
                    flops_per_launch=int(flops), avg_launch_ms=round(mom_ms, 4))

    cpu = None if args.no_cpu else cpu_baseline(args.cpu_side, args.cpu_passes)
    line = {
        "metric": METRIC,
        "value": round(head["rate"], 1),
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-mixed" if mixed else "f64",
        "data": "synthetic",
        "config": {"workload": "turbulence-box VE pass, 200^3 = 8M particles, periodic unit box, Mach 0.3, ng0 100",
                   "particles": head["n"], "formulation": "VE", "precision": "mixed" if mixed else "fp64",
                   "bucket_size": 64, "ngmax": 150},
        "fp64_value": round(res["fp64"]["rate"], 1) if "fp64" in res else None,
        "mixed_over_fp64": round(res["mixed"]["rate"] / res["fp64"]["rate"], 3) if len(res) == 2 else None,
        # the device time of one pass from its first to its last event, median over the timed passes
        "step_ms_median": {f: round(r["step_ms_median"], 4) for f, r in res.items()},
        # particle-updates/s of the pair loops alone (XMass to momentum-energy kernel times)
        "pair_loop_rate": {f: round(r["n"] / (1e-3 * sum(r["kernel_ms"][k] for k in PAIR_KERNELS)), 1)
                           for f, r in res.items()},
        "phase_ms": {f: {k: round(v, 4) for k, v in r["phase_ms"].items()} for f, r in res.items()},
        "kernel_ms": {f: {k: round(v, 4) for k, v in r["kernel_ms"].items()} for f, r in res.items()},
        "mixed_error": ({"reference": "fp64 form of this pass", "fields": field_errors(res["mixed"]["fields"],
                                                                                       res["fp64"]["fields"]),
                         "nc_equal_fraction": float(np.mean(res["mixed"]["nc"] == res["fp64"]["nc"]))}
                        if len(res) == 2 else None),
        "neighbor_pairs": head["pairs"],
        "integrate": {f: dict(ms_per_step=round(r["integrate"]["ms"], 4),
                              kernel_ms={k: round(v, 4) for k, v in r["integrate"]["kernel_ms"].items()},
                              positions_bytes_per_particle=POSITIONS_BYTES[f],
                              positions_gbs=round(POSITIONS_BYTES[f] * r["n"]
                                                  / (r["integrate"]["kernel_ms"]["positions"] * 1e-3) / 1e9, 1),
                              positions_hbm_frac=round(POSITIONS_BYTES[f] * r["n"]
                                                       / (r["integrate"]["kernel_ms"]["positions"] * 1e-3) / 1e9
                                                       / PEAK_HBM_GBS, 4))
                      for f, r in res.items()},
        # the std formulation on the same 8M particles (u-based EOS), per form: ms per pass,
        # particle-updates/s and kernel times; a side measurement, not the headline
        "std_pass": std or None,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
