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

# Algorithmic flops per unit of each kernel of the VE pass, counting + - * as 1, division and
# square root as 1, each table lookup as 7 (index, two loads' difference, slope, offset, fma),
# whichever way the kernel obtains W:
#  search        per candidate distance test: 3 differences, 3 squares, 2 adds (FP64 in both forms,
#                the reference's exact predicate), plus the h iteration's re-searches
#  xmass         per pair 19 (xmassJLoop, sph/include/sph/hydro_ve/xmass_kern.hpp:51-79)
#  ve_def_gradh  per pair 33 (veDefGradhJLoop, hydro_ve/ve_def_gradh_kern.hpp:44-90)
#  iad_divv_curlv per pair 37 + 59 (IADJLoop, hydro_ve/iad_kern.hpp:44-109; divV_curlVJLoop,
#                hydro_ve/divv_curlv_kern.hpp:44-123)
#  av_switches   per pair 58 (AVswitchesJLoop, hydro_ve/av_switches_kern.hpp:44-137)
#  momentum_ve   per pair 142 (momentumAndEnergyJLoop, hydro_ve/momentum_energy_kern.hpp:121-212,
#                avClean = false: separation 3, fold/r2/dist 6, v_ij 3, 1/h_j 1, v1 v2 2, h_j^-3 2,
#                W_i W_j 2+14, IAD projections 2*(3*5+1) = 32, rho_j 2, rv 5, w_ij 1, viscosity 9,
#                signal 4, Atwood 4 + branch 2, a_mom 2, AV terms 2+2+3*4 = 16, AV energy 5,
#                energy 7, momentum 3*6 = 18, accumulations 4), plus 40 per target (rho_i, 1/h_i,
#                h_i^-3, eta_crit, epilogue)
FLOPS_PER_TEST_SEARCH = 8
FLOPS_PER_PAIR = {"xmass": 19, "ve_def_gradh": 33, "iad_divv_curlv": 37 + 59, "av_switches": 58, "momentum_ve": 142}
FLOPS_PER_TARGET = {"momentum_ve": 40}
FLOPS_PER_PAIR_MOM_VE = FLOPS_PER_PAIR["momentum_ve"]

# HBM bytes per launch on the default workload (200^3), from rocprofv3 FETCH_SIZE and WRITE_SIZE
# counter runs scaled by a plain copy kernel of known size (read x 2.0, write x 1.0); see the
# counter summaries under profiles/. Counters cannot be read from inside the timed run, so the
# measured figures are carried here for the default workload only.
MEASURED_TRAFFIC = {("search", 200, "mixed"): 6.034e9, ("search", 200, "fp64"): 7.291e9,
                    ("momentum_ve", 200, "mixed"): 4.987e9, ("momentum_ve", 200, "fp64"): 6.926e9,
                    ("av_switches", 200, "mixed"): 4.815e9, ("av_switches", 200, "fp64"): 6.398e9,
                    ("iad_divv_curlv", 200, "mixed"): 4.297e9, ("iad_divv_curlv", 200, "fp64"): 5.343e9,
                    ("ve_def_gradh", 200, "mixed"): 4.054e9, ("ve_def_gradh", 200, "fp64"): 4.727e9,
                    ("xmass", 200, "mixed"): 3.638e9, ("xmass", 200, "fp64"): 4.143e9}


def parse():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[1])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--side", type=int, default=200, help="lattice side (N = side^3)")
    ap.add_argument("--cpu-passes", type=int, default=2, help="passes of the CPU baseline (median)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-std", action="store_true", help="skip the std-pass side measurement")
    ap.add_argument("--no-side", action="store_true",
                    help="skip the side workloads (Evrard sphere, adjoint pass, checkpoint I/O)")
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
               pairs=int(pairs), search_tests=int(ws.search_tests()),
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


CHECKPOINT_FIELDS = ("x", "y", "z", "h", "m", "u", "vx", "vy", "vz", "x_m1", "y_m1", "z_m1", "du_m1")


def run_checkpoint_io(arr, spec, mixed, reps=2):
    """Side measurement of the H5Part field I/O: the std propagator's checkpoint fields (the conserved
    fields with the x, y, z and du history) of the 8M particles written from HBM into one step of an
    H5Part file and read back into HBM, through the host staging buffer. Wall-clock times of the
    whole call (HDF5 + PCIe), best of `reps`; the read follows the write, so it comes from the page
    cache. None when the I/O library is not built."""
    import shutil
    import tempfile

    import torch

    from sphexa_amd import h5io
    from sphexa_amd.particles import ParticleSet

    if not os.path.exists(h5io.LIB_PATH):
        return None
    a = {k: arr[k] for k in CHECKPOINT_FIELDS if k in arr}
    a["u"] = arr["temp"]
    for k in CHECKPOINT_FIELDS:
        a.setdefault(k, np.zeros_like(arr["x"]))
    ps = ParticleSet.from_host(a, CHECKPOINT_FIELDS, mixed=mixed)
    n = ps.n
    nbytes = sum(ps.t[k][:n].numel() * ps.t[k].element_size() for k in CHECKPOINT_FIELDS)
    back = {k: torch.empty_like(ps.t[k][:n]) for k in CHECKPOINT_FIELDS}
    tmp = tempfile.mkdtemp(prefix="sphx_ckpt_")
    path = os.path.join(tmp, "checkpoint.h5")
    tw, tr = [], []
    try:
        for _ in range(reps):
            if os.path.exists(path):
                os.remove(path)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            with h5io.H5File(path, h5io.WRITE) as f:
                f.add_step(n)
                for k in CHECKPOINT_FIELDS:
                    f.write_field(k, ps.t[k][:n])
            tw.append(time.perf_counter() - t0)
            t0 = time.perf_counter()
            with h5io.H5File(path) as f:
                f.set_step(-1)
                for k in CHECKPOINT_FIELDS:
                    f.read_field(k, out=back[k])
            torch.cuda.synchronize()
            tr.append(time.perf_counter() - t0)
        same = all(torch.equal(back[k], ps.t[k][:n]) for k in CHECKPOINT_FIELDS)
        fsize = os.path.getsize(path)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    del ps, back
    torch.cuda.empty_cache()
    return dict(fields=len(CHECKPOINT_FIELDS), particles=n, field_bytes=nbytes, file_bytes=fsize,
                write_s=round(min(tw), 4), read_s=round(min(tr), 4),
                write_gbs=round(nbytes / min(tw) / 1e9, 3), read_gbs=round(nbytes / min(tr) / 1e9, 3),
                round_trip_bitwise=bool(same))


EVRARD_KERNELS = ("search", "xmass", "ve_def_gradh", "iad_divv_curlv", "av_switches", "momentum_ve", "sfc_keys",
                  "sort", "permute", "tree")


def run_evrard(n, mixed, steps, warmup):
    """Side measurement on the load-balance workload: the Evrard-like sphere (rho ~ 1/r, h spanning
    ~12x, open box [-1,1]^3), VE pass with record chaining; ms per pass, particle-updates/s and
    kernel times, after a first pass that converges h from the initial estimate."""
    import torch

    from sphexa_amd import abi, ic
    from sphexa_amd.lib import Workspace
    from sphexa_amd.particles import VE_FIELDS, ParticleSet, ideal_gas_cv, ve_step

    arr, spec = ic.evrard_sphere(n, seed=11)
    arr["temp"] = arr.pop("u") / ideal_gas_cv(10.0, 5.0 / 3.0)
    ws = Workspace(0)
    box = abi.Box.make(spec["lo"], spec["hi"], abi.OPEN)
    params = abi.Params.defaults(ws.K, hydro_type=abi.SPHX_F32 if mixed else abi.SPHX_F64, minDt=1e-4)
    ps = ParticleSet.from_host(arr, VE_FIELDS, mixed=mixed)
    stream = torch.cuda.current_stream()
    ws.set_profiling(True)
    for _ in range(1 + warmup):
        ve_step(ws, ps, box, params)
    torch.cuda.synchronize()
    ms, kms = [], {k: [] for k in EVRARD_KERNELS}
    for _ in range(steps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        ve_step(ws, ps, box, params)
        e1.record(stream)
        torch.cuda.synchronize()
        ms.append(e0.elapsed_time(e1))
        for k in kms:
            kms[k].append(ws.kernel_ms(k))
    pairs, maxnc = ws.neighbor_stats()
    nc = ps.host("nc")
    t = float(np.mean(ms))
    out = dict(particles=ps.n, ms_per_step=round(t, 4), rate=round(ps.n / (t * 1e-3), 1),
               nc_mean=round(float(nc.mean()), 2), nc_max=int(maxnc), pairs=int(pairs),
               search_tests=int(ws.search_tests()),
               kernel_ms={k: round(float(np.mean(v)), 4) for k, v in kms.items()})
    del ps, ws
    torch.cuda.empty_cache()
    return out


def run_adjoint(side, mixed, steps, warmup):
    """The differentiation workload: the Sedov-like geometry at side^3 particles, std pass with a
    per-particle AV coefficient (forward: density, EOS, IAD, momentum-energy over the lists of one
    search) and its reverse pass through the three adjoint entry points for
    L = sum (rho - rho_t)^2 + 1e-3 sum |a|^2 (rho_t = 1 + 0.05 sin(seeded phase)), as dL/dx, dL/dh
    and dL/dalpha. The search and domain sync are not repeated per step: the lists are the
    forward's checkpoint."""
    import torch

    from sphexa_amd import abi, ic
    from sphexa_amd.lib import Workspace
    from sphexa_amd.particles import STD_FIELDS, ParticleSet

    arr, spec = ic.sedov_lattice(side, seed=42)
    ws = Workspace(0)
    box = abi.Box.make(spec["lo"], spec["hi"], abi.PERIODIC)
    hydro = abi.SPHX_F32 if mixed else abi.SPHX_F64
    params = abi.Params.defaults(ws.K, hydro_type=hydro)
    ps = ParticleSet.from_host(arr, STD_FIELDS + ("alpha",), mixed=mixed)
    n = ps.n
    ws.domain_sync(box, ps.sync_fields(), params, n)
    f = ps.fields()
    ws.find_neighbors(box, f, params, 0, n, True)
    params.flags |= abi.SPHX_FLAG_STD_ALPHA
    dt = ps.hydro_dtype
    rng = np.random.default_rng(5)
    phase = torch.from_numpy(rng.uniform(0, 2 * np.pi, 3)).cuda()
    x = ps.t["x"][:n]
    rho_t = 1.0 + 0.05 * torch.sin(2 * np.pi * x + phase[0])
    w = 1e-3
    z = lambda: torch.zeros(n, dtype=dt, device="cuda")
    g = {k: z() for k in ("x", "y", "z", "h", "rho", "p", "alpha", "c11", "c12", "c13", "c22", "c23", "c33")}
    cot = {k: z() for k in ("rho", "c11", "c12", "c13", "c22", "c23", "c33", "ax", "ay", "az")}
    gm = abi.Adjoint()
    for k, t in g.items():
        setattr(gm, k, t.data_ptr())
    cm = abi.Adjoint()
    for k in ("ax", "ay", "az"):
        setattr(cm, k, cot[k].data_ptr())
    gi = abi.Adjoint()
    for k in ("x", "y", "z", "h", "rho"):
        setattr(gi, k, g[k].data_ptr())
    ci = abi.Adjoint()
    for k in ("c11", "c12", "c13", "c22", "c23", "c33"):
        setattr(ci, k, g[k].data_ptr())  # the momentum adjoint's c_ij gradients are the IAD cotangents
    gd = abi.Adjoint()
    for k in ("x", "y", "z", "h"):
        setattr(gd, k, g[k].data_ptr())
    cd = abi.Adjoint()
    cd.rho = cot["rho"].data_ptr()
    u = ps.t["u"][:n]
    stream = torch.cuda.current_stream()

    def step():
        for t in g.values():
            t.zero_()
        ws.pair("sphx_density_std", box, f, params, 0, n)
        ws.eos("sphx_eos_std", f, params, 0, n)
        ws.pair("sphx_iad_std", box, f, params, 0, n)
        ws.momentum_energy_std(box, f, params, 0, n)
        for k in ("ax", "ay", "az"):
            cot[k].copy_(2.0 * w * ps.t[k][:n])
        ws.vjp("sphx_momentum_energy_std_vjp", box, f, params, 0, n, cm, gm)
        # rho's cotangent: the loss term, and p = rho u (gamma - 1) of the EOS
        g["rho"].add_((2.0 * (ps.t["rho"][:n].double() - rho_t) + g["p"].double() * u * (params.gamma - 1.0))
                      .to(dt))
        ws.vjp("sphx_iad_std_vjp", box, f, params, 0, n, ci, gi)
        cot["rho"].copy_(g["rho"])
        ws.vjp("sphx_density_std_vjp", box, f, params, 0, n, cd, gd)

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    ms = []
    for _ in range(steps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        step()
        e1.record(stream)
        torch.cuda.synchronize()
        ms.append(e0.elapsed_time(e1))
    t = float(np.mean(ms))
    finite = all(bool(torch.isfinite(g[k]).all()) for k in ("x", "y", "z", "h", "alpha"))
    out = dict(particles=n, ms_per_step=round(t, 4), rate=round(n / (t * 1e-3), 1), gradients_finite=finite)
    del ps, ws, g, cot
    torch.cuda.empty_cache()
    return out


def kernel_rooflines(r, mixed):
    """Roofline entry of every kernel of the VE pass: algorithmic flops per launch (per unit above
    times the units of the launch) over the kernel's mean duration from HIP events on its stream,
    against the FP64 vector peak for FP64 work (the search's distance tests in both forms) and the
    FP32 peaks for the FP32 pair loops: 157.3 TFLOP/s with packed FP32 issue, 78.6 TFLOP/s without
    (the library is compiled with packed FP32 off)."""
    out = {}
    side = round(r["n"] ** (1.0 / 3.0))
    form = "mixed" if mixed else "fp64"
    for k in ("search",) + PAIR_KERNELS:
        ms = r["kernel_ms"][k]
        if k == "search":
            flops = FLOPS_PER_TEST_SEARCH * r["search_tests"]
            fp64 = True
            units = dict(candidate_tests=r["search_tests"], tests_per_stored_pair=round(r["search_tests"] / r["pairs"], 2))
        else:
            flops = FLOPS_PER_PAIR[k] * r["pairs"] + FLOPS_PER_TARGET.get(k, 0) * r["n"]
            fp64 = not mixed
            units = dict(pairs=r["pairs"])
        ach = flops / (ms * 1e-3) / 1e12
        e = dict(bound="valu", achieved=round(ach, 3), unit="TFLOP/s", avg_launch_ms=round(ms, 4),
                 flops_per_launch=int(flops), precision="fp64" if fp64 else "fp32", **units)
        if fp64:
            e.update(peak=PEAK_FP64_VECTOR_TFLOPS, frac=round(ach / PEAK_FP64_VECTOR_TFLOPS, 4))
        else:
            e.update(peak=PEAK_FP32_VECTOR_TFLOPS, frac=round(ach / PEAK_FP32_VECTOR_TFLOPS, 4),
                     peak_nonpacked=PEAK_FP64_VECTOR_TFLOPS, frac_nonpacked=round(ach / PEAK_FP64_VECTOR_TFLOPS, 4))
        tr = MEASURED_TRAFFIC.get((k, side, form))
        e["traffic"] = tr
        if tr:
            e["traffic_gbs"] = round(tr / (ms * 1e-3) / 1e9, 1)
        out[k] = e
    return out


def cpu_info():
    """CPU model, physical cores (distinct physical id / core id pairs of /proc/cpuinfo), logical
    CPUs, CPUs this process may run on, and whether the CPU runs x86-64-v4 (AVX-512) code."""
    model, cores, flags = None, set(), set()
    phys = core = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name" and model is None:
                    model = v
                elif k == "physical id":
                    phys = v
                elif k == "core id":
                    core = v
                elif k == "flags" and not flags:
                    flags = set(v.split())
                elif not k and phys is not None:
                    cores.add((phys, core))
                    phys = core = None
        if phys is not None:
            cores.add((phys, core))
    except OSError:
        pass
    v4 = {"avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl"} <= flags
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    return dict(model=model, physical_cores=len(cores) or None, logical_cpus=os.cpu_count(), affinity=affinity,
                x86_64_v4=v4)


def cpu_baseline(arr, spec, passes):
    """The reference's own CPU path (its headers compiled with its Release flags, -O3 -DNDEBUG and
    OpenMP, -march fixed to the x86-64 level this host runs) timed on the same particles as the GPU
    pass; the bit-exact reference build or, without the reference checkout at build time, the oracle
    when that is all there is."""
    from oracle import oracle as O
    from oracle import pipeline
    from sphexa_amd import abi

    info = cpu_info()
    candidates = [("ref_rel_v4" if info["x86_64_v4"] else "ref_rel_v3", "reference", "-O3 -DNDEBUG -fopenmp "
                   + ("-march=x86-64-v4" if info["x86_64_v4"] else "-march=x86-64-v3")),
                  ("ref", "reference", "-O3 -march=x86-64-v3 -fopenmp -ffp-contract=off"),
                  ("orc", "port", "-O3 -march=x86-64-v3 -fopenmp -ffp-contract=off")]
    lib_kind, kind, flags = next(c for c in candidates if O.available(c[0]))
    lib = O.load(lib_kind)
    box = abi.Box.make(spec["lo"], spec["hi"], abi.PERIODIC)
    params = abi.Params.defaults(lib.kernel_K(6.0), minDt=1e-4)
    times, phases = [], []
    for _ in range(passes):
        t0 = time.perf_counter()
        r = pipeline.run(lib, arr, box, params, form="ve")
        times.append(time.perf_counter() - t0)
        phases.append(r["times"])
        del r
    n = len(arr["x"])
    threads = int(os.environ.get("OMP_NUM_THREADS", info["affinity"]))
    return dict(value=round(n / float(np.median(times)), 1), unit="particle-updates/s", cores=threads, kind=kind,
                sample=f"VE FP64 pass (keys, sort, tree, findNeighbors+h-iteration, XMass, VeDefGradh, EOS, "
                       f"IAD+divv/curlv, AVswitches, momentum-energy) of the same turbulence workload and particles "
                       f"as the GPU value, {n} particles, median of {passes} passes",
                threads=threads, build=flags, library=os.path.basename(O.LIBS[lib_kind]), cpu=info,
                pass_s=[round(t, 3) for t in times],
                phase_s={k: round(float(np.median([p[k] for p in phases])), 3) for k in phases[0]})


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: --gpus {args.gpus} is not supported (multi-GPU sharding is a "
              f"design note, not implemented)", file=sys.stderr)
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

    side_runs = {}
    if not args.no_side:
        side_runs["evrard_4M_ve"] = {f: run_evrard(4_000_000, f == "mixed", args.steps, args.warmup) for f in forms}
        side_runs["adjoint_2M_std"] = {f: run_adjoint(126, f == "mixed", args.steps, args.warmup) for f in forms}
        io = {f: run_checkpoint_io(arr, spec, f == "mixed") for f in forms}
        if all(v is not None for v in io.values()):
            side_runs["checkpoint_io_8M"] = io
            print("# checkpoint_io_8M: " + ", ".join(f"{f} write {v['write_gbs']} GB/s read {v['read_gbs']} GB/s"
                                                      for f, v in io.items()), file=sys.stderr, flush=True)
        for name, r in side_runs.items():
            if name == "checkpoint_io_8M":
                continue
            print(f"# {name}: " + ", ".join(f"{f} {v['ms_per_step']:.3f} ms" for f, v in r.items()), file=sys.stderr,
                  flush=True)

    head = res.get("mixed") or res[forms[0]]
    mixed = "mixed" in res
    rl = {f: kernel_rooflines(r, f == "mixed") for f, r in res.items()}
    head_rl = rl["mixed" if mixed else forms[0]]
    dominant = max(head_rl, key=lambda k: head_rl[k]["avg_launch_ms"])
    roofline = dict(kernel=dominant, **head_rl[dominant])

    cpu = None if args.no_cpu else cpu_baseline(arr, spec, args.cpu_passes)
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
        # the kernel with the largest share of the headline pass, then every kernel of the pass per form
        "roofline": roofline,
        "roofline_kernels": rl,
        # the load-balance workload (Evrard-like sphere, 4M, VE), the differentiation workload
        # (2M, std forward + adjoint) and the checkpoint I/O of the 8M particles (H5Part), per form
        "side_workloads": side_runs or None,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
