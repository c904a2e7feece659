"""The C++ host interface (include/sphexa_amd.hpp) driven like the reference propagators
(tests/cpp/dataset_api.cpp) reproduces the CPU oracle's pass."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from sphexa_amd import abi, ic
from sphexa_amd.particles import STD_FIELDS, VE_FIELDS, ParticleSet, ideal_gas_cv, std_step, ve_step

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cpp", "dataset_api")

STD = ("rho", "p", "c", "c11", "c12", "c13", "c22", "c23", "c33", "ax", "ay", "az", "du")
VE = ("xm", "kx", "gradh", "prho", "c", "c11", "c12", "c13", "c22", "c23", "c33", "divv", "curlv", "alpha", "ax", "ay",
      "az", "du")


@pytest.fixture(scope="module")
def ws():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a GPU"
    from sphexa_amd.lib import Workspace

    return Workspace(0)


def _pass_ic(form):
    if form == "std":
        arr, spec = ic.sedov_lattice(16, seed=21)
        arr["vx"] = 0.1 * np.sin(6.0 * arr["x"])
        arr["vy"] = 0.05 * np.cos(5.0 * arr["y"])
        arr["vz"] = 0.02 * arr["z"]
        third = arr["u"]
    else:
        arr, spec = ic.turbulence_box(16, seed=22)
        arr["temp"] = arr.pop("u") / ideal_gas_cv(10.0, 5.0 / 3.0)
        arr["alpha"] = np.random.default_rng(2).uniform(0.05, 1.0, arr["x"].size)
        third = arr["temp"]
    return arr, spec, third


def _run_pass(tmp_path, arr, spec, third, form, calls):
    """writes the IC, runs the driver, returns (fields by name, nc, minDtCourant)"""
    assert os.path.exists(DRIVER), "tests/cpp/dataset_api not built (python -m sphexa_amd.build)"
    n = arr["x"].size
    with open(tmp_path / "in.bin", "wb") as fh:
        fh.write(np.array([n], np.uint64).tobytes())
        fh.write(np.array([1], np.uint32).tobytes())
        fh.write(np.array([spec["lo"], spec["hi"]], np.float64).tobytes())
        for col in (arr["x"], arr["y"], arr["z"], arr["h"], arr["m"], third, arr["vx"], arr["vy"], arr["vz"],
                    arr["alpha"]):
            fh.write(np.ascontiguousarray(col, np.float64).tobytes())
    r = subprocess.run([DRIVER, str(tmp_path / "in.bin"), str(tmp_path / "out.bin"), form, calls], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw = open(tmp_path / "out.bin", "rb").read()
    names = ("x",) + (VE if form.startswith("ve") else STD)
    k = len(names) * n * 8
    vals = np.frombuffer(raw[:k], np.float64).reshape(len(names), n)
    nc = np.frombuffer(raw[k:k + 4 * n], np.uint32)
    dt = float(np.frombuffer(raw[k + 4 * n:], np.float64)[0])
    return dict(zip(names, vals)), nc, dt


@pytest.mark.parametrize("calls", ["cpu", "gpu"])
@pytest.mark.parametrize("form", ["std", "ve"])
def test_cpp_interface_pass(tmp_path, orc, form, calls):
    """calls="cpu": findNeighborsSfc then computeDensity / computeXMass; calls="gpu": the fused
    search + density / XMass, as the reference GPU build runs them"""
    from oracle import pipeline

    arr, spec, third = _pass_ic(form)
    got, nc, dt = _run_pass(tmp_path, arr, spec, third, form, calls)
    box = abi.Box.make(spec["lo"], spec["hi"], abi.PERIODIC)
    o = pipeline.run(orc, arr, box, abi.Params.defaults(orc.kernel_K(6.0), minDt=1e-3), form=form)
    assert np.array_equal(got["x"], o["a"]["x"])
    assert np.array_equal(nc, o["a"]["nc"])
    for name in (VE if form == "ve" else STD):
        ref = o["a"][name]
        err = np.max(np.abs(got[name] - ref)) / max(np.max(np.abs(ref)), 1e-300)
        assert err < 1e-10, (name, err)
    assert abs(dt - o["minDt"]) <= 1e-12 * o["minDt"]


@pytest.mark.parametrize("calls", ["cpu", "gpu"])
@pytest.mark.parametrize("form", ["std", "ve"])
def test_cpp_interface_mixed_pass(tmp_path, orc, ws, form, calls):
    """DeviceDataset<float> (the reference's default HydroType) through the C++ interface: bitwise
    the pass of the C ABI's mixed form driven from Python on the same IC (within 1e-5 of each
    field's largest magnitude after the fused search + first sum), and within the mixed
    form's bars of the FP64 oracle (density <= 1e-5 pointwise, every other output <= 1e-4 of its
    largest magnitude, nc and the SFC order identical)."""
    from oracle import pipeline

    arr, spec, third = _pass_ic(form)
    got, nc, dt = _run_pass(tmp_path, arr, spec, third, form + "32", calls)
    names = VE if form == "ve" else STD
    box = abi.Box.make(spec["lo"], spec["hi"], abi.PERIODIC)

    params = abi.Params.defaults(ws.K, hydro_type=abi.SPHX_F32, minDt=1e-3)
    ps = ParticleSet.from_host(arr, VE_FIELDS if form == "ve" else STD_FIELDS, mixed=True)
    conv, pdt = (ve_step if form == "ve" else std_step)(ws, ps, box, params)
    assert conv
    assert np.array_equal(got["x"], ps.host("x"))
    assert np.array_equal(nc, ps.host("nc"))
    for name in names:
        mine = ps.host(name).astype(np.float64)
        if calls == "cpu":
            assert np.array_equal(got[name], mine), name
        else:
            # the fused search + sum forms r_ij without FMA: the first sum differs within rounding
            assert np.max(np.abs(got[name] - mine)) <= 1e-5 * np.max(np.abs(mine)), name
    assert dt == pdt if calls == "cpu" else abs(dt - pdt) <= 1e-5 * pdt

    o = pipeline.run(orc, arr, box, abi.Params.defaults(orc.kernel_K(6.0), minDt=1e-3), form=form)
    assert np.array_equal(got["x"], o["a"]["x"])
    assert np.array_equal(nc, o["a"]["nc"])
    if form == "std":
        rho_g, rho_o = got["rho"], o["a"]["rho"]
    else:
        rho_g = got["kx"] * ps.host("m") / got["xm"]
        rho_o = o["a"]["kx"] * o["a"]["m"] / o["a"]["xm"]
    rel = np.max(np.abs(rho_g - rho_o) / np.abs(rho_o))
    # the float fields carry float rounding: the driver did not run the FP64 form
    assert 1e-9 < rel <= 1e-5, rel
    for name in names:
        ref = o["a"][name]
        err = np.max(np.abs(got[name] - ref)) / max(np.max(np.abs(ref)), 1e-300)
        assert err <= 1e-4, (name, err)
    assert abs(dt - o["minDt"]) <= 1e-5 * o["minDt"]


@pytest.mark.parametrize("periodic,mixed", [(1, False), (0, False), (1, True)])
def test_cpp_interface_steps(tmp_path, orc, periodic, mixed):
    """Three full std steps through the C++ interface (Domain::sync ... computeMomentumEnergySTD,
    then computeTimestep, computePositions, updateSmoothingLength) against the oracle's three
    steps, both updating the tree once per sync after the first: positions, velocities, h, u and
    the time-step state. In the open box the domain refits
    the box at every sync, and so does the oracle. mixed: DeviceDataset<float> against the FP64
    oracle with the mixed form's step bars (positions 1e-9, h 3e-3 (an FP32 h can move a
    borderline pair across 2h and nc by one), velocities and u 1e-4, time steps 1e-5)."""
    from test_gpu_integrate import _oracle_steps

    arr, spec = ic.sedov_lattice(16, seed=17)
    n = arr["x"].size
    with open(tmp_path / "in.bin", "wb") as fh:
        fh.write(np.array([n], np.uint64).tobytes())
        fh.write(np.array([periodic], np.uint32).tobytes())
        fh.write(np.array([spec["lo"], spec["hi"]], np.float64).tobytes())
        for col in (arr["x"], arr["y"], arr["z"], arr["h"], arr["m"], arr["u"], arr["vx"], arr["vy"], arr["vz"],
                    np.zeros(n)):
            fh.write(np.ascontiguousarray(col, np.float64).tobytes())
    r = subprocess.run([DRIVER, str(tmp_path / "in.bin"), str(tmp_path / "out.bin"), "std32" if mixed else "std", "cpu",
                        "3"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw = np.frombuffer(open(tmp_path / "out.bin", "rb").read(), np.float64)
    names = ("x", "y", "z", "vx", "vy", "vz", "h", "u")
    got = {k: raw[i * n:(i + 1) * n] for i, k in enumerate(names)}
    minDt, minDt_m1, ttot = raw[len(names) * n:len(names) * n + 3]
    q = raw[len(names) * n + 3:]

    box = abi.Box.make(spec["lo"], spec["hi"], abi.PERIODIC if periodic else abi.OPEN)
    for k in ("x_m1", "y_m1", "z_m1", "du_m1"):
        arr[k] = np.zeros(n)
    ts = abi.Timestep.defaults(minDt=1e-4, minDt_m1=1e-4)
    # the C++ Domain converges the tree on its first sync and updates it once on each later one
    a, _ = _oracle_steps(orc, arr, box, abi.Params.defaults(orc.kernel_K(6.0)), ts, 3, np.float64,
                         refit_box=not periodic, tree_update=True)
    tol_dt = 1e-5 if mixed else 1e-12
    assert abs(minDt - ts.minDt) <= tol_dt * ts.minDt and abs(minDt_m1 - ts.minDt_m1) <= tol_dt * ts.minDt_m1
    assert abs(ttot - ts.ttot) <= tol_dt * ts.ttot
    # the conserved quantities of the final state against the oracle's on its own final state
    oq = np.zeros(9)
    orc.conserved_quantities(C.byref(abi.fields_from(a)), C.byref(abi.Params.defaults(1.0)), 0, n, oq.ctypes.data)
    np.testing.assert_allclose([q[0], q[1]], oq[:2], rtol=1e-5 if mixed else 1e-10)
    assert q[2] == 0.0
    # sum nc: a borderline pair may flip in the mixed form
    assert abs(q[9] - oq[8]) <= (1e-4 * oq[8] if mixed else 0)
    og, oo = np.argsort(got["x"], kind="stable"), np.argsort(a["x"], kind="stable")
    for k in names:
        ref = np.asarray(a[k], np.float64)[oo]
        scale = 1.0 if mixed and k in "xyz" else max(np.max(np.abs(ref)), 1e-300)
        err = np.max(np.abs(got[k][og] - ref)) / scale
        tol = ({"x": 1e-9, "y": 1e-9, "z": 1e-9, "h": 3e-3}.get(k, 1e-4)) if mixed else 1e-10
        assert err < tol, (k, err)


@pytest.mark.parametrize("mixed", [False, True])
def test_cpp_interface_block_timesteps(tmp_path, ws, mixed):
    """The block time-step calls of the C++ interface (groupDivvTimestep, groupAccTimestep, storeRung,
    computePositions with rungs, driftPositions; tests/cpp/dataset_api.cpp "bdt") after a VE pass,
    bitwise the same C ABI calls driven from Python with the arguments HydroVeBdtProp passes
    (Krho, etaAcc sqrt(eps) as float; the previous steps per rung), which the GPU integration tests
    check against the oracle."""
    import torch

    from sphexa_amd.particles import VE_FIELDS

    arr, spec, third = _pass_ic("ve")
    n = arr["x"].size
    with open(tmp_path / "in.bin", "wb") as fh:
        fh.write(np.array([n], np.uint64).tobytes())
        fh.write(np.array([1], np.uint32).tobytes())
        fh.write(np.array([spec["lo"], spec["hi"]], np.float64).tobytes())
        for col in (arr["x"], arr["y"], arr["z"], arr["h"], arr["m"], third, arr["vx"], arr["vy"], arr["vz"],
                    arr["alpha"]):
            fh.write(np.ascontiguousarray(col, np.float64).tobytes())
    r = subprocess.run([DRIVER, str(tmp_path / "in.bin"), str(tmp_path / "out.bin"), "ve32" if mixed else "ve",
                        "bdt"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw = open(tmp_path / "out.bin", "rb").read()
    ng = (n + 63) // 64
    gdt_cpp = np.frombuffer(raw[:8 * ng], np.float64)
    rungs_cpp = np.frombuffer(raw[8 * ng:8 * ng + n], np.uint8)
    rest = np.frombuffer(raw[8 * ng + n:], np.float64).reshape(7, n)

    box = abi.Box.make(spec["lo"], spec["hi"], abi.PERIODIC)
    params = abi.Params.defaults(ws.K, hydro_type=abi.SPHX_F32 if mixed else abi.SPHX_F64, minDt=1e-3)
    for k in ("x_m1", "y_m1", "z_m1", "du_m1"):
        arr[k] = np.zeros(n)
    ps = ParticleSet.from_host(arr, VE_FIELDS + ("x_m1", "y_m1", "z_m1", "du_m1"), mixed=mixed)
    conv, _ = ve_step(ws, ps, box, params)
    assert conv
    f = ps.fields()
    gdt = torch.full((ng,), 1e10, dtype=torch.float32, device="cuda")
    ws.group_divv_timestep(f, params, 0, n, float(np.float32(0.06)), gdt.data_ptr())
    ws.group_acc_timestep(f, params, 0, n, float(np.float32(0.2 * np.sqrt(0.005))), gdt.data_ptr())
    rungs = torch.zeros(n, dtype=torch.uint8, device="cuda")
    ws.store_rung(0, n, 1, rungs.data_ptr())
    dt_m1 = [1e-4, 2e-4, 4e-4, 8e-4]
    ws.compute_positions_rungs(box, f, ps.history(), params, 0, n, float(np.float32(2e-4)), dt_m1, rungs.data_ptr())
    ws.drift_positions(f, ps.history(), params, 0, n, float(np.float32(3e-4)), float(np.float32(1e-4)), dt_m1,
                       rungs.data_ptr())
    torch.cuda.synchronize()
    assert np.array_equal(gdt_cpp, gdt.cpu().numpy().astype(np.float64))
    assert np.all(gdt_cpp < 1e10)  # both limits applied
    assert np.array_equal(rungs_cpp, rungs.cpu().numpy()) and np.all(rungs_cpp == 1)
    for row, k in zip(rest, ("x", "y", "z", "vx", "vy", "vz", "temp")):
        assert np.array_equal(row, ps.host(k).astype(np.float64)), k
