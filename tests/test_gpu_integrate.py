"""HIP time integration against the oracle (pinned in tests/test_integrate_oracle.py).

* sphx_compute_positions: positions, velocities, x_m1..z_m1, du_m1 bit for bit, in both forms and
  for open, periodic and fixed boundaries; u / temp bit for bit except where the u < 0 guard calls
  exp (device and host exp may differ in the last bit: 2 ulp allowed there).
* sphx_update_smoothing_length: within 4 ulp (device pow against the C library's).
* sphx_compute_timestep / sphx_rho_timestep: equal to the oracle's (device maxima, host scalars).
* sphx_domain_reorder applies the sync's permutation; sphx_global_box equals the oracle's box update.
* Three full steps of the std propagator (forces, time step, positions, energy, h) on a Sedov-like
  lattice, FP64 and mixed, periodic, and FP64 in an open box refitted at every sync, against the
  oracle's three steps: positions, velocities, u and h after the last step, and the step sizes.
"""
import ctypes as C

import numpy as np
import pytest

from sphexa_amd import abi, ic
from sphexa_amd.particles import STD_FIELDS, ParticleSet, run_steps

from test_integrate_oracle import _positions_case, run_positions

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ws():
    from sphexa_amd.lib import Workspace

    return Workspace(0)


def to_device(d, mixed):
    import torch

    ht = torch.float32 if mixed else torch.float64
    return {k: torch.from_numpy(np.ascontiguousarray(v)).to("cuda").to(torch.float64 if v.dtype == np.float64 and k in
                                                                           abi.REAL_FIELDS else ht)
            for k, v in d.items()}


def ulp_diff(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.max(np.abs(a - b) / np.spacing(np.maximum(np.abs(a), np.abs(b)).astype(a.dtype)))


@pytest.mark.parametrize("mixed", [False, True])
@pytest.mark.parametrize("boundary", [abi.OPEN, abi.PERIODIC, abi.FIXED])
@pytest.mark.parametrize("energy", ["u", "temp", "temp_mui"])
def test_positions_match_oracle(ws, orc, mixed, boundary, energy):
    import torch

    rng = np.random.default_rng(11)
    n = 20000
    hydro = np.float32 if mixed else np.float64
    d = _positions_case(rng, n, hydro, boundary, energy)
    box = abi.Box.make(-0.5, 0.5, boundary)
    params = abi.Params.defaults(1.0, hydro_type=abi.SPHX_F32 if mixed else abi.SPHX_F64)
    dt, dt_m1 = float(np.float32(0.01)), float(np.float32(0.012))
    t = to_device(d, mixed)
    f = abi.fields_from(t)
    hs = abi.History()
    for k in abi.HISTORY_FIELDS:
        if k in t:
            setattr(hs, k, t[k].data_ptr())
    ws.compute_positions(box, f, hs, params, 0, n, dt, dt_m1)
    torch.cuda.synchronize()
    run_positions(orc, d, box, params, n, dt, dt_m1)
    for k in ("x", "y", "z", "vx", "vy", "vz", "x_m1", "y_m1", "z_m1", "du_m1"):
        g = t[k].cpu().numpy()
        assert np.array_equal(g, d[k]), (k, np.max(np.abs(g - d[k])))
    ek = "u" if energy == "u" else "temp"
    g = t[ek].cpu().numpy()
    assert ulp_diff(g, d[ek]) <= 2, ek


@pytest.mark.parametrize("mixed", [False, True])
def test_update_smoothing_length_matches_oracle(ws, orc, mixed):
    import torch

    rng = np.random.default_rng(12)
    n = 50000
    hydro = np.float32 if mixed else np.float64
    d = {"h": rng.uniform(0.005, 0.05, n).astype(hydro), "nc": rng.integers(1, 400, n).astype(np.uint32)}
    params = abi.Params.defaults(1.0, hydro_type=abi.SPHX_F32 if mixed else abi.SPHX_F64)
    th = torch.from_numpy(d["h"].copy()).cuda()
    tn = torch.from_numpy(d["nc"].view(np.int32).copy()).cuda()
    ws.update_smoothing_length(abi.fields_from({"h": th, "nc": tn}), params, 0, n)
    torch.cuda.synchronize()
    orc.update_smoothing_length(C.byref(abi.fields_from(d)), C.byref(params), 0, n)
    assert ulp_diff(th.cpu().numpy(), d["h"]) <= 4


@pytest.mark.parametrize("mixed", [False, True])
def test_timestep_matches_oracle(ws, orc, mixed):
    import torch

    rng = np.random.default_rng(13)
    n = 100003
    hydro = np.float32 if mixed else np.float64
    d = {k: rng.normal(0, 3, n).astype(hydro) for k in ("ax", "ay", "az")}
    d["divv"] = rng.normal(-1, 2, n).astype(hydro)
    params = abi.Params.defaults(1.0, hydro_type=abi.SPHX_F32 if mixed else abi.SPHX_F64)
    t = {k: torch.from_numpy(v).cuda() for k, v in d.items()}
    res = []
    for side, lib, f in (("gpu", None, abi.fields_from(t)), ("orc", orc, abi.fields_from(d))):
        ts = abi.Timestep.defaults(minDt=1e-3, minDt_m1=2e-3, ttot=0.5, minDtCourant=5e-3, g=1.0)
        if lib is None:
            ws.rho_timestep(f, params, 7, n - 5, ts)
            ws.compute_timestep(f, params, 7, n - 5, ts)
        else:
            lib.rho_timestep(C.byref(f), C.byref(params), 7, n - 5, C.byref(ts))
            lib.compute_timestep(C.byref(f), C.byref(params), 7, n - 5, C.byref(ts))
        res.append((ts.minDt, ts.minDt_m1, ts.ttot, ts.minDtRho))
    assert res[0] == res[1]


@pytest.mark.parametrize("bc", [(abi.OPEN,) * 3, (abi.OPEN, abi.PERIODIC, abi.FIXED), (abi.PERIODIC,) * 3])
def test_global_box_matches_oracle(ws, orc, bc):
    """sphx_global_box: the fitting box on the first call, limited shrinking afterwards, periodic
    dimensions untouched; bit for bit against the oracle's globalBox (min and max are exact)."""
    import torch

    rng = np.random.default_rng(15)
    n = 100003
    pts = [rng.normal(0, 1, n), rng.uniform(-3, 2, n), rng.normal(1, 0.2, n)]
    box_g, box_o = abi.Box.make(-4, 4, list(bc)), abi.Box.make(-4, 4, list(bc))
    for call, scale in enumerate((1.0, 0.5, 1.5)):
        p = [np.ascontiguousarray(c * scale) for c in pts]
        t = [torch.from_numpy(c).cuda() for c in p]
        ws.global_box(box_g, *t, n, call == 0)
        orc.global_box(C.byref(box_o), p[0].ctypes.data, p[1].ctypes.data, p[2].ctypes.data, n, int(call == 0))
        for d in range(3):
            assert (box_g.lo[d], box_g.hi[d]) == (box_o.lo[d], box_o.hi[d]), (call, d)
            if bc[d] == abi.PERIODIC:
                assert (box_g.lo[d], box_g.hi[d]) == (-4.0, 4.0)
    e = torch.zeros(1, dtype=torch.float64, device="cuda")
    box_g = abi.Box.make(-1, 1, list(bc))
    ws.global_box(box_g, e, e, e, 0, False)  # no particles: the shrink limit on every refitted side
    for d in range(3):
        assert (box_g.lo[d], box_g.hi[d]) == ((-1.0, 1.0) if bc[d] == abi.PERIODIC else (-0.9, 0.9))


def test_domain_reorder_applies_the_sync_permutation(ws):
    import torch

    arr, spec = ic.sedov_lattice(12, seed=4)
    box = abi.Box.make(spec["lo"], spec["hi"], abi.PERIODIC)
    params = abi.Params.defaults(ws.K)
    ps = ParticleSet.from_host(arr, STD_FIELDS + ("x_m1",))
    n = ps.n
    ps.t["x_m1"][:n] = ps.t["x"][:n].clone()  # the history carries a copy of x ...
    # ... and more arrays of both element sizes, reordered in one batched call
    extra = [ps.t["y"][:n].clone(), ps.t["z"][:n].float(), ps.t["h"][:n].float(), ps.t["m"][:n].clone()]
    ws.domain_sync(box, ps.sync_fields(), params, n, 64)
    ps.reorder_history(ws)
    ws.domain_reorder(extra, n)
    torch.cuda.synchronize()
    assert torch.equal(ps.t["x_m1"][:n], ps.t["x"][:n])  # ... and stays aligned with it
    assert torch.equal(extra[0], ps.t["y"][:n]) and torch.equal(extra[3], ps.t["m"][:n])
    assert torch.equal(extra[1], ps.t["z"][:n].float()) and torch.equal(extra[2], ps.t["h"][:n].float())
    with pytest.raises(Exception):
        ws.domain_reorder(ps.t["x_m1"], n - 1)
    # an array named twice is permuted once; overlapping arrays are refused before any launch
    twice = ps.t["y"][:n].clone()
    once = ps.t["y"][:n].clone()
    ws.domain_reorder([twice, twice], n)
    ws.domain_reorder([once], n)
    torch.cuda.synchronize()
    assert torch.equal(twice, once)
    big = torch.zeros(2 * n, dtype=torch.float64, device="cuda")
    with pytest.raises(Exception):
        ws.domain_reorder([big[:n], big[n // 2: n // 2 + n]], n)


def _oracle_steps(orc, arr, box, params, ts, steps, hydro, refit_box=False, tree_update=False):
    """the oracle's std steps; with tree_update every sync after the first updates the previous tree
    once (Domain::sync) instead of converging a new one"""
    from oracle import pipeline

    from test_integrate_oracle import run_positions as orc_positions

    a = dict(arr)
    n = len(a["x"])
    dts = []
    tree = None
    for step in range(steps):
        if refit_box:
            x, y, z = (np.ascontiguousarray(a[k], np.float64) for k in "xyz")
            orc.global_box(C.byref(box), x.ctypes.data, y.ctypes.data, z.ctypes.data, n, int(step == 0))
        params.minDt = ts.minDt
        o = pipeline.run(orc, a, box, params, form="std", hydro=hydro, prev_tree=tree if tree_update else None)
        tree = o["tree"]
        a = {k: v for k, v in o["a"].items() if v is not None}
        ts.minDtCourant = o["minDt"]
        orc.compute_timestep(C.byref(abi.fields_from(a)), C.byref(params), 0, n, C.byref(ts))
        dt, dt_m1 = float(np.float32(ts.minDt)), float(np.float32(ts.minDt_m1))
        orc_positions(orc, a, box, params, n, dt, dt_m1)
        orc.update_smoothing_length(C.byref(abi.fields_from(a)), C.byref(params), 0, n)
        dts.append(o["minDt"])
    return a, dts


@pytest.mark.parametrize("mixed,boundary", [(False, abi.PERIODIC), (True, abi.PERIODIC), (False, abi.OPEN)])
def test_three_std_steps_match_oracle(ws, orc, mixed, boundary):
    """With open boundaries every sync also refits the box (sphx_global_box against the oracle's
    globalBox), so the keys, the order and the tree follow the reference's moving box."""
    arr, spec = ic.sedov_lattice(16, seed=17)
    n = arr["x"].size
    for k in ("x_m1", "y_m1", "z_m1", "du_m1"):
        arr[k] = np.zeros(n)
    refit = boundary != abi.PERIODIC
    box_o = abi.Box.make(spec["lo"], spec["hi"], boundary)
    steps = 3

    params64 = abi.Params.defaults(ws.K)
    ts_o = abi.Timestep.defaults(minDt=1e-4, minDt_m1=1e-4)
    a, dts_o = _oracle_steps(orc, arr, box_o, params64, ts_o, steps, np.float64, refit_box=refit)

    box = abi.Box.make(spec["lo"], spec["hi"], boundary)
    params = abi.Params.defaults(ws.K, hydro_type=abi.SPHX_F32 if mixed else abi.SPHX_F64)
    ts = abi.Timestep.defaults(minDt=1e-4, minDt_m1=1e-4)
    ps = ParticleSet.from_host(arr, STD_FIELDS + ("x_m1", "y_m1", "z_m1", "du_m1"), mixed=mixed)
    dts = run_steps(ws, ps, box, params, ts, steps, form="std", refit_box=refit)
    if refit:
        # the last step's box, from the same positions on both sides
        assert [box.lo[d] for d in range(3)] == [box_o.lo[d] for d in range(3)]
        assert [box.hi[d] for d in range(3)] == [box_o.hi[d] for d in range(3)]
        assert all(spec["lo"] < box.lo[d] < box.hi[d] < spec["hi"] for d in range(3))  # fitted inside the lattice box
    print(f"{'mixed' if mixed else 'fp64'} Courant steps {dts} oracle {dts_o}; ttot {ts.ttot:.6e} / {ts_o.ttot:.6e}",
          flush=True)
    tol_dt = 1e-12 if not mixed else 1e-5
    np.testing.assert_allclose(dts, dts_o, rtol=tol_dt)
    assert abs(ts.ttot - ts_o.ttot) <= tol_dt * ts_o.ttot

    # both sides hold their particles in SFC order; match them by their x coordinate order
    og = np.argsort(ps.host("x"), kind="stable")
    oo = np.argsort(a["x"], kind="stable")
    scale = {"x": 1.0, "y": 1.0, "z": 1.0}
    for k, tol in (("x", 1e-13), ("y", 1e-13), ("z", 1e-13), ("vx", 1e-9), ("vy", 1e-9), ("vz", 1e-9), ("u", 1e-9),
                   ("h", 1e-12)):
        if mixed:
            # FP32 h can move a borderline pair across 2h, changing nc by one and the next h by
            # about 0.1% (d ln updateH / d nc = -1 / (10 nc))
            tol = {"x": 1e-9, "y": 1e-9, "z": 1e-9, "h": 3e-3}.get(k, 1e-4)
        g = ps.host(k).astype(np.float64)[og]
        o = np.asarray(a[k], np.float64)[oo]
        s = scale.get(k, float(np.max(np.abs(o))))
        err = float(np.max(np.abs(g - o))) / s
        print(f"  {k:3s} max |gpu - oracle| / scale = {err:.2e}", flush=True)
        assert err <= tol, (k, err)


@pytest.mark.parametrize("mixed", [False, True])
def test_positions_explicit_const_cv(ws, mixed):
    """constCv passed in (the reference's computePositionsGpu argument) equal to the value formed
    from muiConst gives the same update bit for bit."""
    import torch

    rng = np.random.default_rng(14)
    n = 5000
    hydro = np.float32 if mixed else np.float64
    d = _positions_case(rng, n, hydro, abi.PERIODIC, "temp")
    box = abi.Box.make(-0.5, 0.5, abi.PERIODIC)
    params = abi.Params.defaults(1.0, hydro_type=abi.SPHX_F32 if mixed else abi.SPHX_F64)
    # idealGasCv(muiConst, gamma) with R and the result in the mass type (eos.hpp:18-23)
    cv = float(hydro(hydro(8.317e7) / hydro(10.0) / (5.0 / 3.0 - 1.0)))
    out = []
    for const_cv in (-1.0, cv):
        t = to_device(d, mixed)
        hs = abi.History()
        for k in abi.HISTORY_FIELDS:
            if k in t:
                setattr(hs, k, t[k].data_ptr())
        ws.compute_positions(box, abi.fields_from(t), hs, params, 0, n, 0.01, 0.012, const_cv)
        torch.cuda.synchronize()
        out.append(t["temp"].cpu().numpy())
    assert np.array_equal(out[0], out[1])


# ---- block time steps ---------------------------------------------------------------------------

def _device_history(t):
    hs = abi.History()
    for k in abi.HISTORY_FIELDS:
        if k in t:
            setattr(hs, k, t[k].data_ptr())
    return hs


@pytest.mark.parametrize("mixed", [False, True])
@pytest.mark.parametrize("energy", ["u", "temp_mui"])
def test_positions_rungs_and_drift_match_oracle(ws, orc, mixed, energy):
    """sphx_compute_positions_rungs (dt_m1 per particle from its rung) and sphx_drift_positions (back
    by dt_back, forward by dt) against the oracle: positions, velocities and the previous-step fields
    bit for bit, energies within 2 ulp (the u < 0 guard's exp)."""
    import torch

    from test_integrate_oracle import RUNG_DT_M1, run_drift, run_positions_rungs

    rng = np.random.default_rng(41)
    n = 20000
    hydro = np.float32 if mixed else np.float64
    params = abi.Params.defaults(1.0, hydro_type=abi.SPHX_F32 if mixed else abi.SPHX_F64)
    box = abi.Box.make(-0.5, 0.5, abi.PERIODIC)
    ek = "u" if energy == "u" else "temp"
    for op in ("positions", "drift"):
        d = _positions_case(rng, n, hydro, abi.PERIODIC, energy)
        rung = rng.integers(0, 4, n).astype(np.uint8)
        t = to_device(d, mixed)
        rung_t = torch.from_numpy(rung).cuda()
        print(f"case {op} mixed={mixed} energy={energy}", flush=True)
        if op == "positions":
            ws.compute_positions_rungs(box, abi.fields_from(t), _device_history(t), params, 0, n, 0.01, RUNG_DT_M1,
                                       rung_t.data_ptr())
            run_positions_rungs(orc, d, box, params, n, 0.01, RUNG_DT_M1, rung)
        else:
            ws.drift_positions(abi.fields_from(t), _device_history(t), params, 0, n, 0.03, 0.0125, RUNG_DT_M1,
                               rung_t.data_ptr())
            run_drift(orc, d, params, n, 0.03, 0.0125, RUNG_DT_M1, rung)
        torch.cuda.synchronize()
        for k in ("x", "y", "z", "vx", "vy", "vz", "x_m1", "y_m1", "z_m1", "du_m1"):
            g = t[k].cpu().numpy()
            assert np.array_equal(g, d[k]), (op, k, np.max(np.abs(g - d[k])))
        assert ulp_diff(t[ek].cpu().numpy(), d[ek]) <= 2, (op, ek)


@pytest.mark.parametrize("mixed", [False, True])
def test_group_timesteps_and_rungs_match_oracle(ws, orc, mixed):
    """sphx_group_divv_timestep and sphx_group_acc_timestep over a group view with groups of 0 to 120
    particles (one wavefront strides over a group larger than 64), sphx_store_rung over those groups
    cut to 64, and all three over the consecutive groups of 64 without a view: groupDt and the rungs
    bit for bit."""
    import torch

    from test_integrate_oracle import _groups_case

    rng = np.random.default_rng(42)
    n = 20000
    hydro = np.float32 if mixed else np.float64
    params = abi.Params.defaults(1.0, hydro_type=abi.SPHX_F32 if mixed else abi.SPHX_F64)
    start, end = _groups_case(rng, n)
    divv = rng.normal(0, 3, n).astype(hydro)
    acc = [rng.normal(0, 5, n).astype(hydro) for _ in range(3)]
    host = {"divv": divv, "ax": acc[0], "ay": acc[1], "az": acc[2]}
    t = {k: torch.from_numpy(v).cuda() for k, v in host.items()}
    f_dev, f_host = abi.fields_from(t), abi.fields_from(host)
    eta = float(np.float32(0.2 * np.sqrt(0.005)))
    for view in (True, False):
        if view:
            s, e = start, end
            s_t = torch.from_numpy(s.astype(np.int32)).cuda()
            e_t = torch.from_numpy(e.astype(np.int32)).cuda()
            g = abi.Groups(firstBody=0, lastBody=n, numGroups=s.size)
            g.groupStart, g.groupEnd = s_t.data_ptr(), e_t.data_ptr()
            ws.set_groups(g)
        else:
            s = np.arange(0, n, 64, dtype=np.uint32)
            e = np.minimum(s + 64, n).astype(np.uint32)
        print(f"case group steps mixed={mixed} view={view} groups={s.size}", flush=True)
        dt0 = rng.uniform(0.01, 0.5, s.size).astype(np.float32)
        for kind in ("divv", "acc"):
            gdt = torch.from_numpy(dt0.copy()).cuda()
            ref = dt0.copy()
            if kind == "divv":
                ws.group_divv_timestep(f_dev, params, 0, n, 0.06, gdt.data_ptr())
                orc.group_divv_timestep(C.byref(f_host), C.byref(params), s.ctypes.data, e.ctypes.data, s.size, 0.06,
                                        ref.ctypes.data)
            else:
                ws.group_acc_timestep(f_dev, params, 0, n, eta, gdt.data_ptr())
                orc.group_acc_timestep(C.byref(f_host), C.byref(params), s.ctypes.data, e.ctypes.data, s.size, eta,
                                       ref.ctypes.data)
            torch.cuda.synchronize()
            np.testing.assert_array_equal(gdt.cpu().numpy(), ref)
        # the rung store, like the elementwise kernels, takes groups of at most 64 particles (one
        # wavefront each, storeRungKernel: one lane per particle)
        e64 = np.minimum(e, s + 64).astype(np.uint32)
        if view:
            e64_t = torch.from_numpy(e64.astype(np.int32)).cuda()
            g.groupEnd = e64_t.data_ptr()
            ws.set_groups(g)
        rungs = torch.full((n,), 9, dtype=torch.uint8, device="cuda")
        ws.store_rung(0, n, 3, rungs.data_ptr())
        ref = np.full(n, 9, np.uint8)
        orc.store_rung(s.ctypes.data, e64.ctypes.data, s.size, 3, ref.ctypes.data)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(rungs.cpu().numpy(), ref)
        ws.set_groups(None)


def test_block_timestep_arguments_are_checked(ws):
    """numRungs outside [1, 4] and a missing groupDt are refused before any launch"""
    import torch

    n = 128
    t = {k: torch.zeros(n, dtype=torch.float64, device="cuda") for k in ("x", "y", "z", "vx", "vy", "vz", "ax", "ay",
                                                                         "az", "x_m1", "y_m1", "z_m1", "divv")}
    params = abi.Params.defaults(1.0)
    with pytest.raises(Exception, match="numRungs"):
        ws.drift_positions(abi.fields_from(t), _device_history(t), params, 0, n, 0.1, 0.0, [0.1] * 5, None)
    with pytest.raises(Exception, match="groupDt"):
        ws.group_divv_timestep(abi.fields_from(t), params, 0, n, 0.06, None)
