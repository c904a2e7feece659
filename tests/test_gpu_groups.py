"""Target groups (cstone::GroupView) on the GPU: spatial groups against the oracle, passes over them,
and pair loops over a subset of groups with per-group time steps (the block time-step propagators'
use of GroupView).

* sphx_compute_groups equals the oracle's computeGroupSplits restatement, entry for entry.
* A std and a VE pass with the spatial groups set give the same neighbour counts, lists and fields,
  bit for bit, as with consecutive groups of 64: a target's result does not depend on its grouping.
* Momentum-energy over every other group updates exactly those groups' targets (the same bits as the
  full pass) and lowers groupDt[g] to the group's smallest Courant step, which equals the minimum over
  the group of the oracle's per-target steps (tsKCourant from its maximum signal velocities).
* The position update over a subset of groups moves only those targets.
"""
import ctypes as C

import numpy as np
import pytest

from sphexa_amd import abi, ic
from sphexa_amd.particles import STD_FIELDS, VE_FIELDS, ParticleSet, ideal_gas_cv, std_step, ve_step

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ws():
    from sphexa_amd.lib import Workspace

    return Workspace(0)


def case(name):
    if name == "sedov16":
        arr, spec = ic.sedov_lattice(16, seed=7)
    elif name == "evrard":
        arr, spec = ic.evrard_sphere(30000, seed=3)
    else:
        arr, spec = ic.turbulence_box(18, seed=2)
        arr["temp"] = arr.pop("u") / ideal_gas_cv(10.0, 5.0 / 3.0)
    box = abi.Box.make(spec["lo"], spec["hi"], abi.PERIODIC if spec["periodic"] else abi.OPEN)
    return arr, box


def device_groups(g):
    import torch

    starts = np.zeros(g.numGroups + 1, np.uint32)
    hip = C.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    torch.cuda.synchronize()
    assert hip.hipMemcpy(starts.ctypes.data, g.groupStart, starts.nbytes, 2) == 0
    return starts


@pytest.mark.parametrize("name", ["sedov16", "evrard", "turb"])
def test_compute_groups_matches_oracle(ws, orc, name):
    from oracle import oracle as O

    arr, box = case(name)
    ps = ParticleSet.from_host(arr, STD_FIELDS)
    params = abi.Params.defaults(ws.K)
    ws.domain_sync(box, ps.sync_fields(), params, ps.n)
    x, y, z = (ps.host(c) for c in "xyz")
    tree = O.Tree(orc, ps.host("keys"), box)
    for first, last in ((0, ps.n), (37, ps.n - 101)):
        g = ws.compute_groups(box, ps.fields(), first, last)
        ref = O.group_splits(orc, tree, box, x, y, z, first, last)
        got = device_groups(g)
        assert (g.firstBody, g.lastBody) == (first, last)
        np.testing.assert_array_equal(got, ref)
        print(f"{name} [{first}, {last}): {g.numGroups} groups, {(last - first + 63) // 64} fixed", flush=True)


def run_pass(ws, form, arr, box, mixed, grouped):
    import torch

    names = STD_FIELDS if form == "std" else VE_FIELDS
    ps = ParticleSet.from_host(arr, names, mixed=mixed)
    params = abi.Params.defaults(ws.K, hydro_type=abi.SPHX_F32 if mixed else abi.SPHX_F64, minDt=1e-4)
    ws.set_groups(None)
    ws.domain_sync(box, ps.sync_fields(), params, ps.n)
    if grouped:
        g = ws.compute_groups(box, ps.fields(), 0, ps.n)
        ws.set_groups(g)
    f = ps.fields()
    conv = ws.find_neighbors(box, f, params, 0, ps.n, True)
    if form == "std":
        ws.pair("sphx_density_std", box, f, params, 0, ps.n)
        ws.eos("sphx_eos_std", f, params, 0, ps.n)
        ws.pair("sphx_iad_std", box, f, params, 0, ps.n)
        dt = ws.momentum_energy_std(box, f, params, 0, ps.n)
    else:
        ws.pair("sphx_xmass", box, f, params, 0, ps.n)
        ws.pair("sphx_ve_def_gradh", box, f, params, 0, ps.n)
        ws.eos("sphx_eos_ve", f, params, 0, ps.n)
        ws.pair("sphx_iad_divv_curlv", box, f, params, 0, ps.n)
        ws.pair("sphx_av_switches", box, f, params, 0, ps.n)
        dt = ws.momentum_energy_ve(box, f, params, 0, ps.n)
    lists = torch.zeros(ps.n * params.ngmax, dtype=torch.int32, device="cuda")
    ws.copy_neighbors(lists, 0, ps.n)
    ws.set_groups(None)
    return conv, dt, ps, lists.cpu().numpy()


@pytest.mark.parametrize("form", ["std", "ve"])
@pytest.mark.parametrize("mixed", [False, True])
def test_pass_over_spatial_groups_is_bitwise_the_same(ws, form, mixed):
    arr, box = case("evrard")
    if form == "ve":
        arr["temp"] = arr.pop("u") / ideal_gas_cv(10.0, 5.0 / 3.0)
    c1, dt1, p1, l1 = run_pass(ws, form, arr, box, mixed, grouped=False)
    c2, dt2, p2, l2 = run_pass(ws, form, arr, box, mixed, grouped=True)
    assert c1 == c2 and dt1 == dt2
    nc = p1.host("nc")
    np.testing.assert_array_equal(nc, p2.host("nc"))
    stored = np.minimum(nc.astype(np.int64) - 1, 150)
    used = (np.arange(150)[None, :] < stored[:, None]).ravel()
    assert np.array_equal(l1[used], l2[used])
    outs = ("rho", "c11", "ax", "ay", "az", "du") if form == "std" else ("kx", "gradh", "prho", "divv", "alpha",
                                                                          "ax", "az", "du")
    for k in outs + ("h",):
        assert np.array_equal(p1.host(k), p2.host(k)), k


def test_subset_groups_and_group_dt(ws, orc):
    """VE momentum-energy over every other spatial group: only those targets change, with the full
    pass's bits, and groupDt[g] = min(initial, the group's smallest Courant step) matches the oracle's
    per-target steps."""
    import torch

    from oracle import pipeline

    arr, box = case("turb")
    params = abi.Params.defaults(ws.K, minDt=1e-4)
    o = pipeline.run(orc, arr, box, params, form="ve")
    maxvs = np.zeros(len(arr["x"]))
    from oracle import oracle as O

    wh, whd = O.tables(orc, 6.0)
    cp = O.CpuPass(orc, box, o["a"], params, o["nbrs"], 0, len(arr["x"]), wh, whd)
    cp.momentum_energy_ve(False, maxvs)

    ps = ParticleSet.from_host(arr, VE_FIELDS)
    conv, dt_full = ve_step(ws, ps, box, params, chain=False)
    assert conv
    full = {k: ps.host(k).copy() for k in ("ax", "ay", "az", "du")}
    g = ws.compute_groups(box, ps.fields(), 0, ps.n)
    starts = torch.from_numpy(device_groups(g).astype(np.int64)).cuda().to(torch.int32)
    sel = np.arange(0, g.numGroups, 2)
    sub_start = starts[torch.from_numpy(sel).cuda()].contiguous()
    sub_end = starts[torch.from_numpy(sel + 1).cuda()].contiguous()
    group_dt = torch.full((len(sel),), 1e10, dtype=torch.float32, device="cuda")
    view = abi.Groups(firstBody=0, lastBody=ps.n, numGroups=len(sel))
    view.groupStart, view.groupEnd, view.groupDt = sub_start.data_ptr(), sub_end.data_ptr(), group_dt.data_ptr()
    for k in ("ax", "ay", "az"):
        ps.t[k].fill_(7.0)
    ps.t["du"].fill_(7.0)
    ws.set_groups(view)
    dt_sub = ws.momentum_energy_ve(box, ps.fields(), params, 0, ps.n)
    ws.set_groups(None)
    s = sub_start.cpu().numpy().astype(np.int64)
    e = sub_end.cpu().numpy().astype(np.int64)
    chosen = np.zeros(ps.n, bool)
    for a, b in zip(s, e):
        chosen[a:b] = True
    for k in ("ax", "ay", "az", "du"):
        got = ps.host(k)
        assert np.array_equal(got[chosen], full[k][chosen]), k
        assert np.all(got[~chosen] == 7.0), k
    # the oracle's per-target step, tsKCourant (sph/include/sph/kernels.hpp:11-16), FP64 form
    h, c = o["a"]["h"], o["a"]["c"]
    v = np.where(maxvs > 0, maxvs, c)
    dt_i = np.float32(params.Kcour) * h / v
    gdt = group_dt.cpu().numpy()
    ref = np.array([np.float32(min(1e10, dt_i[a:b].min())) for a, b in zip(s, e)], np.float32)
    np.testing.assert_allclose(gdt, ref, rtol=1e-6)
    assert np.float32(dt_sub) == gdt.min()
    assert dt_sub >= dt_full


def test_positions_over_subset_groups(ws):
    import torch

    from sphexa_amd.particles import integrate

    arr, box = case("sedov16")
    arr["vx"] = 0.1 * np.sin(6.0 * arr["x"])
    params = abi.Params.defaults(ws.K)
    ps = ParticleSet.from_host(arr, STD_FIELDS + ("x_m1", "y_m1", "z_m1", "du_m1"))
    conv, dtc = std_step(ws, ps, box, params, chain=False)
    before = {k: ps.host(k).copy() for k in ("x", "y", "z", "u")}
    starts = torch.arange(0, ps.n + 1, 64, dtype=torch.int32, device="cuda")
    starts = torch.cat([starts, torch.tensor([ps.n], dtype=torch.int32, device="cuda")]) if starts[-1] != ps.n else starts
    sel = torch.arange(1, starts.numel() - 1, 3, device="cuda")
    view = abi.Groups(firstBody=0, lastBody=ps.n, numGroups=sel.numel())
    gs, ge = starts[sel].contiguous(), starts[sel + 1].contiguous()
    view.groupStart, view.groupEnd = gs.data_ptr(), ge.data_ptr()
    ts = abi.Timestep.defaults(minDt=1e-4, minDt_m1=1e-4, minDtCourant=dtc)
    ws.set_groups(view)
    integrate(ws, ps, box, params, ts)
    ws.set_groups(None)
    chosen = np.zeros(ps.n, bool)
    for a, b in zip(gs.cpu().numpy(), ge.cpu().numpy()):
        chosen[a:b] = True
    moved = ps.host("x") != before["x"]
    assert moved[chosen].mean() > 0.9 and not moved[~chosen].any()
