"""The tree update of the syncs after a Domain's first (SPHX_FLAG_TREE_UPDATE) on the GPU, against the
oracle's updateNsTree (pinned against the reference's focused octree in test_tree_update_oracle.py).

* A sequence of syncs of a collapsing, rotating particle set: after every sync the sorted keys and the
  whole tree (leaves, counts, layout, prefixes, child offsets, parents, internalToLeaf, level ranges,
  centres, sizes) are bit-identical to the oracle's, and the tree changes at most syncs.
* Five full FP64 std steps (forces, time step, positions, h) of a converging flow with the tree updated
  at every sync after the first, as the reference's Domain does: per step the leaves, the neighbour
  counts and every neighbour list entry in order equal the oracle's chain, and the Courant steps agree.
"""
import numpy as np
import pytest

from sphexa_amd import abi, ic
from sphexa_amd.particles import STD_FIELDS, ParticleSet, run_steps, sync_params

from test_gpu_parity import d2h

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ws():
    from sphexa_amd.lib import Workspace

    return Workspace(0)


def collapse(pts, step):
    r2 = (pts ** 2).sum(0)
    pts = pts * (1 - 0.25 * np.exp(-r2 / 0.05))[None, :]
    c, s = np.cos(0.3 + 0.1 * step), np.sin(0.3 + 0.1 * step)
    pts = np.stack([pts[0] * c - pts[1] * s, pts[0] * s + pts[1] * c, pts[2]])
    return (pts + 0.5) % 1.0 - 0.5


def gpu_tree(ws):
    tv = ws.tree_view()
    N, L = tv.numNodes, tv.numLeafNodes
    return dict(leaves=d2h(tv.leaves, L + 1, np.uint64), counts=d2h(tv.counts, L, np.uint32),
                layout=d2h(tv.layout, L + 1, np.uint32), prefixes=d2h(tv.prefixes, N, np.uint64),
                childOffsets=d2h(tv.childOffsets, N, np.int32), internalToLeaf=d2h(tv.internalToLeaf, N, np.int32),
                parents=d2h(tv.parents, max(1, (N - 1) // 8), np.int32), levelRange=d2h(tv.levelRange, 23, np.int32),
                centers=d2h(tv.centers, 3 * N, np.float64).reshape(N, 3),
                sizes=d2h(tv.sizes, 3 * N, np.float64).reshape(N, 3))


def assert_same_tree(g, t, tag):
    assert len(g["leaves"]) == len(t.leaves), (tag, len(g["leaves"]), len(t.leaves))
    for k in ("leaves", "counts", "layout", "prefixes", "childOffsets", "parents", "levelRange", "centers", "sizes"):
        np.testing.assert_array_equal(g[k], getattr(t, k), err_msg=f"{tag}: {k}")
    leaf = t.internalToLeaf >= 0
    np.testing.assert_array_equal(g["internalToLeaf"][leaf], t.internalToLeaf[leaf], err_msg=tag)
    assert np.all(g["internalToLeaf"][~leaf] < 0), tag


def test_sync_tree_update_matches_oracle(ws, orc):
    from oracle import oracle as O

    rng = np.random.default_rng(21)
    n = 40000
    pts = rng.uniform(-0.5, 0.5, (3, n))
    box = abi.Box.make([-0.5] * 3, [0.5] * 3, abi.PERIODIC)
    params = abi.Params.defaults(ws.K)
    t = None
    changes = 0
    for step in range(8):
        arr = {"x": pts[0].copy(), "y": pts[1].copy(), "z": pts[2].copy(), "h": np.full(n, 0.01), "m": np.ones(n)}
        ps = ParticleSet.from_host(arr, ("x", "y", "z", "h", "m", "keys"))
        ws.domain_sync(box, ps.sync_fields(), sync_params(params, step > 0), n, 64)
        keys = np.sort(O.sfc_keys(orc, box, *[np.ascontiguousarray(c) for c in pts]))
        np.testing.assert_array_equal(ps.host("keys"), keys)
        if t is None:
            t = O.Tree(orc, keys, box, 64)
        else:
            before = t.leaves.copy()
            t.update(keys, box, 64)
            changes += int(not np.array_equal(before, t.leaves))
        g = gpu_tree(ws)
        assert_same_tree(g, t, f"sync {step}")
        print(f"sync {step}: {t.numLeafNodes} leaves, largest leaf {int(t.counts.max())}", flush=True)
        pts = collapse(pts, step)
    assert changes >= 5


def test_sync_without_flag_converges(ws, orc):
    """without SPHX_FLAG_TREE_UPDATE (or after sphx_build_tree) a sync converges the tree anew"""
    from oracle import oracle as O

    rng = np.random.default_rng(22)
    n = 20000
    box = abi.Box.make([-0.5] * 3, [0.5] * 3, abi.PERIODIC)
    params = abi.Params.defaults(ws.K)
    pts = rng.uniform(-0.5, 0.5, (3, n))
    for step, flag in enumerate((False, False, True)):
        arr = {"x": pts[0].copy(), "y": pts[1].copy(), "z": pts[2].copy(), "h": np.full(n, 0.01), "m": np.ones(n)}
        ps = ParticleSet.from_host(arr, ("x", "y", "z", "h", "m", "keys"))
        if step == 2:
            ws.build_tree(ps.t["keys"][:n], box, 64)  # a direct build: the next sync has no previous sync tree
        ws.domain_sync(box, ps.sync_fields(), sync_params(params, flag), n, 64)
        keys = np.sort(O.sfc_keys(orc, box, *[np.ascontiguousarray(c) for c in pts]))
        assert_same_tree(gpu_tree(ws), O.Tree(orc, keys, box, 64), f"sync {step}")
        pts = collapse(pts, step) * 0.7


def _oracle_steps_updated(orc, arr, box, params, ts, steps, record):
    import ctypes as C

    from oracle import pipeline
    from test_integrate_oracle import run_positions as orc_positions

    a = dict(arr)
    n = len(a["x"])
    tree = None
    dts = []
    for step in range(steps):
        params.minDt = ts.minDt
        o = pipeline.run(orc, a, box, params, form="std", prev_tree=tree)
        tree = o["tree"]
        record.append((tree.leaves.copy(), o["a"]["nc"].copy(), np.asarray(o["nbrs"]).copy()))
        a = {k: v for k, v in o["a"].items() if v is not None}
        ts.minDtCourant = o["minDt"]
        orc.compute_timestep(C.byref(abi.fields_from(a)), C.byref(params), 0, n, C.byref(ts))
        dt, dt_m1 = float(np.float32(ts.minDt)), float(np.float32(ts.minDt_m1))
        orc_positions(orc, a, box, params, n, dt, dt_m1)
        orc.update_smoothing_length(C.byref(abi.fields_from(a)), C.byref(params), 0, n)
        dts.append(o["minDt"])
    return a, dts


def test_std_steps_with_tree_update_match_oracle(ws, orc):
    import torch

    # 31^3 particles placed at random: about 58 per level-3 leaf, so that many leaves and sibling
    # groups sit close to the split and merge thresholds of bucket size 64
    arr, spec = ic.sedov_lattice(31, seed=23)
    n = arr["x"].size
    rng = np.random.default_rng(5)
    for k in ("x", "y", "z"):
        arr[k] = rng.uniform(-0.5, 0.5, n)
    # a converging, rotating flow in cold gas: large displacements per Courant step
    r = np.sqrt(arr["x"] ** 2 + arr["y"] ** 2 + arr["z"] ** 2)
    arr["vx"] = -2.0 * arr["x"] * np.exp(-r * r / 0.08) - 0.5 * arr["y"]
    arr["vy"] = -2.0 * arr["y"] * np.exp(-r * r / 0.08) + 0.5 * arr["x"]
    arr["vz"] = -2.0 * arr["z"] * np.exp(-r * r / 0.08)
    arr["u"] = np.full(n, 1e-4)
    dt0 = 1e-2
    # the position update moves particles by x_m1 / dt_m1 (the previous step's displacement), so the
    # initial displacements carry the velocity field, as the reference's initial conditions set them
    for k in ("x", "y", "z"):
        arr[k + "_m1"] = arr["v" + k] * dt0
    arr["du_m1"] = np.zeros(n)
    box_o = abi.Box.make(spec["lo"], spec["hi"], abi.PERIODIC)
    steps = 5
    params64 = abi.Params.defaults(ws.K)
    ts_o = abi.Timestep.defaults(minDt=dt0, minDt_m1=dt0)
    rec_o = []
    _, dts_o = _oracle_steps_updated(orc, arr, box_o, params64, ts_o, steps, rec_o)

    box = abi.Box.make(spec["lo"], spec["hi"], abi.PERIODIC)
    params = abi.Params.defaults(ws.K)
    ts = abi.Timestep.defaults(minDt=dt0, minDt_m1=dt0)
    ps = ParticleSet.from_host(arr, STD_FIELDS + ("x_m1", "y_m1", "z_m1", "du_m1"))
    ngmax = params.ngmax
    changes = 0

    def check(step):
        nonlocal changes
        leaves_o, nc_o, nb_o = rec_o[step]
        g = gpu_tree(ws)
        np.testing.assert_array_equal(g["leaves"], leaves_o, err_msg=f"step {step}")
        if step > 0:
            changes += int(not np.array_equal(rec_o[step - 1][0], leaves_o))
        nc = ps.host("nc")
        np.testing.assert_array_equal(nc, nc_o, err_msg=f"step {step}")
        out = torch.zeros(n * ngmax, dtype=torch.int32, device="cuda")
        ws.copy_neighbors(out, 0, n)
        nb = out.cpu().numpy().view(np.uint32).reshape(n, ngmax)
        stored = np.minimum(nc.astype(np.int64) - 1, ngmax)
        bad = np.nonzero((nb != nb_o) & (np.arange(ngmax)[None, :] < stored[:, None]))
        assert bad[0].size == 0, f"step {step}: list mismatch at target {bad[0][0]}"
        print(f"step {step}: {len(leaves_o) - 1} leaves, largest {int(g['counts'].max())}, lists identical",
              flush=True)

    dts = run_steps(ws, ps, box, params, ts, steps, form="std", tree_update=True, on_step=check)
    np.testing.assert_allclose(dts, dts_o, rtol=1e-12)
    assert changes >= 2, "the flow should change the tree during the run"
