"""H5Part-layout I/O with device fields (include/sphexa_amd_io.h): fields written straight from and
read straight into device memory, checked against the host values and against the reference's H5Part
reading the same file; and a checkpoint restart: the std propagator's conserved fields and time-step
state written after two steps (what the reference's checkpoint output holds, README "Restarting from
checkpoint files": the conserved fields with x_m1, y_m1, z_m1, du_m1), loaded into a fresh particle
set, and the third step from the checkpoint equal bit for bit to the third step of the run that wrote
it."""
import os

import numpy as np
import pytest

from sphexa_amd import abi, h5io, ic
from sphexa_amd.particles import STD_FIELDS, ParticleSet, run_steps

import h5part_ref as R

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(h5io.LIB_PATH), reason="libsphexa_amd_io.so not built (no HDF5)")]

HIST = ("x_m1", "y_m1", "z_m1", "du_m1")
CONSERVED = ("x", "y", "z", "h", "m", "u", "vx", "vy", "vz") + HIST


@pytest.fixture(scope="module")
def ws():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a GPU"
    from sphexa_amd.lib import Workspace

    return Workspace(0)


@pytest.mark.parametrize("mixed", [False, True])
def test_device_fields_round_trip(tmp_path, mixed):
    import torch

    arr, spec = ic.sedov_lattice(16, seed=5)
    ps = ParticleSet.from_host(arr, STD_FIELDS, mixed=mixed)
    n = ps.n
    names = [k for k in STD_FIELDS if k in ps.t]
    path = tmp_path / "dev.h5"
    print(f"case device round trip mixed={mixed} n={n}", flush=True)
    with h5io.H5File(path, h5io.WRITE) as f:
        f.add_step(n)
        for k in names:
            f.write_field(k, ps.t[k][:n])
    with h5io.H5File(path) as f:
        assert f.set_step(0) == n
        for k in names:
            out = torch.empty_like(ps.t[k][:n])
            f.read_field(k, out=out)
            torch.cuda.synchronize()
            assert torch.equal(out, ps.t[k][:n]), k
    if R.available():
        ref = R.RefH5Part(path, R.H5PART_READ)
        ref.set_step(0)
        for k in names:
            host = ps.host(k)
            if host.dtype.kind == "u":  # H5Part stores unsigned as the signed type of its size
                host = host.view(np.dtype(f"i{host.dtype.itemsize}"))
            assert np.array_equal(ref.read(k, host.dtype), host), k
        ref.close()


@pytest.mark.parametrize("mixed", [False, True])
def test_restart_from_checkpoint_is_bitwise(tmp_path, ws, mixed):
    import torch

    arr, spec = ic.sedov_lattice(16, seed=9)
    n = arr["x"].size
    arr["vx"] = 0.1 * np.sin(6.0 * arr["x"])
    for k in HIST:
        arr[k] = np.zeros(n)
    box = abi.Box.make(spec["lo"], spec["hi"], abi.PERIODIC)
    params = abi.Params.defaults(ws.K, hydro_type=abi.SPHX_F32 if mixed else abi.SPHX_F64)
    ts = abi.Timestep.defaults(minDt=1e-4, minDt_m1=1e-4)
    ps = ParticleSet.from_host(arr, STD_FIELDS + HIST, mixed=mixed)
    print(f"case restart mixed={mixed} n={n}", flush=True)
    run_steps(ws, ps, box, params, ts, 2, form="std")
    path = tmp_path / "ckpt.h5"
    with h5io.H5File(path, h5io.WRITE) as f:
        f.add_step(n)
        for k, v in (("time", ts.ttot), ("minDt", ts.minDt), ("minDt_m1", ts.minDt_m1), ("iteration", np.int64(3)),
                     ("numParticlesGlobal", np.int64(n)), ("gravConstant", 0.0)):
            f.write_attrib(k, v)
        for k in CONSERVED:
            f.write_field(k, ps.t[k][:n])
    torch.cuda.synchronize()
    run_steps(ws, ps, box, params, ts, 1, form="std", first_sync=False)
    cont = {k: ps.host(k).copy() for k in CONSERVED}

    with h5io.H5File(path) as f:
        assert f.set_step(-1) == n
        ts2 = abi.Timestep.defaults(minDt=float(f.read_attrib("minDt")[0]), minDt_m1=float(f.read_attrib("minDt_m1")[0]))
        ts2.ttot = float(f.read_attrib("time")[0])
        ps2 = ParticleSet(n, STD_FIELDS + HIST, mixed, "cuda")
        for k in CONSERVED:
            f.read_field(k, out=ps2.t[k][:n])
    torch.cuda.synchronize()
    run_steps(ws, ps2, box, params, ts2, 1, form="std", first_sync=False)
    assert ts2.minDt == ts.minDt and ts2.ttot == ts.ttot
    for k in CONSERVED:
        assert np.array_equal(ps2.host(k), cont[k]), k


def test_device_fields_larger_than_a_staging_chunk(tmp_path):
    """device fields of several 16 MiB staging chunks, with a ragged last chunk (3,000,001 doubles are
    two chunks, 9,000,003 floats three): written from and read back into device memory bit for bit,
    and the same values through a host read"""
    import torch

    n = 3_000_001
    g = torch.Generator(device="cuda").manual_seed(11)
    x = torch.rand(n, device="cuda", dtype=torch.float64, generator=g)
    big = torch.rand(9_000_003, device="cuda", dtype=torch.float32, generator=g)
    path = tmp_path / "chunks.h5"
    print(f"case chunked device fields n={n} and {big.numel()}", flush=True)
    with h5io.H5File(path, h5io.WRITE) as f:
        f.add_step(n)
        f.write_field("x", x)
        f.add_step(big.numel())
        f.write_field("h", big)
    with h5io.H5File(path) as f:
        assert f.set_step(0) == n
        back = torch.empty_like(x)
        f.read_field("x", out=back)
        assert torch.equal(back, x)
        assert np.array_equal(f.read_field("x"), x.cpu().numpy())
        assert f.set_step(1) == big.numel()
        back32 = torch.empty_like(big)
        f.read_field("h", out=back32)
        assert torch.equal(back32, big)
