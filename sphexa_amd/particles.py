"""Device particle set named after sphexa::ParticlesData, and the hydro step drivers.

ParticleSet holds one torch tensor per field on the GPU (SoA, SFC order after the domain
sync): x, y, z, u, temp, du in float64 (RealType); the HydroType fields in float64 for the
plain form or float32 for the mixed form (the reference's SphTypes layout,
sph/include/sph/types.hpp:39-46); keys uint64; nc uint32.

The drivers issue the phases in the order of the reference propagators:
  std: HydroProp::computeForces   (main/src/propagator/std_hydro.hpp:124-168)
  VE:  HydroVeProp::computeForces (main/src/propagator/ve_hydro.hpp:130-204)
with the one-rank Domain::sync in front and no halo exchanges.
"""
import numpy as np

from . import abi

STD_FIELDS = ("x", "y", "z", "u", "du", "h", "m", "vx", "vy", "vz", "rho", "p", "c", "c11", "c12", "c13", "c22",
              "c23", "c33", "ax", "ay", "az", "keys", "nc")
VE_FIELDS = ("x", "y", "z", "temp", "du", "h", "m", "vx", "vy", "vz", "prho", "c", "c11", "c12", "c13", "c22",
             "c23", "c33", "xm", "kx", "gradh", "divv", "curlv", "alpha", "ax", "ay", "az", "keys", "nc")
GRADV_FIELDS = ("dV11", "dV12", "dV13", "dV22", "dV23", "dV33")
# the fields a domain sync permutes: HydroProp / HydroVeProp::ConservedFields with x, y, z, h, m
# (main/src/propagator/std_hydro.hpp:73, ve_hydro.hpp:72); every other field is recomputed by the
# pass before it is read. keys receives the sorted SFC keys.
CONSERVED_FIELDS = ("x", "y", "z", "h", "m", "u", "temp", "vx", "vy", "vz", "alpha", "keys")

IDEAL_GAS_R = 8.317e7


def ideal_gas_cv(mui, gamma):
    """idealGasCv (sph/include/sph/eos.hpp:19-23)."""
    return IDEAL_GAS_R / mui / (gamma - 1.0)


class ParticleSet:
    def __init__(self, n, names, mixed=False, device="cuda"):
        import torch

        self.n = n
        self.mixed = mixed
        self.device = device
        self.hydro_dtype = torch.float32 if mixed else torch.float64
        self.t = {}
        for name in names:
            if name in abi.REAL_FIELDS:
                dt = torch.float64
            elif name == "keys":
                dt = torch.int64  # uint64 bit pattern
            elif name == "nc":
                dt = torch.int32  # uint32 bit pattern
            else:
                dt = self.hydro_dtype
            self.t[name] = torch.zeros(max(n, 1), dtype=dt, device=device)

    @classmethod
    def from_host(cls, arrays, names, mixed=False, device="cuda"):
        import torch

        n = len(arrays["x"])
        ps = cls(n, names, mixed, device)
        for name, a in arrays.items():
            if name in ps.t and a is not None:
                ps.t[name][:n] = torch.from_numpy(np.ascontiguousarray(a)).to(ps.t[name].dtype).to(device)
        return ps

    @property
    def hydro_type(self):
        return abi.SPHX_F32 if self.mixed else abi.SPHX_F64

    def fields(self, exclude=()):
        return abi.fields_from({k: v for k, v in self.t.items() if k not in exclude})

    def history(self):
        """abi.History of the set's previous-step tensors (x_m1, y_m1, z_m1, du_m1, mui; absent ones NULL)."""
        h = abi.History()
        for k in abi.HISTORY_FIELDS:
            if k in self.t:
                setattr(h, k, self.t[k].data_ptr())
        return h

    def sync_fields(self):
        """abi.Fields of the conserved tensors only, for sphx_domain_sync."""
        return abi.fields_from({k: v for k, v in self.t.items() if k in CONSERVED_FIELDS})

    def reorder_history(self, ws):
        """Apply the last domain sync's order to the previous-step tensors, as Domain::sync reorders
        every conserved field (one batched call)."""
        ts = [self.t[k] for k in abi.HISTORY_FIELDS if k in self.t]
        if ts:
            ws.domain_reorder(ts, self.n)

    def host(self, name):
        a = self.t[name][: self.n].cpu().numpy()
        if name == "keys":
            return a.view(np.uint64)
        if name == "nc":
            return a.view(np.uint32)
        return a


def search_first_sum(ws, box, f, params, n, rho):
    """Neighbour search with the h/nc iteration, then the first density sum: rho (std,
    computeDensity) or xm (VE, computeXMass), as two kernels. sphx_find_neighbors_xmass runs
    both in one kernel, as the reference GPU computeXMass does
    (sph/include/sph/hydro_ve/xmass_gpu.cu:57-131), and gives the same lists, counts and h; it
    is slower here in both forms, because the search kernel holds fewer waves per SIMD than the
    standalone pair loop and evaluates its hits one lane at a time."""
    converged = ws.find_neighbors(box, f, params, 0, n, True)
    ws.pair("sphx_density_std" if rho else "sphx_xmass", box, f, params, 0, n)
    return converged


def sync_params(params, tree_update):
    """params for sphx_domain_sync: with tree_update, SPHX_FLAG_TREE_UPDATE (update the previous
    sync's tree once, as every Domain::sync after the first does)"""
    if not tree_update:
        return params
    p = abi.Params.from_buffer_copy(params)
    p.flags |= abi.SPHX_FLAG_TREE_UPDATE
    return p


def std_step(ws, ps, box, params, bucket=64, chain=True, tree_update=False):
    """One std-SPH pass; returns (converged, minDtCourant). With chain, the pair passes run with
    SPHX_FLAG_CHAIN (bitwise the same); with tree_update, the sync updates the previous sync's tree
    once instead of converging a new one."""
    n = ps.n
    f = ps.fields()
    cp = abi.chained(params) if chain else params
    ws.domain_sync(box, ps.sync_fields(), sync_params(params, tree_update), n, bucket)
    ps.reorder_history(ws)
    converged = search_first_sum(ws, box, f, cp, n, rho=True)
    ws.eos("sphx_eos_std", f, params, 0, n)
    ws.pair("sphx_iad_std", box, f, cp, 0, n)
    dt = ws.momentum_energy_std(box, f, cp, 0, n)
    return converged, dt


def ve_step(ws, ps, box, params, bucket=64, av_clean=False, ts=None, chain=True, tree_update=False):
    """One VE pass; returns (converged, minDtCourant). With a time-step state ts (abi.Timestep) it
    also sets ts.minDtRho from div v, as HydroVeProp::computeForces does (ve_hydro.hpp:171).
    With chain, the pair passes run with SPHX_FLAG_CHAIN (bitwise the same); with tree_update, the
    sync updates the previous sync's tree once."""
    n = ps.n
    f = ps.fields()
    cp = abi.chained(params) if chain else params
    ws.domain_sync(box, ps.sync_fields(), sync_params(params, tree_update), n, bucket)
    ps.reorder_history(ws)
    converged = search_first_sum(ws, box, f, cp, n, rho=False)
    ws.pair("sphx_ve_def_gradh", box, f, cp, 0, n)
    ws.eos("sphx_eos_ve", f, params, 0, n)
    ws.pair("sphx_iad_divv_curlv", box, f, cp, 0, n)
    if ts is not None:
        ws.rho_timestep(f, params, 0, n, ts)
    ws.pair("sphx_av_switches", box, f, cp, 0, n)
    dt = ws.momentum_energy_ve(box, f, cp, 0, n, av_clean)
    return converged, dt


def integrate(ws, ps, box, params, ts):
    """The integrate phase of HydroProp / HydroVeProp (main/src/propagator/std_hydro.hpp:170-181,
    ve_hydro.hpp:206-217): computeTimestep, computePositions with the step and the previous step
    passed as float like the reference's GPU path, updateSmoothingLength. ts (abi.Timestep) is
    advanced in place; ts.minDtCourant must hold the last momentum-energy pass's value."""
    f = ps.fields()
    ws.compute_timestep(f, params, 0, ps.n, ts)
    dt, dt_m1 = float(np.float32(ts.minDt)), float(np.float32(ts.minDt_m1))
    ws.compute_positions(box, f, ps.history(), params, 0, ps.n, dt, dt_m1)
    ws.update_smoothing_length(f, params, 0, ps.n)


def run_steps(ws, ps, box, params, ts, steps, form="std", av_clean=False, refit_box=False, first_sync=True,
              tree_update=False, on_step=None):
    """`steps` full time steps (forces, then integrate) of the std or VE propagator on one rank.
    params.minDt follows ts.minDt (the AV switches read it). With refit_box, each step's sync first
    updates `box` in place as Domain::sync does (sphx_global_box: the fitting box on the first sync,
    then limited shrinking). With tree_update, every sync after the first updates the previous tree
    once (Domain::sync); the first converges it. on_step(step), if given, runs after each step's
    forces, before its integration. Returns the Courant step of each."""
    out = []
    for step in range(steps):
        first = first_sync and step == 0
        if refit_box:
            ws.global_box(box, ps.t["x"], ps.t["y"], ps.t["z"], ps.n, first)
        params.minDt = ts.minDt
        upd = tree_update and not first
        if form == "std":
            conv, dtc = std_step(ws, ps, box, params, tree_update=upd)
        else:
            conv, dtc = ve_step(ws, ps, box, params, av_clean=av_clean, ts=ts, tree_update=upd)
        if on_step is not None:
            on_step(step)
        if not conv:
            raise RuntimeError("coupled nc/h-updated failed to converge")
        ts.minDtCourant = dtc
        integrate(ws, ps, box, params, ts)
        out.append(dtc)
    return out
