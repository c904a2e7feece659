"""HIP adjoints of the std density, IAD and momentum-energy passes.

* FP64 adjoint == the oracle's adjoint (pinned on CPU by finite differences in
  test_adjoint_oracle.py) on the same inputs and lists, to 1e-9 of each gradient's largest
  magnitude (the summation order of the neighbour-side terms is the only difference).
* Repeated calls give the same bits (the neighbour-side terms are summed per particle in pair
  order, without atomics), in both forms; calls over two halves of the targets add up to the call
  over all of them.
* The gradient of L = sum (rho - rho_t)^2 + w sum |a|^2 through the torch.autograd harness
  (tests/sph_autograd.py) matches central finite differences of L on the GPU (FP64, step 1e-6,
  lists held fixed) on sampled particles, to 1e-4 of the largest magnitude among the sampled
  gradients. The difference is taken per particle term before summing, so the rounding of the
  large total (the blast centre dominates L) does not enter.
* The mixed form's adjoint (FP32 pair arithmetic) against the FP64 one, to 2e-3; at 2M particles
  the whole loss gradient, FP32 forward and backward, against FP64.

Geometry: the Sedov-like jittered lattice in the periodic unit box with a converging flow
v = -0.1 x (so the AV term is active), per-particle AV coefficient alpha = 1.
"""
import numpy as np
import pytest

from sphexa_amd import abi, ic
from sphexa_amd.particles import STD_FIELDS

pytestmark = pytest.mark.gpu

C6 = ("c11", "c12", "c13", "c22", "c23", "c33")
GRAD_OUT = {
    "density": ("sphx_density_std_vjp", "density_std_vjp", ("rho",), ("x", "y", "z", "h")),
    "iad": ("sphx_iad_std_vjp", "iad_std_vjp", C6, ("x", "y", "z", "h", "rho")),
    "momentum": ("sphx_momentum_energy_std_vjp", "momentum_energy_std_vjp", ("ax", "ay", "az", "du"),
                 ("x", "y", "z", "h", "rho", "p", "alpha") + C6),
}


def build(side, mixed, seed=3):
    """Device particle set after one full std pass, with its neighbour lists in the workspace."""
    import torch

    from sphexa_amd import lib
    from sphexa_amd.particles import ParticleSet

    arr, spec = ic.sedov_lattice(side, seed=seed)
    for k in "xyz":
        arr["v" + k] = -0.1 * arr[k]
    arr["alpha"] = np.ones(arr["x"].size)
    ws = lib.Workspace(0)
    hydro = abi.SPHX_F32 if mixed else abi.SPHX_F64
    params = abi.Params.defaults(ws.K, hydro_type=hydro, flags=abi.SPHX_FLAG_STD_ALPHA)
    box = abi.Box.make(spec["lo"], spec["hi"], abi.PERIODIC)
    ps = ParticleSet.from_host(arr, STD_FIELDS + ("alpha",), mixed=mixed)
    n = ps.n
    f = ps.fields()
    ws.domain_sync(box, f, params, n, 64)
    assert ws.find_neighbors(box, f, params, 0, n, True)
    ws.pair("sphx_density_std", box, f, params, 0, n)
    ws.eos("sphx_eos_std", f, params, 0, n)
    ws.pair("sphx_iad_std", box, f, params, 0, n)
    ws.momentum_energy_std(box, f, params, 0, n)
    torch.cuda.synchronize()
    return ws, ps, box, params


@pytest.fixture(scope="module")
def sys64():
    return build(12, mixed=False)


def oracle_vjp(orc, ws, ps, box, params, case, cot):
    """The oracle's adjoint on the device state (host copies, GPU lists in the CPU layout)."""
    import torch

    from oracle import oracle as O

    n = ps.n
    ngmax = params.ngmax
    nb = torch.zeros(n * ngmax, dtype=torch.int32, device="cuda")
    ws.copy_neighbors(nb, 0, n)
    nb = nb.cpu().numpy().view(np.uint32)
    a = {k: ps.host(k).astype(np.float64) if k not in ("nc", "keys") else ps.host(k) for k in ps.t}
    a["alpha"] = ps.host("alpha").astype(np.float64)
    wh, whd = O.tables(orc, 6.0, np.float64)
    p64 = abi.Params.defaults(params.K, flags=params.flags)
    cp = O.CpuPass(orc, box, a, p64, nb, 0, n, wh, whd)
    _, name, _, outs = GRAD_OUT[case]
    grad = {k: np.zeros(n) for k in outs}
    cp.vjp(name, {k: v.astype(np.float64) for k, v in cot.items()}, grad)
    return grad


def gpu_vjp(ws, ps, box, params, case, cot):
    import torch

    from sph_autograd import adjoint

    cname, _, _, outs = GRAD_OUT[case]
    dt = ps.hydro_dtype
    cot_t = {k: torch.from_numpy(v).to(dt).cuda() for k, v in cot.items()}
    out = {k: torch.zeros(ps.n, dtype=dt, device="cuda") for k in outs}
    ws.vjp(cname, box, ps.fields(), params, 0, ps.n, adjoint(cot_t), adjoint(out))
    torch.cuda.synchronize()
    return {k: v.double().cpu().numpy() for k, v in out.items()}


def cotangents(case, n, seed):
    rng = np.random.default_rng(seed)
    return {k: rng.standard_normal(n) for k in GRAD_OUT[case][2]}


@pytest.mark.parametrize("case", sorted(GRAD_OUT))
def test_fp64_adjoint_matches_oracle(orc, sys64, case):
    ws, ps, box, params = sys64
    cot = cotangents(case, ps.n, 1)
    got = gpu_vjp(ws, ps, box, params, case, cot)
    ref = oracle_vjp(orc, ws, ps, box, params, case, cot)
    for k in ref:
        scale = max(np.max(np.abs(ref[k])), 1e-300)
        err = np.max(np.abs(got[k] - ref[k])) / scale
        print(f"{case} d/d{k}: max rel err {err:.2e}", flush=True)
        assert err <= 1e-9, (case, k, err)


@pytest.mark.parametrize("mixed", [False, True])
@pytest.mark.parametrize("case", sorted(GRAD_OUT))
def test_adjoint_is_deterministic(case, mixed):
    ws, ps, box, params = build(12, mixed=mixed, seed=6)
    cot = cotangents(case, ps.n, 5)
    first = gpu_vjp(ws, ps, box, params, case, cot)
    for rep in range(3):
        again = gpu_vjp(ws, ps, box, params, case, cot)
        for k in first:
            assert np.array_equal(first[k], again[k]), (case, mixed, k, rep)


@pytest.mark.parametrize("case", sorted(GRAD_OUT))
def test_adjoint_over_target_halves(sys64, case):
    """the neighbour-side sums of a call cover exactly its targets: two calls over [0, n/2) and
    [n/2, n), accumulating into the same gradients, equal one call over [0, n) to rounding"""
    import torch

    from sph_autograd import adjoint

    ws, ps, box, params = sys64
    cot = cotangents(case, ps.n, 7)
    full = gpu_vjp(ws, ps, box, params, case, cot)
    cname, _, _, outs = GRAD_OUT[case]
    cot_t = {k: torch.from_numpy(v).cuda() for k, v in cot.items()}
    out = {k: torch.zeros(ps.n, dtype=torch.float64, device="cuda") for k in outs}
    half = (ps.n // 2) // 64 * 64
    for lo, hi in ((0, half), (half, ps.n)):
        ws.vjp(cname, box, ps.fields(), params, lo, hi, adjoint(cot_t), adjoint(out))
    torch.cuda.synchronize()
    for k in outs:
        got = out[k].cpu().numpy()
        scale = max(np.max(np.abs(full[k])), 1e-300)
        assert np.max(np.abs(got - full[k])) <= 1e-12 * scale, (case, k)


@pytest.mark.parametrize("case", sorted(GRAD_OUT))
def test_mixed_adjoint_against_fp64(orc, case):
    ws, ps, box, params = build(12, mixed=True)
    cot = cotangents(case, ps.n, 2)
    got = gpu_vjp(ws, ps, box, params, case, cot)
    ref = oracle_vjp(orc, ws, ps, box, params, case, cot)
    for k in ref:
        scale = max(np.max(np.abs(ref[k])), 1e-300)
        err = np.max(np.abs(got[k] - ref[k])) / scale
        rms = np.sqrt(np.mean((got[k] - ref[k]) ** 2)) / scale
        print(f"mixed {case} d/d{k}: max {err:.2e} rms {rms:.2e}", flush=True)
        assert err <= 2e-3, (case, k, err)


def loss_setup(ws, ps, box, params, seed=4, w=1e-3):
    """Harness graph, inputs and the target density 1 + 0.05 sin(2 pi k.x + phase)."""
    import torch

    from sph_autograd import StdGraph

    rng = np.random.default_rng(seed)
    kvec = rng.integers(1, 3, size=3)
    phase = rng.uniform(0, 2 * np.pi)
    x, y, z = (ps.t[k][: ps.n] for k in "xyz")
    rho_t = 1.0 + 0.05 * torch.sin(2 * np.pi * (kvec[0] * x + kvec[1] * y + kvec[2] * z) + phase)
    const = {k: ps.t[k] for k in ("m", "vx", "vy", "vz", "c", "nc")}
    g = StdGraph(ws, box, params, const, ps.n, ps.hydro_dtype)
    u = ps.t["u"][: ps.n]
    return g, rho_t, u, w


def check_against_fd(g, inputs, grads, u, gamma, rho_t, w, sample, tag, eps=1e-6):
    """Central differences of L for each sampled particle and input, per-particle terms
    differenced before the sum; asserts agreement to 1e-4 of the sample's largest gradient.

    The forward interpolates W linearly in a table with spacing 1e-4 in v = r/h while the adjoints
    use the smooth dW/dv; the difference stencil must span about one table interval or more
    (2 eps / h >~ 1e-4) to measure the smooth derivative, which holds for eps = 1e-6 at h ~ 0.011
    (2M particles) and is met at h ~ 0.12 (12^3) with eps = 1e-4."""
    import torch

    from sph_autograd import std_loss_terms

    def terms(name, i, delta):
        vals = {k: v.detach() for k, v in inputs.items()}
        old = float(vals[name][i])
        vals[name][i] += delta
        with torch.no_grad():
            t = std_loss_terms(g, vals["x"], vals["y"], vals["z"], vals["h"], vals["alpha"], u, gamma, rho_t, w)
        vals[name][i] = old
        return t

    for name in inputs:
        fd = np.array([float((terms(name, i, eps) - terms(name, i, -eps)).sum()) / (2 * eps) for i in sample])
        an = grads[name][sample]
        scale = max(np.max(np.abs(fd)), 1e-300)
        err = np.max(np.abs(an - fd)) / scale
        print(f"{tag} dL/d{name}: max |adjoint - FD| / max|FD| over {len(sample)} particles = {err:.2e} "
              f"(sample scale {scale:.3e}, global max|grad| {np.max(np.abs(grads[name])):.3e})", flush=True)
        assert err <= 1e-4, (name, err, an, fd)


def test_loss_gradient_matches_finite_differences(sys64):
    import torch

    from sph_autograd import std_loss

    ws, ps, box, params = sys64
    g, rho_t, u, w = loss_setup(ws, ps, box, params)
    inputs = {k: ps.t[k][: ps.n].clone().requires_grad_(True) for k in ("x", "y", "z", "h", "alpha")}
    L = std_loss(g, inputs["x"], inputs["y"], inputs["z"], inputs["h"], inputs["alpha"], u, params.gamma, rho_t, w)
    L.backward()
    grads = {k: v.grad.double().cpu().numpy() for k, v in inputs.items()}

    rng = np.random.default_rng(9)
    sample = rng.choice(ps.n, 16, replace=False)
    check_against_fd(g, inputs, grads, u, params.gamma, rho_t, w, sample, "12^3", eps=1e-4)


def test_loss_gradient_at_2m_particles():
    """The differentiation workload at scale: 126^3 = 2,000,376 particles, FD on 64 random particles
    (step 1e-6, FP64), and the mixed form's gradient (FP32 forward and backward) against FP64."""
    import torch

    from sph_autograd import std_loss

    def gradient(mixed):
        ws, ps, box, params = build(126, mixed=mixed)
        g, rho_t, u, w = loss_setup(ws, ps, box, params)
        inputs = {k: ps.t[k][: ps.n].clone().requires_grad_(True) for k in ("x", "y", "z", "h", "alpha")}
        L = std_loss(g, inputs["x"], inputs["y"], inputs["z"], inputs["h"], inputs["alpha"], u, params.gamma,
                     rho_t, w)
        L.backward()
        return ws, ps, box, params, g, rho_t, u, w, inputs, float(L.detach())

    ws, ps, box, params, g, rho_t, u, w, inputs, L0 = gradient(False)
    grads = {k: v.grad.double().cpu().numpy() for k, v in inputs.items()}

    rng = np.random.default_rng(64)
    sample = rng.choice(ps.n, 64, replace=False)
    print(f"n = {ps.n}, L = {L0:.6e}", flush=True)
    check_against_fd(g, inputs, grads, u, params.gamma, rho_t, w, sample, "2M")
    del ws

    ws32, ps32, *_, inputs32, L32 = gradient(True)
    print(f"mixed L = {L32:.6e} (FP64 {L0:.6e})", flush=True)
    for name in inputs32:
        a = inputs32[name].grad.double().cpu().numpy()
        scale = np.max(np.abs(grads[name]))
        err = np.max(np.abs(a - grads[name])) / scale
        rms = np.sqrt(np.mean((a - grads[name]) ** 2)) / scale
        local = np.max(np.abs(a[sample] - grads[name][sample])) / np.max(np.abs(grads[name][sample]))
        print(f"2M mixed dL/d{name}: max {err:.2e} rms {rms:.2e} of max|grad|; {local:.2e} of the 64-particle "
              f"sample's scale", flush=True)
        assert err <= 1e-2, (name, err)
