"""The oracle's adjoints of the std density, IAD and momentum-energy loops against central finite
differences of the oracle's forward loops (which are bitwise the reference's), on CPU.

Neighbour lists are held fixed (checkpointed), as the adjoints assume. Checked both as one random
directional derivative over all inputs and per particle on sampled particles; the bar is
1e-4 relative to the largest gradient magnitude of the checked quantity.

The forward interpolates W linearly in a 20000-point table (spacing 1e-4 in v = r/h) and the
adjoints use the smooth dW/dv. A central difference reproduces the smooth derivative once its
stencil spans several table intervals (the interpolation error then enters as O(dx^2)/step), so
the steps below are sized in v: positions 1e-4 with h ~ 0.14 here (7 intervals), h 1e-3 relative.
Quantities that do not enter a table argument use a 1e-6 relative step. The O(step^2) truncation
of the central difference (visible for the IAD inverse) is removed by Richardson extrapolation,
(4 D(step/2) - D(step)) / 3.
"""
import numpy as np
import pytest

from sphexa_amd import abi, ic


TOL = 1e-4
STEP = {"x": 1e-4, "y": 1e-4, "z": 1e-4, "h": 1e-3}  # absolute for positions, relative otherwise
REL_STEP = 1e-6


def central(fn, orc, s, cot, direction):
    """Richardson-extrapolated central difference of fn along `direction` (dict of arrays)."""

    def d(scale):
        plus, minus = copy_arrays(s["a"]), copy_arrays(s["a"])
        for k, v in direction.items():
            plus[k] = plus[k] + scale * v
            minus[k] = minus[k] - scale * v
        return (fn(orc, s, plus, cot) - fn(orc, s, minus, cot)) / (2 * scale)

    return (4 * d(0.5) - d(1.0)) / 3


def step_of(a, k):
    if k in "xyz":
        return STEP[k]
    return STEP.get(k, REL_STEP) * scale_of(a, k)


@pytest.fixture(scope="module", params=["periodic", "compressing"])
def system(orc, request):
    """A jittered lattice, sorted, with neighbour lists and a full std pass.

    periodic: periodic box, random velocities (covers the fold). The AV term switches on at
    w_ij = 0 with a jump in its derivative, so central differences in positions straddling such a
    pair are not smooth; the momentum position checks therefore use
    compressing: open box, v = -k x + omega x x, so w_ij = -k |r_ij| < 0 for every pair."""
    from oracle import oracle as O
    from oracle import pipeline

    arr, spec = ic.sedov_lattice(10, seed=5)
    rng = np.random.default_rng(7)
    n = arr["x"].size
    if request.param == "periodic":
        arr["vx"], arr["vy"], arr["vz"] = (0.3 * rng.standard_normal(n) for _ in range(3))
        box = abi.Box.make(spec["lo"], spec["hi"], abi.PERIODIC)
    else:
        x, y, z = arr["x"], arr["y"], arr["z"]
        arr["vx"], arr["vy"], arr["vz"] = -0.5 * x + 0.2 * y, -0.5 * y - 0.2 * x + 0.1 * z, -0.5 * z - 0.1 * y
        box = abi.Box.make(spec["lo"], spec["hi"], abi.OPEN)
    arr["alpha"] = rng.uniform(0.5, 1.5, n)
    params = abi.Params.defaults(orc.kernel_K(6.0), flags=abi.SPHX_FLAG_STD_ALPHA)
    o = pipeline.run(orc, arr, box, params, form="std")
    wh, whd = O.tables(orc, 6.0, np.float64)
    return dict(a=o["a"], nb=o["nbrs"], box=box, params=params, wh=wh, whd=whd, n=n, kind=request.param)


def checked_inputs(s, case):
    _, _, _, inputs = CASES[case]
    if case == "momentum" and s["kind"] == "periodic":
        return tuple(k for k in inputs if k not in "xyz")
    return inputs


def cpu_pass(orc, s, a):
    from oracle import oracle as O

    return O.CpuPass(orc, s["box"], a, s["params"], s["nb"], 0, s["n"], s["wh"], s["whd"])


def copy_arrays(a):
    return {k: (None if v is None else v.copy()) for k, v in a.items()}


def zeros(n, names):
    return {k: np.zeros(n) for k in names}


# ---- forward functionals L(theta) = <cotangent, output(theta)> ------------------------------------

def density_functional(orc, s, a, g):
    a = copy_arrays(a)
    cpu_pass(orc, s, a).density_std()
    return float(np.dot(g["rho"], a["rho"]))


C6 = ("c11", "c12", "c13", "c22", "c23", "c33")


def iad_functional(orc, s, a, g):
    a = copy_arrays(a)
    cpu_pass(orc, s, a).iad_std()
    return float(sum(np.dot(g[k], a[k]) for k in C6))


def momentum_functional(orc, s, a, g):
    a = copy_arrays(a)
    cpu_pass(orc, s, a).momentum_energy_std()
    return float(sum(np.dot(g[k], a[k]) for k in ("ax", "ay", "az", "du")))


CASES = {
    # name: (functional, adjoint, cotangent names, differentiable inputs)
    "density": (density_functional, "density_std_vjp", ("rho",), ("x", "y", "z", "h")),
    "iad": (iad_functional, "iad_std_vjp", C6, ("x", "y", "z", "h", "rho")),
    "momentum": (momentum_functional, "momentum_energy_std_vjp", ("ax", "ay", "az", "du"),
                 ("x", "y", "z", "h", "rho", "p", "alpha") + C6),
}


def adjoint(orc, s, name, cot, inputs):
    grad = zeros(s["n"], inputs)
    cpu_pass(orc, s, copy_arrays(s["a"])).vjp(name, cot, grad)
    return grad


def scale_of(a, k):
    return float(np.max(np.abs(a[k])))


@pytest.mark.parametrize("case", sorted(CASES))
def test_directional_derivative(orc, system, case):
    fn, name, cot_names, inputs = CASES[case]
    s = system
    rng = np.random.default_rng(11)
    cot = {k: rng.standard_normal(s["n"]) for k in cot_names}
    grad = adjoint(orc, s, name, cot, inputs)

    # random direction, each input scaled to its own step
    d = {k: rng.standard_normal(s["n"]) * step_of(s["a"], k) for k in checked_inputs(s, case)}
    fd = central(fn, orc, s, cot, d)
    an = sum(float(np.dot(grad[k], d[k])) for k in d)
    assert abs(an - fd) <= TOL * abs(fd), (case, an, fd)


@pytest.mark.parametrize("case", sorted(CASES))
def test_per_particle_components(orc, system, case):
    fn, name, cot_names, inputs = CASES[case]
    s = system
    rng = np.random.default_rng(13)
    cot = {k: rng.standard_normal(s["n"]) for k in cot_names}
    grad = adjoint(orc, s, name, cot, inputs)
    for k in checked_inputs(s, case):
        step = step_of(s["a"], k)
        ref = []
        got = []
        for i in rng.choice(s["n"], 6, replace=False):
            e = np.zeros(s["n"])
            e[i] = 1.0
            ref.append(central(fn, orc, s, cot, {k: e * step}) / step)
            got.append(grad[k][i])
        ref, got = np.array(ref), np.array(got)
        bound = TOL * max(np.max(np.abs(grad[k])), 1e-300)
        assert np.max(np.abs(got - ref)) <= bound, (case, k, got, ref)


def test_std_alpha_flag_is_the_reference_loop_at_alpha_one(orc, system):
    """With the per-particle AV coefficient at 1 the flagged loop equals the reference's."""
    s = system
    a1 = copy_arrays(s["a"])
    a1["alpha"] = np.ones(s["n"])
    cpu_pass(orc, s, a1).momentum_energy_std()
    a2 = copy_arrays(s["a"])
    from oracle import oracle as O

    p = abi.Params.defaults(s["params"].K)  # flags = 0: constant alpha = 1
    O.CpuPass(orc, s["box"], a2, p, s["nb"], 0, s["n"], s["wh"], s["whd"]).momentum_energy_std()
    for k in ("ax", "ay", "az", "du"):
        assert np.array_equal(a1[k], a2[k]), k
