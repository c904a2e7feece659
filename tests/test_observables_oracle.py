"""The oracle's conserved quantities (oracle/sphx_oracle.hpp: conservedQuantities, restating
main/src/observables/conserved_quantities.hpp:50-107 with the GPU path's shared cv) against
closed forms: a rigid translation (eKin = M |v|^2 / 2, linmom = M v, angmom = M x_cm x v), a rigid
rotation about z (angmom_z = w I_z, eKin = w^2 I_z / 2, which pins the cross product's indices
and signs), the internal energy from u and from temperature (cv sum m T), and sum nc."""
import ctypes as C

import numpy as np

from sphexa_amd import abi


def run(orc, d, params, first=0, last=None):
    n = len(d["x"])
    out = np.zeros(9)
    orc.conserved_quantities(C.byref(abi.fields_from(d)), C.byref(params), first, n if last is None else last,
                             out.ctypes.data)
    return out


def system(rng, n, hydro=np.float64):
    d = {k: rng.uniform(-1, 1, n) for k in "xyz"}
    d["m"] = rng.uniform(0.5, 1.5, n).astype(hydro)
    return d


def test_rigid_translation(orc):
    rng = np.random.default_rng(1)
    d = system(rng, 4000)
    v = np.array([0.3, -0.2, 0.7])
    for k, c in zip(("vx", "vy", "vz"), v):
        d[k] = np.full(4000, c)
    o = run(orc, d, abi.Params.defaults(1.0))
    M = d["m"].sum()
    xcm = np.array([(d["m"] * d[k]).sum() for k in "xyz"]) / M
    np.testing.assert_allclose(o[0], 0.5 * M * v @ v, rtol=1e-12)
    np.testing.assert_allclose(o[2:5], M * v, rtol=1e-12)
    np.testing.assert_allclose(o[5:8], M * np.cross(xcm, v), rtol=1e-10, atol=1e-12)


def test_rigid_rotation(orc):
    rng = np.random.default_rng(2)
    d = system(rng, 4000)
    w = 1.7
    d["vx"], d["vy"], d["vz"] = -w * d["y"], w * d["x"], np.zeros(4000)
    o = run(orc, d, abi.Params.defaults(1.0))
    Iz = (d["m"] * (d["x"] ** 2 + d["y"] ** 2)).sum()
    np.testing.assert_allclose(o[7], w * Iz, rtol=1e-12)
    np.testing.assert_allclose(o[0], 0.5 * w * w * Iz, rtol=1e-12)
    # L_x = -w sum m x z, L_y = -w sum m y z for v = w z_hat x r
    np.testing.assert_allclose(o[5], -w * (d["m"] * d["x"] * d["z"]).sum(), rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(o[6], -w * (d["m"] * d["y"] * d["z"]).sum(), rtol=1e-10, atol=1e-12)


def test_internal_energy_and_counts(orc):
    rng = np.random.default_rng(3)
    n = 3000
    d = system(rng, n)
    for k in ("vx", "vy", "vz"):
        d[k] = np.zeros(n)
    d["u"] = rng.uniform(0.1, 2, n)
    d["nc"] = rng.integers(1, 200, n).astype(np.uint32)
    p = abi.Params.defaults(1.0)
    o = run(orc, d, p)
    np.testing.assert_allclose(o[1], (d["m"] * d["u"]).sum(), rtol=1e-12)
    assert o[8] == d["nc"].astype(np.float64).sum()
    del d["u"]
    d["temp"] = rng.uniform(1e-4, 1e-2, n)
    o = run(orc, d, p)
    cv = 8.317e7 / 10.0 / (5.0 / 3.0 - 1.0)
    np.testing.assert_allclose(o[1], cv * (d["m"] * d["temp"]).sum(), rtol=1e-12)
    # a sub-range sums only its particles
    o2 = run(orc, d, p, 100, 200)
    np.testing.assert_allclose(o2[1], cv * (d["m"][100:200] * d["temp"][100:200]).sum(), rtol=1e-12)
