"""Pins of the CPU oracle (oracle/sphx_oracle.hpp) -- CPU only.

* the reference's own known-answer tests, transcribed (file:line on each test);
* closed forms and brute force (all-pairs neighbour search, Hilbert continuity/inversion,
  cornerstone tree invariants, Simpson quadrature);
* bitwise equality with the reference headers compiled into oracle/_ref (skipped when that
  wrapper is not built);
* the committed golden vectors in tests/golden (written from oracle/_ref by
  tests/golden/make_golden.py), so the pin holds where the reference checkout is absent.
"""
import ctypes as C
import os

import numpy as np
import pytest

from sphexa_amd import abi, ic

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def ptr(a):
    return a.ctypes.data


# ---- SFC keys --------------------------------------------------------------------------------
def test_hilbert_first_order_curve(orc):
    """domain/test/unit/sfc/hilbert.cpp:83-112: level-1 octants follow {0,1,3,2,6,7,5,4}."""
    hilbert_to_morton = [0, 1, 3, 2, 6, 7, 5, 4]
    half = 1 << 20
    for xi in (0, 1):
        for yi in (0, 1):
            for zi in (0, 1):
                for off in (0, half - 1):
                    ix = np.array([half * xi + off], np.uint32)
                    iy = np.array([half * yi + off], np.uint32)
                    iz = np.array([half * zi + off], np.uint32)
                    k = np.zeros(1, np.uint64)
                    orc.hilbert_encode(ptr(ix), ptr(iy), ptr(iz), ptr(k), 1)
                    octant = int(k[0]) >> 60 & 7
                    assert hilbert_to_morton[octant] == 4 * xi + 2 * yi + zi


def test_hilbert_continuity(orc):
    """domain/test/unit/sfc/hilbert.cpp:116-141: consecutive keys are unit neighbours in space."""
    keys = []
    for level in range(1, 21):
        for octant in range(8 if level > 1 else 7):
            last = (octant + 1) * (1 << (3 * (21 - level))) - 1
            keys += [last, last + 1]
    k = np.array(keys, np.uint64)
    x, y, z = (np.zeros(k.size, np.uint32) for _ in range(3))
    orc.hilbert_decode(ptr(k), ptr(x), ptr(y), ptr(z), k.size)
    d = (np.abs(x[0::2].astype(int) - x[1::2]) + np.abs(y[0::2].astype(int) - y[1::2])
         + np.abs(z[0::2].astype(int) - z[1::2]))
    assert np.all(d == 1)


def test_hilbert_inversion(orc):
    """domain/test/unit/sfc/hilbert.cpp:144-176: decode(encode(p)) == p on random points."""
    rng = np.random.default_rng(0)
    x, y, z = (rng.integers(0, 1 << 21, 5000, dtype=np.uint32) for _ in range(3))
    k = np.zeros(x.size, np.uint64)
    orc.hilbert_encode(ptr(x), ptr(y), ptr(z), ptr(k), x.size)
    a, b, c = (np.zeros(x.size, np.uint32) for _ in range(3))
    orc.hilbert_decode(ptr(k), ptr(a), ptr(b), ptr(c), k.size)
    assert np.array_equal(a, x) and np.array_equal(b, y) and np.array_equal(c, z)
    # bijection on a full level-2 grid: 64 distinct keys, all below 8^2 after scaling
    g = np.arange(4, dtype=np.uint32) << 19
    gx, gy, gz = (v.ravel().copy() for v in np.meshgrid(g, g, g, indexing="ij"))
    kk = np.zeros(64, np.uint64)
    orc.hilbert_encode(ptr(gx), ptr(gy), ptr(gz), ptr(kk), 64)
    assert sorted((kk >> np.uint64(57)).tolist()) == list(range(64))


def test_key_center_octants(orc):
    """domain/test/unit/sfc/sfc.cpp:38-60: the box centre belongs to the octant farthest from the origin."""
    from oracle import oracle as O

    box = abi.Box.make(-1.0, 1.0, abi.OPEN)
    pts = np.array([0.0, 1.0, -1e-40, -1.0])
    k = O.sfc_keys(orc, box, pts, pts.copy(), pts.copy())
    top = (k >> np.uint64(60)) & np.uint64(7)
    assert top[0] == top[1] and top[2] == top[3] and top[0] != top[2]


# ---- distances, quadrature, tables -------------------------------------------------------------
def test_distance_sq_pbc_literals(orc):
    """domain/test/unit/neighbors/findneighbors.cpp:27-41."""
    b = abi.Box.make(0.0, 10.0, abi.OPEN)
    assert orc.distance_sq_pbc(C.byref(b), 1, 0, 0, 9, 0, 0) == 64.0
    assert orc.distance_sq_pbc(C.byref(b), 9, 0, 0, 1, 0, 0) == 64.0
    assert orc.distance_sq_pbc(C.byref(b), 9, 9, 9, 1, 1, 1) == 192.0
    b = abi.Box.make(0.0, 10.0, abi.PERIODIC)
    assert orc.distance_sq_pbc(C.byref(b), 1, 0, 0, 9, 0, 0) == 4.0
    assert orc.distance_sq_pbc(C.byref(b), 9, 0, 0, 1, 0, 0) == 4.0
    assert orc.distance_sq_pbc(C.byref(b), 9, 9, 9, 1, 1, 1) == 12.0


def test_simpson_and_normalization(orc):
    """sph/test/table_creation.cpp:16-40: Simpson of sin to 1e-14; K within 1e-4 of SPHYNX's fit;
    plus an independent high-order quadrature of 1/K."""
    assert abs(orc.simpson_sin() - 1.0) < 1e-14
    K = orc.kernel_K(6.0)
    assert abs(orc.sphynx_K(6.0) - K) < 1e-4
    from scipy.integrate import quad

    f = lambda v: 4 * np.pi * v * v * (np.sinc(v / 2.0)) ** 6  # np.sinc(x) = sin(pi x)/(pi x)
    I, _ = quad(f, 0, 2, epsabs=1e-15, epsrel=1e-14, limit=200)
    assert abs(1.0 / I - K) / K < 1e-12


def test_table_lookup_is_linear_interpolation(orc):
    """table_lookup.hpp:10-26: exact at nodes, linear in between, zero beyond the support."""
    from oracle import oracle as O

    wh, whd = O.tables(orc)
    dx = 2.0 / 19999
    for i in (0, 1, 777, 19998):
        # i*dx*invDx may truncate to i-1: the interpolant is then evaluated at the far end of
        # the previous interval, equal to the node value up to rounding
        assert orc.lookup_f64(ptr(wh), i * dx) == pytest.approx(wh[i], rel=1e-9, abs=1e-300)
        mid = (i + 0.5) * dx
        assert orc.lookup_f64(ptr(wh), mid) == pytest.approx(0.5 * (wh[i] + wh[i + 1]), rel=1e-12)
    assert orc.lookup_f64(ptr(wh), 2.0) == 0.0
    assert orc.lookup_f64(ptr(wh), 2.5) == 0.0
    v = np.linspace(0.01, 1.99, 7)
    assert np.allclose(wh[np.rint(v / dx).astype(int)], np.sinc(np.rint(v / dx) * dx / 2) ** 6, rtol=1e-12)
    # derivative table vs central difference of the exact kernel
    s = lambda t: np.sinc(t / 2) ** 6
    for i in (100, 5000, 15000):
        t = i * dx
        assert whd[i] == pytest.approx((s(t + 1e-6) - s(t - 1e-6)) / 2e-6, rel=1e-6, abs=1e-9)


# ---- tree ------------------------------------------------------------------------------------
def _tree_for(orc, n, gaussian, seed=42, bucket=64):
    from oracle import oracle as O

    rng = np.random.default_rng(seed)
    if gaussian:
        pts = np.clip(rng.normal(0.5, 0.2, (3, n)), 0, 1 - 1e-12)
    else:
        pts = rng.uniform(0, 1, (3, n))
    box = abi.Box.make(0.0, 1.0, abi.OPEN)
    keys = O.sfc_keys(orc, box, *[np.ascontiguousarray(p) for p in pts])
    sk, order = O.sort_keys(orc, keys)
    return O.Tree(orc, sk, box, bucket), sk


@pytest.mark.parametrize("gaussian", [False, True])
def test_tree_invariants(orc, gaussian):
    """Cornerstone fixed point: leaves tile the key space with power-of-8 ranges, every leaf holds
    <= bucket particles, every internal node > bucket, numInternal = (L-1)/7, and the linked
    octree's children are the 8 equal sub-ranges of their parent."""
    t, keys = _tree_for(orc, 20000, gaussian)
    L = t.numLeafNodes
    leaves = t.leaves.astype(object)
    assert leaves[0] == 0 and leaves[-1] == 1 << 63
    ranges = np.diff(t.leaves)
    assert all(int(r) & (int(r) - 1) == 0 and (int(r).bit_length() - 1) % 3 == 0 for r in ranges)
    assert np.all(t.counts <= 64)
    assert np.array_equal(np.searchsorted(keys, t.leaves[:-1]), t.layout[:-1])
    assert t.numInternalNodes == (L - 1) // 7 and t.numNodes == L + t.numInternalNodes
    internal = np.where(t.childOffsets > 0)[0]
    assert internal.size == t.numInternalNodes
    for node in internal:
        p = int(t.prefixes[node])
        plen = p.bit_length() - 1
        key = (p ^ (1 << plen)) << (63 - plen)
        rng = 1 << (63 - plen)
        lo, hi = np.searchsorted(keys, [key, key + rng])
        assert hi - lo > 64
        kids = t.prefixes[t.childOffsets[node]:t.childOffsets[node] + 8].astype(object)
        for c, kp in enumerate(kids):
            assert kp == (1 << (plen + 3)) | (((key + c * (rng >> 3)) >> (60 - plen)))
    # levelRange brackets the placeholder-bit levels
    for lvl in range(22):
        lo, hi = t.levelRange[lvl], t.levelRange[lvl + 1]
        assert all((int(p).bit_length() - 1) // 3 == lvl for p in t.prefixes[lo:hi])


def test_tree_root_only(orc):
    """domain/test/unit/tree/octree.cpp: fewer than bucket particles give the root as the only leaf."""
    t, _ = _tree_for(orc, 40, False)
    assert t.numLeafNodes == 1 and t.numNodes == 1 and t.childOffsets[0] == 0
    assert t.prefixes[0] == 1 and t.layout.tolist() == [0, 40]


def test_node_centers_closed_form(orc):
    """nodeFpCenters: the root of [0,1]^3 is centred at 0.5 with half size 0.5, and every node's
    box contains the particles of its key range."""
    from oracle import oracle as O

    arr, spec = ic.sedov_lattice(10, seed=3)
    box = abi.Box.make(-0.5, 0.5, abi.PERIODIC)
    keys = O.sfc_keys(orc, box, arr["x"], arr["y"], arr["z"])
    sk, order = O.sort_keys(orc, keys)
    t = O.Tree(orc, sk, box)
    assert np.allclose(t.centers[0], 0.0) and np.allclose(t.sizes[0], 0.5)
    x = np.stack([arr[k][order] for k in "xyz"], 1)
    i2l = t.internalToLeaf
    for node in np.where(i2l >= 0)[0]:
        lf = i2l[node]
        p = x[t.layout[lf]:t.layout[lf + 1]]
        if len(p):
            assert np.all(np.abs(p - t.centers[node]) <= t.sizes[node] * (1 + 1e-12))


# ---- neighbour search: brute force (domain/test/unit/neighbors/findneighbors.cpp:43-133) --------
def brute_force(x, y, z, h, box, ngmax):
    """all2allNeighbors (domain/test/unit/neighbors/all_to_all.hpp:27-65) in numpy: exact IEEE
    element-wise arithmetic, rint folding on periodic axes."""
    n = x.size
    L = np.array(box.lengths)
    pbc = np.array(box.periodic, dtype=np.float64)
    lists, counts = [], np.zeros(n, np.uint32)
    P = np.stack([x, y, z], 1)
    for i in range(n):
        d = P[i] - P
        d = d - pbc * L * np.rint(d * (1.0 / L))
        d2 = d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]
        r = 2 * h[i]
        hit = np.where(d2 < r * r)[0]
        hit = hit[hit != i]
        counts[i] = hit.size
        lists.append(hit[:ngmax])
    return lists, counts


@pytest.mark.parametrize("radius,n", [(0.124, 2500), (0.0624, 2500), (3.0, 500)])
@pytest.mark.parametrize("limits", [(0, 1, 0, 1, 0, 1), (-1.2, 0.23, -0.213, 3.213, -5.1, 1.23)])
@pytest.mark.parametrize("bc", [abi.OPEN, abi.PERIODIC])
@pytest.mark.parametrize("gaussian", [False, True])
def test_neighbors_vs_brute_force(orc, radius, n, limits, bc, gaussian):
    from oracle import oracle as O

    rng = np.random.default_rng(42)
    lo = np.array(limits[0::2])
    hi = np.array(limits[1::2])
    if gaussian:
        u = np.clip(rng.normal(0.5, 0.2, (n, 3)), 0.0, 1.0 - 1e-12)
    else:
        u = rng.uniform(0, 1, (n, 3))
    pts = lo + u * (hi - lo)
    box = abi.Box.make(lo, hi, bc)
    x, y, z = (np.ascontiguousarray(pts[:, d]) for d in range(3))
    keys = O.sfc_keys(orc, box, x, y, z)
    sk, order = O.sort_keys(orc, keys)
    x, y, z = x[order].copy(), y[order].copy(), z[order].copy()
    t = O.Tree(orc, sk, box)
    h = np.full(n, radius / 2)
    ngmax = n
    a = dict(x=x, y=y, z=z, h=h, nc=np.zeros(n, np.uint32))
    p = abi.Params.defaults(1.0, ngmax=ngmax, ng0=1)
    nb, _ = O.find_neighbors(orc, t, box, a, p, 0, n, iterate_h=False)
    ref_lists, ref_counts = brute_force(x, y, z, h, box, ngmax)
    assert np.array_equal(a["nc"] - 1, ref_counts)
    for i in range(n):
        assert np.array_equal(np.sort(nb[i, :a["nc"][i] - 1]), np.sort(ref_lists[i]))


# ---- SPH known answers (sph/test/std.cpp, sph/test/ve.cpp) ---------------------------------------
def _kat_fields(n, data):
    a = {k: np.zeros(n) for k in abi.HYDRO_FIELDS}
    a.update({k: np.asarray(v, np.float64).copy() for k, v in data.items()})
    a["du"] = np.zeros(n)
    a["tdpdTrho"] = None
    return a


def test_std_iad_and_momentum_kat(orc):
    """sph/test/std.cpp:43-127 (5 particles, K = sphynx_3D_k(6), open box [0,6])."""
    from oracle import oracle as O

    data = dict(x=[1.0, 1.1, 3.2, 1.3, 2.4], y=[1.1, 1.2, 1.3, 4.4, 5.5], z=[1.2, 2.3, 1.4, 1.5, 1.6],
                h=[5.0, 5.1, 5.2, 5.3, 5.4], m=[1.1, 1.2, 1.3, 1.4, 1.5], rho=[0.014, 0.015, 0.016, 0.017, 0.018],
                vx=[0.010, -0.020, 0.030, -0.040, 0.050], vy=[-0.011, 0.021, -0.031, 0.041, -0.051],
                vz=[0.091, -0.081, 0.071, -0.061, 0.055], c=[0.4, 0.5, 0.6, 0.7, 0.8], p=[0.2, 0.3, 0.4, 0.5, 0.6],
                c11=[0.21, 0.27, 0.10, 0.45, 0.46], c12=[-0.22, -0.29, -0.11, -0.44, -0.47],
                c13=[-0.23, -0.31, -0.12, -0.43, -0.48], c22=[0.24, 0.32, 0.13, 0.42, 0.49],
                c23=[-0.25, -0.33, -0.14, -0.41, -0.50], c33=[0.26, 0.34, 0.15, 0.40, 0.51])
    a = _kat_fields(5, data)
    a["nc"] = np.array([5, 0, 0, 0, 0], np.uint32)
    nb = np.array([[1, 2, 3, 4]], np.uint32)
    box = abi.Box.make(0.0, 6.0, abi.OPEN)
    p = abi.Params.defaults(orc.sphynx_K(6.0), ngmax=4)
    wh, whd = O.tables(orc)
    cp = O.CpuPass(orc, box, a, p, nb, 0, 1, wh, whd)

    b = {k: v.copy() if v is not None else None for k, v in a.items()}
    cpi = O.CpuPass(orc, box, b, p, nb, 0, 1, wh, whd)
    cpi.iad_std()
    exp = [0.68826690779384281, -0.12963692768970825, -0.20435302538490346, 0.39616100688793993,
           -0.16797800827029263, 1.9055087813473524]
    got = [b[k][0] for k in ("c11", "c12", "c13", "c22", "c23", "c33")]
    assert np.allclose(got, exp, atol=1e-8, rtol=0)

    maxvs = np.zeros(5)
    cp.momentum_energy_std(maxvs)
    assert a["ax"][0] == pytest.approx(14.407211846688075, abs=1.3e-7)
    assert a["ay"][0] == pytest.approx(-1.2396802157028355, abs=1.4e-7)
    assert a["az"][0] == pytest.approx(15.596554152643426, abs=2.15e-7)
    assert a["du"][0] == pytest.approx(-0.40541191600274296, abs=1e-8)
    assert maxvs[0] == pytest.approx(1.4112466828564341, abs=1e-10)


def _ve_kat_setup(orc):
    """sph/test/ve.cpp:51-100: 99 particles from the reference's example_data.txt, neighbours of
    particle 0 are 1..98, K = sphynx_3D_k(6), open box [-1e9, 1e9]."""
    cols = ["x", "y", "z", "vx", "vy", "vz", "h", "c", "c11", "c12", "c13", "c22", "c23", "c33", "p", "gradh", "rho0",
            "sumwhrho0", "sumwh", "dvxdx", "dvxdy", "dvxdz", "dvydx", "dvydy", "dvydz", "dvzdx", "dvzdy", "dvzdz",
            "alpha", "u", "divv"]
    raw = np.loadtxt(os.path.join(GOLDEN, "ve_example_data.txt"))
    assert raw.shape == (99, 31)
    d = {c: raw[:, i].copy() for i, c in enumerate(cols)}
    n = 99
    K = orc.sphynx_K(6.0)
    mpart = 3.781038064465603e26
    m = np.full(n, mpart)
    xm = mpart / d["rho0"]
    kx = K * xm / d["h"] ** 3
    prho = d["p"] / (kx * m * m * d["gradh"])
    a = _kat_fields(n, {k: d[k] for k in ("x", "y", "z", "vx", "vy", "vz", "h", "c", "c11", "c12", "c13", "c22",
                                          "c23", "c33", "p", "gradh", "alpha", "divv")})
    a.update(m=m, xm=xm, kx=kx, prho=prho)
    a["dV11"] = d["dvxdx"]
    a["dV12"] = d["dvxdy"] + d["dvydx"]
    a["dV13"] = d["dvxdz"] + d["dvzdx"]
    a["dV22"] = d["dvydy"]
    a["dV23"] = d["dvydz"] + d["dvzdy"]
    a["dV33"] = d["dvzdz"]
    a["nc"] = np.zeros(n, np.uint32)
    a["nc"][0] = 99
    nb = np.arange(1, 99, dtype=np.uint32)[None, :]
    box = abi.Box.make(-1e9, 1e9, abi.OPEN)
    p = abi.Params.defaults(K, ngmax=98, minDt=0.3)
    return a, nb, box, p, d


def test_ve_kats(orc):
    """sph/test/ve.cpp:102-233: AVswitches, divv/curlv, IAD, momentum-energy (+/- AV cleaning),
    VeDefGradh and XMass of particle 0 against the reference's literal values."""
    from oracle import oracle as O

    wh, whd = O.tables(orc)

    def fresh():
        a, nb, box, p, d = _ve_kat_setup(orc)
        return a, O.CpuPass(orc, box, a, p, nb, 0, 1, wh, whd), d

    a, cp, d = fresh()
    cp.av_switches()
    assert a["alpha"][0] == pytest.approx(0.93941905320351171, abs=2e-9)

    a, cp, d = fresh()
    cp.iad_divv_curlv()
    iad_exp = [1.9296619855715329e-18, -1.7838691836843698e-20, -1.2892885646884301e-20, 1.9482845913025683e-18,
               1.635410357476855e-20, 1.9246939006338132e-18]
    got = [a[k][0] for k in ("c11", "c12", "c13", "c22", "c23", "c33")]
    assert np.allclose(got, iad_exp, atol=1e-10, rtol=0)

    # divV_curlVJLoop on the data file's c_ij, as the reference test feeds them
    a, cp, d = fresh()
    for k in ("divv", "curlv", "dV11", "dV12", "dV13", "dV22", "dV23", "dV33"):
        a[k] = np.zeros(99)
    orc.divv_curlv(C.byref(cp.box), C.byref(abi.fields_from(a)), C.byref(cp.p), 0, 1, ptr(cp.nb), ptr(wh))
    exp = dict(divv=3.3760353440920682e-2, curlv=3.7836647734377962e-2, dV11=0.0013578323369918166,
               dV12=0.02465266861727711, dV13=-0.0046604174274769167, dV22=0.022556438947324862,
               dV23=0.0097704904179710741, dV33=0.0098460821566040066)
    for k, v in exp.items():
        assert a[k][0] == pytest.approx(v, abs=2e-9), k

    for clean, ref in ((True, (-505548.68073726865, 303384.91384746187, -1767463.9739728321, 8.5525242525359648e12)),
                       (False, (-521261.07791667967, -74471.016515749841, -1730426.827721074, 7.1838438980436924e12))):
        a, cp, d = fresh()
        maxvs = np.zeros(99)
        cp.momentum_energy_ve(clean, maxvs)
        tol = (0.023, 0.053, 0.043, 7.1e5) if clean else (0.022, 0.064, 0.042, 3.1e5)
        assert a["ax"][0] == pytest.approx(ref[0], abs=tol[0])
        assert a["ay"][0] == pytest.approx(ref[1], abs=tol[1])
        assert a["az"][0] == pytest.approx(ref[2], abs=tol[2])
        assert a["du"][0] == pytest.approx(ref[3], abs=tol[3])
        assert maxvs[0] == pytest.approx(26490876.319252387, abs=1e-6)

    a, cp, d = fresh()
    cp.ve_def_gradh()
    density = a["kx"][0] * a["m"][0] / a["xm"][0]
    assert density == pytest.approx(3.4662283566584293e1, abs=8e-7)
    assert a["gradh"][0] == pytest.approx(0.98699067585409861, abs=5e-7)
    assert a["kx"][0] == pytest.approx(1.0042661134076782, abs=3e-7)

    a, cp, d = fresh()
    xm0 = a["xm"].copy()
    cp.xmass()
    rho0 = a["m"][0] / a["xm"][0]
    assert rho0 == pytest.approx(34.515038498081417, abs=7.33e-7)
    assert a["xm"][0] == pytest.approx(a["m"][0] / d["rho0"][0], rel=1e-7)
    assert xm0[1] == a["xm"][1]


# ---- bitwise against the compiled reference ------------------------------------------------------
@pytest.mark.parametrize("form", ["std", "ve", "ve_clean"])
@pytest.mark.parametrize("hydro", [np.float64, np.float32])
def test_oracle_equals_reference_pass(orc, ref, form, hydro):
    from oracle import pipeline
    from sphexa_amd.particles import ideal_gas_cv

    if form == "std":
        arr, spec = ic.sedov_lattice(14, seed=5)
        arr["vx"] = 0.1 * np.sin(6.0 * arr["x"])
    else:
        arr, spec = ic.turbulence_box(14, seed=6)
        arr["temp"] = arr.pop("u") / ideal_gas_cv(10.0, 5.0 / 3.0)
        arr["alpha"] = np.random.default_rng(1).uniform(0.05, 1.0, arr["x"].size)
    box = abi.Box.make(spec["lo"], spec["hi"], abi.PERIODIC)
    ht = abi.SPHX_F32 if hydro == np.float32 else abi.SPHX_F64
    p = abi.Params.defaults(orc.kernel_K(6.0), hydro_type=ht, minDt=1e-3)
    kw = dict(form="std" if form == "std" else "ve", av_clean=form == "ve_clean", hydro=hydro)
    o = pipeline.run(orc, arr, box, p, **kw)
    r = pipeline.run(ref, arr, box, p, **kw)
    assert np.array_equal(o["keys"], r["keys"]) and np.array_equal(o["order"], r["order"])
    assert np.array_equal(o["a"]["nc"], r["a"]["nc"])
    assert np.array_equal(o["nbrs"], r["nbrs"])
    outs = pipeline.STD_OUT if form == "std" else pipeline.VE_OUT + (pipeline.GRADV if form == "ve_clean" else ())
    for k in outs:
        assert np.array_equal(o["a"][k], r["a"][k]), k
    assert o["minDt"] == r["minDt"]


def _eos_inputs(form, hydro, n=4000, seed=21):
    rng = np.random.default_rng(seed)
    a = {"rho": rng.uniform(0.2, 5.0, n).astype(hydro), "p": np.zeros(n, hydro), "c": np.full(n, -1.0, hydro),
         "temp": rng.uniform(1.0, 2.0, n)}
    if form == "ve":
        a.update(m=rng.uniform(0.5, 1.5, n).astype(hydro) * hydro(1e-4), kx=rng.uniform(0.5, 2.0, n).astype(hydro),
                 xm=rng.uniform(0.5, 2.0, n).astype(hydro) * hydro(1e-4), gradh=rng.uniform(0.3, 1.2, n).astype(hydro),
                 prho=np.zeros(n, hydro))
        a["rho"] = np.zeros(n, hydro)
    return a


def _run_eos(lib, form, hydro, eos):
    a = _eos_inputs(form, hydro)
    ht = abi.SPHX_F32 if hydro == np.float32 else abi.SPHX_F64
    p = abi.Params.defaults(1.0, hydro_type=ht, eos=eos, soundSpeedConst=0.7, polytropic_const=2.5,
                            polytropic_index=1.4)
    f = abi.fields_from(a)
    (lib.eos_std if form == "std" else lib.eos_ve)(C.byref(f), C.byref(p), 0, a["temp"].size)
    return a, p


@pytest.mark.parametrize("form", ["std", "ve"])
@pytest.mark.parametrize("eos", [abi.SPHX_EOS_ISOTHERMAL, abi.SPHX_EOS_POLYTROPIC])
def test_eos_variants_closed_form(orc, form, eos):
    """The isothermal (p = rho c0^2, c = c0) and polytropic (p = K rho^n, c = sqrt(n p / rho)) EOS of
    sph/include/sph/eos.hpp:58-86 in FP64, with rho = kx m / xm and prho = p / (kx m^2 gradh) for VE;
    temp is zeroed."""
    a, p = _run_eos(orc, form, np.float64, eos)
    rho = a["kx"] * a["m"] / a["xm"] if form == "ve" else a["rho"]
    if eos == abi.SPHX_EOS_ISOTHERMAL:
        pe, ce = rho * 0.7 * 0.7, np.full(rho.size, 0.7)
    else:
        pe = 2.5 * rho ** 1.4
        ce = np.sqrt(1.4 * pe / rho)
    tol = 4e-16 if eos == abi.SPHX_EOS_ISOTHERMAL else 1e-15  # pow, then a division: a few ulp
    np.testing.assert_allclose(a["c"], ce, rtol=tol, atol=0)
    if form == "ve":
        np.testing.assert_allclose(a["prho"], pe / (a["kx"] * a["m"] * a["m"] * a["gradh"]), rtol=tol, atol=0)
        np.testing.assert_allclose(a["rho"], rho, rtol=0, atol=0)
    else:
        np.testing.assert_allclose(a["p"], pe, rtol=tol, atol=0)
    assert not a["temp"].any()


@pytest.mark.parametrize("form", ["std", "ve"])
@pytest.mark.parametrize("eos", [abi.SPHX_EOS_ISOTHERMAL, abi.SPHX_EOS_POLYTROPIC])
@pytest.mark.parametrize("hydro", [np.float64, np.float32])
def test_eos_variants_equal_reference(orc, ref, form, eos, hydro):
    """Oracle == the reference's own CPU loops (computeEOS_HydroStd / computeEOS branches), bit for bit.
    The reference's std isothermal CPU loop leaves c unchanged (hydro_std/eos.hpp:82-97); the oracle, like
    the reference's GPU kernel (hydro_std/eos_gpu.cu:73-82), stores c = c0 there."""
    o, p = _run_eos(orc, form, hydro, eos)
    r, _ = _run_eos(ref, form, hydro, eos)
    names = ("p", "c", "temp") if form == "std" else ("prho", "c", "rho", "temp")
    for k in names:
        if form == "std" and eos == abi.SPHX_EOS_ISOTHERMAL and k == "c":
            assert np.all(r["c"] == hydro(-1.0)) and np.all(o["c"] == hydro(p.soundSpeedConst))
            continue
        assert np.array_equal(o[k], r[k]), k


def _groups_for(orc, x, y, z, box, first=0, last=None, tol=2.0):
    from oracle import oracle as O

    x, y, z = (np.ascontiguousarray(c, np.float64) for c in (x, y, z))
    keys = O.sfc_keys(orc, box, x, y, z)
    sk, order = O.sort_keys(orc, keys)
    xs, ys, zs = (np.ascontiguousarray(c[order]) for c in (x, y, z))
    tree = O.Tree(orc, sk, box)
    last = xs.size if last is None else last
    return O.group_splits(orc, tree, box, xs, ys, zs, first, last, tol), (xs, ys, zs), tree


def test_group_splits_lattice_has_no_splits(orc):
    """On a perfect lattice consecutive particles in SFC order are lattice neighbours, far closer than
    twice the smallest leaf side: the groups are the fixed groups of 64 (computeGroupSplits,
    domain/include/cstone/traversal/groups_gpu.cuh:149-216)."""
    side = 16
    g1 = (np.arange(side) + 0.5) / side - 0.5
    x, y, z = (c.ravel().copy() for c in np.meshgrid(g1, g1, g1, indexing="ij"))
    box = abi.Box.make(-0.5, 0.5, abi.PERIODIC)
    g, _, _ = _groups_for(orc, x, y, z, box)
    np.testing.assert_array_equal(g, np.append(np.arange(0, x.size, 64), x.size))
    g, _, _ = _groups_for(orc, x, y, z, box, first=10, last=1000)
    np.testing.assert_array_equal(g, np.append(np.arange(10, 1000, 64), 1000))


def test_group_splits_at_cluster_gaps(orc):
    """Three small lattice clusters (5^3 particles, spacing 0.002) far apart: each holds more than a
    bucket, so its leaves are a few spacings wide, and consecutive particles inside a cluster stay within
    twice the smallest leaf side while the gaps between clusters are hundreds of times wider. Every
    fixed group is split exactly where the SFC order passes from one cluster to the next."""
    g1 = (np.arange(5) - 2.0) * 0.002
    cube = np.stack([c.ravel() for c in np.meshgrid(g1, g1, g1, indexing="ij")], 1)
    centres = np.array([[0.1, 0.1, 0.1], [0.85, 0.2, 0.7], [0.3, 0.9, 0.6]])
    pts = np.concatenate([c + cube for c in centres])
    box = abi.Box.make(0.0, 1.0, abi.OPEN)
    g, (xs, ys, zs), _ = _groups_for(orc, pts[:, 0], pts[:, 1], pts[:, 2], box)
    n = xs.size
    which = np.argmin(((np.stack([xs, ys, zs], 1)[:, None, :] - centres[None]) ** 2).sum(2), 1)
    changes = 1 + np.nonzero(which[1:] != which[:-1])[0]
    expected = np.unique(np.concatenate([np.arange(0, n, 64), changes, [n]]))
    np.testing.assert_array_equal(g, expected)


def test_group_splits_invariants(orc):
    """On the uneven Evrard-like sphere: starts increase, every fixed-group start is a group start, no
    group holds more than 64 particles, the last entry is the range end, and a split leaves a gap wider
    than the split distance."""
    arr, spec = ic.evrard_sphere(20000, seed=4)
    box = abi.Box.make(spec["lo"], spec["hi"], abi.OPEN)
    g, (xs, ys, zs), tree = _groups_for(orc, arr["x"], arr["y"], arr["z"], box)
    n = xs.size
    assert g[0] == 0 and g[-1] == n and np.all(np.diff(g) > 0) and np.max(np.diff(g)) <= 64
    assert set(range(0, n, 64)) <= set(g.tolist())
    assert len(g) - 1 > (n + 63) // 64  # the sphere's outskirts split
    L = 2.0
    for s in set(g[1:-1].tolist()) - set(range(0, n, 64)):
        d = np.sqrt((xs[s] - xs[s - 1]) ** 2 + (ys[s] - ys[s - 1]) ** 2 + (zs[s] - zs[s - 1]) ** 2) / L
        assert d > 2.0 * 2.0 ** -21  # at least twice the finest possible leaf


def test_oracle_tree_equals_reference(orc, ref):
    from oracle import oracle as O

    for gaussian in (False, True):
        t1, keys = _tree_for(orc, 30000, gaussian, seed=9)
        t2 = O.Tree(ref, keys, abi.Box.make(0.0, 1.0, abi.OPEN))
        for f in ("leaves", "counts", "layout", "prefixes", "childOffsets", "parents", "levelRange", "centers",
                  "sizes"):
            assert np.array_equal(getattr(t1, f), getattr(t2, f)), f
        leaf = t1.internalToLeaf >= 0
        assert np.array_equal(t1.internalToLeaf[leaf], t2.internalToLeaf[leaf])


# ---- golden vectors written from the compiled reference (tests/golden/make_golden.py) -----------
GOLDEN_CASES = {
    "sedov10_std_f64": ("std", False, np.float64, True),
    "sedov10_std_f32": ("std", False, np.float32, True),
    "turb10_ve_f64": ("ve", False, np.float64, True),
    "turb10_veclean_f64": ("ve", True, np.float64, True),
    "turb10_ve_f32": ("ve", False, np.float32, True),
    "evrard1k_search_f64": ("search", False, np.float64, False),
}


@pytest.mark.parametrize("name", sorted(GOLDEN_CASES))
def test_oracle_matches_golden(orc, name):
    from oracle import pipeline

    form, clean, hydro, periodic = GOLDEN_CASES[name]
    g = np.load(os.path.join(GOLDEN, name + ".npz"))
    arr = {k[3:]: g[k] for k in g.files if k.startswith("in_")}
    if periodic:
        box = abi.Box.make(-0.5, 0.5, abi.PERIODIC)
    else:
        box = abi.Box.make(-1.0, 1.0, abi.OPEN)
    ht = abi.SPHX_F32 if hydro == np.float32 else abi.SPHX_F64
    p = abi.Params.defaults(orc.kernel_K(6.0), hydro_type=ht, minDt=1e-3)
    r = pipeline.run(orc, arr, box, p, form=form, av_clean=clean, hydro=hydro)
    assert np.array_equal(r["keys"], g["keys"]) and np.array_equal(r["order"], g["order"])
    t = r["tree"]
    for f in ("leaves", "counts", "layout", "prefixes", "childOffsets", "levelRange", "centers", "sizes"):
        assert np.array_equal(getattr(t, f), g["tree_" + f]), f
    nc = r["a"]["nc"]
    assert np.array_equal(nc, g["nc"]) and np.array_equal(r["a"]["h"], g["h_final"])
    counts = np.minimum(nc.astype(np.int64) - 1, p.ngmax)
    flat = np.concatenate([r["nbrs"][i, :counts[i]] for i in range(nc.size)])
    assert np.array_equal(flat, g["nbrs_flat"])
    outs = {"std": pipeline.STD_OUT, "ve": pipeline.VE_OUT + (pipeline.GRADV if clean else ()), "search": ()}[form]
    for k in outs:
        assert np.array_equal(r["a"][k], g[k]), k
    if "minDt" in g.files:
        assert r["minDt"] == float(g["minDt"])
