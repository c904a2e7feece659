"""C ABI and host-side logic -- CPU only (no kernel is launched)."""
import ctypes as C
import math
import os
import re

import numpy as np
import pytest

from sphexa_amd import abi, ic

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "sphexa_amd.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(sphx_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    from sphexa_amd import lib

    L = lib.load()
    names = declared_symbols()
    assert len(names) >= 20
    for name in names:
        assert hasattr(L, name), f"{name} declared in include/sphexa_amd.h but not exported"
    assert sorted(lib.EXPORTED) == names


def test_struct_layouts_match_header():
    """The ctypes mirrors have the field order and sizes of the C structs."""
    text = open(HEADER).read()
    body = text[text.index("typedef struct sphx_fields"):text.index("} sphx_fields_t;")]
    names = re.findall(r"\*\s*(\w+);", body)
    assert tuple(names) == abi.ALL_FIELDS
    assert C.sizeof(abi.Fields) == 8 * len(abi.ALL_FIELDS)
    assert C.sizeof(abi.Box) == 6 * 8 + 3 * 4 + 4  # padded to 8
    assert C.sizeof(abi.Params) == 11 * 8 + 6 * 4 + 3 * 8
    assert C.sizeof(abi.TreeView) == 4 * 4 + 10 * 8
    body = text[text.index("typedef struct sphx_adjoint"):text.index("} sphx_adjoint_t;")]
    names = tuple(re.findall(r"\*?(\w+)[,;]", body.split("{", 1)[1]))
    assert names == abi.ADJOINT_FIELDS
    assert C.sizeof(abi.Adjoint) == 8 * len(abi.ADJOINT_FIELDS)
    assert "int32_t  flags;" in text and abi.Params.flags.offset == 11 * 8 + 3 * 4
    # the params' EOS members follow flags; the group view is 4 uint32 then 3 pointers
    names = re.findall(r"^\s*(?:double|int32_t|uint32_t)\s+([\w, ]+);",
                       text[text.index("typedef struct sphx_params"):text.index("} sphx_params_t;")], re.M)
    flat = [n.strip() for group in names for n in group.split(",")]
    assert flat == [f[0] for f in abi.Params._fields_]
    body = text[text.index("typedef struct sphx_groups"):text.index("} sphx_groups_t;")]
    decls = re.findall(r"^\s*(?:const uint32_t\*|float\*|uint32_t)\s+([\w, ]+);", body, re.M)
    assert [n.strip() for d in decls for n in d.split(",")] == [f[0] for f in abi.Groups._fields_]
    assert C.sizeof(abi.Groups) == 4 * 4 + 3 * 8


def test_null_workspace_is_rejected_without_touching_the_gpu():
    from sphexa_amd import lib

    L = lib.load()
    box = abi.Box.make(0, 1)
    f = abi.Fields()
    p = abi.Params.defaults(1.0)
    assert L.sphx_find_neighbors(None, C.byref(box), C.byref(f), C.byref(p), 0, 0, 1) == abi.SPHX_ERR_ARG
    assert L.sphx_density_std(None, C.byref(box), C.byref(f), C.byref(p), 0, 0) == abi.SPHX_ERR_ARG
    assert L.sphx_workspace_destroy(None) == abi.SPHX_OK
    assert L.sphx_last_error(None) == b"null workspace"
    assert L.sphx_workspace_create(None, 0, None) == abi.SPHX_ERR_ARG


def test_product_does_not_link_the_oracle():
    """The HIP library must not depend on the CPU oracle (no fallback path)."""
    import subprocess

    from sphexa_amd import lib

    out = subprocess.run(["ldd", lib.LIB_PATH], capture_output=True, text=True).stdout
    assert "oracle" not in out and "sphx_ref" not in out
    syms = subprocess.run(["nm", "-D", lib.LIB_PATH], capture_output=True, text=True).stdout
    assert " orc_" not in syms and " ref_" not in syms


# ---- initial conditions ------------------------------------------------------------------------
def test_sedov_ic():
    arr, spec = ic.sedov_lattice(12, seed=3)
    n = 12 ** 3
    assert arr["x"].size == n and spec["periodic"]
    for k in "xyz":
        assert arr[k].min() > -0.5 and arr[k].max() < 0.5
    assert np.isclose(arr["m"].sum(), 1.0)
    assert np.isclose(np.sum(arr["m"] * (arr["u"] - 1e-5)), 1.0)
    h0 = 0.5 * (3 * 100 / (4 * math.pi * n)) ** (1 / 3)
    assert np.allclose(arr["h"], h0)
    a2, _ = ic.sedov_lattice(12, seed=3)
    assert np.array_equal(arr["x"], a2["x"])  # seeded


def test_turbulence_ic_is_solenoidal_at_mach_03():
    arr, spec = ic.turbulence_box(16, seed=5, jitter=0.0)
    cs = math.sqrt(5 / 3 * 2 / 3 * 1.0)
    vrms = math.sqrt(np.mean(arr["vx"] ** 2 + arr["vy"] ** 2 + arr["vz"] ** 2))
    assert abs(vrms / cs - 0.3) < 1e-12
    # spectral divergence on the unjittered lattice vanishes
    s = 16
    v = [arr[k].reshape(s, s, s) for k in ("vx", "vy", "vz")]
    k = np.fft.fftfreq(s, 1.0 / s)
    kx, ky, kz = np.meshgrid(k, k, k, indexing="ij")
    div = kx * np.fft.fftn(v[0]) + ky * np.fft.fftn(v[1]) + kz * np.fft.fftn(v[2])
    power = np.abs(np.fft.fftn(v[0])) ** 2 + np.abs(np.fft.fftn(v[1])) ** 2
    assert np.max(np.abs(div)) < 1e-9 * np.sqrt(power.max())
    # only modes with 1 <= |k| <= 2 are populated
    kk = np.sqrt(kx ** 2 + ky ** 2 + kz ** 2)
    p = np.abs(np.fft.fftn(v[0])) ** 2
    assert p[(kk < 1) | (kk > 2.001)].max() < 1e-20 * p.max()


def test_evrard_ic_density_profile():
    arr, spec = ic.evrard_sphere(200000, seed=1)
    r = np.sqrt(arr["x"] ** 2 + arr["y"] ** 2 + arr["z"] ** 2)
    assert r.max() <= 1.0 and not spec["periodic"]
    # M(<r) = r^2 for rho ~ 1/r
    for q in (0.25, 0.5, 0.75):
        assert abs(np.mean(r < q) - q * q) < 5e-3
