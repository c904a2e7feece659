"""Fields and attributes in the H5Part layout (include/sphexa_amd_io.h, sphexa_amd/io/h5part.cpp) against
the reference's own H5Part (extern/h5part/H5Part.c built into oracle/_ref/libh5part_ref.so) -- CPU only.

* A file the library writes (two steps, fields of every element type, step and file attributes)
  reads back through the reference's H5Part with the same steps, particle counts, dataset names,
  values (bit for bit) and attribute types and values, and the other way round.
* Appending adds Step#k after the existing k steps (H5PartWriter::addStep on an existing path).
* Reading converts field types through HDF5 (H5PartReadData*), and attributes follow readAttribute
  (h5part_wrapper.hpp:183-222): the sizes must match, floating types convert both ways, a type
  converts to a wider one; anything else, a missing attribute, field or step is an error.
* The C ABI exports every symbol the header declares."""
import os
import re
import subprocess

import numpy as np
import pytest

from sphexa_amd import h5io

import h5part_ref as R

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

pytestmark = pytest.mark.skipif(not os.path.exists(h5io.LIB_PATH), reason="libsphexa_amd_io.so not built (no HDF5)")
needs_ref = pytest.mark.skipif(not R.available(), reason="oracle/_ref/libh5part_ref.so not built (make -C oracle h5ref)")

FIELDS = {"x": np.float64, "h": np.float32, "nc": np.int32, "id": np.int64, "rung": np.int8}


def _fields(rng, n, shift=0.0):
    return {"x": rng.uniform(-0.5, 0.5, n) + shift, "h": rng.uniform(0.01, 0.03, n).astype(np.float32),
            "nc": rng.integers(0, 150, n).astype(np.int32), "id": rng.integers(0, 1 << 40, n).astype(np.int64),
            "rung": rng.integers(0, 4, n).astype(np.int8)}


def _attribs(k):
    return {"time": np.array([0.25 * k]), "iteration": np.array([k], np.int64), "gamma": np.array([5 / 3], np.float32),
            "numParticlesGlobal": np.array([1000 + k], np.int64), "box": np.array([-0.5, 0.5, -0.5, 0.5, 0.0, 1.0]),
            "eosChoice": np.array([2], np.int8)}


def test_exports_every_declared_symbol():
    names = re.findall(r"\b(sphx_h5_\w+)\s*\(", open(os.path.join(ROOT, "include", "sphexa_amd_io.h")).read())
    out = subprocess.run(["nm", "-D", "--defined-only", h5io.LIB_PATH], capture_output=True, text=True).stdout
    for name in set(names):
        assert re.search(rf"\bT {name}\b", out), name


@needs_ref
def test_library_file_reads_with_reference(tmp_path):
    rng = np.random.default_rng(1)
    path = tmp_path / "lib.h5"
    steps = [(_fields(rng, 1000), _attribs(0)), (_fields(rng, 777, 1.0), _attribs(1))]
    with h5io.H5File(path, h5io.WRITE) as f:
        f.write_attrib("Kcour", 0.2, step=False)
        f.write_attrib("sincIndex", np.array([6.0, 1.0]), step=False)
        for flds, attrs in steps:
            f.add_step(len(flds["x"]))
            for k, v in attrs.items():
                f.write_attrib(k, v)
            for k, v in flds.items():
                f.write_field(k, v)
    r = R.RefH5Part(path, R.H5PART_READ)
    assert r.num_steps() == 2
    fa = r.attribs(step=False)
    assert fa["Kcour"] == (r.type_of(np.float64), 1) and fa["sincIndex"][1] == 2
    assert np.array_equal(r.read_attrib("sincIndex", np.float64, 2, step=False), [6.0, 1.0])
    for s, (flds, attrs) in enumerate(steps):
        r.set_step(s)
        assert r.num_particles() == len(flds["x"])
        assert sorted(r.dataset_names()) == sorted(flds)
        for k, v in flds.items():
            assert np.array_equal(r.read(k, v.dtype), v), k
        info = r.attribs()
        assert sorted(info) == sorted(attrs)
        for k, v in attrs.items():
            assert info[k] == (r.type_of(v.dtype), v.size), k
            assert np.array_equal(r.read_attrib(k, v.dtype, v.size), v), k
    r.close()


@needs_ref
def test_reference_file_reads_with_library(tmp_path):
    rng = np.random.default_rng(2)
    path = tmp_path / "ref.h5"
    steps = [(_fields(rng, 500), _attribs(0)), (_fields(rng, 64), _attribs(1)), (_fields(rng, 1), _attribs(2))]
    w = R.RefH5Part(path, R.H5PART_WRITE)
    w.write_attrib("muiConst", np.array([10.0]), step=False)
    for s, (flds, attrs) in enumerate(steps):
        w.set_step(s)
        w.set_num_particles(len(flds["x"]))
        for k, v in attrs.items():
            w.write_attrib(k, v)
        for k, v in flds.items():
            w.write(k, v)
    w.close()
    with h5io.H5File(path) as f:
        assert f.num_steps() == 3
        assert f.file_attributes() == ["muiConst"] and f.read_attrib("muiConst", step=False)[0] == 10.0
        for s, (flds, attrs) in enumerate(steps):
            assert f.set_step(s) == len(flds["x"])
            assert sorted(f.fields()) == sorted(flds) and sorted(f.step_attributes()) == sorted(attrs)
            for k, v in flds.items():
                assert np.array_equal(f.read_field(k, v.dtype), v), k
            for k, v in attrs.items():
                assert f.attrib_info(k) == (h5io._FROM_NUMPY[v.dtype], v.size)
                assert np.array_equal(f.read_attrib(k), v), k
        # a negative step is the last one (H5PartReader::setStep)
        assert f.set_step(-1) == 1


@needs_ref
def test_append_adds_the_next_step(tmp_path):
    rng = np.random.default_rng(3)
    path = tmp_path / "append.h5"
    a, b = _fields(rng, 200), _fields(rng, 300)
    with h5io.H5File(path, h5io.WRITE) as f:
        f.add_step(200)
        f.write_field("x", a["x"])
    with h5io.H5File(path, h5io.APPEND) as f:
        assert f.num_steps() == 1
        f.add_step(300)
        f.write_field("x", b["x"])
        f.write_attrib("time", 1.0)
    r = R.RefH5Part(path, R.H5PART_READ)
    assert r.num_steps() == 2
    r.set_step(1)
    assert r.num_particles() == 300 and np.array_equal(r.read("x", np.float64), b["x"])
    r.close()
    # appending with the reference's H5Part after the library's steps
    w = R.RefH5Part(path, R.H5PART_APPEND)
    w.set_step(2)
    w.set_num_particles(5)
    w.write("x", np.arange(5.0))
    w.close()
    with h5io.H5File(path) as f:
        assert f.num_steps() == 3 and f.set_step(2) == 5
        assert np.array_equal(f.read_field("x"), np.arange(5.0))


def test_conversions_and_errors(tmp_path):
    path = tmp_path / "conv.h5"
    x = np.linspace(0, 1, 50)
    with h5io.H5File(path, h5io.WRITE) as f:
        f.add_step(50)
        f.write_field("x", x)
        f.write_field("n", np.arange(50, dtype=np.int32))
        f.write_attrib("time", 0.1)
        f.write_attrib("gamma", np.float32(1.4))
        f.write_attrib("iteration", np.int32(7))
        f.write_attrib("count", np.int64(1 << 40))
    with h5io.H5File(path) as f:
        f.set_step(0)
        # fields convert through HDF5, like H5PartReadDataFloat32 on a double dataset
        assert np.array_equal(f.read_field("x", np.float32), x.astype(np.float32))
        assert np.array_equal(f.read_field("n", np.float64), np.arange(50.0))
        # attributes: floating both ways, widening, nothing else
        assert f.read_attrib("time", np.float32)[0] == np.float32(0.1)
        assert f.read_attrib("gamma", np.float64)[0] == np.float64(np.float32(1.4))
        assert f.read_attrib("iteration", np.int64)[0] == 7
        assert f.read_attrib("iteration", np.float64)[0] == 7.0
        assert f.read_attrib("count", np.float64)[0] == float(1 << 40)
        for name, dt in (("time", np.int64), ("count", np.int32), ("gamma", np.int32)):
            with pytest.raises(h5io.H5Error) as e:
                f.read_attrib(name, dt)
            assert e.value.code == h5io.ERR_TYPE, name
        with pytest.raises(h5io.H5Error) as e:
            f.read_attrib("time", np.float64, count=2)
        assert e.value.code == h5io.ERR_TYPE and "size mismatch" in str(e.value)
        for call in (lambda: f.read_attrib("minDt"), lambda: f.read_field("vx"), lambda: f.set_step(4)):
            with pytest.raises(h5io.H5Error) as e:
                call()
            assert e.value.code == h5io.ERR_MISSING
        with pytest.raises(h5io.H5Error) as e:
            f.add_step(10)
        assert e.value.code == h5io.ERR_ARG
    with pytest.raises(h5io.H5Error) as e:
        h5io.H5File(tmp_path / "missing.h5")
    assert e.value.code == h5io.ERR_MISSING


@needs_ref
def test_cpp_interface_checkpoint(tmp_path):
    """tests/cpp/h5_api: a Dataset's attributes (loadOrStoreAttributes, particles_data.hpp:147-205)
    and fields through H5PartWriter / H5PartReader of include/sphexa_amd_io.hpp, twice into one file;
    the reference's H5Part then finds the two steps and the attribute set under the reference's names
    and types."""
    drv = os.path.join(ROOT, "tests", "cpp", "h5_api")
    assert os.path.exists(drv), "tests/cpp/h5_api not built (python -m sphexa_amd.build)"
    path = tmp_path / "ckpt.h5"
    r = subprocess.run([drv, str(path)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
    ref = R.RefH5Part(path, R.H5PART_READ)
    assert ref.num_steps() == 2
    ref.set_step(1)
    info = ref.attribs()
    expect = {"iteration": np.int64, "numParticlesGlobal": np.int64, "ng0": np.int32, "ngmax": np.int32,
              "time": np.float64, "minDt": np.float64, "minDt_m1": np.float64, "gravConstant": np.float64,
              "gamma": np.float64, "eosChoice": np.int32, "soundSpeedConst": np.float32, "sincIndex": np.float64}
    for k, t in expect.items():
        assert info[k] == (ref.type_of(t), 1), k
    assert ref.read_attrib("iteration", np.int64, 1)[0] == 43 and ref.read_attrib("time", np.float64, 1)[0] == 0.125
    assert sorted(ref.dataset_names()) == ["h", "nc", "temp", "x"] and ref.num_particles() == 1000
    ref.close()


@needs_ref
def test_template_block_and_settings_from_reference_file(tmp_path):
    """readTemplateBlock and readFileAttributes (main/src/init/utils.hpp:74-121) on a file the
    reference's H5Part wrote: the block is x, y, z of the last step, read as double from float
    datasets too; the settings are the one-value file attributes as doubles, overriding defaults,
    and an attribute of several values is left out."""
    rng = np.random.default_rng(4)
    path = tmp_path / "glass.h5"
    blocks = [rng.uniform(-0.5, 0.5, (3, 100)), rng.uniform(-0.5, 0.5, (3, 125)).astype(np.float32)]
    w = R.RefH5Part(path, R.H5PART_WRITE)
    w.write_attrib("gamma", np.array([1.4]), step=False)
    w.write_attrib("ng0", np.array([120], np.int32), step=False)
    w.write_attrib("eosChoice", np.array([1], np.int8), step=False)
    w.write_attrib("box", np.array([-0.5, 0.5]), step=False)
    for s, b in enumerate(blocks):
        w.set_step(s)
        w.set_num_particles(b.shape[1])
        for k, v in zip("xyz", b):
            w.write(k, v)
    w.close()
    x, y, z = h5io.read_template_block(path)
    for got, want in zip((x, y, z), blocks[1]):
        assert got.dtype == np.float64 and np.array_equal(got, want.astype(np.float64))
    settings = h5io.read_file_attributes(path, {"gamma": 5 / 3, "Kcour": 0.2})
    assert settings == {"gamma": 1.4, "ng0": 120.0, "eosChoice": 1.0, "Kcour": 0.2}


@needs_ref
def test_long_names_and_write_errors(tmp_path):
    """attribute names of any length list and read back; a field name over 64 characters is stored
    as its first 63, on writing and reading alike (H5Part.c:906-921), so the reference's H5Part reads
    it under the full name too. Writing in a file opened for reading, or over a dataset of the step
    with another particle count, is an error."""
    path = tmp_path / "names.h5"
    long_field, long_attr = "f" * 300, "a" * 280
    with h5io.H5File(path, h5io.WRITE) as f:
        f.add_step(10)
        f.write_field(long_field, np.arange(10.0))
        f.write_attrib(long_attr, 2.5)
        f.write_attrib("file_" + "b" * 260, np.int32(3), step=False)
    with h5io.H5File(path) as f:
        f.set_step(0)
        assert f.fields() == [long_field[:63]] and f.step_attributes() == [long_attr]
        assert f.file_attributes() == ["file_" + "b" * 260]
        assert np.array_equal(f.read_field(long_field), np.arange(10.0))
        with pytest.raises(h5io.H5Error) as e:
            f.write_field("x", np.zeros(10))
        assert e.value.code == h5io.ERR_ARG
    r = R.RefH5Part(path, R.H5PART_READ)
    r.set_step(0)
    assert np.array_equal(r.read(long_field, np.float64), np.arange(10.0))
    r.close()
    with h5io.H5File(path, h5io.APPEND) as f:
        f.add_step(4)
        f.write_field("x", np.zeros(4))
        f.set_step(0)  # back to the first step, whose datasets hold 10 particles
        f.write_field(long_field, np.ones(10))
    # a step whose datasets differ in length (H5PartSetNumParticles between writes): the step's count
    # is its first dataset's, and writing over a dataset of another length is refused
    mixed = tmp_path / "ragged.h5"
    w = R.RefH5Part(mixed, R.H5PART_WRITE)
    w.set_step(0)
    w.set_num_particles(4)
    w.write("x", np.zeros(4))
    w.set_num_particles(6)
    w.write("y", np.zeros(6))
    w.close()
    with h5io.H5File(mixed, h5io.APPEND) as f:
        assert f.set_step(0) == 4
        f.write_field("x", np.ones(4))
        with pytest.raises(h5io.H5Error) as e:
            f.write_field("y", np.ones(4))
        assert e.value.code == h5io.ERR_ARG and "another particle count" in str(e.value)
        with pytest.raises(h5io.H5Error) as e:
            f.read_field("y")
        assert e.value.code == h5io.ERR_ARG
