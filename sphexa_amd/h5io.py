"""ctypes binding of libsphexa_amd_io.so (include/sphexa_amd_io.h): particle fields and attributes in
the H5Part layout of HDF5 files, the format of the reference's H5PartWriter / H5PartReader
(main/src/io/ifile_io_hdf5.cpp). Fields are numpy arrays (host) or torch tensors (host or device)."""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("SPHEXA_AMD_IO_LIB") or os.path.join(HERE, "libsphexa_amd_io.so")

READ, WRITE, APPEND = 1, 2, 4
F64, F32, I32, I64, I8 = 0, 1, 2, 3, 4
OK, ERR_ARG, ERR_IO, ERR_MISSING, ERR_TYPE = 0, -1, -2, -3, -4

NUMPY = {F64: np.float64, F32: np.float32, I32: np.int32, I64: np.int64, I8: np.int8}
_FROM_NUMPY = {np.dtype(np.float64): F64, np.dtype(np.float32): F32, np.dtype(np.int32): I32,
               np.dtype(np.uint32): I32, np.dtype(np.int64): I64, np.dtype(np.uint64): I64,
               np.dtype(np.int8): I8, np.dtype(np.uint8): I8}

_P = C.c_void_p
_SIGS = {
    "sphx_h5_last_error": (C.c_char_p, []),
    "sphx_h5_open": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(_P)]),
    "sphx_h5_close": (C.c_int, [_P]),
    "sphx_h5_num_steps": (C.c_int, [_P, C.POINTER(C.c_int64)]),
    "sphx_h5_add_step": (C.c_int, [_P, C.c_uint64]),
    "sphx_h5_set_step": (C.c_int, [_P, C.c_int64, C.POINTER(C.c_uint64)]),
    "sphx_h5_write_field": (C.c_int, [_P, C.c_char_p, C.c_int, _P]),
    "sphx_h5_read_field": (C.c_int, [_P, C.c_char_p, C.c_int, _P]),
    "sphx_h5_write_attrib": (C.c_int, [_P, C.c_int, C.c_char_p, C.c_int, _P, C.c_uint64]),
    "sphx_h5_read_attrib": (C.c_int, [_P, C.c_int, C.c_char_p, C.c_int, _P, C.c_uint64]),
    "sphx_h5_count": (C.c_int, [_P, C.c_int, C.c_int, C.POINTER(C.c_int64)]),
    "sphx_h5_name": (C.c_int, [_P, C.c_int, C.c_int, C.c_int64, C.c_char_p, C.c_size_t]),
    "sphx_h5_attrib_info": (C.c_int, [_P, C.c_int, C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_uint64)]),
}

_lib = None


class H5Error(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} is not built (python -m sphexa_amd.build, needs HDF5)")
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name)
            fn.restype, fn.argtypes = res, args
        _lib = lib
    return _lib


def _check(code):
    if code != OK:
        raise H5Error(code, load().sphx_h5_last_error().decode())


def _buffer(arr):
    """(pointer, element type, count) of a contiguous numpy array or torch tensor"""
    if hasattr(arr, "data_ptr"):
        import torch

        assert arr.is_contiguous()
        dt = np.dtype(str(arr.dtype).replace("torch.", ""))
        return arr.data_ptr(), _FROM_NUMPY[dt], arr.numel()
    assert arr.flags.c_contiguous
    return arr.ctypes.data, _FROM_NUMPY[arr.dtype], arr.size


class H5File:
    """One open file; the writer side mirrors H5PartWriter (addStep, stepAttribute, fileAttribute,
    writeField), the reader side H5PartReader (setStep, stepAttributes, readField, ...)."""

    def __init__(self, path, mode=READ):
        lib = load()
        h = _P()
        _check(lib.sphx_h5_open(str(path).encode(), mode, C.byref(h)))
        self.h = h
        self.n = 0

    def close(self):
        if self.h:
            _check(load().sphx_h5_close(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def num_steps(self):
        v = C.c_int64()
        _check(load().sphx_h5_num_steps(self.h, C.byref(v)))
        return v.value

    def add_step(self, n):
        _check(load().sphx_h5_add_step(self.h, n))
        self.n = n

    def set_step(self, step=-1):
        n = C.c_uint64()
        _check(load().sphx_h5_set_step(self.h, step, C.byref(n)))
        self.n = n.value
        return self.n

    def write_field(self, name, arr):
        p, t, count = _buffer(arr)
        assert count == self.n, f"field {name}: {count} values for a step of {self.n} particles"
        _check(load().sphx_h5_write_field(self.h, name.encode(), t, p))

    def read_field(self, name, dtype=np.float64, out=None):
        """into `out` (numpy or torch, host or device) or a new numpy array of dtype"""
        if out is None:
            out = np.empty(self.n, dtype)
        p, t, count = _buffer(out)
        assert count == self.n
        _check(load().sphx_h5_read_field(self.h, name.encode(), t, p))
        return out

    def write_attrib(self, name, values, step=True):
        a = np.ascontiguousarray(np.atleast_1d(values))
        p, t, count = _buffer(a)
        _check(load().sphx_h5_write_attrib(self.h, int(step), name.encode(), t, p, count))

    def attrib_info(self, name, step=True):
        t, n = C.c_int(), C.c_uint64()
        _check(load().sphx_h5_attrib_info(self.h, int(step), name.encode(), C.byref(t), C.byref(n)))
        return t.value, n.value

    def read_attrib(self, name, dtype=None, count=None, step=True):
        t, n = self.attrib_info(name, step)
        out = np.empty(n if count is None else count, NUMPY[t] if dtype is None else dtype)
        p, tt, c = _buffer(out)
        _check(load().sphx_h5_read_attrib(self.h, int(step), name.encode(), tt, p, c))
        return out

    def names(self, kind, step=True):
        lib = load()
        k = C.c_int64()
        _check(lib.sphx_h5_count(self.h, int(step), kind, C.byref(k)))
        out = []
        for i in range(k.value):
            size = 256
            while True:
                buf = C.create_string_buffer(size)
                code = lib.sphx_h5_name(self.h, int(step), kind, i, buf, size)
                if code != ERR_ARG or size >= 1 << 20:
                    break
                size *= 4  # the name did not fit
            _check(code)
            out.append(buf.value.decode())
        return out

    def fields(self):
        return self.names(0)

    def step_attributes(self):
        return self.names(1)

    def file_attributes(self):
        return self.names(1, step=False)


def read_template_block(path):
    """x, y, z (float64) of the last step of `path`: readTemplateBlock (main/src/init/utils.hpp:74-87),
    how the reference's initialisers load a glass block."""
    with H5File(path) as f:
        f.set_step(-1)
        return tuple(f.read_field(k, np.float64) for k in ("x", "y", "z"))


def read_file_attributes(path, settings=None):
    """readFileAttributes (main/src/init/utils.hpp:90-121): every file attribute of `path` holding one
    value, read as a double into `settings` (a dict, overriding what it holds); returns `settings`."""
    settings = {} if settings is None else settings
    with H5File(path) as f:
        if f.num_steps() > 0:
            f.set_step(-1)
        for name in f.file_attributes():
            if f.attrib_info(name, step=False)[1] == 1:
                settings[name] = float(f.read_attrib(name, np.float64, step=False)[0])
    return settings
