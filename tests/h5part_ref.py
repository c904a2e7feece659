"""The reference's H5Part (extern/h5part/H5Part.c, built by `make -C oracle h5ref` into
oracle/_ref/libh5part_ref.so against the image's HDF5), driven through ctypes as the reference's
h5part_wrapper.hpp drives it: the checker of the library's H5Part-layout files."""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "_ref", "libh5part_ref.so")
HDF5_LIB = "/opt/conda/lib/libhdf5.so"

H5PART_READ, H5PART_WRITE, H5PART_APPEND = 0x01, 0x02, 0x04
_I64 = C.c_int64
_P = C.c_void_p


def available():
    return os.path.exists(LIB) and os.path.exists(HDF5_LIB)


class RefH5Part:
    """a file opened with H5PartOpenFile (serial, as openH5Part does on one rank)"""

    _lib = None
    _types = None

    @classmethod
    def lib(cls):
        if cls._lib is None:
            hdf5 = C.CDLL(HDF5_LIB, mode=C.RTLD_GLOBAL)
            hdf5.H5open()
            lib = C.CDLL(LIB)
            sigs = {
                "H5PartOpenFile": (_P, [C.c_char_p, C.c_char]),
                "H5PartCloseFile": (_I64, [_P]),
                "H5PartSetStep": (_I64, [_P, _I64]),
                "H5PartSetNumParticles": (_I64, [_P, _I64]),
                "H5PartGetNumSteps": (_I64, [_P]),
                "H5PartGetNumParticles": (_I64, [_P]),
                "H5PartGetNumDatasets": (_I64, [_P]),
                "H5PartGetDatasetName": (_I64, [_P, _I64, C.c_char_p, _I64]),
                "H5PartWriteStepAttrib": (_I64, [_P, C.c_char_p, _I64, _P, _I64]),
                "H5PartWriteFileAttrib": (_I64, [_P, C.c_char_p, _I64, _P, _I64]),
                "H5PartGetNumStepAttribs": (_I64, [_P]),
                "H5PartGetNumFileAttribs": (_I64, [_P]),
                "H5PartGetStepAttribInfo": (_I64, [_P, _I64, C.c_char_p, _I64, C.POINTER(_I64), C.POINTER(_I64)]),
                "H5PartGetFileAttribInfo": (_I64, [_P, _I64, C.c_char_p, _I64, C.POINTER(_I64), C.POINTER(_I64)]),
                "H5PartReadStepAttrib": (_I64, [_P, C.c_char_p, _P]),
                "H5PartReadFileAttrib": (_I64, [_P, C.c_char_p, _P]),
                "H5PartSetVerbosityLevel": (_I64, [_I64]),
            }
            for t in ("Float64", "Float32", "Int64", "Int32", "Int8"):
                sigs[f"H5PartWriteData{t}"] = (_I64, [_P, C.c_char_p, _P])
                sigs[f"H5PartReadData{t}"] = (_I64, [_P, C.c_char_p, _P])
            for name, (res, args) in sigs.items():
                fn = getattr(lib, name)
                fn.restype, fn.argtypes = res, args
            lib.H5PartSetVerbosityLevel(0)
            # H5PART_FLOAT64 ... H5PART_CHAR are the native HDF5 type ids (H5Part.h:58-63)
            g = lambda s: C.c_int64.in_dll(hdf5, s).value  # noqa: E731
            cls._types = {np.dtype(np.float64): g("H5T_NATIVE_DOUBLE_g"), np.dtype(np.float32): g("H5T_NATIVE_FLOAT_g"),
                          np.dtype(np.int64): g("H5T_NATIVE_INT64_g"), np.dtype(np.int32): g("H5T_NATIVE_INT32_g"),
                          np.dtype(np.int8): g("H5T_NATIVE_SCHAR_g")}
            cls._lib = lib
        return cls._lib

    DATA = {np.dtype(np.float64): "Float64", np.dtype(np.float32): "Float32", np.dtype(np.int64): "Int64",
            np.dtype(np.int32): "Int32", np.dtype(np.int8): "Int8"}

    def __init__(self, path, flags):
        self.L = self.lib()
        self.f = self.L.H5PartOpenFile(str(path).encode(), bytes([flags]))
        assert self.f, f"H5PartOpenFile({path}) failed"

    def close(self):
        self.L.H5PartCloseFile(self.f)

    def num_steps(self):
        return self.L.H5PartGetNumSteps(self.f)

    def set_step(self, step):
        assert self.L.H5PartSetStep(self.f, step) == 0

    def set_num_particles(self, n):
        assert self.L.H5PartSetNumParticles(self.f, n) == 0

    def num_particles(self):
        return self.L.H5PartGetNumParticles(self.f)

    def dataset_names(self):
        out = []
        for i in range(self.L.H5PartGetNumDatasets(self.f)):
            buf = C.create_string_buffer(256)
            self.L.H5PartGetDatasetName(self.f, i, buf, 256)
            out.append(buf.value.decode())
        return out

    def write(self, name, arr):
        arr = np.ascontiguousarray(arr)
        fn = getattr(self.L, f"H5PartWriteData{self.DATA[arr.dtype]}")
        assert fn(self.f, name.encode(), arr.ctypes.data) == 0

    def read(self, name, dtype):
        out = np.empty(self.num_particles(), dtype)
        fn = getattr(self.L, f"H5PartReadData{self.DATA[out.dtype]}")
        assert fn(self.f, name.encode(), out.ctypes.data) == 0
        return out

    def write_attrib(self, name, values, step=True):
        a = np.ascontiguousarray(np.atleast_1d(values))
        fn = self.L.H5PartWriteStepAttrib if step else self.L.H5PartWriteFileAttrib
        assert fn(self.f, name.encode(), self._types[a.dtype], a.ctypes.data, a.size) == 0

    def attribs(self, step=True):
        """{name: (type id, count)} of the step's or the file's attributes"""
        n = (self.L.H5PartGetNumStepAttribs if step else self.L.H5PartGetNumFileAttribs)(self.f)
        info = self.L.H5PartGetStepAttribInfo if step else self.L.H5PartGetFileAttribInfo
        out = {}
        for i in range(n):
            buf, t, k = C.create_string_buffer(256), _I64(), _I64()
            info(self.f, i, buf, 256, C.byref(t), C.byref(k))
            out[buf.value.decode()] = (t.value, k.value)
        return out

    def read_attrib(self, name, dtype, count, step=True):
        out = np.empty(count, dtype)
        fn = self.L.H5PartReadStepAttrib if step else self.L.H5PartReadFileAttrib
        assert fn(self.f, name.encode(), out.ctypes.data) == 0
        return out

    def type_of(self, dtype):
        return self._types[np.dtype(dtype)]
