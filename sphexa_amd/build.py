"""Build libsphexa_amd.so (gfx950) in-tree from sphexa_amd/csrc/*.hip.

    python -m sphexa_amd.build            # compile every translation unit, then link

Each .hip file is compiled to an object in parallel with hipcc, then linked into
sphexa_amd/libsphexa_amd.so. Packed FP32 instructions are disabled for all device code
(-packed-fp32-ops); FP64 neighbour predicates use explicit per-function contraction pragmas.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libsphexa_amd.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SPHX_ARCH", "gfx950")
FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", f"-I{os.path.join(ROOT, 'include')}",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-munsafe-fp-atomics", "-Wno-unused-result"]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def compile_one(src):
    obj = os.path.join(OBJ, os.path.basename(src)[:-4] + ".o")
    cmd = [HIPCC, *FLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


CPP_DRIVER_SRC = os.path.join(ROOT, "tests", "cpp", "dataset_api.cpp")
CPP_DRIVER = os.path.join(ROOT, "tests", "cpp", "dataset_api")


def build_cpp_driver():
    """tests/cpp/dataset_api: the C++ interface (include/sphexa_amd.hpp) driven like the
    reference propagators; run by the GPU tests."""
    cmd = [HIPCC, "-O2", "-std=c++17", f"--offload-arch={ARCH}", f"-I{os.path.join(ROOT, 'include')}",
           CPP_DRIVER_SRC, f"-L{HERE}", "-lsphexa_amd", "-Wl,-rpath,$ORIGIN/../../sphexa_amd", "-o", CPP_DRIVER]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"C++ driver build failed:\n{r.stdout}\n{r.stderr}")
    return CPP_DRIVER


HDF5_ROOT = os.environ.get("SPHX_HDF5_ROOT", "/opt/conda")
IO_SRC = os.path.join(HERE, "io", "h5part.cpp")
IO_LIB = os.path.join(HERE, "libsphexa_amd_io.so")


def build_io():
    """sphexa_amd/libsphexa_amd_io.so: fields and attributes in the H5Part layout of HDF5 files
    (include/sphexa_amd_io.h), linked against the image's HDF5; skipped when HDF5 is absent."""
    if not os.path.exists(os.path.join(HDF5_ROOT, "include", "hdf5.h")):
        print(f"HDF5 not found under {HDF5_ROOT}: libsphexa_amd_io.so not built")
        return None
    lib = os.path.join(HDF5_ROOT, "lib")
    cmd = [HIPCC, "-O2", "-std=c++17", "-fPIC", "-shared", f"-I{os.path.join(ROOT, 'include')}",
           f"-I{os.path.join(HDF5_ROOT, 'include')}", IO_SRC, f"-L{lib}", "-lhdf5", f"-Wl,-rpath,{lib}",
           "-o", IO_LIB]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"I/O library build failed:\n{r.stdout}\n{r.stderr}")
    # tests/cpp/h5_api: the C++ I/O interface (include/sphexa_amd_io.hpp), run by the CPU tests
    drv = os.path.join(ROOT, "tests", "cpp", "h5_api")
    cmd = [HIPCC, "-O2", "-std=c++17", f"-I{os.path.join(ROOT, 'include')}", drv + ".cpp", f"-L{HERE}",
           "-lsphexa_amd_io", "-Wl,-rpath,$ORIGIN/../../sphexa_amd", "-o", drv]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"tests/cpp/h5_api build failed:\n{r.stdout}\n{r.stderr}")
    return IO_LIB


def build(verbose=True):
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    workers = min(len(srcs), max(1, (os.cpu_count() or 4) // 2), 8)
    with cf.ThreadPoolExecutor(workers) as ex:
        objs = list(ex.map(compile_one, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_cpp_driver()
    io = build_io()
    if verbose:
        print(f"built {LIB}, {CPP_DRIVER}" + (f" and {io}" if io else ""))
    return LIB


if __name__ == "__main__":
    try:
        build()
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
