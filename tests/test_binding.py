"""The drop-in binding (sphexa_amd/binding/sph_gpu.cpp) against the reference's own headers -- CPU only.

1. The binding compiles as host C++ against sph/include/sph/sph_gpu.hpp and the reference's
   ParticlesData<GpuTag> (the reference's USE_CUDA build on HIP).
2. It defines every template that sph_gpu.hpp declares (sph_gpu.hpp:11-98).
3. A caller translation unit that includes the reference's public drivers (hydro_std and hydro_ve
   drivers, the EOS selectors, positions.hpp, update_h.hpp, find_neighbors.hpp) and calls each one for
   ParticlesData<GpuTag> leaves only undefined sph:: symbols that the binding object defines: the
   reference linked with the binding in place of its hydro .cu files has no undefined plug-in.

Both translation units are compiled with g++ and the HIP headers (host code only). The reference's
util/tuple.hpp specialises std::tuple_element for thrust::tuple when CUDART_VERSION < 12040 or under
hipcc; the rocThrust shipped here already provides those specialisations, so the compile passes
CUDART_VERSION=12040, the reference's own switch for a thrust that has them.
Skipped where the reference checkout is absent.
"""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
BINDING = os.path.join(ROOT, "sphexa_amd", "binding", "sph_gpu.cpp")

pytestmark = pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "sph", "include")) or not shutil.which("g++"),
                                reason="needs the reference checkout and g++")

FLAGS = ["-std=c++20", "-D__HIP_PLATFORM_AMD__", "-DUSE_CUDA", "-DCUDART_VERSION=12040", "-I/opt/rocm/include",
         f"-I{REF}/domain/include", f"-I{REF}/sph/include", f"-I{ROOT}/include", "-w"]

CALLER = r"""
#include <iostream>
#include "sph/particles_data.hpp"
#include "sph/find_neighbors.hpp"
#include "sph/hydro_std/density.hpp"
#include "sph/hydro_std/eos.hpp"
#include "sph/hydro_std/iad.hpp"
#include "sph/hydro_std/momentum_energy.hpp"
#include "sph/hydro_ve/xmass.hpp"
#include "sph/hydro_ve/ve_def_gradh.hpp"
#include "sph/hydro_ve/eos.hpp"
#include "sph/hydro_ve/iad_divv_curlv.hpp"
#include "sph/hydro_ve/av_switches.hpp"
#include "sph/hydro_ve/momentum_energy.hpp"
#include "sph/positions.hpp"
#include "sph/update_h.hpp"

using Data = sphexa::ParticlesData<cstone::GpuTag>;

// the calls of HydroProp / HydroVeProp::computeForces and integrate
// (main/src/propagator/std_hydro.hpp:139-181, ve_hydro.hpp:147-217)
void drive(Data& d, const cstone::GroupView& g, const cstone::Box<double>& b, float* groupDt)
{
    sph::findNeighborsSfc(0, 1, d, b);
    sph::computeDensity(g, d, b);
    sph::computeEOS_HydroStd(0, 1, d);
    sph::computeIAD(g, d, b);
    sph::computeMomentumEnergySTD(g, d, b);
    sph::computeXMass(g, d, b);
    sph::computeVeDefGradh(g, d, b);
    sph::computeEOS(0, 1, d);
    sph::computeIadDivvCurlv(g, d, b);
    sph::computeAVswitches(g, d, b);
    sph::computeMomentumEnergy<false>(g, groupDt, d, b);
    sph::computeMomentumEnergy<true>(g, groupDt, d, b);
    sph::computePositions(g, d, b, 1e-3f, {1e-3f, 1e-3f, 1e-3f, 1e-3f});
    sph::updateSmoothingLength(g, d);
}
"""


def compile_obj(src, out):
    r = subprocess.run(["g++", *FLAGS, "-c", src, "-o", out], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]


def symbols(obj, kinds):
    r = subprocess.run(["nm", obj], capture_output=True, text=True, check=True)
    out = set()
    for line in r.stdout.splitlines():
        parts = line.split()
        if len(parts) >= 2 and parts[-2] in kinds:
            out.add(parts[-1])
    return out


def demangle(names):
    r = subprocess.run(["c++filt"], input="\n".join(sorted(names)), capture_output=True, text=True, check=True)
    return r.stdout.splitlines()


@pytest.fixture(scope="module")
def objects(tmp_path_factory):
    d = tmp_path_factory.mktemp("binding")
    binding = str(d / "sph_gpu.o")
    compile_obj(BINDING, binding)
    caller_src = d / "caller.cpp"
    caller_src.write_text(CALLER)
    caller = str(d / "caller.o")
    compile_obj(str(caller_src), caller)
    return binding, caller


def test_binding_defines_every_plugin_template(objects):
    binding, _ = objects
    with open(os.path.join(REF, "sph", "include", "sph", "sph_gpu.hpp")) as fh:
        decl = fh.read()
    names = set(re.findall(r"^\s*(?:extern\s+)?(?:void|unsigned)\s+(\w+)\s*\(", decl, re.M))
    names |= set(re.findall(r"^\s*extern\s+void\s+(\w+)\s*\(", decl, re.M))
    assert {"computeXMass", "computeMomentumEnergy", "computeIsothermalEOS", "nsGroupSize"} <= names
    defined = "\n".join(demangle(symbols(binding, {"T", "W"})))
    missing = [n for n in sorted(names) if re.search(rf"\bsph::(cuda::)?{n}\b", defined) is None]
    assert not missing, missing


def test_reference_callers_resolve_against_binding(objects):
    binding, caller = objects
    undefined = symbols(caller, {"U"})
    defined = symbols(binding, {"T", "W"})
    plugin = [s for s in undefined if s.startswith("_ZN3sph") or "3sph" in s[:12]]
    missing = [s for s in demangle(set(plugin) - defined)]
    assert not missing, missing
    assert len(plugin) >= 14, demangle(plugin)
