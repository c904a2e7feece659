"""ctypes binding of libsphexa_amd.so (the C ABI in include/sphexa_amd.h).

torch provides device memory and the stream; every computation runs in the HIP library.
Loading fails loudly if the library has not been built: there is no fallback path.
"""
import ctypes as C
import os

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("SPHEXA_AMD_LIB") or os.path.join(HERE, "libsphexa_amd.so")

_lib = None

_P = C.c_void_p
_BOXP = C.POINTER(abi.Box)
_FP = C.POINTER(abi.Fields)
_PP = C.POINTER(abi.Params)
_AP = C.POINTER(abi.Adjoint)
_U32 = C.c_uint32
_SZ = C.c_size_t

_TSP = C.POINTER(abi.Timestep)
_HP = C.POINTER(abi.History)
_SIGS = {
    "sphx_workspace_create": [C.POINTER(_P), C.c_int, _P],
    "sphx_workspace_destroy": [_P],
    "sphx_workspace_set_kernel": [_P, C.c_double, C.POINTER(C.c_double)],
    "sphx_compute_sfc_keys": [_P, _BOXP, _P, _P, _P, _P, _SZ],
    "sphx_sort_keys": [_P, _P, _P, _SZ],
    "sphx_sort_keys_noalloc": [C.c_int, _P, _P, _P, _SZ],
    "sphx_gather": [_P, _P, _P, _P, _SZ, C.c_int],
    "sphx_build_tree": [_P, _P, _SZ, _U32, _BOXP],
    "sphx_tree_view": [_P, C.POINTER(abi.TreeView)],
    "sphx_domain_sync": [_P, _BOXP, _FP, _PP, _SZ, _U32],
    "sphx_global_box": [_P, _BOXP, _P, _P, _P, _SZ, C.c_int],
    "sphx_domain_reorder": [_P, C.POINTER(C.c_void_p), C.POINTER(C.c_int), _SZ, _SZ],
    "sphx_find_neighbors": [_P, _BOXP, _FP, _PP, _U32, _U32, C.c_int],
    "sphx_find_neighbors_xmass": [_P, _BOXP, _FP, _PP, _U32, _U32, C.c_int, C.c_int],
    "sphx_copy_neighbors": [_P, _P, _U32, _U32],
    "sphx_neighbor_stats": [_P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint32)],
    "sphx_search_tests": [_P, C.POINTER(C.c_uint64)],
    "sphx_compute_groups": [_P, _BOXP, _FP, _U32, _U32, C.c_float, C.POINTER(abi.Groups)],
    "sphx_set_groups": [_P, C.POINTER(abi.Groups)],
    "sphx_density_std": [_P, _BOXP, _FP, _PP, _U32, _U32],
    "sphx_eos_std": [_P, _FP, _PP, _U32, _U32],
    "sphx_iad_std": [_P, _BOXP, _FP, _PP, _U32, _U32],
    "sphx_momentum_energy_std": [_P, _BOXP, _FP, _PP, _U32, _U32, C.POINTER(C.c_double)],
    "sphx_xmass": [_P, _BOXP, _FP, _PP, _U32, _U32],
    "sphx_ve_def_gradh": [_P, _BOXP, _FP, _PP, _U32, _U32],
    "sphx_eos_ve": [_P, _FP, _PP, _U32, _U32],
    "sphx_iad_divv_curlv": [_P, _BOXP, _FP, _PP, _U32, _U32],
    "sphx_av_switches": [_P, _BOXP, _FP, _PP, _U32, _U32],
    "sphx_momentum_energy_ve": [_P, _BOXP, _FP, _PP, _U32, _U32, C.c_int, C.POINTER(C.c_double)],
    "sphx_workspace_set_profiling": [_P, C.c_int],
    "sphx_kernel_ms": [_P, C.c_char_p, C.POINTER(C.c_double)],
    "sphx_rho_timestep": [_P, _FP, _PP, _U32, _U32, _TSP],
    "sphx_compute_timestep": [_P, _FP, _PP, _U32, _U32, _TSP],
    "sphx_compute_positions": [_P, _BOXP, _FP, _HP, _PP, _U32, _U32, C.c_double, C.c_double, C.c_double],
    "sphx_compute_positions_rungs": [_P, _BOXP, _FP, _HP, _PP, _U32, _U32, C.c_double, _P, C.c_int, _P,
                                     C.c_double],
    "sphx_drift_positions": [_P, _FP, _HP, _PP, _U32, _U32, C.c_double, C.c_double, _P, C.c_int, _P, C.c_double],
    "sphx_group_divv_timestep": [_P, _FP, _PP, _U32, _U32, C.c_float, _P],
    "sphx_group_acc_timestep": [_P, _FP, _PP, _U32, _U32, C.c_float, _P],
    "sphx_store_rung": [_P, _U32, _U32, C.c_uint8, _P],
    "sphx_update_smoothing_length": [_P, _FP, _PP, _U32, _U32],
    "sphx_conserved_quantities": [_P, _FP, _PP, _U32, _U32, C.c_double, C.POINTER(abi.Conserved)],
    "sphx_density_std_vjp": [_P, _BOXP, _FP, _PP, _U32, _U32, _AP, _AP],
    "sphx_iad_std_vjp": [_P, _BOXP, _FP, _PP, _U32, _U32, _AP, _AP],
    "sphx_momentum_energy_std_vjp": [_P, _BOXP, _FP, _PP, _U32, _U32, _AP, _AP],
}

EXPORTED = tuple(_SIGS) + ("sphx_last_error",)


class SphxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"sphexa_amd error {code}: {msg}")
        self.code = code


def load():
    """Load the HIP library (raises if it is missing: no CPU fallback exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(f"{LIB_PATH} is not built; run `python -m sphexa_amd.build`")
    lib = C.CDLL(LIB_PATH)
    for name, args in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_int
    lib.sphx_last_error.argtypes = [_P]
    lib.sphx_last_error.restype = C.c_char_p
    _lib = lib
    return lib


class Workspace:
    """Owns a sphx_workspace on one device, bound to a HIP stream (default: torch's current)."""

    def __init__(self, device=0, stream=None, sinc_index=6.0):
        import torch

        self.lib = load()
        self.device = device
        if stream is None:
            stream = torch.cuda.current_stream(device).cuda_stream
        self.stream = stream
        h = _P()
        self._check(self.lib.sphx_workspace_create(C.byref(h), device, _P(stream)), None)
        self.handle = h
        K = C.c_double()
        self.call("sphx_workspace_set_kernel", sinc_index, C.byref(K))
        self.K = K.value

    def _check(self, code, handle):
        if code != abi.SPHX_OK:
            msg = self.lib.sphx_last_error(handle).decode() if handle else "workspace creation failed"
            raise SphxError(code, msg)

    def call(self, name, *args, allow=()):
        code = getattr(self.lib, name)(self.handle, *args)
        if code != abi.SPHX_OK and code not in allow:
            self._check(code, self.handle)
        return code

    def close(self):
        if getattr(self, "handle", None):
            self.lib.sphx_workspace_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- thin wrappers -----------------------------------------------------------------------
    def sfc_keys(self, box, x, y, z, keys):
        self.call("sphx_compute_sfc_keys", C.byref(box), x.data_ptr(), y.data_ptr(), z.data_ptr(), keys.data_ptr(),
                  x.numel())

    def sort_keys(self, keys, order):
        self.call("sphx_sort_keys", keys.data_ptr(), order.data_ptr(), keys.numel())

    def build_tree(self, sorted_keys, box, bucket=64):
        self.call("sphx_build_tree", sorted_keys.data_ptr(), sorted_keys.numel(), bucket, C.byref(box))

    def tree_view(self):
        tv = abi.TreeView()
        self.call("sphx_tree_view", C.byref(tv))
        return tv

    def domain_sync(self, box, fields, params, n, bucket=64):
        self.call("sphx_domain_sync", C.byref(box), C.byref(fields), C.byref(params), n, bucket)

    def find_neighbors(self, box, fields, params, first, last, iterate_h=True):
        """Returns True when every particle's h/nc iteration converged."""
        code = self.call("sphx_find_neighbors", C.byref(box), C.byref(fields), C.byref(params), first, last,
                         int(iterate_h), allow=(abi.SPHX_ERR_NOT_CONVERGED,))
        return code == abi.SPHX_OK

    def find_neighbors_xmass(self, box, fields, params, first, last, iterate_h=True, rho=False):
        """Search fused with XMass (xm) or, with rho=True, the std density; True when converged."""
        code = self.call("sphx_find_neighbors_xmass", C.byref(box), C.byref(fields), C.byref(params), first, last,
                         int(iterate_h), 2 if rho else 1, allow=(abi.SPHX_ERR_NOT_CONVERGED,))
        return code == abi.SPHX_OK

    def copy_neighbors(self, out, first, last):
        self.call("sphx_copy_neighbors", out.data_ptr(), first, last)

    def neighbor_stats(self):
        s, m = C.c_uint64(), C.c_uint32()
        self.call("sphx_neighbor_stats", C.byref(s), C.byref(m))
        return s.value, m.value

    def compute_groups(self, box, fields, first, last, tol_factor=2.0):
        """Spatial target groups of [first, last) (sphx_compute_groups); an abi.Groups view into the
        workspace."""
        g = abi.Groups()
        self.call("sphx_compute_groups", C.byref(box), C.byref(fields), first, last, tol_factor, C.byref(g))
        return g

    def set_groups(self, groups):
        """Target groups of the calls that follow (abi.Groups or None for consecutive groups of 64)."""
        self.call("sphx_set_groups", C.byref(groups) if groups is not None else None)

    def search_tests(self):
        """candidate distance tests of the last search (sphx_search_tests)"""
        t = C.c_uint64()
        self.call("sphx_search_tests", C.byref(t))
        return t.value

    def pair(self, name, box, fields, params, first, last):
        self.call(name, C.byref(box), C.byref(fields), C.byref(params), first, last)

    def eos(self, name, fields, params, first, last):
        self.call(name, C.byref(fields), C.byref(params), first, last)

    def momentum_energy_std(self, box, fields, params, first, last):
        dt = C.c_double()
        self.call("sphx_momentum_energy_std", C.byref(box), C.byref(fields), C.byref(params), first, last,
                  C.byref(dt))
        return dt.value

    def set_profiling(self, on=True):
        self.call("sphx_workspace_set_profiling", int(on))

    def kernel_ms(self, kernel):
        """Duration (ms) of the last timed launch of `kernel` (see sphx_kernel_ms)."""
        ms = C.c_double()
        self.call("sphx_kernel_ms", kernel.encode(), C.byref(ms))
        return ms.value

    def vjp(self, name, box, fields, params, first, last, cot, grad):
        """Adjoint entry point `name`: cotangents `cot`, gradients added into `grad` (abi.Adjoint)."""
        self.call(name, C.byref(box), C.byref(fields), C.byref(params), first, last, C.byref(cot), C.byref(grad))

    def global_box(self, box, x, y, z, n, first_call):
        """Update box (abi.Box, in place) as Domain::sync does: refit the non-periodic dimensions to
        the particles, with limited shrinking unless first_call."""
        self.call("sphx_global_box", C.byref(box), C.c_void_p(x.data_ptr()), C.c_void_p(y.data_ptr()),
                  C.c_void_p(z.data_ptr()), n, int(first_call))
        return box

    def domain_reorder(self, tensors, n):
        """Apply the last domain sync's permutation to more per-particle tensors (one or a list)."""
        if not isinstance(tensors, (list, tuple)):
            tensors = [tensors]
        k = len(tensors)
        ptrs = (C.c_void_p * k)(*[t.data_ptr() for t in tensors])
        sizes = (C.c_int * k)(*[t.element_size() for t in tensors])
        self.call("sphx_domain_reorder", ptrs, sizes, k, n)

    def rho_timestep(self, fields, params, first, last, ts):
        """ts.minDtRho = Krho / |max divv| (sphx_rho_timestep)."""
        self.call("sphx_rho_timestep", C.byref(fields), C.byref(params), first, last, C.byref(ts))

    def compute_timestep(self, fields, params, first, last, ts):
        """Advance the time-step state ts (abi.Timestep) as computeTimestep does."""
        self.call("sphx_compute_timestep", C.byref(fields), C.byref(params), first, last, C.byref(ts))

    def compute_positions(self, box, fields, history, params, first, last, dt, dt_m1, const_cv=-1.0):
        """Press positions and Adams-Bashforth energy; const_cv < 0: cv from params.muiConst."""
        self.call("sphx_compute_positions", C.byref(box), C.byref(fields), C.byref(history), C.byref(params), first,
                  last, dt, dt_m1, const_cv)

    def compute_positions_rungs(self, box, fields, history, params, first, last, dt, dt_m1_rungs, rung_ptr,
                                const_cv=-1.0):
        """compute_positions with the previous step per rung (a sequence of floats) and a device
        uint8 rung per particle (pointer, or None)."""
        arr = (C.c_float * len(dt_m1_rungs))(*dt_m1_rungs)
        self.call("sphx_compute_positions_rungs", C.byref(box), C.byref(fields), C.byref(history), C.byref(params),
                  first, last, dt, arr, len(dt_m1_rungs), rung_ptr, const_cv)

    def drift_positions(self, fields, history, params, first, last, dt, dt_back, dt_m1_rungs, rung_ptr,
                        const_cv=-1.0):
        """drift the particles of a block time-step hierarchy back by dt_back and forward by dt"""
        arr = (C.c_float * len(dt_m1_rungs))(*dt_m1_rungs)
        self.call("sphx_drift_positions", C.byref(fields), C.byref(history), C.byref(params), first, last, dt, dt_back,
                  arr, len(dt_m1_rungs), rung_ptr, const_cv)

    def group_divv_timestep(self, fields, params, first, last, krho, group_dt_ptr):
        self.call("sphx_group_divv_timestep", C.byref(fields), C.byref(params), first, last, krho, group_dt_ptr)

    def group_acc_timestep(self, fields, params, first, last, eta_acc, group_dt_ptr):
        self.call("sphx_group_acc_timestep", C.byref(fields), C.byref(params), first, last, eta_acc, group_dt_ptr)

    def store_rung(self, first, last, rung, rungs_ptr):
        self.call("sphx_store_rung", first, last, rung, rungs_ptr)

    def conserved_quantities(self, fields, params, first, last, cv=-1.0):
        """Energies and momenta over [first, last) (sphx_conserved_quantities) as a dict; cv < 0: from
        params.muiConst."""
        out = abi.Conserved()
        self.call("sphx_conserved_quantities", C.byref(fields), C.byref(params), first, last, cv, C.byref(out))
        return out.as_dict()

    def update_smoothing_length(self, fields, params, first, last):
        self.call("sphx_update_smoothing_length", C.byref(fields), C.byref(params), first, last)

    def momentum_energy_ve(self, box, fields, params, first, last, av_clean=False):
        dt = C.c_double()
        self.call("sphx_momentum_energy_ve", C.byref(box), C.byref(fields), C.byref(params), first, last,
                  int(av_clean), C.byref(dt))
        return dt.value
