"""ctypes mirrors of the C ABI structs declared in include/sphexa_amd.h.

Pure data-layout definitions (no library is loaded here), shared by the product
wrapper (sphexa_amd.lib) and the CPU oracle loader so that one particle view can be
handed to both sides.
"""
import ctypes as C

SPHX_OK = 0
SPHX_ERR_HIP = 1
SPHX_ERR_ARG = 2
SPHX_ERR_NOT_CONVERGED = 3
SPHX_ERR_STATE = 4
SPHX_ERR_ALLOC = 5

SPHX_F64 = 0
SPHX_F32 = 1

OPEN = 0
PERIODIC = 1

# HydroType fields of sphexa::ParticlesData (sph/include/sph/particles_data.hpp:220-248) that the
# hot path reads or writes, in the order of sphx_fields_t.
REAL_FIELDS = ("x", "y", "z", "u", "temp", "du")
HYDRO_FIELDS = ("h", "m", "vx", "vy", "vz", "rho", "p", "c", "prho", "tdpdTrho",
                "c11", "c12", "c13", "c22", "c23", "c33", "xm", "kx", "gradh", "divv", "curlv",
                "alpha", "ax", "ay", "az", "dV11", "dV12", "dV13", "dV22", "dV23", "dV33")
INT_FIELDS = ("keys", "nc")
ALL_FIELDS = REAL_FIELDS + HYDRO_FIELDS + INT_FIELDS


class Box(C.Structure):
    """cstone::Box<double> (domain/include/cstone/sfc/box.hpp:95-174)."""
    _fields_ = [("lo", C.c_double * 3), ("hi", C.c_double * 3), ("boundary", C.c_int32 * 3)]

    @classmethod
    def make(cls, lo, hi, boundary=OPEN):
        lo = [lo] * 3 if isinstance(lo, (int, float)) else list(lo)
        hi = [hi] * 3 if isinstance(hi, (int, float)) else list(hi)
        bc = [boundary] * 3 if isinstance(boundary, int) else list(boundary)
        b = cls()
        for d in range(3):
            b.lo[d], b.hi[d], b.boundary[d] = float(lo[d]), float(hi[d]), int(bc[d])
        return b

    @property
    def lengths(self):
        return [self.hi[d] - self.lo[d] for d in range(3)]

    @property
    def periodic(self):
        return [self.boundary[d] == PERIODIC for d in range(3)]


class Fields(C.Structure):
    _fields_ = ([(n, C.c_void_p) for n in REAL_FIELDS] + [(n, C.c_void_p) for n in HYDRO_FIELDS]
                + [("keys", C.c_void_p), ("nc", C.c_void_p)])


class Params(C.Structure):
    """Dataset scalars (sph/include/sph/particles_data.hpp:91-146)."""
    _fields_ = [("K", C.c_double), ("Kcour", C.c_double), ("gamma", C.c_double), ("muiConst", C.c_double),
                ("minDt", C.c_double), ("alphamin", C.c_double), ("alphamax", C.c_double),
                ("decay_constant", C.c_double), ("Atmin", C.c_double), ("Atmax", C.c_double),
                ("ramp", C.c_double), ("ng0", C.c_uint32), ("ngmax", C.c_uint32),
                ("hydro_type", C.c_int32), ("flags", C.c_int32), ("eos", C.c_int32), ("reserved", C.c_int32),
                ("soundSpeedConst", C.c_double), ("polytropic_const", C.c_double), ("polytropic_index", C.c_double)]

    @classmethod
    def defaults(cls, K, hydro_type=SPHX_F64, **kw):
        """Reference defaults: ng0 100, ngmax 150, Kcour 0.2, gamma 5/3, muiConst 10,
        alphamin 0.05, alphamax 1, decay 0.2, Atmin 0.1, Atmax 0.2, ramp 1/(Atmax-Atmin); ideal-gas EOS,
        soundSpeedConst 1, polytropic_const 1, polytropic_index 5/3 (particles_data.hpp:91-146)."""
        p = cls(K=K, Kcour=0.2, gamma=5.0 / 3.0, muiConst=10.0, minDt=1e-12, alphamin=0.05, alphamax=1.0,
                decay_constant=0.2, Atmin=0.1, Atmax=0.2, ramp=1.0 / (0.2 - 0.1), ng0=100, ngmax=150,
                hydro_type=hydro_type, flags=0, eos=SPHX_EOS_IDEAL_GAS, soundSpeedConst=1.0, polytropic_const=1.0,
                polytropic_index=5.0 / 3.0)
        for k, v in kw.items():
            setattr(p, k, v)
        return p


SPHX_EOS_IDEAL_GAS, SPHX_EOS_ISOTHERMAL, SPHX_EOS_POLYTROPIC = 0, 1, 2
SPHX_FLAG_STD_ALPHA = 1
SPHX_FLAG_CHAIN = 2
SPHX_FLAG_TREE_UPDATE = 4


def chained(params):
    """A copy of params with SPHX_FLAG_CHAIN: XMass, the IAD passes and AV switches write the next
    pass's neighbour records in their epilogues (bitwise the same results, fewer pack kernels)."""
    p = Params.from_buffer_copy(params)
    p.flags |= SPHX_FLAG_CHAIN
    return p

FIXED = 2  # cstone::BoundaryType::fixed (box.hpp:80-85)

HISTORY_FIELDS = ("x_m1", "y_m1", "z_m1", "du_m1", "mui")


class History(C.Structure):
    """sphx_history_t: previous-step fields of the integrator (XM1Type = HydroType; mui optional)."""
    _fields_ = [(n, C.c_void_p) for n in HISTORY_FIELDS]


class Timestep(C.Structure):
    """sphx_timestep_t: the global time-step scalars (particles_data.hpp:96-141)."""
    _fields_ = [(n, C.c_double) for n in ("minDt", "minDt_m1", "ttot", "minDtCourant", "minDtRho", "maxDtIncrease",
                                          "g", "etaAcc", "eps", "Krho")]

    @classmethod
    def defaults(cls, **kw):
        """Reference defaults: minDt = minDt_m1 = 1e-12, ttot 0, minDtCourant = minDtRho = inf,
        maxDtIncrease 1.1, g 0, etaAcc 0.2, eps 0.005, Krho 0.06."""
        t = cls(minDt=1e-12, minDt_m1=1e-12, ttot=0.0, minDtCourant=float("inf"), minDtRho=float("inf"),
                maxDtIncrease=1.1, g=0.0, etaAcc=0.2, eps=0.005, Krho=0.06)
        for k, v in kw.items():
            setattr(t, k, v)
        return t

ADJOINT_FIELDS = ("x", "y", "z", "h", "rho", "p", "alpha", "c11", "c12", "c13", "c22", "c23", "c33", "ax", "ay", "az",
                  "du")


class Adjoint(C.Structure):
    """sphx_adjoint_t: cotangent / gradient arrays of the adjoint entry points."""
    _fields_ = [(n, C.c_void_p) for n in ADJOINT_FIELDS]


class Groups(C.Structure):
    """sphx_groups_t: cstone::GroupView (device arrays) plus the optional per-group time step."""
    _fields_ = [("firstBody", C.c_uint32), ("lastBody", C.c_uint32), ("numGroups", C.c_uint32),
                ("reserved", C.c_uint32), ("groupStart", C.c_void_p), ("groupEnd", C.c_void_p),
                ("groupDt", C.c_void_p)]

    @classmethod
    def of(cls, starts, first, last, group_dt=None):
        """a view of a device tensor `starts` holding the group starts followed by `last`"""
        g = cls(firstBody=first, lastBody=last, numGroups=starts.numel() - 1)
        g.groupStart = starts.data_ptr()
        g.groupEnd = starts.data_ptr() + starts.element_size()
        g.groupDt = group_dt.data_ptr() if group_dt is not None else None
        return g


class TreeView(C.Structure):
    _fields_ = [("numLeafNodes", C.c_int32), ("numInternalNodes", C.c_int32), ("numNodes", C.c_int32),
                ("reserved", C.c_int32), ("prefixes", C.c_void_p), ("childOffsets", C.c_void_p),
                ("parents", C.c_void_p), ("internalToLeaf", C.c_void_p), ("levelRange", C.c_void_p),
                ("leaves", C.c_void_p), ("counts", C.c_void_p), ("layout", C.c_void_p),
                ("centers", C.c_void_p), ("sizes", C.c_void_p)]


def fields_from(arrays):
    """Build a Fields struct from a mapping name -> array exposing an address
    (numpy: .ctypes.data, torch: .data_ptr()); missing names stay NULL."""
    f = Fields()
    for name in ALL_FIELDS:
        a = arrays.get(name)
        if a is None:
            continue
        addr = a.data_ptr() if hasattr(a, "data_ptr") else a.ctypes.data
        setattr(f, name, addr)
    return f


class Conserved(C.Structure):
    """sphx_conserved_t: totals of sphx_conserved_quantities."""
    _fields_ = [("eKin", C.c_double), ("eInt", C.c_double), ("linmom", C.c_double * 3), ("angmom", C.c_double * 3),
                ("ncsum", C.c_double)]

    def as_dict(self):
        return {"eKin": self.eKin, "eInt": self.eInt, "linmom": list(self.linmom), "angmom": list(self.angmom),
                "ncsum": self.ncsum}
