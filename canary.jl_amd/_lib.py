"""ctypes binding of libcanary_dg.so (declarations mirror include/canary_dg.h).

The library is loaded from this package directory; a missing library is an
error, never a silent fallback.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# CANARY_DG_LIB points at another build of the same library (kernel tuning runs)
LIBPATH = os.environ.get("CANARY_DG_LIB") or os.path.join(HERE, "libcanary_dg.so")

# error codes (canary_dg.h)
OK, ERR_ARG, ERR_UNSUPPORTED, ERR_LAUNCH, ERR_BC, ERR_VMAP = 0, 1, 2, 3, 4, 5
ERR_NOCONVERGENCE, ERR_LOBATTO_N1, ERR_ORDERING, ERR_ORIENTATION = 6, 7, 8, 9
ERR_NOMEM, ERR_CAPACITY = 10, 11
EULER, SWE, EULER_TE = 0, 1, 2  # EULER_TE: 2D total-energy equations (set3c)
FP64, FP32, FP64_FMA, FP64_MFMA = 0, 1, 2, 3  # FP64_MFMA: contractions on the matrix cores
PATH_KERNELS, PATH_FUSED, PATH_FACE_UPDATE = 0, 1, 2
REDUCE_WORK = 1024  # CANARY_REDUCE_WORK
ENDPT_NEITHER, ENDPT_LEFT, ENDPT_RIGHT, ENDPT_BOTH = 0, 1, 2, 3

ERRNAMES = {
    ERR_ARG: "bad argument", ERR_UNSUPPORTED: "unsupported (eqn, dim, N, prec)",
    ERR_LAUNCH: "HIP launch/runtime error", ERR_BC: "invalid boundary condition",
    ERR_VMAP: "vmap index out of range", ERR_NOCONVERGENCE: "QL iteration did not converge",
    ERR_LOBATTO_N1: "Must have at least two points for both ends.",
    ERR_ORDERING: "Problem finding vertex ordering", ERR_ORIENTATION: "unsupported face orientation",
    ERR_NOMEM: "out of memory", ERR_CAPACITY: "request exceeds a fixed capacity",
}


class CanaryError(RuntimeError):
    def __init__(self, code, what=""):
        self.code = code
        super().__init__("%s failed: %s (code %d)" % (what, ERRNAMES.get(code, "error"), code))


class Phys(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("gamma", "p0", "R_gas", "c_p", "c_v", "gravity", "swe_gravity", "delta_nl")]


class Bubble(C.Structure):
    _fields_ = [("theta0", C.c_double), ("dtheta", C.c_double), ("rc", C.c_double),
                ("xc", C.c_double), ("yc", C.c_double), ("zc", C.c_double),
                ("spherical", C.c_int), ("jitter", C.c_double), ("seed", C.c_uint64)]


class Vortex(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("beta", "x0", "y0", "u0", "v0", "rho_inf", "p_inf", "gamma", "Lx", "Ly")]


class BatchSolve(C.Structure):
    _fields_ = [("Q", C.c_void_p * 2), ("rhs", C.c_void_p), ("vgeo", C.c_void_p),
                ("sgeo", C.c_void_p), ("D", C.c_void_p), ("vmapM", C.c_void_p),
                ("vmapP", C.c_void_p), ("elemtobndy", C.c_void_p), ("nreal", C.c_int64),
                ("nelem_total", C.c_int64), ("dt", C.c_double), ("phys", Phys),
                ("N", C.c_int32), ("advection", C.c_int32), ("pface", C.c_void_p)]


class P32Solve(C.Structure):
    _fields_ = [("q", C.c_void_p * 2), ("rhs", C.c_void_p), ("bg32", C.c_void_p),
                ("vg", C.c_void_p), ("sg", C.c_void_p), ("nbr", C.c_void_p),
                ("D", C.c_void_p), ("nreal", C.c_int64), ("dt", C.c_double), ("phys", Phys),
                ("N", C.c_int32), ("dim", C.c_int32)]


class BatchItem(C.Structure):
    _fields_ = [("solve", C.c_int32), ("N", C.c_int16), ("epw", C.c_int16), ("e0", C.c_int64)]


class BatchClass(C.Structure):
    _fields_ = [("items", C.c_void_p), ("nitems", C.c_int64), ("N", C.c_int32),
                ("epw", C.c_int32)]


BATCH_DROPIN, BATCH_P32, BATCH_DROPIN_ALL, BATCH_PREPARED, BATCH_PREPARED_ALL = 0, 1, 2, 3, 4


_lib = None

P = C.c_void_p
I64 = C.c_int64
D = C.c_double
INT = C.c_int

_SIGS = {
    "canary_phys_default": (None, [C.POINTER(Phys)]),
    "canary_volumerhs": (INT, [INT, INT, INT, INT, P, P, P, P, I64, C.POINTER(Phys), P]),
    "canary_facerhs": (INT, [INT, INT, INT, INT, P, P, P, I64, I64, P, P, P, C.POINTER(Phys), P]),
    "canary_facerhs_geo": (INT, [INT, INT, INT, INT, P, P, P, P, I64, I64, P, P, P,
                                 C.POINTER(Phys), P]),
    "canary_updatesolution": (INT, [INT, INT, INT, INT, P, P, P, I64, D, D, D, INT, P]),
    "canary_stage_fused": (INT, [INT, INT, INT, INT, P, P, P, P, P, P, I64, I64, P, P, P,
                                 D, D, D, INT, C.POINTER(Phys), P]),
    "canary_facerhs_update": (INT, [INT, INT, INT, INT, P, P, P, P, P, I64, I64, P, P, P,
                                    D, D, D, INT, C.POINTER(Phys), P]),
    "canary_lsrk54": (INT, [INT, INT, INT, INT, INT, I64, D, P, P, P, P, P, P, I64, I64, P, P,
                            P, C.POINTER(Phys), INT, INT, C.POINTER(INT), P]),
    "canary_lsrk54_coefs": (None, [P, P, P]),
    "canary_pow_register": (INT, [D, P]),
    "canary_prepare_geometry": (INT, [INT, INT, I64, I64, P, P, P, P, P, P, P, P, P]),
    "canary_stage_prepared": (INT, [INT, INT, INT, INT, P, P, P, P, P, P, P, I64, I64, D, D, D,
                                    INT, C.POINTER(Phys), P]),
    "canary_lsrk54_prepared": (INT, [INT, INT, INT, INT, I64, D, P, P, P, P, P, P, P, I64,
                                     C.POINTER(Phys), INT, INT, C.POINTER(INT), P]),
    "canary_pack_nodes": (INT, [INT, INT, INT, P, P, P, I64, I64, INT, P]),
    "canary_unpack_nodes": (INT, [INT, INT, INT, P, P, P, I64, I64, INT, P]),
    "canary_batch_run": (INT, [INT, INT, INT, INT, P, P, INT, I64, INT, P]),
    "canary_check_errors": (INT, [P, P]),
    "canary_perturb32_prepare": (INT, [INT, INT, I64, I64, P, P, P, P, P, P, P, C.POINTER(Phys),
                                       P, P, P, P, P, P, P]),
    "canary_stage_perturb32": (INT, [INT, INT, P, P, P, P, P, P, P, P, I64, D, D, D,
                                     C.POINTER(Phys), P]),
    "canary_perturb32_reconstruct": (INT, [INT, INT, I64, P, P, P, P]),
    "canary_perturb_mixed_prepare": (INT, [INT, INT, I64, I64, P, P, P, P, P, P, P,
                                           C.POINTER(Phys), P, P, P, P, P, P, P]),
    "canary_stage_perturb_mixed": (INT, [INT, INT, P, P, P, P, P, P, P, P, I64, D, D, D,
                                         C.POINTER(Phys), P]),
    "canary_p32_batch_variant": (INT, [INT, INT, INT, C.POINTER(INT), C.POINTER(INT)]),
    "canary_p32_batch_stage": (INT, [INT, INT, INT, P, P, I64, INT, D, D, P]),
    "canary_batch_variant": (INT, [INT, INT, INT, INT, INT, C.POINTER(INT), C.POINTER(INT)]),
    "canary_batch_stage": (INT, [INT, INT, INT, INT, INT, P, P, I64, INT, D, D, P]),
    "canary_batch_all_variant": (INT, [INT, INT, INT, INT, C.POINTER(INT), C.POINTER(INT)]),
    "canary_batch_stage_all": (INT, [INT, INT, INT, P, P, I64, INT, D, D, P]),
    "canary_batch_stage_prepared": (INT, [INT, INT, INT, INT, INT, P, P, I64, INT, D, D, P]),
    "canary_batch_stage_all_prepared": (INT, [INT, INT, INT, P, P, I64, INT, D, D, P]),
    "canary_device_count": (INT, []),
    "canary_stream_create": (INT, [C.POINTER(P)]),
    "canary_stream_destroy": (INT, [P]),
    "canary_computegeometry_device": (INT, [INT, INT, I64, P, P, P, P, P]),
    "canary_fillsendQ": (INT, [INT, INT, INT, P, P, P, I64, I64, INT, P]),
    "canary_transferrecvQ": (INT, [INT, INT, INT, P, P, I64, I64, INT, P]),
    "canary_l2energysquared_device": (INT, [INT, INT, INT, INT, P, P, I64, P, P, P]),
    "canary_cfl_device": (INT, [INT, INT, INT, P, P, I64, C.POINTER(Phys), P, P, P]),
    "canary_cfl_device_noalloc": (INT, [INT, INT, INT, P, P, I64, C.POINTER(Phys), P, P]),
    "canary_courant_device": (INT, [INT, INT, P, P, I64, C.POINTER(Phys), D, P, P, P]),
    "canary_ic_bubble_device": (INT, [INT, INT, I64, P, C.POINTER(Bubble), C.POINTER(Phys), P, P]),
    "canary_ic_vortex2d_device": (INT, [INT, I64, P, C.POINTER(Vortex), D, P, P]),
    "canary_gauss_rule": (INT, [INT, INT, INT, P, P]),
    "canary_lglpoints": (INT, [INT, P, P]),
    "canary_lgpoints": (INT, [INT, P, P]),
    "canary_baryweights": (INT, [INT, P, P]),
    "canary_spectralderivative": (INT, [INT, P, P]),
    "canary_interpolationmatrix": (INT, [INT, P, INT, P, P]),
    "canary_hilbertcode": (INT, [INT, P, INT, P]),
    "canary_centroidtocode": (INT, [INT, INT, I64, P, P]),
    "canary_vertsortandorder": (INT, [INT, P, P, P]),
    "canary_brickmesh": (P, [INT, P, P, P, I64, I64, P, P]),
    "canary_mesh_from_arrays": (P, [INT, I64, P, P, P, I64, P, P]),
    "canary_mesh_partition": (INT, [P]),
    "canary_mesh_connect": (INT, [P]),
    "canary_mesh_size": (I64, [P, INT]),
    "canary_mesh_get": (INT, [P, C.c_char_p, P]),
    "canary_mesh_free": (None, [P]),
    "canary_mappings": (INT, [INT, INT, I64, P, P, P, P, P]),
    "canary_creategrid": (INT, [INT, INT, I64, P, P, P]),
    "canary_computemetric": (INT, [INT, INT, I64, P, P, P, P]),
    "canary_computegeometry": (INT, [INT, INT, I64, P, P, P, P]),
    "canary_hash_jitter": (D, [C.c_uint64, D, D, D]),
    "canary_pow_cr": (D, [D, D]),
    "canary_pow_cr_table": (D, [D, D]),
    "canary_ic_bubble": (INT, [INT, INT, I64, P, C.POINTER(Bubble), C.POINTER(Phys), P]),
    "canary_ic_vortex2d": (INT, [INT, I64, P, C.POINTER(Vortex), D, P]),
    "canary_ic_swe_bumps": (INT, [INT, I64, P, D, INT, P, D, D, P]),
    "canary_cfl": (D, [INT, INT, INT, P, P, I64, C.POINTER(Phys)]),
    "canary_l2energysquared": (D, [INT, INT, INT, INT, P, P, I64]),
}


def lib():
    """Load libcanary_dg.so (built by build.py); raise if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError("libcanary_dg.so is not built: run `python canary.jl_amd/build.py`")
        L = C.CDLL(LIBPATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name, None)
            if f is None:
                continue
            f.restype = res
            f.argtypes = args
        _lib = L
        # Each unit's code object is loaded on the GPU at the unit's first use;
        # a load that lands inside timed or concurrent work slows that work
        # (measured: the next 100 fused stages 3-5% slower).  Reading every
        # unit's status word once uses them all now.
        if L.canary_device_count() > 0:
            L.canary_check_errors(None, (C.c_int64 * 4)())
    return _lib


def exported_symbols():
    return list(_SIGS)


def check(rc, what):
    if rc != 0:
        raise CanaryError(rc, what)


def default_phys(**kw):
    p = Phys()
    lib().canary_phys_default(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, float(v))
    return p
