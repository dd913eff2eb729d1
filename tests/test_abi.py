"""The C-ABI library loads and exports every function include/*.h declares
(no compute calls: these run without a GPU)."""
import ctypes
import glob
import os
import re

import canary_amd._lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(canary_\w+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = L.lib()
    names = declared_functions()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_bindings_cover_header():
    assert set(declared_functions()) <= set(L.exported_symbols())


def test_device_count_without_gpu_is_safe():
    n = L.lib().canary_device_count()
    assert n >= 0


def test_device_entry_points_reject_bad_arguments_before_launch():
    lib = L.lib()
    p = L.default_phys()
    # unsupported (eqn, dim), N and precision are rejected on the host
    assert lib.canary_volumerhs(1, 3, 4, 0, None, None, None, None, 1, ctypes.byref(p), None) == L.ERR_UNSUPPORTED
    assert lib.canary_volumerhs(0, 3, 9, 0, None, None, None, None, 1, ctypes.byref(p), None) == L.ERR_UNSUPPORTED
    assert lib.canary_volumerhs(0, 3, 4, 7, None, None, None, None, 1, ctypes.byref(p), None) == L.ERR_UNSUPPORTED
    # null pointers with work to do
    assert lib.canary_volumerhs(0, 3, 4, 0, None, None, None, None, 1, ctypes.byref(p), None) == L.ERR_ARG
    assert lib.canary_facerhs(0, 3, 4, 0, None, None, None, 4, 2, None, None, None, ctypes.byref(p), None) == L.ERR_ARG
    # zero elements is a no-op
    assert lib.canary_updatesolution(0, 2, 3, 0, None, None, None, 0, 0.0, 0.0, 0.0, 0, None) == 0
    # the contracted FP64 form has every operation; the matrix-core one (3D
    # Euler only) the volume term and fused stages, so its facerhs and
    # updatesolution are refused by the unit, after the argument checks
    for prec in (L.FP64_FMA, L.FP64_MFMA):
        assert lib.canary_volumerhs(0, 3, 4, prec, None, None, None, None, 1, ctypes.byref(p), None) == L.ERR_ARG
        assert lib.canary_updatesolution(0, 3, 3, prec, None, None, None, 1, 0.0, 0.0, 0.0, 0, None) == L.ERR_ARG
        assert lib.canary_facerhs(0, 3, 4, prec, None, None, None, 4, 4, None, None, None, ctypes.byref(p), None) == L.ERR_ARG
    assert lib.canary_updatesolution(0, 2, 3, L.FP64_MFMA, None, None, None, 1, 0.0, 0.0, 0.0, 0, None) == L.ERR_UNSUPPORTED
    assert lib.canary_facerhs(1, 2, 4, L.FP64_MFMA, None, None, None, 4, 4, None, None, None, ctypes.byref(p), None) == L.ERR_UNSUPPORTED
    for eqn, dim in ((0, 2), (1, 2), (2, 2)):
        assert lib.canary_volumerhs(eqn, dim, 4, L.FP64_MFMA, None, None, None, None, 1, ctypes.byref(p), None) == L.ERR_UNSUPPORTED
        assert lib.canary_stage_fused(eqn, dim, 4, L.FP64_MFMA, None, None, None, None, None, None, 1, 1,
                                      None, None, None, 0.0, 0.0, 0.0, 0, ctypes.byref(p), None) == L.ERR_UNSUPPORTED
    for eqn, dim in ((0, 3), (0, 2), (1, 2)):
        assert lib.canary_stage_fused(eqn, dim, 4, L.FP64_FMA, None, None, None, None, None, None, 0, 0,
                                      None, None, None, 0.0, 0.0, 0.0, 0, ctypes.byref(p), None) == 0
        assert lib.canary_stage_fused(eqn, dim, 4, L.FP64_FMA, None, None, None, None, None, None, 1, 1,
                                      None, None, None, 0.0, 0.0, 0.0, 0, ctypes.byref(p), None) == L.ERR_ARG
    assert lib.canary_stage_fused(0, 3, 8, L.FP64_FMA, None, None, None, None, None, None, 1, 1,
                                  None, None, None, 0.0, 0.0, 0.0, 0, ctypes.byref(p), None) == L.ERR_UNSUPPORTED
    for prec in (L.FP64, L.FP64_FMA):
        assert lib.canary_facerhs_update(0, 3, 4, prec, None, None, None, None, None, 0, 0, None,
                                         None, None, 0.0, 0.0, 0.0, 0, ctypes.byref(p), None) == 0
        assert lib.canary_facerhs_update(1, 2, 4, prec, None, None, None, None, None, 3, 3, None,
                                         None, None, 0.0, 0.0, 0.0, 0, ctypes.byref(p), None) == L.ERR_ARG
    assert lib.canary_facerhs_update(0, 3, 4, L.FP32, None, None, None, None, None, 3, 3, None,
                                     None, None, 0.0, 0.0, 0.0, 0, ctypes.byref(p), None) == L.ERR_UNSUPPORTED


def test_lsrk54_tables_match_the_host_ones():
    import canary_amd.kernels as K
    import canary_amd.problems as P
    assert K.lsrk54_coefs() == P.rk_coefs()
    lib = L.lib()
    alt = ctypes.c_int(7)
    p = L.default_phys()
    # argument checks before any launch
    assert lib.canary_lsrk54(0, 3, 4, 0, 1, 5, 0.1, None, None, None, None, None, None, 1, 1,
                             None, None, None, ctypes.byref(p), 0, 0, ctypes.byref(alt),
                             None) == L.ERR_ARG  # fused path needs Qalt
    assert lib.canary_lsrk54(0, 3, 4, L.FP32, 0, 5, 0.1, None, None, None, None, None, None,
                             1, 1, None, None, None, ctypes.byref(p), 0, 0, ctypes.byref(alt),
                             None) == L.ERR_UNSUPPORTED  # FP64 forms only
    assert lib.canary_lsrk54(0, 3, 4, 0, 0, 0, 0.1, None, None, None, None, None, None, 1, 1,
                             None, None, None, ctypes.byref(p), 0, 0, ctypes.byref(alt),
                             None) == 0 and alt.value == 0  # no steps


def test_batch_run_rejects_bad_arguments_before_launch():
    lib = L.lib()
    cls = (L.BatchClass * 1)()
    assert lib.canary_batch_run(7, 0, 3, 0, None, cls, 1, 1, 0, None) == L.ERR_ARG  # kind
    assert lib.canary_batch_run(L.BATCH_DROPIN, 0, 3, 0, None, cls, 1, 1, 2, None) == L.ERR_ARG
    assert lib.canary_batch_run(L.BATCH_P32, 0, 3, 0, None, cls, -1, 1, 0, None) == L.ERR_ARG
    assert lib.canary_batch_run(L.BATCH_DROPIN, 0, 3, 0, None, cls, 1, 1, 0, None) == L.ERR_ARG
    assert lib.canary_batch_run(L.BATCH_DROPIN, 0, 3, 0, None, cls, 1, 0, 0, None) == 0  # no steps
    assert lib.canary_batch_run(L.BATCH_P32, 0, 3, 0, None, None, 0, 3, 1, None) == 0  # no classes
    # a class with no compiled (N, epw) variant is refused before anything runs
    # (the items pointer is only checked for null, never read on the host)
    solves = ctypes.c_void_p(64)
    for kind, N, epw in ((L.BATCH_DROPIN, 4, 999), (L.BATCH_DROPIN, 9, 1), (L.BATCH_P32, 4, 999)):
        cls = (L.BatchClass * 2)()
        cls[0].items, cls[0].nitems, cls[0].N, cls[0].epw = 64, 0, 4, 1
        cls[1].items, cls[1].nitems, cls[1].N, cls[1].epw = 64, 5, N, epw
        assert lib.canary_batch_run(kind, 0, 3, 0, solves, cls, 2, 1, 0, None) == L.ERR_UNSUPPORTED
    cls = (L.BatchClass * 1)()
    cls[0].items, cls[0].nitems, cls[0].N, cls[0].epw = None, 5, 4, 1
    assert lib.canary_batch_run(L.BATCH_DROPIN, 0, 3, 0, solves, cls, 1, 1, 0, None) == L.ERR_ARG
    # the single-launch kind: an (eqn, dim, prec) without the kernel is refused
    cls[0].items = 64
    assert lib.canary_batch_run(L.BATCH_DROPIN_ALL, 1, 3, 0, solves, cls, 1, 1, 0,
                                None) == L.ERR_UNSUPPORTED  # 3D shallow water


def test_single_launch_variants():
    """Each order's single-launch variant is its fullest 256-lane one; 3D
    orders 6 and 7 (343 and 512 nodes) have none."""
    lib = L.lib()
    e, t = ctypes.c_int(), ctypes.c_int()
    want2 = {1: 64, 2: 28, 3: 16, 4: 10, 5: 7, 6: 5, 7: 4}
    want3 = {1: 32, 2: 9, 3: 4, 4: 2, 5: 1}
    for prec in (L.FP64, L.FP64_FMA):
        for N in range(1, 8):
            assert lib.canary_batch_all_variant(L.SWE, 2, prec, N, ctypes.byref(e),
                                                ctypes.byref(t)) == 0
            assert (e.value, t.value) == (want2[N], 256)
            rc = lib.canary_batch_all_variant(L.EULER, 3, prec, N, ctypes.byref(e),
                                              ctypes.byref(t))
            if N <= 5:
                assert rc == 0 and (e.value, t.value) == (want3[N], 256)
            else:
                assert rc == L.ERR_UNSUPPORTED
    assert lib.canary_batch_all_variant(L.EULER, 3, L.FP32, 4, ctypes.byref(e),
                                        ctypes.byref(t)) == L.ERR_UNSUPPORTED
    assert lib.canary_batch_stage_all(L.SWE, 2, L.FP64, None, None, 0, 0, 0.0, 0.0, None) == 0
    assert lib.canary_batch_stage_all(L.SWE, 2, L.FP64, None, None, 3, 0, 0.0, 0.0,
                                      None) == L.ERR_ARG
    assert lib.canary_batch_stage_all(L.SWE, 2, L.FP64, None, None, 1, 2, 0.0, 0.0,
                                      None) == L.ERR_ARG


def test_phys_defaults_are_the_reference_rationals():
    p = L.default_phys()
    assert (p.gamma, p.p0, p.R_gas, p.c_p, p.c_v, p.gravity) == (14 / 10, 100000.0, 28717 / 100,
                                                                  100467 / 100, 7175 / 10, 10.0)


def test_no_kernel_is_instantiated_in_two_units():
    """Each unit has its own code object, but a kernel's host launch stub is
    one (weak) symbol across the library: a kernel instantiated in two units
    (say the exact-order and the contracted one) launches whichever unit's
    code the loader picked.  Every stub must come from one unit."""
    import shutil
    import subprocess
    import pytest
    build = os.path.join(ROOT, "canary.jl_amd", "_build")
    objs = sorted(glob.glob(os.path.join(build, "hip_*.o")))
    if not objs or shutil.which("nm") is None:
        pytest.skip("no built objects or no nm")
    owner = {}
    for o in objs:
        out = subprocess.run(["nm", "--defined-only", o], capture_output=True, text=True,
                             check=True).stdout
        for line in out.splitlines():
            parts = line.split()
            if len(parts) == 3 and "__device_stub__" in parts[2]:
                owner.setdefault(parts[2], set()).add(os.path.basename(o))
    assert len(owner) > 100
    shared = {s: u for s, u in owner.items() if len(u) > 1}
    assert not shared, list(shared.items())[:3]


def test_lsrk54_refuses_ghost_elements_before_launch():
    """The library loop exchanges no ghosts: nelem_total != nreal is refused
    on the host, whatever the path (nothing is launched or dereferenced)."""
    lib = L.lib()
    p = L.default_phys()
    alt = ctypes.c_int(-1)
    dummy = ctypes.c_void_p(0x1000)
    for path in (L.PATH_KERNELS, L.PATH_FUSED, L.PATH_FACE_UPDATE):
        rc = lib.canary_lsrk54(0, 3, 4, L.FP64, path, 1, 0.1, dummy, ctypes.c_void_p(0x2000), dummy,
                               dummy, dummy, dummy, 8, 9, None, dummy, dummy, ctypes.byref(p), 0, 0,
                               ctypes.byref(alt), None)
        assert rc == L.ERR_ARG, path
    # the shallow-water velocity freeze is refused for the Euler sets
    rc = lib.canary_lsrk54(0, 3, 4, L.FP64, L.PATH_FUSED, 1, 0.1, dummy, ctypes.c_void_p(0x2000),
                           dummy, dummy, dummy, dummy, 8, 8, None, dummy, dummy, ctypes.byref(p), 1,
                           0, ctypes.byref(alt), None)
    assert rc == L.ERR_ARG

