"""The C-ABI library builds, loads on the CPU and exports every entry point
that include/iemic_amd.h declares (no device calls)."""
import ctypes
import os
import re

from gpu_helpers import ROOT, iemic_amd


def declared_functions():
    txt = open(os.path.join(ROOT, "include", "iemic_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(iemic_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(iemic_amd.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_signatures_cover_the_header():
    lib = iemic_amd.load()
    for n in declared_functions():
        assert getattr(lib, n).argtypes is not None, n


def test_argument_checks_without_a_gpu():
    """Entry points reject bad arguments before touching the device."""
    import ctypes as C
    import iemic_amd
    lib = iemic_amd.load()
    dummy = (C.c_void_p * 2)(C.c_void_p(1), C.c_void_p(1))
    ptrs = (C.c_void_p * 2)(C.c_void_p(8), C.c_void_p(16))
    rc = lib.iemic_newton_step_many(dummy, 2, ptrs, ptrs, None, None, 1)
    assert rc != 0 and b"twice" in lib.iemic_last_error()
    assert lib.iemic_newton_step_many(None, 0, None, None, None, None, 0) != 0


def test_default_parameters_follow_the_reference():
    """iemic_continuation_defaults follows Continuation::getDefaultInitParameters
    (Continuation.H:1330-1385): initial step 1e-2, minimum 1e-8, maximum 1e3,
    epsilon increment 1e-5, maximum Newton iterations 7, optimal 3.5, Newton
    tolerance 1e-4, state tangent scaling 1, predictor bound 1e3, parameter
    'Combined Forcing' (COMB, index 19 of par.F90).  The solver defaults
    solve to Ocean::solve's 1e-8."""
    import iemic_amd
    cp = iemic_amd.Ocean.continuation_params()
    assert (cp.par, cp.ds, cp.ds_min, cp.ds_max) == (19, 1e-2, 1e-8, 1e3)
    assert (cp.eps, cp.newton_maxit, cp.opt_newton, cp.newton_tol) == (1e-5, 7, 3.5, 1e-4)
    assert (cp.tan_scaling, cp.predictor_bound) == (1.0, 1e3)
    sp = iemic_amd.Ocean.solver_params()
    assert sp.tol == 1e-8 and sp.precision == 0
