"""ctypes binding of libiemic_amd.so (include/iemic_amd.h).

`Ocean` mirrors the reference's Model/Ocean facade (src/utils/Model.H:55-113,
src/ocean/Ocean.H:175): computeRHS, computeJacobian, applyMatrix,
applyMassMat, buildPreconditioner, applyPrecon, solve, getPar/setPar.  Vectors
cross the boundary as float64 numpy arrays in the reference's unknown order;
the *_dev entry points are reachable through `lib` for device-resident
callers.  Loading fails loudly when the HIP library is missing: there is no
CPU fallback.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.normpath(os.path.join(_HERE, "..", "..", "libiemic_amd.so"))

PAR_INDEX = {
    "AL_T": 1, "Rayleigh-Number": 2, "Vertical Ekman-Number": 3, "Horizontal Ekman-Number": 4,
    "Rossby-Number": 5, "MIXP": 6, "RESC": 7, "SPL1": 8, "Salinity Homotopy": 9,
    "Solar Forcing": 10, "Horizontal Peclet-Number": 11, "Vertical Peclet-Number": 12,
    "P_VC": 13, "LAMB": 14, "Salinity Forcing": 15, "Wind Forcing": 16,
    "Temperature Forcing": 17, "Nonlinear Factor": 18, "Combined Forcing": 19, "ARCL": 20,
    "NLES": 21, "IFRICB": 22, "CONT": 23, "Energy": 24, "ALPC": 25, "CMPR": 26,
    "Flux Perturbation": 27, "Salinity Perturbation": 28, "MKAP": 29, "SPL2": 30,
}
# short Fortran names of par.F90:32-67
PAR_SHORT = dict(AL_T=1, RAYL=2, EK_V=3, EK_H=4, ROSB=5, MIXP=6, RESC=7, SPL1=8, HMTP=9,
                 SUNP=10, PE_H=11, PE_V=12, P_VC=13, LAMB=14, SALT=15, WIND=16, TEMP=17,
                 BIOT=18, COMB=19, ARCL=20, NLES=21, IFRICB=22, CONT=23, ENER=24, ALPC=25,
                 CMPR=26, FPER=27, SPER=28, MKAP=29, SPL2=30)


class Config(C.Structure):
    _fields_ = [("n", C.c_int), ("m", C.c_int), ("l", C.c_int),
                ("xmin_deg", C.c_double), ("xmax_deg", C.c_double),
                ("ymin_deg", C.c_double), ("ymax_deg", C.c_double),
                ("periodic", C.c_int), ("hdim", C.c_double), ("qz", C.c_double),
                ("alphaT", C.c_double), ("alphaS", C.c_double),
                ("ih", C.c_int), ("vmix", C.c_int), ("tap", C.c_int),
                ("rho_mixing", C.c_int), ("coriolis_on", C.c_int),
                ("tres", C.c_int), ("sres", C.c_int), ("iza", C.c_int), ("ite", C.c_int),
                ("its", C.c_int), ("coupled_t", C.c_int), ("coupled_s", C.c_int),
                ("forcing_type", C.c_int), ("int_sign", C.c_int), ("nic", C.c_int),
                ("mic", C.c_int)]


class SolverParams(C.Structure):
    _fields_ = [("tol", C.c_double), ("restart", C.c_int), ("maxit", C.c_int),
                ("precision", C.c_int), ("verbose", C.c_int)]


class SolveInfo(C.Structure):
    _fields_ = [("iterations", C.c_int), ("restarts", C.c_int), ("rel_residual", C.c_double),
                ("t_precon_setup", C.c_double), ("t_solve", C.c_double), ("converged", C.c_int)]


class ContinuationParams(C.Structure):
    _fields_ = [("par", C.c_int), ("target", C.c_double), ("ds", C.c_double), ("ds_min", C.c_double),
                ("ds_max", C.c_double), ("max_steps", C.c_int), ("newton_maxit", C.c_int),
                ("newton_tol", C.c_double), ("eps", C.c_double), ("tan_scaling", C.c_double),
                ("opt_newton", C.c_double), ("predictor_bound", C.c_double)]


class ContinuationInfo(C.Structure):
    _fields_ = [("steps", C.c_int), ("rejected", C.c_int), ("newton_iterations", C.c_int),
                ("linear_iterations", C.c_int), ("reached", C.c_int), ("par", C.c_double),
                ("seconds", C.c_double)]


PREC_ROW_LU, PREC_COLUMN_ILU = 0, 1
DIAG_MEAN_SST, DIAG_MAX_PSIM = 0, 1


class PrecondInfo(C.Structure):
    _fields_ = [("kind", C.c_int), ("blocks", C.c_int), ("max_block", C.c_int), ("pins", C.c_int),
                ("levels", C.c_int), ("factor_bytes", C.c_double)]


_P = C.POINTER
_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built (run i-emic_amd/build.py)")
    lib = C.CDLL(LIB_PATH)
    vp, dp, ip = C.c_void_p, _P(C.c_double), _P(C.c_int)
    sig = {
        "iemic_last_error": ([], C.c_char_p),
        "iemic_device_count": ([ip], C.c_int),
        "iemic_create": ([_P(Config), ip, dp, dp, dp, dp, dp, _P(vp)], C.c_int),
        "iemic_destroy": ([vp], C.c_int),
        "iemic_set_par": ([vp, C.c_int, C.c_double], C.c_int),
        "iemic_get_par": ([vp, C.c_int, dp], C.c_int),
        "iemic_ndim": ([vp, ip], C.c_int),
        "iemic_graph_nnz": ([vp, ip], C.c_int),
        "iemic_rowintcon": ([vp, ip], C.c_int),
        "iemic_intcond_coef": ([vp, dp], C.c_int),
        "iemic_forcing": ([vp, dp], C.c_int),
        "iemic_set_intcond_correction": ([vp, dp], C.c_int),
        "iemic_set_intcond_correction_value": ([vp, C.c_double], C.c_int),
        "iemic_rhs": ([vp, dp, dp], C.c_int),
        "iemic_rhs_dev": ([vp, vp, vp], C.c_int),
        "iemic_rhs_dev_noalloc": ([vp, vp, vp], C.c_int),
        "iemic_thcm_get_array_sizes": ([vp, ip, ip], C.c_int),
        "iemic_mrilu_create": ([ip, C.c_int, C.c_int, ip, ip, dp], C.c_int),
        "iemic_mrilu_set_params": ([C.c_int] * 6 + [C.c_double] * 3 + [C.c_int] + [C.c_double] * 7 + [C.c_int]
                                   + [C.c_double] * 4 + [C.c_int] * 2, C.c_int),
        "iemic_mrilu_compute": ([C.c_int], C.c_int),
        "iemic_mrilu_apply": ([C.c_int, C.c_int, dp, dp], C.c_int),
        "iemic_mrilu_info": ([C.c_int, ip, ip, ip], C.c_int),
        "iemic_mrilu_destroy": ([ip], C.c_int),
        "iemic_thcm_set_pointers": ([vp, C.c_int, C.c_int, ip, ip, dp, dp, ip, ip, dp], C.c_int),
        "iemic_thcm_rhs": ([vp, dp, dp], C.c_int),
        "iemic_thcm_matrix": ([vp, dp], C.c_int),
        "iemic_jacobian": ([vp, dp], C.c_int),
        "iemic_jacobian_dev": ([vp, vp], C.c_int),
        "iemic_jacobian_csr": ([vp, ip, ip, dp, dp], C.c_int),
        "iemic_apply_matrix": ([vp, dp, dp], C.c_int),
        "iemic_apply_matrix_dev": ([vp, vp, vp], C.c_int),
        "iemic_apply_matrix_transpose_dev": ([vp, vp, vp], C.c_int),
        "iemic_apply_mass": ([vp, dp, dp], C.c_int),
        "iemic_pressure_projection": ([vp, dp], C.c_int),
        "iemic_solve_multi": ([vp, vp, C.c_int, dp, dp, vp], C.c_int),
        "iemic_solve_multi_dev": ([vp, vp, C.c_int, vp, vp, vp], C.c_int),
        "iemic_solve_multi_transpose": ([vp, vp, C.c_int, dp, dp, vp], C.c_int),
        "iemic_solve_multi_transpose_dev": ([vp, vp, C.c_int, vp, vp, vp], C.c_int),
        "iemic_solve_block": ([vp, vp, C.c_int, dp, dp, vp], C.c_int),
        "iemic_solve_block_dev": ([vp, vp, C.c_int, vp, vp, vp], C.c_int),
        "iemic_solve_block_transpose": ([vp, vp, C.c_int, dp, dp, vp], C.c_int),
        "iemic_solve_block_transpose_dev": ([vp, vp, C.c_int, vp, vp, vp], C.c_int),
        "iemic_block_gram_dev": ([C.c_int, C.c_int, C.c_int, vp, vp, vp], C.c_int),
        "iemic_block_update_dev": ([C.c_int, C.c_int, C.c_int, C.c_double, vp, vp, C.c_double, vp], C.c_int),
        "iemic_vec_axpby_dev": ([C.c_longlong, C.c_double, vp, C.c_double, vp], C.c_int),
        "iemic_vec_dot_dev": ([C.c_longlong, vp, vp, dp], C.c_int),
        "iemic_apply_precon_multi_dev": ([vp, C.c_int, vp, vp], C.c_int),
        "iemic_newton_dev": ([vp, vp, C.c_double, C.c_double, C.c_int, vp, ip, dp], C.c_int),
        "iemic_continuation_defaults": ([vp], None),
        "iemic_set_precond_reuse": ([vp, C.c_int], C.c_int),
        "iemic_apply_tuning": ([vp, C.c_int, ip], C.c_int),
        "iemic_spmm_tuning": ([vp, C.c_int, ip], C.c_int),
        "iemic_read_mask": ([C.c_char_p, C.c_int, C.c_int, C.c_int, ip], C.c_int),
        "iemic_read_state": ([C.c_char_p, dp, C.c_int64, dp, ip], C.c_int),
        "iemic_write_state": ([C.c_char_p, dp, C.c_int64, dp], C.c_int),
        "iemic_newton_step_many": ([vp, C.c_int, vp, vp, vp, vp, C.c_int], C.c_int),
        "iemic_continuation_dev": ([vp, vp, vp, vp, dp, dp, vp], C.c_int),
        "iemic_pressure_projection_dev": ([vp, vp], C.c_int),
        "iemic_set_preconditioner": ([vp, C.c_int], C.c_int),
        "iemic_preconditioner_info": ([vp, _P(PrecondInfo)], C.c_int),
        "iemic_build_preconditioner": ([vp], C.c_int),
        "iemic_apply_precon": ([vp, dp, dp], C.c_int),
        "iemic_apply_precon_dev": ([vp, vp, vp], C.c_int),
        "iemic_apply_precon_fp32": ([vp, dp, dp], C.c_int),
        "iemic_solver_defaults": ([_P(SolverParams)], None),
        "iemic_solve": ([vp, _P(SolverParams), dp, dp, _P(SolveInfo)], C.c_int),
        "iemic_solve_dev": ([vp, _P(SolverParams), vp, vp, _P(SolveInfo)], C.c_int),
        "iemic_solve_transpose": ([vp, _P(SolverParams), dp, dp, _P(SolveInfo)], C.c_int),
        "iemic_solve_transpose_dev": ([vp, _P(SolverParams), vp, vp, _P(SolveInfo)], C.c_int),
        "iemic_apply_precon_transpose_dev": ([vp, vp, vp], C.c_int),
        "iemic_diagnostic": ([vp, C.c_int, dp, dp, dp], C.c_int),
        "iemic_dev_alloc": ([_P(vp), C.c_longlong], C.c_int),
        "iemic_dev_free": ([vp], C.c_int),
        "iemic_memcpy_h2d": ([vp, vp, C.c_longlong], C.c_int),
        "iemic_memcpy_d2h": ([vp, vp, C.c_longlong], C.c_int),
        "iemic_synchronize": ([vp], C.c_int),
        "iemic_last_kernel_ms": ([vp, C.c_int, dp], C.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    _lib = lib
    return lib


class IemicError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"iemic error {code}: {msg}")
        self.code = code


def _check(rc):
    if rc != 0:
        raise IemicError(rc, load().iemic_last_error().decode())
    return rc


def _d(a):
    return a.ctypes.data_as(_P(C.c_double))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def newton_step_many(oceans, dxs, ddxs, params=None, concurrent=True):
    """One Newton step (F, J, preconditioner, J dx = F) for each model, on
    one host thread and stream each (iemic_newton_step_many); returns the
    SolveInfo of each."""
    q = len(oceans)
    hs = (C.c_void_p * q)(*[o.h.value for o in oceans])
    xs = (C.c_void_p * q)(*[d.ptr.value for d in dxs])
    ys = (C.c_void_p * q)(*[d.ptr.value for d in ddxs])
    infos = (SolveInfo * q)()
    p = params if params is not None else oceans[0].solver_params()
    _check(load().iemic_newton_step_many(hs, q, xs, ys, C.byref(p), infos, int(bool(concurrent))))
    return list(infos)


def read_mask(path, n, m, l):
    """A data/mkmask land mask as iemic_create takes it (iemic_read_mask)."""
    out = np.zeros((n + 2) * (m + 2) * (l + 2), np.int32)
    _check(load().iemic_read_mask(str(path).encode(), n, m, l, out.ctypes.data_as(_P(C.c_int))))
    return out


def read_state(path, ndim, parameters=True):
    """A state file in the reference's EpetraExt-HDF5 layout
    (iemic_read_state, after Model::loadStateFromFile): returns (x, pars)
    with pars {par.F90 index: value} for the parameters the file holds (the
    file must then have a /Parameters group, as the reference requires);
    parameters=False reads the state alone."""
    x = np.zeros(ndim)
    par = np.zeros(31)
    found = np.zeros(31, np.int32)
    pp = par.ctypes.data_as(_P(C.c_double)) if parameters else None
    fp = found.ctypes.data_as(_P(C.c_int)) if parameters else None
    _check(load().iemic_read_state(str(path).encode(), x.ctypes.data_as(_P(C.c_double)), ndim, pp, fp))
    return x, {i: float(par[i]) for i in range(1, 31) if found[i]}


def write_state(path, x, pars=None):
    """Writes x and the parameters {par.F90 index: value} (all 30, or none)
    in the reference's state-file layout (iemic_write_state, after
    Model::saveStateToFile)."""
    x = np.ascontiguousarray(x, np.float64)
    pp = None
    if pars is not None:
        if sorted(pars) != list(range(1, 31)):
            raise ValueError("write_state needs all 30 parameters (par.F90 indices 1..30)")
        arr = np.zeros(31)
        for i, v in pars.items():
            arr[i] = v
        pp = arr.ctypes.data_as(_P(C.c_double))
    _check(load().iemic_write_state(str(path).encode(), x.ctypes.data_as(_P(C.c_double)), x.size, pp))


class MriluPrec:
    """The MRILU plug-in interface (iemic_mrilu_*): create from a CSR matrix,
    compute, apply, destroy; block ILU(0) on the GPU."""

    DEFAULTS = dict(blocksize=1, cutmck=0, scarow=0, xactelm=0, clsonce=0, nlsfctr=1.0, epsw=1.0e-2,
                    elmfctr=0.1, gusmod=0, gusfctr=1.0, redfctr=0.8, schtol=1.0e-6, denslim=1.0e-3,
                    globfrac=1.0, locfrac=1.0, sparslim=0.675, ilutype=9, droptol=1.0e-4, compfct=1.0,
                    cpivtol=0.1, lutol=1.0e-14, singlu=1, outlev=2)

    def __init__(self, A, **params):
        """A: a scipy.sparse matrix (any format, converted to CSR)."""
        import scipy.sparse as sp
        self.lib = load()
        A = sp.csr_matrix(A)
        A.sort_indices()
        self.n = A.shape[0]
        self.id = C.c_int(0)
        beg = A.indptr.astype(np.int32)
        jco = A.indices.astype(np.int32)
        co = A.data.astype(np.float64)
        ip = _P(C.c_int)
        _check(self.lib.iemic_mrilu_create(C.byref(self.id), self.n, len(co), beg.ctypes.data_as(ip),
                                           jco.ctypes.data_as(ip), _d(co)))
        p = dict(self.DEFAULTS)
        p.update(params)
        _check(self.lib.iemic_mrilu_set_params(self.id.value, *[p[k] for k in self.DEFAULTS]))

    def compute(self):
        _check(self.lib.iemic_mrilu_compute(self.id.value))

    def apply(self, v):
        v = _f64(v)
        out = np.empty(self.n)
        _check(self.lib.iemic_mrilu_apply(self.id.value, self.n, _d(v), _d(out)))
        return out

    def info(self):
        b, lv, fl = C.c_int(), C.c_int(), C.c_int()
        _check(self.lib.iemic_mrilu_info(self.id.value, C.byref(b), C.byref(lv), C.byref(fl)))
        return b.value, lv.value, fl.value

    def close(self):
        if self.id.value:
            self.lib.iemic_mrilu_destroy(C.byref(self.id))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceArray:
    """A float64 device buffer owned by the library's allocator."""

    def __init__(self, n):
        self.n = int(n)
        self.ptr = C.c_void_p()
        _check(load().iemic_dev_alloc(C.byref(self.ptr), 8 * self.n))

    def upload(self, a):
        a = _f64(a)
        assert a.size == self.n
        _check(load().iemic_memcpy_h2d(self.ptr, a.ctypes.data_as(C.c_void_p), 8 * self.n))
        return self

    def download(self):
        out = np.empty(self.n)
        _check(load().iemic_memcpy_d2h(out.ctypes.data_as(C.c_void_p), self.ptr, 8 * self.n))
        return out

    def view(self, start, n):
        """A non-owning view of elements [start, start + n)."""
        assert 0 <= start and start + n <= self.n
        v = DeviceArray.__new__(DeviceArray)
        v.n = int(n)
        v.ptr = C.c_void_p(self.ptr.value + 8 * int(start))
        v._owner = self
        return v

    def free(self):
        if self.ptr and not hasattr(self, "_owner"):
            load().iemic_dev_free(self.ptr)
        self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Ocean:
    """One THCM ocean model on the GPU (the reference's Ocean/THCM pair)."""

    def __init__(self, cfg, landm, taux=None, tauy=None, tatm=None, emip=None, spert=None,
                 pars=None):
        lib = load()
        self.lib = lib
        self.cfg = cfg
        landm = np.ascontiguousarray(landm, dtype=np.int32)
        fields = [None if f is None else _f64(f) for f in (taux, tauy, tatm, emip, spert)]
        self._keep = fields
        h = C.c_void_p()
        _check(lib.iemic_create(C.byref(cfg), landm.ctypes.data_as(_P(C.c_int)),
                                *[None if f is None else _d(f) for f in fields], C.byref(h)))
        self.h = h
        n = C.c_int()
        _check(lib.iemic_ndim(h, C.byref(n)))
        self.ndim = n.value
        for k, v in (pars or {}).items():
            self.setPar(k, v)

    def close(self):
        if getattr(self, "h", None):
            self.lib.iemic_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- parameters (THCM::setParameter / getParameter)
    @staticmethod
    def _idx(name):
        if isinstance(name, int):
            return name
        return PAR_INDEX.get(name, PAR_SHORT.get(name))

    def setPar(self, name, value):
        _check(self.lib.iemic_set_par(self.h, self._idx(name), float(value)))

    def getPar(self, name):
        v = C.c_double()
        _check(self.lib.iemic_get_par(self.h, self._idx(name), C.byref(v)))
        return v.value

    def loadStateFromFile(self, path):
        """Model::loadStateFromFile (Model.C:15-110): the state of a
        reference state file, with the parameters the file holds set on this
        model; returns the state (host array)."""
        x, pars = read_state(path, self.ndim)
        for i, v in pars.items():
            self.setPar(i, v)
        return x

    def saveStateToFile(self, path, x):
        """Model::saveStateToFile (Model.C:112-160): x and this model's 30
        parameters in the reference's state-file layout."""
        write_state(path, x, {i: self.getPar(i) for i in range(1, 31)})

    # ---- structure
    def rowintcon(self):
        r = C.c_int()
        _check(self.lib.iemic_rowintcon(self.h, C.byref(r)))
        return r.value

    def graph_nnz(self):
        r = C.c_int()
        _check(self.lib.iemic_graph_nnz(self.h, C.byref(r)))
        return r.value

    def intcond_coef(self):
        out = np.zeros(self.ndim)
        _check(self.lib.iemic_intcond_coef(self.h, _d(out)))
        return out

    def forcing(self):
        out = np.zeros(self.ndim)
        _check(self.lib.iemic_forcing(self.h, _d(out)))
        return out

    def setIntCondCorrection(self, x=None, value=None):
        if x is not None:
            x = _f64(x)
            _check(self.lib.iemic_set_intcond_correction(self.h, _d(x)))
        else:
            _check(self.lib.iemic_set_intcond_correction_value(self.h, float(value)))

    # ---- Model interface
    def computeRHS(self, x):
        x = _f64(x)
        F = np.empty(self.ndim)
        _check(self.lib.iemic_rhs(self.h, _d(x), _d(F)))
        return F

    def computeJacobian(self, x):
        x = _f64(x)
        _check(self.lib.iemic_jacobian(self.h, _d(x)))

    def getJacobian(self):
        """(rowptr, col, val, diagB) in the reference's CSR order."""
        nnz = self.graph_nnz()
        rp = np.empty(self.ndim + 1, np.int32)
        ci = np.empty(nnz, np.int32)
        val = np.empty(nnz)
        dB = np.empty(self.ndim)
        _check(self.lib.iemic_jacobian_csr(self.h, rp.ctypes.data_as(_P(C.c_int)),
                                           ci.ctypes.data_as(_P(C.c_int)), _d(val), _d(dB)))
        return rp, ci, val, dB

    def applyMatrix(self, v):
        v = _f64(v)
        out = np.empty(self.ndim)
        _check(self.lib.iemic_apply_matrix(self.h, _d(v), _d(out)))
        return out

    def applyMatrixTranspose_dev(self, dv, dout):
        _check(self.lib.iemic_apply_matrix_transpose_dev(self.h, dv.ptr, dout.ptr))

    def applyMatrixTranspose(self, v):
        dv = DeviceArray(self.ndim).upload(v)
        do = DeviceArray(self.ndim)
        _check(self.lib.iemic_apply_matrix_transpose_dev(self.h, dv.ptr, do.ptr))
        _check(self.lib.iemic_synchronize(self.h))
        return do.download()

    def applyMassMat(self, v):
        v = _f64(v)
        out = np.empty(self.ndim)
        _check(self.lib.iemic_apply_mass(self.h, _d(v), _d(out)))
        return out

    def pressureProjection(self, v):
        """v - (v.s1) s1 - (v.s2) s2 over the checkerboard pressure modes
        (Ocean::pressureProjection); returns a new array."""
        v = np.array(v, dtype=np.float64, copy=True)
        _check(self.lib.iemic_pressure_projection(self.h, _d(v)))
        return v

    def applyTuning(self, nrhs=1):
        """Block target the apply tuned for nrhs right-hand sides (0: none)."""
        b = C.c_int()
        _check(self.lib.iemic_apply_tuning(self.h, int(nrhs), C.byref(b)))
        return b.value

    def spmmTuning(self, nrhs):
        """Columns per thread the SpMM chose for nrhs columns (0: none)."""
        b = C.c_int()
        _check(self.lib.iemic_spmm_tuning(self.h, int(nrhs), C.byref(b)))
        return b.value

    def setPreconditionerReuse(self, max_age):
        """Keep row-LU factors for up to max_age newer Jacobians."""
        _check(self.lib.iemic_set_precond_reuse(self.h, int(max_age)))

    def setPreconditioner(self, kind):
        _check(self.lib.iemic_set_preconditioner(self.h, int(kind)))

    def preconditionerInfo(self):
        info = PrecondInfo()
        _check(self.lib.iemic_preconditioner_info(self.h, C.byref(info)))
        return info

    def buildPreconditioner(self):
        _check(self.lib.iemic_build_preconditioner(self.h))

    def applyPrecon(self, v):
        v = _f64(v)
        out = np.empty(self.ndim)
        _check(self.lib.iemic_apply_precon(self.h, _d(v), _d(out)))
        return out

    def applyPreconFP32(self, v):
        v = _f64(v)
        out = np.empty(self.ndim)
        _check(self.lib.iemic_apply_precon_fp32(self.h, _d(v), _d(out)))
        return out

    @staticmethod
    def solver_params(**kw):
        p = SolverParams()
        load().iemic_solver_defaults(C.byref(p))
        for k, v in kw.items():
            setattr(p, k, v)
        return p

    def solve(self, b, params=None, raise_on_noconv=True):
        """Ocean::solve: J x = b from x0 = 0; returns (x, SolveInfo)."""
        b = _f64(b)
        x = np.zeros(self.ndim)
        info = SolveInfo()
        p = params if params is not None else self.solver_params()
        rc = self.lib.iemic_solve(self.h, C.byref(p), _d(b), _d(x), C.byref(info))
        if rc != 0 and (raise_on_noconv or rc != -6):
            _check(rc)
        return x, info

    # ---- device-resident variants (DeviceArray arguments)
    def computeRHS_dev(self, dx, dF):
        _check(self.lib.iemic_rhs_dev(self.h, dx.ptr, dF.ptr))

    def computeRHS_dev_noalloc(self, dx, dF):
        """F(x) into dF with no device allocation (iemic_rhs_dev_noalloc)."""
        _check(self.lib.iemic_rhs_dev_noalloc(self.h, dx.ptr, dF.ptr))

    def fortranRHS(self, un):
        """rhs_: the Fortran right-hand side B (iemic_thcm_rhs)."""
        un = _f64(un)
        B = np.empty(self.ndim)
        _check(self.lib.iemic_thcm_rhs(self.h, _d(un), _d(B)))
        return B

    def fortranMatrix(self, un):
        """matrix_: (begA, jcoA, coA, coB) in the Fortran's 1-based CSR layout,
        through iemic_thcm_get_array_sizes / set_pointers / matrix."""
        un = _f64(un)
        nr, nz = C.c_int(), C.c_int()
        _check(self.lib.iemic_thcm_get_array_sizes(self.h, C.byref(nr), C.byref(nz)))
        beg = np.zeros(nr.value + 1, np.int32)
        jco = np.zeros(nz.value, np.int32)
        co = np.zeros(nz.value)
        coB = np.zeros(nr.value)
        ip = _P(C.c_int)
        _check(self.lib.iemic_thcm_set_pointers(self.h, nr, nz, beg.ctypes.data_as(ip), jco.ctypes.data_as(ip),
                                                _d(co), _d(coB), None, None, None))
        _check(self.lib.iemic_thcm_matrix(self.h, _d(un)))
        n = beg[-1] - 1
        return beg, jco[:n].copy(), co[:n].copy(), coB

    def computeJacobian_dev(self, dx):
        _check(self.lib.iemic_jacobian_dev(self.h, dx.ptr))

    def applyMatrix_dev(self, dv, dout):
        _check(self.lib.iemic_apply_matrix_dev(self.h, dv.ptr, dout.ptr))

    def applyPrecon_dev(self, dv, dout):
        _check(self.lib.iemic_apply_precon_dev(self.h, dv.ptr, dout.ptr))

    def solve_dev(self, db, dx, params=None, raise_on_noconv=True, transpose=False):
        """J x = b (J^T x = b if transpose) on device arrays."""
        info = SolveInfo()
        p = params if params is not None else self.solver_params()
        fn = self.lib.iemic_solve_transpose_dev if transpose else self.lib.iemic_solve_dev
        rc = fn(self.h, C.byref(p), db.ptr, dx.ptr, C.byref(info))
        if rc != 0 and (raise_on_noconv or rc != -6):
            _check(rc)
        return info

    def synchronize(self):
        _check(self.lib.iemic_synchronize(self.h))

    def solveTranspose(self, b, params=None, raise_on_noconv=True):
        """Adjoint solve J^T x = b from x0 = 0; returns (x, SolveInfo)."""
        b = _f64(b)
        x = np.zeros(self.ndim)
        info = SolveInfo()
        p = params if params is not None else self.solver_params()
        rc = self.lib.iemic_solve_transpose(self.h, C.byref(p), _d(b), _d(x), C.byref(info))
        if rc != 0 and (raise_on_noconv or rc != -6):
            _check(rc)
        return x, info

    def newton(self, x0, tol=1e-10, step_tol=1e-13, maxit=40, params=None, raise_on_noconv=True):
        """Damped Newton on device state (iemic_newton_dev); returns
        (x, iterations, ||F||)."""
        dx = DeviceArray(self.ndim).upload(x0)
        its, fn = C.c_int(), C.c_double()
        p = params if params is not None else self.solver_params(tol=1e-10, restart=50, maxit=200)
        rc = self.lib.iemic_newton_dev(self.h, C.byref(p), tol, step_tol, maxit, dx.ptr, C.byref(its), C.byref(fn))
        if rc != 0 and (raise_on_noconv or rc != -6):
            _check(rc)
        return dx.download(), its.value, fn.value

    @staticmethod
    def continuation_params(**kw):
        p = ContinuationParams()
        load().iemic_continuation_defaults(C.byref(p))
        for k, v in kw.items():
            setattr(p, k, v)
        return p

    def continuation(self, x0, cparams, params=None):
        """Pseudo-arclength continuation from the steady state x0 at the
        current parameter value (iemic_continuation_dev); returns (x,
        parameter history, ||x|| history, ContinuationInfo)."""
        dx = DeviceArray(self.ndim).upload(x0)
        ph = np.full(cparams.max_steps + 1, np.nan)
        nh = np.full(cparams.max_steps + 1, np.nan)
        info = ContinuationInfo()
        p = params if params is not None else self.solver_params(tol=1e-10, restart=50, maxit=200)
        _check(self.lib.iemic_continuation_dev(self.h, C.byref(p), C.byref(cparams), dx.ptr, _d(ph), _d(nh),
                                               C.byref(info)))
        k = info.steps + 1
        return dx.download(), ph[:k], nh[:k], info

    def solveMulti(self, B, params=None, raise_on_noconv=True, transpose=False, block=False):
        """J X = B (J^T X = B if transpose) for the columns of B (ndim x
        nrhs) in one batched solve; returns (X, [SolveInfo per column]).
        block: block FGMRES over one shared Krylov space (at most 32
        columns) instead of one FGMRES per column in lock step."""
        B = np.asarray(B, dtype=np.float64)
        if B.ndim == 1:
            B = B[:, None]
        nrhs = B.shape[1]
        Bc = np.ascontiguousarray(B.T)            # column q contiguous
        Xc = np.zeros_like(Bc)
        infos = (SolveInfo * nrhs)()
        p = params if params is not None else self.solver_params()
        if block:
            fn = self.lib.iemic_solve_block_transpose if transpose else self.lib.iemic_solve_block
        else:
            fn = self.lib.iemic_solve_multi_transpose if transpose else self.lib.iemic_solve_multi
        rc = fn(self.h, C.byref(p), nrhs, _d(Bc), _d(Xc), infos)
        if rc != 0 and (raise_on_noconv or rc != -6):
            _check(rc)
        return Xc.T.copy(), list(infos)

    def solve_multi_dev(self, dB, dX, nrhs, params=None, raise_on_noconv=True, transpose=False, block=False):
        """J X = B (J^T X = B if transpose) for nrhs columns stored one after
        another in device arrays (block: block FGMRES, at most 32 columns)."""
        infos = (SolveInfo * nrhs)()
        p = params if params is not None else self.solver_params()
        if block:
            fn = self.lib.iemic_solve_block_transpose_dev if transpose else self.lib.iemic_solve_block_dev
        else:
            fn = self.lib.iemic_solve_multi_transpose_dev if transpose else self.lib.iemic_solve_multi_dev
        rc = fn(self.h, C.byref(p), nrhs, dB.ptr, dX.ptr, infos)
        if rc != 0 and (raise_on_noconv or rc != -6):
            _check(rc)
        return list(infos)

    def applyPreconMulti(self, V):
        """M^{-1} applied to the columns of V (ndim x nrhs), one batched sweep."""
        V = np.asarray(V, dtype=np.float64)
        nrhs = V.shape[1]
        dv = DeviceArray(self.ndim * nrhs).upload(np.ascontiguousarray(V.T).reshape(-1))
        do = DeviceArray(self.ndim * nrhs)
        _check(self.lib.iemic_apply_precon_multi_dev(self.h, nrhs, dv.ptr, do.ptr))
        _check(self.lib.iemic_synchronize(self.h))
        return do.download().reshape(nrhs, self.ndim).T.copy()

    def applyPreconTranspose(self, v):
        dv = DeviceArray(self.ndim).upload(v)
        do = DeviceArray(self.ndim)
        _check(self.lib.iemic_apply_precon_transpose_dev(self.h, dv.ptr, do.ptr))
        _check(self.lib.iemic_synchronize(self.h))
        return do.download()

    def diagnostic(self, which, x):
        """(g(x), dg/dx) for DIAG_MEAN_SST or DIAG_MAX_PSIM."""
        x = _f64(x)
        v = C.c_double()
        g = np.empty(self.ndim)
        _check(self.lib.iemic_diagnostic(self.h, int(which), _d(x), C.byref(v), _d(g)))
        return v.value, g

    def kernel_ms(self, which):
        v = C.c_double()
        _check(self.lib.iemic_last_kernel_ms(self.h, which, C.byref(v)))
        return v.value
