"""Shared test infrastructure: ctypes bindings to the CPU oracle
(oracle/libthcm_oracle.so), to the reference build (oracle/_ref/libthcm_ref.so,
present only where /root/reference was available at build time) and the
seeded test cases.

The oracle and the reference build are checkers only: nothing in the product
imports this module.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "libthcm_oracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libthcm_ref.so")
GOLDEN = os.path.join(ROOT, "tests", "golden")

OCEAN, LAND, PERIO = 0, 1, 3
UU, VV, WW, PP, TT, SS = range(1, 7)
# parameter indices of par.F90:32-67
PAR = dict(AL_T=1, RAYL=2, EK_V=3, EK_H=4, ROSB=5, MIXP=6, RESC=7, SPL1=8, HMTP=9,
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


def make_config(**kw):
    """THCM::getDefaultInitParameters defaults (THCM.C:2697-2765)."""
    d = dict(n=16, m=16, l=16, xmin_deg=286.0, xmax_deg=350.0, ymin_deg=10.0, ymax_deg=74.0,
             periodic=0, hdim=4000.0, qz=1.0, alphaT=1.0e-4, alphaS=7.6e-4, ih=0, vmix=1,
             tap=1, rho_mixing=1, coriolis_on=1, tres=1, sres=1, iza=2, ite=1, its=1,
             coupled_t=0, coupled_s=0, forcing_type=0, int_sign=-1, nic=-1, mic=-1)
    d.update(kw)
    return Config(**d)


def config_dict(cfg):
    return {f: getattr(cfg, f) for f, _ in Config._fields_}


_P = C.POINTER


def _ptr(a, t):
    return a.ctypes.data_as(_P(t))


def load_oracle():
    lib = C.CDLL(ORACLE_SO)
    lib.tho_create.restype = C.c_void_p
    lib.tho_create.argtypes = [_P(Config), _P(C.c_int)] + [_P(C.c_double)] * 5
    lib.tho_destroy.argtypes = [C.c_void_p]
    lib.tho_set_par.argtypes = [C.c_void_p, C.c_int, C.c_double]
    lib.tho_get_par.argtypes = [C.c_void_p, C.c_int]
    lib.tho_get_par.restype = C.c_double
    for f in ("tho_ndim", "tho_rowintcon", "tho_graph_nnz"):
        getattr(lib, f).argtypes = [C.c_void_p]
    lib.tho_graph.argtypes = [C.c_void_p, _P(C.c_int), _P(C.c_int)]
    lib.tho_intcond_coef.argtypes = [C.c_void_p, _P(C.c_double)]
    lib.tho_landm.argtypes = [C.c_void_p, _P(C.c_int)]
    lib.tho_forcing.argtypes = [C.c_void_p, _P(C.c_double)]
    lib.tho_rhs.argtypes = [C.c_void_p, _P(C.c_double), C.c_double, _P(C.c_double)]
    lib.tho_jacobian.argtypes = [C.c_void_p, _P(C.c_double), _P(C.c_double), _P(C.c_double)]
    lib.tho_mix.argtypes = [C.c_void_p, _P(C.c_double), _P(C.c_double)]
    return lib


def ref_available():
    return os.path.exists(REF_SO)


_ref_lib = None


def load_ref():
    global _ref_lib
    if _ref_lib is None:
        lib = C.CDLL(REF_SO)
        lib.thcmref_reference_landm.argtypes = [_P(Config), C.c_int, C.c_int, C.c_char_p, _P(C.c_int)]
        lib.thcmref_init.argtypes = [_P(Config), _P(C.c_int)] + [_P(C.c_double)] * 5
        lib.thcmref_set_par.argtypes = [C.c_int, C.c_double]
        lib.thcmref_get_par.argtypes = [C.c_int]
        lib.thcmref_get_par.restype = C.c_double
        lib.thcmref_intcond_coef.argtypes = [_P(C.c_double)]
        lib.thcmref_rhs.argtypes = [_P(C.c_double), C.c_double, _P(C.c_double)]
        lib.thcmref_jacobian.argtypes = [_P(C.c_double), _P(C.c_int), _P(C.c_int),
                                         _P(C.c_double), _P(C.c_double)]
        lib.thcmref_fortran_csr.argtypes = [_P(C.c_int), _P(C.c_int), _P(C.c_double)]
        lib.thcmref_psim.argtypes = [_P(C.c_double), _P(C.c_double)]
        lib.thcmref_fortran_rhs.argtypes = [_P(C.c_double), _P(C.c_double)]
        lib.thcmref_fortran_coB.argtypes = [_P(C.c_double)]
        lib.thcmref_psim.restype = C.c_double
        _ref_lib = lib
    return _ref_lib


class Case:
    """One evaluation point: configuration, land mask, forcing fields,
    continuation parameters, state x and the integral correction."""

    def __init__(self, name, cfg, landm, fields, pars, x, intcorr=0.0):
        self.name, self.cfg, self.landm = name, cfg, np.ascontiguousarray(landm, np.int32)
        self.fields = {k: (None if v is None else np.ascontiguousarray(v, np.float64))
                       for k, v in fields.items()}
        self.pars = dict(pars)
        self.x = np.ascontiguousarray(x, np.float64)
        self.intcorr = float(intcorr)

    @property
    def ndim(self):
        return 6 * self.cfg.n * self.cfg.m * self.cfg.l


def _fld(case, k):
    v = case.fields.get(k)
    return None if v is None else _ptr(v, C.c_double)


class OracleModel:
    def __init__(self, case):
        self.lib = load_oracle()
        self.case = case
        self.h = self.lib.tho_create(C.byref(case.cfg), _ptr(case.landm, C.c_int),
                                     *[_fld(case, k) for k in ("taux", "tauy", "tatm", "emip", "spert")])
        assert self.h, "oracle rejected the configuration"
        for k, v in case.pars.items():
            assert self.lib.tho_set_par(self.h, PAR[k], v) == 0
        self.ndim = self.lib.tho_ndim(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.tho_destroy(self.h)
            self.h = None

    def par(self, name):
        return self.lib.tho_get_par(self.h, PAR[name])

    def set_par(self, name, value):
        assert self.lib.tho_set_par(self.h, PAR[name], float(value)) == 0

    def rowintcon(self):
        return self.lib.tho_rowintcon(self.h)

    def graph(self):
        nnz = self.lib.tho_graph_nnz(self.h)
        rp = np.zeros(self.ndim + 1, np.int32)
        ci = np.zeros(nnz, np.int32)
        self.lib.tho_graph(self.h, _ptr(rp, C.c_int), _ptr(ci, C.c_int))
        return rp, ci

    def coef(self):
        c = np.zeros(self.ndim)
        self.lib.tho_intcond_coef(self.h, _ptr(c, C.c_double))
        return c

    def landm(self):
        c = self.case.cfg
        out = np.zeros((c.n + 2) * (c.m + 2) * (c.l + 2), np.int32)
        self.lib.tho_landm(self.h, _ptr(out, C.c_int))
        return out

    def forcing(self):
        f = np.zeros(self.ndim)
        self.lib.tho_forcing(self.h, _ptr(f, C.c_double))
        return f

    def rhs(self, x=None, intcorr=None):
        x = self.case.x if x is None else np.ascontiguousarray(x, np.float64)
        ic = self.case.intcorr if intcorr is None else intcorr
        F = np.zeros(self.ndim)
        rc = self.lib.tho_rhs(self.h, _ptr(x, C.c_double), ic, _ptr(F, C.c_double))
        assert rc == 0, rc
        return F

    def jacobian(self, x=None):
        x = self.case.x if x is None else np.ascontiguousarray(x, np.float64)
        rp, ci = self.graph()
        val = np.zeros(len(ci))
        dB = np.zeros(self.ndim)
        rc = self.lib.tho_jacobian(self.h, _ptr(x, C.c_double), _ptr(val, C.c_double),
                                   _ptr(dB, C.c_double))
        assert rc >= 0, rc
        return rp, ci, val, dB, rc

    def jacobian_rc(self, x=None):
        """jacobian() without the success check: (rowptr, col, val, diagB, rc)."""
        x = self.case.x if x is None else np.ascontiguousarray(x, np.float64)
        rp, ci = self.graph()
        val = np.zeros(len(ci))
        dB = np.zeros(self.ndim)
        rc = self.lib.tho_jacobian(self.h, _ptr(x, C.c_double), _ptr(val, C.c_double), _ptr(dB, C.c_double))
        return rp, ci, val, dB, rc

    def mix(self, x=None):
        x = self.case.x if x is None else np.ascontiguousarray(x, np.float64)
        out = np.zeros(self.ndim)
        assert self.lib.tho_mix(self.h, _ptr(x, C.c_double), _ptr(out, C.c_double)) == 0
        return out


def on_big_stack(fn, *args, stack_bytes=1 << 32):
    """Run fn on a thread with a large stack: the reference Fortran keeps its
    per-call work arrays (tens of 27-point atom arrays) on the stack."""
    import threading
    out = {}

    def body():
        try:
            out["r"] = fn(*args)
        except BaseException as e:  # re-raised in the caller
            out["e"] = e
    old = threading.stack_size()
    threading.stack_size(min(stack_bytes, (1 << 32) - 4096))
    try:
        th = threading.Thread(target=body)
        th.start()
        th.join()
    finally:
        threading.stack_size(old)
    if "e" in out:
        raise out["e"]
    return out["r"]


class RefModel:
    """The reference Fortran (one process-wide instance, as THCM is a singleton)."""

    def __init__(self, case):
        self.lib = load_ref()
        self.case = case
        on_big_stack(self.lib.thcmref_init, C.byref(case.cfg), _ptr(case.landm, C.c_int),
                     *[_fld(case, k) for k in ("taux", "tauy", "tatm", "emip", "spert")])
        for k, v in case.pars.items():
            on_big_stack(self.lib.thcmref_set_par, PAR[k], v)
        self.ndim = self.lib.thcmref_ndim()

    def par(self, name):
        return self.lib.thcmref_get_par(PAR[name])

    def rowintcon(self):
        return self.lib.thcmref_rowintcon()

    def coef(self):
        c = np.zeros(self.ndim)
        self.lib.thcmref_intcond_coef(_ptr(c, C.c_double))
        return c

    def rhs(self, x=None, intcorr=None):
        x = self.case.x if x is None else np.ascontiguousarray(x, np.float64)
        ic = self.case.intcorr if intcorr is None else intcorr
        F = np.zeros(self.ndim)
        on_big_stack(self.lib.thcmref_rhs, _ptr(x, C.c_double), ic, _ptr(F, C.c_double))
        return F

    def jacobian(self, x=None):
        x = self.case.x if x is None else np.ascontiguousarray(x, np.float64)
        nnz = self.lib.thcmref_graph_nnz()
        rp = np.zeros(self.ndim + 1, np.int32)
        ci = np.zeros(nnz, np.int32)
        val = np.zeros(nnz)
        dB = np.zeros(self.ndim)
        dropped = on_big_stack(self.lib.thcmref_jacobian, _ptr(x, C.c_double), _ptr(rp, C.c_int),
                               _ptr(ci, C.c_int), _ptr(val, C.c_double), _ptr(dB, C.c_double))
        return rp, ci, val, dB, dropped

    def psim(self, x=None):
        """(max psiM, psiM(0:m, 0:l) j fastest) of OceanGrid::recomputePsiM."""
        x = self.case.x if x is None else np.ascontiguousarray(x, np.float64)
        c = self.case.cfg
        out = np.zeros((c.m + 1) * (c.l + 1))
        mx = on_big_stack(self.lib.thcmref_psim, _ptr(x, C.c_double), _ptr(out, C.c_double))
        return mx, out

    def fortran_rhs(self, x=None):
        """rhs_ itself: the Fortran right-hand side B (usrc.F90:523-603)."""
        x = self.case.x if x is None else np.ascontiguousarray(x, np.float64)
        B = np.zeros(self.ndim)
        on_big_stack(self.lib.thcmref_fortran_rhs, _ptr(x, C.c_double), _ptr(B, C.c_double))
        return B

    def fortran_coB(self):
        out = np.zeros(self.ndim)
        self.lib.thcmref_fortran_coB(_ptr(out, C.c_double))
        return out

    def fortran_csr(self):
        nz = self.lib.thcmref_fortran_csr_nnz()
        beg = np.zeros(self.ndim + 1, np.int32)
        jco = np.zeros(nz, np.int32)
        co = np.zeros(nz)
        self.lib.thcmref_fortran_csr(_ptr(beg, C.c_int), _ptr(jco, C.c_int), _ptr(co, C.c_double))
        return beg, jco, co

    @staticmethod
    def reference_landm(cfg, itopo, rd_mask=0, maskfile=None):
        lib = load_ref()
        out = np.zeros((cfg.n + 2) * (cfg.m + 2) * (cfg.l + 2), np.int32)
        lib.thcmref_reference_landm(C.byref(cfg), itopo, rd_mask,
                                    None if maskfile is None else maskfile.encode(),
                                    _ptr(out, C.c_int))
        return out


# --------------------------------------------------------------------------
# Land masks, forcing and states

def read_mask_file(path, n, m, l):
    """The reference's land-mask text format (topo.F90:50-62): for each level
    k = 0..l+1 a header line, then rows j = m+1 .. 0 of n+2 digits
    (Fortran's blank padding of short records read as 0)."""
    landm = np.full((l + 2, m + 2, n + 2), LAND, np.int32)
    with open(path) as f:
        lines = f.read().splitlines()
    p = 0
    for k in range(l + 2):
        p += 1
        for j in range(m + 1, -1, -1):
            row = lines[p]
            p += 1
            for i in range(n + 2):
                # a short record pads with blanks, read as 0 (format i1)
                landm[k, j, i] = int(row[i]) if i < len(row) and row[i] != " " else 0
    # land inversion fix of topo.F90:97-105
    for i in range(1, n + 1):
        for j in range(1, m + 1):
            for k in range(l, 1, -1):
                if landm[k, j, i] == LAND and landm[k - 1, j, i] == OCEAN:
                    landm[k - 1, j, i] = LAND
    return landm.reshape(-1)


def box_mask(n, m, l, periodic=False):
    """itopo = 1 (topo.F90:177-187): ocean everywhere, land border."""
    landm = np.full((l + 2, m + 2, n + 2), LAND, np.int32)
    landm[1:l + 1, 1:m + 1, 1:n + 1] = OCEAN
    if periodic:
        both = (landm[:, :, 1] == OCEAN) & (landm[:, :, n] == OCEAN)
        landm[:, :, 0][both] = PERIO
        landm[:, :, n + 1][both] = PERIO
    return landm.reshape(-1)


def lat_deg(cfg):
    """cell-centre and v-point latitudes in degrees (grid.F90:35-43)."""
    dy = (cfg.ymax_deg - cfg.ymin_deg) / cfg.m
    y = cfg.ymin_deg + (np.arange(1, cfg.m + 1) - 0.5) * dy
    yv = cfg.ymin_deg + np.arange(1, cfg.m + 1) * dy
    return y, yv


def analytic_forcing(cfg):
    """The seeded-basin forcing: tau_x = -cos(2 pi (lat-10)/64) at v-point
    latitudes, T* = 15 cos(pi (lat-42)/64) (tatm = T* - t0, t0 = 15) and
    salt flux 0.1 cos(pi (lat-10)/64) at cell-centre latitudes."""
    y, yv = lat_deg(cfg)
    n = cfg.n
    taux = np.repeat((-np.cos(2 * np.pi * (yv - 10.0) / 64.0))[:, None], n, 1)
    tstar = 15.0 * np.cos(np.pi * (y - 42.0) / 64.0)
    tatm = np.repeat((tstar - 15.0)[:, None], n, 1)
    emip = np.repeat((0.1 * np.cos(np.pi * (y - 10.0) / 64.0))[:, None], n, 1)
    return dict(taux=taux.reshape(-1), tauy=np.zeros(n * cfg.m), tatm=tatm.reshape(-1),
                emip=emip.reshape(-1), spert=np.zeros(n * cfg.m))


def seeded_state(cfg, landm, seed, noise=1e-3):
    """Rest velocities, T linear in depth from T*-15 (k = l) to 4-15 (k = 1),
    S = 0 (anomalies), N(0, noise) on every unknown, surface w zeroed
    (Ocean::inspectVector, Ocean.C:427-487)."""
    n, m, l = cfg.n, cfg.m, cfg.l
    y, _ = lat_deg(cfg)
    tstar = 15.0 * np.cos(np.pi * (y - 42.0) / 64.0) - 15.0
    x = np.zeros((l, m, n, 6))
    for k in range(l):
        f = k / max(l - 1, 1)
        x[k, :, :, TT - 1] = ((4.0 - 15.0) * (1 - f) + tstar * f)[:, None]
    rng = np.random.default_rng(seed)
    x += rng.normal(0.0, noise, x.shape)
    x[l - 1, :, :, WW - 1] = 0.0
    return x.reshape(-1)


def box16_case(seed=11):
    """The 16x16x12 flat-bottom box with analytic forcing at COMB = 1."""
    cfg = make_config(n=16, m=16, l=12, sres=0, iza=3, ite=2, its=2)
    landm = box_mask(16, 16, 12)
    x = seeded_state(cfg, landm, seed)
    pars = dict(COMB=1.0, WIND=1.0, TEMP=1.0, SALT=1.0)
    return Case("box16", cfg, landm, analytic_forcing(cfg), pars, x)


def rand_state(cfg, seed, amp=1.0):
    """A seeded state far from rest, to exercise every nonlinear term:
    velocities, pressure and salinity noise, T stratified from -11 (bottom) to
    0 (surface) plus N(0, amp) noise, so that parts of the column are
    statically unstable and the convective mixing is active."""
    rng = np.random.default_rng(seed)
    n, m, l = cfg.n, cfg.m, cfg.l
    x = np.zeros((l, m, n, 6))
    x[..., 0] = rng.normal(0, 0.1 * amp, (l, m, n))
    x[..., 1] = rng.normal(0, 0.1 * amp, (l, m, n))
    x[..., 2] = rng.normal(0, 0.01 * amp, (l, m, n))
    x[..., 3] = rng.normal(0, 0.1 * amp, (l, m, n))
    x[..., 4] = np.linspace(-11, 0, l)[:, None, None] + rng.normal(0, 1.0 * amp, (l, m, n))
    x[..., 5] = rng.normal(0, 0.3 * amp, (l, m, n))
    return x.reshape(-1)


def checkerboard_modes(cfg, landm):
    """The two normalised checkerboard pressure modes s1 (i + j even) and s2
    (odd) over the wet cells (Ocean.C:2193-2223, THCM.C:2802-2841); landm is
    the bordered (l+2)(m+2)(n+2) mask."""
    n, m, l = cfg.n, cfg.m, cfg.l
    wet = np.asarray(landm).reshape(l + 2, m + 2, n + 2)[1:-1, 1:-1, 1:-1] == 0
    k, j, i = np.nonzero(wet)
    rows = 6 * ((k * m + j) * n + i) + 3
    s = np.zeros((2, 6 * n * m * l))
    s[(i + j) % 2, rows] = 1.0
    return s / np.sqrt(s.sum(axis=1))[:, None]


def pressure_projection(cfg, landm, v):
    """Ocean::pressureProjection (Ocean.C:2180-2235): v - (v.s1) s1 - (v.s2) s2."""
    s = checkerboard_modes(cfg, landm)
    return v - s.T @ (s @ v)


def golden_names():
    return sorted(f[:-4] for f in os.listdir(GOLDEN) if f.endswith(".npz"))


def load_golden(name):
    """(Case, expected dict) of tests/golden/<name>.npz."""
    import json
    z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    cfg = Config(**json.loads(str(z["cfg"])))
    fields = {k[4:]: z[k] for k in z.files if k.startswith("fld_")}
    case = Case(name, cfg, z["landm"], fields, json.loads(str(z["pars"])), z["x"],
                float(z["intcorr"]))
    exp = {k: z[k] for k in z.files}
    return case, exp


def expected_rows(exp, ndim):
    """Row indices covered by a fixture's Jacobian (all rows if not sampled)."""
    return exp["J_rows"] if "J_rows" in exp else np.arange(ndim, dtype=np.int32)


def restrict_rows(rp, ci, val, rows):
    """CSR of the given rows only (row pointer, columns, values)."""
    lens = rp[rows + 1] - rp[rows]
    sub = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    idx = np.concatenate([np.arange(rp[r], rp[r + 1]) for r in rows]) if len(rows) else np.zeros(0, int)
    return sub, ci[idx], val[idx]
