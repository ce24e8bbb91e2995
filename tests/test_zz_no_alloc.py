"""iemic_rhs_dev_noalloc, F(x) with no device allocation, against
iemic_rhs_dev through the C ABI: two calls in a row on different models, the
second with a smaller input, each equal to iemic_rhs_dev's F in every row
and, in the integral-condition row (a reduction summed in another order),
to 1e-12 of the sum of its terms' magnitudes (the bar the oracle parity
tests hold that row to); sizes above IEMIC_NOALLOC_MAX_NDIM are refused."""
import ctypes as C

import numpy as np
import pytest

import iemic_amd
from gpu_helpers import gpu_ocean
from thcm_cases import load_golden, seeded_state

pytestmark = pytest.mark.gpu

MAX_NDIM = 4194304


def _check_model(oc, x, intcorr):
    lib = iemic_amd.load()
    dx = iemic_amd.DeviceArray(oc.ndim).upload(x)
    dF_ref = iemic_amd.DeviceArray(oc.ndim).upload(np.full(oc.ndim, np.nan))
    dF = iemic_amd.DeviceArray(oc.ndim).upload(np.full(oc.ndim, np.nan))
    assert lib.iemic_rhs_dev(oc.h, dx.ptr, dF_ref.ptr) == 0
    rc = lib.iemic_rhs_dev_noalloc(oc.h, dx.ptr, dF.ptr)
    assert rc == 0, lib.iemic_last_error()
    oc.synchronize()
    F_ref, F = dF_ref.download(), dF.download()
    ric = oc.rowintcon()
    rows = np.arange(oc.ndim) != ric
    assert np.array_equal(F[rows], F_ref[rows])
    if ric >= 0:
        scale = np.sum(np.abs(oc.intcond_coef() * x)) + abs(intcorr)
        print(f"ndim {oc.ndim}: integral row {F[ric]:.17e} vs {F_ref[ric]:.17e} (scale {scale:.3e})")
        assert abs(F[ric] - F_ref[ric]) <= 1e-12 * scale
    return F


def test_noalloc_rhs_matches_rhs_dev_twice_second_smaller():
    big, _ = load_golden("global96")
    big.x = seeded_state(big.cfg, big.landm, 11)
    small, _ = load_golden("natl8")
    sizes = []
    for case in (big, small):
        oc = gpu_ocean(case)
        assert oc.rowintcon() >= 0 or case is small
        _check_model(oc, case.x, case.intcorr)
        sizes.append(oc.ndim)
        oc.close()
    assert sizes[1] < sizes[0]


def test_noalloc_rhs_refuses_more_than_its_maximum():
    from iemic_amd import basins
    cfg_d, lm, frc, x0, pars = basins.make_box(220, 100, 32)
    cfg = iemic_amd.Config(**{f: cfg_d[f] for f, _ in iemic_amd.Config._fields_})
    oc = iemic_amd.Ocean(cfg, lm, frc["taux"], frc["tauy"], frc["tatm"], frc["emip"], frc["spert"])
    assert oc.ndim > MAX_NDIM
    lib = iemic_amd.load()
    dx = iemic_amd.DeviceArray(oc.ndim)
    dF = iemic_amd.DeviceArray(oc.ndim)
    rc = lib.iemic_rhs_dev_noalloc(oc.h, dx.ptr, dF.ptr)
    msg = lib.iemic_last_error().decode()
    print(f"ndim {oc.ndim}: rc {rc}, {msg}")
    assert rc == -1 and "exceeds the maximum" in msg
    oc.close()
