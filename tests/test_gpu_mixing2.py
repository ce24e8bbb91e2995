"""Mixing = 2 on the GPU: the reference's own F and J sequence on the natl8
basin (tests/golden/mix2, written by the reference Fortran), one model
evaluated as Ocean::computeRHS then Ocean::computeJacobian do at each state
while the T and S mixing switch on and off.  F and J equal the reference bit
for bit (the integral row of F, a device reduction, to 1e-12 of the sum of
its terms' magnitudes); at the state whose pattern was built with a zero S
field the Jacobian is refused, as the reference's entries there depend on
its DSM column grouping."""
import json
import os

import numpy as np
import pytest

import iemic_amd
from gpu_helpers import gpu_ocean
from thcm_cases import GOLDEN, Case, Config

pytestmark = pytest.mark.gpu


def test_gpu_follows_the_reference_mixing2_sequence():
    z = np.load(os.path.join(GOLDEN, "mix2", "natl8_sequence.npz"), allow_pickle=False)
    cfg = Config(**json.loads(str(z["cfg"])))
    case = Case("natl8_mix2", cfg, z["landm"], {}, json.loads(str(z["pars"])), z["x0"], float(z["intcorr"]))
    oc = gpu_ocean(case)
    ric = oc.rowintcon()
    coef = oc.intcond_coef()
    rows = np.arange(oc.ndim) != ric
    for q, lab in enumerate(str(v) for v in z["labels"]):
        x = z[f"x{q}"]
        F = oc.computeRHS(x)
        print(f"{lab}: F", flush=True)
        assert np.array_equal(F[rows], z[f"F{q}"][rows]), lab
        assert abs(F[ric] - z[f"F{q}"][ric]) <= 1e-12 * (np.sum(np.abs(coef * x)) + abs(case.intcorr))
        if lab == "t_only":
            with pytest.raises(iemic_amd.IemicError) as e:
                oc.computeJacobian(x)
            assert e.value.code == -2
            print(f"{lab}: J refused: {e.value}")
            continue
        oc.computeJacobian(x)
        rp, ci, val, _ = oc.getJacobian()
        assert np.array_equal(rp, z["J_rowptr"]) and np.array_equal(ci, z["J_col"])
        d = np.abs(val - z[f"J{q}"]).max()
        print(f"{lab}: J max difference {d:.3e}")
        assert np.array_equal(val, z[f"J{q}"]), lab
    oc.close()
