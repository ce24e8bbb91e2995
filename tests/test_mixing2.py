"""Mixing = 2, the adaptive mixing partition of THCM (mix_imp.f:131-227,
934-1049; usrc.F90:489-563): before every F and J evaluation
(Ocean::computeRHS / computeJacobian call THCM::fixMixing(0), Ocean.C:1281,
1302) the T and S mixing are switched on or off by whether the T and S
fields are nonzero, and the finite-difference mixing Jacobian's pattern is
rebuilt when the switches change while T is on.  The oracle follows the
sequence of the reference's own F and J (tests/golden/mix2, written by the
reference Fortran) bit for bit, and the switching is pinned by what it must
reduce to: no mixing at all leaves F equal to Mixing = 0's."""
import json
import os

import numpy as np
import pytest

from thcm_cases import GOLDEN, Case, Config, OracleModel, config_dict

SEQ = os.path.join(GOLDEN, "mix2", "natl8_sequence.npz")


def _seq():
    z = np.load(SEQ, allow_pickle=False)
    cfg = Config(**json.loads(str(z["cfg"])))
    case = Case("natl8_mix2", cfg, z["landm"], {}, json.loads(str(z["pars"])), z["x0"], float(z["intcorr"]))
    return case, z


def test_oracle_follows_the_reference_sequence():
    case, z = _seq()
    orc = OracleModel(case)
    labels = [str(v) for v in z["labels"]]
    for q, lab in enumerate(labels):
        x = z[f"x{q}"]
        np.testing.assert_array_equal(orc.rhs(x), z[f"F{q}"], err_msg=lab)
        if lab == "t_only":
            # pattern built with S = 0: the reference's entries then depend on
            # its DSM column grouping (the unpatterned S columns are perturbed
            # with the T groups); the oracle, like the GPU library, refuses
            rp, ci, val, dB, rc = orc.jacobian_rc(x)
            assert rc == -2
            continue
        rp, ci, val, dB, rc = orc.jacobian(x)
        np.testing.assert_array_equal(rp, z["J_rowptr"])
        np.testing.assert_array_equal(val, z[f"J{q}"], err_msg=lab)


def test_switched_off_mixing_is_mixing_zero():
    """With neither field nonzero, and before any pattern exists, F and J
    equal those of Mixing = 0 at the same state; with T off (S only, pattern
    kept) the T rows of F equal Mixing = 0's."""
    case, z = _seq()
    labels = [str(v) for v in z["labels"]]
    cfg0 = Config(**config_dict(case.cfg))
    cfg0.vmix = 0
    case0 = Case("natl8_mix0", cfg0, case.landm, {}, case.pars, case.x, case.intcorr)
    o0 = OracleModel(case0)
    xn = z[f"x{labels.index('neither')}"]
    np.testing.assert_array_equal(z[f"F{labels.index('neither')}"], o0.rhs(xn))
    np.testing.assert_array_equal(z[f"J{labels.index('neither')}"], o0.jacobian(xn)[2])
    xs = z[f"x{labels.index('s_only')}"]
    Ts = np.arange(case.ndim) % 6 == 4
    np.testing.assert_array_equal(z[f"F{labels.index('s_only')}"][Ts], o0.rhs(xs)[Ts])
    # a fresh Mixing = 2 model whose first state has T = 0 never builds a
    # pattern: no mixing in F at all
    fresh = OracleModel(case)
    np.testing.assert_array_equal(fresh.rhs(xs), o0.rhs(xs))
