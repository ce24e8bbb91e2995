"""The land-mask text reader of the library (iemic_read_mask, after readmask
topo.F90:41-134) against the masks the reference's own readmask produced
for the golden fixtures, from the same data/mkmask files (copied into
tests/golden/mkmask as data).  Host code only: no GPU."""
import os

import numpy as np
import pytest

import iemic_amd
from thcm_cases import GOLDEN, LAND, OCEAN, load_golden, read_mask_file

MASKS = [("natl8", "mask_natl8"), ("gateway16", "mask_gateway"), ("global96", "mask_global_96x38x12")]


@pytest.mark.parametrize("name,fname", MASKS)
def test_read_mask_equals_the_reference_readmask(name, fname):
    case, _ = load_golden(name)
    c = case.cfg
    got = iemic_amd.read_mask(os.path.join(GOLDEN, "mkmask", fname), c.n, c.m, c.l)
    np.testing.assert_array_equal(got, np.asarray(case.landm).reshape(-1))


def _write(tmp_path, text):
    p = tmp_path / "mask"
    p.write_text(text)
    return p


def test_short_rows_read_as_ocean_and_inversions_are_filled(tmp_path):
    """n = m = 1, l = 2; rows are written j = 2, 1, 0.  Level 2, row 1 is
    the short record '11' (its third value pads as blank = 0 = ocean);
    level 1, row 1 holds ocean at i = 1, under land at level 2, which the
    land-inversion fix turns into land."""
    levels = []
    for k in range(4):
        rows = ["111", "111", "111"]
        if k == 1:
            rows[1] = "101"
        if k == 2:
            rows[1] = "11"
        levels.append("level\n" + "\n".join(rows))
    p = _write(tmp_path, "\n".join(levels) + "\n")
    lm = iemic_amd.read_mask(p, 1, 1, 2).reshape(4, 3, 3)
    assert lm[2, 1, 2] == OCEAN
    assert lm[2, 1, 1] == LAND and lm[1, 1, 1] == LAND
    np.testing.assert_array_equal(lm, read_mask_file(str(p), 1, 1, 2).reshape(4, 3, 3))


def test_truncated_or_malformed_files_are_errors(tmp_path):
    with pytest.raises(iemic_amd.IemicError):
        iemic_amd.read_mask(_write(tmp_path, "level\n111\n111\n"), 1, 1, 2)
    bad = "\n".join(["level\n111\n1x1\n111"] * 4) + "\n"
    with pytest.raises(iemic_amd.IemicError):
        iemic_amd.read_mask(_write(tmp_path, bad), 1, 1, 2)
    with pytest.raises(iemic_amd.IemicError):
        iemic_amd.read_mask(tmp_path / "missing", 1, 1, 2)
