"""State files in the reference's EpetraExt-HDF5 layout (iemic_read_state /
iemic_write_state, after Model::loadStateFromFile / saveStateToFile,
Model.C:15-160), against the reference's own stored state
test/ocean/ocean_reference.h5 (copied into tests/golden/reft as data) and
its independent conversion by h5py (tests/golden/make_reft_state.py).
Host code only: no GPU."""
import os

import numpy as np
import pytest

import iemic_amd
from thcm_cases import GOLDEN

H5 = os.path.join(GOLDEN, "reft", "ocean_reference.h5")
NDIM = 16 * 16 * 16 * 6


def _converted():
    return np.load(os.path.join(GOLDEN, "reft", "ocean_reference_state.npz"), allow_pickle=False)


def test_reads_the_reference_state_file():
    """The state equals h5py's reading bit for bit; the parameters are the
    file's under THCM::int2par's labels.  This file predates two of those
    labels ("Flux Perturbation", stored as "FPER", and "SPL2"), which the
    reference's loader skips, so the reader reports them absent."""
    x, pars = iemic_amd.read_state(H5, NDIM)
    z = _converted()
    np.testing.assert_array_equal(x, z["x"])
    assert sorted(pars) == [i for i in range(1, 31) if i not in (27, 30)]
    for i, v in pars.items():
        assert v == z["par"][i]
    assert pars[19] == 0.020000184160660548   # Combined Forcing


def test_write_then_read_is_exact(tmp_path):
    z = _converted()
    rng = np.random.default_rng(5)
    x = rng.normal(size=NDIM)
    pars = {i: float(z["par"][i]) for i in range(1, 31)}
    p = tmp_path / "state.h5"
    iemic_amd.write_state(p, x, pars)
    x2, p2 = iemic_amd.read_state(p, NDIM)
    np.testing.assert_array_equal(x2, x)
    assert p2 == pars
    # without parameters: the state alone; asking for parameters is then an
    # error, as Model::loadStateFromFile stops without a <Parameters> group
    iemic_amd.write_state(p, x)
    x3, p3 = iemic_amd.read_state(p, NDIM, parameters=False)
    np.testing.assert_array_equal(x3, x)
    assert p3 == {}
    with pytest.raises(iemic_amd.IemicError) as e:
        iemic_amd.read_state(p, NDIM)
    assert e.value.code == -7 and "Parameters" in str(e.value)


def test_concurrent_first_loads(tmp_path):
    """Reads from several threads at once (in a fresh process the first of
    them race to load the HDF5 library, which is loaded once)."""
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(8) as ex:
        outs = list(ex.map(lambda _: iemic_amd.read_state(H5, NDIM)[0], range(16)))
    for o in outs[1:]:
        np.testing.assert_array_equal(o, outs[0])


def test_errors(tmp_path):
    with pytest.raises(iemic_amd.IemicError) as e:
        iemic_amd.read_state(H5, NDIM - 6)   # another grid's length
    assert e.value.code == -1
    with pytest.raises(iemic_amd.IemicError) as e:
        iemic_amd.read_state(tmp_path / "missing.h5", NDIM)
    assert e.value.code == -7
    bad = tmp_path / "not_hdf5.h5"
    bad.write_text("not an HDF5 file")
    with pytest.raises(iemic_amd.IemicError) as e:
        iemic_amd.read_state(bad, NDIM)
    assert e.value.code == -7
    with pytest.raises(ValueError):
        iemic_amd.write_state(tmp_path / "x.h5", np.zeros(6), {19: 1.0})
