"""Converts the reference's stored steady state test/ocean/ocean_reference.h5
(the EpetraExt-HDF5 file that the reference test reft_ocean compares its
continuation result with, src/tests/reft_ocean.C:59-89) into the data
fixture tests/golden/reft/ocean_reference_state.npz.

Needs h5py, which only the conda interpreter of this container has:

    /opt/conda/bin/python3.9 tests/golden/make_reft_state.py

The file is read as data (datasets and scalars only).  Stored: the state
vector (State/Values, reference ordering 6((kM + j)N + i) + var), the 30
THCM parameters by index (the file stores them under the labels of
THCM::par2int, THCM.C:1848-1887; SPL2 is not stored and takes the value of
test/ocean/reft_ocean_params.xml), the global land mask (MaskGlobal/Global,
(n+2)(m+2)(l+2) with borders) and the 2-norms of the u, v, T, S parts of the
state that reft_ocean checks.
"""
import os
import sys

import h5py
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/test/ocean/ocean_reference.h5"

# THCM.C:1848-1887 labels -> par.F90 indices
LABELS = {"AL_T": 1, "Rayleigh-Number": 2, "Vertical Ekman-Number": 3, "Horizontal Ekman-Number": 4,
          "Rossby-Number": 5, "MIXP": 6, "RESC": 7, "SPL1": 8, "Salinity Homotopy": 9,
          "Solar Forcing": 10, "Horizontal Peclet-Number": 11, "Vertical Peclet-Number": 12,
          "P_VC": 13, "LAMB": 14, "Salinity Forcing": 15, "Wind Forcing": 16,
          "Temperature Forcing": 17, "Nonlinear Factor": 18, "Combined Forcing": 19, "ARCL": 20,
          "NLES": 21, "IFRICB": 22, "CONT": 23, "Energy": 24, "ALPC": 25, "CMPR": 26, "FPER": 27,
          "Salinity Perturbation": 28, "MKAP": 29, "SPL2": 30}
SPL2_XML = 0.01   # test/ocean/reft_ocean_params.xml, Starting Parameters


def main():
    f = h5py.File(SRC, "r")
    x = np.asarray(f["State/Values"][()], dtype=np.float64).reshape(-1)
    assert int(f["State/GlobalLength"][()]) == x.size == 16 * 16 * 16 * 6
    par = np.full(31, np.nan)
    for k in f["Parameters"]:
        if k in LABELS:   # "Time" is the model time, not a THCM parameter
            par[LABELS[k]] = float(f["Parameters"][k][()])
    if np.isnan(par[30]):
        par[30] = SPL2_XML
    assert not np.isnan(par[1:]).any(), "a parameter is missing"
    mask = np.asarray(f["MaskGlobal/Global"][()], dtype=np.int32)
    norms = np.array([np.linalg.norm(x[v::6]) for v in (0, 1, 4, 5)])
    out = os.path.join(HERE, "reft", "ocean_reference_state.npz")
    np.savez_compressed(out, x=x, par=par, mask=mask, norms_uvTS=norms)
    print(out, "COMB", par[19], "norms u v T S", norms)


if __name__ == "__main__":
    sys.exit(main())
