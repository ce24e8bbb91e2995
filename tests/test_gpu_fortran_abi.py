"""The Fortran interface THCM.C binds (THCM.C:49-176): rhs_ and matrix_ with
caller-allocated 1-based CSR buffers (m_mat::get_array_sizes /
set_pointers, mat.F90:58-104), through the C ABI, against what the reference
Fortran itself writes into those buffers (tests/golden/fortran, written by
tests/golden/make_golden.py): B, begA, jcoA, coA and coB bit for bit."""
import os

import numpy as np
import pytest

from gpu_helpers import gpu_ocean
from thcm_cases import GOLDEN, load_golden

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["natl8", "gateway16"])
def test_rhs_and_matrix_in_the_fortran_layout(name):
    case, _ = load_golden(name)
    z = np.load(os.path.join(GOLDEN, "fortran", name + ".npz"), allow_pickle=False)
    oc = gpu_ocean(case)
    B = oc.fortranRHS(case.x)
    assert np.array_equal(B, z["B"])
    beg, jco, co, coB = oc.fortranMatrix(case.x)
    print(f"{name}: nnz {len(co)} (reference {len(z['coA'])})")
    assert np.array_equal(beg, z["begA"])
    assert np.array_equal(jco, z["jcoA"])
    assert np.array_equal(co, z["coA"])
    assert np.array_equal(coB, z["coB"])
    oc.close()
