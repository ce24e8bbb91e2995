"""Layer 4 of the boundary: the GMRES template contract of I-EMIC
(GMRESSolverDecl.H:4-16).  A GMRES written only against that contract
(tests/gmres_contract/contract_gmres.cpp: Model::applyMatrix/applyPrecon,
Vector copy/assignment/update/norm/dot/scale, as GMRESSolver.H uses them)
is instantiated with include/iemic_gmres.hpp's iemic::GpuModel and
std::shared_ptr<iemic::DeviceVector>, the types the reference's
GMRESSolver<Model, VectorPointer> would take, and solves J x = b on the GPU
model to the reference's FP64 bar (Ocean.C:1129-1146: the true residual,
recomputed on the host from the exported Jacobian), in as many iterations
as the library's own FGMRES."""
import ctypes as C
import os

import numpy as np
import pytest
import scipy.sparse as sp

import iemic_amd
from gpu_helpers import gpu_ocean
from thcm_cases import load_golden, seeded_state

HARNESS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "gmres_contract", "libcontract_gmres.so")


def _harness():
    iemic_amd.load()   # the product library first: the harness resolves to the same one
    lib = C.CDLL(HARNESS)
    dp = C.POINTER(C.c_double)
    lib.contract_gmres_solve.argtypes = [C.c_void_p, dp, dp, C.c_double, C.c_int, C.c_int, C.POINTER(C.c_int),
                                         C.POINTER(C.c_double)]
    lib.contract_gmres_solve.restype = C.c_int
    return lib


def test_contract_harness_is_built_and_exports_its_solver():
    assert os.path.exists(HARNESS), "run __graft_entry__.build()"
    assert hasattr(_harness(), "contract_gmres_solve")


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["natl8", "gateway16", "global96"])
def test_contract_gmres_on_the_gpu_model(name):
    case, _ = load_golden(name)
    if name in ("gateway16", "global96"):
        case.x = seeded_state(case.cfg, case.landm, 11)
    oc = gpu_ocean(case)
    oc.computeJacobian(case.x)
    rp, ci, val, _ = oc.getJacobian()
    J = sp.csr_matrix((val, ci, rp), shape=(oc.ndim, oc.ndim))
    b = np.ascontiguousarray(J @ np.random.default_rng(9).normal(size=oc.ndim))
    x = np.zeros(oc.ndim)
    its, res = C.c_int(), C.c_double()
    dp = C.POINTER(C.c_double)
    rc = _harness().contract_gmres_solve(oc.h, b.ctypes.data_as(dp), x.ctypes.data_as(dp), 1e-10, 50, 200,
                                         C.byref(its), C.byref(res))
    rel = np.linalg.norm(b - J @ x) / np.linalg.norm(b)
    _, info = oc.solve(b, oc.solver_params(tol=1e-10, restart=50, maxit=200))
    print(f"{name}: contract GMRES {its.value} iterations, true residual {rel:.2e} (its own {res.value:.2e}); "
          f"library FGMRES {info.iterations}")
    assert rc == 0
    assert rel <= 1e-10
    assert its.value <= info.iterations + 1
    oc.close()
