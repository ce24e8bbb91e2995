"""Writes the golden fixtures of tests/golden/*.npz from the reference build.

Run from the repository root in the container that holds /root/reference,
after `make -C oracle` has built oracle/_ref/libthcm_ref.so:

    python tests/golden/make_golden.py

Each fixture holds the inputs of one evaluation point (configuration, land
mask, surface forcing fields, continuation parameters, state, integral
correction) and what the reference THCM Fortran -- driven by the one-rank
restatement of the THCM.C glue in oracle/ref_harness.c -- returns for them:
F(x) as Ocean::computeRHS gives it, the Jacobian on the fixed graph (row
pointer, columns, values; all rows, or a seeded row sample for the larger
grids), diag(B), the integral-condition coefficients and row.  Land masks of
the natl8 / gateway / global cases are the reference's own data files
(data/mkmask), read by the reference's readmask.  Each fixture also holds the
meridional overturning streamfunction psiM(0:m, 0:l) of its state and its
maximum, as OceanGrid::recomputePsiM forms them with the reference's usol1D
and compute_psim (thcm_utils.F90:95-118, 190-254).
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from thcm_cases import (Case, RefModel, analytic_forcing, box16_case, config_dict,  # noqa: E402
                        make_config, rand_state)


def row_sample(ndim, frac, seed):
    rng = np.random.default_rng(seed)
    rows = np.sort(rng.choice(ndim, size=max(1, int(frac * ndim)), replace=False))
    return rows.astype(np.int32)


def write(case, row_frac=None):
    ref = RefModel(case)
    F = ref.rhs()
    rp, ci, val, dB, dropped = ref.jacobian()
    assert dropped == 0, "reference produced entries outside the fixed graph"
    out = dict(cfg=json.dumps(config_dict(case.cfg)), pars=json.dumps(case.pars),
               landm=case.landm, x=case.x, intcorr=np.float64(case.intcorr),
               F=F, diagB=dB, coef=ref.coef(), rowintcon=np.int32(ref.rowintcon()))
    # OceanGrid::recomputePsiM through the reference's usol1D and compute_psim
    out["psim_max"], out["psim"] = ref.psim()
    for k, v in case.fields.items():
        if v is not None:
            out["fld_" + k] = v
    if row_frac is None:
        out.update(J_rowptr=rp, J_col=ci, J_val=val)
    else:
        rows = row_sample(case.ndim, row_frac, 17)
        if ref.rowintcon() >= 0 and ref.rowintcon() not in rows:
            rows = np.sort(np.append(rows, np.int32(ref.rowintcon())))
        lens = rp[rows + 1] - rp[rows]
        sub_rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
        idx = np.concatenate([np.arange(rp[r], rp[r + 1]) for r in rows])
        out.update(J_rows=rows, J_rowptr=sub_rp, J_col=ci[idx], J_val=val[idx])
    path = os.path.join(HERE, case.name + ".npz")
    np.savez_compressed(path, **out)
    print(f"{path}: ndim={case.ndim} |F|max={np.abs(F).max():.3e} "
          f"|J|max={np.abs(val).max():.3e} size={os.path.getsize(path)} B")


def cases():
    yield box16_case(), None
    cfg = make_config(n=8, m=8, l=4, rho_mixing=0, sres=0)
    landm = RefModel.reference_landm(cfg, 0, 1, "mask_natl8")
    yield Case("natl8", cfg, landm, {}, dict(COMB=0.7, WIND=1.0, SALT=1.0, TEMP=10.0),
               rand_state(cfg, 3)), None
    cfg = make_config(n=8, m=8, l=4, rho_mixing=1, sres=1)
    yield Case("natl8_eos", cfg, landm, {}, dict(COMB=1.0, WIND=1.0, SALT=1.0, TEMP=10.0,
                                                  NLES=1.0), rand_state(cfg, 4, 2.0)), None
    cfg = make_config(n=16, m=16, l=16, periodic=1, xmin_deg=0, xmax_deg=360, ymin_deg=-80,
                      ymax_deg=80, hdim=5000, qz=2.25, sres=0)
    landm = RefModel.reference_landm(cfg, 0, 1, "mask_gateway")
    yield Case("gateway16", cfg, landm, {}, dict(COMB=1.0, WIND=1.0, SALT=1.0, TEMP=10.0),
               rand_state(cfg, 5)), None
    cfg = make_config(n=96, m=38, l=12, periodic=1, xmin_deg=0, xmax_deg=360, ymin_deg=-81,
                      ymax_deg=81, hdim=5000, qz=2.25, sres=0, ih=1, iza=3, ite=2, its=2)
    landm = RefModel.reference_landm(cfg, 0, 1, "mask_global_96x38x12")
    yield Case("global96", cfg, landm, analytic_forcing(cfg),
               dict(COMB=1.0, WIND=1.0, SALT=1.0, TEMP=1.0), rand_state(cfg, 6)), 0.02
    yield reft16_case(), 0.25
    yield reft16_mix2_case(), 0.25


def reft16_case():
    """The configuration of the reference test reft_ocean
    (test/ocean/reft_ocean_params.xml: 16x16x16 gateway mask over
    300-340E x 20-60N, periodic, topography from the mask, idealised
    wind/temperature/salinity forcing of Forcing Type 2, restoring T and S)
    at the steady state that the test stores (test/ocean/ocean_reference.h5,
    converted to reft/ocean_reference_state.npz by make_reft_state.py), with the stored
    parameters.  The test runs Mixing = 2, whose right-hand side at this state
    is bitwise the one of Mixing = 1 (vmix_control switches both mixing terms
    on once T and S are nonzero, mix_imp.f:131-166; only the grouping of the
    finite-difference mixing Jacobian differs, vmix_el_2), so the fixture uses
    Mixing = 1, the form the GPU library implements."""
    from thcm_cases import PAR
    z = np.load(os.path.join(HERE, "reft", "ocean_reference_state.npz"), allow_pickle=False)
    cfg = make_config(n=16, m=16, l=16, xmin_deg=300.0, xmax_deg=340.0, ymin_deg=20.0, ymax_deg=60.0,
                      periodic=1, hdim=4000.0, qz=1.0, vmix=1, tap=1, rho_mixing=0, tres=1, sres=1,
                      ite=1, its=1, iza=2, forcing_type=2)
    landm = RefModel.reference_landm(cfg, 0, 1, "mask_gateway")
    assert np.array_equal(landm, z["mask"]), "mask_gateway differs from the stored MaskGlobal"
    name = {v: k for k, v in PAR.items()}
    pars = {name[i]: float(z["par"][i]) for i in range(1, 31)}
    return Case("reft16", cfg, landm, {}, pars, z["x"])


def mixing2_sequence():
    """Mixing = 2 (the adaptive partition of mix_imp.f:131-227) on the natl8
    basin: one fresh model evaluated, as Ocean::computeRHS and then
    Ocean::computeJacobian do, at a sequence of states that switches the
    T / S mixing on and off: T only (pattern built without salinity), T and
    S (pattern rebuilt), S only (T switched off, pattern kept), neither, T and
    S again.  The T-only Jacobian is stored as the reference gives it; its
    entries depend on the reference's DSM column grouping."""
    cfg = make_config(n=8, m=8, l=4, rho_mixing=0, sres=0, vmix=2)
    landm = RefModel.reference_landm(cfg, 0, 1, "mask_natl8")
    case = Case("natl8_mix2", cfg, landm, {}, dict(COMB=0.7, WIND=1.0, SALT=1.0, TEMP=10.0), rand_state(cfg, 8))
    x = case.x
    var = np.arange(x.size) % 6
    states = [("t_only", np.where(var == 5, 0.0, x)), ("t_and_s", x), ("s_only", np.where(var == 4, 0.0, x)),
              ("neither", np.where((var == 4) | (var == 5), 0.0, x)), ("t_and_s_again", 0.5 * x)]
    ref = RefModel(case)
    out = dict(cfg=json.dumps(config_dict(cfg)), pars=json.dumps(case.pars), landm=landm,
               intcorr=np.float64(case.intcorr), labels=np.array([k for k, _ in states]))
    for q, (lab, xs) in enumerate(states):
        out[f"x{q}"] = xs
        out[f"F{q}"] = ref.rhs(xs)
        rp, ci, val, dB, dropped = ref.jacobian(xs)
        assert dropped == 0
        out[f"J{q}"] = val
    out["J_rowptr"], out["J_col"] = rp, ci
    path = os.path.join(HERE, "mix2", "natl8_sequence.npz")
    np.savez_compressed(path, **out)
    print(f"{path}: {len(states)} states, size={os.path.getsize(path)} B")


def fortran_layout():
    """The Fortran interface itself on the natl8 and gateway16 fixtures:
    rhs_'s B, and after matrix_ the caller-allocated buffers that set_pointers
    registered (1-based begA / jcoA / coA in fillcolA's order, coB)."""
    for name in ("natl8", "gateway16"):
        case, _ = load_golden_case(name)
        ref = RefModel(case)
        B = ref.fortran_rhs()
        ref.jacobian()
        beg, jco, co = ref.fortran_csr()
        path = os.path.join(HERE, "fortran", name + ".npz")
        np.savez_compressed(path, B=B, begA=beg, jcoA=jco, coA=co, coB=ref.fortran_coB())
        print(f"{path}: nnz={len(co)} size={os.path.getsize(path)} B")


def load_golden_case(name):
    from thcm_cases import load_golden
    return load_golden(name)


def reft16_mix2_case():
    """reft16 with the reference test's own Mixing = 2."""
    c = reft16_case()
    c.cfg.vmix = 2
    c.name = "reft16_mix2"
    return c


if __name__ == "__main__":
    only = set(sys.argv[1:])   # case names to write (all when none are given)
    for case, frac in cases():
        if not only or case.name in only:
            write(case, frac)
    if not only or "natl8_mix2" in only:
        mixing2_sequence()
    if not only or "fortran" in only:
        fortran_layout()
