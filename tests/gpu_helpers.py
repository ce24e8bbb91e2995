"""Helpers shared by the GPU tests: the product library's Python binding and
builders that turn a test Case into a GPU Ocean."""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "i-emic_amd", "python"))

import iemic_amd  # noqa: E402


def gpu_ocean(case):
    cfg = iemic_amd.Config(**{f: getattr(case.cfg, f) for f, _ in iemic_amd.Config._fields_})
    f = case.fields
    oc = iemic_amd.Ocean(cfg, case.landm, f.get("taux"), f.get("tauy"), f.get("tatm"),
                         f.get("emip"), f.get("spert"))
    for k, v in case.pars.items():
        oc.setPar(k, v)
    oc.setIntCondCorrection(value=case.intcorr)
    return oc
