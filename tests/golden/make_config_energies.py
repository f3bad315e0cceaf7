"""Tightly converged FP64 CCD energies of the benchmark systems, from the CPU
oracle (the reference's damped-Jacobi iteration run to max|dT| < 1e-11).

Writes tests/golden/config_energies.json.  The reference ships no CCD output
for these systems, so these oracle values are the FP64 fixed point that the
GPU solves (DIIS, mixed precision, batched, served) are checked against.

    python tests/golden/make_config_energies.py [name ...]
"""
from __future__ import annotations

import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from tests.helpers import info_for, oracle_for  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
PATH = os.path.join(HERE, "config_energies.json")
TOL = 1e-11


def main(names):
    data = json.load(open(PATH)) if os.path.exists(PATH) else {}
    for name in names:
        info = info_for(name)
        orc = oracle_for(info)
        e2, *_ = orc.mp2()
        t = time.time()
        ref = orc.ccd()                       # the reference's own stopping rule
        tight = orc.ccd(tolmax=TOL, maxiter=3000)
        data[name] = {
            "nao": orc.nao, "nocc": orc.nocc, "ehf": orc.ehf, "e2": e2,
            "ecorr_reference_mode": ref["ecorr"], "niter_reference_mode": ref["niter"],
            "maxres_reference_mode": ref["maxres"],
            "ecorr_tight": tight["ecorr"], "niter_tight": tight["niter"], "tol_dt_tight": TOL,
            "oracle_seconds": time.time() - t,
        }
        print(name, data[name], flush=True)
        with open(PATH, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main(sys.argv[1:] or ["n14_cut4", "n14_cut5", "n54_cut9", "n38_cut9", "n14_cut25",
                          "n54_cut16"])
