"""Writes quadpack_golden.json: DQNG / DQAGI results of the reference's own
QUADPACK (src/quadpack.f90, compiled into oracle/_ref by oracle/Makefile) for
the SRM integrand srm_fn (src/scarflib_srm.f90:975-995) at the parameter sets
of the benchmark configurations and the reference's driver examples.

Run in the container that holds the reference tree:
    make -C oracle ref && python tests/golden/make_quadpack_golden.py
"""
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
from oracle_lib import RefQuadpack  # noqa: E402
from test_oracle_quadpack import CASES  # noqa: E402


def main():
    ref = RefQuadpack()
    cases = []
    for acf, ndim, hurst, k0, k1 in CASES:
        for single in (0, 1):
            nu = hurst + ndim / 2.0
            if single:
                nu = float(np.float32(nu))
            for finite in (1, 0):
                v, err, neval, ier = ref.intg(acf, ndim, nu, single, finite, k0, k1)
                cases.append(dict(acf=acf, ndim=ndim, nu=nu, single=single, finite=finite, k0=k0, k1=k1, value=v,
                                  abserr=err, neval=neval, ier=ier))
    out = {"source": "reference QUADPACK (src/quadpack.f90) via oracle/_ref/libref_quadpack.so",
           "integrand": "srm_fn(x) = x * psd(real(x)) [* x in 3D]; psd: 1/(1+x^2)^nu (acf 0), exp(-x^2/4) (acf 1)",
           "tolerances": {"epsabs": 1e-8, "epsrel": 1e-5, "dqagi_limit": 500}, "cases": cases}
    (HERE / "quadpack_golden.json").write_text(json.dumps(out, indent=1))
    print(len(cases), "cases")


if __name__ == "__main__":
    main()
