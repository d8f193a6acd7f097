"""Write tests/golden/scf_c10h22_oracle.json: the energies of the oracle's
FP64 SCF (pscf/scf.c:348-459 restated in oracle/gtf_oracle.c) on all-trans
C10H22 / cc-pVDZ, zero starting density, Schwarz tolerance 1e-11, 15
iterations.  The GPU SCF is compared with it iteration by iteration.

Run:  python tests/golden/make_oracle_scf.py   (a few minutes on 8 cores)
"""
import json
import os
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from gtfock_amd import geometry  # noqa: E402

GBS = os.path.join(ROOT, "gtfock_amd", "data", "cc-pvdz.gbs")


def main():
    with tempfile.TemporaryDirectory() as d:
        xyz = os.path.join(d, "c10h22.xyz")
        with open(xyz, "w") as fh:
            fh.write(geometry.to_xyz(geometry.alkane(10)))
        t0 = time.time()
        e, _, _ = oracle.Basis(GBS, xyz).scf(15, tol=1e-11)
        dt = time.time() - t0
    out = {"molecule": "all-trans C10H22 (C-C 1.54, C-H 1.09 A), cc-pVDZ", "tolscr": 1e-11,
           "start": "zero density (core-Hamiltonian guess)", "iterations": len(e),
           "energies": [float(x) for x in e], "seconds": round(dt, 1),
           "command": "python tests/golden/make_oracle_scf.py"}
    with open(os.path.join(HERE, "scf_c10h22_oracle.json"), "w") as fh:
        json.dump(out, fh, indent=1)
    print("C10H22 oracle SCF: %d iterations, E = %.12f, %.0f s" % (len(e), e[-1], dt))


if __name__ == "__main__":
    main()
