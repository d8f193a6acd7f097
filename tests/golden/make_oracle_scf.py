"""Write tests/golden/scf_c10h22_oracle.json: the energies of the oracle's
FP64 SCF (pscf/scf.c:348-459 restated in oracle/gtf_oracle.c) on all-trans
C10H22 / cc-pVDZ, zero starting density, Schwarz tolerance 1e-11, 15
iterations.  The GPU SCF is compared with it iteration by iteration.

The same SCF rerun with other OpenMP thread counts (another FP64 summation
order in the oracle's Fock build) is stored beside it: the reference
algorithm's trajectory amplifies such rounding differences by up to ~1e5
in its middle iterations (DIIS on nearly dependent error vectors) before
converging to the same energy, and the GPU comparison allows that spread.

Run:  python tests/golden/make_oracle_scf.py              (all; ~20 min on 8 cores)
      python tests/golden/make_oracle_scf.py --add-reruns 3 5
                                        (keep the energies, add the reruns)
"""
import json
import os
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from gtfock_amd import geometry  # noqa: E402

GBS = os.path.join(ROOT, "gtfock_amd", "data", "cc-pvdz.gbs")
OUT = os.path.join(HERE, "scf_c10h22_oracle.json")
RERUN_THREADS = (3, 5)


def run_scf():
    with tempfile.TemporaryDirectory() as d:
        xyz = os.path.join(d, "c10h22.xyz")
        with open(xyz, "w") as fh:
            fh.write(geometry.to_xyz(geometry.alkane(10)))
        t0 = time.time()
        e, _, _ = oracle.Basis(GBS, xyz).scf(15, tol=1e-11)
        return [float(x) for x in e], time.time() - t0


def rerun(threads):
    """the SCF in a child process with OMP_NUM_THREADS = threads"""
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    out = subprocess.run([sys.executable, os.path.abspath(__file__), "--child"], env=env, check=True,
                         capture_output=True, text=True).stdout
    return json.loads(out.strip().splitlines()[-1])


def main(argv):
    if argv[:1] == ["--child"]:
        e, _ = run_scf()
        print(json.dumps(e))
        return
    if argv[:1] == ["--add-reruns"]:
        with open(OUT) as fh:
            out = json.load(fh)
        threads = [int(t) for t in argv[1:]] or list(RERUN_THREADS)
    else:
        e, dt = run_scf()
        out = {"molecule": "all-trans C10H22 (C-C 1.54, C-H 1.09 A), cc-pVDZ", "tolscr": 1e-11,
               "start": "zero density (core-Hamiltonian guess)", "iterations": len(e),
               "energies": e, "seconds": round(dt, 1), "threads": os.environ.get("OMP_NUM_THREADS", "default"),
               "command": "python tests/golden/make_oracle_scf.py"}
        threads = list(RERUN_THREADS)
    reruns = out.get("reruns", {})
    for t in threads:
        reruns[str(t)] = rerun(t)
        d = [abs(a - b) for a, b in zip(reruns[str(t)], out["energies"])]
        print("rerun with %d threads: max |dE| %.2e, final %.2e" % (t, max(d), d[-1]))
    out["reruns"] = reruns
    with open(OUT, "w") as fh:
        json.dump(out, fh, indent=1)
    print("C10H22 oracle SCF: %d iterations, E = %.12f" % (out["iterations"], out["energies"][-1]))


if __name__ == "__main__":
    main(sys.argv[1:])
