"""GTFock's own SCF driver, pscf/scf.c compiled unchanged and linked against
libgtfock_hip.so (oracle/_ref/scf, built by `make -C oracle ref` where the
GTFock sources are present), run on the GPU: its call sequence --
CInt_loadBasisSet, PFock_create, the initial guess through
PFock_putDenMat / PFock_commitDenMats, PFock_computeFock, PFock_getMat,
compute_diis, compute_purification -- goes through the drop-in C ABI.

Bars: the energies it prints (10 decimals) match the oracle's SCF iteration
by iteration within 1e-9 Eh: water (10 iterations, the oracle run here) and
C10H22 (15 iterations, tests/golden/scf_c10h22_oracle.json; converged
energy within 1e-9 Eh, the middle iterations within ten times the spread of
the oracle's own reruns with other thread counts, as the reference
algorithm amplifies FP64 rounding there)."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle

gpu = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
DRIVER = os.path.join(ROOT, "oracle", "_ref", "scf")

pytestmark = pytest.mark.skipif(not os.path.exists(DRIVER), reason="oracle/_ref/scf not built (GTFock sources absent)")


def run_driver(gbs, xyz, niters):
    r = subprocess.run([DRIVER, gbs, xyz, "1", "1", "1", "4", str(niters)], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    e = [float(x) for x in re.findall(r"^\s+energy (-?[0-9.]+) \(", r.stdout, re.M)]
    return np.array(e), r.stdout


@gpu
def test_reference_driver_water(gbs, mol_dir):
    xyz = str(mol_dir / "water.xyz")
    e, out = run_driver(gbs, xyz, 10)
    eo, _, _ = oracle.Basis(gbs, xyz).scf(10)
    n = min(len(e), len(eo))
    print(out[-600:])
    print("water, GTFock scf.c on the GPU: %d iterations, E = %.10f; oracle %.10f; max |dE| %.2e" % (
        len(e), e[-1], eo[n - 1], np.abs(e[:n] - eo[:n]).max()))
    assert len(e) == len(eo)
    assert np.abs(e - eo).max() < 1e-9


@gpu
def test_reference_driver_c10h22(gbs, mol_dir, c10h22_gold):
    gold, eo, tol = c10h22_gold
    e, out = run_driver(gbs, str(mol_dir / "c10h22.xyz"), gold["iterations"])
    assert len(e) == len(eo)
    d = np.abs(e - eo)
    print("C10H22, GTFock scf.c on the GPU: %d iterations, E = %.10f; oracle %.10f; |dE| per iteration %s" % (
        len(e), e[-1], eo[-1], " ".join("%.1e" % x for x in d)))
    assert np.all(d <= tol)
    assert d[-1] < 1e-9
