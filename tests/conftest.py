import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI library)")


@pytest.fixture(scope="session")
def gbs():
    return os.path.join(ROOT, "gtfock_amd", "data", "cc-pvdz.gbs")


@pytest.fixture(scope="session")
def sto3g():
    return os.path.join(ROOT, "gtfock_amd", "data", "sto-3g.gbs")


@pytest.fixture(scope="session")
def mol_dir(tmp_path_factory):
    """Write the workload geometries once per session."""
    from gtfock_amd import geometry as g
    d = tmp_path_factory.mktemp("mol")
    mols = {
        "water": g.water(),
        "methane": g.alkane(1),
        "ethane": g.alkane(2),
        "butane": g.alkane(4),
        "c10h22": g.alkane(10),
        "benzene": g.graphene_flake(1),
        "c24h12": g.graphene_flake(2),
        "c54h18": g.graphene_flake(3),
    }
    for k, v in mols.items():
        (d / (k + ".xyz")).write_text(g.to_xyz(v))
    (d / "h2.xyz").write_text("2\n\nH 0 0 0\nH 0 0 %.12f\n" % (1.4 * 0.529177210903))
    return d


@pytest.fixture(scope="session")
def c10h22_gold():
    """The oracle's 15-iteration C10H22 SCF energies and, per iteration, the
    tolerance of a comparison with them: 1e-9 Eh, or ten times the spread
    of the oracle's own reruns with other OpenMP thread counts where that is
    larger (the reference algorithm's middle iterations amplify FP64
    summation-order differences; its converged energy does not move)."""
    import json
    import numpy as np
    with open(os.path.join(ROOT, "tests", "golden", "scf_c10h22_oracle.json")) as fh:
        gold = json.load(fh)
    e = np.array(gold["energies"])
    spread = np.zeros_like(e)
    for r in gold.get("reruns", {}).values():
        spread = np.maximum(spread, np.abs(np.array(r) - e))
    return gold, e, np.maximum(1e-9, 10.0 * spread)
