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


class HipRuntime:
    """The HIP runtime the library itself runs on (found in this process's
    maps after the library is loaded), for device buffers and memory
    queries in the tests: the one runtime every GPU object of the test
    shares."""

    def __init__(self):
        import ctypes
        import gtfock_amd
        gtfock_amd.lib()
        path = None
        with open("/proc/self/maps") as fh:
            for line in fh:
                if "libamdhip64.so" in line:
                    path = line.split()[-1]
                    break
        assert path, "libamdhip64 is not loaded"
        self.path = path
        self.rt = ctypes.CDLL(path)
        self.rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.rt.hipFree.argtypes = [ctypes.c_void_p]
        self.rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.rt.hipMemGetInfo.argtypes = [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]
        self._c = ctypes

    def _ok(self, st, what):
        assert st == 0, "%s failed (%d)" % (what, st)

    def malloc(self, nbytes):
        p = self._c.c_void_p()
        self._ok(self.rt.hipMalloc(self._c.byref(p), nbytes), "hipMalloc")
        return p.value

    def free(self, p):
        self._ok(self.rt.hipFree(p), "hipFree")

    def upload(self, p, a):
        """host array -> device (a contiguous numpy array)"""
        self._ok(self.rt.hipMemcpy(p, a.ctypes.data, a.nbytes, 1), "hipMemcpy H2D")

    def download(self, a, p):
        self._ok(self.rt.hipMemcpy(a.ctypes.data, p, a.nbytes, 2), "hipMemcpy D2H")

    def sync(self):
        self._ok(self.rt.hipDeviceSynchronize(), "hipDeviceSynchronize")

    def free_bytes(self):
        free, total = self._c.c_size_t(), self._c.c_size_t()
        self._ok(self.rt.hipMemGetInfo(self._c.byref(free), self._c.byref(total)), "hipMemGetInfo")
        return free.value


@pytest.fixture(scope="session")
def hip():
    return HipRuntime()
