"""MI355X-native GTFock Fock build and purification -- Python host mirror.

The product is the C-ABI library ``gtfock_amd/_lib/libgtfock_hip.so`` (HIP
kernels for gfx950 behind GTFock's ``PFock_*``, ``CInt_*`` and purif API).
This module binds it with ctypes and mirrors the reference's call sequence
(pscf/scf.c) in Python: the same names, argument meaning and status codes.
It never computes anything itself and raises if the library is missing.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libgtfock_hip.so")
DATA_DIR = os.path.join(_HERE, "data")
# set-up phases timed by PFock.setup_times / Batch.setup_times
SETUP_PHASES = ("basis", "pairs", "schwarz", "sig_order", "row_count", "tiles")
# fields of a quartet class's tile shape (gtf_pfock_tile_shape)
from .tuning import SHAPE_KEYS, flat as _flat_shapes  # noqa: E402
CSRC_DIR = os.path.join(_HERE, "csrc")

PFOCK_STATUS = {0: "SUCCESS", 1: "INIT_FAILED", 2: "ALLOC_FAILED", 3: "INVALID_VALUE",
                4: "EXECUTION_FAILED", 5: "INTERNAL_ERROR"}
MAT_D, MAT_F, MAT_J, MAT_K = 0, 1, 2, 3
PAIR_TYPES = ["ss", "ps", "pp", "ds", "dp", "dd"]


def class_names():
    out = []
    for a in range(6):
        for b in range(a + 1):
            out.append("(%s|%s)" % (PAIR_TYPES[a], PAIR_TYPES[b]))
    return out


def build(jobs=8):
    """Compile libgtfock_hip.so for gfx950 with hipcc (make -C csrc)."""
    subprocess.check_call(["make", "-s", "-j%d" % jobs, "-C", CSRC_DIR])


_lib = None
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_long)
_vp = ctypes.c_void_p


class GtfError(RuntimeError):
    pass


def lib():
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise GtfError("libgtfock_hip.so is not built (%s); run gtfock_amd.build()" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    sig = {
        "CInt_createBasisSet": ([ctypes.POINTER(_vp)], ctypes.c_int),
        "CInt_destroyBasisSet": ([_vp], ctypes.c_int),
        "CInt_loadBasisSet": ([_vp, ctypes.c_char_p, ctypes.c_char_p], ctypes.c_int),
        "CInt_packBasisSet": ([_vp, ctypes.POINTER(_vp), _ip], ctypes.c_int),
        "CInt_unpackBasisSet": ([_vp, _vp], ctypes.c_int),
        "CInt_getNumShells": ([_vp], ctypes.c_int),
        "CInt_getNumFuncs": ([_vp], ctypes.c_int),
        "CInt_getNumAtoms": ([_vp], ctypes.c_int),
        "CInt_getNumOccOrb": ([_vp], ctypes.c_int),
        "CInt_getMaxShellDim": ([_vp], ctypes.c_int),
        "CInt_getShellDim": ([_vp, ctypes.c_int], ctypes.c_int),
        "CInt_getFuncStartInd": ([_vp, ctypes.c_int], ctypes.c_int),
        "CInt_getFuncEndInd": ([_vp, ctypes.c_int], ctypes.c_int),
        "CInt_getAtomStartInd": ([_vp, ctypes.c_int], ctypes.c_int),
        "CInt_getTotalCharge": ([_vp], ctypes.c_int),
        "CInt_getNneutral": ([_vp], ctypes.c_int),
        "CInt_getNucEnergy": ([_vp], ctypes.c_double),
        "CInt_getInitialGuess": ([_vp, ctypes.c_int, ctypes.POINTER(_dp), _ip, _ip], None),
        "CInt_createERD": ([_vp, ctypes.POINTER(_vp), ctypes.c_int], ctypes.c_int),
        "CInt_destroyERD": ([_vp], ctypes.c_int),
        "CInt_computeShellQuartet": ([_vp, _vp, ctypes.c_int] + [ctypes.c_int] * 4 + [ctypes.POINTER(_dp), _ip],
                                     ctypes.c_int),
        "CInt_createOED": ([_vp, ctypes.POINTER(_vp)], ctypes.c_int),
        "CInt_destroyOED": ([_vp], ctypes.c_int),
        "CInt_computePairOvl": ([_vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_dp), _ip], ctypes.c_int),
        "CInt_computePairCoreH": ([_vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_dp), _ip], ctypes.c_int),
        "PFock_create": ([_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                          ctypes.c_int, ctypes.POINTER(_vp)], ctypes.c_int),
        "PFock_destroy": ([_vp], ctypes.c_int),
        "PFock_setNumDenMat": ([ctypes.c_int, _vp], ctypes.c_int),
        "PFock_putDenMat": ([ctypes.c_int] * 5 + [_dp, ctypes.c_int, _vp], ctypes.c_int),
        "PFock_fillDenMat": ([ctypes.c_double, ctypes.c_int, _vp], ctypes.c_int),
        "PFock_commitDenMats": ([_vp], ctypes.c_int),
        "PFock_sync": ([_vp], ctypes.c_int),
        "PFock_getMat": ([_vp, ctypes.c_int, ctypes.c_int] + [ctypes.c_int] * 5 + [_dp], ctypes.c_int),
        "PFock_computeFock": ([_vp, _vp], ctypes.c_int),
        "PFock_createCoreHMat": ([_vp, _vp], ctypes.c_int),
        "PFock_destroyCoreHMat": ([_vp], ctypes.c_int),
        "PFock_getCoreHMat": ([_vp] + [ctypes.c_int] * 5 + [_dp], ctypes.c_int),
        "PFock_createOvlMat": ([_vp, _vp], ctypes.c_int),
        "PFock_destroyOvlMat": ([_vp], ctypes.c_int),
        "PFock_getOvlMat": ([_vp] + [ctypes.c_int] * 5 + [_dp], ctypes.c_int),
        "PFock_getOvlMat2": ([_vp] + [ctypes.c_int] * 5 + [_dp], ctypes.c_int),
        "PFock_getMemorySize": ([_vp, _dp], ctypes.c_int),
        "PFock_getStatistics": ([_vp], ctypes.c_int),
        "gtf_pfock_stats": ([_vp, _dp], ctypes.c_int),
        "gtf_pfock_fused_densities": ([_vp], ctypes.c_int),
        "gtf_pfock_balance": ([_vp, _dp], ctypes.c_int),
        "gtf_pfock_setup_times": ([_vp, _dp], ctypes.c_int),
        "gtf_pfock_csr": ([_vp, _ip, _ip, _ip, _dp], ctypes.c_int),
        "gtf_pfock_class_counts": ([_vp, _lp], ctypes.c_int),
        "gtf_pfock_quartets": ([_vp, _ip, ctypes.c_long], ctypes.c_long),
        "gtf_pfock_device_ptrs": ([_vp, ctypes.c_int, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                   ctypes.POINTER(_vp)], ctypes.c_int),
        "gtf_pfock_stream": ([_vp], _vp),
        "gtf_pfock_class_info": ([_vp, ctypes.c_int, _ip], ctypes.c_int),
        "gtf_pfock_class_work": ([_vp, _lp, _dp], ctypes.c_int),
        "gtf_pfock_set_timing": ([_vp, ctypes.c_int], ctypes.c_int),
        "gtf_pfock_set_serial": ([_vp, ctypes.c_int], ctypes.c_int),
        "gtf_pfock_set_precision": ([_vp, ctypes.c_double], ctypes.c_int),
        "gtf_pfock_set_prim_tol": ([_vp, ctypes.c_double], ctypes.c_int),
        "gtf_pfock_prim_census": ([_vp, _dp, _dp], ctypes.c_int),
        "gtf_pfock_mixed_counts": ([_vp, _lp, _dp], ctypes.c_int),
        "gtf_pfock_class_times": ([_vp, _dp, _lp], ctypes.c_int),
        "gtf_pfock_device_bytes": ([_vp], ctypes.c_double),
        "gtf_pfock_autotune": ([_vp, ctypes.c_int, _dp], ctypes.c_int),
        "gtf_pfock_tile_shape": ([_vp, ctypes.c_int, _ip], ctypes.c_int),
        "gtf_pfock_set_tile_shape": ([_vp, ctypes.c_int, _ip], ctypes.c_int),
        "gtf_pfock_set_tile_shapes": ([_vp, _ip], ctypes.c_int),
        "gtf_batch_set_tile_shapes": ([_vp, _ip], ctypes.c_int),
        "gtf_server_set_tile_shapes": ([_vp, _ip], ctypes.c_int),
        "gtf_server_set_tile_shapes_range": ([_vp, ctypes.c_int, ctypes.c_int, _ip], ctypes.c_int),
        "gtf_server_set_large_slots": ([_vp, ctypes.c_int], ctypes.c_int),
        "gtf_server_set_purif_precision": ([_vp, ctypes.c_int, ctypes.c_double], ctypes.c_int),
        "gtf_server_set_join_gap": ([_vp, ctypes.c_int], ctypes.c_int),
        "gtf_server_set_large_yield": ([_vp, ctypes.c_double], ctypes.c_int),
        "gtf_server_join_stats": ([_vp, _lp, _lp], ctypes.c_int),
        "gtf_batch_scf_begin": ([_vp, ctypes.c_int, ctypes.c_double, _vp, _vp], ctypes.c_int),
        "gtf_batch_scf_step": ([_vp], ctypes.c_int),
        "gtf_batch_scf_results": ([_vp, _dp, _ip, _dp], ctypes.c_int),
        "gtf_batch_regroup": ([_vp, ctypes.POINTER(_vp), ctypes.c_int, ctypes.POINTER(_vp), _ip], ctypes.c_int),
        "gtf_batch_autotune": ([_vp, ctypes.c_int, _dp], ctypes.c_int),
        "gtf_batch_tile_shape": ([_vp, ctypes.c_int, _ip], ctypes.c_int),
        "gtf_batch_set_tile_shape": ([_vp, ctypes.c_int, _ip], ctypes.c_int),
        "gtf_erd_quartets": ([_vp, ctypes.c_int, _ip, _dp], ctypes.c_int),
        "gtf_dgemm_host": ([ctypes.c_int, _dp, _dp, _dp, ctypes.c_int], ctypes.c_int),
        "gtf_gemm_host": ([ctypes.c_int, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int], ctypes.c_int),
        "gtf_purif_set_precision": ([_vp, ctypes.c_int, ctypes.c_double], ctypes.c_int),
        "gtf_purif_fp32_iters": ([_vp], ctypes.c_int),
        "gtf_purif_set_slices": ([_vp, ctypes.c_int], ctypes.c_int),
        "gtf_purif_set_split_residual": ([_vp, ctypes.c_double], ctypes.c_int),
        "gtf_batch_set_purif_residual": ([_vp, ctypes.c_double], ctypes.c_int),
        "gtf_tile_shapes_check": ([_ip], ctypes.c_int),
        "gtf_batch_job_bytes": ([_vp], ctypes.c_double),
        "gtf_gemm_split_host": ([ctypes.c_int, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int], ctypes.c_int),
        "gtf_batch_set_purif_precision": ([_vp, ctypes.c_int, ctypes.c_double], ctypes.c_int),
        "gtf_batch_set_purif_slices": ([_vp, ctypes.c_int], ctypes.c_int),
        "create_purif": ([_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int], _vp),
        "destroy_purif": ([_vp], None),
        "compute_purification": ([_vp, _dp, _dp], ctypes.c_int),
        "compute_diis": ([_vp, _vp, _dp, _dp, ctypes.c_int], None),
        "gtf_batch_create": ([ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_double, ctypes.POINTER(_vp)],
                             ctypes.c_int),
        "gtf_batch_destroy": ([_vp], None),
        "gtf_batch_layout": ([_vp, _ip, _lp], ctypes.c_int),
        "gtf_batch_total": ([_vp], ctypes.c_long),
        "gtf_batch_set_precision": ([_vp, ctypes.c_double], ctypes.c_int),
        "gtf_batch_set_prim_tol": ([_vp, ctypes.c_double], ctypes.c_int),
        "gtf_batch_use_graph": ([_vp, ctypes.c_int], ctypes.c_int),
        "gtf_batch_put_density": ([_vp, ctypes.c_int, _dp], ctypes.c_int),
        "gtf_batch_compute_fock": ([_vp], ctypes.c_int),
        "gtf_batch_get_fock": ([_vp, ctypes.c_int, _dp], ctypes.c_int),
        "gtf_batch_get_density": ([_vp, ctypes.c_int, _dp], ctypes.c_int),
        "gtf_batch_device_ptrs": ([_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp)], ctypes.c_int),
        "gtf_batch_stream": ([_vp], _vp),
        "gtf_batch_stats": ([_vp, _dp], ctypes.c_int),
        "gtf_batch_setup_times": ([_vp, _dp], ctypes.c_int),
        "gtf_batch_class_work": ([_vp, _lp, _dp, _lp, _dp], ctypes.c_int),
        "gtf_batch_scf": ([_vp, ctypes.c_int, ctypes.c_double, _dp, _ip, _dp], ctypes.c_int),
        "gtf_batch_done_times": ([_vp, _dp], ctypes.c_int),
        "gtf_server_create": ([ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                               ctypes.c_int, ctypes.c_double, ctypes.POINTER(_vp)], ctypes.c_int),
        "gtf_server_submit": ([_vp, _vp], ctypes.c_long),
        "gtf_server_poll": ([_vp, ctypes.c_long, _dp], ctypes.c_int),
        "gtf_server_wait": ([_vp], ctypes.c_int),
        "gtf_server_stats": ([_vp, _dp], ctypes.c_int),
        "gtf_server_destroy": ([_vp], None),
    }
    # a library older than this binding lacks some entry points: only calls
    # to those fail (AttributeError); tests/test_host.py checks that the
    # built library exports every symbol the headers declare
    for name, (args, res) in sig.items():
        if not hasattr(L, name):
            MISSING_SYMBOLS.append(name)
            continue
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


MISSING_SYMBOLS = []


def exported_symbols():
    """Names the C headers under include/ declare (parsed from the headers)."""
    import re
    names = []
    inc = os.path.join(os.path.dirname(_HERE), "include")
    for h in sorted(os.listdir(inc)):
        if not h.endswith(".h"):
            continue
        txt = open(os.path.join(inc, h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", txt):
            n = m.group(1)
            if n.startswith(("PFock_", "CInt_", "gtf_", "create_purif", "destroy_purif", "compute_")):
                names.append(n)
    return sorted(set(names))


def _check(st, what):
    if st != 0:
        raise GtfError("%s failed: %s" % (what, PFOCK_STATUS.get(st, st)))


def _p(a):
    return a.ctypes.data_as(_dp)


class BasisSet:
    """CInt basis set (libcint/include/CInt.h:45-145)."""

    def __init__(self, gbs, xyz):
        L = lib()
        self.h = _vp()
        _check(L.CInt_createBasisSet(ctypes.byref(self.h)), "CInt_createBasisSet")
        _check(L.CInt_loadBasisSet(self.h, gbs.encode(), xyz.encode()), "CInt_loadBasisSet")
        self.nshells = L.CInt_getNumShells(self.h)
        self.nbf = L.CInt_getNumFuncs(self.h)
        self.natoms = L.CInt_getNumAtoms(self.h)
        self.nocc = L.CInt_getNumOccOrb(self.h)
        self.charge = L.CInt_getTotalCharge(self.h)
        self.shell_dim = [L.CInt_getShellDim(self.h, s) for s in range(self.nshells)]
        self.shell_fstart = [L.CInt_getFuncStartInd(self.h, s) for s in range(self.nshells)] + [self.nbf]

    def nuc_energy(self):
        return lib().CInt_getNucEnergy(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().CInt_destroyBasisSet(self.h)
            self.h = None


class PFock:
    """The PFock engine (pfock/pfock.h:241-565) on one GPU."""

    def __init__(self, basis, nprow=1, npcol=1, ntasks=4, tolscr=1e-11, max_numdmat=1, symm=1):
        self.basis = basis
        self.h = _vp()
        _check(lib().PFock_create(basis.h, nprow, npcol, ntasks, tolscr, max_numdmat, symm,
                                  ctypes.byref(self.h)), "PFock_create")
        self.nbf = basis.nbf

    def close(self):
        if getattr(self, "h", None):
            lib().PFock_destroy(self.h)
            self.h = None

    __del__ = close

    def set_num_den_mat(self, n):
        return lib().PFock_setNumDenMat(n, self.h)

    def put_den_mat(self, D, index=0):
        D = np.ascontiguousarray(D, dtype=np.float64)
        n = self.nbf
        _check(lib().PFock_putDenMat(0, n - 1, 0, n - 1, n, _p(D), index, self.h), "PFock_putDenMat")

    def fill_den_mat(self, value, index=0):
        _check(lib().PFock_fillDenMat(value, index, self.h), "PFock_fillDenMat")

    def commit(self):
        _check(lib().PFock_commitDenMats(self.h), "PFock_commitDenMats")

    def compute_fock(self):
        _check(lib().PFock_computeFock(self.basis.h, self.h), "PFock_computeFock")

    def get_mat(self, mtype=MAT_F, index=0):
        n = self.nbf
        out = np.zeros((n, n))
        _check(lib().PFock_getMat(self.h, mtype, index, 0, n - 1, 0, n - 1, n, _p(out)), "PFock_getMat")
        return out

    def fock(self, D, index=0):
        """putDenMat + commit + computeFock + getMat(F): G(D) = 2J - K."""
        self.put_den_mat(D, index)
        self.commit()
        self.compute_fock()
        return self.get_mat(MAT_F, index)

    def _one(self, create, get):
        L = lib()
        _check(getattr(L, create)(self.h, self.basis.h), create)
        n = self.nbf
        out = np.zeros((n, n))
        _check(getattr(L, get)(self.h, 0, n - 1, 0, n - 1, n, _p(out)), get)
        return out

    def overlap(self):
        return self._one("PFock_createOvlMat", "PFock_getOvlMat")

    def x_matrix(self):
        return self._one("PFock_createOvlMat", "PFock_getOvlMat2")

    def core_hamiltonian(self):
        return self._one("PFock_createCoreHMat", "PFock_getCoreHMat")

    def stats(self):
        out = np.zeros(8)
        lib().gtf_pfock_stats(self.h, _p(out))
        keys = ["usq", "uitl", "nnz", "maxvalue", "gpu_seconds", "wall_seconds", "create_seconds", "device_bytes"]
        return dict(zip(keys, out.tolist()))

    def balance(self):
        """Load balance of the last build: slowest persistent workgroup over
        the mean (summed over the tile launches), tiles taken after each
        workgroup's first, workgroups."""
        out = np.zeros(3)
        _check(lib().gtf_pfock_balance(self.h, _p(out)), "gtf_pfock_balance")
        return dict(load_balance=out[0], steals=int(out[1]), workgroups=int(out[2]))

    def setup_times(self):
        """Wall seconds of PFock_create's set-up phases."""
        out = np.zeros(6)
        if lib().gtf_pfock_setup_times(self.h, _p(out)) != 6:
            raise GtfError("gtf_pfock_setup_times failed")
        return dict(zip(SETUP_PHASES, out.tolist()))

    def csr(self):
        ns = self.basis.nshells
        ptr = np.zeros(ns + 1, dtype=np.int32)
        cap = ns * ns
        sid = np.zeros(cap, dtype=np.int32)
        srid = np.zeros(cap, dtype=np.int32)
        sval = np.zeros(cap)
        nnz = lib().gtf_pfock_csr(self.h, ptr.ctypes.data_as(_ip), sid.ctypes.data_as(_ip),
                                  srid.ctypes.data_as(_ip), _p(sval))
        return dict(nnz=nnz, shellptr=ptr, shellid=sid[:nnz], shellrid=srid[:nnz], shellvalue=sval[:nnz])

    def class_counts(self):
        c = np.zeros(21, dtype=np.int64)
        lib().gtf_pfock_class_counts(self.h, c.ctypes.data_as(_lp))
        return c

    def quartets(self):
        n = lib().gtf_pfock_quartets(self.h, None, 0)
        q = np.zeros((max(n, 1), 4), dtype=np.int32)
        lib().gtf_pfock_quartets(self.h, q.ctypes.data_as(_ip), n)
        return q[:n]

    def class_work(self):
        nq = np.zeros(21, dtype=np.int64)
        npq = np.zeros(21)
        lib().gtf_pfock_class_work(self.h, nq.ctypes.data_as(_lp), _p(npq))
        return nq, npq

    def set_precision(self, tau):
        """Mixed precision: FP32 for quartets with Schwarz bound < tau (0: FP64)."""
        _check(lib().gtf_pfock_set_precision(self.h, float(tau)), "gtf_pfock_set_precision")

    def set_prim_tol(self, tol):
        """Primitive screening: skip primitive quartets with Schwarz bound < tol (0: none)."""
        _check(lib().gtf_pfock_set_prim_tol(self.h, float(tol)), "gtf_pfock_set_prim_tol")

    def prim_census(self):
        """Per class: primitive quartets the kernels evaluate (FP64 tiles, FP32 tiles)."""
        hi, lo = np.zeros(21), np.zeros(21)
        if lib().gtf_pfock_prim_census(self.h, _p(hi), _p(lo)) != 21:
            raise GtfError("gtf_pfock_prim_census failed")
        return hi, lo

    def mixed_counts(self):
        nq = np.zeros(21, dtype=np.int64)
        npq = np.zeros(21)
        lib().gtf_pfock_mixed_counts(self.h, nq.ctypes.data_as(_lp), _p(npq))
        return nq, npq

    def set_timing(self, on=True):
        _check(lib().gtf_pfock_set_timing(self.h, 1 if on else 0), "gtf_pfock_set_timing")

    def class_kernel(self, cls):
        """Name of the kernel that runs quartet class cls (FP64 part)."""
        info = np.zeros(11, dtype=np.int32)
        _check(lib().gtf_pfock_class_info(self.h, cls, info.ctypes.data_as(_ip)), "gtf_pfock_class_info")
        la, lb, lc, ld, nc = (int(x) for x in info[:5])
        blds = "true" if info[10] else "false"
        if info[6] > 0 and info[8]:
            return "eri_jk_tile_t_kernel<%d, %d, %d, %d, %d, double, %s>" % (la, lb, lc, ld, int(info[9]), blds)
        if info[6] > 0:
            return "eri_jk_tile_kernel<%d, %d, %d, %d, %d, double, %s>" % (la, lb, lc, ld, nc, blds)
        return "eri_kernel<%d, %d, %d, %d, %d, 0>" % (la, lb, lc, ld, nc)

    def autotune(self, reps=3):
        """Tile autotuner on density 0 (gtf_pfock_autotune): returns
        {"start_ms", "tuned_ms", "candidates"}; the tuned shapes stay in
        effect for later builds."""
        ms = np.zeros(3)
        _check(lib().gtf_pfock_autotune(self.h, int(reps), _p(ms)), "gtf_pfock_autotune")
        return {"start_ms": float(ms[0]), "tuned_ms": float(ms[1]), "candidates": int(ms[2])}

    def tile_shape(self, cls):
        """{units, unit_pq, tile_pq, grid, exchanged, boys_lds} of quartet class cls."""
        sh = np.zeros(len(SHAPE_KEYS), dtype=np.int32)
        st = lib().gtf_pfock_tile_shape(self.h, cls, sh.ctypes.data_as(_ip))
        _check(0 if st == len(SHAPE_KEYS) else st, "gtf_pfock_tile_shape")
        return dict(zip(SHAPE_KEYS, (int(x) for x in sh)))

    def set_tile_shape(self, cls, shape):
        sh = np.array([shape[k] for k in SHAPE_KEYS], dtype=np.int32)
        _check(lib().gtf_pfock_set_tile_shape(self.h, cls, sh.ctypes.data_as(_ip)), "gtf_pfock_set_tile_shape")

    def tile_shapes(self):
        return [self.tile_shape(c) for c in range(21)]

    def set_tile_shapes(self, shapes):
        """All 21 class shapes at once (e.g. a table from gtfock_amd.tuning)."""
        sh = np.array(_flat_shapes(shapes), dtype=np.int32)
        _check(lib().gtf_pfock_set_tile_shapes(self.h, sh.ctypes.data_as(_ip)), "gtf_pfock_set_tile_shapes")

    def set_serial(self, on):
        """Class kernels one after another (per-kernel measurement)."""
        _check(lib().gtf_pfock_set_serial(self.h, 1 if on else 0), "gtf_pfock_set_serial")

    def class_times(self):
        ms = np.zeros(21)
        n = np.zeros(21, dtype=np.int64)
        lib().gtf_pfock_class_times(self.h, _p(ms), n.ctypes.data_as(_lp))
        return ms, n

    def device_ptrs(self, index=0):
        d, j, k = _vp(), _vp(), _vp()
        _check(lib().gtf_pfock_device_ptrs(self.h, index, ctypes.byref(d), ctypes.byref(j), ctypes.byref(k)),
               "gtf_pfock_device_ptrs")
        return d.value, j.value, k.value

    def stream(self):
        return lib().gtf_pfock_stream(self.h)


class ERD:
    """CInt ERD handle: shell-quartet ERIs computed by the GPU kernels."""

    def __init__(self, basis, nthreads=1):
        self.basis = basis
        self.h = _vp()
        _check(lib().CInt_createERD(basis.h, ctypes.byref(self.h), nthreads), "CInt_createERD")

    def __del__(self):
        if getattr(self, "h", None):
            lib().CInt_destroyERD(self.h)
            self.h = None

    def quartet(self, M, N, P, Q):
        ptr = _dp()
        n = ctypes.c_int()
        _check(lib().CInt_computeShellQuartet(self.basis.h, self.h, 0, M, N, P, Q, ctypes.byref(ptr),
                                              ctypes.byref(n)), "CInt_computeShellQuartet")
        return np.ctypeslib.as_array(ptr, shape=(n.value,)).copy()

    def quartets(self, q):
        q = np.ascontiguousarray(q, dtype=np.int32).reshape(-1, 4)
        d = np.array(self.basis.shell_dim)
        sizes = d[q[:, 0]] * d[q[:, 1]] * d[q[:, 2]] * d[q[:, 3]]
        out = np.zeros(int(sizes.sum()))
        _check(lib().gtf_erd_quartets(self.h, len(q), q.ctypes.data_as(_ip), _p(out)), "gtf_erd_quartets")
        offs = np.concatenate([[0], np.cumsum(sizes)])
        return [out[offs[i]:offs[i + 1]] for i in range(len(q))]


def dgemm(A, B, transB=False, fp32=False, split=0):
    """MFMA GEMM on the GPU: A @ B, or A @ B.T when transB; FP64, FP32 on
    rounded operands when fp32, or split BF16 with `split` exact slices
    (3..6) of every operand row when split > 0."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    B = np.ascontiguousarray(B, dtype=np.float64)
    n = A.shape[0]
    C = np.zeros((n, n))
    if split:
        _check(lib().gtf_gemm_split_host(n, _p(A), _p(B), _p(C), 1 if transB else 0, int(split)),
               "gtf_gemm_split_host")
        return C
    _check(lib().gtf_gemm_host(n, _p(A), _p(B), _p(C), 1 if transB else 0, 1 if fp32 else 0), "gtf_gemm_host")
    return C


# purification product modes of gtf_purif_set_precision / gtf_batch_set_purif_precision
PURIF_MODES = {"fp64": 0, "fp32": 1, "split": 2}
# ||D - D^2||_F below which the split-BF16 purification finishes in FP64
PURIF_SPLIT_UNTIL = 1e-6


def _purif_mode(mode):
    if isinstance(mode, str):
        return PURIF_MODES[mode]
    if isinstance(mode, bool):
        return 1 if mode else 0
    return int(mode)


class Purif:
    """purif driver (pscf/purif.h:78-86) on one GPU."""

    def __init__(self, basis):
        self.h = lib().create_purif(basis.h, 1, 1, 1)
        if not self.h:
            raise GtfError("create_purif failed")
        self.n = basis.nbf

    def close(self):
        if getattr(self, "h", None):
            lib().destroy_purif(self.h)
            self.h = None

    __del__ = close

    def _fields(self):
        return ctypes.cast(self.h, ctypes.POINTER(PurifT)).contents

    def set_xs(self, X, S):
        f = self._fields()
        n = self.n
        ctypes.memmove(f.X_block, np.ascontiguousarray(X).ctypes.data, 8 * n * n)
        ctypes.memmove(f.S_block, np.ascontiguousarray(S).ctypes.data, 8 * n * n)

    def diis(self, pfock, D, F, it):
        D = np.ascontiguousarray(D, dtype=np.float64)
        F = np.array(F, dtype=np.float64, order="C")
        lib().compute_diis(pfock.h if pfock is not None else None, self.h, _p(D), _p(F), it)
        return F

    def set_precision(self, mode, until=None):
        """Reduced-precision purification products (gtf_purif_set_precision):
        mode "fp64" / "fp32" (True) / "split", until the error drops below
        `until` (default 1e-3 for FP32, PURIF_SPLIT_UNTIL for split)."""
        m = _purif_mode(mode)
        if until is None:
            until = PURIF_SPLIT_UNTIL if m == 2 else 1e-3
        _check(lib().gtf_purif_set_precision(self.h, m, until), "gtf_purif_set_precision")

    def set_slices(self, ns):
        """Slices per operand of the split-BF16 products (3..6, default 5)."""
        _check(lib().gtf_purif_set_slices(self.h, int(ns)), "gtf_purif_set_slices")

    def set_split_residual(self, resid):
        """Split products only while the SCF's last DIIS residual is at least
        resid (gtf_purif_set_split_residual; 0: in every SCF iteration)."""
        _check(lib().gtf_purif_set_split_residual(self.h, float(resid)), "gtf_purif_set_split_residual")

    def gpu_seconds(self):
        """device time of the last purification (its loop and X D X, HIP events)"""
        return float(self._fields().timedgemm)

    def fp32_iters(self):
        """iterations of the last purification run in the reduced precision"""
        return lib().gtf_purif_fp32_iters(self.h)

    def purify(self, F):
        F = np.ascontiguousarray(F, dtype=np.float64)
        D = np.zeros((self.n, self.n))
        it = lib().compute_purification(self.h, _p(F), _p(D))
        if it < 0:
            raise GtfError("compute_purification failed")
        return D, it


class Batch:
    """Many molecules per Fock build and per SCF iteration (gtf_batch_*).

    Matrices are packed: molecule m is the nbf[m] x nbf[m] block at
    offset[m] of one array.  ``fock(Ds)`` takes and returns a list of
    per-molecule matrices; G = 2J(D) - K(D) as PFock_getMat(F)."""

    def __init__(self, bases, tolscr=1e-11, tau=0.0, graph=True):
        L = lib()
        self.bases = list(bases)
        self.nmol = len(self.bases)
        arr = (_vp * self.nmol)(*[b.h for b in self.bases])
        self.h = _vp()
        st = L.gtf_batch_create(arr, self.nmol, tolscr, ctypes.byref(self.h))
        if st != 0:
            self.h = None
            raise GtfError("gtf_batch_create failed (%d)" % st)
        self.nbf = np.zeros(self.nmol, dtype=np.int32)
        self.offset = np.zeros(self.nmol, dtype=np.int64)
        L.gtf_batch_layout(self.h, self.nbf.ctypes.data_as(_ip), self.offset.ctypes.data_as(_lp))
        self.total = int(L.gtf_batch_total(self.h))
        if tau > 0:
            self.set_precision(tau)
        if not graph:
            L.gtf_batch_use_graph(self.h, 0)

    def close(self):
        if getattr(self, "h", None):
            lib().gtf_batch_destroy(self.h)
            self.h = None

    __del__ = close

    def _st(self, st, what):
        if st != 0:
            raise GtfError("%s failed (%d)" % (what, st))

    def set_precision(self, tau):
        self._st(lib().gtf_batch_set_precision(self.h, tau), "gtf_batch_set_precision")

    def set_prim_tol(self, tol):
        self._st(lib().gtf_batch_set_prim_tol(self.h, float(tol)), "gtf_batch_set_prim_tol")

    def use_graph(self, on):
        self._st(lib().gtf_batch_use_graph(self.h, 1 if on else 0), "gtf_batch_use_graph")

    def set_purif_precision(self, mode, until=None):
        """as Purif.set_precision, per molecule of the batch"""
        m = _purif_mode(mode)
        if until is None:
            until = PURIF_SPLIT_UNTIL if m == 2 else 1e-3
        self._st(lib().gtf_batch_set_purif_precision(self.h, m, until), "gtf_batch_set_purif_precision")

    def set_purif_slices(self, ns):
        self._st(lib().gtf_batch_set_purif_slices(self.h, int(ns)), "gtf_batch_set_purif_slices")

    def set_purif_residual(self, resid):
        """as Purif.set_split_residual, per molecule of the batch"""
        self._st(lib().gtf_batch_set_purif_residual(self.h, float(resid)), "gtf_batch_set_purif_residual")

    def autotune(self, reps=3):
        """Tile autotuner on the batch's density (gtf_batch_autotune)."""
        ms = np.zeros(3)
        self._st(lib().gtf_batch_autotune(self.h, int(reps), _p(ms)), "gtf_batch_autotune")
        return {"start_ms": float(ms[0]), "tuned_ms": float(ms[1]), "candidates": int(ms[2])}

    def tile_shape(self, cls):
        sh = np.zeros(len(SHAPE_KEYS), dtype=np.int32)
        st = lib().gtf_batch_tile_shape(self.h, cls, sh.ctypes.data_as(_ip))
        self._st(0 if st == len(SHAPE_KEYS) else st, "gtf_batch_tile_shape")
        return dict(zip(SHAPE_KEYS, (int(x) for x in sh)))

    def tile_shapes(self):
        return [self.tile_shape(c) for c in range(21)]

    def set_tile_shape(self, cls, shape):
        sh = np.array([shape[k] for k in SHAPE_KEYS], dtype=np.int32)
        self._st(lib().gtf_batch_set_tile_shape(self.h, cls, sh.ctypes.data_as(_ip)), "gtf_batch_set_tile_shape")

    def set_tile_shapes(self, shapes):
        sh = np.array(_flat_shapes(shapes), dtype=np.int32)
        self._st(lib().gtf_batch_set_tile_shapes(self.h, sh.ctypes.data_as(_ip)), "gtf_batch_set_tile_shapes")

    def pack(self, mats):
        out = np.zeros(self.total)
        for m, A in enumerate(mats):
            n = int(self.nbf[m])
            out[self.offset[m]:self.offset[m] + n * n] = np.ascontiguousarray(A, dtype=np.float64).reshape(-1)
        return out

    def unpack(self, packed):
        return [packed[self.offset[m]:self.offset[m] + int(self.nbf[m]) ** 2].reshape(int(self.nbf[m]),
                                                                                       int(self.nbf[m])).copy()
                for m in range(self.nmol)]

    def put_density(self, Ds):
        buf = self.pack(Ds)
        self._st(lib().gtf_batch_put_density(self.h, -1, _p(buf)), "gtf_batch_put_density")

    def compute(self):
        self._st(lib().gtf_batch_compute_fock(self.h), "gtf_batch_compute_fock")

    def get_fock(self):
        buf = np.zeros(self.total)
        self._st(lib().gtf_batch_get_fock(self.h, -1, _p(buf)), "gtf_batch_get_fock")
        return self.unpack(buf)

    def get_density(self):
        buf = np.zeros(self.total)
        self._st(lib().gtf_batch_get_density(self.h, -1, _p(buf)), "gtf_batch_get_density")
        return self.unpack(buf)

    def fock(self, Ds):
        self.put_density(Ds)
        self.compute()
        return self.get_fock()

    def stats(self):
        out = np.zeros(11)
        lib().gtf_batch_stats(self.h, _p(out))
        keys = ["nquartets", "nprimquartets", "nq_fp32", "npq_fp32", "gpu_ms", "graph", "device_bytes",
                "create_s", "scf_fock_s", "scf_diis_s", "scf_purif_s"]
        return dict(zip(keys, out.tolist()))

    def setup_times(self):
        """Wall seconds of the set-up phases: creation (as PFock's), then the
        SCF arrays with the one-electron matrices, and X = S^-1/2."""
        out = np.zeros(8)
        self._st(0 if lib().gtf_batch_setup_times(self.h, _p(out)) == 8 else -3, "gtf_batch_setup_times")
        return dict(zip(SETUP_PHASES + ("scf_oneel", "scf_xmat"), out.tolist()))

    def class_work(self):
        nq = np.zeros(21, dtype=np.int64)
        npq = np.zeros(21)
        nq32 = np.zeros(21, dtype=np.int64)
        npq32 = np.zeros(21)
        lib().gtf_batch_class_work(self.h, nq.ctypes.data_as(_lp), _p(npq), nq32.ctypes.data_as(_lp), _p(npq32))
        return nq, npq, nq32, npq32

    def scf_begin(self, max_iters=30, ediff=1e-11):
        """Step-wise SCF (gtf_batch_scf_begin); then scf_step() per iteration."""
        self._max_iters = int(max_iters)
        self._st(lib().gtf_batch_scf_begin(self.h, self._max_iters, ediff, None, None), "gtf_batch_scf_begin")

    def scf_step(self):
        """One SCF iteration of the running molecules; returns how many still run."""
        r = lib().gtf_batch_scf_step(self.h)
        if r < 0:
            raise GtfError("gtf_batch_scf_step failed (%d)" % r)
        return r

    def scf_results(self):
        """(energies, builds (0 while running), history[nmol, max_iters])"""
        mi = self._max_iters
        e = np.zeros(self.nmol)
        it = np.zeros(self.nmol, dtype=np.int32)
        hist = np.zeros(self.nmol * mi)
        lib().gtf_batch_scf_results(self.h, _p(e), it.ctypes.data_as(_ip), _p(hist))
        return e, it, hist.reshape(self.nmol, mi)

    def regroup(self, add=()):
        """Continuous batching (gtf_batch_regroup): a new Batch of this one's
        running molecules, SCF state carried over, plus the BasisSets `add`
        starting their SCF.  Returns (batch, origin); origin[k] is the index
        in this batch, or -(j + 1) for add[j].  This batch stays valid."""
        add = list(add)
        arr = (_vp * max(1, len(add)))(*[b.h for b in add])
        h = _vp()
        origin = np.zeros(self.nmol + len(add), dtype=np.int32)
        self._st(lib().gtf_batch_regroup(self.h, arr, len(add), ctypes.byref(h), origin.ctypes.data_as(_ip)),
                 "gtf_batch_regroup")
        nb = Batch.__new__(Batch)
        nb.h = h
        n = int(lib().gtf_batch_layout(h, None, None))
        origin = origin[:n]
        nb.bases = [self.bases[o] if o >= 0 else add[-o - 1] for o in origin]
        nb.nmol = n
        nb.nbf = np.zeros(n, dtype=np.int32)
        nb.offset = np.zeros(n, dtype=np.int64)
        lib().gtf_batch_layout(h, nb.nbf.ctypes.data_as(_ip), nb.offset.ctypes.data_as(_lp))
        nb.total = int(lib().gtf_batch_total(h))
        nb._max_iters = self._max_iters
        return nb, origin

    def scf(self, max_iters=30, ediff=1e-11, history=False):
        """Batched RHF SCF; returns (energies, builds used[, history])."""
        e = np.zeros(self.nmol)
        it = np.zeros(self.nmol, dtype=np.int32)
        hist = np.full(self.nmol * max_iters, np.nan)
        self._st(lib().gtf_batch_scf(self.h, max_iters, ediff, _p(e), it.ctypes.data_as(_ip), _p(hist)),
                 "gtf_batch_scf")
        if history:
            return e, it, hist.reshape(self.nmol, max_iters)
        return e, it


class Server:
    """SCF request server (gtf_server_*): jobs admitted by a device-memory
    budget into cohorts that run as batches, several cohorts at once."""

    JOB_KEYS = ["energy", "iters", "t_submit", "t_start", "t_done", "cohort", "mem_est"]

    def __init__(self, hbm_budget_gb=64.0, max_cohorts=4, max_jobs_per_cohort=64, tolscr=1e-11, tau=0.0,
                 max_iters=30, ediff=1e-8, shapes=None, purif="fp64", join_gap=1):
        self.h = _vp()
        st = lib().gtf_server_create(hbm_budget_gb * 2.0 ** 30, max_cohorts, max_jobs_per_cohort, tolscr, tau,
                                     max_iters, ediff, ctypes.byref(self.h))
        if st != 0:
            self.h = None
            raise GtfError("gtf_server_create failed (%d)" % st)
        self._keep = {}
        if shapes is not None:
            self.set_tile_shapes(shapes)
        if purif != "fp64":
            self.set_purif_precision(purif)
        if join_gap != 1:
            self.set_join_gap(join_gap)

    def set_purif_precision(self, mode, until=None):
        """Purification products of the cohorts started from now on, as
        Purif.set_precision ("fp64", "fp32", "split")."""
        m = _purif_mode(mode)
        if until is None:
            until = PURIF_SPLIT_UNTIL if m == 2 else 1e-3
        st = lib().gtf_server_set_purif_precision(self.h, m, until)
        if st != 0:
            raise GtfError("gtf_server_set_purif_precision failed (%d)" % st)

    def set_join_gap(self, iters):
        """SCF iterations between admissions into a running cohort (< 0: off)."""
        st = lib().gtf_server_set_join_gap(self.h, int(iters))
        if st != 0:
            raise GtfError("gtf_server_set_join_gap failed (%d)" % st)

    def set_large_yield(self, factor):
        """How long large-molecule cohorts defer an iteration while small ones
        run, in units of their last iteration (0: never)."""
        st = lib().gtf_server_set_large_yield(self.h, float(factor))
        if st != 0:
            raise GtfError("gtf_server_set_large_yield failed (%d)" % st)

    def join_stats(self):
        """(jobs admitted into running cohorts, regroups)"""
        j, r = ctypes.c_long(), ctypes.c_long()
        lib().gtf_server_join_stats(self.h, ctypes.byref(j), ctypes.byref(r))
        return j.value, r.value

    def set_tile_shapes(self, shapes):
        """Tile shapes of the cohorts started from now on (None: defaults)."""
        if shapes is None:
            st = lib().gtf_server_set_tile_shapes(self.h, None)
        else:
            sh = np.array(_flat_shapes(shapes), dtype=np.int32)
            st = lib().gtf_server_set_tile_shapes(self.h, sh.ctypes.data_as(_ip))
        if st != 0:
            raise GtfError("gtf_server_set_tile_shapes failed (%d)" % st)

    def set_large_slots(self, n):
        """Cohorts of molecules with 384 or more basis functions run at once."""
        st = lib().gtf_server_set_large_slots(self.h, int(n))
        if st != 0:
            raise GtfError("gtf_server_set_large_slots failed (%d)" % st)

    def set_tile_shapes_range(self, nbf_min, nbf_max, shapes):
        """Shapes for cohorts whose largest molecule has nbf_min..nbf_max basis
        functions (the last matching range wins)."""
        sh = np.array(_flat_shapes(shapes), dtype=np.int32)
        st = lib().gtf_server_set_tile_shapes_range(self.h, int(nbf_min), int(nbf_max), sh.ctypes.data_as(_ip))
        if st != 0:
            raise GtfError("gtf_server_set_tile_shapes_range failed (%d)" % st)

    def submit(self, basis):
        jid = int(lib().gtf_server_submit(self.h, basis.h))
        if jid < 0:
            raise GtfError("gtf_server_submit failed (%d)" % jid)
        self._keep[jid] = basis  # the basis must outlive the job
        return jid

    def poll(self, jid):
        out = np.zeros(7)
        st = lib().gtf_server_poll(self.h, jid, _p(out))
        d = dict(zip(self.JOB_KEYS, out.tolist()))
        d["status"] = st
        if st == 2 or st < 0:
            self._keep.pop(jid, None)  # the server no longer reads the basis
        return d

    def wait(self):
        lib().gtf_server_wait(self.h)

    def stats(self):
        out = np.zeros(8)
        lib().gtf_server_stats(self.h, _p(out))
        return dict(zip(["jobs", "queued", "running", "cohorts", "peak_running", "bytes_in_use", "peak_bytes",
                         "budget"], out.tolist()))

    def close(self):
        if getattr(self, "h", None):
            lib().gtf_server_destroy(self.h)
            self.h = None

    __del__ = close


class PurifT(ctypes.Structure):
    """Field layout of purif_t (include/purif.h)."""
    _fields_ = [(n, ctypes.c_int) for n in (
        "runpurif", "rundgemm", "nbf", "nobtls", "np_purif", "nprow_purif", "npcol_purif", "npgrd_purif",
        "srow_purif", "scol_purif", "nrows_purif", "ncols_purif", "ldx", "meshsize")] + [
        ("H_block", _dp), ("X_block", _dp), ("S_block", _dp), ("F_block", _dp), ("D_block", _dp),
        ("len_diis", ctypes.c_int), ("bmax_id", ctypes.c_int), ("bmax", ctypes.c_double),
        ("b_mat", ctypes.c_double * 121),
        ("timedgemm", ctypes.c_double), ("timepdgemm", ctypes.c_double), ("timetr", ctypes.c_double),
        ("timepass", ctypes.c_double), ("dev", _vp)]


def initial_guess(basis):
    """The starting density of pscf/scf.c:34-78: the atomic blocks of
    CInt_getInitialGuess (read from <Element>.dat next to the basis file;
    zero where there is none) on the block diagonal, times R/2 with
    R = (N_neutral - Q) / N_neutral for a molecule of charge Q."""
    L = lib()
    n = basis.nbf
    D = np.zeros((n, n))
    for a in range(basis.natoms):
        g = ctypes.POINTER(ctypes.c_double)()
        s, e = ctypes.c_int(), ctypes.c_int()
        L.CInt_getInitialGuess(basis.h, a, ctypes.byref(g), ctypes.byref(s), ctypes.byref(e))
        m = e.value - s.value + 1
        D[s.value:e.value + 1, s.value:e.value + 1] = np.ctypeslib.as_array(g, shape=(m * m,)).reshape(m, m)
    nn, q = L.CInt_getNneutral(basis.h), L.CInt_getTotalCharge(basis.h)
    R = (nn - q) / float(nn) if (q != 0 and nn != 0) else 1.0
    return D * (R / 2.0)


def scf(gbs, xyz, niters, tolscr=1e-11, verbose=False, tau=0.0, purif_fp32=False, guess=False, purif=None,
        ediff=1e-11, stats=None, shapes=None, purif_ns=None, purif_resid=None):
    """pscf/scf.c:189-474 on one GPU.  The initial density is zero
    (core-Hamiltonian guess), or with guess=True the atomic-block guess of
    scf.c's initial_guess (zero for atoms without a .dat file).  tau > 0
    selects the mixed-precision Fock build; purif ("fp64", "fp32", "split";
    purif_fp32=True is "fp32") the purification products, purif_ns and
    purif_resid the split products' slices and DIIS-residual switch (None:
    the library defaults).  The SCF stops
    after the build whose energy differs from the previous one by less than
    ediff (scf.c: 1e-11).  stats (a dict, may be None) receives the
    purification iteration counts and the seconds spent in Fock builds and
    in purifications; shapes: tile shapes of the Fock builds (None: the
    built-in defaults).  Returns (energies, last D, last F = H + G(D))."""
    import time
    basis = BasisSet(gbs, xyz)
    pf = PFock(basis, tolscr=tolscr)
    if tau > 0:
        pf.set_precision(tau)
    if shapes is not None:
        pf.set_tile_shapes(shapes)
    pu = Purif(basis)
    if purif is None:
        purif = "fp32" if purif_fp32 else "fp64"
    if purif != "fp64":
        pu.set_precision(purif)
    if purif_ns is not None:
        pu.set_slices(purif_ns)
    if purif_resid is not None:
        pu.set_split_residual(purif_resid)
    S = pf.overlap()
    X = pf.x_matrix()
    H = pf.core_hamiltonian()
    pu.set_xs(X, S)
    n = basis.nbf
    D = initial_guess(basis) if guess else np.zeros((n, n))
    enuc = basis.nuc_energy()
    e0 = -1.0
    energies = []
    F = None
    Dlast = D
    for it in range(niters):
        t0 = time.perf_counter()
        G = pf.fock(D)
        if stats is not None:
            stats["fock_s"] = stats.get("fock_s", 0.0) + time.perf_counter() - t0
        F = G + H
        e = float(np.sum(D * (H + F)))
        energies.append(e + enuc)
        Dlast = D
        if verbose:
            print("  iter %d energy %.10f" % (it, e + enuc), flush=True)
        if it > 0 and abs(e - e0) < ediff:
            break
        e0 = e
        Fo = pu.diis(pf, D, F, it)
        t0 = time.perf_counter()
        D, pit = pu.purify(Fo)
        if stats is not None:
            stats["purif_s"] = stats.get("purif_s", 0.0) + time.perf_counter() - t0
            stats.setdefault("purif_iters", []).append(pit)
            stats.setdefault("purif_low_iters", []).append(pu.fp32_iters())
    pu.close()
    pf.close()
    return np.array(energies), Dlast, F
