"""ctypes binding of the CPU oracle (oracle/gtf_oracle.c).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by __graft_entry__.smoke() and
by bench.py's cpu_baseline leg, always as the checker or the timed CPU
baseline, never as part of the product path.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build():
    """Compile the oracle with its Makefile (gcc + OpenMP)."""
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.ob_load.restype = ctypes.c_void_p
        L.ob_load.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
        L.ob_free.argtypes = [ctypes.c_void_p]
        for name in ("ob_nshells", "ob_nbf", "ob_natoms", "ob_nocc", "ob_max_shell_dim"):
            getattr(L, name).argtypes = [ctypes.c_void_p]
            getattr(L, name).restype = ctypes.c_int
        for name in ("ob_shell_L", "ob_shell_nprim", "ob_shell_fstart", "ob_shell_atom", "ob_shell_dim"):
            getattr(L, name).argtypes = [ctypes.c_void_p, ctypes.c_int]
            getattr(L, name).restype = ctypes.c_int
        L.ob_nuc_energy.argtypes = [ctypes.c_void_p]
        L.ob_nuc_energy.restype = ctypes.c_double
        L.ob_atom_xyz.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
        L.ob_boys.argtypes = [ctypes.c_int, ctypes.c_double, _dp]
        L.ob_eri_quartet.argtypes = [ctypes.c_void_p] + [ctypes.c_int] * 4 + [_dp]
        L.ob_eri_quartet.restype = ctypes.c_int
        for name in ("ob_overlap", "ob_kinetic", "ob_potential", "ob_core_hamiltonian"):
            getattr(L, name).argtypes = [ctypes.c_void_p, _dp]
        L.ob_screening.argtypes = [ctypes.c_void_p, ctypes.c_double, _dp, _dp, _ip, _ip, _ip, _dp]
        L.ob_screening.restype = ctypes.c_int
        L.ob_fock_build.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int, _dp, _dp, _dp, _dp]
        L.ob_enumerate_quartets.argtypes = [ctypes.c_void_p, ctypes.c_double, _ip, ctypes.c_long]
        L.ob_enumerate_quartets.restype = ctypes.c_long
        L.ob_gemm_abt.argtypes = [ctypes.c_int, _dp, _dp, _dp]
        L.ob_inv_sqrt.argtypes = [ctypes.c_int, _dp, _dp]
        L.ob_purification.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp]
        L.ob_purification.restype = ctypes.c_int
        L.ob_diis_create.argtypes = [ctypes.c_int]
        L.ob_diis_create.restype = ctypes.c_void_p
        L.ob_diis_free.argtypes = [ctypes.c_void_p]
        L.ob_diis_step.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp, ctypes.c_int]
        L.ob_scf.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int, _dp, _dp, _dp]
        L.ob_scf.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp)


def _ipp(a):
    return a.ctypes.data_as(_ip)


def boys(mmax, T):
    out = np.zeros(mmax + 1)
    lib().ob_boys(mmax, float(T), _p(out))
    return out


class Basis:
    """A basis set on a molecule, as the oracle sees it."""

    def __init__(self, gbs, xyz):
        self.h = lib().ob_load(gbs.encode(), xyz.encode())
        if not self.h:
            raise RuntimeError("oracle could not load %s / %s" % (gbs, xyz))
        L = lib()
        self.nshells = L.ob_nshells(self.h)
        self.nbf = L.ob_nbf(self.h)
        self.natoms = L.ob_natoms(self.h)
        self.nocc = L.ob_nocc(self.h)
        self.shell_L = [L.ob_shell_L(self.h, s) for s in range(self.nshells)]
        self.shell_nprim = [L.ob_shell_nprim(self.h, s) for s in range(self.nshells)]
        self.shell_fstart = [L.ob_shell_fstart(self.h, s) for s in range(self.nshells)] + [self.nbf]
        self.shell_atom = [L.ob_shell_atom(self.h, s) for s in range(self.nshells)]

    def __del__(self):
        if getattr(self, "h", None):
            lib().ob_free(self.h)
            self.h = None

    def nuc_energy(self):
        return lib().ob_nuc_energy(self.h)

    def atom_xyz(self, a):
        out = np.zeros(3)
        lib().ob_atom_xyz(self.h, a, _p(out))
        return out

    def eri_quartet(self, M, N, P, Q):
        """Integrals in the reference layout, reshaped to [iQ, iP, iN, iM]."""
        d = [2 * self.shell_L[s] + 1 for s in (M, N, P, Q)]
        out = np.zeros(d[0] * d[1] * d[2] * d[3])
        lib().ob_eri_quartet(self.h, M, N, P, Q, _p(out))
        return out.reshape(d[3], d[2], d[1], d[0])

    def eri_tensor(self):
        """Full (mn|pq) tensor, chemists' order, for small molecules."""
        n = self.nbf
        g = np.zeros((n, n, n, n))
        fs = self.shell_fstart
        for M in range(self.nshells):
            for N in range(self.nshells):
                for P in range(self.nshells):
                    for Q in range(self.nshells):
                        blk = self.eri_quartet(M, N, P, Q)  # [q,p,n,m]
                        g[fs[M]:fs[M + 1], fs[N]:fs[N + 1], fs[P]:fs[P + 1], fs[Q]:fs[Q + 1]] = \
                            blk.transpose(3, 2, 1, 0)
        return g

    def _mat(self, fn):
        out = np.zeros((self.nbf, self.nbf))
        getattr(lib(), fn)(self.h, _p(out))
        return out

    def overlap(self):
        return self._mat("ob_overlap")

    def kinetic(self):
        return self._mat("ob_kinetic")

    def potential(self):
        return self._mat("ob_potential")

    def core_hamiltonian(self):
        return self._mat("ob_core_hamiltonian")

    def screening(self, tol):
        ns = self.nshells
        v = np.zeros(ns * ns)
        vmax = np.zeros(1)
        ptr = np.zeros(ns + 1, dtype=np.int32)
        sid = np.zeros(ns * ns, dtype=np.int32)
        srid = np.zeros(ns * ns, dtype=np.int32)
        sval = np.zeros(ns * ns)
        nnz = lib().ob_screening(self.h, tol, _p(v), _p(vmax), _ipp(ptr), _ipp(sid), _ipp(srid), _p(sval))
        return dict(nnz=nnz, vpair=v.reshape(ns, ns), vmax=vmax[0], shellptr=ptr,
                    shellid=sid[:nnz].copy(), shellrid=srid[:nnz].copy(), shellvalue=sval[:nnz].copy())

    def fock(self, D, tol=1e-11, nblocks=4):
        """(Jacc, Kacc, stats): symmetrised 2J(D) and -K(D) accumulators."""
        n = self.nbf
        D = np.ascontiguousarray(D, dtype=np.float64)
        J = np.zeros((n, n))
        K = np.zeros((n, n))
        st = np.zeros(5)
        lib().ob_fock_build(self.h, tol, nblocks, _p(D), _p(J), _p(K), _p(st))
        return J, K, dict(usq=st[0], uitl=st[1], nnz=st[2], screen_s=st[3], build_s=st[4])

    def quartets(self, tol=1e-11):
        n = lib().ob_enumerate_quartets(self.h, tol, None, 0)
        q = np.zeros((max(n, 1), 4), dtype=np.int32)
        lib().ob_enumerate_quartets(self.h, tol, _ipp(q), n)
        return q[:n]

    def scf(self, niters, tol=1e-11):
        n = self.nbf
        e = np.zeros(niters)
        D = np.zeros((n, n))
        F = np.zeros((n, n))
        k = lib().ob_scf(self.h, tol, niters, _p(e), _p(D), _p(F))
        return e[:k], D, F


def gemm_abt(A, B):
    n = A.shape[0]
    A = np.ascontiguousarray(A, dtype=np.float64)
    B = np.ascontiguousarray(B, dtype=np.float64)
    C = np.zeros((n, n))
    lib().ob_gemm_abt(n, _p(A), _p(B), _p(C))
    return C


def inv_sqrt(S):
    n = S.shape[0]
    S = np.ascontiguousarray(S, dtype=np.float64)
    X = np.zeros((n, n))
    lib().ob_inv_sqrt(n, _p(S), _p(X))
    return X


def purification(F, X, nocc):
    n = F.shape[0]
    F = np.ascontiguousarray(F, dtype=np.float64)
    X = np.ascontiguousarray(X, dtype=np.float64)
    D = np.zeros((n, n))
    it = lib().ob_purification(n, nocc, _p(F), _p(X), _p(D))
    return D, it


class Diis:
    def __init__(self, n):
        self.n = n
        self.h = lib().ob_diis_create(n)

    def __del__(self):
        if getattr(self, "h", None):
            lib().ob_diis_free(self.h)
            self.h = None

    def step(self, X, S, D, F, it):
        """Returns the orthogonalised / extrapolated Fock matrix."""
        F = np.array(F, dtype=np.float64, order="C")
        X = np.ascontiguousarray(X, dtype=np.float64)
        S = np.ascontiguousarray(S, dtype=np.float64)
        D = np.ascontiguousarray(D, dtype=np.float64)
        lib().ob_diis_step(self.h, _p(X), _p(S), _p(D), _p(F), it)
        return F
