"""Atomic density guesses, the <Element>.dat files that CInt_getInitialGuess
reads from beside a basis file and pscf/scf.c:34-78 places on the block
diagonal of the starting density.

Each block is the total density D (nfa x nfa, tr(D S) = Z) of the free
neutral atom from a restricted SCF with fractional, spherically averaged
occupations: electrons fill the orbitals by energy, two per orbital, shared
equally inside a degenerate set (carbon: 1s2 2s2 and 2/3 in each 2p).  The
Fock builds, overlap, core Hamiltonian and S^-1/2 come from this library on
the GPU.
"""
import os
import tempfile

import numpy as np

ELEMENTS = {"H": 1, "C": 6, "N": 7, "O": 8}


def occupations(eps, nelec, tol=1e-6):
    """Aufbau over orbital energies eps (ascending), equal shares inside a
    degenerate set (energies within tol)."""
    occ = np.zeros(len(eps))
    left = float(nelec)
    i = 0
    while left > 1e-12 and i < len(eps):
        j = i
        while j + 1 < len(eps) and abs(eps[j + 1] - eps[i]) < tol:
            j += 1
        put = min(2.0 * (j - i + 1), left)
        occ[i:j + 1] = put / (j - i + 1)
        left -= put
        i = j + 1
    return occ


def atomic_density(gbs, element, max_iter=300, tol=1e-11):
    """Total density of the free atom in basis file gbs; returns (D, builds)."""
    import gtfock_amd as g
    z = ELEMENTS[element]
    with tempfile.TemporaryDirectory() as d:
        xyz = os.path.join(d, "atom.xyz")
        with open(xyz, "w") as fh:
            fh.write("1\n\n%s 0.0 0.0 0.0\n" % element)
        b = g.BasisSet(gbs, xyz)
        pf = g.PFock(b)
        H, X = pf.core_hamiltonian(), pf.x_matrix()
        half = np.zeros((b.nbf, b.nbf))  # D/2, the closed-shell convention of the Fock build
        it = 0
        for it in range(1, max_iter + 1):
            F = H + pf.fock(half)
            eps, V = np.linalg.eigh(X @ F @ X)
            C = X @ V
            new = (C * (occupations(eps, z) / 2.0)) @ C.T
            step = float(np.abs(new - half).max())
            # damped: fractional occupations can oscillate between builds
            half = new if it == 1 else 0.5 * (half + new)
            if step < tol:
                break
        pf.close()
    return 2.0 * half, it


def write_dat(gbs, outdir, elements=("H", "C", "N", "O")):
    """<Element>.dat for each element, in outdir; returns the paths."""
    os.makedirs(outdir, exist_ok=True)
    paths = []
    for el in elements:
        D, _ = atomic_density(gbs, el)
        p = os.path.join(outdir, el + ".dat")
        np.savetxt(p, D.reshape(-1), fmt="%.16e")
        paths.append(p)
    return paths
