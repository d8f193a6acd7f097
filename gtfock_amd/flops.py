"""Algorithmic FLOP counts of the Fock build, per quartet class.

The count is an operation count of the reference integral scheme the CPU
oracle runs -- Obara-Saika vertical recursion on (e0|f0)^(m) per primitive
quartet, contraction, Head-Gordon-Pople horizontal recursion and the
spherical transform per contracted quartet -- made once by walking the same
loops and counting each +, -, *, / (and each exp / sqrt / Boys evaluation as
a fixed cost).  It is fixed by the class and independent of how the GPU
computes the integrals, so "achieved FLOP/s" compares implementations on the
same work.  Digestion adds 12 FLOPs per integral per density (the six
multiply-adds of update_F, pfock/fock_task.c:79-113).

  flops(class) = C_prim(class) * primitive quartets + C_contr(class) * quartets
"""
from functools import lru_cache

PAIR_L = [(0, 0), (1, 0), (1, 1), (2, 0), (2, 1), (2, 2)]


def _ncu(l):
    return 0 if l < 0 else (l + 1) * (l + 2) * (l + 3) // 6


def _nc(l):
    return (l + 1) * (l + 2) // 2


def _carts(lmax):
    out = []
    for l in range(lmax + 1):
        for lx in range(l, -1, -1):
            for lz in range(0, l - lx + 1):
                out.append((lx, l - lx - lz, lz))
    return out


def _first_dir(c):
    return 0 if c[0] else (1 if c[1] else 2)


def _sph_nnz(l):
    return {0: [1], 1: [1, 1, 1], 2: [1, 1, 3, 1, 2]}[l]


@lru_cache(maxsize=None)
def class_cost(la, lb, lc, ld):
    """(C_prim, C_contr) for the ordered class (la lb | lc ld)."""
    lab, lcd = la + lb, lc + ld
    L = lab + lcd
    E = _carts(lab)
    Fc = _carts(lcd)
    # per primitive quartet: geometry and prefactors (W, WP, WQ, QC, |PQ|^2,
    # rho, prefactor: 60), the Boys function (15 + 3L), pref * F_m (L + 1)
    prim = 60 + 15 + 3 * L + (L + 1)
    for e in E[1:]:
        d = _first_dir(e)
        nm = L - sum(e) + 1
        prim += nm * (3 + (5 if e[d] > 1 else 0))
    for f in Fc[1:]:
        d = _first_dir(f)
        for e in E:
            nm = L - sum(e) - sum(f) + 1
            if nm <= 0:
                continue
            prim += nm * (3 + (5 if f[d] > 1 else 0) + (3 if e[d] > 0 else 0))
    ne_used = _ncu(lab) - _ncu(la - 1)
    nf_used = _ncu(lcd) - _ncu(lc - 1)
    prim += ne_used * nf_used  # contraction adds
    # per contracted quartet: bra HRR over the used ket components, ket HRR
    # over the bra cartesian pairs, then the spherical transform
    contr = 0
    for l2 in range(1, lb + 1):
        for _ in range(_nc(l2)):
            contr += 2 * (_ncu(lab - l2) - _ncu(la - 1)) * nf_used
    for l2 in range(1, ld + 1):
        for _ in range(_nc(l2)):
            contr += 2 * (_ncu(lcd - l2) - _ncu(lc - 1)) * _nc(la) * _nc(lb)
    nnz = 1
    for l in (la, lb, lc, ld):
        nnz *= sum(_sph_nnz(l))
    contr += 2 * nnz
    return prim, contr


def class_index_types(cls):
    A = 0
    while (A + 1) * (A + 2) // 2 <= cls:
        A += 1
    return A, cls - A * (A + 1) // 2


def class_L(cls):
    A, B = class_index_types(cls)
    return PAIR_L[A] + PAIR_L[B]


def nints(cls):
    la, lb, lc, ld = class_L(cls)
    return (2 * la + 1) * (2 * lb + 1) * (2 * lc + 1) * (2 * ld + 1)


def class_flops(cls, nquartets, nprimquartets, ndens=1):
    """Algorithmic FLOPs of one launch of class cls (ERIs + digestion)."""
    cp, cc = class_cost(*class_L(cls))
    return cp * nprimquartets + cc * nquartets + 12.0 * nints(cls) * nquartets * ndens


def digestion_bytes(cls, nquartets, ndens=1):
    """Bytes per launch of the fused digestion: per quartet and density, the
    six D blocks are read and the six F blocks read and written, 8 bytes per
    element: 8 * 3 * (dMdN + dPdQ + dMdP + dNdQ + dMdQ + dNdP)."""
    la, lb, lc, ld = class_L(cls)
    dM, dN, dP, dQ = (2 * l + 1 for l in (la, lb, lc, ld))
    per = 8 * 3 * (dM * dN + dP * dQ + dM * dP + dN * dQ + dM * dQ + dN * dP)
    return per * nquartets * ndens
