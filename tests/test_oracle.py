"""Pins of the CPU oracle (CPU only).

The reference ships no golden vectors or known-answer tests for this path
(its integrals are closed OptErd archives), so the oracle is pinned by:
  * Boys F_m(T) against mpmath 1F1 at 40 digits (tests/golden/boys.npz);
  * ERIs against an independent McMurchie-Davidson implementation with
    mpmath Boys values (tests/golden/eri_water.npz, every ordered L class);
  * S, T, V against the same independent code (tests/golden/oneel_water.npz);
  * the Szabo-Ostlund H2/STO-3G tables (tests/golden/h2_sto3g.json);
  * brute-force J/K and Schwarz values from the full ERI tensor (einsum);
  * closed forms and invariants: (ss|ss) formula, 8-fold symmetry, Schwarz
    inequality, rotation/translation invariance, eigh projector, Pulay DIIS.
Each check is chosen so that a dropped term, a wrong sign or a transposed
index fails it.
"""
import json
import math
import os

import numpy as np
import pytest

import oracle
from gtfock_amd import geometry

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


@pytest.fixture(scope="module")
def water(gbs, mol_dir):
    return oracle.Basis(gbs, str(mol_dir / "water.xyz"))


def test_boys_against_mpmath():
    g = np.load(os.path.join(GOLD, "boys.npz"))
    for i, T in enumerate(g["T"]):
        F = oracle.boys(16, T)
        np.testing.assert_allclose(F, g["F"][i], rtol=3e-15, atol=0)


def test_boys_special_values():
    F = oracle.boys(8, 0.0)
    np.testing.assert_allclose(F, [1.0 / (2 * m + 1) for m in range(9)], rtol=1e-15)
    T = 2.5
    assert abs(oracle.boys(0, T)[0] - 0.5 * math.sqrt(math.pi / T) * math.erf(math.sqrt(T))) < 1e-15


def test_eri_against_independent_md(water):
    g = np.load(os.path.join(GOLD, "eri_water.npz"))
    q, off, val = g["quartets"], g["offsets"], g["values"]
    Ls = set()
    for i, (M, N, P, Q) in enumerate(q):
        mine = water.eri_quartet(M, N, P, Q).reshape(-1)
        np.testing.assert_allclose(mine, val[off[i]:off[i + 1]], rtol=0, atol=1e-13)
        Ls.add(tuple(water.shell_L[s] for s in (M, N, P, Q)))
    assert len(Ls) == 81  # every ordered (La, Lb, Lc, Ld) with L <= 2


def test_one_electron_against_independent_md(water):
    g = np.load(os.path.join(GOLD, "oneel_water.npz"))
    np.testing.assert_allclose(water.overlap(), g["S"], atol=1e-13)
    np.testing.assert_allclose(water.kinetic(), g["T"], atol=1e-13)
    np.testing.assert_allclose(water.potential(), g["V"], atol=1e-12)


def test_h2_sto3g_textbook(sto3g, mol_dir):
    ref = json.load(open(os.path.join(GOLD, "h2_sto3g.json")))
    tol = ref["tolerance"]
    b = oracle.Basis(sto3g, str(mol_dir / "h2.xyz"))
    S, T, V = b.overlap(), b.kinetic(), b.potential()
    assert abs(S[0, 1] - ref["S12"]) < tol
    assert abs(T[0, 0] - ref["T11"]) < tol and abs(T[0, 1] - ref["T12"]) < tol
    # V_11 = V_11^(1) + V_11^(2) = V_11^(1) + V_22^(1) by symmetry; V_12 = 2 V_12^(1)
    assert abs(V[0, 0] - (ref["V11_1"] + ref["V22_1"])) < tol
    assert abs(V[0, 1] - 2 * ref["V12_1"]) < tol
    H = T + V
    assert abs(H[0, 0] - ref["H11"]) < tol and abs(H[0, 1] - ref["H12"]) < tol
    assert abs(b.eri_quartet(0, 0, 0, 0).item() - ref["eri_1111"]) < tol
    assert abs(b.eri_quartet(0, 0, 1, 1).item() - ref["eri_1122"]) < tol
    assert abs(b.eri_quartet(1, 0, 0, 0).item() - ref["eri_2111"]) < tol
    assert abs(b.eri_quartet(1, 0, 1, 0).item() - ref["eri_2121"]) < tol
    e, _, _ = b.scf(10)
    assert abs(e[-1] - ref["E_total"]) < tol


def test_ssss_closed_form(mol_dir, tmp_path):
    """(ss|ss) of single normalised s primitives:
    2 pi^(5/2) / (p q sqrt(p+q)) exp(-ab/p AB^2 - cd/q CD^2) F0(pq/(p+q) PQ^2) N_a N_b N_c N_d."""
    gbs = tmp_path / "s.gbs"
    gbs.write_text("spherical\n\n****\nH     0\nS   1   1.00\n  0.8  1.0\n****\n"
                   "C     0\nS   1   1.00\n  1.7  1.0\n****\n")
    xyz = tmp_path / "hc.xyz"
    xyz.write_text("2\n\nH 0.1 -0.2 0.3\nC 0.9 0.5 -0.4\n")
    b = oracle.Basis(str(gbs), str(xyz))
    A, B = b.atom_xyz(0), b.atom_xyz(1)
    a, c = 0.8, 1.7
    N = lambda x: (2 * x / math.pi) ** 0.75
    p, q = 2 * a, a + c
    P, Q = A, (a * A + c * B) / q
    val = (2 * math.pi ** 2.5 / (p * q * math.sqrt(p + q)) * math.exp(-a * c / q * np.sum((A - B) ** 2))
           * oracle.boys(0, p * q / (p + q) * np.sum((P - Q) ** 2))[0] * N(a) ** 3 * N(c))
    assert abs(b.eri_quartet(0, 0, 0, 1).item() - val) < 1e-14


def test_eri_permutational_symmetry_and_schwarz(water):
    g = water.eri_tensor()
    for perm in [(1, 0, 2, 3), (0, 1, 3, 2), (2, 3, 0, 1), (3, 2, 1, 0)]:
        assert np.abs(g - g.transpose(perm)).max() < 1e-13
    n = water.nbf
    diag = np.einsum("ijij->ij", g)
    lhs = g.reshape(n * n, n * n) ** 2
    rhs = np.outer(diag.reshape(-1), diag.reshape(-1))
    assert np.all(lhs <= rhs * (1 + 1e-12) + 1e-14)
    # the pair Gram matrix (mn|pq) is positive semidefinite
    assert np.linalg.eigvalsh(g.reshape(n * n, n * n)).min() > -1e-10


def test_jk_digestion_equals_bruteforce(water):
    """update_F multiplicities + symmetrisation give 2J - K of the full tensor."""
    g = water.eri_tensor()
    rng = np.random.default_rng(5)
    A = rng.normal(size=(water.nbf, water.nbf))
    D = A + A.T
    Jacc, Kacc, st = water.fock(D, tol=1e-11)
    J = np.einsum("mnpq,pq->mn", g, D)
    K = np.einsum("mpnq,pq->mn", g, D)
    assert np.abs(Jacc - 2 * J).max() < 1e-12
    assert np.abs(Kacc + K).max() < 1e-12


def test_quartet_count_without_screening(water):
    """With tol = 0 every shell pair survives and the checkerboard enumeration
    yields each unordered pair of pairs once: npairs (npairs + 1) / 2."""
    ns = water.nshells
    npairs = ns * (ns + 1) // 2
    q = water.quartets(tol=0.0)
    assert len(q) == npairs * (npairs + 1) // 2
    canon = {tuple(sorted([tuple(sorted(x[:2])), tuple(sorted(x[2:]))])) for x in q.tolist()}
    assert len(canon) == len(q)


def test_screening_values_bruteforce(water):
    g = water.eri_tensor()
    s = water.screening(1e-11)
    fs = water.shell_fstart
    ns = water.nshells
    for M in range(ns):
        for N in range(ns):
            blk = g[fs[M]:fs[M + 1], fs[N]:fs[N + 1], fs[M]:fs[M + 1], fs[N]:fs[N + 1]]
            v = np.abs(np.einsum("ijij->ij", blk)).max()
            assert abs(s["vpair"][M, N] - v) < 1e-14
    assert s["vmax"] == s["vpair"].max()
    # checkerboard rule: exactly one orientation of every significant pair
    kept = set(zip(s["shellrid"].tolist(), s["shellid"].tolist()))
    for (A, B) in kept:
        if A != B:
            assert (B, A) not in kept


def test_purification_is_projector(water):
    S, H = water.overlap(), water.core_hamiltonian()
    w, U = np.linalg.eigh(S)
    X = U @ np.diag(w ** -0.5) @ U.T
    np.testing.assert_allclose(oracle.inv_sqrt(S), X, atol=1e-10)
    Fo = X @ H @ X
    D, it = oracle.purification(Fo, X, water.nocc)
    e, C = np.linalg.eigh(Fo)
    P = C[:, :water.nocc] @ C[:, :water.nocc].T
    assert np.abs(D - X @ P @ X).max() < 1e-9
    assert 0 < it < 200


def test_gemm_abt_convention():
    rng = np.random.default_rng(2)
    A, B = rng.normal(size=(7, 7)), rng.normal(size=(7, 7))
    assert np.abs(oracle.gemm_abt(A, B) - A @ B.T).max() < 1e-13


def test_diis_pulay_system():
    """The extrapolation coefficients solve min |sum c_i e_i| s.t. sum c_i = 1
    (Pulay 1980): with the oracle's stored error vectors the Lagrange
    solution from numpy must reproduce the extrapolated F."""
    n = 6
    rng = np.random.default_rng(9)
    X = np.eye(n)
    S = np.eye(n)
    d = oracle.Diis(n)
    errs, Fs = [], []
    for it in range(5):
        A = rng.normal(size=(n, n))
        F = A + A.T
        B = rng.normal(size=(n, n))
        D = B + B.T
        out = d.step(X, S, D, F, it)
        if it > 1:
            E = F @ D @ S
            E = E - E.T
            errs.append(-(X @ E @ X))
            Fs.append(X @ F @ X)
            m = len(errs)
            Bm = -np.ones((m + 1, m + 1))
            Bm[m, m] = 0
            for i in range(m):
                for j in range(m):
                    Bm[i, j] = np.sum(errs[i] * errs[j])
            rhs = np.zeros(m + 1)
            rhs[m] = -1
            c = np.linalg.solve(Bm, rhs)
            assert abs(c[:m].sum() - 1) < 1e-12
            ref = sum(ci * Fi for ci, Fi in zip(c[:m], Fs))
            assert np.abs(out - ref).max() < 1e-10
        else:
            assert np.abs(out - X @ F @ X).max() < 1e-12


def test_scf_energy_rigid_motion_invariant(gbs, tmp_path):
    atoms = geometry.water()
    rng = np.random.default_rng(4)
    R = geometry.random_rotation(rng)
    moved = [(s, tuple(R @ np.array(p) + np.array([1.3, -0.7, 2.1]))) for s, p in atoms]
    e = []
    for k, mol in enumerate([atoms, moved]):
        f = tmp_path / ("w%d.xyz" % k)
        f.write_text(geometry.to_xyz(mol))
        e.append(oracle.Basis(gbs, str(f)).scf(12)[0][-1])
    assert abs(e[0] - e[1]) < 1e-9
    assert abs(e[0] - (-76.0267987)) < 1e-6  # RHF/cc-pVDZ at this geometry
