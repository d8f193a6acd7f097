"""GPU parity of the hot path against the CPU oracle, through the C-ABI.

Bars (from the task statement): Fock elements within 1e-10 Eh max-abs of
the FP64 oracle; screening CSR and quartet sets identical (index work is
bit-exact); ERIs within 1e-12 absolute (two independent algorithms:
McMurchie-Davidson on the GPU, Obara-Saika/Head-Gordon-Pople in the oracle).
"""
import numpy as np
import pytest

import oracle

gpu = pytest.mark.gpu
FOCK_TOL = 1e-10


def _rand_sym(n, seed, scale=0.1):
    rng = np.random.default_rng(seed)
    A = rng.normal(size=(n, n)) * scale
    return A + A.T


@pytest.fixture(scope="module")
def g():
    import gtfock_amd
    gtfock_amd.lib()
    return gtfock_amd


@gpu
def test_eri_all_water_quartets(g, gbs, mol_dir):
    xyz = str(mol_dir / "water.xyz")
    b = g.BasisSet(gbs, xyz)
    ob = oracle.Basis(gbs, xyz)
    erd = g.ERD(b)
    ns = b.nshells
    q = np.array([(M, N, P, Q) for M in range(ns) for N in range(ns) for P in range(ns) for Q in range(ns)],
                 dtype=np.int32)
    res = erd.quartets(q)
    worst = 0.0
    for (M, N, P, Q), r in zip(q, res):
        ref = ob.eri_quartet(M, N, P, Q).reshape(-1)
        worst = max(worst, float(np.abs(r - ref).max()))
    print("water: %d quartets, max |gpu - oracle| = %.3e" % (len(q), worst))
    assert worst < 1e-12


@gpu
def test_eri_single_quartet_api(g, gbs, mol_dir):
    xyz = str(mol_dir / "water.xyz")
    b = g.BasisSet(gbs, xyz)
    ob = oracle.Basis(gbs, xyz)
    erd = g.ERD(b)
    for quartet in [(5, 5, 5, 5), (3, 8, 5, 0), (11, 2, 4, 9), (0, 0, 0, 0)]:
        r = erd.quartet(*quartet)
        ref = ob.eri_quartet(*quartet).reshape(-1)
        assert np.abs(r - ref).max() < 1e-12


@gpu
@pytest.mark.parametrize("mol", ["water", "ethane", "c10h22"])
def test_screening_csr_matches_oracle(g, gbs, mol_dir, mol):
    xyz = str(mol_dir / (mol + ".xyz"))
    b = g.BasisSet(gbs, xyz)
    pf = g.PFock(b, tolscr=1e-11)
    c = pf.csr()
    ref = oracle.Basis(gbs, xyz).screening(1e-11)
    assert c["nnz"] == ref["nnz"]
    np.testing.assert_array_equal(c["shellptr"], ref["shellptr"])
    np.testing.assert_array_equal(c["shellid"], ref["shellid"])
    np.testing.assert_array_equal(c["shellrid"], ref["shellrid"])
    np.testing.assert_allclose(c["shellvalue"], ref["shellvalue"], rtol=1e-11, atol=1e-15)
    pf.close()


@gpu
@pytest.mark.parametrize("mol", ["water", "ethane", "butane"])
def test_quartet_set_matches_reference_enumeration(g, gbs, mol_dir, mol):
    """The GPU compaction yields exactly the unique quartet set of the
    reference task loop (pfock/fock_task.c:166-192)."""
    xyz = str(mol_dir / (mol + ".xyz"))
    b = g.BasisSet(gbs, xyz)
    pf = g.PFock(b, tolscr=1e-11)
    mine = pf.quartets()
    ref = oracle.Basis(gbs, xyz).quartets(1e-11)

    def canon(q):
        M, N, P, Q = q
        a, bb = (M, N) if M >= N else (N, M)
        c, d = (P, Q) if P >= Q else (Q, P)
        return tuple(sorted([(a, bb), (c, d)]))

    s1 = set(map(canon, mine.tolist()))
    s2 = set(map(canon, ref.tolist()))
    assert len(mine) == len(ref) == len(s1) == len(s2)
    assert s1 == s2
    pf.close()


@gpu
@pytest.mark.parametrize("mol", ["water", "butane"])
def test_fock_matches_oracle_random_density(g, gbs, mol_dir, mol):
    xyz = str(mol_dir / (mol + ".xyz"))
    b = g.BasisSet(gbs, xyz)
    ob = oracle.Basis(gbs, xyz)
    D = _rand_sym(b.nbf, 11)
    pf = g.PFock(b, tolscr=1e-11)
    F = pf.fock(D)
    J = pf.get_mat(g.MAT_J)
    K = pf.get_mat(g.MAT_K)
    Jr, Kr, st = ob.fock(D)
    print(mol, "max|dF| = %.3e" % np.abs(F - (Jr + Kr)).max())
    assert np.abs(J - Jr).max() < FOCK_TOL
    assert np.abs(K - Kr).max() < FOCK_TOL
    assert np.abs(F - (Jr + Kr)).max() < FOCK_TOL
    s = pf.stats()
    assert s["usq"] == st["usq"]
    assert s["uitl"] == st["uitl"]
    pf.close()


@gpu
def test_fock_c10h22_matches_oracle(g, gbs, mol_dir):
    """Config 2 molecule, realistic density: the first SCF density,
    X purify(X H X) X, which the SCF reaches after its core-Hamiltonian step."""
    xyz = str(mol_dir / "c10h22.xyz")
    b = g.BasisSet(gbs, xyz)
    ob = oracle.Basis(gbs, xyz)
    S, H = ob.overlap(), ob.core_hamiltonian()
    X = oracle.inv_sqrt(S)
    D, _ = oracle.purification(X @ H @ X, X, ob.nocc)
    pf = g.PFock(b, tolscr=1e-11)
    F = pf.fock(D)
    Jr, Kr, st = ob.fock(D)
    err = np.abs(F - (Jr + Kr)).max()
    print("c10h22 max|dF| = %.3e, usq %.0f" % (err, st["usq"]))
    assert err < FOCK_TOL
    assert pf.stats()["usq"] == st["usq"]
    pf.close()


@gpu
def test_primitive_screening_against_oracle(g, gbs, mol_dir):
    """Primitive screening (default 1e-15) against the oracle, which
    evaluates every primitive quartet: the default stays two orders inside
    the Fock bar, a higher threshold skips more and errs more (each skipped
    primitive quartet is below its Schwarz bound), 0 turns it off, and a
    negative threshold is refused."""
    xyz = str(mol_dir / "butane.xyz")
    b = g.BasisSet(gbs, xyz)
    ob = oracle.Basis(gbs, xyz)
    D = _rand_sym(b.nbf, 5)
    Jr, Kr, _ = ob.fock(D)
    pf = g.PFock(b, tolscr=1e-11)
    err = {}
    for tol in (None, 0.0, 1e-15, 1e-12, 1e-9):
        if tol is not None:
            pf.set_prim_tol(tol)
        err[tol] = np.abs(pf.fock(D) - (Jr + Kr)).max()
        print("prim_tol %s: max|dF| = %.3e" % (tol, err[tol]))
    assert err[None] == err[1e-15] or abs(err[None] - err[1e-15]) < 1e-13
    assert err[0.0] < FOCK_TOL
    assert err[1e-15] < FOCK_TOL / 100
    assert err[1e-9] > err[1e-15]
    assert err[1e-9] < 1e-5
    with pytest.raises(g.GtfError):
        pf.set_prim_tol(-1.0)
    pf.close()


@gpu
def test_primitive_census(g, gbs, mol_dir):
    """The census of evaluated primitive quartets: with no primitive
    threshold it equals the screened quartets' primitive quartets class by
    class (FP64 and FP32 tiles together, with mixed precision too); with the
    default threshold it is smaller in the screened classes (total angular
    momentum <= 3) and unchanged in the others."""
    b = g.BasisSet(gbs, str(mol_dir / "c10h22.xyz"))
    pf = g.PFock(b)
    nq, npq = pf.class_work()
    names = g.class_names()
    low_l = np.array([sum({"s": 0, "p": 1, "d": 2}[ch] for ch in nm if ch in "spd") <= 3 for nm in names])
    pf.set_prim_tol(0.0)
    hi, lo = pf.prim_census()
    assert np.array_equal(hi + lo, npq) and not lo.any()
    pf.set_precision(1e-4)
    hi, lo = pf.prim_census()
    assert np.array_equal(hi + lo, npq) and lo.sum() > 0
    pf.set_precision(0.0)
    pf.set_prim_tol(1e-15)
    hi, lo = pf.prim_census()
    print("evaluated / all primitive quartets: %.4f" % (hi.sum() / npq.sum()))
    assert np.array_equal(hi[~low_l], npq[~low_l])
    assert (hi[low_l] <= npq[low_l]).all() and hi[low_l].sum() < npq[low_l].sum()
    pf.close()


@gpu
def test_one_electron_matches_oracle(g, gbs, mol_dir):
    xyz = str(mol_dir / "butane.xyz")
    b = g.BasisSet(gbs, xyz)
    ob = oracle.Basis(gbs, xyz)
    pf = g.PFock(b)
    S = pf.overlap()
    H = pf.core_hamiltonian()
    X = pf.x_matrix()
    assert np.abs(S - ob.overlap()).max() < 1e-12
    assert np.abs(H - ob.core_hamiltonian()).max() < 1e-11
    assert np.abs(X - oracle.inv_sqrt(ob.overlap())).max() < 1e-9
    assert np.abs(X @ S @ X - np.eye(b.nbf)).max() < 1e-10
    pf.close()


@gpu
@pytest.mark.parametrize("mask", ["0", "0xffffffff"])
def test_fock_kernel_variants_match_oracle(g, gbs, mol_dir, monkeypatch, mask):
    """Every class through the standard tile kernels (mask 0) and through the
    bra/ket-exchanged ones wherever a class has one (all bits), FP64 and the
    FP32 part of the mixed build."""
    monkeypatch.setenv("GTF_JK_EXCHANGED", mask)
    xyz = str(mol_dir / "butane.xyz")
    b = g.BasisSet(gbs, xyz)
    D = _rand_sym(b.nbf, 21)
    pf = g.PFock(b, tolscr=1e-11)
    F = pf.fock(D)
    Jr, Kr, _ = oracle.Basis(gbs, xyz).fock(D)
    err = np.abs(F - (Jr + Kr)).max()
    pf.set_precision(1e-3)
    F32 = pf.fock(D)
    err32 = np.abs(F32 - F).max()
    print("mask %s: max|dF| vs oracle %.2e; mixed (tau 1e-3) vs FP64 %.2e" % (mask, err, err32))
    assert err < FOCK_TOL
    assert err32 < 1e-6
    pf.close()
