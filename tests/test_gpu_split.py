"""Split-BF16 MFMA products (exact BF16 slices, FP64 residuals) and the
purification that uses them, through the C-ABI.

Bars: integer matrices whose rows fit one slice multiply exactly; random
matrices with row scales over eight decades stay inside the rigorous bound
n (ns + 5) 2^(-7 ns) rowmax|A| colmax|B| (+ FP64 rounding); purification in
the split mode ends idempotent with tr(DS) = nocc to 1e-9, within 1e-7 of
the FP64 projector with 5 slices and 1e-9 with 6 (the purification's middle
iterations amplify product errors between the states at the gap); SCFs
with split purification, which switches to FP64 purification once the DIIS
residual is below 1e-3, converge in the FP64 SCF's iteration count (+1) to
its energy within 1e-9 Eh, the converged tail within 1e-6 Eh per iteration."""
import numpy as np
import pytest

gpu = pytest.mark.gpu


@pytest.fixture(scope="module")
def g():
    import gtfock_amd
    gtfock_amd.lib()
    return gtfock_amd


def _bound(A, B, ns):
    n = A.shape[1]
    ra = np.abs(A).max(axis=1)
    cb = np.abs(B).max(axis=0)
    return n * (ns + 5) * 2.0 ** (-7 * ns) * np.outer(ra, cb) + 1e-15 * (np.abs(A) @ np.abs(B))


@gpu
@pytest.mark.parametrize("n", [64, 130, 250])
def test_split_gemm_exact_on_small_integers(g, n):
    """Rows of integers below 2^7 are one exact slice: the product is exact."""
    rng = np.random.default_rng(n)
    A = rng.integers(-127, 128, size=(n, n)).astype(np.float64)
    B = rng.integers(-127, 128, size=(n, n)).astype(np.float64)
    # scale rows of A and columns of B by powers of two (still exact)
    A *= 2.0 ** rng.integers(-20, 20, size=(n, 1))
    B *= 2.0 ** rng.integers(-20, 20, size=(1, n))
    for ns in (3, 5):
        assert np.array_equal(g.dgemm(A, B, split=ns), A @ B)
        assert np.array_equal(g.dgemm(A, B.T.copy(), transB=True, split=ns), A @ B)


@gpu
@pytest.mark.parametrize("n", [100, 250, 846, 1464])
def test_split_gemm_error_bound(g, n):
    rng = np.random.default_rng(n + 11)
    A = rng.normal(size=(n, n)) * 10.0 ** rng.uniform(-6, 2, size=(n, 1))
    B = rng.normal(size=(n, n)) * 10.0 ** rng.uniform(-6, 2, size=(1, n))
    ref = A @ B
    errs = []
    for ns in (3, 4, 5, 6):
        C = g.dgemm(A, B, split=ns)
        err = np.abs(C - ref)
        assert (err <= _bound(A, B, ns)).all(), "ns %d: bound exceeded" % ns
        Ct = g.dgemm(A, B.T.copy(), transB=True, split=ns)
        assert np.array_equal(C, Ct)
        errs.append(float((err / np.outer(np.abs(A).max(1), np.abs(B).max(0))).max()))
    print("n %d: max error / (rowmax colmax) for 3..6 slices: %s" % (n, " ".join("%.1e" % e for e in errs)))
    # each slice adds 7 bits
    assert errs[0] > errs[1] > errs[2] > errs[3]


@gpu
@pytest.mark.parametrize("mol", ["water", "c10h22"])
def test_split_purification_matches_fp64(g, gbs, mol_dir, mol):
    xyz = str(mol_dir / (mol + ".xyz"))
    b = g.BasisSet(gbs, xyz)
    pf = g.PFock(b)
    S, X, H = pf.overlap(), pf.x_matrix(), pf.core_hamiltonian()
    pf.close()
    Fo = X @ H @ X
    p64 = g.Purif(b)
    p64.set_xs(X, S)
    D64, it64 = p64.purify(Fo)
    ps = g.Purif(b)
    ps.set_xs(X, S)
    ps.set_precision("split")
    for ns, bar in ((6, 1e-9), (5, 1e-7)):
        ps.set_slices(ns)
        Ds, its = ps.purify(Fo)
        nlow = ps.fp32_iters()
        idem = np.abs(Ds @ S @ Ds - Ds).max()
        tr = np.trace(Ds @ S)
        d = np.abs(Ds - D64).max()
        print("%s, %d slices: FP64 %d iterations; split %d (%d split BF16); max|dD| %.2e, |DSD - D| %.2e, "
              "tr DS - nocc %.2e" % (mol, ns, it64, its, nlow, d, idem, tr - b.nocc))
        assert 0 < nlow <= its
        assert d < bar
        assert idem < 1e-9
        assert abs(tr - b.nocc) < 1e-9
    # the plain FP64 mode is untouched by the split scratch
    ps.set_precision("fp64")
    D2, _ = ps.purify(Fo)
    assert np.abs(D2 - D64).max() < 1e-12
    p64.close()
    ps.close()


@gpu
@pytest.mark.parametrize("mol", ["butane", "c10h22"])
def test_split_purification_scf_energy(g, gbs, mol_dir, mol):
    """Whole SCF with split-BF16 purification: the split products run until
    the DIIS residual drops below 1e-3, FP64 purification after; the SCF
    converges in the FP64 SCF's iteration count (+1) to its energy within
    1e-9 Eh, the converged tail (FP64 energy changes below 1e-5 Eh) within
    1e-6 Eh per iteration (earlier iterations are printed: the reference
    algorithm's middle iterations amplify any perturbation, see the oracle's
    own reruns in tests/golden/scf_c10h22_oracle.json).  With the switch off
    (split in every iteration) the energy changes stall above ediff: the
    split products' rounding sets their floor."""
    xyz = str(mol_dir / (mol + ".xyz"))
    e64, _, _ = g.scf(gbs, xyz, 30)
    st = {}
    es, _, _ = g.scf(gbs, xyz, 30, purif="split", stats=st)
    n = min(len(e64), len(es))
    d = np.abs(e64[:n] - es[:n])
    low = st["purif_low_iters"]
    print("%s SCF: FP64 %.12f (%d), split purification %.12f (%d); |dE| per iteration %s; "
          "split iterations per purification %s" % (mol, e64[-1], len(e64), es[-1], len(es),
                                                    " ".join("%.0e" % x for x in d), low))
    assert len(e64) < 30 and len(e64) <= len(es) <= len(e64) + 1
    assert abs(es[-1] - e64[-1]) < 1e-9
    tail = [k for k in range(1, n) if abs(e64[k] - e64[k - 1]) < 1e-5]
    assert len(tail) >= 2 and max(d[k] for k in tail) < 1e-6
    assert low[0] > 0 and low[-1] == 0
    if mol == "butane":
        ea, _, _ = g.scf(gbs, xyz, 30, purif="split", purif_resid=0.0)
        print("split in every iteration: %d iterations, final |dE| %.1e" % (len(ea), abs(ea[-1] - e64[-1])))
        assert len(ea) > len(e64)


@gpu
def test_batch_split_purification(g, gbs, mol_dir):
    from gtfock_amd import geometry as geo
    paths = [str(mol_dir / "water.xyz"), str(mol_dir / "butane.xyz"), str(mol_dir / "ethane.xyz")]
    bases = [g.BasisSet(gbs, p) for p in paths]
    b64 = g.Batch(bases)
    e64, it64 = b64.scf(max_iters=20, ediff=1e-10)
    bs = g.Batch(bases)
    bs.set_purif_precision("split")
    es, its = bs.scf(max_iters=20, ediff=1e-10)
    d = np.abs(e64 - es).max()
    print("batched SCF, split purification: max |dE| %.2e; builds FP64 %s, split %s" % (d, list(it64), list(its)))
    assert d < 1e-9
    assert all(0 < a <= b <= a + 1 for a, b in zip(it64, its))
    b64.close()
    bs.close()
