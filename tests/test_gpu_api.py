"""PFock API semantics through the C ABI (pfock/pfock.h:241-441): inclusive
block ranges with strides, commit rules, fill, several density matrices,
the J / K / F accessors, error codes, and a molecule whose d classes are
all empty (H2, cc-pVDZ)."""
import ctypes

import numpy as np
import pytest

import oracle

gpu = pytest.mark.gpu
SUCCESS, INVALID, EXEC_FAILED = 0, 3, 4


@pytest.fixture(scope="module")
def g():
    import gtfock_amd
    gtfock_amd.lib()
    return gtfock_amd


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _rand_sym(n, seed):
    rng = np.random.default_rng(seed)
    A = rng.normal(size=(n, n)) * 0.05
    return A + A.T


@gpu
def test_blocks_strides_and_commit_rules(g, gbs, mol_dir):
    L = g.lib()
    xyz = str(mol_dir / "water.xyz")
    b = g.BasisSet(gbs, xyz)
    n = b.nbf
    D = _rand_sym(n, 3)
    pf = g.PFock(b)
    ref = pf.fock(D)
    # the same density put as four blocks from a wider host array (stride n + 7)
    h = n // 2
    wide = np.zeros((n, n + 7))
    wide[:, :n] = D
    for rs, re, cs, ce in [(0, h - 1, 0, h - 1), (0, h - 1, h, n - 1), (h, n - 1, 0, h - 1), (h, n - 1, h, n - 1)]:
        # the block's first element inside the wide array, rows n + 7 apart
        ptr = ctypes.cast(wide.ctypes.data + 8 * (rs * (n + 7) + cs), ctypes.POINTER(ctypes.c_double))
        st = L.PFock_putDenMat(rs, re, cs, ce, n + 7, ptr, 0, pf.h)
        assert st == SUCCESS
    assert L.PFock_commitDenMats(pf.h) == SUCCESS
    # putting after commit is refused until the next computeFock
    assert L.PFock_putDenMat(0, n - 1, 0, n - 1, n, _p(np.ascontiguousarray(D)), 0, pf.h) == EXEC_FAILED
    assert L.PFock_computeFock(b.h, pf.h) == SUCCESS
    # F read back as a block with a stride
    out = np.zeros((n - 3, n + 5))
    assert L.PFock_getMat(pf.h, g.MAT_F, 0, 2, n - 2, 1, n - 1, n + 5, _p(out)) == SUCCESS
    assert np.abs(out[:, :n - 1] - ref[2:n - 1, 1:n]).max() < 1e-13
    # after computeFock the density may be put again
    assert L.PFock_putDenMat(0, n - 1, 0, n - 1, n, _p(np.ascontiguousarray(D)), 0, pf.h) == SUCCESS
    pf.close()


@gpu
def test_error_codes(g, gbs, mol_dir):
    L = g.lib()
    b = g.BasisSet(gbs, str(mol_dir / "water.xyz"))
    n = b.nbf
    pf = g.PFock(b)
    D = np.zeros((n, n))
    assert L.PFock_putDenMat(0, n - 1, 0, n - 1, n, _p(D), 5, pf.h) == INVALID     # bad index
    assert L.PFock_putDenMat(0, n, 0, n - 1, n, _p(D), 0, pf.h) == INVALID         # row out of range
    assert L.PFock_putDenMat(3, 2, 0, n - 1, n, _p(D), 0, pf.h) == INVALID         # empty range
    assert L.PFock_putDenMat(0, n - 1, 0, n - 1, n - 1, _p(D), 0, pf.h) == INVALID  # stride too small
    out = np.zeros((n, n))
    assert L.PFock_getMat(pf.h, 9, 0, 0, n - 1, 0, n - 1, n, _p(out)) == INVALID   # bad matrix type
    pf.close()
    h = ctypes.c_void_p()
    assert L.PFock_create(b.h, 1, 1, 4, -1.0, 1, 1, ctypes.byref(h)) == INVALID     # negative tolerance
    assert L.PFock_create(b.h, 2, 1, 4, 1e-11, 1, 1, ctypes.byref(h)) == INVALID    # process grid > 1


@gpu
def test_fill_two_densities_and_jk_accessors(g, gbs, mol_dir):
    L = g.lib()
    xyz = str(mol_dir / "water.xyz")
    b = g.BasisSet(gbs, xyz)
    ob = oracle.Basis(gbs, xyz)
    n = b.nbf
    pf = g.PFock(b, max_numdmat=2)
    L.PFock_setNumDenMat(2, pf.h)
    D1 = _rand_sym(n, 11)
    assert L.PFock_putDenMat(0, n - 1, 0, n - 1, n, _p(np.ascontiguousarray(D1)), 0, pf.h) == SUCCESS
    assert L.PFock_fillDenMat(0.01, 1, pf.h) == SUCCESS
    assert L.PFock_commitDenMats(pf.h) == SUCCESS
    assert L.PFock_computeFock(b.h, pf.h) == SUCCESS
    for idx, D in ((0, D1), (1, np.full((n, n), 0.01))):
        Jr, Kr, _ = ob.fock(D)
        J = pf.get_mat(g.MAT_J, idx)
        K = pf.get_mat(g.MAT_K, idx)
        F = pf.get_mat(g.MAT_F, idx)
        assert np.abs(J - Jr).max() < 1e-10
        assert np.abs(K - Kr).max() < 1e-10
        assert np.abs(F - (J + K)).max() < 1e-13
    pf.close()


@gpu
def test_h2_without_d_classes(g, gbs, mol_dir):
    xyz = str(mol_dir / "h2.xyz")
    b = g.BasisSet(gbs, xyz)
    ob = oracle.Basis(gbs, xyz)
    pf = g.PFock(b)
    counts = pf.class_counts()
    names = g.class_names()
    assert all(counts[c] == 0 for c in range(21) if "d" in names[c])
    D = _rand_sym(b.nbf, 5)
    Jr, Kr, _ = ob.fock(D)
    assert np.abs(pf.fock(D) - (Jr + Kr)).max() < 1e-12
    e, _, _ = g.scf(gbs, xyz, 10)
    eo, _, _ = ob.scf(10)
    m = min(len(e), len(eo))
    assert np.abs(e[:m] - eo[:m]).max() < 1e-9
    pf.close()


@gpu
def test_several_densities_in_one_pass(g, gbs, mol_dir):
    """Three committed densities of C10H22 are digested from each integral in
    one build (update_F's num_dmat loop); each result equals the build of
    that density alone."""
    L = g.lib()
    b = g.BasisSet(gbs, str(mol_dir / "c10h22.xyz"))
    n = b.nbf
    Ds = [_rand_sym(n, 20 + k) for k in range(3)]
    single = []
    p1 = g.PFock(b)
    for D in Ds:
        single.append(p1.fock(D))
    p1.close()
    pf = g.PFock(b, max_numdmat=3)
    L.PFock_setNumDenMat(3, pf.h)
    for k, D in enumerate(Ds):
        assert L.PFock_putDenMat(0, n - 1, 0, n - 1, n, _p(np.ascontiguousarray(D)), k, pf.h) == SUCCESS
    assert L.PFock_commitDenMats(pf.h) == SUCCESS
    assert L.PFock_computeFock(b.h, pf.h) == SUCCESS
    assert L.gtf_pfock_fused_densities(pf.h) == 3
    for k in range(3):
        F = pf.get_mat(g.MAT_F, k)
        err = np.abs(F - single[k]).max()
        print("density %d: max |F_fused - F_alone| %.2e" % (k, err))
        assert err < 1e-11
    pf.close()


@gpu
def test_statistics_report_measured_balance(g, gbs, mol_dir, capfd):
    """PFock_getStatistics (pfock/pfock.c:1918-2010) prints the load balance
    and steals of the last build, measured per persistent workgroup."""
    b = g.BasisSet(gbs, str(mol_dir / "c10h22.xyz"))
    pf = g.PFock(b)
    D = _rand_sym(b.nbf, 7)
    pf.fock(D)
    bal = pf.balance()
    print(bal)
    assert bal["workgroups"] > 0
    assert 1.0 <= bal["load_balance"] < 4.0
    assert bal["steals"] >= 0
    capfd.readouterr()
    assert g.lib().PFock_getStatistics(pf.h) == SUCCESS
    out = capfd.readouterr().out
    assert "load blance = %.3f" % bal["load_balance"] in out
    pf.close()


@gpu
def test_setup_phase_times(g, gbs, mol_dir):
    """The set-up phase timers cover PFock_create's and a batch's set-up:
    every phase measured, their sum within the wall time of the call."""
    b = g.BasisSet(gbs, str(mol_dir / "c10h22.xyz"))
    pf = g.PFock(b)
    ph = pf.setup_times()
    print(ph)
    assert tuple(ph) == g.SETUP_PHASES
    assert all(v >= 0.0 for v in ph.values()) and ph["pairs"] > 0.0 and ph["tiles"] > 0.0
    assert sum(ph.values()) <= pf.stats()["create_seconds"] + 1e-3
    pf.close()
    w = g.BasisSet(gbs, str(mol_dir / "water.xyz"))
    bt = g.Batch([b, w, b])
    bt.scf_begin(max_iters=2, ediff=1e-8)
    bph = bt.setup_times()
    print(bph)
    assert len(bph) == len(g.SETUP_PHASES) + 2
    assert bph["scf_oneel"] > 0.0 and bph["scf_xmat"] > 0.0
    assert sum(bph[k] for k in g.SETUP_PHASES) <= bt.stats()["create_s"] + 1e-3
    bt.close()
