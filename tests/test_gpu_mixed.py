"""Mixed-precision Fock build: FP32 ERIs for quartets whose Schwarz bound is
below tau, FP64 for the rest, FP64 J/K accumulation.  The bar is the
total-energy error against the FP64 path: <= 1e-6 Eh per SCF iteration."""
import numpy as np
import pytest

gpu = pytest.mark.gpu


@pytest.fixture(scope="module")
def g():
    import gtfock_amd
    gtfock_amd.lib()
    return gtfock_amd


@gpu
def test_mixed_fock_close_to_fp64(g, gbs, mol_dir):
    xyz = str(mol_dir / "c10h22.xyz")
    b = g.BasisSet(gbs, xyz)
    pf = g.PFock(b)
    rng = np.random.default_rng(2)
    S, X, H = pf.overlap(), pf.x_matrix(), pf.core_hamiltonian()
    pu = g.Purif(b)
    pu.set_xs(X, S)
    D, _ = pu.purify(X @ H @ X)
    F64 = pf.fock(D)
    pf.set_precision(1e-4)
    nq_lo, _ = pf.mixed_counts()
    assert nq_lo.sum() > 0
    F32 = pf.fock(D)
    err = np.abs(F32 - F64).max()
    de = abs(np.sum(D * (F32 - F64)))
    print("mixed: %d FP32 quartets, max|dF| %.2e, |dE| %.2e" % (nq_lo.sum(), err, de))
    assert err < 1e-5
    assert de < 1e-6
    pf.set_precision(0.0)
    assert np.abs(pf.fock(D) - F64).max() < 1e-12
    pu.close()
    pf.close()


@gpu
def test_mixed_scf_energy_error(g, gbs, mol_dir):
    xyz = str(mol_dir / "butane.xyz")
    e64, _, _ = g.scf(gbs, xyz, 12)
    e32, _, _ = g.scf(gbs, xyz, 12, tau=1e-4)
    n = min(len(e64), len(e32))
    print("butane dE per iteration", np.abs(e64[:n] - e32[:n]))
    assert np.abs(e64[:n] - e32[:n]).max() < 1e-6
