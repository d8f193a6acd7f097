"""Host-side checks that need no GPU: the C-ABI library loads and exports
every symbol include/*.h declares, the basis-set API, the host error paths,
the geometry generators, and that GTFock's own SCF driver compiles against
these headers."""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle
from gtfock_amd import geometry

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def g():
    import gtfock_amd
    gtfock_amd.lib()
    return gtfock_amd


def test_library_exports_every_declared_symbol(g):
    names = g.exported_symbols()
    assert len(names) > 50
    L = ctypes.CDLL(g.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_basis_api_matches_oracle(g, gbs, mol_dir):
    for mol in ["water", "c10h22", "c54h18"]:
        xyz = str(mol_dir / (mol + ".xyz"))
        b = g.BasisSet(gbs, xyz)
        ob = oracle.Basis(gbs, xyz)
        assert (b.nshells, b.nbf, b.natoms, b.nocc) == (ob.nshells, ob.nbf, ob.natoms, ob.nocc)
        assert b.shell_fstart == ob.shell_fstart
        assert abs(b.nuc_energy() - ob.nuc_energy()) < 1e-9
    sizes = {"water": (12, 24, 5), "c10h22": (126, 250, 41), "c54h18": (378, 846, 171)}
    for mol, (ns, nbf, nocc) in sizes.items():
        b = g.BasisSet(gbs, str(mol_dir / (mol + ".xyz")))
        assert (b.nshells, b.nbf, b.nocc) == (ns, nbf, nocc)


def test_pack_unpack_roundtrip(g, gbs, mol_dir):
    L = g.lib()
    b = g.BasisSet(gbs, str(mol_dir / "ethane.xyz"))
    buf = ctypes.c_void_p()
    size = ctypes.c_int()
    assert L.CInt_packBasisSet(b.h, ctypes.byref(buf), ctypes.byref(size)) == 0
    h2 = ctypes.c_void_p()
    assert L.CInt_createBasisSet(ctypes.byref(h2)) == 0
    assert L.CInt_unpackBasisSet(h2, buf) == 0
    assert L.CInt_getNumFuncs(h2) == b.nbf
    assert L.CInt_getNumShells(h2) == b.nshells
    assert abs(L.CInt_getNucEnergy(h2) - b.nuc_energy()) < 1e-12
    L.CInt_destroyBasisSet(h2)
    ctypes.CDLL(None).free(buf)


def test_imported_basis_guess_is_read_only_zero(g):
    """CInt_importBasisSet: an imported basis has no guess blocks, so
    CInt_getInitialGuess hands out a zero block without writing the basis;
    element numbers outside the table are rejected."""
    L = g.lib()
    h = ctypes.c_void_p()
    assert L.CInt_createBasisSet(ctypes.byref(h)) == 0
    I = ctypes.c_int * 2
    D = ctypes.c_double * 2
    Zs, X, Y, Z = I(1, 1), D(0.0, 0.0), D(0.0, 0.0), D(0.0, 1.4)
    spa, pps, Ls = I(1, 1), I(1, 1), I(0, 0)
    cc, al = D(1.0, 1.0), D(0.5, 0.5)
    args = [h, 2, Zs, X, Y, Z, 2, 2, 1, spa, pps, Ls, cc, al]
    assert L.CInt_importBasisSet(*args) == 0
    guess = ctypes.POINTER(ctypes.c_double)()
    s, e = ctypes.c_int(), ctypes.c_int()
    for a in range(2):
        L.CInt_getInitialGuess(h, a, ctypes.byref(guess), ctypes.byref(s), ctypes.byref(e))
        assert (s.value, e.value) == (a, a) and guess[0] == 0.0
    bad = list(args)
    bad[2] = I(1, 99)
    assert L.CInt_importBasisSet(*bad) == 3  # CINT_STATUS_INVALID_VALUE
    L.CInt_destroyBasisSet(h)


def test_charge_line_and_guess(g, gbs, tmp_path):
    xyz = tmp_path / "oh.xyz"
    xyz.write_text("2\n-1\nO 0 0 0\nH 0 0 0.97\n")
    b = g.BasisSet(gbs, str(xyz))
    assert b.charge == -1
    assert b.nocc == 5
    L = g.lib()
    guess = ctypes.POINTER(ctypes.c_double)()
    s, e = ctypes.c_int(), ctypes.c_int()
    L.CInt_getInitialGuess(b.h, 1, ctypes.byref(guess), ctypes.byref(s), ctypes.byref(e))
    assert (s.value, e.value) == (14, 18)
    # no <Element>.dat next to the basis file: zero guess (core-Hamiltonian start)
    assert all(guess[i] == 0.0 for i in range(25))


def test_bad_inputs_are_rejected(g, gbs, tmp_path):
    L = g.lib()
    h = ctypes.c_void_p()
    L.CInt_createBasisSet(ctypes.byref(h))
    assert L.CInt_loadBasisSet(h, gbs.encode(), b"/nonexistent.xyz") == 6  # FILEIO_FAILED
    bad = tmp_path / "x.xyz"
    bad.write_text("1\n\nXe 0 0 0\n")
    assert L.CInt_loadBasisSet(h, gbs.encode(), str(bad).encode()) in (3, 6)
    L.CInt_destroyBasisSet(h)


def test_pfock_create_fails_cleanly_without_gpu(g, gbs, mol_dir):
    """Without a visible device the engine refuses with INIT_FAILED instead of
    falling back to a CPU path; argument errors are reported first."""
    b = g.BasisSet(gbs, str(mol_dir / "water.xyz"))
    L = g.lib()
    pf = ctypes.c_void_p()
    ndev = ctypes.c_int(0)
    try:
        hip = ctypes.CDLL("libamdhip64.so")
        hip.hipGetDeviceCount(ctypes.byref(ndev))
    except OSError:
        pass
    if ndev.value > 0:
        pytest.skip("a GPU is visible; covered by the GPU tests")
    st = L.PFock_create(b.h, 1, 1, 4, 1e-11, 1, 1, ctypes.byref(pf))
    assert st == 1 and not pf.value


def test_geometries():
    w = geometry.water()
    O, H1, H2 = (np.array(p) for _, p in w)
    assert abs(np.linalg.norm(H1 - O) - 0.9572) < 1e-9
    ang = np.degrees(np.arccos(np.dot(H1 - O, H2 - O) / 0.9572 ** 2))
    assert abs(ang - 104.52) < 1e-9
    for n in (1, 2, 5, 10, 20):
        a = geometry.alkane(n)
        assert geometry.formula(a) == ("C%dH%d" % (n, 2 * n + 2)).replace("C1H", "CH")
        xyz = np.array([p for _, p in a])
        d = np.linalg.norm(xyz[:, None] - xyz[None], axis=-1) + np.eye(len(a)) * 9
        assert d.min() > 1.05
    for r, f in ((2, "C24H12"), (3, "C54H18"), (4, "C96H24")):
        assert geometry.formula(geometry.graphene_flake(r)) == f
    rng = np.random.default_rng(1)
    wc = geometry.water_cluster(8, rng)
    Os = np.array([p for s, p in wc if s == "O"])
    d = np.linalg.norm(Os[:, None] - Os[None], axis=-1) + np.eye(8) * 9
    assert d.min() >= 2.7


@pytest.mark.skipif(not os.path.isdir("/root/reference/pscf"), reason="GTFock sources not present")
def test_reference_scf_driver_compiles_against_headers(tmp_path):
    """GTFock's own pscf/scf.c compiles unchanged against include/ (API
    compatibility; compiled only, in this container, never copied)."""
    mpi_inc = "/opt/conda/include"
    if not os.path.exists(os.path.join(mpi_inc, "mpi.h")) or not shutil.which("gcc"):
        pytest.skip("no mpi.h / gcc")
    out = tmp_path / "scf.o"
    # a temporary copy, so that its quoted includes resolve to include/ and
    # not to the reference headers beside it
    src = tmp_path / "scf.c"
    shutil.copy("/root/reference/pscf/scf.c", src)
    r = subprocess.run(["gcc", "-std=gnu99", "-fopenmp", "-c", str(src),
                        "-I", os.path.join(ROOT, "include"), "-I", mpi_inc, "-o", str(out)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_atomic_guess_blocks_from_dat_files(g, gbs, tmp_path):
    """The reference's atomic density guesses (data/opt-cc-pvdz/*.dat, kept
    as fixtures) read through CInt_getInitialGuess land on each atom's
    diagonal block, symmetric, with tr(D_AA S_AA) equal to the atom's
    nuclear charge under this library's shell order and normalisation; the
    scf.c starting density is half of them (R = 1, neutral), and for a
    charged molecule R = (N_neutral - Q) / N_neutral."""
    import shutil
    import numpy as np
    import oracle
    from gtfock_amd import geometry
    golden = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "guess")
    for f in ("H.dat", "C.dat", "O.dat"):
        shutil.copy(os.path.join(golden, f), str(tmp_path / f))
    shutil.copy(gbs, str(tmp_path / "cc-pvdz.gbs"))
    basis_file = str(tmp_path / "cc-pvdz.gbs")
    for name, atoms, z in (("ch4", geometry.alkane(1), [6, 1, 1, 1, 1]), ("h2o", geometry.water(), [8, 1, 1])):
        xyz = tmp_path / (name + ".xyz")
        xyz.write_text(geometry.to_xyz(atoms))
        b = g.BasisSet(basis_file, str(xyz))
        S = oracle.Basis(basis_file, str(xyz)).overlap()
        D0 = g.initial_guess(b)
        assert np.abs(D0 - D0.T).max() < 1e-15
        L = g.lib()
        for a, za in enumerate(z):
            s, e = ctypes.c_int(), ctypes.c_int()
            gp = ctypes.POINTER(ctypes.c_double)()
            L.CInt_getInitialGuess(b.h, a, ctypes.byref(gp), ctypes.byref(s), ctypes.byref(e))
            blk = slice(s.value, e.value + 1)
            assert abs(np.trace(2.0 * D0[blk, blk] @ S[blk, blk]) - za) < 1e-6
        # only the atoms' own blocks are set
        mask = np.zeros_like(D0, dtype=bool)
        for a in range(len(z)):
            s, e = ctypes.c_int(), ctypes.c_int()
            gp = ctypes.POINTER(ctypes.c_double)()
            L.CInt_getInitialGuess(b.h, a, ctypes.byref(gp), ctypes.byref(s), ctypes.byref(e))
            mask[s.value:e.value + 1, s.value:e.value + 1] = True
        assert not D0[~mask].any()
        assert abs(np.trace(D0 @ S) - sum(z) / 2.0) < 1e-6
    # OH-: N_neutral = 9 electrons of the neutral atoms, Q = -1, R = 10 / 9,
    # so the starting density holds the anion's 5 occupied orbitals
    xyz = tmp_path / "oh.xyz"
    xyz.write_text("2\n-1\nO 0 0 0\nH 0 0 0.97\n")
    b = g.BasisSet(basis_file, str(xyz))
    assert (g.lib().CInt_getNneutral(b.h), g.lib().CInt_getTotalCharge(b.h)) == (9, -1)
    S = oracle.Basis(basis_file, str(xyz)).overlap()
    assert abs(np.trace(g.initial_guess(b) @ S) - 5.0) < 1e-6


def test_bench_contract_on_the_host():
    """bench.py refuses more than one GPU with a message (the build runs on
    one GPU) and lists the options the benchmark contract relies on."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr and r.stdout == ""
    h = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--help"], capture_output=True, text=True,
                       timeout=120)
    assert h.returncode == 0
    for opt in ("--gpus", "--steps", "--warmup", "--tune", "--guess", "--batch", "--serve"):
        assert opt in h.stdout
