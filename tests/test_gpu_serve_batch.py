"""Batched Fock builds and batched SCF through the C-ABI (gtf_batch_*).

Bars: a molecule's G = 2J - K from the batch equals the single-molecule
PFock build and the FP64 oracle within 1e-10 Eh; HIP-graph replays equal
the eager build; batched SCF energies equal the single-molecule SCF within
1e-9 Eh; the mixed-precision batch stays within 1e-6 Eh of FP64."""
import os

import numpy as np
import pytest

import oracle

gpu = pytest.mark.gpu
FOCK_TOL = 1e-10


@pytest.fixture(scope="module")
def g():
    import gtfock_amd
    gtfock_amd.lib()
    return gtfock_amd


@pytest.fixture(scope="module")
def small_batch(tmp_path_factory):
    from gtfock_amd import geometry as geo
    d = tmp_path_factory.mktemp("batch")
    rng = np.random.default_rng(5)
    mols = [("water", geo.water()), ("methane", geo.alkane(1)), ("ethane", geo.alkane(2)),
            ("w3", geo.water_cluster(3, rng)), ("propane", geo.alkane(3))]
    paths = []
    for name, atoms in mols:
        p = str(d / (name + ".xyz"))
        open(p, "w").write(geo.to_xyz(atoms))
        paths.append(p)
    return paths


def _rand_sym(n, seed, scale=0.05):
    rng = np.random.default_rng(seed)
    A = rng.normal(size=(n, n)) * scale
    return A + A.T


@gpu
def test_batch_fock_matches_single_and_oracle(g, gbs, small_batch):
    bases = [g.BasisSet(gbs, p) for p in small_batch]
    Ds = [_rand_sym(b.nbf, 100 + i) for i, b in enumerate(bases)]
    bt = g.Batch(bases, graph=False)
    Gs = bt.fock(Ds)
    for i, (b, p) in enumerate(zip(bases, small_batch)):
        pf = g.PFock(b, tolscr=1e-11)
        G1 = pf.fock(Ds[i])
        pf.close()
        err = np.abs(Gs[i] - G1).max()
        print("%s: nbf %d, max|G_batch - G_single| %.2e" % (os.path.basename(p), b.nbf, err))
        assert err < 1e-12
    # the oracle on the two smallest
    for i in (0, 3):
        Jr, Kr, _ = oracle.Basis(gbs, small_batch[i]).fock(Ds[i])
        assert np.abs(Gs[i] - (Jr + Kr)).max() < FOCK_TOL
    bt.close()


@gpu
def test_batch_order_independent(g, gbs, small_batch):
    bases = [g.BasisSet(gbs, p) for p in small_batch]
    Ds = [_rand_sym(b.nbf, 200 + i) for i, b in enumerate(bases)]
    a = g.Batch(bases)
    Ga = a.fock(Ds)
    perm = [4, 2, 0, 3, 1]
    b = g.Batch([bases[k] for k in perm])
    Gb = b.fock([Ds[k] for k in perm])
    for j, k in enumerate(perm):
        assert np.abs(Gb[j] - Ga[k]).max() < 1e-12
    a.close()
    b.close()


@gpu
def test_batch_graph_replays_equal_eager(g, gbs, small_batch):
    bases = [g.BasisSet(gbs, p) for p in small_batch]
    Ds = [_rand_sym(b.nbf, 300 + i) for i, b in enumerate(bases)]
    bt = g.Batch(bases, graph=False)
    ref = bt.fock(Ds)
    bt.use_graph(True)
    for rep in range(4):
        # eager work between replays: a host round trip of the density
        Ds2 = [D * (1.0 + 0.0 * rep) for D in Ds]
        G = bt.fock(Ds2)
        assert bt.stats()["graph"] == 1.0
        worst = max(np.abs(x - y).max() for x, y in zip(G, ref))
        print("replay %d: max|G_graph - G_eager| %.2e" % (rep, worst))
        assert worst < 1e-12
    # a different density through the same graph
    Ds3 = [0.5 * D for D in Ds]
    G3 = bt.fock(Ds3)
    for x, y in zip(G3, ref):
        assert np.abs(x - 0.5 * y).max() < 1e-12
    bt.close()


@gpu
def test_batch_primitive_threshold_reaches_the_graph(g, gbs, small_batch):
    """gtf_batch_set_prim_tol takes effect in the captured graph (it is
    captured again): 0 and the default agree to rounding, a high threshold
    visibly changes G, and going back to 0 restores it."""
    bases = [g.BasisSet(gbs, p) for p in small_batch]
    Ds = [_rand_sym(b.nbf, 400 + i) for i, b in enumerate(bases)]
    bt = g.Batch(bases)
    G_def = bt.fock(Ds)
    assert bt.stats()["graph"] == 1.0
    bt.set_prim_tol(0.0)
    G0 = bt.fock(Ds)
    bt.set_prim_tol(1e-9)
    G9 = bt.fock(Ds)
    bt.set_prim_tol(0.0)
    G0b = bt.fock(Ds)
    assert bt.stats()["graph"] == 1.0
    d_def = max(np.abs(x - y).max() for x, y in zip(G_def, G0))
    d9 = max(np.abs(x - y).max() for x, y in zip(G9, G0))
    d0 = max(np.abs(x - y).max() for x, y in zip(G0b, G0))
    print("max|G - G(prim_tol 0)|: default %.2e, 1e-9 %.2e, 0 again %.2e" % (d_def, d9, d0))
    assert d_def < 1e-12 and d0 < 1e-12
    assert 1e-12 < d9 < 1e-5
    with pytest.raises(g.GtfError):
        bt.set_prim_tol(-1.0)
    bt.close()


@gpu
def test_batch_scf_matches_single_scf(g, gbs, small_batch):
    bases = [g.BasisSet(gbs, p) for p in small_batch]
    bt = g.Batch(bases)
    e, it, hist = bt.scf(max_iters=12, ediff=0.0, history=True)
    for i, p in enumerate(small_batch):
        e1, _, _ = g.scf(gbs, p, 12)
        n = min(len(e1), 12)
        d = np.abs(hist[i, :n] - e1[:n]).max()
        print("%s: E %.10f, max |dE| vs single SCF over %d iterations %.2e" % (os.path.basename(p), e[i], n, d))
        assert d < 1e-9
    bt.close()


@gpu
def test_batch_scf_water_matches_oracle(g, gbs, small_batch):
    b = g.BasisSet(gbs, small_batch[0])
    bt = g.Batch([b, g.BasisSet(gbs, small_batch[2])])
    e, it, hist = bt.scf(max_iters=10, ediff=0.0, history=True)
    eo, _, _ = oracle.Basis(gbs, small_batch[0]).scf(10)
    n = len(eo)
    d = np.abs(hist[0, :n] - eo).max()
    print("water batched SCF %.12f, oracle %.12f, max |dE| over %d iterations %.2e" % (e[0], eo[-1], n, d))
    assert d < 1e-9
    bt.close()


@gpu
def test_batch_mixed_energy_error(g, gbs, small_batch):
    bases = [g.BasisSet(gbs, p) for p in small_batch]
    b64 = g.Batch(bases)
    e64, _ = b64.scf(max_iters=15, ediff=1e-10)
    b32 = g.Batch(bases, tau=1e-4)
    e32, _ = b32.scf(max_iters=15, ediff=1e-10)
    st = b32.stats()
    print("mixed batch: FP32 quartets %d of %d; max |dE| %.2e" % (st["nq_fp32"], st["nquartets"],
                                                                   np.abs(e64 - e32).max()))
    assert np.abs(e64 - e32).max() < 1e-6
    b64.close()
    b32.close()


@gpu
def test_server_alone_equals_in_mix(g, gbs, small_batch):
    """A job's energy is the same when it runs alone and when it shares a
    cohort (and the device) with other jobs."""
    bases = [g.BasisSet(gbs, p) for p in small_batch]
    # alone: one job per cohort, one cohort at a time
    s1 = g.Server(hbm_budget_gb=8.0, max_cohorts=1, max_jobs_per_cohort=1, max_iters=20, ediff=1e-10)
    ids = [s1.submit(b) for b in bases]
    s1.wait()
    alone = [s1.poll(i) for i in ids]
    s1.close()
    # in the mix: all jobs together, several cohorts overlapping
    s2 = g.Server(hbm_budget_gb=8.0, max_cohorts=3, max_jobs_per_cohort=3, max_iters=20, ediff=1e-10)
    ids = [s2.submit(b) for b in bases + bases]
    s2.wait()
    mix = [s2.poll(i) for i in ids]
    st = s2.stats()
    s2.close()
    for k, r in enumerate(alone + mix):
        print("job %d: status %d, %d builds, cohort %d, E %.12f" % (k, r["status"], r["iters"], r["cohort"], r["energy"]))
    assert all(r["status"] == 2 for r in alone + mix)
    assert st["cohorts"] >= 2
    for k, r in enumerate(mix):
        a = alone[k % len(bases)]
        print("job %d: alone %.12f (%d builds), mix %.12f (%d builds, cohort %d)" % (
            k, a["energy"], a["iters"], r["energy"], r["iters"], r["cohort"]))
        assert abs(r["energy"] - a["energy"]) < 1e-9
        assert r["iters"] == a["iters"]


@gpu
def test_server_budget_admission(g, gbs, small_batch):
    """Under a budget that holds little more than the largest job every job
    still completes, the budget is never exceeded, and a job larger than the
    whole budget fails instead of blocking the queue."""
    bases = [g.BasisSet(gbs, p) for p in small_batch[:3]]
    big = g.BasisSet(gbs, small_batch[4])
    s = g.Server(hbm_budget_gb=8.0, max_cohorts=1, max_jobs_per_cohort=8, max_iters=2, ediff=1e-9)
    est = [s.poll(s.submit(b))["mem_est"] for b in bases + [big]]
    s.wait()
    s.close()
    budget = 1.05 * max(est[:3])
    assert est[3] > budget
    s = g.Server(hbm_budget_gb=budget / 2 ** 30, max_cohorts=4, max_jobs_per_cohort=8, max_iters=15, ediff=1e-9)
    ids = [s.submit(b) for b in bases] + [s.submit(big)] + [s.submit(b) for b in bases]
    s.wait()
    res = [s.poll(i) for i in ids]
    st = s.stats()
    s.close()
    print([r["status"] for r in res], st)
    assert 0 < st["peak_bytes"] <= st["budget"]
    assert [r["status"] for r in res] == [2, 2, 2, -2, 2, 2, 2]


@gpu
def test_job_memory_estimates_cover_real_allocations(g, gbs, small_batch, mol_dir, hip):
    """The server admits jobs by gtf_batch_job_bytes: everything a batch
    allocates (device bytes counted cumulatively, so the setup temporaries
    freed after gtf_batch_create count too) while it runs SCF iterations
    with split-BF16 purification stays within the sum of its molecules'
    estimates, for the small molecules and for C54H18.  The drop in free
    device memory is printed beside it: it also holds what the HIP runtime
    and rocSOLVER keep for themselves, outside the library's accounting."""
    for paths in (small_batch, [str(mol_dir / "c54h18.xyz")]):
        bases = [g.BasisSet(gbs, p) for p in paths]
        est = sum(g.lib().gtf_batch_job_bytes(b.h) for b in bases)
        free0 = hip.free_bytes()
        bt = g.Batch(bases, tau=1e-4)
        bt.set_purif_precision("split")
        bt.scf_begin(max_iters=3, ediff=1e-12)
        for _ in range(3):
            if bt.scf_step() <= 0:
                break
        free1 = hip.free_bytes()
        tracked = bt.stats()["device_bytes"]
        bt.close()
        print("%d molecules (nbf %s): estimate %.1f MB, allocated %.1f MB, free memory drop %.1f MB" % (
            len(bases), [b.nbf for b in bases], est / 1e6, tracked / 1e6, (free0 - free1) / 1e6))
        assert 0 < tracked <= est


@gpu
def test_server_tuned_shapes_keep_energies(g, gbs, small_batch):
    """Cohorts running with the shipped tuning tables (by molecule size, as
    the serving benchmark sets them) give the energies of cohorts running
    with the built-in shapes: a tile shape changes only summation order."""
    from gtfock_amd import tuning
    bases = [g.BasisSet(gbs, p) for p in small_batch]
    small, _ = tuning.load("batch_mixed")
    large, _ = tuning.load("c54h18_mixed")
    res = []
    for ranges in ([], [(0, 1 << 30, small), (60, 1 << 30, large)]):
        s = g.Server(hbm_budget_gb=8.0, max_cohorts=2, max_jobs_per_cohort=2, tau=1e-4, max_iters=20, ediff=1e-10)
        for lo, hi, sh in ranges:
            s.set_tile_shapes_range(lo, hi, sh)
        ids = [s.submit(b) for b in bases]
        s.wait()
        res.append([s.poll(i) for i in ids])
        s.close()
    for k, (a, b) in enumerate(zip(*res)):
        print("job %d (nbf %d): default %.12f (%d), tuned %.12f (%d)" % (
            k, bases[k].nbf, a["energy"], a["iters"], b["energy"], b["iters"]))
        assert a["status"] == 2 and b["status"] == 2
        assert abs(a["energy"] - b["energy"]) < 1e-8


@gpu
def test_batched_scf_starts_from_the_atomic_guess(g, gbs, mol_dir, tmp_path):
    """With <Element>.dat guesses beside the basis file the batched SCF
    starts, as scf.c does, from the atomic blocks (R/2 scaled) and matches
    the single-molecule SCF from the same start, molecule by molecule."""
    import os
    import shutil
    golden = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "guess")
    for f in ("H.dat", "C.dat", "O.dat"):
        shutil.copy(os.path.join(golden, f), str(tmp_path / f))
    basis_file = str(tmp_path / "cc-pvdz.gbs")
    shutil.copy(gbs, basis_file)
    xyzs = [str(mol_dir / (n + ".xyz")) for n in ("water", "methane", "ethane")]
    bt = g.Batch([g.BasisSet(basis_file, x) for x in xyzs])
    e, it = bt.scf(max_iters=30, ediff=1e-11)
    bt.close()
    bz = g.Batch([g.BasisSet(gbs, x) for x in xyzs])
    ez, itz = bz.scf(max_iters=30, ediff=1e-11)
    bz.close()
    for k, x in enumerate(xyzs):
        es, _, _ = g.scf(basis_file, x, 30, guess=True)
        print("%s: batch from guess %.10f (%d builds), single from guess %.10f (%d), batch from zero %.10f (%d)" % (
            os.path.basename(x), e[k], it[k], es[-1], len(es), ez[k], itz[k]))
        assert abs(e[k] - es[-1]) < 1e-9
        assert abs(e[k] - ez[k]) < 1e-8


@gpu
def test_batch_regroup_carries_scf_state(g, gbs, small_batch):
    """A batch regrouped after three iterations -- its running molecules'
    SCF state carried into a new batch, a new molecule starting there --
    gives every molecule the energies and build counts of an SCF that ran
    without the regroup."""
    bases = [g.BasisSet(gbs, p) for p in small_batch]
    first, newcomer = bases[1:4], bases[4]
    ref = g.Batch(first + [newcomer])
    e_ref, it_ref = ref.scf(max_iters=25, ediff=1e-10)
    ref.close()
    bt = g.Batch(first)
    bt.scf_begin(max_iters=25, ediff=1e-10)
    for _ in range(3):
        assert bt.scf_step() == len(first)
    nb, origin = bt.regroup([newcomer])
    bt.close()
    assert list(origin) == [0, 1, 2, -1]
    while nb.scf_step() > 0:
        pass
    e, it, hist = nb.scf_results()
    nb.close()
    for k in range(4):
        print("molecule %d: regrouped %.12f (%d builds), plain %.12f (%d builds)" % (k, e[k], it[k], e_ref[k], it_ref[k]))
    assert np.abs(e - e_ref).max() < 1e-9
    assert list(it) == list(it_ref)
    # the carried molecules kept their first three energies
    assert not np.isnan(hist[:3, :3]).any()


@gpu
def test_server_job_joins_running_cohort(g, gbs, tmp_path):
    """Jobs submitted while a cohort of their size class runs join it between
    iterations (same cohort, no new one) and get the energies and build
    counts they get alone."""
    import time
    from gtfock_amd import geometry as geo
    mols = [("c10h22", geo.alkane(10)), ("c8h18", geo.alkane(8)), ("w9", geo.water_cluster(9, np.random.default_rng(9)))]
    paths = []
    for name, atoms in mols:
        p = str(tmp_path / (name + ".xyz"))
        open(p, "w").write(geo.to_xyz(atoms))
        paths.append(p)
    bases = [g.BasisSet(gbs, p) for p in paths]
    assert all(192 <= b.nbf < 384 for b in bases)  # one size class
    alone = []
    for b in bases:
        s = g.Server(hbm_budget_gb=32.0, max_cohorts=1, max_jobs_per_cohort=1, max_iters=30, ediff=1e-9,
                     join_gap=-1)
        jid = s.submit(b)
        s.wait()
        alone.append(s.poll(jid))
        s.close()
    s = g.Server(hbm_budget_gb=32.0, max_cohorts=4, max_jobs_per_cohort=8, max_iters=30, ediff=1e-9)
    ids = [s.submit(bases[0])]
    while s.poll(ids[0])["status"] == 0:
        time.sleep(0.001)
    ids += [s.submit(b) for b in bases[1:]]
    s.wait()
    mix = [s.poll(i) for i in ids]
    joins, regroups = s.join_stats()
    st = s.stats()
    s.close()
    for k, (a, r) in enumerate(zip(alone, mix)):
        print("job %d: alone %.12f (%d builds), joined %.12f (%d builds, cohort %d, started %.3f s after the first)" % (
            k, a["energy"], a["iters"], r["energy"], r["iters"], r["cohort"], r["t_start"] - mix[0]["t_start"]))
    print("joins %d, regroups %d, cohorts %d" % (joins, regroups, st["cohorts"]))
    assert all(r["status"] == 2 for r in mix)
    assert joins == 2 and st["cohorts"] == 1
    assert all(r["t_start"] < mix[0]["t_done"] for r in mix[1:])
    for a, r in zip(alone, mix):
        assert abs(a["energy"] - r["energy"]) < 1e-9
        assert a["iters"] == r["iters"]
