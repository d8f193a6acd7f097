"""Tile shapes and the tile autotuner (gtf_pfock_autotune, gtf_batch_autotune)
through the C-ABI.

A tile shape only changes how the screened quartets of a class are cut into
tiles and launched, so every shape must give the oracle's Fock matrix
(1e-10 Eh max-abs) and the default shapes' matrix up to FP64 summation
order.  The extreme shapes below reach the corner cases of the cut: one ket
per tile, tiles whose J_PQ strip no longer fits in LDS (the class then runs
the row kernel), and one persistent workgroup draining every tile; each
with the Boys table read from HBM and staged in LDS.
"""
import glob
import os

import numpy as np
import pytest

import oracle

gpu = pytest.mark.gpu
FOCK_TOL = 1e-10
ORDER_TOL = 1e-11  # same quartets, another FP64 summation order

EXTREME_SHAPES = {
    "finest": dict(units=1, unit_pq=1, tile_pq=1, grid=2048, boys_lds=0),
    "finest_lds": dict(units=1, unit_pq=1, tile_pq=1, grid=2048, boys_lds=1),
    "coarsest": dict(units=4096, unit_pq=256, tile_pq=1 << 22, grid=8192, boys_lds=0),
    "coarsest_lds": dict(units=4096, unit_pq=256, tile_pq=1 << 22, grid=8192, boys_lds=1),
    "one_workgroup_lds": dict(units=1024, unit_pq=64, tile_pq=65536, grid=1, boys_lds=1),
    "small": dict(units=256, unit_pq=16, tile_pq=16384, grid=512, boys_lds=0),
}


def _rand_sym(n, seed, scale=0.1):
    rng = np.random.default_rng(seed)
    A = rng.normal(size=(n, n)) * scale
    return A + A.T


@pytest.fixture(scope="module")
def g():
    import gtfock_amd
    gtfock_amd.lib()
    return gtfock_amd


@pytest.fixture(scope="module")
def butane(gbs, mol_dir):
    xyz = str(mol_dir / "butane.xyz")
    ob = oracle.Basis(gbs, xyz)
    D = _rand_sym(ob.nbf, 31)
    Jr, Kr, _ = ob.fock(D)
    return xyz, D, Jr + Kr


@gpu
@pytest.mark.parametrize("name", sorted(EXTREME_SHAPES))
def test_every_tile_shape_matches_oracle(g, gbs, butane, name):
    xyz, D, Fr = butane
    b = g.BasisSet(gbs, xyz)
    pf = g.PFock(b, tolscr=1e-11)
    F0 = pf.fock(D)
    for cls in range(21):
        sh = dict(EXTREME_SHAPES[name])
        sh["exchanged"] = pf.tile_shape(cls)["exchanged"]
        sh["fp32"] = cls % 2  # every other class stays FP64 in the mixed build below
        pf.set_tile_shape(cls, sh)
        assert pf.tile_shape(cls) == sh
    print("shape %s, case butane seed 31" % name, flush=True)
    F = pf.fock(D)
    err = np.abs(F - Fr).max()
    print("shape %s: max|dF| vs oracle %.2e, vs default shapes %.2e" % (name, err, np.abs(F - F0).max()))
    assert err < FOCK_TOL
    assert np.abs(F - F0).max() < ORDER_TOL
    # the FP32 tiles of the mixed build are cut by the same shapes
    pf.set_precision(1e-3)
    F32 = pf.fock(D)
    assert np.abs(F32 - F0).max() < 1e-6
    pf.close()


@gpu
def test_exchanged_kernels_through_the_shape_api(g, gbs, butane):
    xyz, D, Fr = butane
    b = g.BasisSet(gbs, xyz)
    pf = g.PFock(b, tolscr=1e-11)
    flipped = 0
    for cls in range(21):
        sh = pf.tile_shape(cls)
        sh["exchanged"] = 1 - sh["exchanged"]
        try:
            pf.set_tile_shape(cls, sh)
            flipped += 1
        except g.GtfError:
            # no exchanged kernel for this class: it stays as it was
            assert sh["exchanged"] == 1
            assert pf.tile_shape(cls)["exchanged"] == 0
    assert flipped >= 2
    F = pf.fock(D)
    assert np.abs(F - Fr).max() < FOCK_TOL
    pf.close()


@gpu
def test_invalid_tile_shapes_are_rejected(g, gbs, butane):
    xyz, D, Fr = butane
    b = g.BasisSet(gbs, xyz)
    pf = g.PFock(b, tolscr=1e-11)
    good = pf.tile_shape(0)
    for bad in [dict(good, units=0), dict(good, grid=-1), dict(good, exchanged=2), dict(good, boys_lds=2),
                dict(good, fp32=-1),
                dict(good, exchanged=1)]:  # class 0, (ss|ss), has no exchanged kernel
        with pytest.raises(g.GtfError):
            pf.set_tile_shape(0, bad)
    with pytest.raises(g.GtfError):
        pf.tile_shape(21)
    assert pf.tile_shape(0) == good
    assert np.abs(pf.fock(D) - Fr).max() < FOCK_TOL
    pf.close()


@gpu
def test_pfock_autotune_keeps_results(g, gbs, butane):
    xyz, D, Fr = butane
    b = g.BasisSet(gbs, xyz)
    pf = g.PFock(b, tolscr=1e-11)
    F0 = pf.fock(D)
    before = [pf.tile_shape(c) for c in range(21)]
    print("autotune, case butane seed 31", flush=True)
    r = pf.autotune(reps=2)
    after = [pf.tile_shape(c) for c in range(21)]
    changed = sum(x != y for x, y in zip(before, after))
    print("autotune: %.3f -> %.3f ms, %d candidates, %d classes changed" %
          (r["start_ms"], r["tuned_ms"], r["candidates"], changed))
    assert r["candidates"] > 21
    assert 0 < r["tuned_ms"] <= r["start_ms"]
    # the build's own accumulators are untouched by the tuning launches
    np.testing.assert_array_equal(pf.get_mat(g.MAT_F), F0)
    F1 = pf.fock(D)
    assert np.abs(F1 - F0).max() < ORDER_TOL
    assert np.abs(F1 - Fr).max() < FOCK_TOL
    pf.close()


@gpu
def test_batch_autotune_and_shape_transfer(g, gbs, mol_dir):
    names = ["water", "ethane", "butane"]
    bases = [g.BasisSet(gbs, str(mol_dir / (n + ".xyz"))) for n in names]
    Ds = [_rand_sym(b.nbf, 40 + k) for k, b in enumerate(bases)]
    bt = g.Batch(bases)
    G0 = bt.fock(Ds)
    print("batch autotune, case water/ethane/butane", flush=True)
    r = bt.autotune(reps=2)
    assert 0 < r["tuned_ms"] <= r["start_ms"]
    G1 = bt.fock(Ds)  # the graph is captured again over the new tiles
    for a, c in zip(G0, G1):
        assert np.abs(a - c).max() < ORDER_TOL
    # shapes tuned on one batch applied to another batch of the same molecules
    shapes = bt.tile_shapes()
    bt.close()
    bt2 = g.Batch(bases)
    for cls, sh in enumerate(shapes):
        bt2.set_tile_shape(cls, sh)
    assert bt2.tile_shapes() == shapes
    G2 = bt2.fock(Ds)
    for a, c in zip(G0, G2):
        assert np.abs(a - c).max() < ORDER_TOL
    bt2.close()


@gpu
def test_committed_tuning_tables_match_oracle(g, gbs, butane):
    """Every shape table the library ships applies and keeps oracle parity."""
    from gtfock_amd import tuning
    xyz, D, Fr = butane
    b = g.BasisSet(gbs, xyz)
    pf = g.PFock(b, tolscr=1e-11)
    names = sorted(os.path.basename(p)[:-5] for p in glob.glob(os.path.join(tuning.TUNING_DIR, "*.json")))
    assert names
    for name in names:
        shapes, _ = tuning.load(name)
        print("table %s, case butane seed 31" % name, flush=True)
        pf.set_tile_shapes(shapes)
        assert pf.tile_shapes() == shapes
        assert np.abs(pf.fock(D) - Fr).max() < FOCK_TOL
    pf.close()


@gpu
def test_fp32_off_keeps_classes_in_fp64(g, gbs, butane):
    """With fp32 = 0 on every class a mixed-precision build evaluates no
    quartet in FP32 and equals the FP64 build; with it on again the FP32
    share returns."""
    xyz, D, Fr = butane
    b = g.BasisSet(gbs, xyz)
    pf = g.PFock(b, tolscr=1e-11)
    F64 = pf.fock(D)
    shapes = pf.tile_shapes()
    pf.set_precision(1e-3)
    nq_lo, _ = pf.mixed_counts()
    assert nq_lo.sum() > 0
    pf.set_tile_shapes([dict(sh, fp32=0) for sh in shapes])
    nq_lo, _ = pf.mixed_counts()
    assert nq_lo.sum() == 0
    F = pf.fock(D)
    assert np.abs(F - F64).max() < ORDER_TOL
    assert np.abs(F - Fr).max() < FOCK_TOL
    pf.set_tile_shapes(shapes)
    nq_lo, _ = pf.mixed_counts()
    assert nq_lo.sum() > 0
    pf.close()
