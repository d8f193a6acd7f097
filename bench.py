"""Fock-build throughput on one MI355X.

Workload (default): all-trans C10H22 / cc-pVDZ (126 shells, 250 basis
functions), Schwarz tolerance 1e-11, plain FP64 GPU drop-in.  One step is one
PFock_computeFock -- screened ERIs of every class with fused J/K digestion and
the symmetrisation -- with the density already resident in HBM (put and
committed before the timed region).  The density is the first SCF density
X purify(X H X) X, computed on the GPU.

Prints ONE JSON line: value = Fock builds per second, with the roofline of
the dominant kernel (event-timed per class launch on the engine's stream) and
the CPU oracle timed on this host's cores as the reported baseline.
"""
import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = "Fock builds/s FP64 vs mixed at ΔE≤1e-6 Eh; serve p50/p99 SCF latency"
FP64_PEAK = 78.6e12
HBM_PEAK = 8.0e12


def device_sync():
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
            return
    except Exception:
        pass
    import ctypes
    ctypes.CDLL("libamdhip64.so").hipDeviceSynchronize()


def molecule(name):
    from gtfock_amd import geometry
    if name == "c10h22":
        return geometry.alkane(10)
    if name == "water":
        return geometry.water()
    if name == "c54h18":
        return geometry.graphene_flake(3)
    if name.startswith("alkane"):
        return geometry.alkane(int(name[6:]))
    raise SystemExit("unknown workload %s" % name)


def first_density(g, pf, basis):
    S = pf.overlap()
    X = pf.x_matrix()
    H = pf.core_hamiltonian()
    pu = g.Purif(basis)
    pu.set_xs(X, S)
    D, _ = pu.purify(X @ H @ X)
    pu.close()
    return D


def cpu_baseline(gbs, xyz, D, tol):
    """The oracle's FP64 Fock build (OpenMP, the reference's task structure)
    on this host's cores, one full build of the same molecule and density."""
    import oracle
    b = oracle.Basis(gbs, xyz)
    t0 = time.perf_counter()
    b.fock(D, tol=tol)
    dt = time.perf_counter() - t0
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": 1.0 / dt, "unit": "Fock builds/s", "cores": threads, "kind": "port",
            "sample": "one full FP64 Fock build of the same molecule and density (%.1f s)" % dt}


def batch_leg(g, gbs, nmol, seed, tau, steps, scf_iters):
    """Many molecules per build: the seeded alkane / water-cluster batch.
    Molecule-builds per second of (a) one PFock per molecule, FP64, built one
    after another, (b) the batch, FP64, one HIP graph per build, (c) the
    batch with FP32 ERIs below the Schwarz bound tau.  The density of each
    molecule is its first SCF density.  With scf_iters > 0 the batched SCF is
    run to convergence in FP64 and in mixed precision and the largest
    total-energy difference is reported."""
    from gtfock_amd import geometry
    tmp = tempfile.mkdtemp()
    mols = geometry.batch_molecules(nmol, seed)
    bases = []
    for k, (name, atoms) in enumerate(mols):
        p = os.path.join(tmp, "m%03d.xyz" % k)
        with open(p, "w") as fh:
            fh.write(geometry.to_xyz(atoms))
        bases.append(g.BasisSet(gbs, p))
    nbf = [b.nbf for b in bases]
    t0 = time.perf_counter()
    bt = g.Batch(bases)
    t_create = time.perf_counter() - t0
    bt.scf(max_iters=2, ediff=0.0)  # leaves the first SCF density on the device
    Ds = bt.get_density()
    # (b) batched FP64, graph replays
    bt.compute()
    device_sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        bt.compute()
    device_sync()
    t_b64 = (time.perf_counter() - t0) / steps
    G64 = bt.get_fock()
    # (c) batched mixed
    bt.set_precision(tau)
    bt.compute()
    device_sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        bt.compute()
    device_sync()
    t_b32 = (time.perf_counter() - t0) / steps
    G32 = bt.get_fock()
    st = bt.stats()
    dF = max(float(np.abs(a - b).max()) for a, b in zip(G32, G64))
    dE1 = max(float(abs(np.sum(D * (a - b)))) for D, a, b in zip(Ds, G32, G64))
    bt.close()
    # (a) one PFock per molecule, FP64, sequential
    t_single = 0.0
    for b, D in zip(bases, Ds):
        pf = g.PFock(b)
        pf.put_den_mat(D)
        pf.commit()
        pf.compute_fock()
        device_sync()
        t0 = time.perf_counter()
        for _ in range(max(1, steps // 4)):
            pf.compute_fock()
        device_sync()
        t_single += (time.perf_counter() - t0) / max(1, steps // 4)
        pf.close()
    out = {"molecules": nmol, "seed": seed, "nbf_min": min(nbf), "nbf_max": max(nbf),
           "nbf_mean": float(np.mean(nbf)), "create_s": round(t_create, 3),
           "unbatched_fp64": nmol / t_single, "batched_fp64": nmol / t_b64, "batched_mixed": nmol / t_b32,
           "unit": "molecule Fock builds/s", "tau": tau,
           "mixed_vs_unbatched_fp64": t_single / t_b32, "batched_fp64_vs_unbatched": t_single / t_b64,
           "fp32_quartet_fraction": st["nq_fp32"] / max(1.0, st["nquartets"]),
           "max_abs_dF": dF, "max_first_order_dE_Eh": dE1}
    if scf_iters > 0:
        b64 = g.Batch(bases)
        e64, it64 = b64.scf(max_iters=scf_iters, ediff=1e-8)
        b64.close()
        b32 = g.Batch(bases, tau=tau)
        t0 = time.perf_counter()
        e32, it32 = b32.scf(max_iters=scf_iters, ediff=1e-8)
        t_scf = time.perf_counter() - t0
        b32.close()
        out.update({"scf_max_abs_dE_Eh": float(np.abs(e32 - e64).max()),
                    "scf_builds_mean": float(np.mean(it32)), "scf_mixed_s": t_scf})
    return out


def load_traffic(workload, cls_name):
    path = os.path.join(ROOT, "profiles", "eri_traffic_%s.json" % workload)
    if not os.path.exists(path):
        return None
    try:
        t = json.load(open(path))
        return t.get(cls_name)
    except Exception:
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c10h22")
    ap.add_argument("--tol", type=float, default=1e-11)
    ap.add_argument("--tau", type=float, default=1e-4, help="Schwarz-bound threshold of the FP32 ERIs")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--classes", action="store_true", help="print the per-class breakdown to stderr")
    ap.add_argument("--batch", type=int, default=32, help="molecules of the batched leg (0: skip)")
    ap.add_argument("--batch-seed", type=int, default=20261202)
    ap.add_argument("--batch-scf", type=int, default=0, help="SCF iterations of the batched energy check")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported "
              "(multi-GPU sharding is design-only)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import gtfock_amd as g
    from gtfock_amd import flops, geometry

    gbs = os.path.join(g.DATA_DIR, "cc-pvdz.gbs")
    tmp = tempfile.mkdtemp()
    xyz = os.path.join(tmp, args.workload + ".xyz")
    with open(xyz, "w") as fh:
        fh.write(geometry.to_xyz(molecule(args.workload)))

    basis = g.BasisSet(gbs, xyz)
    t0 = time.perf_counter()
    pf = g.PFock(basis, tolscr=args.tol)
    t_create = time.perf_counter() - t0
    D = first_density(g, pf, basis)
    pf.put_den_mat(D)
    pf.commit()
    for _ in range(args.warmup):
        pf.compute_fock()
    pf.set_timing(True)
    device_sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        pf.compute_fock()
    device_sync()
    t1 = time.perf_counter()
    ms_step = (t1 - t0) * 1e3 / args.steps
    value = args.steps / (t1 - t0)

    ms, launches = pf.class_times()
    nq, npq = pf.class_work()
    names = g.class_names()
    rows = []
    for c in range(21):
        if launches[c] == 0:
            continue
        avg_s = ms[c] / launches[c] * 1e-3
        fl = flops.class_flops(c, nq[c], npq[c])
        rows.append((ms[c], c, avg_s, fl))
    rows.sort(reverse=True)
    total_fl = sum(flops.class_flops(c, nq[c], npq[c]) for c in range(21))
    if args.classes:
        for tot, c, avg_s, fl in rows:
            print("%-10s q=%10d pq=%12.4g  %8.3f ms  %7.2f TF/s" %
                  (names[c], nq[c], npq[c], avg_s * 1e3, fl / avg_s / 1e12), file=sys.stderr)
    _, dom, dom_s, dom_fl = rows[0]
    achieved = dom_fl / dom_s
    roofline = {"bound": "mfma", "peak_kind": "FP64 (vector and matrix peaks are both 78.6 TF/s on MI355X; "
                "the ERI kernels run on the FP64 VALU)",
                "kernel": pf.class_kernel(dom), "class": names[dom],
                "achieved": achieved / 1e12, "peak": FP64_PEAK / 1e12, "unit": "TFLOP/s",
                "frac": achieved / FP64_PEAK, "traffic": load_traffic(args.workload, names[dom]),
                "algorithmic_flops_per_launch": dom_fl, "avg_launch_ms": dom_s * 1e3,
                "build_tflops": total_fl / (ms_step * 1e-3) / 1e12}

    # mixed precision on the same molecule and density: FP32 ERIs for quartets
    # whose Schwarz bound is below tau, FP64 J/K accumulation
    F64 = pf.get_mat(g.MAT_F)
    pf.set_timing(False)
    pf.set_precision(args.tau)
    for _ in range(args.warmup):
        pf.compute_fock()
    device_sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        pf.compute_fock()
    device_sync()
    t1 = time.perf_counter()
    F32 = pf.get_mat(g.MAT_F)
    nq_lo, _ = pf.mixed_counts()
    mixed = {"value": args.steps / (t1 - t0), "unit": "Fock builds/s", "tau": args.tau,
             "ms_per_step": (t1 - t0) * 1e3 / args.steps,
             "speedup_vs_fp64": (args.steps / (t1 - t0)) / value,
             "fp32_quartet_fraction": float(nq_lo.sum()) / max(1.0, float(nq.sum())),
             "max_abs_dF": float(np.abs(F32 - F64).max()),
             "dE_Eh": float(abs(np.sum(D * (F32 - F64))))}
    pf.set_precision(0.0)

    batched = None
    if args.batch > 0:
        batched = batch_leg(g, gbs, args.batch, args.batch_seed, args.tau, max(2, args.steps // 4), args.batch_scf)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(gbs, xyz, D, args.tol)

    st = pf.stats()
    out = {
        "metric": METRIC, "value": value, "unit": "Fock builds/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic geometry (all-trans C10H22, ideal bond lengths), cc-pVDZ",
        "config": {"workload": "%s/cc-pVDZ plain FP64 Fock build" % args.workload, "nbf": basis.nbf,
                   "nshells": basis.nshells, "tolscr": args.tol, "quartets": int(st["usq"]),
                   "integrals": int(st["uitl"]), "create_s": round(t_create, 3)},
        "roofline": roofline,
        "mixed": mixed,
        "batched": batched,
        "cpu_baseline": cpu,
    }
    pf.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
