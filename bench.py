"""Fock-build throughput on one MI355X.

Workload (default): all-trans C10H22 / cc-pVDZ (126 shells, 250 basis
functions), Schwarz tolerance 1e-11, plain FP64 GPU drop-in.  One step is one
PFock_computeFock -- screened ERIs of every class with fused J/K digestion and
the symmetrisation -- with the density already resident in HBM (put and
committed before the timed region).  The density is the first SCF density
X purify(X H X) X, computed on the GPU.  The tile shapes of the quartet
classes are the autotuner's: by default from the library's tuning table for
the workload (written by `bench.py --tune live --write-shapes DIR`), with
--tune live from the autotuner run on this molecule and density before the
timed region, with --tune off the built-in defaults.

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


def log(msg):
    """Progress of the legs, on stderr (the JSON line alone goes to stdout)."""
    print("[bench %s] %s" % (time.strftime("%H:%M:%S"), msg), file=sys.stderr, flush=True)


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


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(gbs, xyz, D, tol, builds=3):
    """The oracle's FP64 Fock build (OpenMP, the reference's task structure)
    on this host's cores: `builds` builds of the same molecule and density;
    the value is the median of the builds' quartet loops + reduction +
    symmetrisation, the screening (CSR) pass of each build timed apart and
    left out, as GTFock screens once per PFock_create."""
    import oracle
    b = oracle.Basis(gbs, xyz)
    build_s, screen_s = [], []
    for _ in range(builds):
        _, _, st = b.fock(D, tol=tol)
        build_s.append(st["build_s"])
        screen_s.append(st["screen_s"])
    dt = float(np.median(build_s))
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": 1.0 / dt, "unit": "Fock builds/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
            "sample": "%d full FP64 Fock builds of the same molecule and density with %d OpenMP threads; median "
                      "build %.2f s (builds %s s), screening pass %.2f s per build timed apart, not counted" % (
                          builds, threads, dt, ", ".join("%.2f" % x for x in build_s), float(np.median(screen_s)))}


def batch_bases(g, gbs, nmol, seed):
    """BasisSets of the seeded alkane / water-cluster batch distribution."""
    from gtfock_amd import geometry
    tmp = tempfile.mkdtemp()
    bases = []
    for k, (name, atoms) in enumerate(geometry.batch_molecules(nmol, seed)):
        p = os.path.join(tmp, "m%03d.xyz" % k)
        with open(p, "w") as fh:
            fh.write(geometry.to_xyz(atoms))
        bases.append(g.BasisSet(gbs, p))
    return bases


class TileShapes:
    """Where each leg's tile shapes come from (--tune): "db" the library's
    tuning table for the leg when there is one (written earlier by the
    autotuner, gtfock_amd/data/tuning/<name>.json), else the autotuner run
    now; "live" the autotuner run now; "off" the built-in defaults.  With
    write_dir, shapes tuned now are also written there as tables."""

    def __init__(self, mode, write_dir, command):
        self.mode, self.write_dir, self.command = mode, write_dir, command

    def get(self, name, live):
        """(shapes or None, record); live() runs the autotuner and returns
        (shapes, autotuner result, description of what it was tuned on)."""
        from gtfock_amd import tuning
        if self.mode == "off":
            return None, {"source": "built-in defaults"}
        if self.mode == "db":
            shapes, rec = tuning.load(name)
            if shapes is not None:
                return shapes, {"source": "tuning table gtfock_amd/data/tuning/%s.json" % name,
                                "tuned": rec["tuned"]}
        t0 = time.perf_counter()
        shapes, res, on = live()
        res = dict(res, tuned_on=on)
        res.setdefault("seconds", time.perf_counter() - t0)
        if self.write_dir:
            tuning.save(name, shapes, dict(res, command=self.command, date=time.strftime("%Y-%m-%d")),
                        directory=self.write_dir)
        return shapes, {"source": "autotuner run in this process", "result": res}


def tune_on_sample(g, gbs, nmol, seed, tau, reps):
    """Tile shapes from the autotuner, run on a sample batch of nmol
    molecules drawn from the batch distribution with another seed (tuning
    the full batch would cost more than one of its SCF runs; the shapes
    carry over between batches of one distribution).  FP64 shapes, and with
    tau > 0 the shapes of the mixed form."""
    bases = batch_bases(g, gbs, nmol, seed)
    bt = g.Batch(bases)
    bt.scf(max_iters=2, ediff=0.0)
    out = {}
    on = "%d-molecule sample of the batch distribution, seed %d" % (nmol, seed)
    for key, t in (("fp64", 0.0), ("mixed", tau)):
        if key == "mixed":
            bt.set_precision(t)
        t0 = time.perf_counter()
        r = bt.autotune(reps=reps)
        r["seconds"] = time.perf_counter() - t0
        out[key] = (bt.tile_shapes(), r, on + (", tau %g" % t if t else ""))
    bt.close()
    return out


def time_batch(bt, steps):
    bt.compute()
    device_sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        bt.compute()
    device_sync()
    return (time.perf_counter() - t0) / steps


def scf_logged(bt, max_iters, ediff, tag):
    """Batched SCF one iteration at a time (gtf_batch_scf_step) with a
    progress line per iteration; returns (energies, builds) as Batch.scf."""
    bt.scf_begin(max_iters=max_iters, ediff=ediff)
    it = 0
    while True:
        left = bt.scf_step()
        it += 1
        log("%s: iteration %d, %d molecules still running" % (tag, it, left))
        if left == 0:
            break
    e, n, _ = bt.scf_results()
    return e, n


def batch_leg(g, gbs, nmol, seed, tau, steps, scf_iters, tune_nmol, tune_reps, tiles):
    """Many molecules per build: the seeded alkane / water-cluster batch.
    Molecule-builds per second of (a) one PFock per molecule, FP64, built one
    after another, (b) the batch, FP64, one HIP graph per build, with the
    default tile shapes and with the autotuner's, (c) the batch with FP32
    ERIs below the Schwarz bound tau (autotuned shapes).  The density of
    each molecule is its first SCF density.  With scf_iters > 0 the batched
    SCF is run to convergence in FP64 and in mixed precision and the largest
    total-energy difference is reported."""
    from gtfock_amd import geometry
    bases = batch_bases(g, gbs, nmol, seed)
    nbf = [b.nbf for b in bases]
    sample = {}

    def live(key):
        if not sample:
            log("batched: autotuner on a %d-molecule sample" % tune_nmol)
            sample.update(tune_on_sample(g, gbs, tune_nmol, seed + 1, tau, tune_reps))
        return sample[key]

    tuned = None
    if tune_nmol > 0 and tune_reps > 0:
        tuned = {"fp64": tiles.get("batch", lambda: live("fp64")),
                 "mixed": tiles.get("batch_mixed", lambda: live("mixed"))}
        if tuned["fp64"][0] is None:
            tuned = None
    t0 = time.perf_counter()
    bt = g.Batch(bases)
    t_create = time.perf_counter() - t0
    bt.scf(max_iters=2, ediff=0.0)  # leaves the first SCF density on the device
    Ds = bt.get_density()
    log("batched: timed builds")
    # (b) batched FP64, graph replays, default then tuned shapes
    t_b64_default = time_batch(bt, steps)
    if tuned:
        bt.set_tile_shapes(tuned["fp64"][0])
        t_b64 = time_batch(bt, steps)
    else:
        t_b64 = t_b64_default
    G64 = bt.get_fock()
    # (c) batched mixed
    bt.set_precision(tau)
    if tuned:
        bt.set_tile_shapes(tuned["mixed"][0])
    t_b32 = time_batch(bt, steps)
    G32 = bt.get_fock()
    st = bt.stats()
    dF = max(float(np.abs(a - b).max()) for a, b in zip(G32, G64))
    dE1 = max(float(abs(np.sum(D * (a - b)))) for D, a, b in zip(Ds, G32, G64))
    bt.close()
    # (a) one PFock per molecule, FP64, sequential, default shapes
    log("batched: one PFock per molecule")
    t_single = 0.0
    for b, D in zip(bases, Ds):
        pf = g.PFock(b)
        pf.put_den_mat(D)
        pf.commit()
        pf.compute_fock()
        device_sync()
        t0 = time.perf_counter()
        for _ in range(2):
            pf.compute_fock()
        device_sync()
        t_single += (time.perf_counter() - t0) / 2
        pf.close()
    out = {"molecules": nmol, "seed": seed, "nbf_min": min(nbf), "nbf_max": max(nbf),
           "nbf_mean": float(np.mean(nbf)), "create_s": round(t_create, 3),
           "unbatched_fp64": nmol / t_single, "batched_fp64": nmol / t_b64,
           "batched_fp64_default_shapes": nmol / t_b64_default, "batched_mixed": nmol / t_b32,
           "unit": "molecule Fock builds/s", "tau": tau,
           "mixed_vs_unbatched_fp64": t_single / t_b32, "batched_fp64_vs_unbatched": t_single / t_b64,
           "mixed_vs_batched_fp64": t_b64 / t_b32,
           "fp32_quartet_fraction": st["nq_fp32"] / max(1.0, st["nquartets"]),
           "max_abs_dF": dF, "max_first_order_dE_Eh": dE1}
    if tuned:
        out["tile_shapes"] = {"fp64": tuned["fp64"][1], "mixed": tuned["mixed"][1]}
    log("batched: builds timed; SCF energy check")
    if scf_iters > 0:
        # a molecule counts when its SCF converged (|dE| < 1e-8 Eh between
        # builds) within scf_iters builds; the reference algorithm (zero initial
        # density, trace-correcting purification, DIIS) does not converge for
        # every geometry, and those are reported, not compared
        b64 = g.Batch(bases)
        if tuned:
            b64.set_tile_shapes(tuned["fp64"][0])
        t0 = time.perf_counter()
        e64, it64 = scf_logged(b64, scf_iters, 1e-8, "batched SCF FP64")
        t_scf64 = time.perf_counter() - t0
        b64.close()
        b32 = g.Batch(bases, tau=tau)
        b32.set_purif_precision("split")
        if tuned:
            b32.set_tile_shapes(tuned["mixed"][0])
        t0 = time.perf_counter()
        e32, it32 = scf_logged(b32, scf_iters, 1e-8, "batched SCF mixed")
        t_scf = time.perf_counter() - t0
        b32.close()
        # energies are compared where both forms converged; a molecule that
        # converged in FP64 only is listed, with both runs' build counts
        conv64 = (it64 > 0) & (it64 < scf_iters)
        conv32 = (it32 > 0) & (it32 < scf_iters)
        both = conv64 & conv32
        de = np.abs(e32 - e64)
        names = [n for n, _ in geometry.batch_molecules(nmol, seed)]
        only64 = [{"molecule": names[k], "nbf": nbf[k], "builds_fp64": int(it64[k]), "builds_mixed": int(it32[k]),
                   "E_fp64": float(e64[k]), "E_mixed": float(e32[k])} for k in np.flatnonzero(conv64 & ~conv32)]
        # a molecule over 1e-6 Eh is run again, alone, in both forms: when its
        # FP64 SCF does not reproduce itself either (several SCF solutions
        # within reach of the FP64 summation-order noise), it says nothing
        # about the mixed form and is listed apart
        outliers, ill = [], set()
        for k in np.flatnonzero(both & (de > 1e-6)):
            rerun = {}
            for key, t, sh in (("fp64", 0.0, tuned["fp64"][0] if tuned else None),
                               ("mixed", tau, tuned["mixed"][0] if tuned else None)):
                bk = g.Batch([bases[k]], tau=t)
                if t > 0:
                    bk.set_purif_precision("split")
                if sh is not None:
                    bk.set_tile_shapes(sh)
                e, it = bk.scf(max_iters=scf_iters, ediff=1e-8)
                bk.close()
                rerun[key] = (float(e[0]), int(it[0]))
            rec = {"molecule": names[k], "nbf": nbf[k], "E_fp64": float(e64[k]), "builds_fp64": int(it64[k]),
                   "E_mixed": float(e32[k]), "builds_mixed": int(it32[k]),
                   "E_fp64_rerun": rerun["fp64"][0], "builds_fp64_rerun": rerun["fp64"][1],
                   "E_mixed_rerun": rerun["mixed"][0], "builds_mixed_rerun": rerun["mixed"][1]}
            rec["fp64_reproducible"] = abs(rec["E_fp64_rerun"] - rec["E_fp64"]) < 1e-6
            if not rec["fp64_reproducible"]:
                ill.add(int(k))
            outliers.append(rec)
        ok = both.copy()
        ok[list(ill)] = False
        out.update({"scf_converged": int(both.sum()), "scf_not_converged_fp64": int((~conv64).sum()),
                    "scf_converged_fp64_only": only64,
                    "scf_max_abs_dE_Eh": float(de[ok].max()) if ok.any() else None,
                    "scf_compared": int(ok.sum()),
                    "scf_over_1e-6": outliers,
                    "scf_builds_mean": float(np.mean(it64[both])) if both.any() else None,
                    "scf_fp64_s": t_scf64, "scf_mixed_s": t_scf})
    return out


def serve_leg(g, gbs, njobs, rate, seed, tau, max_rings, budget_gb, cohorts, max_iters, ref, shapes=(), large=0,
              purif="split", join_gap=1):
    """SCF requests with seeded Poisson arrivals (rate jobs/s) submitted in
    real time to the request server (mixed precision, HBM-budget admission,
    several cohorts on the device at once).  Time to converge = done time -
    submit time per job; throughput = jobs / (last done - first submit).
    With ref, the same molecules are run afterwards as one FP64 batch and
    the largest total-energy difference over jobs converged in both is
    reported."""
    from gtfock_amd import geometry
    tmp = tempfile.mkdtemp()
    mols = geometry.serve_molecules(njobs, seed, max_rings)
    bases = []
    for k, (name, atoms) in enumerate(mols):
        p = os.path.join(tmp, "j%03d.xyz" % k)
        with open(p, "w") as fh:
            fh.write(geometry.to_xyz(atoms))
        bases.append(g.BasisSet(gbs, p))
    rng = np.random.default_rng(seed + 1)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=njobs))
    srv = g.Server(hbm_budget_gb=budget_gb, max_cohorts=cohorts, max_jobs_per_cohort=64, tau=tau,
                   max_iters=max_iters, ediff=1e-8, purif=purif, join_gap=join_gap)
    for lo, hi, sh in shapes:
        srv.set_tile_shapes_range(lo, hi, sh)
    if large > 0:
        srv.set_large_slots(large)
    t0 = time.time()
    ids = []
    for b, at in zip(bases, arrivals):
        dt = t0 + at - time.time()
        if dt > 0:
            time.sleep(dt)
        ids.append(srv.submit(b))
    log("serve: %d jobs submitted" % njobs)
    done = set()
    while len(done) < njobs:
        time.sleep(0.05)
        for k, i in enumerate(ids):
            if k in done:
                continue
            r = srv.poll(i)
            if r["status"] not in (0, 1):
                done.add(k)
                log("serve: job %d (%s, nbf %d) status %d after %.2f s, %d builds" % (
                    k, mols[k][0], bases[k].nbf, r["status"], r["t_done"] - r["t_submit"], r["iters"]))
    srv.wait()
    res = [srv.poll(i) for i in ids]
    st = srv.stats()
    joins, regroups = srv.join_stats()
    srv.close()
    ok = [r for r in res if r["status"] == 2]
    lat = np.array([r["t_done"] - r["t_submit"] for r in ok])
    span = max(r["t_done"] for r in ok) - min(r["t_submit"] for r in ok)
    nbf = [b.nbf for b in bases]
    out = {"jobs": njobs, "completed": len(ok), "rate_jobs_per_s": rate, "seed": seed,
           "flake_rings_max": max_rings, "nbf_max": max(nbf), "nbf_mean": float(np.mean(nbf)),
           "tau": tau, "purification": purif, "join_gap": join_gap, "joins": joins, "regroups": regroups,
           "hbm_budget_gb": budget_gb, "max_cohorts": cohorts, "cohorts": int(st["cohorts"]),
           "peak_running_cohorts": int(st["peak_running"]), "peak_gb": st["peak_bytes"] / 2 ** 30,
           "throughput_jobs_per_s": len(ok) / span, "p50_s": float(np.percentile(lat, 50)),
           "p99_s": float(np.percentile(lat, 99)), "mean_builds": float(np.mean([r["iters"] for r in ok])),
           "not_converged": int(sum(1 for r in ok if r["iters"] >= max_iters or r["iters"] < 0))}
    if ref:
        # one FP64 SCF per distinct geometry (the stream repeats molecules;
        # water clusters of one size differ in their random placement)
        geo = [geometry.to_xyz(atoms) for _, atoms in mols]
        first = {}
        for k, x in enumerate(geo):
            first.setdefault(x, k)
        keys = sorted(first.values())
        log("serve: FP64 reference batch of the %d distinct molecules" % len(keys))
        bt = g.Batch([bases[k] for k in keys])
        e64u, it64u = scf_logged(bt, max_iters, 1e-8, "serve FP64 reference")
        bt.close()
        ref64 = {geo[k]: (e64u[i], it64u[i]) for i, k in enumerate(keys)}
        e64 = [ref64[x][0] for x in geo]
        it64 = [ref64[x][1] for x in geo]
        errs = [abs(r["energy"] - e64[k]) for k, r in enumerate(res)
                if r["status"] == 2 and 0 < r["iters"] < max_iters and 0 < it64[k] < max_iters]
        out["max_abs_dE_vs_fp64_Eh"] = float(max(errs)) if errs else None
        out["dE_compared_jobs"] = len(errs)
    return out


def big_leg(g, gbs, name, tau, steps, scf_iters, tune, purif_reps=0):
    """The mixed-precision single-molecule workload: C54H18 / cc-pVDZ (846
    basis functions).  FP64 and mixed Fock builds/s on the first SCF density
    (tile shapes from the library's tuning tables for it when tune == "db"),
    the purification of the first Fock matrix FP64 vs split BF16, and the SCF
    FP64 vs mixed (FP32 ERIs below tau, split-BF16 purification) with the
    energy difference of every iteration."""
    from gtfock_amd import geometry, tuning
    tmp = tempfile.mkdtemp()
    xyz = os.path.join(tmp, name + ".xyz")
    with open(xyz, "w") as fh:
        fh.write(geometry.to_xyz(molecule(name)))
    basis = g.BasisSet(gbs, xyz)
    sh64 = shm = None
    if tune == "db":
        sh64, _ = tuning.load(name)
        shm, _ = tuning.load(name + "_mixed")
    pf = g.PFock(basis)
    S, X, H = pf.overlap(), pf.x_matrix(), pf.core_hamiltonian()
    D = first_density(g, pf, basis)
    pf.put_den_mat(D)
    pf.commit()
    out = {"workload": "%s/cc-pVDZ" % name, "nbf": basis.nbf, "tau": tau,
           "tile_shapes": "tuning tables %s, %s_mixed" % (name, name) if sh64 is not None else "built-in defaults"}

    def timed(shapes):
        if shapes is not None:
            pf.set_tile_shapes(shapes)
        pf.compute_fock()
        device_sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            pf.compute_fock()
        device_sync()
        return steps / (time.perf_counter() - t0)

    log("%s: %d timed FP64 builds" % (name, steps))
    v64 = timed(sh64)
    F64 = pf.get_mat(g.MAT_F)
    pf.set_precision(tau)
    log("%s: %d timed mixed builds" % (name, steps))
    vm = timed(shm)
    Fm = pf.get_mat(g.MAT_F)
    nq_lo, _ = pf.mixed_counts()
    nq, _ = pf.class_work()
    pf.close()
    out.update({"fp64_builds_per_s": v64, "mixed_builds_per_s": vm, "mixed_speedup": vm / v64,
                "fp32_quartet_fraction": float(nq_lo.sum()) / max(1.0, float(nq.sum())),
                "max_abs_dF_mixed": float(np.abs(Fm - F64).max()),
                "first_order_dE_mixed_Eh": float(abs(np.sum(D * (Fm - F64))))})
    # purification of the first Fock matrix: FP64 vs split BF16
    Fo = X @ (H + F64) @ X
    pur = {}
    Ds = {}
    for mode in ("fp64", "split"):
        pu = g.Purif(basis)
        pu.set_xs(X, S)
        if mode != "fp64":
            pu.set_precision(mode)
        pu.purify(Fo)
        t0 = time.perf_counter()
        Dm, it = pu.purify(Fo)
        dt = time.perf_counter() - t0
        pur[mode] = {"seconds": dt, "gpu_seconds": pu.gpu_seconds(), "iterations": it,
                     "low_iterations": pu.fp32_iters()}
        Ds[mode] = Dm
        pu.close()
    pur["split_speedup"] = pur["fp64"]["seconds"] / pur["split"]["seconds"]
    pur["split_gpu_speedup"] = pur["fp64"]["gpu_seconds"] / max(1e-12, pur["split"]["gpu_seconds"])
    pur["max_abs_dD"] = float(np.abs(Ds["split"] - Ds["fp64"]).max())
    for mode in ("fp64", "split") if purif_reps > 0 else ():
        # further purifications of the same matrix, for kernel profiles
        pu = g.Purif(basis)
        pu.set_xs(X, S)
        if mode != "fp64":
            pu.set_precision(mode)
        for _ in range(purif_reps):
            pu.purify(Fo)
        pu.close()
    out["purification"] = pur
    if scf_iters > 0:
        log("%s: SCF FP64" % name)
        st64, stm = {}, {}
        e64, _, _ = g.scf(gbs, xyz, scf_iters, ediff=1e-8, stats=st64, shapes=sh64)
        log("%s: SCF mixed (FP32 ERIs below %g, split-BF16 purification)" % (name, tau))
        em, Dm, _ = g.scf(gbs, xyz, scf_iters, ediff=1e-8, tau=tau, purif="split", stats=stm, shapes=shm)
        n = min(len(e64), len(em))
        d = np.abs(em[:n] - e64[:n])
        tail = [k for k in range(1, n) if abs(e64[k] - e64[k - 1]) < 1e-6]
        DS = Dm @ S
        out["scf"] = {"ediff": 1e-8, "builds_fp64": len(e64), "builds_mixed": len(em),
                      "E_fp64_Eh": float(e64[-1]), "E_mixed_Eh": float(em[-1]),
                      "dE_final_Eh": float(abs(em[-1] - e64[-1])),
                      "dE_per_iteration_Eh": [float(x) for x in d],
                      "dE_converged_tail_max_Eh": float(max(d[k] for k in tail)) if tail else None,
                      "idempotency_max_abs": float(np.abs(DS @ DS - DS).max()),
                      "trace_DS_minus_nocc": float(np.trace(DS) - basis.nocc),
                      "fock_s_fp64": st64.get("fock_s"), "fock_s_mixed": stm.get("fock_s"),
                      "purif_s_fp64": st64.get("purif_s"), "purif_s_mixed": stm.get("purif_s"),
                      "split_iterations_per_purification": stm.get("purif_low_iters")}
    return out


def copy_probe(mib, reps=10):
    """Device-to-device copies of mib MiB by a plain elementwise kernel: each
    reads and writes mib MiB, the yardstick that scales rocprofv3's
    FETCH_SIZE / WRITE_SIZE to bytes."""
    import torch
    n = mib * (1 << 20) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    print(json.dumps({"copy_MiB": mib, "reps": reps, "ms_per_copy": dt * 1e3,
                      "GB_s_read_plus_write": 2 * mib * (1 << 20) / dt / 1e9}), flush=True)


def load_traffic(workload, kernel):
    """HBM bytes per launch of a kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE passes (profiles/eri_traffic_<workload>.json)."""
    path = os.path.join(ROOT, "profiles", "eri_traffic_%s.json" % workload)
    if not os.path.exists(path):
        return None
    try:
        return json.load(open(path))["kernels"].get(kernel)
    except Exception:
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c10h22")
    ap.add_argument("--tol", type=float, default=1e-11)
    ap.add_argument("--prim-tol", type=float, default=1e-15,
                    help="primitive screening threshold of the builds (the library's default; 0: none)")
    ap.add_argument("--tau", type=float, default=1e-4, help="Schwarz-bound threshold of the FP32 ERIs")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--classes", action="store_true", help="print the per-class breakdown to stderr")
    ap.add_argument("--batch", type=int, default=256, help="molecules of the batched leg (0: skip)")
    ap.add_argument("--batch-tune", type=int, default=32,
                    help="molecules of the sample batch the tile autotuner runs on (0: default shapes)")
    ap.add_argument("--autotune-reps", type=int, default=3,
                    help="launches per autotuner candidate, best taken")
    ap.add_argument("--tune", choices=["db", "live", "off"], default="db",
                    help="tile shapes: the library's tuning tables (autotuner run now where a leg has none), "
                         "the autotuner run now, or the built-in defaults")
    ap.add_argument("--write-shapes", default=None,
                    help="directory to write the shapes tuned in this run to, as tuning tables")
    ap.add_argument("--batch-seed", type=int, default=20261202)
    ap.add_argument("--batch-scf", type=int, default=30, help="SCF iterations of the batched energy check (0: skip)")
    ap.add_argument("--serve", type=int, default=32, help="jobs of the serving leg (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=50.0, help="Poisson arrival rate, jobs/s")
    ap.add_argument("--serve-rings", type=int, default=4,
                    help="largest graphene flake (rings) in the request stream (4: C96H24, 1464 basis functions)")
    ap.add_argument("--serve-budget", type=float, default=128.0, help="HBM budget of the server, GiB")
    ap.add_argument("--serve-cohorts", type=int, default=16,
                    help="cohorts the server runs at once (16: a request starts as soon as it arrives at this load; "
                         "4 queued small requests behind flake cohorts, p50 9.6 s instead of 1.3-2.5 s)")
    ap.add_argument("--serve-large-grid", type=int, default=0,
                    help="cap the persistent workgroups per class launch of flake cohorts (0: as tuned)")
    ap.add_argument("--serve-large", type=int, default=2,
                    help="of those, cohorts of molecules with 384 or more basis functions")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision leg")
    ap.add_argument("--guess", choices=["atomic", "zero"], default="atomic",
                    help="starting density of the batched and served SCFs: scf.c's atomic blocks, generated by "
                         "gtfock_amd.guess, or zero")
    ap.add_argument("--big", default="c54h18",
                    help="molecule of the mixed-precision single-molecule leg ('none': skip)")
    ap.add_argument("--big-steps", type=int, default=4, help="timed Fock builds per form of that leg")
    ap.add_argument("--big-scf", type=int, default=40,
                    help="SCF iterations (at most) of that leg's FP64 vs mixed comparison (0: skip)")
    ap.add_argument("--purif-reps", type=int, default=0,
                    help="extra purifications of that leg's first Fock matrix per mode (kernel profiles)")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only time device copies of this many MiB (the reference kernel that calibrates the "
                         "FETCH_SIZE / WRITE_SIZE counters) and exit")
    ap.add_argument("--serve-purif", choices=["fp64", "fp32", "split"], default="split",
                    help="purification products of the served SCFs")
    ap.add_argument("--serve-join-gap", type=int, default=1,
                    help="SCF iterations between admissions into running cohorts (< 0: off)")
    ap.add_argument("--scf-compare", type=int, default=0,
                    help="SCF iterations of a per-iteration FP64 vs mixed energy comparison on the workload molecule")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported "
              "(multi-GPU sharding is design-only)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.copy_probe > 0:
        copy_probe(args.copy_probe)
        return

    import gtfock_amd as g
    from gtfock_amd import flops, geometry

    gbs = os.path.join(g.DATA_DIR, "cc-pvdz.gbs")
    tmp = tempfile.mkdtemp()
    xyz = os.path.join(tmp, args.workload + ".xyz")
    with open(xyz, "w") as fh:
        fh.write(geometry.to_xyz(molecule(args.workload)))

    basis = g.BasisSet(gbs, xyz)
    log("%s: PFock_create" % args.workload)
    t0 = time.perf_counter()
    pf = g.PFock(basis, tolscr=args.tol)
    t_create = time.perf_counter() - t0
    pf.set_prim_tol(args.prim_tol)
    D = first_density(g, pf, basis)
    pf.put_den_mat(D)
    pf.commit()
    # tile shapes of the autotuner, tuned on this molecule and density
    tiles = TileShapes(args.tune, args.write_shapes, " ".join(["python", "bench.py"] + sys.argv[1:]))

    def live_molecule(what):
        log("%s: tile autotuner (%s)" % (args.workload, what))
        r = pf.autotune(reps=args.autotune_reps)
        return pf.tile_shapes(), r, "%s, first SCF density, %s" % (args.workload, what)

    sh, tune = tiles.get(args.workload, lambda: live_molecule("FP64"))
    tune = {"fp64": tune}
    if sh is not None:
        pf.set_tile_shapes(sh)
    for _ in range(args.warmup):
        pf.compute_fock()
    pf.set_timing(True)
    log("%s: %d timed FP64 builds" % (args.workload, args.steps))
    device_sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        pf.compute_fock()
    device_sync()
    t1 = time.perf_counter()
    ms_step = (t1 - t0) * 1e3 / args.steps
    value = args.steps / (t1 - t0)

    ms, launches = pf.class_times()
    nq, npq_all = pf.class_work()
    # the primitive quartets the kernels evaluate at the build's primitive
    # threshold (the algorithmic FLOPs count executed work only)
    npq, npq_lo = pf.prim_census()
    npq = npq + npq_lo
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
    roofline = {"bound": "valu_fp64", "peak_kind": "FP64 VALU, 78.6 TF/s on MI355X (the ERI kernels run on the "
                "vector ALUs, not on the matrix cores)",
                "kernel": pf.class_kernel(dom), "class": names[dom],
                "achieved": achieved / 1e12, "peak": FP64_PEAK / 1e12, "unit": "TFLOP/s",
                "frac": achieved / FP64_PEAK, "traffic": load_traffic(args.workload, pf.class_kernel(dom)),
                "traffic_source": "copy-calibrated rocprofv3 FETCH_SIZE + WRITE_SIZE per launch of this kernel "
                                  "on this workload (profiles/eri_traffic_%s.json)" % args.workload,
                "algorithmic_flops_per_launch": dom_fl, "avg_launch_ms": dom_s * 1e3,
                "build_tflops": total_fl / (ms_step * 1e-3) / 1e12,
                "prim_tol": args.prim_tol,
                "primitive_quartets_evaluated": float(npq.sum()),
                "primitive_quartets_screened_quartets": float(npq_all.sum())}

    # the same kernels each alone on the GPU (class launches serialised), after
    # the timed region: the per-kernel efficiency without sharing the device
    pf.set_serial(True)
    pf.set_timing(True)
    for _ in range(max(2, args.steps // 4)):
        pf.compute_fock()
    ms1, n1 = pf.class_times()
    pf.set_timing(False)
    pf.set_serial(False)
    iso = []
    for c in range(21):
        if n1[c]:
            t1 = ms1[c] / n1[c] * 1e-3
            iso.append((ms1[c], c, t1, flops.class_flops(c, nq[c], npq[c])))
    iso.sort(reverse=True)
    _, idom, idom_s, idom_fl = iso[0]
    roofline["isolated"] = {
        "note": "class kernels launched one after another, same density, after the timed region",
        "kernel": pf.class_kernel(idom), "class": names[idom], "avg_launch_ms": idom_s * 1e3,
        "achieved": idom_fl / idom_s / 1e12, "frac": idom_fl / idom_s / FP64_PEAK,
        "dominant_kernel_same_as_timed": idom == dom,
        "build_ms_serial": sum(x[2] for x in iso) * 1e3,
        "build_tflops_serial": total_fl / sum(x[2] for x in iso) / 1e12}

    mixed = None
    if not args.no_mixed:
        # mixed precision on the same molecule and density: FP32 ERIs for quartets
        # whose Schwarz bound is below tau, FP64 J/K accumulation
        F64 = pf.get_mat(g.MAT_F)
        pf.set_timing(False)
        pf.set_precision(args.tau)
        sh, tune["mixed"] = tiles.get(args.workload + "_mixed", lambda: live_molecule("tau %g" % args.tau))
        if sh is not None:
            pf.set_tile_shapes(sh)
        log("%s: %d timed mixed-precision builds" % (args.workload, args.steps))
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

    scf_cmp = None
    if args.scf_compare > 0:
        log("%s: SCF FP64 vs mixed, %d iterations" % (args.workload, args.scf_compare))
        b64 = g.Batch([basis])
        e64, it64, h64 = b64.scf(max_iters=args.scf_compare, ediff=0.0, history=True)
        b64.close()
        b32 = g.Batch([basis], tau=args.tau)
        e32, it32, h32 = b32.scf(max_iters=args.scf_compare, ediff=0.0, history=True)
        b32.close()
        d = np.abs(h32[0] - h64[0])
        scf_cmp = {"iterations": args.scf_compare, "tau": args.tau, "E_fp64_final": float(h64[0][-1]),
                   "dE_final_Eh": float(d[-1]), "dE_per_iteration_Eh": [float(x) for x in d],
                   "dE_last_iterations_max_Eh": float(d[len(d) // 2:].max()),
                   "E_change_last_iteration_fp64_Eh": float(abs(h64[0][-1] - h64[0][-2]))}

    # basis directory of the SCF legs: the basis file with the library's
    # atomic guesses beside it (CInt_getInitialGuess reads them from there)
    scf_gbs = gbs
    if args.guess == "atomic" and (args.serve > 0 or args.batch > 0):
        import shutil
        from gtfock_amd import guess
        log("atomic guesses for H, C, O")
        gdir = tempfile.mkdtemp()
        scf_gbs = os.path.join(gdir, os.path.basename(gbs))
        shutil.copy(gbs, scf_gbs)
        guess.write_dat(gbs, gdir, ("H", "C", "O"))

    served = None
    if args.serve > 0:
        log("serve leg: %d jobs at %.0f jobs/s" % (args.serve, args.serve_rate))
        from gtfock_amd import tuning
        # cohorts run with mixed-form shapes: those tuned on the batch
        # distribution, and for flake cohorts (384 basis functions or more)
        # those tuned on C54H18
        ranges, used = [], []
        if args.tune == "db":
            for name, lo, hi in (("batch_mixed", 0, 1 << 30), ("c54h18_mixed", 384, 1 << 30)):
                sh, _ = tuning.load(name)
                if sh is not None:
                    if lo > 0 and args.serve_large_grid > 0:
                        # flake cohorts hold fewer persistent workgroups
                        sh = [dict(x, grid=min(x["grid"], args.serve_large_grid)) for x in sh]
                        name += " (grid <= %d)" % args.serve_large_grid
                    ranges.append((lo, hi, sh))
                    used.append("%s for nbf %d.." % (name, lo))
        served = serve_leg(g, scf_gbs, args.serve, args.serve_rate, 20261203, args.tau, args.serve_rings,
                           args.serve_budget, args.serve_cohorts, 40, True, ranges, args.serve_large,
                           args.serve_purif, args.serve_join_gap)
        served["large_slots"] = args.serve_large
        served["initial_guess"] = args.guess
        served["tile_shapes"] = ("tuning tables: " + ", ".join(used)) if used else "built-in defaults"

    big = None
    if args.big != "none":
        log("mixed-precision leg: %s" % args.big)
        big = big_leg(g, gbs, args.big, args.tau, args.big_steps, args.big_scf, args.tune, args.purif_reps)

    batched = None
    if args.batch > 0:
        log("batched leg: %d molecules" % args.batch)
        batched = batch_leg(g, scf_gbs, args.batch, args.batch_seed, args.tau, max(2, args.steps // 4), args.batch_scf,
                            args.batch_tune, max(1, args.autotune_reps - 1), tiles)
        batched["initial_guess"] = args.guess

    cpu = None
    if not args.no_cpu:
        log("CPU baseline: oracle Fock builds")
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
        "tile_shapes": tune,
        "mixed": mixed,
        "mixed_c54h18": big,
        "batched": batched,
        "serve": served,
        "scf_compare": scf_cmp,
        "cpu_baseline": cpu,
    }
    pf.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
