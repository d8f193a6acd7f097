"""Host tests of the request server's scheduling (gtfock_amd/csrc/serve.cpp)
built against a stand-in batch backend (tests/sched_stub.cpp, where a
molecule's SCF runs in 10 ms iterations): size-class cohorts, per-job
publication, the memory budget, the cohort limit, jobs that can never fit,
cancellation, and continuous batching (jobs joining running cohorts between
iterations, finished jobs returning memory early).  No GPU is used."""
import ctypes
import os
import subprocess
import time

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("sched") / "libsched.so")
    subprocess.check_call(["g++", "-O1", "-std=c++17", "-fPIC", "-shared", "-o", out,
                           os.path.join(ROOT, "gtfock_amd", "csrc", "serve.cpp"), os.path.join(HERE, "sched_stub.cpp"),
                           "-lpthread"])
    L = ctypes.CDLL(out)
    vp, dp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)
    L.gtf_server_create.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_int, ctypes.c_double, ctypes.POINTER(vp)]
    L.gtf_server_submit.argtypes = [vp, vp]
    L.gtf_server_submit.restype = ctypes.c_long
    L.gtf_server_poll.argtypes = [vp, ctypes.c_long, dp]
    L.gtf_server_wait.argtypes = [vp]
    L.gtf_server_stats.argtypes = [vp, dp]
    L.gtf_server_destroy.argtypes = [vp]
    L.stub_molecule.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double]
    L.stub_molecule.restype = vp
    L.stub_cohort_sizes.argtypes = [ctypes.POINTER(ctypes.c_int), ctypes.c_int]
    L.gtf_server_set_join_gap.argtypes = [vp, ctypes.c_int]
    L.gtf_server_join_stats.argtypes = [vp, ctypes.POINTER(ctypes.c_long), ctypes.POINTER(ctypes.c_long)]
    return L


def server(L, budget, cohorts, jobs_per_cohort):
    h = ctypes.c_void_p()
    assert L.gtf_server_create(budget, cohorts, jobs_per_cohort, 1e-11, 0.0, 30, 1e-8, ctypes.byref(h)) == 0
    return h


def poll(L, h, i):
    out = (ctypes.c_double * 7)()
    st = L.gtf_server_poll(h, i, out)
    return st, list(out)


def stats(L, h):
    out = (ctypes.c_double * 8)()
    L.gtf_server_stats(h, out)
    return dict(zip(["jobs", "queued", "running", "cohorts", "peak_running", "bytes", "peak_bytes", "budget"], out))


def test_size_classes_and_early_results(lib):
    L = lib
    L.stub_reset()
    h = server(L, 1e12, 4, 64)
    small = [L.stub_molecule(30, 0.05, -1.0 - k) for k in range(6)]
    big = [L.stub_molecule(800, 0.6, -100.0 - k) for k in range(2)]
    ids = [L.gtf_server_submit(h, b) for b in big[:1] + small + big[1:]]
    time.sleep(0.3)
    # the small jobs, in a cohort of their own, are done while the large ones run
    sts = [poll(L, h, i)[0] for i in ids]
    assert sts[1:7] == [2] * 6
    assert sts[0] in (0, 1) and sts[7] in (0, 1)
    L.gtf_server_wait(h)
    res = [poll(L, h, i) for i in ids]
    assert all(st == 2 for st, _ in res)
    assert [r[0] for _, r in res] == [-100.0] + [-1.0 - k for k in range(6)] + [-101.0]
    sizes = (ctypes.c_int * 16)()
    n = L.stub_cohort_sizes(sizes, 16)
    # molecules of one size class share cohorts; none mixes 30 and 800 functions
    assert sum(sizes[:n]) == 8
    cohort_of = {i: int(poll(L, h, i)[1][5]) for i in ids}
    assert not ({cohort_of[i] for i in ids[1:7]} & {cohort_of[ids[0]], cohort_of[ids[7]]})
    L.gtf_server_destroy(h)


def test_budget_and_cohort_limit(lib):
    L = lib
    L.stub_reset()
    one = 1000.0 * 100 * 100
    h = server(L, 2.5 * one, 2, 64)  # room for two jobs at a time
    mols = [L.stub_molecule(100, 0.05, -float(k)) for k in range(7)]
    ids = [L.gtf_server_submit(h, b) for b in mols]
    L.gtf_server_wait(h)
    st = stats(L, h)
    assert all(poll(L, h, i)[0] == 2 for i in ids)
    assert st["peak_bytes"] <= st["budget"]
    assert st["peak_running"] <= 2 and L.stub_max_alive() <= 2
    assert st["bytes"] == 0.0
    L.gtf_server_destroy(h)


def test_job_larger_than_budget_fails_alone(lib):
    L = lib
    L.stub_reset()
    h = server(L, 1000.0 * 200 * 200, 4, 64)
    ok = L.stub_molecule(100, 0.02, -1.0)
    huge = L.stub_molecule(400, 0.02, -2.0)
    ids = [L.gtf_server_submit(h, huge), L.gtf_server_submit(h, ok)]
    L.gtf_server_wait(h)
    assert poll(L, h, ids[0])[0] == -2
    assert poll(L, h, ids[1])[0] == 2
    L.gtf_server_destroy(h)


def test_destroy_cancels_queued_jobs(lib):
    L = lib
    L.stub_reset()
    h = server(L, 1e12, 1, 1)  # one job at a time
    mols = [L.stub_molecule(50, 0.2, -1.0) for _ in range(4)]
    ids = [L.gtf_server_submit(h, b) for b in mols]
    time.sleep(0.05)
    L.gtf_server_destroy(h)  # waits for the running job, cancels the rest
    assert L.stub_created() == 1


def test_large_molecules_take_at_most_half_the_slots(lib):
    L = lib
    L.stub_reset()
    h = server(L, 1e13, 4, 1)  # one job per cohort
    big = [L.stub_molecule(800, 0.4, -10.0) for _ in range(4)]
    small = [L.stub_molecule(30, 0.05, -1.0) for _ in range(4)]
    t0 = time.time()
    ids_big = [L.gtf_server_submit(h, b) for b in big]
    ids_small = [L.gtf_server_submit(h, b) for b in small]
    L.gtf_server_wait(h)
    done_small = max(poll(L, h, i)[1][4] for i in ids_small) - t0
    done_big = sorted(poll(L, h, i)[1][4] - t0 for i in ids_big)
    # two large cohorts at a time; the small jobs use the other two slots, so
    # they are all done before the first large job ends, and the large jobs
    # end in two waves one large SCF time (0.4 s) apart.  Ordering and gaps,
    # not absolute times, so a busy host does not fail the test.
    assert done_small < done_big[0]
    assert done_big[1] - done_big[0] < 0.15
    assert done_big[2] - done_big[1] > 0.25
    assert stats(L, h)["peak_running"] <= 4
    L.gtf_server_destroy(h)


def _vm_kib():
    with open("/proc/self/status") as f:
        for line in f:
            if line.startswith("VmSize:"):
                return int(line.split()[1])
    return 0


def test_ended_cohort_threads_are_reclaimed(lib):
    # one cohort per job, one at a time: 64 cohort threads start and end over
    # the server's life.  A thread that has ended but is never joined keeps
    # its stack mapped (8 MiB by default), so reclaiming them keeps the
    # process's address space flat.
    L = lib
    L.stub_reset()
    h = server(L, 1e12, 1, 1)
    assert L.gtf_server_set_join_gap(h, -1) == 0  # a cohort never takes a second job
    mols = [L.stub_molecule(30, 0.001, -1.0) for _ in range(4)]
    ids = [L.gtf_server_submit(h, mols[k % 4]) for k in range(8)]
    L.gtf_server_wait(h)
    base = _vm_kib()
    ids += [L.gtf_server_submit(h, mols[k % 4]) for k in range(64)]
    L.gtf_server_wait(h)
    grown = _vm_kib() - base
    assert all(poll(L, h, i)[0] == 2 for i in ids)
    assert stats(L, h)["cohorts"] == 72
    assert grown < 64 * 1024, grown  # far below 64 unjoined 8 MiB stacks
    L.gtf_server_destroy(h)


def test_tile_shapes_reach_the_cohorts_that_start_after(lib):
    L = lib
    L.stub_reset()
    L.gtf_server_set_tile_shapes.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
    h = server(L, 1e12, 1, 1)  # one job per cohort
    mols = [L.stub_molecule(30, 0.01, -1.0) for _ in range(4)]
    ids = [L.gtf_server_submit(h, mols[0])]
    L.gtf_server_wait(h)
    assert L.stub_shapes_set() == 0  # built-in defaults: nothing set
    shapes = (ctypes.c_int * (21 * 7))(*([777, 64, 65536, 2048, 0, 0, 1] * 21))
    assert L.gtf_server_set_tile_shapes(h, shapes) == 0
    ids += [L.gtf_server_submit(h, m) for m in mols[1:3]]
    L.gtf_server_wait(h)
    assert L.stub_shapes_set() == 2 and L.stub_shape0_units() == 777
    assert L.gtf_server_set_tile_shapes(h, None) == 0
    ids.append(L.gtf_server_submit(h, mols[3]))
    L.gtf_server_wait(h)
    assert L.stub_shapes_set() == 2
    assert all(poll(L, h, i)[0] == 2 for i in ids)
    L.gtf_server_destroy(h)


def test_tile_shapes_by_molecule_size(lib):
    L = lib
    L.stub_reset()
    L.gtf_server_set_tile_shapes.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
    L.gtf_server_set_tile_shapes_range.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                                   ctypes.POINTER(ctypes.c_int)]
    L.stub_shape_units.argtypes = [ctypes.POINTER(ctypes.c_int), ctypes.c_int]

    def shapes(units):
        return (ctypes.c_int * (21 * 7))(*([units, 64, 65536, 2048, 0, 0, 1] * 21))

    h = server(L, 1e12, 1, 1)  # one job per cohort, one cohort at a time
    assert L.gtf_server_set_tile_shapes(h, shapes(111)) == 0            # every size
    assert L.gtf_server_set_tile_shapes_range(h, 0, 47, shapes(222)) == 0
    assert L.gtf_server_set_tile_shapes_range(h, 400, 100000, shapes(333)) == 0
    assert L.gtf_server_set_tile_shapes_range(h, 50, 10, shapes(444)) == -3  # empty range
    # invalid shapes are refused when set, not when a cohort starts
    assert L.gtf_server_set_tile_shapes_range(h, 0, 10, shapes(0)) == -3
    bad = shapes(555)
    bad[7 * 3 + 6] = 2  # fp32 flag of class 3
    assert L.gtf_server_set_tile_shapes_range(h, 0, 10, bad) == -3
    assert L.gtf_server_set_tile_shapes(h, bad) == -3
    mols = [L.stub_molecule(n, 0.01, -1.0) for n in (30, 100, 800)]
    for m in mols:
        L.gtf_server_submit(h, m)
        L.gtf_server_wait(h)
    out = (ctypes.c_int * 8)()
    n = L.stub_shape_units(out, 8)
    assert list(out[:n]) == [222, 111, 333]
    # clearing: later cohorts run with the built-in defaults
    assert L.gtf_server_set_tile_shapes(h, None) == 0
    L.gtf_server_submit(h, mols[2])
    L.gtf_server_wait(h)
    assert L.stub_shape_units(out, 8) == 3
    L.gtf_server_destroy(h)


def test_large_slots_setting(lib):
    """With 8 cohort slots and one large slot, large molecules run one at a
    time while small ones use the other slots."""
    L = lib
    L.stub_reset()
    L.gtf_server_set_large_slots.argtypes = [ctypes.c_void_p, ctypes.c_int]
    h = server(L, 1e13, 8, 1)
    assert L.gtf_server_set_large_slots(h, 0) == -3
    assert L.gtf_server_set_large_slots(h, 1) == 0
    big = [L.stub_molecule(800, 0.2, -10.0) for _ in range(3)]
    small = [L.stub_molecule(30, 0.05, -1.0) for _ in range(4)]
    ids_big = [L.gtf_server_submit(h, b) for b in big]
    ids_small = [L.gtf_server_submit(h, b) for b in small]
    L.gtf_server_wait(h)
    spans = sorted((poll(L, h, i)[1][3], poll(L, h, i)[1][4]) for i in ids_big)
    # one large cohort at a time: each starts after the previous one ended
    assert spans[1][0] >= spans[0][1] - 0.01 and spans[2][0] >= spans[1][1] - 0.01
    done_small = max(poll(L, h, i)[1][4] for i in ids_small)
    assert done_small < spans[0][1]  # the small jobs did not wait for them
    L.gtf_server_destroy(h)


def join_stats(L, h):
    j, r = ctypes.c_long(), ctypes.c_long()
    assert L.gtf_server_join_stats(h, ctypes.byref(j), ctypes.byref(r)) == 0
    return j.value, r.value


def test_jobs_join_a_running_cohort(lib):
    """Jobs that arrive while a cohort of their size class runs join it at
    its next iteration instead of waiting for it or opening a new cohort."""
    L = lib
    L.stub_reset()
    h = server(L, 1e12, 4, 64)
    first = L.stub_molecule(30, 0.4, -1.0)
    later = [L.stub_molecule(30, 0.1, -2.0 - k) for k in range(3)]
    ids = [L.gtf_server_submit(h, first)]
    time.sleep(0.1)
    ids += [L.gtf_server_submit(h, m) for m in later]
    L.gtf_server_wait(h)
    res = [poll(L, h, i) for i in ids]
    assert all(st == 2 for st, _ in res)
    assert [r[0] for _, r in res] == [-1.0, -2.0, -3.0, -4.0]
    cohorts = {int(r[5]) for _, r in res}
    assert cohorts == {0} and stats(L, h)["cohorts"] == 1
    joins, regroups = join_stats(L, h)
    assert joins == 3 and 1 <= regroups <= 3
    # the newcomers ran beside the first job: they started and ended before it ended
    t_done_first = res[0][1][4]
    assert all(r[3] < t_done_first and r[4] < t_done_first for _, r in res[1:])
    L.gtf_server_destroy(h)


def test_finished_jobs_return_memory_between_iterations(lib):
    """Finished jobs' memory returns to the budget between iterations, not
    when their cohort ends: a queued job that needs it starts while the long
    job of that cohort still runs, and the budget holds throughout (both
    batches of a regroup counted)."""
    L = lib
    L.stub_reset()
    one = 1000.0 * 30 * 30
    h = server(L, 5.0 * one, 4, 64)
    first = [L.stub_molecule(30, 0.05, -1.0 - k) for k in range(3)] + [L.stub_molecule(30, 0.6, -9.0)]
    ids = [L.gtf_server_submit(h, m) for m in first]
    time.sleep(0.02)
    late = L.stub_molecule(42, 0.05, -5.0)  # about two jobs' memory: fits only after the short jobs leave
    ids.append(L.gtf_server_submit(h, late))
    L.gtf_server_wait(h)
    res = [poll(L, h, i) for i in ids]
    st = stats(L, h)
    assert all(s_ == 2 for s_, _ in res)
    assert st["peak_bytes"] <= st["budget"]
    t_done_long = res[3][1][4]
    assert res[4][1][3] < t_done_long  # the late job started while the long job ran
    assert join_stats(L, h)[1] >= 1
    L.gtf_server_destroy(h)


def test_join_gap_off_keeps_cohorts_closed(lib):
    L = lib
    L.stub_reset()
    h = server(L, 1e12, 4, 64)
    assert L.gtf_server_set_join_gap(h, -1) == 0
    first = L.stub_molecule(30, 0.2, -1.0)
    later = L.stub_molecule(30, 0.05, -2.0)
    ids = [L.gtf_server_submit(h, first)]
    time.sleep(0.05)
    ids.append(L.gtf_server_submit(h, later))
    L.gtf_server_wait(h)
    assert join_stats(L, h)[0] == 0
    assert int(poll(L, h, ids[0])[1][5]) != int(poll(L, h, ids[1])[1][5])
    L.gtf_server_destroy(h)


def test_large_cohorts_yield_to_small_ones(lib):
    """While small-molecule cohorts run, a large-molecule cohort waits before
    each iteration for at most large_yield times its last iteration: with
    the default (2) a large SCF shorter than the small ones ends after them,
    with 0 it runs at full speed and ends first; it always completes."""
    L = lib
    L.gtf_server_set_large_yield.argtypes = [ctypes.c_void_p, ctypes.c_double]
    out = {}
    for factor in (0.0, 2.0):
        L.stub_reset()
        h = server(L, 1e13, 4, 1)
        assert L.gtf_server_set_large_yield(h, -1.0) == -3
        assert L.gtf_server_set_large_yield(h, factor) == 0
        t0 = time.time()
        big = L.gtf_server_submit(h, L.stub_molecule(800, 0.2, -10.0))
        small = [L.gtf_server_submit(h, L.stub_molecule(30, 0.4, -1.0)) for _ in range(2)]
        L.gtf_server_wait(h)
        st, r = poll(L, h, big)
        assert st == 2
        out[factor] = (r[4] - t0, max(poll(L, h, i)[1][4] for i in small) - t0)
        L.gtf_server_destroy(h)
    print(out)
    assert out[0.0][0] < out[0.0][1]
    assert out[2.0][0] > out[2.0][1]

