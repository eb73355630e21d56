"""Device-resident contigs and the continuous-batching serving loop (GPU)."""
import numpy as np
import pytest

import mhm2_staging_amd as m
from golden.kat_cases import MHM2_DEFAULT_SCORING
from test_gpu_parity import assert_same, gpu_table, oracle_table

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def serving_set():
    cs = m.contigs(nctgs=4000, seed=51)
    res = m.ResidentContigs(cs, device=0)
    yield cs, res
    res.close()


def test_resident_windows_equal_string_windows_and_oracle(gpu_aligner, serving_set):
    cs, res = serving_set
    wb = m.window_pairs(cs, 3000, seed=52)
    sb = wb.as_strings(cs)
    want = oracle_table(sb, MHM2_DEFAULT_SCORING)
    al = gpu_aligner
    al.set_scoring(MHM2_DEFAULT_SCORING)
    al.set_orientation(True)
    for mode in (m.MODE_FAST, m.MODE_PLAIN):
        al.set_mode(mode)
        print(f"[gpu] resident windows mode={mode}", flush=True)
        al.stage_windows(res, wb)
        al.align_staged()
        al.wait()
        assert_same(gpu_table(al.results()), want, sb)
        al.stage(sb)
        al.align_staged()
        al.wait()
        assert_same(gpu_table(al.results()), want, sb)
    al.set_mode(m.MODE_FAST)


def test_server_results_same_alone_and_in_the_mix(serving_set):
    cs, res = serving_set
    sizes = [1000, 7000, 2500, 15000, 600, 4000]
    reqs = [m.window_pairs(cs, n, seed=60 + i) for i, n in enumerate(sizes)]
    # alone: a fresh server, one request at a time
    alone = []
    srv = m.Server(res, scoring=MHM2_DEFAULT_SCORING, slots=1)
    for r in reqs:
        rid = srv.submit(r)
        alone.append(srv.results(rid)[0])
    srv.close()
    # in the mix: all submitted back to back on 4 slots, overlapping on the device
    srv = m.Server(res, scoring=MHM2_DEFAULT_SCORING, slots=4)
    ids = [srv.submit(r) for r in reqs]
    mixed = [srv.results(i) for i in ids]
    srv.close()
    overlapped = sum(1 for a, b in zip(mixed, mixed[1:]) if b[1]["admit_s"] < a[1]["done_s"])
    assert overlapped > 0, "no two requests were on the device together"
    for r, a, (b, t) in zip(reqs, alone, mixed):
        assert (a == b).all()
        assert t["submit_s"] <= t["admit_s"] <= t["launched_s"] <= t["finished_s"] <= t["done_s"]
        assert t["device_ms"] > 0
    # and they are the SSW answers
    for r, a in zip(reqs[:3], alone[:3]):
        sb = r.as_strings(cs)
        assert_same(gpu_table(a), oracle_table(sb, MHM2_DEFAULT_SCORING), sb)


def test_server_hbm_budget_admission_in_order(serving_set):
    """A budget that fits one request at a time: requests queue, run one by one, in arrival order."""
    cs, res = serving_set
    reqs = [m.window_pairs(cs, 3000, seed=80 + i) for i in range(4)]
    # the budget: one and a half times one request's admission estimate
    srv = m.Server(res, scoring=MHM2_DEFAULT_SCORING, slots=1)
    est = max(srv.results(srv.submit(r))[1]["job_est_bytes"] for r in reqs)
    srv.close()
    srv = m.Server(res, scoring=MHM2_DEFAULT_SCORING, slots=4, hbm_budget=int(1.5 * est))
    ids = [srv.submit(r) for r in reqs]
    outs = [srv.results(i) for i in ids]
    srv.close()
    admits = [t["admit_s"] for _, t in outs]
    dones = [t["done_s"] for _, t in outs]
    assert admits == sorted(admits)
    for k in range(1, len(outs)):
        assert admits[k] >= dones[k - 1] - 1e-6  # serialized by the budget
    for r, (o, _) in zip(reqs[:2], outs[:2]):
        sb = r.as_strings(cs)
        assert_same(gpu_table(o), oracle_table(sb, MHM2_DEFAULT_SCORING), sb)


def test_server_merges_queued_requests(serving_set):
    """Requests that queue while a job runs share the next job; each one's
    results equal the request aligned alone, and are SSW's."""
    cs, res = serving_set
    sizes = [20000] + [300 + 97 * i for i in range(15)]
    reqs = [m.window_pairs(cs, n, seed=100 + i) for i, n in enumerate(sizes)]
    al = m.Aligner(device=0)
    al.set_scoring(MHM2_DEFAULT_SCORING)
    alone = []
    for r in reqs:
        al.stage_windows(res, r)
        al.align_staged()
        al.wait()
        alone.append(al.results())
    al.close()
    srv = m.Server(res, scoring=MHM2_DEFAULT_SCORING, slots=1)
    ids = [srv.submit(r) for r in reqs]
    outs = [srv.results(i) for i in ids]
    srv.close()
    admits = [t["admit_s"] for _, t in outs]
    assert admits == sorted(admits)
    shared = sum(1 for a, b in zip(admits, admits[1:]) if a == b)
    print(f"[gpu] {shared} requests admitted together with their predecessor", flush=True)
    assert shared > 0, "no queued requests were merged"
    for a, (b, _) in zip(alone, outs):
        assert (gpu_table(a) == gpu_table(b)).all()
    for r, (o, _) in list(zip(reqs, outs))[1:4]:
        sb = r.as_strings(cs)
        assert_same(gpu_table(o), oracle_table(sb, MHM2_DEFAULT_SCORING), sb)


def test_server_score2_and_cigars_match_the_batch_api(gpu_aligner, serving_set):
    """With second-best scores and CIGARs on, each request's extras equal what
    the batch API computes for the same windows, also when queued requests
    share one job (one slot, submitted back to back)."""
    cs, res = serving_set
    reqs = [m.window_pairs(cs, n, seed=90 + i) for i, n in enumerate([800, 3000, 150, 2200])]
    al = gpu_aligner
    al.set_scoring(MHM2_DEFAULT_SCORING)
    al.set_mode(m.MODE_FAST)
    al.set_orientation(True)
    al.set_score2(True)
    al.set_cigar(True)
    want = []
    try:
        for r in reqs:
            al.stage_windows(res, r)
            al.align_staged()
            al.wait()
            want.append((al.results(), al.score2(), al.cigars()))
    finally:
        al.set_score2(False)
        al.set_cigar(False)
    srv = m.Server(res, scoring=MHM2_DEFAULT_SCORING, slots=1)
    srv.set_outputs(score2=True, cigar=True)
    ids = [srv.submit(r) for r in reqs]
    for r, rid, (wres, ws2, (wm, woff, wops)) in zip(reqs, ids, want):
        got, _ = srv.results(rid)
        s2, mism, off, ops = srv.extras(rid, r.npairs)
        assert (got == wres).all()
        assert np.array_equal(s2, ws2)
        assert np.array_equal(mism, wm) and np.array_equal(off, woff) and np.array_equal(ops, wops)
    # outputs off again: plain results, no extras held
    srv.set_outputs(False, False)
    rid = srv.submit(reqs[0])
    got, _ = srv.results(rid)
    assert (got == want[0][0]).all()
    assert srv.extras(rid, reqs[0].npairs) == (None, None, None, None)
    srv.close()


@pytest.mark.parametrize("npairs,cigar", [(1000, False), (20000, False), (5000, True)])
def test_server_job_estimate_bounds_its_allocations(serving_set, npairs, cigar):
    """The admission estimate of a job (the bytes counted against hbm_budget while it runs) bounds what the
    same batch allocates on a context that holds no buffers yet, and follows the allocation rule closely."""
    cs, res = serving_set
    r = m.window_pairs(cs, npairs, seed=300 + npairs)
    srv = m.Server(res, scoring=MHM2_DEFAULT_SCORING, slots=1)
    srv.set_outputs(cigar=cigar)
    _, t = srv.results(srv.submit(r))
    srv.close()
    al = m.Aligner(device=0, scoring=MHM2_DEFAULT_SCORING)
    al.set_cigar(cigar)
    al.stage_windows(res, r)
    al.align_staged()
    al.wait()
    st = al.stats()
    al.close()
    print(f"[gpu] {npairs} pairs cigar={cigar}: allocated {st['device_bytes']} estimate {t['job_est_bytes']:.0f}",
          flush=True)
    assert st["cigar_reruns"] == 0
    assert 0 < st["device_bytes"] <= t["job_est_bytes"]
    assert t["job_est_bytes"] <= 1.3 * st["device_bytes"] + (2 << 20)


def test_server_rejects_a_request_larger_than_its_budget(serving_set):
    cs, res = serving_set
    srv = m.Server(res, scoring=MHM2_DEFAULT_SCORING, slots=1, hbm_budget=2 << 20)
    with pytest.raises(m.MhmswError) as e:
        srv.submit(m.window_pairs(cs, 20000, seed=7))
    assert e.value.code == -1 and "more than the server's HBM budget" in str(e.value)
    # a small one still runs
    small = m.window_pairs(cs, 200, seed=8)
    got, _ = srv.results(srv.submit(small))
    srv.close()
    sb = small.as_strings(cs)
    assert_same(gpu_table(got), oracle_table(sb, MHM2_DEFAULT_SCORING), sb)
