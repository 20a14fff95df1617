"""Request loop (uc_server) on the device: every job's text equals what a
lone context produces for the same k-mers (bit-exact bytes, same seed and
bucket count), admission keeps the workers' memory within the budget, and a
job larger than the whole budget fails with UC_E_NOMEM."""
import threading

import numpy as np
import pytest

import oracle
import unitigcoarsening_amd as uc

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def jobs_mix():
    out = []
    for i, (glen, k, l) in enumerate([(20000, 31, 11), (3000, 15, 7), (60000, 63, 13), (500, 31, 11),
                                      (40000, 47, 11), (8000, 31, 9), (1, 31, 11), (100000, 31, 11)]):
        km = uc.medium_model(100 + i, glen, k) if glen > 1 else np.zeros((0, 2), np.uint32)
        out.append((km, k, l))
    return out


def test_estimate_bounds_held_bytes():
    # the last job: k - l = 2 makes most k-mers cross (x/n near its bound of
    # one), at a size where the fixed part of the estimate is under a tenth
    for glen, k, l in [(5000, 15, 7), (50000, 31, 11), (200000, 31, 11), (80000, 63, 11), (120000, 47, 11),
                       (3_000_000, 15, 13)]:
        km = uc.medium_model(7, glen, k)
        log(f"estimate n={km.shape[0]} k={k} l={l}")
        with uc.Compactor() as c:
            c.compact(km, k, l)
            held = c.held_bytes()
            est = uc.job_bytes_estimate(km.shape[0], k)
            b = c.bucket_info()
            log(f"  x/n = {b.n_cross / max(b.n_kmers, 1):.2f}, held {held}, estimate {est}")
            assert 0 < held <= est, (glen, k, held, est)
            if glen >= 3_000_000:
                assert (64 << 20) < est // 10 and b.n_cross > b.n_kmers // 2
            c.trim()
            assert c.held_bytes() == 0
            # the context works again after a trim
            text = c.compact(km, k, l)
            assert text.count(b"\n") > 0


def test_server_matches_lone_context():
    mix = jobs_mix()
    expect = []
    with uc.Compactor() as c:
        for km, k, l in mix:
            expect.append(c.compact(km, k, l))
    with uc.Server(n_slots=3) as srv:
        ids = [srv.submit(km, k, l) for km, k, l in mix]
        log(f"submitted {len(ids)} jobs")
        for (km, k, l), jid, want in zip(mix, ids, expect):
            text, qms, rms = srv.result(jid)
            assert text == want, (km.shape, k, l)
            assert qms >= 0 and rms >= 0
        assert srv.stats()["jobs_done"] == len(mix)


def test_server_texts_match_oracle():
    rng = np.random.default_rng(3)
    with uc.Server(n_slots=2) as srv:
        cases = []
        for i in range(6):
            k = int(rng.choice([15, 31, 63]))
            km = uc.synth_kmers(900 + i, int(rng.integers(100, 3000)), k, repeat_frac=0.1)
            cases.append((km, k, srv.submit(km, k, 7 if k == 15 else 11)))
        for km, k, jid in cases:
            text, _, _ = srv.result(jid)
            want, _, _ = oracle.textbook(km, k)
            assert uc.sorted_lines(text) == uc.sorted_lines(want)


def test_server_budget_admission_and_threads():
    mix = jobs_mix()
    largest = max(uc.job_bytes_estimate(km.shape[0], k) for km, k, _ in mix)
    budget = int(largest * 1.5)
    expect = []
    with uc.Compactor() as c:
        for km, k, l in mix:
            expect.append(c.compact(km, k, l))
    results = {}
    with uc.Server(n_slots=4, hbm_budget=budget) as srv:
        def client(idx):
            for i in idx:
                km, k, l = mix[i]
                results[i] = srv.result(srv.submit(km, k, l))[0]
        ts = [threading.Thread(target=client, args=(list(range(j, len(mix), 3)),)) for j in range(3)]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        st = srv.stats()
        log(f"budget {budget} peak {st['peak_held_bytes']}")
        assert st["jobs_done"] == len(mix)
        assert st["peak_held_bytes"] <= budget
    for i, want in enumerate(expect):
        assert results[i] == want, i


def test_server_rejects_job_over_budget():
    km = uc.medium_model(5, 50000, 31)
    est = uc.job_bytes_estimate(km.shape[0], 31)
    with uc.Server(n_slots=1, hbm_budget=est // 2) as srv:
        jid = srv.submit(km, 31, 11)
        assert srv.poll(jid) == uc.UC_JOB_FAILED
        with pytest.raises(uc.UcError) as e:
            srv.result(jid)
        assert e.value.status == uc.UC_E_NOMEM
        small = uc.medium_model(6, 100, 31)
        if uc.job_bytes_estimate(small.shape[0], 31) <= est // 2:
            text, _, _ = srv.result(srv.submit(small, 31, 11))
            with uc.Compactor() as c:
                assert text == c.compact(small, 31, 11)
