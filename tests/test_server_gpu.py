"""Serving engine through the C ABI: budget admission, ring reuse with
wrap-around, continuous batching, and every job's answers bit-exact against
the oracle's monotone stack."""
import sys

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _log(*a):
    print("[server]", *a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    assert torch.cuda.is_available()
    return torch


def _jobs(torch, lengths, seed):
    kinds = ("random", "descending", "ascending", "dup16")
    rng = np.random.default_rng(seed)
    host, dev = [], []
    for j, n in enumerate(lengths):
        a = oracle.fill(kinds[rng.integers(len(kinds))], int(n), seed=seed * 1000 + j) \
            if n else np.zeros(0, np.int32)
        host.append(a)
        dev.append(torch.from_numpy(a).cuda() if n else None)
    return host, dev


def _answers(s, jid):
    return s.answers(jid).cpu().numpy()


def test_batch_of_jobs_all_correct(torch_mod):
    import apsep_amd as A
    torch = torch_mod
    rng = np.random.default_rng(1)
    lengths = list(rng.integers(0, 200000, size=60)) + [0, 1, 2, 3, 4, 5, 4096, 4097]
    host, dev = _jobs(torch, lengths, 1)
    torch.cuda.synchronize()
    with A.Server(1 << 30, 1 << 24, 64) as s:
        ids = []
        for k, n in enumerate(lengths):
            jid = s.submit(dev[k].data_ptr() if n else 0, int(n))
            assert jid is not None
            ids.append(jid)
            if k % 7 == 0:
                s.poll()
        s.drain()
        for k, jid in enumerate(ids):
            info = s.info(jid)
            assert info.state == A.JOB_DONE and info.done_us >= info.submit_us
            got = _answers(s, jid)
            assert np.array_equal(got, oracle.apsep(host[k])), "job %d n=%d" % (k, lengths[k])
            s.release(jid)


def test_result_alone_equals_result_in_mix(torch_mod):
    """Each job served alone (a server with nothing else) and served among
    others (submitted between them, batched with them) gives the same
    answer bits, and both are the oracle's."""
    import apsep_amd as A
    torch = torch_mod
    rng = np.random.default_rng(4)
    lengths = [int(x) for x in rng.integers(1, 300000, size=24)]
    host, dev = _jobs(torch, lengths, 4)
    torch.cuda.synchronize()
    probe = [0, 5, 11, 23]
    alone = {}
    for k in probe:
        with A.Server(1 << 28, 1 << 22, 8) as s:
            jid = s.submit(dev[k].data_ptr(), lengths[k])
            s.drain()
            alone[k] = _answers(s, jid)
            s.release(jid)
    with A.Server(1 << 30, 1 << 24, 16) as s:
        ids = []
        for k, n in enumerate(lengths):
            ids.append(s.submit(dev[k].data_ptr(), n))
            if k % 3 == 0:
                s.poll()
        s.drain()
        for k in probe:
            mixed = _answers(s, ids[k])
            _log("alone vs mix", k, lengths[k])
            assert np.array_equal(mixed, alone[k]), "job %d: mix differs from alone" % k
            assert np.array_equal(mixed, oracle.apsep(host[k])), "job %d vs oracle" % k
        for jid in ids:
            s.release(jid)


def test_inclusive_predicate_server(torch_mod):
    """A server created for "<=": every job's answers are the oracle's
    "<=" answers (ties answer to the nearest equal key)."""
    import apsep_amd as A
    torch = torch_mod
    rng = np.random.default_rng(9)
    lengths = [int(x) for x in rng.integers(1, 120000, size=20)] + [1, 2, 4096, 4097]
    host, dev = _jobs(torch, lengths, 9)
    torch.cuda.synchronize()
    with A.Server(1 << 28, 1 << 22, 16, incl=True) as s:
        ids = [s.submit(dev[k].data_ptr(), n) for k, n in enumerate(lengths)]
        assert all(j is not None for j in ids)
        s.drain()
        for k, jid in enumerate(ids):
            got = _answers(s, jid)
            assert np.array_equal(got, oracle.apsep(host[k], True)), "job %d n=%d" % (k, lengths[k])
            s.release(jid)


def test_budget_admission_and_wraparound(torch_mod):
    """A pool that holds a few jobs at a time: submissions are refused until
    finished jobs are released; slots wrap around the ring."""
    import apsep_amd as A
    torch = torch_mod
    rng = np.random.default_rng(2)
    lengths = list(rng.integers(1000, 150000, size=80))
    host, dev = _jobs(torch, lengths, 2)
    torch.cuda.synchronize()
    pool = 4 * 400000  # 400k answers
    with A.Server(pool, 400000, 16) as s:
        pending = list(range(len(lengths)))
        live = []
        refused = 0
        checked = 0
        while pending or live:
            while pending:
                k = pending[0]
                jid = s.submit(dev[k].data_ptr(), int(lengths[k]))
                if jid is None:
                    refused += 1
                    break
                live.append((k, jid))
                pending.pop(0)
            s.poll()
            # release finished jobs in order
            while live:
                k, jid = live[0]
                info = s.info(jid)
                if info.state != A.JOB_DONE:
                    if not pending:
                        s.drain()
                    break
                got = _answers(s, jid)
                assert np.array_equal(got, oracle.apsep(host[k])), "job %d" % k
                checked += 1
                s.release(jid)
                live.pop(0)
        _log("refused", refused, "checked", checked)
        assert checked == len(lengths) and refused > 0


def test_oversized_job_rejected(torch_mod):
    import apsep_amd as A
    torch = torch_mod
    x = torch.zeros(5000, dtype=torch.int32, device="cuda")
    with A.Server(1 << 20, 4096, 8) as s:
        with pytest.raises(A.ApsepError) as e:
            s.submit(x.data_ptr(), 5000)
        assert e.value.status == A.ERR_CAPACITY


def test_bulk_submit_and_release(torch_mod):
    import apsep_amd as A
    torch = torch_mod
    rng = np.random.default_rng(3)
    lengths = list(rng.integers(0, 90000, size=50))
    host, dev = _jobs(torch, lengths, 3)
    torch.cuda.synchronize()
    ptrs = np.array([d.data_ptr() if d is not None else 0 for d in dev], dtype=np.uint64)
    with A.Server(1 << 28, 1 << 22, 32) as s:
        first, acc = s.submit_many(ptrs, lengths)
        assert acc == len(lengths)
        s.drain()
        assert s.done_upto() == first + acc
        sub, done = s.job_times(first, acc)
        assert np.all(done >= sub) and np.all(sub > 0)
        for k in range(acc):
            assert np.array_equal(_answers(s, first + k), oracle.apsep(host[k])), "job %d" % k
        s.release_upto(first + acc)
        # the pool is empty again: a job as large as the batch limit fits
        big = torch.zeros(1 << 22, dtype=torch.int32, device="cuda")
        assert s.submit(big.data_ptr(), 1 << 22) is not None
        s.drain()


def test_c5_served_stream(torch_mod):
    """The served configuration: Poisson arrivals, lengths log-uniform in
    [2^10, 2^24] (numpy seed), classes 60/20/20 random / descending /
    ascending.  The pool is sized so that admission refuses jobs while the
    early ones run (a refused job waits in the client's backlog and is
    resubmitted after a poll); every job's answers are checked against the
    oracle."""
    import time
    import apsep_amd as A
    torch = torch_mod
    rng = np.random.default_rng(5)
    J = 256
    lengths = np.floor(np.exp(rng.uniform(np.log(1 << 10), np.log(1 << 24), size=J))).astype(np.int64)
    cls = rng.choice(3, size=J, p=[0.6, 0.2, 0.2])
    kinds = ("random", "descending", "ascending")
    host = [oracle.fill(kinds[cls[j]], int(lengths[j]), seed=500 + j) for j in range(J)]
    dev = [torch.from_numpy(h).cuda() for h in host]
    torch.cuda.synchronize()
    # offered jobs/s: the whole stream (about 4e8 keys) arrives within a few
    # milliseconds, faster than it is served, so the pool fills up
    rate = 100000.0
    arrive = np.cumsum(rng.exponential(1.0 / rate, size=J))
    pool = 128 << 20  # 32M answers: room for two of the largest jobs at once
    assert pool >= 2 * 4 * int(lengths.max())
    _log("C5 stream", J, "jobs,", int(lengths.sum()), "keys, pool", pool)
    ids = [None] * J
    got = [None] * J
    refused = 0
    with A.Server(pool, 1 << 26, 1024) as s:
        t0 = time.monotonic()
        nxt = done = 0
        while done < J:
            now = time.monotonic() - t0
            assert now < 300, "served stream stalled (%d of %d jobs done)" % (done, J)
            while nxt < J and arrive[nxt] <= now:
                jid = s.submit(dev[nxt].data_ptr(), int(lengths[nxt]))
                if jid is None:
                    refused += 1
                    break  # backlog: retry after the next poll
                ids[nxt] = jid
                nxt += 1
            s.poll()
            # finished jobs hand their answers back and free their slots
            while done < nxt and s.info(ids[done]).state == A.JOB_DONE:
                info = s.info(ids[done])
                assert info.n == lengths[done]
                got[done] = _answers(s, ids[done])
                s.release(ids[done])
                done += 1
    assert refused >= 1, "the pool budget never refused a job"
    _log("C5 refused", refused, "submissions")
    for j in range(J):
        if not np.array_equal(got[j], oracle.apsep(host[j])):
            raise AssertionError("C5 job %d (n=%d, %s) differs from the oracle"
                                 % (j, lengths[j], kinds[cls[j]]))
