"""GPU: the serving loop -- resident tables, budgeted admission, continuous
batching.  Each request's result is the oracle's, alone and in the mix."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def words(p):
    return (p["r_rid"].astype(np.uint64) << np.uint64(32)) | p["s_rid"].astype(np.uint64)


def make_workload(sj, seed, ntab=4, nreq=40):
    rng = np.random.default_rng(seed)
    tables = [sj.gen_r(n, 500 + i) for i, n in enumerate([1 << 12, 1 << 14, 1 << 16, 50_000][:ntab])]
    reqs = []
    for i in range(nreq):
        t = int(rng.integers(0, ntab))
        n = int(2 ** rng.uniform(8, 16))
        reqs.append((t, sj.gen_s(n, tables[t], float(rng.uniform(0, 1.5)), 900 + i)))
    return tables, reqs


def test_alone_equals_mix_equals_oracle(gpu, orc):
    sj = gpu
    tables, reqs = make_workload(sj, 1)
    alone = []
    sv = sj.Server(1 << 30, slots=1)
    tids = [sv.add_table(R) for R in tables]
    for t, S in reqs:
        out, _ = sv.wait(sv.submit(tids[t], S))
        alone.append(np.sort(words(out)))
    sv.close()
    sv = sj.Server(1 << 30, slots=3)
    tids = [sv.add_table(R) for R in tables]
    ids = [sv.submit(tids[t], S) for t, S in reqs]   # burst: they queue and batch
    mixed = [np.sort(words(sv.wait(i)[0])) for i in ids]
    st = sv.stats()
    sv.close()
    assert st["requests"] == len(reqs)
    assert st["batches"] < len(reqs)   # requests were batched together
    for (t, S), a, m in zip(reqs, alone, mixed):
        want = words(orc.expected_pairs(tables[t], S))
        assert np.array_equal(a, want)
        assert np.array_equal(m, want)


def test_budget_admission(gpu, orc):
    sj = gpu
    tables, reqs = make_workload(sj, 2, ntab=2, nreq=12)
    sv = sj.Server(4 << 20, slots=2)   # room for about one small request at a time
    tids = [sv.add_table(R) for R in tables]
    big = sj.gen_s(1 << 20, tables[0], 0.0, 7)
    with pytest.raises(sj.SpjError) as e:
        sv.submit(tids[0], big)
    assert e.value.code == -7
    ids = [sv.submit(tids[t], S[:2048]) for t, S in reqs]
    for (t, S), i in zip(reqs, ids):
        out, lat = sv.wait(i)
        assert lat >= 0
        assert np.array_equal(np.sort(words(out)), words(orc.expected_pairs(tables[t], S[:2048])))
    sv.close()
