"""GPU parity at the configured sizes of the batched and serving workloads:
every instance / request against the oracle's closed-form checksum."""
import ctypes
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_batched_4096_mixed_joins_match_the_oracle(gpu, orc):
    """4096 instances, |R| log-uniform in [2^10, 2^20] (unique keys), R:S from
    {1, 2, 4, 8}, S Zipf U[0, 1.25]: one batched call and one call per
    instance both give each instance the oracle's count and pair checksum."""
    sj = gpu
    n = 4096
    rng = np.random.default_rng(4096)
    nRs = np.floor(2.0 ** rng.uniform(10, 20, n)).astype(np.int64)
    nSs = nRs * rng.choice([1, 2, 4, 8], n)
    zipfs = rng.uniform(0, 1.25, n)
    totR, totS = int(nRs.sum()), int(nSs.sum())
    offR = np.concatenate([[0], np.cumsum(nRs)[:-1]])
    offS = np.concatenate([[0], np.cumsum(nSs)[:-1]])
    print(f"case batched 4096: sum |R| {totR}, sum |S| {totS}", flush=True)
    dR, dS, dO = sj.DeviceArray(totR * 8), sj.DeviceArray(totS * 8), sj.DeviceArray(totS * 8)
    want = []
    t0 = time.time()
    try:
        for i in range(n):
            R = sj.gen_r(int(nRs[i]), 10_000 + i)
            S = sj.gen_s(int(nSs[i]), R, float(zipfs[i]), 20_000 + i)
            sj.check(sj.lib().spj_h2d(dR.ptr + int(offR[i]) * 8, R.ctypes.data, R.nbytes))
            sj.check(sj.lib().spj_h2d(dS.ptr + int(offS[i]) * 8, S.ctypes.data, S.nbytes))
            want.append(orc.join_checksum(R, S))
        print(f"  data and oracle in {time.time() - t0:.1f}s", flush=True)
        P = ctypes.c_void_p
        rp = (P * n)(*[dR.ptr + int(o) * 8 for o in offR])
        sp = (P * n)(*[dS.ptr + int(o) * 8 for o in offS])
        op = (P * n)(*[dO.ptr + int(o) * 8 for o in offS])
        nr = (ctypes.c_uint64 * n)(*[int(x) for x in nRs])
        ns = (ctypes.c_uint64 * n)(*[int(x) for x in nSs])
        cnt = (ctypes.c_uint64 * n)()
        sj.check(sj.lib().spj_join_batch_device(n, ctypes.cast(rp, P), ctypes.cast(nr, P), ctypes.cast(sp, P),
                                                ctypes.cast(ns, P), ctypes.cast(op, P), ctypes.cast(cnt, P), None,
                                                None))
        got = [(int(cnt[i]),) + sj.checksum_device(int(op[i]), int(cnt[i])) for i in range(n)]
        bad = [i for i in range(n) if got[i] != want[i]]
        assert not bad, (len(bad), bad[:5])
        print("  batched ok", flush=True)
        c1 = ctypes.c_uint64()
        for i in range(n):
            sj.check(sj.lib().spj_join_device(rp[i], nr[i], sp[i], ns[i], op[i], ns[i], ctypes.byref(c1), None, None))
            got = (int(c1.value),) + sj.checksum_device(int(op[i]), int(c1.value))
            assert got == want[i], i
        print("  one per call ok", flush=True)
    finally:
        for d in (dR, dS, dO):
            d.free()


def test_serving_16_resident_tables_match_the_oracle(gpu, orc):
    """16 resident tables, |R| log-uniform in [2^16, 2^24]; probe requests of
    |S| log-uniform in [2^12, 2^22], Zipf U[0, 1.5], each against a random
    table.  Each request's result has the oracle's count and pair checksum
    alone (one request at a time) and in the mix (all submitted at once to a
    server with four slots and an HBM budget that holds only some of them)."""
    sj = gpu
    rng = np.random.default_rng(16)
    sizes = np.floor(2.0 ** rng.uniform(16, 24, 16)).astype(np.int64)
    tables = [sj.gen_r(int(nr), 30_000 + i) for i, nr in enumerate(sizes)]
    reqs = []
    for i in range(40):
        t = int(rng.integers(0, 16))
        ns = int(2 ** rng.uniform(12, 22))
        reqs.append((t, sj.gen_s(ns, tables[t], float(rng.uniform(0, 1.5)), 40_000 + i)))
    print(f"case serving: tables {sizes.sum()} tuples, {len(reqs)} requests, "
          f"{sum(len(s) for _, s in reqs)} probe tuples", flush=True)
    want = [orc.join_checksum(tables[t], S) for t, S in reqs]

    def check(out, k):
        assert (len(out),) + orc.pairs_checksum(out) == want[k], k

    sv = sj.Server(8 << 30, slots=1)
    tids = [sv.add_table(R) for R in tables]
    for k, (t, S) in enumerate(reqs):
        out, _ = sv.wait(sv.submit(tids[t], S))
        check(out, k)
    sv.close()
    print("  alone ok", flush=True)
    sv = sj.Server(256 << 20, slots=4)   # a few of the largest requests at a time
    tids = [sv.add_table(R) for R in tables]
    ids = [sv.submit(tids[t], S) for t, S in reqs]
    for k, i in enumerate(ids):
        out, lat = sv.wait(i)
        assert lat >= 0
        check(out, k)
    st = sv.stats()
    sv.close()
    assert st["requests"] == len(reqs) and st["batches"] < len(reqs)
    print(f"  mix ok: {st}", flush=True)
