"""The autotuner on the GPU: it times candidates on device relations, stores
the winner for the size / skew bucket, and later calls with a default
config run with it (results still equal the oracle's)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def words(p):
    return (p["r_rid"].astype(np.uint64) << np.uint64(32)) | p["s_rid"].astype(np.uint64)


@pytest.mark.parametrize("nR,ratio,zipf", [(1 << 20, 4, 0.5), (1 << 19, 8, 1.2), ((1 << 22) + 1, 2, 0.0)])
def test_autotune_stores_a_winner_that_later_calls_use(gpu, orc, nR, ratio, zipf, tmp_path):
    sj = gpu
    L = sj.lib()
    sj.check(L.spj_tuner_clear())
    R = sj.gen_r(nR, 50 + nR)
    S = sj.gen_s(nR * ratio, R, zipf, 51 + nR)
    nS = len(S)
    dR, dS, dO = sj.DeviceArray.from_numpy(R), sj.DeviceArray.from_numpy(S), sj.DeviceArray(nS * 8)
    try:
        print(f"case autotune nR={nR} nS={nS} zipf={zipf}", flush=True)
        before = sj.tuner_get(nR, nS, zipf)
        ms = sj.autotune(dR, nR, dS, nS, dO, zipf, reps=2)
        assert 0 < ms < 1e3
        assert L.spj_tuner_entries() == 1
        won = sj.tuner_get(nR, nS, zipf)
        print(f"  default {before}\n  tuned {won} in {ms:.3f} ms", flush=True)
        assert won["algorithm"] in (0, 1)
        if won["algorithm"] == 0:
            assert won["scatter_ipt"] in (8, 10, 12) and won["table_log2"] in (12, 13)
            assert won["items_s_max"] in (1 << 16, 1 << 17, 1 << 18)
        # a later default call runs with the entry: the no-partition join
        # partitions nothing, the partitioned one partitions all of R
        n, t = sj.join_device(dR, nR, dS, nS, dO, None, timed=True)
        assert (t["r_partitioned"] == 0) == (won["algorithm"] == 1)
        got = dO.to_numpy(sj.PAIR, n)
        assert np.array_equal(sj.sort_pairs(got), words(orc.expected_pairs(R, S)))
        # the table survives a save / clear / load
        path = str(tmp_path / "tuned.txt")
        sj.check(L.spj_tuner_save(path.encode()))
        sj.check(L.spj_tuner_clear())
        sj.check(L.spj_tuner_load(path.encode()))
        assert sj.tuner_get(nR, nS, zipf) == won
    finally:
        sj.check(L.spj_tuner_clear())
        for d in (dR, dS, dO):
            d.free()


def test_no_partition_size_limit_is_an_argument_error(gpu):
    """The no-partition join refuses |R| > 2^30 up front (SPJ_EINVAL), before
    any device work; the pointers are never read."""
    sj = gpu
    import ctypes
    cnt = ctypes.c_uint64()
    d = sj.DeviceArray(64)
    try:
        cfg = sj.config(algorithm=sj.ALG_NPHJ)
        rc = sj.lib().spj_join_device(d.ptr, (1 << 30) + 1, d.ptr, 4, d.ptr, 4, ctypes.byref(cnt), ctypes.byref(cfg),
                                      None)
        assert rc == -1 and b"2^30" in sj.lib().spj_last_error()
    finally:
        d.free()
