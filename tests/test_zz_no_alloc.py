"""The no-allocation entry point (spj_join_device_noalloc) against the
allocating one it varies (spj_join_device with the no-partition join) and
the oracle: two calls in a row, the second smaller, then the limits."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def words(p):
    return (p["r_rid"].astype(np.uint64) << np.uint64(32)) | p["s_rid"].astype(np.uint64)


def noalloc(sj, dR, nR, dS, nS, dO, cap=None):
    cnt = ctypes.c_uint64()
    sj.check(sj.lib().spj_join_device_noalloc(dR.ptr if nR else None, nR, dS.ptr if nS else None, nS, dO.ptr,
                                              nS if cap is None else cap, ctypes.byref(cnt)))
    return cnt.value


def test_noalloc_matches_the_allocating_join(gpu, orc):
    sj = gpu
    for nR, nS, zipf, seed in [(1 << 20, 1 << 22, 1.0, 1), (3000, 11000, 0.4, 2), (1 << 20, 1 << 21, 0.0, 3)]:
        print(f"case noalloc nR={nR} nS={nS} zipf={zipf}", flush=True)
        R = sj.gen_r(nR, 900 + seed)
        S = sj.gen_s(nS, R, zipf, 950 + seed)
        S["key"][::97] += np.uint32(1 << 21)   # some S keys without an R partner
        dR = sj.DeviceArray.from_numpy(R)
        dS = sj.DeviceArray.from_numpy(S)
        dA = sj.DeviceArray(nS * 8)
        dB = sj.DeviceArray(nS * 8)
        try:
            n_na = noalloc(sj, dR, nR, dS, nS, dA)
            n_ref, _ = sj.join_device(dR, nR, dS, nS, dB, sj.config(algorithm=sj.ALG_NPHJ))
            got = dA.to_numpy(sj.PAIR, n_na)
            ref = dB.to_numpy(sj.PAIR, n_ref)
        finally:
            for d in (dR, dS, dA, dB):
                d.free()
        assert n_na == n_ref
        assert np.array_equal(sj.sort_pairs(got), sj.sort_pairs(ref))
        exp = orc.expected_pairs(R, S)
        assert np.array_equal(sj.sort_pairs(got), words(exp))


def test_noalloc_limits_and_errors(gpu, orc):
    sj = gpu
    R = sj.gen_r(4000, 5)
    S = sj.gen_s(9000, R, 0.8, 6)
    dR = sj.DeviceArray.from_numpy(R)
    dS = sj.DeviceArray.from_numpy(S)
    dO = sj.DeviceArray(len(S) * 8)
    try:
        # |R| above the table's limit, and an output too small: refused before any GPU work
        with pytest.raises(sj.SpjError) as e:
            noalloc(sj, dR, (1 << 20) + 1, dS, len(S), dO)
        assert e.value.code == -1 and "limit" in str(e.value)
        with pytest.raises(sj.SpjError) as e:
            noalloc(sj, dR, len(R), dS, len(S), dO, cap=len(S) - 1)
        assert e.value.code == -1
        # empty S, then a duplicate R key, then a clean call again
        assert noalloc(sj, dR, len(R), dS, 0, dO) == 0
        R2 = R.copy()
        R2["key"][1] = R2["key"][2]
        dR2 = sj.DeviceArray.from_numpy(R2)
        with pytest.raises(sj.SpjError) as e:
            noalloc(sj, dR2, len(R2), dS, len(S), dO)
        assert e.value.code == -4
        dR2.free()
        n = noalloc(sj, dR, len(R), dS, len(S), dO)
        assert np.array_equal(sj.sort_pairs(dO.to_numpy(sj.PAIR, n)), words(orc.expected_pairs(R, S)))
    finally:
        for d in (dR, dS, dO):
            d.free()
