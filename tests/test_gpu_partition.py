"""GPU parity of the global partition pass (GLB_PARTITION): the device's
partition of every tuple is the reference's glb_partition_hash(key) &
(2^bits - 1) (glb_partition_count.c:192-195), checked per partition against
the oracle's SuperFastHash (pinned by the reference's own hash.c)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def words(t):
    return (t["key"].astype(np.uint64) << np.uint64(32)) | t["rid"].astype(np.uint64)


@pytest.mark.parametrize("bits", [0, 1, 6, 10, 12])
def test_partition_contents_match_the_reference_hash(gpu, orc, bits):
    """The small workload's R (2^20 keys, a permutation of 1..2^20) and S
    (2^21 keys drawn uniformly from R's): each partition's tuple set equals
    the set of tuples whose oracle SuperFastHash has that low-bit digit."""
    sj = gpu
    R = sj.gen_r(1 << 20, 1)
    S = sj.gen_s(1 << 21, R, 0.0, 2)
    F = 1 << bits
    for rel in (R, S):
        print(f"case partition bits={bits} n={len(rel)}", flush=True)
        out, pst = sj.partition(rel, bits)
        digit = orc.hash_batch("glb_partition_hash", rel["key"]) & np.uint32(F - 1)
        want = np.bincount(digit, minlength=F)
        assert pst[0] == 0 and np.array_equal(np.diff(pst.astype(np.int64)), want)
        order = np.argsort(digit, kind="stable")
        exp_w = words(rel[order])
        got_w = words(out)
        for p in range(F):
            lo, hi = int(pst[p]), int(pst[p + 1])
            assert np.array_equal(np.sort(got_w[lo:hi]), np.sort(exp_w[lo:hi])), p


def test_partition_skips_key_zero_and_edges(gpu, orc):
    """Key-0 tuples are not partitioned (the reference skips them); empty and
    one-tuple relations; every digit of 12 bits."""
    sj = gpu
    rng = np.random.default_rng(3)
    rel = np.zeros(50000, dtype=sj.TUPLE)
    rel["key"] = rng.integers(0, 1 << 32, len(rel), dtype=np.uint64).astype(np.uint32)
    rel["key"][::7] = 0
    rel["rid"] = np.arange(len(rel), dtype=np.uint32)
    out, pst = sj.partition(rel, 12)
    nz = rel[rel["key"] != 0]
    digit = orc.hash_batch("glb_partition_hash", nz["key"]) & np.uint32(4095)
    assert np.array_equal(np.diff(pst.astype(np.int64)), np.bincount(digit, minlength=4096))
    assert int(pst[-1]) == len(nz)
    assert np.array_equal(np.sort(words(out)), np.sort(words(nz)))
    for n in (0, 1):
        out, pst = sj.partition(rel[:n], 3)
        assert int(pst[-1]) == int((rel["key"][:n] != 0).sum())
