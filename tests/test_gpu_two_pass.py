"""GPU parity of the two-pass radix partition (radix_passes = 2): the first
pass partitions R and S by radix_bits of SuperFastHash, the second splits
every partition by the next bits and joins it.  Results against the oracle
(bit-exact sorted multiset of (R rid, S rid) pairs, exact count)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def words(p):
    return (p["r_rid"].astype(np.uint64) << np.uint64(32)) | p["s_rid"].astype(np.uint64)


def check_equal(sj, orc, R, S, pairs):
    exp = orc.expected_pairs(R, S)
    assert len(pairs) == len(exp), (len(pairs), len(exp))
    assert np.array_equal(sj.sort_pairs(pairs), words(exp))


@pytest.mark.parametrize("nR,ratio,zipf,bits,alg,heavy", [
    (1 << 16, 4, 0.0, 1, 0, 0),
    (1 << 18, 4, 0.5, 3, 0, 1),
    (1 << 18, 8, 1.5, 2, 0, 1),     # heavy keys joined in the first pass's S scatter
    (1 << 19, 2, 1.0, 6, 0, 0),     # no heavy handling: hot keys go through both passes
    (100_003, 3, 0.25, 4, 2, 1),    # sort-merge in the second pass
    (1 << 17, 4, 0.8, 12, 0, 1),    # 4096 first-pass partitions: tiny second-pass instances
    (5000, 2, 0.0, 3, 3, 0),        # nested loop in the second pass
])
def test_two_pass_matches_oracle(gpu, orc, nR, ratio, zipf, bits, alg, heavy):
    sj = gpu
    R = sj.gen_r(nR, 300 + nR)
    S = sj.gen_s(nR * ratio, R, zipf, 400 + nR)
    print(f"case two passes nR={nR} nS={len(S)} zipf={zipf} bits={bits} alg={alg} heavy={heavy}", flush=True)
    cfg = sj.config(algorithm=alg, radix_passes=2, radix_bits=bits, heavy=heavy)
    pairs, t = sj.join(R, S, cfg, timed=True)
    check_equal(sj, orc, R, S, pairs)
    assert t["passes"] == 2
    assert t["r_partitioned"] == nR
    if heavy and zipf >= 1.0:
        assert t["heavy_keys"] > 0 and t["s_partitioned"] < len(S)


def test_two_pass_unmatched_zero_keys_and_holes(gpu, orc):
    """S keys missing from R and key-0 tuples on both sides: second-pass
    instances leave holes, which the packing closes; heavy pairs of the
    first pass follow the second pass's pairs."""
    sj = gpu
    rng = np.random.default_rng(17)
    nR, nS = 1 << 17, 1 << 20
    R = sj.gen_r(nR, 51)
    S = sj.gen_s(nS, R, 1.2, 52)
    R["key"][::3] = 0
    miss = rng.random(nS) < 0.1
    S["key"][miss] = S["key"][miss] + np.uint32(1 << 24)
    S["key"][rng.random(nS) < 0.01] = 0
    for heavy in (0, 1):
        print(f"case two passes unmatched heavy={heavy}", flush=True)
        pairs, t = sj.join(R, S, sj.config(radix_passes=2, radix_bits=5, heavy=heavy), timed=True)
        check_equal(sj, orc, R, S, pairs)
        assert t["passes"] == 2


@pytest.mark.parametrize("layout,pct", [(1, 0), (1, 60), (0, 0)])
def test_two_pass_first_pass_layouts(gpu, orc, layout, pct):
    """The first pass in cells (its partitions read as segments by the
    second), in cells too small (an overflow redoes it counted), and
    counted: the same exact result."""
    sj = gpu
    R = sj.gen_r(1 << 20, 31)   # 32 partitions: 4K-tuple segments (cells only from 4K on)
    S = sj.gen_s(1 << 21, R, 0.6, 32)
    print(f"case two passes layout={layout} pct={pct}", flush=True)
    cfg = sj.config(radix_passes=2, radix_bits=5, part_layout=layout, cell_capacity_pct=pct if pct else -1)
    pairs, t = sj.join(R, S, cfg, timed=True)
    check_equal(sj, orc, R, S, pairs)
    assert t["passes"] == 2 and t["r_partitioned"] == len(R)
    if layout == 1 and pct == 0:
        assert t["hist_ms"] == 0.0   # no histogram pass in either pass


def test_two_pass_duplicate_r_key_is_an_error(gpu):
    sj = gpu
    R = sj.gen_r(1 << 15, 61)
    R["key"][100] = R["key"][2000]
    S = sj.gen_s(1 << 16, sj.gen_r(1 << 15, 61), 0.0, 62)
    print("case two passes duplicate key", flush=True)
    with pytest.raises(sj.SpjError) as e:
        sj.join(R, S, sj.config(radix_passes=2, radix_bits=3))
    assert e.value.code == -4


def test_two_pass_same_as_one_pass(gpu, orc):
    """The same join in one pass and in two: identical pair multisets."""
    sj = gpu
    R = sj.gen_r(1 << 20, 71)
    S = sj.gen_s(1 << 22, R, 0.7, 72)
    print("case one pass vs two", flush=True)
    one, t1 = sj.join(R, S, sj.config(radix_bits=8), timed=True)
    two, t2 = sj.join(R, S, sj.config(radix_bits=8, radix_passes=2), timed=True)
    assert t1["passes"] == 1 and t2["passes"] == 2
    assert np.array_equal(sj.sort_pairs(one), sj.sort_pairs(two))
    check_equal(sj, orc, R, S, two)


def test_two_pass_argument_errors(gpu):
    sj = gpu
    R = sj.gen_r(1 << 12, 81)
    S = sj.gen_s(1 << 13, R, 0.0, 82)
    for cfg in (sj.config(radix_passes=3), sj.config(radix_passes=2, radix_bits=0),
                sj.config(radix_passes=2, algorithm=sj.ALG_NPHJ)):
        with pytest.raises(sj.SpjError) as e:
            sj.join(R, S, cfg)
        assert e.value.code == -1
    with pytest.raises(sj.SpjError) as e:   # batches partition in one pass
        sj.join_batch([R, R], [S, S], sj.config(radix_passes=2))
    assert e.value.code == -1


def test_two_pass_auto_at_2_28(gpu, orc):
    """|R| = 2^28: one pass of 12 bits would leave partitions of four LDS
    tables, so AUTO partitions twice.  The device checksum of the pairs and
    the count equal the oracle's closed form."""
    sj = gpu
    nR, nS = 1 << 28, 1 << 28
    R = sj.gen_r(nR, 91)
    S = sj.gen_s(nS, R, 0.5, 92)
    want = orc.join_checksum(R, S)
    print("case AUTO two passes at 2^28 x 2^28", flush=True)
    dR = sj.DeviceArray.from_numpy(R)
    dS = sj.DeviceArray.from_numpy(S)
    del R, S
    dO = sj.DeviceArray(nS * 8)
    try:
        n, t = sj.join_device(dR, nR, dS, nS, dO, None, timed=True)
        got = (n,) + sj.checksum_device(dO.ptr, n)
        assert t["passes"] == 2 and t["hist_ms"] == 0.0   # both passes in cells
        assert tuple(got) == tuple(want)
    finally:
        dR.free()
        dS.free()
        dO.free()
