"""CPU: the oracle, pinned against the reference's own hash functions, brute
force, closed forms and its restatement of the reference join."""
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


def load(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


def words(p):
    return (p["r_rid"].astype(np.uint64) << np.uint64(32)) | p["s_rid"].astype(np.uint64)


def test_crc_table_closed_form_and_reference(orc):
    t = orc.crc_table()
    # reflected CRC-32 spot values
    assert t[0] == 0 and t[1] == 0x77073096 and t[128] == 0xEDB88320 and t[255] == 0x2D02EF8D
    # byte-equal to the reference's crctab (captured from its compiled hash.c)
    assert t.tolist() == load("hash_kat.json")["crctab"]
    # CRC-32 of "123456789" is 0xCBF43926 with this table (the standard check value)
    c = 0xFFFFFFFF
    for b in b"123456789":
        c = int(t[(c ^ b) & 0xFF]) ^ (c >> 8)
    assert c ^ 0xFFFFFFFF == 0xCBF43926


@pytest.mark.parametrize("name", ["glb_partition_hash", "join_hash", "local_partition_hash", "join_hash2"])
def test_hashes_match_reference_kat(orc, name):
    kat = load("hash_kat.json")
    got = orc.hash_batch(name, np.array(kat["keys"], dtype=np.uint32))
    assert got.tolist() == kat[name]


def test_hashes_match_reference_build_when_present(orc):
    if not os.path.exists(orc.REF_HASH):
        pytest.skip("oracle/_ref not built (reference tree absent)")
    import ctypes

    lib = ctypes.CDLL(orc.REF_HASH)
    rng = np.random.default_rng(3)
    keys = rng.integers(0, 1 << 32, 4096, dtype=np.uint64).astype(np.uint32)
    for name in orc.HASHES:
        f = getattr(lib, name)
        f.restype = ctypes.c_uint32
        f.argtypes = [ctypes.c_uint32]
        want = [f(int(k)) for k in keys]
        assert orc.hash_batch(name, keys).tolist() == want, name


def test_superfasthash_published_steps(orc):
    """SuperFastHash (Hsieh) on 4 bytes, written out from the published
    algorithm in plain Python: a different implementation of the same steps."""
    M = 0xFFFFFFFF

    def sfh(key):
        data = key.to_bytes(4, "little")
        h = 4
        h = (h + (data[0] | data[1] << 8)) & M
        tmp = (((data[2] | data[3] << 8) << 11) ^ h) & M
        h = ((h << 16) ^ tmp) & M
        h = (h + (h >> 11)) & M
        for sh, op in ((3, "x<"), (5, "+>"), (4, "x<"), (17, "+>"), (25, "x<"), (6, "+>")):
            h = (h ^ (h << sh)) & M if op == "x<" else (h + (h >> sh)) & M
        return h

    for k in [0, 1, 77, 65535, 65536, 123456789, 0xFFFFFFFF]:
        assert orc.hash_batch("glb_partition_hash", np.array([k], dtype=np.uint32))[0] == sfh(k)


def brute_join(R, S):
    out = []
    for sk, sr in zip(S["key"].tolist(), S["rid"].tolist()):
        if sk == 0:
            continue
        for rk, rr in zip(R["key"].tolist(), R["rid"].tolist()):
            if rk == sk:
                out.append((rr << 32) | sr)
    return np.array(sorted(out), dtype=np.uint64)


def test_expected_pairs_brute_force(orc):
    rng = np.random.default_rng(11)
    for trial in range(20):
        nR, nS = int(rng.integers(0, 64)), int(rng.integers(0, 256))
        R = np.zeros(nR, dtype=orc.TUPLE)
        R["key"] = rng.choice(np.arange(0, 200, dtype=np.uint32), nR, replace=False)
        R["rid"] = rng.integers(0, 1 << 32, nR, dtype=np.uint64).astype(np.uint32)
        S = np.zeros(nS, dtype=orc.TUPLE)
        S["key"] = rng.integers(0, 220, nS).astype(np.uint32)
        S["rid"] = np.arange(nS, dtype=np.uint32)
        assert np.array_equal(words(orc.expected_pairs(R, S)), brute_join(R, S))


def test_golden_join_fixture(orc):
    g = load("join_small.json")
    R = np.array([tuple(x) for x in g["R"]], dtype=orc.TUPLE)
    S = np.array([tuple(x) for x in g["S"]], dtype=orc.TUPLE)
    want = np.array([(a << 32) | b for a, b in g["pairs"]], dtype=np.uint64)
    assert np.array_equal(words(orc.expected_pairs(R, S)), want)
    assert np.array_equal(brute_join(R, S), want)


def test_inverse_permutation_closed_form(orc, sj):
    R = sj.gen_r(1 << 14, 5)
    S = sj.gen_s(1 << 15, R, 1.1, 6)
    inv = np.empty(len(R) + 1, dtype=np.uint64)
    inv[R["key"]] = R["rid"]
    want = np.sort((inv[S["key"]] << np.uint64(32)) | S["rid"].astype(np.uint64))
    assert np.array_equal(words(orc.expected_pairs(R, S)), want)
    c, s, x = orc.join_checksum(R, S)
    assert c == len(S) and (s, x) == orc.pairs_checksum(orc.expected_pairs(R, S))


@pytest.mark.parametrize("design", [dict(), dict(bank_set=8), dict(bank_set=16), dict(bank_set=32),
                                    dict(bank_set=64), dict(rank_set=2), dict(rank_set=4, bank_set=8),
                                    dict(rank_num=4), dict(join_alg=2), dict(join_alg=2, bank_set=16)])
def test_cpu_restatement_equals_closed_form(orc, sj, design):
    R = sj.gen_r(1 << 17, 21)
    S = sj.gen_s(1 << 18, R, 1.0, 22)
    got = orc.spid_join_cpu(R, S, threads=4, **design)
    assert len(got) == len(S)
    assert np.array_equal(words(orc.sort_pairs(got)), words(orc.expected_pairs(R, S)))


def test_cpu_restatement_nested_loop_small(orc, sj):
    R = sj.gen_r(2000, 31)
    S = sj.gen_s(5000, R, 0.3, 32)
    got = orc.spid_join_cpu(R, S, join_alg=3, threads=4)
    assert np.array_equal(words(orc.sort_pairs(got)), words(orc.expected_pairs(R, S)))


def test_cpu_restatement_unmatched_and_zero(orc, sj):
    R = sj.gen_r(1 << 14, 41)
    S = sj.gen_s(1 << 15, R, 0.0, 42)
    R["key"][::5] = 0
    S["key"][::7] += np.uint32(1 << 20)
    got = orc.spid_join_cpu(R, S, threads=4)
    assert np.array_equal(words(orc.sort_pairs(got)), words(orc.expected_pairs(R, S)))


def test_cpu_restatement_detects_duplicates(orc, sj):
    R = sj.gen_r(1 << 12, 51)
    R["key"][7] = R["key"][8]
    S = sj.gen_s(1 << 12, sj.gen_r(1 << 12, 51), 0.0, 52)
    with pytest.raises(RuntimeError):
        orc.spid_join_cpu(R, S, threads=2)


def test_planted_mistake_is_caught(orc, sj):
    """A join that drops one pair, or swaps one rid, must fail the multiset
    comparison the GPU tests use."""
    R = sj.gen_r(4096, 61)
    S = sj.gen_s(8192, R, 0.5, 62)
    exp = orc.expected_pairs(R, S)
    bad = exp.copy()
    bad["r_rid"][100] ^= 1
    assert not np.array_equal(words(orc.sort_pairs(bad)), words(exp))
    assert orc.pairs_checksum(bad) != orc.pairs_checksum(exp)
    assert len(exp[1:]) != len(exp)
