"""GPU parity: the HIP join through the C ABI against the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def words(p):
    return (p["r_rid"].astype(np.uint64) << np.uint64(32)) | p["s_rid"].astype(np.uint64)


def check_equal(sj, orc, R, S, pairs):
    exp = orc.expected_pairs(R, S)
    assert len(pairs) == len(exp)
    assert np.array_equal(sj.sort_pairs(pairs), words(exp))


@pytest.mark.parametrize("nR,ratio,zipf", [(1 << 12, 2, 0.0), (1 << 16, 4, 0.5), (1 << 20, 2, 0.0),
                                           (1 << 18, 8, 1.5)])
def test_hash_join_matches_oracle(gpu, orc, nR, ratio, zipf):
    sj = gpu
    R = sj.gen_r(nR, 100 + nR)
    S = sj.gen_s(nR * ratio, R, zipf, 200 + nR)
    pairs, t = sj.join(R, S, timed=True)
    print(f"nR={nR} nS={len(S)} zipf={zipf} times={t}", flush=True)
    check_equal(sj, orc, R, S, pairs)
