"""GPU parity over a seeded sweep of configurations: sizes from 1 to 2^18,
R:S ratios, skew, every join algorithm, radix fan-outs, scatter rounds, LDS
table sizes, S-range splits, partition layouts and heavy-key handling, each
checked bit-exact against the oracle (sorted multiset of pairs and count);
single joins and one batch of all cases."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def words(p):
    return (p["r_rid"].astype(np.uint64) << np.uint64(32)) | p["s_rid"].astype(np.uint64)


def cases():
    rng = np.random.default_rng(20261204)
    out = []
    for i in range(36):
        nR = int(2 ** rng.uniform(0, 18))
        nS = int(nR * rng.choice([0.5, 1, 2, 3, 8]))
        kw = dict(algorithm=int(rng.choice([-1, 0, 1, 2, 3])) if nR <= 4096 else int(rng.choice([-1, 0, 1, 2])),
                  heavy=int(rng.choice([-1, 0, 1])), scatter_ipt=int(rng.choice([4, 8, 10, 12])),
                  table_log2=int(rng.choice([12, 13])), part_layout=int(rng.choice([-1, 0, 1])))
        if rng.random() < 0.4:
            kw["radix_bits"] = int(rng.integers(0, 13))
        if rng.random() < 0.3:
            kw["items_s_max"] = int(rng.choice([512, 4096, 65536]))
        out.append((i, nR, nS, float(rng.choice([0.0, 0.5, 1.0, 1.5])), kw))
    return out


CASES = cases()


@pytest.mark.parametrize("i,nR,nS,zipf,kw", CASES, ids=[f"case{c[0]}" for c in CASES])
def test_random_config_matches_oracle(gpu, orc, i, nR, nS, zipf, kw):
    sj = gpu
    R = sj.gen_r(nR, 7000 + i)
    S = sj.gen_s(nS, R, zipf, 8000 + i) if nS else np.zeros(0, dtype=R.dtype)
    print(f"case {i}: nR={nR} nS={nS} zipf={zipf} {kw}", flush=True)
    pairs, _ = sj.join(R, S, sj.config(**kw))
    exp = orc.expected_pairs(R, S)
    assert len(pairs) == len(exp)
    assert np.array_equal(sj.sort_pairs(pairs), words(exp))


def test_random_configs_as_one_batch(gpu, orc):
    """The same relations in one batch (one configuration for all: the
    batch takes a single config)."""
    sj = gpu
    Rs, Ss = [], []
    for i, nR, nS, zipf, _ in CASES:
        R = sj.gen_r(nR, 7000 + i)
        Rs.append(R)
        Ss.append(sj.gen_s(nS, R, zipf, 8000 + i) if nS else np.zeros(0, dtype=R.dtype))
    print(f"case batch of {len(Rs)}", flush=True)
    outs, _ = sj.join_batch(Rs, Ss)
    for R, S, p in zip(Rs, Ss, outs):
        exp = orc.expected_pairs(R, S)
        assert len(p) == len(exp)
        assert np.array_equal(sj.sort_pairs(p), words(exp))
