"""GPU parity: the HIP join through the C ABI against the oracle (bit-exact
sorted multiset of (R rid, S rid) pairs, exact count)."""
import re
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def words(p):
    return (p["r_rid"].astype(np.uint64) << np.uint64(32)) | p["s_rid"].astype(np.uint64)


def check_equal(sj, orc, R, S, pairs):
    exp = orc.expected_pairs(R, S)
    assert len(pairs) == len(exp), (len(pairs), len(exp))
    assert np.array_equal(sj.sort_pairs(pairs), words(exp))


@pytest.mark.parametrize("nR,ratio,zipf", [(1 << 12, 2, 0.0), (1 << 16, 4, 0.5), (1 << 20, 2, 0.0),
                                           (1 << 18, 8, 1.5), (1 << 20, 4, 1.0), (100_003, 3, 0.25)])
def test_hash_join_matches_oracle(gpu, orc, nR, ratio, zipf):
    sj = gpu
    R = sj.gen_r(nR, 100 + nR)
    S = sj.gen_s(nR * ratio, R, zipf, 200 + nR)
    print(f"case nR={nR} nS={len(S)} zipf={zipf}", flush=True)
    pairs, t = sj.join(R, S, sj.config(algorithm=0), timed=True)
    print(f"  times={t}", flush=True)
    check_equal(sj, orc, R, S, pairs)


@pytest.mark.parametrize("alg", [2, 3])
def test_sort_merge_and_nested_loop(gpu, orc, alg):
    sj = gpu
    nR = (1 << 16) if alg == 2 else 3000
    R = sj.gen_r(nR, 7 + alg)
    S = sj.gen_s(nR * 2, R, 0.8, 9 + alg)
    print(f"case alg={alg} nR={nR}", flush=True)
    pairs, _ = sj.join(R, S, sj.config(algorithm=alg))
    check_equal(sj, orc, R, S, pairs)


@pytest.mark.parametrize("alg", [0, 2, 3])
@pytest.mark.parametrize("layout", [0, 1])
@pytest.mark.parametrize("heavy", [0, 1])
def test_partition_layouts(gpu, orc, alg, layout, heavy):
    """R and S partitioned counted (histogram pass) or into fixed-capacity
    cells (no histogram pass), under every join algorithm, with skew, with
    and without heavy keys (whose R rids the R cell scatter records)."""
    sj = gpu
    nR = {0: 1 << 17, 2: 1 << 15, 3: 2000}[alg]
    R = sj.gen_r(nR, 70 + alg)
    S = sj.gen_s(nR * 4, R, 1.1, 80 + alg)
    print(f"case layout={layout} alg={alg} heavy={heavy} nR={nR}", flush=True)
    pairs, t = sj.join(R, S, sj.config(algorithm=alg, part_layout=layout, heavy=heavy), timed=True)
    check_equal(sj, orc, R, S, pairs)
    if layout == 0:
        assert t["s_cells"] == 0 and t["r_cells"] == 0
    if heavy:
        assert t["heavy_keys"] > 0


@pytest.mark.parametrize("pct", [10, 60, 100])
def test_cell_overflow_redone_counted(gpu, orc, pct):
    """Cells too small for their tuples: the scatter flags the overflow and
    the join is redone with counted partitions, with the same result."""
    sj = gpu
    R = sj.gen_r(1 << 18, 91)
    S = sj.gen_s(1 << 21, R, 0.6, 92)
    print(f"case cell pct={pct}", flush=True)
    pairs, t = sj.join(R, S, sj.config(part_layout=1, cell_capacity_pct=pct), timed=True)
    check_equal(sj, orc, R, S, pairs)
    assert t["s_cells"] == 0 and t["r_cells"] == 0   # redone counted
    pairs, t = sj.join(R, S, sj.config(algorithm=0), timed=True)   # automatic capacity: no overflow
    check_equal(sj, orc, R, S, pairs)
    assert t["s_cells"] == 1 and t["r_cells"] == 1   # R in cells at any size (unique keys)
    pairs, t = sj.join(R, S, sj.config(heavy=0), timed=True)   # no heavy handling: S stays counted
    check_equal(sj, orc, R, S, pairs)
    assert t["s_cells"] == 0
    pairs, t = sj.join(R, S, sj.config(part_layout=1), timed=True)
    check_equal(sj, orc, R, S, pairs)
    assert t["s_cells"] == 1 and t["r_cells"] == 1


def test_unmatched_and_zero_keys(gpu, orc):
    """S keys missing from R, key-0 tuples on both sides (skipped, as the
    reference treats key 0 as empty), heavy keys without an R partner."""
    sj = gpu
    rng = np.random.default_rng(5)
    nR, nS = 1 << 18, 1 << 21
    R = sj.gen_r(nR, 21)
    S = sj.gen_s(nS, R, 1.2, 22)
    # drop every 3rd R tuple's key to 0 and shift some S keys out of range
    R["key"][::3] = 0
    miss = rng.random(nS) < 0.1
    S["key"][miss] = S["key"][miss] + np.uint32(1 << 24)
    S["key"][rng.random(nS) < 0.01] = 0
    print("case unmatched", flush=True)
    for heavy in (0, 1):
        pairs, _ = sj.join(R, S, sj.config(heavy=heavy))
        check_equal(sj, orc, R, S, pairs)


def test_tiny_and_empty(gpu, orc):
    sj = gpu
    R = sj.gen_r(5, 1)
    S = sj.gen_s(17, R, 0.0, 2)
    E = np.zeros(0, dtype=sj.TUPLE)
    for cfg in (None, sj.config(algorithm=0)):
        pairs, _ = sj.join(R, S, cfg)
        check_equal(sj, orc, R, S, pairs)
        pairs, _ = sj.join(E, S, cfg)
        assert len(pairs) == 0
        pairs, _ = sj.join(R, E, cfg)
        assert len(pairs) == 0


def test_duplicate_r_key_is_an_error(gpu):
    sj = gpu
    R = sj.gen_r(1 << 12, 3)
    R["key"][10] = R["key"][11]
    S = sj.gen_s(1 << 13, sj.gen_r(1 << 12, 3), 0.0, 4)
    for cfg in (None, sj.config(algorithm=0)):
        with pytest.raises(sj.SpjError) as e:
            sj.join(R, S, cfg)
        assert e.value.code == -4


@pytest.mark.parametrize("bits", [2, 4, 6])
def test_sort_merge_long_runs_and_duplicates(gpu, orc, bits):
    """Sort-merge with R partitions longer than one sorted run (2^18 R tuples
    over 4 partitions: four 16K runs each), around one run (16 partitions of
    16K +- 128: about half add a short second run to the same pass) and
    below it (64), and a duplicate R key reported as an error."""
    sj = gpu
    R = sj.gen_r(1 << 18, 95 + bits)
    S = sj.gen_s(1 << 19, R, 1.0, 96)
    print(f"case smj bits={bits}", flush=True)
    pairs, _ = sj.join(R, S, sj.config(algorithm=2, radix_bits=bits))
    check_equal(sj, orc, R, S, pairs)
    R2 = R.copy()
    R2["key"][500] = R2["key"][9000]
    with pytest.raises(sj.SpjError) as e:
        sj.join(R2, S, sj.config(algorithm=2, radix_bits=8))
    assert e.value.code == -4


@pytest.mark.parametrize("nR,ratio,zipf", [(1 << 12, 2, 0.0), (1 << 16, 4, 1.0), (1 << 20, 2, 0.5),
                                           (100_003, 3, 1.5), (1, 5, 0.0), (1 << 21, 8, 0.0)])
def test_no_partition_hash_join(gpu, orc, nR, ratio, zipf):
    """join_alg 1 (NPHJ): one HBM table over R, probed by every S tuple."""
    sj = gpu
    R = sj.gen_r(nR, 300 + nR)
    S = sj.gen_s(nR * ratio, R, zipf, 400 + nR)
    print(f"case nphj nR={nR} nS={len(S)} zipf={zipf}", flush=True)
    pairs, t = sj.join(R, S, sj.config(algorithm=1), timed=True)
    check_equal(sj, orc, R, S, pairs)
    assert t["heavy_keys"] == 0 and t["s_partitioned"] == 0 and t["r_partitioned"] == 0


def test_no_partition_hash_join_edges(gpu, orc):
    """Empty R or S, S keys absent from R (unmatched S tuples leave no
    holes), key 0 in both, and a duplicate R key (an error)."""
    sj = gpu
    cfg = sj.config(algorithm=1)
    R = sj.gen_r(5000, 81)
    S = sj.gen_s(20000, R, 0.7, 82)
    empty = np.zeros(0, dtype=R.dtype)
    for r, s_ in [(empty, S), (R, empty)]:
        pairs, _ = sj.join(r, s_, cfg)
        assert len(pairs) == 0
    other = sj.gen_r(7000, 83)   # keys 1..7000: those above 5000 are not in R
    S2 = S.copy()
    S2["key"][::3] = other["key"][: len(S2[::3])]
    S2["key"][5] = 0
    R2 = R.copy()
    R2["key"][7] = 0
    pairs, _ = sj.join(R2, S2, cfg)
    check_equal(sj, orc, R2, S2, pairs)
    R3 = R.copy()
    R3["key"][10] = R3["key"][11]
    with pytest.raises(sj.SpjError) as e:
        sj.join(R3, S, cfg)
    assert e.value.code == -4


def test_no_partition_hash_join_batch_and_instance(gpu, orc):
    """NPHJ instances of mixed sizes in one batch, and JoinInstance with
    join_alg 1."""
    sj = gpu
    rng = np.random.default_rng(19)
    Rs, Ss = [], []
    for i in range(24):
        nR = int(2 ** rng.uniform(4, 17))
        R = sj.gen_r(nR, 500 + i)
        Rs.append(R)
        Ss.append(sj.gen_s(nR * int(rng.choice([1, 2, 4, 8])), R, float(rng.uniform(0, 1.25)), 600 + i))
    outs, _ = sj.join_batch(Rs, Ss, sj.config(algorithm=1))
    for R, S, p in zip(Rs, Ss, outs):
        check_equal(sj, orc, R, S, p)
    ji = sj.JoinInstance(16, 1, 16)
    ji.load_column(Rs[3], "left")
    ji.load_column(Ss[3], "right")
    d = sj.JoinInstance.design(join_alg=1)
    ji.start(d)
    check_equal(sj, orc, Rs[3], Ss[3], ji.finish())
    ji.close()


def test_auto_picks_no_partition_join_for_small_calls(gpu, orc):
    """AUTO runs the no-partition join (no partition pass: r_partitioned 0)
    for a small call (one join of |R| <= 2^21 and |R| |S| <= 2^44, a batch
    of total |R| <= 2^20) with no partition parameter set; otherwise the
    partitioned join."""
    sj = gpu
    R = sj.gen_r(1 << 20, 171)
    S = sj.gen_s(1 << 21, R, 1.2, 172)
    print("case auto single", flush=True)
    for cfg, parted in ((None, False), (sj.config(), False), (sj.config(radix_bits=9), True),
                        (sj.config(heavy=1), True), (sj.config(algorithm=0), True)):
        pairs, t = sj.join(R, S, cfg, timed=True)
        check_equal(sj, orc, R, S, pairs)
        assert (t["r_partitioned"] == len(R)) == parted, (cfg, t)
    # above 2^21 R tuples, or |R| |S| above 2^44: partitioned
    for nr, ns in (((1 << 21) + 1, 1 << 22), (1 << 19, (1 << 25) + 64)):
        R2 = sj.gen_r(nr, 173)
        S2 = sj.gen_s(ns, R2, 0.5, 174)
        pairs, t = sj.join(R2, S2, timed=True)
        check_equal(sj, orc, R2, S2, pairs)
        assert t["r_partitioned"] == len(R2), (nr, ns)
    # a batch is judged by its total |R|
    Rs = [sj.gen_r(1 << 18, 180 + i) for i in range(5)]
    Ss = [sj.gen_s(1 << 19, Rs[i], 0.8, 190 + i) for i in range(5)]
    print("case auto batch", flush=True)
    for k, parted in ((4, False), (5, True)):
        outs, t = sj.join_batch(Rs[:k], Ss[:k], timed=True)
        for R_, S_, p in zip(Rs[:k], Ss[:k], outs):
            check_equal(sj, orc, R_, S_, p)
        assert (t["r_partitioned"] > 0) == parted, (k, t)


@pytest.mark.parametrize("bits", [0, 3, 9, 12])
def test_radix_bits_sweep(gpu, orc, bits):
    """Any radix fan-out gives the same multiset (large partitions are split
    into several LDS tables)."""
    sj = gpu
    R = sj.gen_r(1 << 17, 31)
    S = sj.gen_s(1 << 18, R, 0.5, 32)
    print(f"case bits={bits}", flush=True)
    pairs, _ = sj.join(R, S, sj.config(radix_bits=bits))
    check_equal(sj, orc, R, S, pairs)


@pytest.mark.parametrize("copies", [2, 3, 5])
def test_duplicate_r_key_in_large_partition(gpu, copies):
    """A key repeated `copies` times among 2^18 R tuples (partitions of
    thousands of tuples, its bucket holding at least `copies` entries) is
    reported as a duplicate."""
    sj = gpu
    R = sj.gen_r(1 << 18, 11 + copies)
    R["key"][1000:1000 + copies] = R["key"][77]
    S = sj.gen_s(1 << 19, sj.gen_r(1 << 18, 11 + copies), 0.0, 12)
    with pytest.raises(sj.SpjError) as e:
        sj.join(R, S, sj.config(algorithm=0))
    assert e.value.code == -4


def test_batch_matches_single(gpu, orc):
    sj = gpu
    rng = np.random.default_rng(9)
    Rs, Ss = [], []
    for i in range(40):
        nR = int(2 ** rng.uniform(10, 16))
        ratio = int(rng.choice([1, 2, 4, 8]))
        R = sj.gen_r(nR, 1000 + i)
        S = sj.gen_s(nR * ratio, R, float(rng.uniform(0, 1.25)), 2000 + i)
        Rs.append(R)
        Ss.append(S)
    outs, _ = sj.join_batch(Rs, Ss, sj.config(algorithm=0))
    for R, S, p in zip(Rs, Ss, outs):
        check_equal(sj, orc, R, S, p)


def test_batch_many_tiny_and_empty(gpu, orc):
    """Hundreds of instances, many of them tiny or empty on one side, next to
    a skewed one with heavy keys: tiles of many instances share the scatter
    groups and the join's work queue."""
    sj = gpu
    rng = np.random.default_rng(77)
    Rs, Ss = [], []
    for i in range(300):
        nR = int(rng.choice([0, 1, 2, 7, 33, 100, 1000, 5000]))
        nS = int(rng.choice([0, 1, 3, 50, 999, 4096, 20000]))
        R = sj.gen_r(nR, 3000 + i)
        S = sj.gen_s(nS, R, float(rng.uniform(0, 1.5)), 4000 + i) if nR else np.zeros(0, dtype=sj.TUPLE)
        Rs.append(R)
        Ss.append(S)
    Rs.append(sj.gen_r(1 << 18, 5000))
    Ss.append(sj.gen_s(1 << 21, Rs[-1], 1.5, 5001))
    print("case batch many tiny", flush=True)
    outs, _ = sj.join_batch(Rs, Ss, sj.config(algorithm=0))
    for R, S, p in zip(Rs, Ss, outs):
        check_equal(sj, orc, R, S, p)


def test_batch_cell_overflow_redone(gpu, orc):
    """A batch whose instances all use cells too small for them: the batch is
    redone with counted partitions and every instance's result is exact."""
    sj = gpu
    rng = np.random.default_rng(19)
    Rs, Ss = [], []
    for i in range(12):
        nR = int(2 ** rng.uniform(12, 17))
        R = sj.gen_r(nR, 600 + i)
        Rs.append(R)
        Ss.append(sj.gen_s(nR * int(rng.choice([1, 4])), R, float(rng.uniform(0, 1.2)), 700 + i))
    print("case batch cell overflow", flush=True)
    outs, t = sj.join_batch(Rs, Ss, sj.config(part_layout=1, cell_capacity_pct=60), timed=True)
    for R, S, p in zip(Rs, Ss, outs):
        check_equal(sj, orc, R, S, p)
    assert t["s_cells"] == 0


def test_join_instance_lifecycle_cells(gpu, orc):
    """The asynchronous JoinInstance path on sizes that use the cell layout
    with heavy keys (R scattered first, recording their rids)."""
    sj = gpu
    R = sj.gen_r(1 << 20, 43)
    S = sj.gen_s(1 << 22, R, 1.1, 44)
    ji = sj.JoinInstance(16, 1, 16)
    ji.load_column(R, "left")
    ji.load_column(S, "right")
    d = sj.JoinInstance.design(join_alg=0, bank_set=8)
    ji.prepare(d)
    ji.start(d)
    pairs = ji.finish()
    check_equal(sj, orc, R, S, pairs)
    t = ji.phase_times()
    assert t["s_cells"] == 1 and t["r_cells"] == 1 and t["heavy_keys"] > 0
    ji.close()


def test_join_instance_lifecycle(gpu, orc):
    """LoadColumn / PrepareJoin / StartJoin (async) / FinishJoin, query.cc."""
    sj = gpu
    R = sj.gen_r(1 << 16, 41)
    S = sj.gen_s(1 << 18, R, 1.0, 42)
    ji = sj.JoinInstance(16, 1, 16)
    ji.load_column(R, "left")
    ji.load_column(S, "right")
    d = sj.JoinInstance.design(join_alg=0, bank_set=8)
    ji.prepare(d)
    ji.start(d)
    pairs = ji.finish()
    check_equal(sj, orc, R, S, pairs)
    assert ji.phase_times()["total_ms"] > 0
    with pytest.raises(sj.SpjError):
        ji.finish()   # FinishJoin without StartJoin
    ji.close()


@pytest.mark.parametrize("nR,nS,zipf", [(5000, 12345, 0.3), ((1 << 22) + 3, (1 << 24) + 5, 1.0), (777, 0, 0.0)])
def test_host_paths_staged_copies(gpu, orc, nR, nS, zipf):
    """Host relations through the page-locked staging ring (chunks of 32 MiB,
    four slots): below one chunk, several ring turns with a ragged last
    chunk, and an empty S; both the one-call host join and JoinInstance."""
    sj = gpu
    R = sj.gen_r(nR, 71 + nR)
    S = sj.gen_s(nS, R, zipf, 72 + nS) if nS else np.zeros(0, dtype=R.dtype)
    print(f"case host nR={nR} nS={nS}", flush=True)
    check_equal(sj, orc, R, S, sj.join_host(R, S))
    ji = sj.JoinInstance(16, 1, 16)
    ji.load_column(R, "left")
    ji.load_column(S, "right")
    d = sj.JoinInstance.design(join_alg=0)
    ji.start(d)
    check_equal(sj, orc, R, S, ji.finish())
    ji.close()


def parse_timeline(text):
    """RecordLogs lines: (rank_group, rank, index, start, stop, operator, inputs)
    per plan node, then (bytes, packets) per partitioned relation."""
    nodes, rnc = [], []
    for line in text.splitlines():
        if line.startswith("rnc_xfer_bytes"):
            f = line.split()
            assert f[0] == "rnc_xfer_bytes" and f[2] == "rnc_num_packets", line
            rnc.append((int(f[1]), int(f[3])))
            continue
        rg, rank, idx, a, b, rest = line.split("-", 5)
        op, inputs = rest.split("-", 1)
        assert re.fullmatch(r"\d+\.\d{6}", a) and re.fullmatch(r"\d+\.\d{6}", b), line   # %f
        nodes.append((int(rg), int(rank), int(idx), float(a), float(b), op, inputs))
    return nodes, rnc


def node_inputs(n):
    if "input" in n:
        return n["input"]
    return ",".join(n.get("inputs", []))


@pytest.mark.parametrize("replicate", [False, True])
def test_join_instance_plan_and_timeline(gpu, orc, replicate):
    """A plan document set on the instance (the two plans of execution.hpp)
    runs the join (same pairs as the oracle) and RecordLogs prints one line
    per plan node, in plan order, inside the join's wall time."""
    import test_plan

    sj = gpu
    R = sj.gen_r(1 << 18, 45)
    S = sj.gen_s(1 << 20, R, 1.2, 46)
    ji = sj.JoinInstance(16, 1, 16)
    ji.load_column(R, "left")
    ji.load_column(S, "right")
    tree = test_plan.bg_tree() if replicate else test_plan.plain_tree()
    ji.set_query_plan(test_plan.doc(tree))
    d = sj.JoinInstance.design(join_alg=0)   # bank_set 1: the plan decides
    print(f"case plan replicate={replicate}", flush=True)
    t0 = time.clock_gettime_ns(time.CLOCK_MONOTONIC)
    ji.start(d)
    pairs = ji.finish()
    wall_ms = (time.clock_gettime_ns(time.CLOCK_MONOTONIC) - t0) / 1e6
    check_equal(sj, orc, R, S, pairs)
    pt = ji.phase_times()
    if replicate:
        assert pt["heavy_keys"] > 0
    text = ji.timeline(t0, rank_group=3)
    print(text, flush=True)
    nodes, rnc = parse_timeline(text)
    assert [n[5] for n in nodes] == [n["operator"] for n in tree]
    assert [n[6] for n in nodes] == [node_inputs(n) for n in tree]
    assert [n[:3] for n in nodes] == [(3, 0, i) for i in range(len(tree))]
    for n in nodes:
        assert 0.0 <= n[3] <= n[4] <= wall_ms + 0.01, n
    by_op = {}
    for n in nodes:
        by_op.setdefault(n[5], []).append(n)
    sends = by_op["HOST_FUNC_SEND_DATA_OPT"]
    join = [n for n in nodes if n[5].endswith("_JOIN")][0]
    recv = by_op["HOST_FUNC_RECV_DATA_OPT"][0]
    assert max(s[4] for s in sends) <= join[3] + 1e-3 <= recv[3] + 2e-3
    assert join[4] > join[3]
    assert rnc == [(8 * pt["r_partitioned"], pt["r_partitioned"]), (8 * pt["s_partitioned"], pt["s_partitioned"])]
    assert pt["r_partitioned"] == len(R) and 0 < pt["s_partitioned"] <= len(S)
    # without a plan of its own the instance takes the design's: bank_set 8
    # selects the replicating plan
    ji.set_query_plan(None)
    d8 = sj.JoinInstance.design(join_alg=0, bank_set=8)
    ji.start(d8)
    check_equal(sj, orc, R, S, ji.finish())
    ops = [n[5] for n in parse_timeline(ji.timeline(t0))[0]]
    assert ops == [n["operator"] for n in test_plan.bg_tree()]
    ji.close()


def test_join_instance_plan_errors_and_nphj(gpu, orc):
    """A malformed plan is refused without disturbing the instance; a plain
    plan runs join_alg 1 (no-partition join); the timeline of a join whose R
    and S were scattered by one launch gives both partition nodes the same
    interval."""
    import test_plan

    sj = gpu
    R = sj.gen_r(1 << 12, 91)
    S = sj.gen_s(1 << 13, R, 0.9, 92)   # S below 2^14: no heavy keys, so R and S scatter in one launch
    ji = sj.JoinInstance(16, 1, 16)
    ji.load_column(R, "left")
    ji.load_column(S, "right")
    for bad in ["{", test_plan.doc([test_plan.send("left")])]:
        with pytest.raises(sj.SpjError) as e:
            ji.set_query_plan(bad)
        assert e.value.code == -1
    ji.set_query_plan(test_plan.doc(test_plan.plain_tree()))
    ji.start(sj.JoinInstance.design(join_alg=1))
    check_equal(sj, orc, R, S, ji.finish())
    assert ji.phase_times()["s_partitioned"] == 0
    ji.start(sj.JoinInstance.design(join_alg=0))   # small relations: one scatter launch for R and S
    check_equal(sj, orc, R, S, ji.finish())
    nodes, _ = parse_timeline(ji.timeline(0))
    parts = [n for n in nodes if n[5] == "COMPOUND_FUNC_GLB_PARTITION"]
    assert len(parts) == 2 and parts[0][3:5] == parts[1][3:5]
    ji.close()


def test_batch_mixed_sizes(gpu, orc):
    """A batch where one instance has 256 partitions and heavy keys, next to
    small instances: tiles of different instances share scatter workgroups
    and XCD groups."""
    sj = gpu
    Rs = [sj.gen_r(1 << 22, 51), sj.gen_r(5000, 52), sj.gen_r(70000, 53)]
    Ss = [sj.gen_s(1 << 23, Rs[0], 1.5, 61), sj.gen_s(20000, Rs[1], 0.0, 62), sj.gen_s(280000, Rs[2], 1.0, 63)]
    print("case batch mixed sizes", flush=True)
    outs, _ = sj.join_batch(Rs, Ss)
    for R, S, p in zip(Rs, Ss, outs):
        check_equal(sj, orc, R, S, p)


@pytest.mark.parametrize("nR,nS,zipf,alg", [(1 << 26, 1 << 27, 0.5, -1), (1 << 26, 1 << 28, 0.0, -1),
                                              (1 << 26, 1 << 28, 1.0, -1), (1 << 26, 1 << 28, 1.5, -1),
                                              (1 << 24, 1 << 26, 1.0, 1), (1 << 26, 1 << 27, 0.5, 2)])
def test_full_size_checksum(gpu, orc, nR, nS, zipf, alg):
    """Full-size parity through a size-independent property: the pair count
    and the order-independent checksum of the GPU output (computed on the
    device) equal the oracle's closed-form checksum."""
    sj = gpu
    print(f"case full nR={nR} nS={nS} zipf={zipf} alg={alg}", flush=True)
    R = sj.gen_r(nR, 1000)
    S = sj.gen_s(nS, R, zipf, 1001)
    want = orc.join_checksum(R, S)
    dR = sj.DeviceArray.from_numpy(R)
    dS = sj.DeviceArray.from_numpy(S)
    del S
    dO = sj.DeviceArray(nS * 8)
    try:
        n, _ = sj.join_device(dR, nR, dS, nS, dO, sj.config(algorithm=alg) if alg >= 0 else None)
        got = (n,) + sj.checksum_device(dO.ptr, n)
    finally:
        dR.free()
        dS.free()
        dO.free()
    assert got == want


def hottest_keys(S, k):
    keys, counts = np.unique(S["key"], return_counts=True)
    return keys[np.argsort(-counts)[:k]]


@pytest.mark.parametrize("layout", [0, 1])
def test_heavy_keys_with_full_range_r_rids(gpu, orc, layout):
    """R rids are full-range uint32 values (kv_pair_t.rvalue): the hottest
    keys of a Zipf 1.5 S, which the heavy path joins inside the scatter, have
    R rids 0xFFFFFFFF and other extreme values; every other rid is random.
    Counted and cell layouts, heavy handling on."""
    sj = gpu
    rng = np.random.default_rng(2024)
    nR = 1 << 20
    R = sj.gen_r(nR, 131)
    S = sj.gen_s(1 << 22, R, 1.5, 132)
    R["rid"] = rng.integers(0, 1 << 32, nR, dtype=np.uint64).astype(np.uint32)
    hot = hottest_keys(S, 8)
    pos = {int(k): i for i, k in enumerate(R["key"])}
    for j, k in enumerate(hot):
        R["rid"][pos[int(k)]] = [0xFFFFFFFF, 0, 0xFFFFFFFE, 0x80000000][j % 4]
    print(f"case heavy full-range rids layout={layout}", flush=True)
    pairs, t = sj.join(R, S, sj.config(algorithm=0, heavy=1, part_layout=layout), timed=True)
    check_equal(sj, orc, R, S, pairs)
    assert t["heavy_keys"] >= 8
    assert t["s_partitioned"] < len(S) // 2   # the hot keys went through the heavy path
    # the same with a heavy key that has no R tuple (its S tuples are dropped)
    R2 = R[R["key"] != hot[1]]
    pairs, _ = sj.join(R2, S, sj.config(algorithm=0, heavy=1, part_layout=layout))
    check_equal(sj, orc, R2, S, pairs)


@pytest.mark.parametrize("bits", [0, 1, 2])
def test_sort_merge_duplicate_across_runs(gpu, orc, bits):
    """Sort-merge partitions of several sorted runs (2^17 R tuples over 1, 2
    or 4 partitions: 2 to 8 passes of 16K): a key repeated in two different
    runs of one partition is a duplicate (SPJ_EDUPKEY), as is one repeated
    inside a run; without it the result matches the oracle."""
    sj = gpu
    R = sj.gen_r(1 << 17, 140 + bits)
    S = sj.gen_s(1 << 18, R, 0.3, 150 + bits)
    print(f"case smj runs bits={bits}", flush=True)
    pairs, _ = sj.join(R, S, sj.config(algorithm=2, radix_bits=bits))
    check_equal(sj, orc, R, S, pairs)
    # partition of each key (the oracle's SuperFastHash, low bits)
    part = orc.hash_batch("glb_partition_hash", R["key"]) & np.uint32((1 << bits) - 1)
    p0 = np.flatnonzero(part == 0)
    assert len(p0) > 16384 + 1024   # more than one pass
    # Input tiles of 8192 tuples are dealt to 8 scatter groups in turn, and a
    # partition holds group 0's tuples first and group 7's last: a key early
    # in the input and one late in it sit in the partition's first and last
    # runs.  The second pair may share a run.
    for a, b in [(p0[10], p0[-10]), (p0[5], p0[len(p0) // 2])]:
        R2 = R.copy()
        R2["key"][b] = R2["key"][a]
        with pytest.raises(sj.SpjError) as e:
            sj.join(R2, S, sj.config(algorithm=2, radix_bits=bits))
        assert e.value.code == -4, (a, b)


def sub_table(h, bits, Q):
    """The sub-table of an R tuple with SuperFastHash h in a partition split
    Q ways (high hash bits, multiplied by 0x9E3779B1; the top of the 64-bit
    product with Q).  Used only to build inputs that crowd one sub-table."""
    x = ((h.astype(np.uint64) >> np.uint64(bits)) * np.uint64(0x9E3779B1)) & np.uint64(0xFFFFFFFF)
    return (x * np.uint64(Q)) >> np.uint64(32)


@pytest.mark.parametrize("layout", [0, 1])
def test_oversized_partition_resplit_and_sort_merge_fallback(gpu, orc, layout):
    """R partitions larger than one LDS table, with keys chosen so that a
    sub-table overflows: once (the partition is split again, into twice the
    sub-tables) and for good (every key in the same eighth of the hash
    range: the partition goes to sort-merge).  Same multiset as the oracle,
    and a duplicate key in the sort-merge fallback is still reported."""
    sj = gpu
    cand = np.arange(1, 1 << 22, dtype=np.uint32)
    h = orc.hash_batch("glb_partition_hash", cand)
    # radix_bits 0: one partition; Q starts at ceil(n / 16384)
    # (a) 20000 keys, all in sub-table 0 of 2 and of 4, spread over 8
    #     sub-tables below: re-split twice, then fits
    s8 = sub_table(h, 0, 8)
    a = np.concatenate([cand[s8 == 0][:10000], cand[s8 == 1][:10000]])
    # (b) 20000 keys in sub-table 0 of 8 (and so of 2 and 4): sort-merge
    b = cand[s8 == 0][:20000]
    rng = np.random.default_rng(7)
    for name, keys in (("resplit", a), ("fallback", b)):
        keys = rng.permutation(keys)
        R = np.zeros(len(keys), dtype=sj.TUPLE)
        R["key"] = keys
        R["rid"] = np.arange(len(keys), dtype=np.uint32)
        S = np.zeros(len(keys) * 3, dtype=sj.TUPLE)
        S["key"] = keys[rng.integers(0, len(keys), len(S))]
        S["rid"] = np.arange(len(S), dtype=np.uint32)
        print(f"case oversized {name} layout={layout}", flush=True)
        cfg = sj.config(algorithm=0, radix_bits=0, part_layout=layout, heavy=0)
        pairs, _ = sj.join(R, S, cfg)
        check_equal(sj, orc, R, S, pairs)
        R2 = R.copy()
        R2["key"][3] = R2["key"][len(R2) - 5]
        with pytest.raises(sj.SpjError) as e:
            sj.join(R2, S, cfg)
        assert e.value.code == -4


@pytest.mark.parametrize("layout", [0, 1])
def test_sort_merge_key_range_buckets(gpu, orc, layout):
    """Sort-merge of partitions that fit one table: R is counting-sorted into
    key-range buckets and each bucket sorted in place.  Keys at the ends of
    the uint32 range (1 and 0xFFFFFFFF), S keys below and above R's range,
    a partition whose keys crowd a few buckets (the items go to the run
    sort), and a duplicate inside one range bucket."""
    sj = gpu
    rng = np.random.default_rng(321)
    n = 1 << 18
    keys = rng.permutation(np.unique(rng.integers(2, 1 << 31, 2 * n, dtype=np.uint64)))[:n].astype(np.uint32)
    keys[:2] = [1, 0xFFFFFFFF]
    R = np.zeros(n, dtype=sj.TUPLE)
    R["key"] = keys
    R["rid"] = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
    S = np.zeros(4 * n, dtype=sj.TUPLE)
    S["key"] = keys[rng.integers(0, n, len(S))]
    S["key"][::50] = rng.integers(1 << 31, 1 << 32, len(S[::50]), dtype=np.uint64).astype(np.uint32)   # mostly absent
    S["rid"] = np.arange(len(S), dtype=np.uint32)
    cfg = sj.config(algorithm=2, part_layout=layout)
    print(f"case smj range layout={layout}", flush=True)
    pairs, _ = sj.join(R, S, cfg)
    check_equal(sj, orc, R, S, pairs)
    # clustered keys: a dense run of consecutive keys next to a few far ones
    # puts thousands of keys in one range bucket of each partition
    R2 = R.copy()
    R2["key"][: n // 2] = np.arange(1000, 1000 + n // 2, dtype=np.uint32)
    far = rng.permutation(np.unique(rng.integers(1 << 30, 1 << 31, 2 * n, dtype=np.uint64)))[: n - n // 2]
    R2["key"][n // 2:] = far.astype(np.uint32)
    S2 = S.copy()
    S2["key"] = R2["key"][rng.integers(0, n, len(S2))]
    pairs, _ = sj.join(R2, S2, cfg)
    check_equal(sj, orc, R2, S2, pairs)
    # a duplicate: the same key twice lands in the same range bucket
    for R3 in (R.copy(), R2.copy()):
        R3["key"][7] = R3["key"][n - 3]
        with pytest.raises(sj.SpjError) as e:
            sj.join(R3, S, cfg)
        assert e.value.code == -4


def range_buckets(keys, nb=8192):
    """Key-range bucket of each R key of a sort-merge partition: (key - kmin)
    * floor(nb 2^32 / (kmax - kmin + 1)) >> 32.  Used only to check that an
    input reaches the bucket counts a test is about."""
    k = keys.astype(np.uint64)
    lo, hi = int(k.min()), int(k.max())
    scale = np.uint64(min((nb << 32) // (hi - lo + 1), 0xFFFFFFFF))
    return ((k - np.uint64(lo)) * scale) >> np.uint64(32)


def test_sort_merge_many_long_buckets(gpu, orc):
    """One sort-merge partition (radix bits 0, 15K R tuples, one table)
    whose keys come in clusters of 3 to 8 consecutive keys, one cluster per
    key-range bucket, next to buckets of one and two keys: more buckets of
    three or more entries than the list that deals them to the threads
    holds (3584), so some are sorted where they are found.  Full-range R
    rids; a duplicate inside a long bucket is SPJ_EDUPKEY wherever it is
    sorted."""
    sj = gpu
    rng = np.random.default_rng(77)
    sizes = np.concatenate([np.full(1850, 3), np.full(1850, 4), 5 + np.arange(120) % 4,
                            np.full(400, 2), np.full(400, 1)])
    sizes = sizes[rng.permutation(len(sizes))]
    # cluster c starts at 2000 c + 10: bucket width ~1000 keys, so a cluster
    # of at most 8 consecutive keys shares a bucket with no other cluster
    keys = np.concatenate([np.arange(s, dtype=np.uint32) + np.uint32(2000 * c + 10) for c, s in enumerate(sizes)])
    n = len(keys)
    assert n <= 17400
    counts = np.bincount(range_buckets(keys).astype(np.int64))
    assert (counts >= 3).sum() > 3584 and counts.max() <= 8
    R = np.zeros(n, dtype=sj.TUPLE)
    perm = rng.permutation(n)
    R["key"] = keys[perm]
    R["rid"] = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
    S = np.zeros(4 * n, dtype=sj.TUPLE)
    S["key"] = R["key"][rng.integers(0, n, len(S))]
    S["key"][::9] += np.uint32(1000)   # between clusters: unmatched
    S["rid"] = np.arange(len(S), dtype=np.uint32)
    cfg = sj.config(algorithm=2, radix_bits=0)
    print("case smj long buckets", flush=True)
    pairs, _ = sj.join(R, S, cfg)
    check_equal(sj, orc, R, S, pairs)
    # duplicates in long buckets at the start, middle and end of the range
    for c in (3, len(sizes) // 2, len(sizes) - 3):
        if sizes[c] < 3:
            c += 1 if sizes[c + 1] >= 3 else 2
        first = int(sizes[:c].sum())
        R2 = R.copy()
        i0, i1 = np.flatnonzero(perm == first)[0], np.flatnonzero(perm == first + 2)[0]
        R2["key"][i1] = R2["key"][i0]
        with pytest.raises(sj.SpjError) as e:
            sj.join(R2, S, cfg)
        assert e.value.code == -4, c


@pytest.mark.parametrize("ranks,alg", [(16, 0), (4, 2), (8, 3)])
def test_rank_padded_result_layout(gpu, orc, ranks, alg):
    """FinishJoin in the reference's gathered layout: one buffer per rank of
    64 DPU blocks, each the pairs whose S tuple's global partition
    glb_partition_hash(key) & (64 ranks - 1) is that DPU, then (0, 0) pairs up
    to the rank's largest block (queryHostFunc.cc:1058-1126, finish_join.c)."""
    sj = gpu
    nR = 1 << 18 if alg != 3 else 1 << 16
    R = sj.gen_r(nR, 210 + ranks)
    S = sj.gen_s(2 * nR, R, 0.6, 211 + ranks)
    S["key"][::37] += np.uint32(1 << 24)   # unmatched: no pair, no slot
    D = 64 * ranks
    order = np.argsort(R["key"], kind="stable")
    rk = R["key"][order]
    pos = np.minimum(np.searchsorted(rk, S["key"]), len(rk) - 1)
    hit = (rk[pos] == S["key"]) & (S["key"] != 0)
    dpu = orc.hash_batch("glb_partition_hash", S["key"][hit]) & np.uint32(D - 1)
    want = (R["rid"][order][pos[hit]].astype(np.uint64) << np.uint64(32)) | S["rid"][hit].astype(np.uint64)
    count = np.bincount(dpu, minlength=D)
    ji = sj.JoinInstance(ranks, 1, ranks)
    ji.load_column(R, "left")
    ji.load_column(S, "right")
    ji.set_result_layout(sj.JoinInstance.RESULT_RANK_PADDED)
    print(f"case rank-padded ranks={ranks} alg={alg}", flush=True)
    ji.start(sj.JoinInstance.design(join_alg=alg))
    bufs = ji.finish_buffers()
    assert len(bufs) == ranks
    by_dpu = np.argsort(dpu, kind="stable")
    starts = np.concatenate([[0], np.cumsum(count)])
    for r in range(ranks):
        block = int(count[r * 64:(r + 1) * 64].max())
        assert len(bufs[r]) == 64 * block, r
        w = words(bufs[r])
        for i in range(64):
            d = r * 64 + i
            c = int(count[d])
            blk = w[i * block:(i + 1) * block]
            assert np.array_equal(np.sort(blk[:c]), np.sort(want[by_dpu[starts[d]:starts[d + 1]]])), d
            assert not blk[c:].any(), d
    # the exact layout is back on request; replication and no-partition joins refuse the padded one
    ji.set_result_layout(sj.JoinInstance.RESULT_EXACT)
    ji.start(sj.JoinInstance.design(join_alg=alg))
    check_equal(sj, orc, R, S, ji.finish())
    ji.set_result_layout(sj.JoinInstance.RESULT_RANK_PADDED)
    for d in (sj.JoinInstance.design(join_alg=0, bank_set=8), sj.JoinInstance.design(join_alg=1)):
        with pytest.raises(sj.SpjError) as e:
            ji.start(d)
        assert e.value.code == -1
    ji.close()
