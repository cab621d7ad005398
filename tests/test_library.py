"""CPU: libspidjoin.so loads and exports the C ABI; host logic that needs no
GPU (generator, argument checks, life-cycle errors, tuner table)."""
import ctypes
import json
import os
import re

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def declared_functions():
    src = open(os.path.join(ROOT, "include", "spidjoin.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(spj_[a-z0-9_]+)\s*\(", src)))


def test_exports_every_declared_symbol(sj):
    names = declared_functions()
    assert len(names) >= 30
    lib = sj.lib()
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(sj.EXPORTS) == names


def test_generator_properties(sj):
    R = sj.gen_r(1 << 16, 3)
    assert np.array_equal(np.sort(R["key"]), np.arange(1, (1 << 16) + 1, dtype=np.uint32))
    assert np.array_equal(R["rid"], np.arange(1 << 16, dtype=np.uint32))
    for z in (0.0, 0.5, 1.5):
        S = sj.gen_s(1 << 17, R, z, 4)
        assert np.isin(S["key"], R["key"]).all()
        assert np.array_equal(S["rid"], np.arange(1 << 17, dtype=np.uint32))
    # Zipf 1.5 over 2^16 ranks: the top key holds 1 / H(2^16, 1.5) of S
    S = sj.gen_s(1 << 20, R, 1.5, 5)
    top = np.bincount(S["key"]).max() / len(S)
    h = (np.arange(1, (1 << 16) + 1, dtype=np.float64) ** -1.5).sum()
    assert abs(top - 1 / h) < 0.01
    # uniform: no key far above its expectation
    S = sj.gen_s(1 << 20, R, 0.0, 6)
    assert np.bincount(S["key"]).max() < 16 * (1 << 20) / (1 << 16)


def test_generator_deterministic_and_thread_independent(sj):
    R = sj.gen_r(1 << 15, 9)
    a = sj.gen_s(300_000, R, 1.2, 10, threads=1)
    b = sj.gen_s(300_000, R, 1.2, 10, threads=7)
    assert np.array_equal(a, b)
    g = json.load(open(os.path.join(HERE, "golden", "gen_prefix.json")))
    for seed in (1, 1000):
        R = sj.gen_r(1 << 16, seed)
        assert R["key"][:16].tolist() == g[f"R_{seed}"]
        for z in (0.0, 0.5, 1.5):
            S = sj.gen_s(1 << 17, R, z, seed + 1)
            assert S["key"][:16].tolist() + S["key"][-4:].tolist() == g[f"S_{seed}_{z}"]


def test_config_defaults(sj):
    c = sj.config()
    assert c.algorithm == -1 and c.radix_bits == -1 and c.tile_tuples == -1


def test_instance_argument_and_state_errors(sj):
    ji = sj.JoinInstance(16, 1, 16)
    with pytest.raises(sj.SpjError) as e:
        ji.load_column(np.zeros(4, dtype=sj.TUPLE), "middle")
    assert e.value.code == -1
    with pytest.raises(sj.SpjError) as e:
        ji.finish()
    assert e.value.code == -6
    with pytest.raises(sj.SpjError) as e:
        ji.prepare()   # before both LoadColumn calls
    assert e.value.code == -6
    ji.close()


def test_tuner_table_roundtrip(sj, tmp_path):
    L = sj.lib()
    sj.check(L.spj_tuner_clear())
    c = sj.config(radix_bits=9, scatter_ipt=8, table_log2=11)
    sj.check(L.spj_tuner_set(1 << 22, 1 << 24, 1.0, ctypes.byref(c)))
    assert L.spj_tuner_entries() == 1
    got = sj.Config()
    sj.check(L.spj_tuner_get((1 << 22) - 5, (1 << 24) - 20, 1.1, ctypes.byref(got)))
    assert (got.radix_bits, got.scatter_ipt, got.table_log2) == (9, 8, 11)
    p = str(tmp_path / "tune.txt")
    sj.check(L.spj_tuner_save(p.encode()))
    sj.check(L.spj_tuner_clear())
    assert L.spj_tuner_entries() == 0
    sj.check(L.spj_tuner_load(p.encode()))
    assert L.spj_tuner_entries() == 1
    sj.check(L.spj_tuner_clear())
    bad = sj.config(radix_bits=20)
    assert L.spj_tuner_set(1 << 20, 1 << 21, 0.0, ctypes.byref(bad)) == -1
