#!/usr/bin/env python3
"""Writes the committed test fixtures (run in the development container,
where the reference tree is mounted and oracle/_ref is built):

  hash_kat.json   -- outputs of the reference's own hash functions
                     (src/dpu/src/hash.c compiled from the reference sources by
                     oracle/Makefile into oracle/_ref/libref_hash.so) on a fixed
                     key set, plus its crctab words;
  join_small.json -- a small seeded join (R 300 x S 1000, Zipf 0.9) from the
                     library's generator and its expected sorted pairs from the
                     closed-form oracle;
  gen_prefix.json -- the first tuples the generator emits for fixed seeds.

The reference itself is never imported or run; only the hash functions
compiled from its sources are called, through ctypes.
"""
import ctypes
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "spid-join_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle  # noqa: E402
import spidjoin as sj  # noqa: E402


def kat_keys():
    rng = np.random.default_rng(20240607)
    fixed = [0, 1, 2, 3, 4, 5, 255, 256, 257, 65535, 65536, 65537, (1 << 20), (1 << 24) - 1, (1 << 26),
             0x7FFFFFFF, 0x80000000, 0xDEADBEEF, 0xFFFFFFFE, 0xFFFFFFFF]
    return fixed + [int(x) for x in rng.integers(0, 1 << 32, 236, dtype=np.uint64)]


def hash_kat():
    lib = ctypes.CDLL(oracle.REF_HASH)
    names = ["glb_partition_hash", "join_hash", "local_partition_hash", "join_hash2"]
    for n in names:
        getattr(lib, n).restype = ctypes.c_uint32
        getattr(lib, n).argtypes = [ctypes.c_uint32]
    keys = kat_keys()
    out = {"source": "reference src/dpu/src/hash.c compiled by oracle/Makefile (oracle/_ref/libref_hash.so)",
           "keys": keys}
    for n in names:
        out[n] = [int(getattr(lib, n)(k)) for k in keys]
    tab = (ctypes.c_uint32 * 256).in_dll(lib, "crctab")
    out["crctab"] = [int(tab[i]) for i in range(256)]
    return out


def join_small():
    R = sj.gen_r(300, 77)
    S = sj.gen_s(1000, R, 0.9, 78)
    exp = oracle.expected_pairs(R, S)
    return {"R": [[int(a), int(b)] for a, b in zip(R["key"], R["rid"])],
            "S": [[int(a), int(b)] for a, b in zip(S["key"], S["rid"])],
            "pairs": [[int(a), int(b)] for a, b in zip(exp["r_rid"], exp["s_rid"])]}


def gen_prefix():
    out = {}
    for seed in (1, 1000):
        R = sj.gen_r(1 << 16, seed)
        out[f"R_{seed}"] = [int(x) for x in R["key"][:16]]
        for z in (0.0, 0.5, 1.5):
            S = sj.gen_s(1 << 17, R, z, seed + 1)
            out[f"S_{seed}_{z}"] = [int(x) for x in S["key"][:16]] + [int(x) for x in S["key"][-4:]]
    return out


if __name__ == "__main__":
    if not os.path.exists(oracle.REF_HASH):
        sys.exit("oracle/_ref/libref_hash.so is not built (make -C oracle; needs the reference tree)")
    for name, fn in (("hash_kat.json", hash_kat), ("join_small.json", join_small), ("gen_prefix.json", gen_prefix)):
        with open(os.path.join(HERE, name), "w") as f:
            json.dump(fn(), f, separators=(",", ":"))
        print("wrote", name)
