"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker / reported CPU baseline.  The
product (libspidjoin.so and spid-join_amd/) never imports it.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_HASH = os.path.join(HERE, "_ref", "libref_hash.so")

TUPLE = np.dtype([("key", "<u4"), ("rid", "<u4")])
PAIR = np.dtype([("r_rid", "<u4"), ("s_rid", "<u4")])

_l = None


class Design(ctypes.Structure):
    _fields_ = [("rank_num", ctypes.c_int), ("rank_set", ctypes.c_int), ("bank_set", ctypes.c_int),
                ("join_alg", ctypes.c_int), ("threads", ctypes.c_int)]


def lib():
    global _l
    if _l is None:
        if not os.path.exists(LIB):
            raise OSError(f"{LIB} not built (make -C oracle)")
        L = ctypes.CDLL(LIB)
        P, U = ctypes.c_void_p, ctypes.c_uint64
        L.or_crc_table.restype = ctypes.POINTER(ctypes.c_uint32)
        for f in ("or_glb_partition_hash", "or_join_hash", "or_local_partition_hash", "or_join_hash2"):
            getattr(L, f).restype = ctypes.c_uint32
            getattr(L, f).argtypes = [ctypes.c_uint32]
        L.or_hash_batch.argtypes = [ctypes.c_int, P, P, U]
        L.or_join_expected.restype = ctypes.c_int64
        L.or_join_expected.argtypes = [P, U, P, U, P, U]
        L.or_pairs_checksum.argtypes = [P, U, ctypes.POINTER(U), ctypes.POINTER(U)]
        L.or_join_checksum.restype = ctypes.c_int
        L.or_join_checksum.argtypes = [P, U, P, U, ctypes.c_int, ctypes.POINTER(U), ctypes.POINTER(U),
                                       ctypes.POINTER(U)]
        L.or_sort_pairs.argtypes = [P, U]
        L.or_spid_join_cpu.restype = ctypes.c_int64
        L.or_spid_join_cpu.argtypes = [P, U, P, U, ctypes.POINTER(Design), P, U]
        _l = L
    return _l


def crc_table():
    t = lib().or_crc_table()
    return np.array([t[i] for i in range(256)], dtype=np.uint32)


HASHES = {"glb_partition_hash": 0, "join_hash": 1, "local_partition_hash": 2, "join_hash2": 3}


def hash_batch(which, keys):
    keys = np.ascontiguousarray(keys, dtype=np.uint32)
    out = np.empty_like(keys)
    lib().or_hash_batch(HASHES[which], keys.ctypes.data, out.ctypes.data, len(keys))
    return out


def expected_pairs(R, S):
    """Sorted expected (R rid, S rid) pairs of the equi-join."""
    out = np.empty(max(len(S), 1) * 2 if len(R) else 1, dtype=PAIR)
    n = lib().or_join_expected(R.ctypes.data, len(R), S.ctypes.data, len(S), out.ctypes.data, len(out))
    if n < 0:
        raise RuntimeError(f"or_join_expected failed: {n}")
    return out[:n]


def pairs_checksum(p):
    s, x = ctypes.c_uint64(), ctypes.c_uint64()
    p = np.ascontiguousarray(p)
    lib().or_pairs_checksum(p.ctypes.data, len(p), ctypes.byref(s), ctypes.byref(x))
    return s.value, x.value


def join_checksum(R, S, threads=0):
    c, s, x = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
    rc = lib().or_join_checksum(R.ctypes.data, len(R), S.ctypes.data, len(S), threads, ctypes.byref(c),
                                ctypes.byref(s), ctypes.byref(x))
    if rc:
        raise RuntimeError("or_join_checksum failed")
    return c.value, s.value, x.value


def sort_pairs(p):
    p = np.ascontiguousarray(p).copy()
    lib().or_sort_pairs(p.ctypes.data, len(p))
    return p


def spid_join_cpu(R, S, rank_num=16, rank_set=1, bank_set=1, join_alg=0, threads=0):
    """CPU restatement of the reference join; returns pairs (per-DPU order)."""
    d = Design(rank_num, rank_set, bank_set, join_alg, threads)
    out = np.empty(max(len(S), 1), dtype=PAIR)
    n = lib().or_spid_join_cpu(R.ctypes.data, len(R), S.ctypes.data, len(S), ctypes.byref(d), out.ctypes.data,
                               len(out))
    if n < 0:
        raise RuntimeError(f"or_spid_join_cpu failed: {n}")
    return out[:n]
