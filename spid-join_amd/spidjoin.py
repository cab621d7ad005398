"""ctypes binding of libspidjoin.so (the C ABI in include/spidjoin.h).

This is the Python host of the library: tests, bench.py and Python callers
drive the HIP join through it.  It never computes a join itself; if the
shared library is missing it raises instead of falling back to anything.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libspidjoin.so")

TUPLE = np.dtype([("key", "<u4"), ("rid", "<u4")])   # kv_pair_t
PAIR = np.dtype([("r_rid", "<u4"), ("s_rid", "<u4")])

ALG_AUTO, ALG_PHJ, ALG_NPHJ, ALG_SMJ, ALG_NL = -1, 0, 1, 2, 3

ERRORS = {0: "OK", -1: "EINVAL", -2: "EHIP", -3: "ENOMEM", -4: "EDUPKEY", -5: "EOVERFLOW",
          -6: "ESTATE", -7: "EBUSY"}

# Every function the C header declares (checked by the CPU tests).
EXPORTS = [
    "spj_config_default", "spj_last_error", "spj_init", "spj_device_info", "spj_dev_alloc", "spj_dev_free",
    "spj_h2d", "spj_d2h", "spj_dev_memset", "spj_d2d", "spj_pairs_checksum_device", "spj_dev_sync", "spj_join_device",
    "spj_partition_device", "spj_join_device_noalloc", "spj_graph_stats", "spj_join_host",
    "spj_join_batch_device", "spj_join_design_default", "spj_create", "spj_load_column", "spj_prepare",
    "spj_start", "spj_finish", "spj_execute", "spj_last_phase_times", "spj_set_query_plan", "spj_plan_check",
    "spj_timeline", "spj_set_result_layout", "spj_finish_buffers", "spj_free_result", "spj_destroy",
    "spj_tuner_set", "spj_tuner_get", "spj_tuner_clear", "spj_tuner_entries", "spj_tuner_save",
    "spj_tuner_load", "spj_autotune", "spj_gen_r", "spj_gen_s", "spj_server_create", "spj_server_add_table",
    "spj_server_submit", "spj_server_wait", "spj_server_stats", "spj_server_destroy",
]


class SpjError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Config(ctypes.Structure):
    _fields_ = [("algorithm", ctypes.c_int32), ("radix_bits", ctypes.c_int32), ("table_log2", ctypes.c_int32),
                ("heavy", ctypes.c_int32), ("heavy_min_count", ctypes.c_int32), ("scatter_ipt", ctypes.c_int32),
                ("tile_tuples", ctypes.c_int64), ("items_s_max", ctypes.c_int64),
                ("part_layout", ctypes.c_int32), ("cell_capacity_pct", ctypes.c_int32),
                ("radix_passes", ctypes.c_int32)]


class PhaseTimes(ctypes.Structure):
    _fields_ = [(n, ctypes.c_float) for n in
                ("heavy_ms", "hist_ms", "scan_ms", "scatter_ms", "items_ms", "join_ms", "total_ms")] + [
        ("heavy_keys", ctypes.c_uint32), ("r_partitioned", ctypes.c_uint64), ("s_partitioned", ctypes.c_uint64),
        ("s_cells", ctypes.c_uint32), ("r_cells", ctypes.c_uint32), ("passes", ctypes.c_uint32)]

    def as_dict(self):
        return {n: (float(getattr(self, n)) if t is ctypes.c_float else int(getattr(self, n)))
                for n, t in self._fields_}


class JoinDesign(ctypes.Structure):
    _fields_ = [("input_size", ctypes.c_int64), ("rank_num", ctypes.c_int32), ("zipf_factor", ctypes.c_int32),
                ("rs_ratio", ctypes.c_int32), ("rank_set", ctypes.c_int32), ("bank_set", ctypes.c_int32),
                ("join_alg", ctypes.c_int32)]


_lib = None
P = ctypes.c_void_p
U64 = ctypes.c_uint64


def lib():
    """Load libspidjoin.so (raises OSError when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OSError(f"{LIB_PATH} not built: run __graft_entry__.build() or make -C spid-join_amd/csrc")
    L = ctypes.CDLL(LIB_PATH)
    sig = {
        "spj_config_default": (None, [ctypes.POINTER(Config)]),
        "spj_last_error": (ctypes.c_char_p, []),
        "spj_init": (ctypes.c_int, [ctypes.c_int]),
        "spj_device_info": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int), ctypes.POINTER(U64), ctypes.c_char_p,
                                           ctypes.c_int]),
        "spj_dev_alloc": (ctypes.c_int, [ctypes.POINTER(P), U64]),
        "spj_dev_free": (ctypes.c_int, [P]),
        "spj_h2d": (ctypes.c_int, [P, P, U64]),
        "spj_d2h": (ctypes.c_int, [P, P, U64]),
        "spj_dev_memset": (ctypes.c_int, [P, ctypes.c_int, U64]),
        "spj_dev_sync": (ctypes.c_int, []),
        "spj_d2d": (ctypes.c_int, [P, P, U64, ctypes.POINTER(ctypes.c_float)]),
        "spj_pairs_checksum_device": (ctypes.c_int, [P, U64, ctypes.POINTER(U64), ctypes.POINTER(U64)]),
        "spj_join_device": (ctypes.c_int, [P, U64, P, U64, P, U64, ctypes.POINTER(U64), ctypes.POINTER(Config),
                                           ctypes.POINTER(PhaseTimes)]),
        "spj_partition_device": (ctypes.c_int, [P, U64, ctypes.c_int, P, P]),
        "spj_join_device_noalloc": (ctypes.c_int, [P, U64, P, U64, P, U64, ctypes.POINTER(U64)]),
        "spj_graph_stats": (ctypes.c_int, [ctypes.POINTER(U64), ctypes.POINTER(U64)]),
        "spj_join_host": (ctypes.c_int, [P, U64, P, U64, P, U64, ctypes.POINTER(U64), ctypes.POINTER(Config)]),
        "spj_join_batch_device": (ctypes.c_int, [ctypes.c_int, P, P, P, P, P, P, ctypes.POINTER(Config),
                                                 ctypes.POINTER(PhaseTimes)]),
        "spj_join_design_default": (None, [ctypes.POINTER(JoinDesign)]),
        "spj_create": (P, [ctypes.c_int, ctypes.c_int, ctypes.c_int]),
        "spj_load_column": (ctypes.c_int, [P, P, ctypes.c_int64, ctypes.c_char_p]),
        "spj_prepare": (ctypes.c_int, [P, ctypes.POINTER(JoinDesign)]),
        "spj_start": (ctypes.c_int, [P, ctypes.POINTER(JoinDesign)]),
        "spj_finish": (ctypes.c_int, [P, ctypes.POINTER(P), ctypes.POINTER(U64)]),
        "spj_execute": (ctypes.c_int, [P, ctypes.c_char_p, ctypes.POINTER(JoinDesign), ctypes.POINTER(P),
                                       ctypes.POINTER(U64)]),
        "spj_last_phase_times": (ctypes.c_int, [P, ctypes.POINTER(PhaseTimes)]),
        "spj_set_query_plan": (ctypes.c_int, [P, ctypes.c_char_p]),
        "spj_plan_check": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
        "spj_timeline": (ctypes.c_int, [P, ctypes.c_int64, ctypes.c_int64, ctypes.c_int, ctypes.c_char_p, U64,
                                        ctypes.POINTER(U64)]),
        "spj_set_result_layout": (ctypes.c_int, [P, ctypes.c_int]),
        "spj_finish_buffers": (ctypes.c_int, [P, P, P, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
        "spj_free_result": (None, [P]),
        "spj_destroy": (None, [P]),
        "spj_tuner_set": (ctypes.c_int, [U64, U64, ctypes.c_double, ctypes.POINTER(Config)]),
        "spj_tuner_get": (ctypes.c_int, [U64, U64, ctypes.c_double, ctypes.POINTER(Config)]),
        "spj_tuner_clear": (ctypes.c_int, []),
        "spj_tuner_entries": (ctypes.c_int, []),
        "spj_tuner_save": (ctypes.c_int, [ctypes.c_char_p]),
        "spj_tuner_load": (ctypes.c_int, [ctypes.c_char_p]),
        "spj_autotune": (ctypes.c_int, [P, U64, P, U64, P, ctypes.c_double, ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_float)]),
        "spj_server_create": (P, [U64, ctypes.c_int]),
        "spj_server_add_table": (ctypes.c_int, [P, P, U64, ctypes.POINTER(ctypes.c_int)]),
        "spj_server_submit": (ctypes.c_int, [P, ctypes.c_int, P, U64, ctypes.POINTER(U64)]),
        "spj_server_wait": (ctypes.c_int, [P, U64, ctypes.POINTER(P), ctypes.POINTER(U64),
                                           ctypes.POINTER(ctypes.c_double)]),
        "spj_server_stats": (ctypes.c_int, [P, ctypes.POINTER(U64), ctypes.POINTER(U64)]),
        "spj_server_destroy": (None, [P]),
        "spj_gen_r": (ctypes.c_int, [P, U64, U64]),
        "spj_gen_s": (ctypes.c_int, [P, U64, P, U64, ctypes.c_double, U64, ctypes.c_int]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(rc):
    if rc != 0:
        raise SpjError(rc, lib().spj_last_error().decode(errors="replace"))
    return rc


def config(**kw):
    c = Config()
    lib().spj_config_default(ctypes.byref(c))
    for k, v in kw.items():
        if not hasattr(c, k):
            raise KeyError(k)
        setattr(c, k, v)
    return c


# ---------------------------------------------------------------- generator
def gen_r(n, seed):
    R = np.empty(n, dtype=TUPLE)
    check(lib().spj_gen_r(R.ctypes.data, n, seed))
    return R


def gen_s(n, R, zipf, seed, threads=0):
    S = np.empty(n, dtype=TUPLE)
    check(lib().spj_gen_s(S.ctypes.data, n, R.ctypes.data, len(R), float(zipf), seed, threads))
    return S


# ---------------------------------------------------------------- device memory
class DeviceArray:
    """A device allocation owned by this object (hipMalloc through the C ABI)."""

    def __init__(self, nbytes):
        p = P()
        check(lib().spj_dev_alloc(ctypes.byref(p), max(int(nbytes), 1)))
        self.ptr = p.value
        self.nbytes = int(nbytes)

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes)
        if a.nbytes:
            check(lib().spj_h2d(d.ptr, a.ctypes.data, a.nbytes))
        return d

    def to_numpy(self, dtype, count):
        out = np.empty(count, dtype=dtype)
        nb = count * np.dtype(dtype).itemsize
        if nb > self.nbytes:
            raise ValueError("read past the allocation")
        if nb:
            check(lib().spj_d2h(out.ctypes.data, self.ptr, nb))
        return out

    def free(self):
        if self.ptr:
            check(lib().spj_dev_free(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def checksum_device(ptr, n):
    """(sum, xor) of the pairs at device address ptr (see spj_pairs_checksum_device)."""
    s, x = U64(), U64()
    check(lib().spj_pairs_checksum_device(ptr, n, ctypes.byref(s), ctypes.byref(x)))
    return s.value, x.value


def autotune(dR, nR, dS, nS, dOut, zipf, reps=3):
    """spj_autotune on device relations; returns the winner's median ms."""
    ms = ctypes.c_float()
    check(lib().spj_autotune(dR.ptr, nR, dS.ptr, nS, dOut.ptr, float(zipf), int(reps), ctypes.byref(ms)))
    return ms.value


def tuner_get(nR, nS, zipf):
    """The parameters a single call on relations of this shape would run with."""
    c = Config()
    check(lib().spj_tuner_get(nR, nS, float(zipf), ctypes.byref(c)))
    return {n: getattr(c, n) for n, _ in Config._fields_}


def graph_stats():
    """(captures, replays) of launch sequences replayed as HIP graphs."""
    c, r = U64(), U64()
    check(lib().spj_graph_stats(ctypes.byref(c), ctypes.byref(r)))
    return c.value, r.value


def init(device=0):
    check(lib().spj_init(device))


def device_info():
    cus = ctypes.c_int()
    lds = U64()
    arch = ctypes.create_string_buffer(64)
    check(lib().spj_device_info(ctypes.byref(cus), ctypes.byref(lds), arch, 64))
    return {"cus": cus.value, "lds_per_block": lds.value, "arch": arch.value.decode()}


def join_device(dR, nR, dS, nS, dOut, cfg=None, timed=False):
    """Join device-resident relations; returns (count, phase-times dict or None)."""
    cnt = U64()
    t = PhaseTimes()
    check(lib().spj_join_device(dR.ptr if nR else None, nR, dS.ptr if nS else None, nS, dOut.ptr, nS,
                                ctypes.byref(cnt), ctypes.byref(cfg) if cfg is not None else None,
                                ctypes.byref(t) if timed else None))
    return cnt.value, (t.as_dict() if timed else None)


def join(R, S, cfg=None, timed=False):
    """Join host numpy relations on the GPU; returns (pairs, times)."""
    dR = DeviceArray.from_numpy(R)
    dS = DeviceArray.from_numpy(S)
    dO = DeviceArray(max(len(S), 1) * 8)
    try:
        n, t = join_device(dR, len(R), dS, len(S), dO, cfg, timed)
        return dO.to_numpy(PAIR, n), t
    finally:
        dR.free()
        dS.free()
        dO.free()


def partition(rel, radix_bits):
    """The global partition pass on the GPU (spj_partition_device): returns
    (partitioned tuples, partition starts of 2^radix_bits + 1 entries)."""
    n = len(rel)
    F = 1 << radix_bits
    d_in = DeviceArray.from_numpy(rel)
    d_out = DeviceArray(max(n, 1) * 8)
    d_pst = DeviceArray((F + 1) * 4)
    try:
        check(lib().spj_partition_device(d_in.ptr if n else None, n, radix_bits, d_out.ptr, d_pst.ptr))
        pst = d_pst.to_numpy(np.uint32, F + 1)
        return d_out.to_numpy(TUPLE, int(pst[-1])), pst
    finally:
        d_in.free()
        d_out.free()
        d_pst.free()


def join_host(R, S, cfg=None):
    """spj_join_host: host relations in, exact pairs out (staged copies)."""
    out = np.empty(max(len(S), 1), dtype=PAIR)
    n = U64()
    c = ctypes.byref(cfg) if cfg is not None else None
    check(lib().spj_join_host(R.ctypes.data if len(R) else None, len(R), S.ctypes.data if len(S) else None, len(S),
                              out.ctypes.data, len(S), ctypes.byref(n), c))
    return out[:n.value]


def join_batch(Rs, Ss, cfg=None, timed=False):
    """Join many instances in one launch sequence; returns ([pairs_i], times)."""
    n = len(Rs)
    dRs = [DeviceArray.from_numpy(r) for r in Rs]
    dSs = [DeviceArray.from_numpy(s) for s in Ss]
    dOs = [DeviceArray(max(len(s), 1) * 8) for s in Ss]
    arr = lambda vals: (ctypes.c_uint64 * n)(*vals)
    rp = (ctypes.c_void_p * n)(*[d.ptr for d in dRs])
    sp = (ctypes.c_void_p * n)(*[d.ptr for d in dSs])
    op = (ctypes.c_void_p * n)(*[d.ptr for d in dOs])
    nr = arr([len(r) for r in Rs])
    ns = arr([len(s) for s in Ss])
    cnt = (ctypes.c_uint64 * n)()
    t = PhaseTimes()
    try:
        check(lib().spj_join_batch_device(n, ctypes.cast(rp, P), ctypes.cast(nr, P), ctypes.cast(sp, P),
                                          ctypes.cast(ns, P), ctypes.cast(op, P), ctypes.cast(cnt, P),
                                          ctypes.byref(cfg) if cfg is not None else None,
                                          ctypes.byref(t) if timed else None))
        return [dOs[i].to_numpy(PAIR, cnt[i]) for i in range(n)], (t.as_dict() if timed else None)
    finally:
        for d in dRs + dSs + dOs:
            d.free()


class JoinInstance:
    """pidjoin::JoinInstance through the C ABI (query.cc life cycle)."""

    def __init__(self, ranks=16, rank_groups=1, ranks_per_group=16):
        self.h = lib().spj_create(ranks, rank_groups, ranks_per_group)
        if not self.h:
            raise SpjError(-1, lib().spj_last_error().decode())
        self._keep = {}

    @staticmethod
    def design(**kw):
        d = JoinDesign()
        lib().spj_join_design_default(ctypes.byref(d))
        for k, v in kw.items():
            setattr(d, k, v)
        return d

    def load_column(self, arr, name):
        arr = np.ascontiguousarray(arr, dtype=TUPLE)
        self._keep[name] = arr
        check(lib().spj_load_column(self.h, arr.ctypes.data if len(arr) else None, len(arr), name.encode()))

    def prepare(self, design=None):
        check(lib().spj_prepare(self.h, ctypes.byref(design or self.design())))

    def start(self, design=None):
        check(lib().spj_start(self.h, ctypes.byref(design or self.design())))

    def finish(self):
        p = P()
        n = U64()
        check(lib().spj_finish(self.h, ctypes.byref(p), ctypes.byref(n)))
        try:
            out = np.empty(n.value, dtype=PAIR)
            if n.value:
                ctypes.memmove(out.ctypes.data, p.value, n.value * 8)
            return out
        finally:
            lib().spj_free_result(p)

    RESULT_EXACT, RESULT_RANK_PADDED = 0, 1

    def set_result_layout(self, layout):
        check(lib().spj_set_result_layout(self.h, int(layout)))

    def finish_buffers(self, cap=64):
        """FinishJoin as buffers: a list of pair arrays, one per buffer (their
        lengths are the byte counts / 8, padding included)."""
        bufs = (P * cap)()
        nbytes = (U64 * cap)()
        n = ctypes.c_int()
        check(lib().spj_finish_buffers(self.h, ctypes.cast(bufs, P), ctypes.cast(nbytes, P), cap, ctypes.byref(n)))
        out = []
        for i in range(n.value):
            a = np.empty(nbytes[i] // 8, dtype=PAIR)
            if nbytes[i]:
                ctypes.memmove(a.ctypes.data, bufs[i], nbytes[i])
            lib().spj_free_result(bufs[i])
            out.append(a)
        return out

    def phase_times(self):
        t = PhaseTimes()
        check(lib().spj_last_phase_times(self.h, ctypes.byref(t)))
        return t.as_dict()

    def set_query_plan(self, plan_json):
        """The join plan StartJoin runs (None: the design's choice)."""
        check(lib().spj_set_query_plan(self.h, None if plan_json is None else plan_json.encode()))

    def timeline(self, start_ns=0, rank_group=0):
        """RecordLogs text of the last join; start_ns is a CLOCK_MONOTONIC instant
        (time.clock_gettime_ns(time.CLOCK_MONOTONIC))."""
        n = U64()
        sec, nsec = divmod(int(start_ns), 10**9)
        check(lib().spj_timeline(self.h, sec, nsec, rank_group, None, 0, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value + 1)
        check(lib().spj_timeline(self.h, sec, nsec, rank_group, buf, n.value + 1, ctypes.byref(n)))
        return buf.value.decode()

    def close(self):
        if self.h:
            lib().spj_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class Server:
    """Serving loop: resident tables, HBM-budget admission, continuous batching."""

    def __init__(self, hbm_budget_bytes, slots=2):
        self.h = lib().spj_server_create(int(hbm_budget_bytes), int(slots))
        if not self.h:
            raise SpjError(-1, lib().spj_last_error().decode())
        self._keep = {}

    def add_table(self, R):
        R = np.ascontiguousarray(R, dtype=TUPLE)
        tid = ctypes.c_int()
        check(lib().spj_server_add_table(self.h, R.ctypes.data if len(R) else None, len(R), ctypes.byref(tid)))
        return tid.value

    def submit(self, table, S):
        S = np.ascontiguousarray(S, dtype=TUPLE)
        rid = U64()
        check(lib().spj_server_submit(self.h, int(table), S.ctypes.data if len(S) else None, len(S),
                                      ctypes.byref(rid)))
        self._keep[rid.value] = S
        return rid.value

    def wait(self, rid):
        p = P()
        n = U64()
        lat = ctypes.c_double()
        try:
            check(lib().spj_server_wait(self.h, rid, ctypes.byref(p), ctypes.byref(n), ctypes.byref(lat)))
            out = np.empty(n.value, dtype=PAIR)
            if n.value:
                ctypes.memmove(out.ctypes.data, p.value, n.value * 8)
            return out, lat.value
        finally:
            if p.value:
                lib().spj_free_result(p)
            self._keep.pop(rid, None)

    def stats(self):
        b, r = U64(), U64()
        check(lib().spj_server_stats(self.h, ctypes.byref(b), ctypes.byref(r)))
        return {"batches": b.value, "requests": r.value}

    def close(self):
        if self.h:
            lib().spj_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def sort_pairs(p):
    """Sort pairs by (r_rid, s_rid) for multiset comparison."""
    w = (p["r_rid"].astype(np.uint64) << np.uint64(32)) | p["s_rid"].astype(np.uint64)
    w.sort()
    return w


def plan_check(plan_json):
    """Validate a join plan document; returns (number of nodes, replicates)."""
    n, r = ctypes.c_int(), ctypes.c_int()
    check(lib().spj_plan_check(plan_json.encode(), ctypes.byref(n), ctypes.byref(r)))
    return n.value, bool(r.value)
