#!/usr/bin/env python3
"""Benchmark: the radix-partitioned equi-join on one MI355X.

Default line (one step = one full join on HBM-resident inputs):
  |R| = 2^26 unique uint32 keys (permutation of 1..2^26, rid = row),
  |S| = 2^28 keys drawn uniformly (Zipf 0) from R's keys, hash join.
It prints one JSON line: joined tuples/s = (|R| + |S|) / step time, the HBM
roofline of the dominant kernel (HIP events around each phase of every timed
join), and a CPU baseline (the CPU restatement of the reference join on the
reference's CPU-runnable workload, 2^20 x 2^21, on the host cores).
`extra` holds the same join at Zipf 1.0 and 1.5 (and the Zipf 1.5 / Zipf 0
time ratio), the 2^26 x 2^27 Zipf 0.5 hash and sort-merge joins, and the
batched workload of 4096 mixed joins (--no-batched skips it).

--headline-only runs the default workload alone (for profiling);
--batched and --serving print their own lines instead.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "spid-join_amd"))

import spidjoin as sj  # noqa: E402

HBM_PEAK = 8.0e12   # MI355X HBM3E, B/s


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def run_workload(nR, nS, zipf, steps, warmup, cfg, seed, keep=None):
    """Generate, upload, warm up, then time `steps` joins.  Each timed join
    records HIP events between its phases on the join's own stream; the
    per-phase times are averaged over the timed joins.  Returns a dict."""
    t0 = time.time()
    if keep is not None and keep.get("key") == (nR, seed):
        R = keep["R"]
        dR = keep["dR"]
    else:
        if keep is not None and keep.get("dR") is not None:
            keep["dR"].free()
        R = sj.gen_r(nR, seed)
        dR = sj.DeviceArray.from_numpy(R)
        if keep is not None:
            keep.update(key=(nR, seed), R=R, dR=dR)
    skey = (nS, zipf, seed)
    if keep is not None and keep.get("skey") == skey:
        dS = keep["dS"]
    else:
        if keep is not None and keep.get("dS") is not None:
            keep["dS"].free()
            keep["dS"] = None
        S = sj.gen_s(nS, R, zipf, seed + 1 + int(zipf * 100))
        dS = sj.DeviceArray.from_numpy(S)
        del S
        if keep is not None:
            keep.update(skey=skey, dS=dS)
    dO = sj.DeviceArray(nS * 8)
    log(f"[bench] nR={nR} nS={nS} zipf={zipf}: data ready in {time.time() - t0:.1f}s")
    try:
        for _ in range(warmup):
            n, _t = sj.join_device(dR, nR, dS, nS, dO, cfg, timed=False)
            if n != nS:
                raise RuntimeError(f"join produced {n} pairs, expected {nS}")
        sj.check(sj.lib().spj_dev_sync())
        phases = []
        t1 = time.perf_counter()
        for _ in range(steps):
            n, t = sj.join_device(dR, nR, dS, nS, dO, cfg, timed=True)
            phases.append(t)
            if n != nS:
                raise RuntimeError(f"join produced {n} pairs, expected {nS}")
        sj.check(sj.lib().spj_dev_sync())
        wall = time.perf_counter() - t1
    finally:
        if keep is None:
            dS.free()
        dO.free()
    ms = wall / steps * 1e3
    avg = {k: float(np.mean([p[k] for p in phases])) for k in phases[0]}
    log(f"[bench]   {ms:.3f} ms/step; phases {json.dumps({k: round(v, 4) for k, v in avg.items()})}")
    return {"nR": nR, "nS": nS, "zipf": zipf, "ms_per_step": ms, "phases": avg, "pairs": int(n)}


def roofline(res, workload):
    """Dominant kernel and its achieved algorithmic GB/s.

    Algorithmic bytes per join (the phase's launches together: the scatter
    is one launch, or two -- R, then S -- when R has cells and heavy keys):
      k_hist    reads every tuple of a counted relation
                once (cell relations skip it):          8 (|R_counted| + |S_counted|)
      k_scatter reads every tuple, writes it to its
                partition (or a heavy pair to out):     16 (|R| + |S|)
      k_join    reads the partitioned R and S, writes
                one pair per partitioned S tuple:       8 (|R'| + |S'|) + 8 |S'|
    (|R'|, |S'|: tuples that went through partitions; M = |S| here.)"""
    p = res["phases"]
    nR, nS = res["nR"], res["nS"]
    rp, spn = p["r_partitioned"], p["s_partitioned"]
    counted = (0 if p.get("r_cells") else nR) + (0 if p.get("s_cells") else nS)
    kern = {
        "k_hist": (p["hist_ms"], 8 * counted),
        "k_scatter": (p["scatter_ms"], 16 * (nR + nS)),
        "k_join": (p["join_ms"], 8 * (rp + spn) + 8 * spn),
    }
    name = max(kern, key=lambda k: kern[k][0])
    ms, byts = kern[name]
    ach = byts / (ms * 1e-3)
    b_alg = 24 * (nR + nS) + 8 * res["pairs"]
    return name, {"bound": "hbm", "achieved": round(ach / 1e9, 1), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                  "frac": round(ach / HBM_PEAK, 4), "traffic": measured_traffic(name, workload)}, {
        k: {"ms": round(v[0], 4), "alg_bytes": v[1], "GB/s": round(v[1] / (v[0] * 1e-3) / 1e9, 1) if v[0] > 0 else 0.0}
        for k, v in kern.items()}, b_alg


TRAFFIC_FILE = "c3_traffic.json"   # counters of the default workload


def measured_traffic(kernel, workload):
    """HBM bytes of `kernel` per join (FETCH_SIZE + WRITE_SIZE over its
    dispatches of one join, scaled by the plain-copy calibration) from the
    committed rocprofv3 counter summary of this same workload
    (profiles/c3_traffic.json), or None when the summary is of another
    workload."""
    path = os.path.join(ROOT, "profiles", TRAFFIC_FILE)
    try:
        with open(path) as f:
            doc = json.load(f)
        if doc.get("workload") != workload:
            return None
        per = doc["per_join"]
        # template instances carry their argument (k_scatter<12>)
        key = next(k for k in per if k == kernel or k.startswith(kernel + "<"))
        t = per[key]
        return int(t["fetch_bytes"] + t["write_bytes"])
    except (OSError, KeyError, ValueError, StopIteration):
        return None


def cpu_baseline(seconds=10.0):
    """The oracle's CPU restatement of the reference join (global SFH partition
    into 1024 virtual DPUs, P_loc local partitions, 4096-slot linear probing)
    on the reference's CPU-runnable workload: |R| = 2^20 (permutation of
    1..2^20), |S| = 2^21 uniform over R's keys, on the host cores."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    nR, nS = 1 << 20, 1 << 21
    R = sj.gen_r(nR, 7)
    S = sj.gen_s(nS, R, 0.0, 8)
    oracle.spid_join_cpu(R, S, threads=threads)   # warm-up
    ts = []
    t_end = time.time() + seconds
    while time.time() < t_end or len(ts) < 5:
        t = time.perf_counter()
        out = oracle.spid_join_cpu(R, S, threads=threads)
        ts.append(time.perf_counter() - t)
        if len(ts) >= 200:
            break
    if len(out) != nS:
        raise RuntimeError("CPU baseline produced a wrong count")
    med = float(np.median(ts))
    cpu = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"value": round((nR + nS) / med, 1), "unit": "joined tuples/s", "cores": threads, "kind": "port",
            "sample": f"|R|=2^20, |S|=2^21, Zipf 0 (the reference's CPU-runnable workload), median of {len(ts)} "
                      f"runs ({med * 1e3:.2f} ms each); {cpu}"}


def copy_calibration(nbytes=1 << 30, reps=3):
    """Device-to-device copies of nbytes: each reads nbytes and writes nbytes."""
    import ctypes
    a = sj.DeviceArray(nbytes)
    b = sj.DeviceArray(nbytes)
    ms = ctypes.c_float()
    out = []
    for _ in range(reps):
        sj.check(sj.lib().spj_d2d(b.ptr, a.ptr, nbytes, ctypes.byref(ms)))
        out.append(ms.value)
    a.free()
    b.free()
    print(json.dumps({"copy_bytes": nbytes, "ms": out, "GB/s_read_plus_write": [round(2 * nbytes / (m * 1e-3) / 1e9, 1)
                                                                              for m in out]}), flush=True)


def batched_run(n, max_log2r, steps, warmup):
    """Batched mixed joins: n instances with |R| log-uniform in
    [2^10, 2^max_log2r], R:S from {1,2,4,8}, Zipf U[0, 1.25].  One batched
    launch sequence (persistent join kernel over one work queue) against one
    join per call, with identical per-instance results (count and
    pair-multiset checksum).  Returns a dict."""
    import ctypes
    rng = np.random.default_rng(4096)
    nRs = np.floor(2.0 ** rng.uniform(10, max_log2r, n)).astype(np.int64)
    ratios = rng.choice([1, 2, 4, 8], n)
    zipfs = rng.uniform(0, 1.25, n)
    nSs = nRs * ratios
    totR, totS = int(nRs.sum()), int(nSs.sum())
    t0 = time.time()
    dR = sj.DeviceArray(totR * 8)
    dS = sj.DeviceArray(totS * 8)
    dO = sj.DeviceArray(totS * 8)
    offR = np.concatenate([[0], np.cumsum(nRs)[:-1]])
    offS = np.concatenate([[0], np.cumsum(nSs)[:-1]])
    for i in range(n):
        R = sj.gen_r(int(nRs[i]), 10_000 + i)
        S = sj.gen_s(int(nSs[i]), R, float(zipfs[i]), 20_000 + i)
        sj.check(sj.lib().spj_h2d(dR.ptr + int(offR[i]) * 8, R.ctypes.data, R.nbytes))
        sj.check(sj.lib().spj_h2d(dS.ptr + int(offS[i]) * 8, S.ctypes.data, S.nbytes))
    log(f"[bench] batched: {n} instances, |R| total {totR}, |S| total {totS}; data ready in {time.time() - t0:.1f}s")
    P = ctypes.c_void_p
    rp = (P * n)(*[dR.ptr + int(o) * 8 for o in offR])
    sp = (P * n)(*[dS.ptr + int(o) * 8 for o in offS])
    op = (P * n)(*[dO.ptr + int(o) * 8 for o in offS])
    nr = (ctypes.c_uint64 * n)(*[int(x) for x in nRs])
    ns = (ctypes.c_uint64 * n)(*[int(x) for x in nSs])
    cnt = (ctypes.c_uint64 * n)()
    L = sj.lib()

    def run_batch():
        sj.check(L.spj_join_batch_device(n, ctypes.cast(rp, P), ctypes.cast(nr, P), ctypes.cast(sp, P),
                                         ctypes.cast(ns, P), ctypes.cast(op, P), ctypes.cast(cnt, P), None, None))

    def checksums():
        return [sj.checksum_device(int(op[i]), int(cnt[i])) for i in range(n)]

    try:
        for _ in range(warmup):
            run_batch()
        sj.check(L.spj_dev_sync())
        blat = []
        t1 = time.perf_counter()
        for _ in range(steps):   # each call returns when its batch is done
            ta = time.perf_counter()
            run_batch()
            blat.append(time.perf_counter() - ta)
        sj.check(L.spj_dev_sync())
        t_batch = (time.perf_counter() - t1) / steps
        counts_b = [int(cnt[i]) for i in range(n)]
        sums_b = checksums()
        # one join per call
        lat = []
        one = sj.lib().spj_join_device
        c1 = ctypes.c_uint64()
        for i in range(min(n, 16)):   # warm-up
            sj.check(one(rp[i], nr[i], sp[i], ns[i], op[i], ns[i], ctypes.byref(c1), None, None))
        sj.check(L.spj_dev_sync())
        counts_1 = []
        t2 = time.perf_counter()
        for i in range(n):
            ta = time.perf_counter()
            sj.check(one(rp[i], nr[i], sp[i], ns[i], op[i], ns[i], ctypes.byref(c1), None, None))
            lat.append(time.perf_counter() - ta)
            counts_1.append(int(c1.value))
        t_single = time.perf_counter() - t2
        for i in range(n):
            cnt[i] = counts_1[i]
        sums_1 = checksums()
    finally:
        dR.free()
        dS.free()
        dO.free()
    identical = counts_b == counts_1 and sums_b == sums_1 and counts_b == [int(x) for x in nSs]
    lat_ms = np.array(lat) * 1e3
    blat_ms = np.array(blat) * 1e3
    tuples = totR + totS
    return {
        "instances": n, "max_log2r": max_log2r, "sum_R": totR, "sum_S": totS,
        "batch_ms": round(t_batch * 1e3, 3), "batch_tuples_per_s": round(tuples / t_batch, 1),
        "batch_latency_ms_p50": round(float(np.percentile(blat_ms, 50)), 3),
        "batch_latency_ms_p99": round(float(np.percentile(blat_ms, 99)), 3),
        "one_join_per_call": {"tuples_per_s": round(tuples / t_single, 1), "total_ms": round(t_single * 1e3, 2),
                              "latency_ms_p50": round(float(np.percentile(lat_ms, 50)), 4),
                              "latency_ms_p99": round(float(np.percentile(lat_ms, 99)), 4)},
        "speedup_batched_over_one_per_call": round(t_single / t_batch, 2),
        "identical_per_instance_results": bool(identical),
    }


def batched(args):
    r = batched_run(args.instances, args.max_log2r, args.steps, args.warmup)
    out = {
        "metric": "batched joined tuples/s (sum |R|+|S| over instances)",
        "value": r["batch_tuples_per_s"], "unit": "tuples/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": r["batch_ms"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "uint32",
        "data": "synthetic (seeded: |R| log-uniform, R:S in {1,2,4,8}, Zipf U[0,1.25])",
        "config": {"workload": f"batched mixed joins, {r['instances']} instances, |R| in [2^10, 2^{args.max_log2r}]",
                   "instances": r["instances"], "sum_R": r["sum_R"], "sum_S": r["sum_S"]},
        "batched": r,
    }
    print(json.dumps(out), flush=True)
    if not r["identical_per_instance_results"]:
        sys.exit(1)


def serving(args):
    """Serving loop: 16 resident tables (|R| log-uniform [2^16, 2^24]), probe
    requests with |S| log-uniform [2^12, 2^22], Zipf U[0, 1.5], random table,
    arriving as a Poisson process at each swept rate.  Reports p50 / p99
    latency (submission to result on the host) and probe tuples/s."""
    import threading
    rng = np.random.default_rng(16)
    t0 = time.time()
    ntab = 16
    sizes = np.floor(2.0 ** rng.uniform(16, 24, ntab)).astype(np.int64)
    tables = [sj.gen_r(int(n), 30_000 + i) for i, n in enumerate(sizes)]
    pool = []
    for i in range(args.requests):
        t = int(rng.integers(0, ntab))
        n = int(2 ** rng.uniform(12, 22))
        pool.append((t, sj.gen_s(n, tables[t], float(rng.uniform(0, 1.5)), 40_000 + i)))
    sv = sj.Server(args.budget_gb << 30, slots=args.slots)
    tids = [sv.add_table(R) for R in tables]
    log(f"[bench] serving: {ntab} tables (sum |R| {int(sizes.sum())}), {len(pool)} requests "
        f"(sum |S| {sum(len(s) for _, s in pool)}); ready in {time.time() - t0:.1f}s")
    # warm-up: the largest requests, submitted together so that every slot
    # runs some and sizes its buffers before the timed sweep
    big = sorted(range(len(pool)), key=lambda k: -len(pool[k][1]))[:2 * args.slots]
    for rid in [sv.submit(tids[pool[k][0]], pool[k][1]) for k in big]:
        sv.wait(rid)
    results = []
    for rate in args.rates:
        gaps = rng.exponential(1.0 / rate, len(pool))
        ids = [None] * len(pool)

        def client():
            nxt = time.perf_counter()
            for k, (t, S) in enumerate(pool):
                nxt += gaps[k]
                d = nxt - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                ids[k] = sv.submit(tids[t], S)

        th = threading.Thread(target=client)
        ts = time.perf_counter()
        th.start()
        lats, probe = [], 0
        for k in range(len(pool)):
            while ids[k] is None:
                time.sleep(1e-4)
            out, lat = sv.wait(ids[k])
            if len(out) != len(pool[k][1]):
                raise RuntimeError("serving result has a wrong count")
            lats.append(lat)
            probe += len(pool[k][1])
        th.join()
        wall = time.perf_counter() - ts
        st = sv.stats()
        lats = np.array(lats)
        r = {"rate_req_per_s": rate, "p50_ms": round(float(np.percentile(lats, 50)), 3),
             "p99_ms": round(float(np.percentile(lats, 99)), 3), "probe_tuples_per_s": round(probe / wall, 1),
             "wall_s": round(wall, 3)}
        log(f"[bench]   rate {rate}/s: {r}")
        results.append(r)
    st = sv.stats()
    sv.close()
    best = max(results, key=lambda r: r["probe_tuples_per_s"])
    out = {
        "metric": "serving probe tuples/s (p50/p99 latency per rate)", "value": best["probe_tuples_per_s"],
        "unit": "tuples/s", "n_gpus": 1, "steps": len(pool), "warmup": 8, "ms_per_step": None,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "uint32",
        "data": "synthetic (seeded tables and Poisson request stream)",
        "config": {"workload": "serving loop: 16 resident tables, Poisson probe requests",
                   "tables": ntab, "requests": len(pool), "budget_gb": args.budget_gb, "slots": args.slots},
        "sweep": results, "batches": st["batches"], "batched_requests": st["requests"],
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    # the GPU's clocks take ~8 joins of sustained load to settle (per-call
    # times fall ~7% over them), so the default warm-up covers that
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--headline-only", action="store_true",
                    help="only the default workload (2^26 x 2^28, Zipf 0), e.g. under a profiler")
    ap.add_argument("--no-batched", action="store_true", help="leave the batched workload out of the default line")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 1 GiB device-to-device copies (the plain copy kernel that scales the "
                         "FETCH_SIZE / WRITE_SIZE counters) and print their bandwidth")
    ap.add_argument("--batched", action="store_true", help="batched mixed-joins workload instead of the default")
    ap.add_argument("--instances", type=int, default=4096)
    ap.add_argument("--max-log2r", type=float, default=20.0)
    ap.add_argument("--serving", action="store_true", help="serving-loop workload instead of the default")
    ap.add_argument("--requests", type=int, default=256)
    ap.add_argument("--rates", type=float, nargs="+", default=[50, 200, 800, 3200])
    ap.add_argument("--slots", type=int, default=4)
    ap.add_argument("--budget-gb", type=int, default=16)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)

    sj.init(0)
    info = sj.device_info()
    log(f"[bench] device {info}")
    if args.copy_calibration:
        copy_calibration()
        return
    if args.batched:
        batched(args)
        return
    if args.serving:
        serving(args)
        return
    nR, nS, zipf = 1 << 26, 1 << 28, 0.0
    workload = "hash equi-join |R|=2^26 x |S|=2^28, Zipf 0, inputs resident in HBM"
    keep = {}
    try:
        res = run_workload(nR, nS, zipf, args.steps, args.warmup, None, 1000, keep)
        kname, roof, kernels, b_alg = roofline(res, workload)
        ms = res["ms_per_step"]
        value = (nR + nS) / (ms * 1e-3)
        b_min = 8 * (nR + nS) + 8 * res["pairs"]   # compulsory: read R and S once, write the pairs
        measured = {}
        for k in ("k_scatter", "k_join"):   # rocprof HBM-side bytes per join over the phase's event time
            tb = measured_traffic(k, workload)
            if tb is not None and kernels[k]["ms"] > 0:
                measured[k] = {"bytes": tb, "GB/s": round(tb / (kernels[k]["ms"] * 1e-3) / 1e9, 1)}
        extra = {"kernels": kernels, "dominant_kernel": kname,
                 "join_B_alg_bytes": b_alg,
                 "join_roofline_frac": round(b_alg / (ms * 1e-3) / HBM_PEAK, 4),
                 "join_B_min_bytes": b_min,
                 "join_B_min_frac": round(b_min / (ms * 1e-3) / HBM_PEAK, 4),
                 "measured_traffic": measured,
                 "phases_ms": {k: round(v, 4) for k, v in res["phases"].items()}}
        if not args.headline_only:
            sk = {"0.0": {"ms_per_step": round(ms, 4), "roofline_frac": extra["join_roofline_frac"]}}
            for z in (1.0, 1.5):
                r = run_workload(nR, nS, z, args.steps, args.warmup, None, 1000, keep)
                bal = 24 * (nR + nS) + 8 * r["pairs"]
                _, rz, kz, _ = roofline(r, None)
                sk[str(z)] = {"ms_per_step": round(r["ms_per_step"], 4),
                              "tuples_per_s": round((nR + nS) / (r["ms_per_step"] * 1e-3), 1),
                              "roofline_frac": round(bal / (r["ms_per_step"] * 1e-3) / HBM_PEAK, 4),
                              "heavy_keys": r["phases"]["heavy_keys"],
                              "s_partitioned": r["phases"]["s_partitioned"],
                              "kernels": kz}
            sk["zipf1.5_over_zipf0"] = round(sk["1.5"]["ms_per_step"] / ms, 3)
            sk["zipf1.0_over_zipf0"] = round(sk["1.0"]["ms_per_step"] / ms, 3)
            extra["skew_2^26x2^28"] = sk
            c2 = {}
            for name, cfg in (("hash", None), ("sort_merge", sj.config(algorithm=sj.ALG_SMJ))):
                r = run_workload(nR, 1 << 27, 0.5, max(5, args.steps // 2), args.warmup, cfg, 1000, keep)
                _, _, kz, bal = roofline(r, None)
                c2[name] = {"ms_per_step": round(r["ms_per_step"], 4),
                            "tuples_per_s": round((nR + (1 << 27)) / (r["ms_per_step"] * 1e-3), 1),
                            "roofline_frac": round(bal / (r["ms_per_step"] * 1e-3) / HBM_PEAK, 4),
                            "kernels": kz}
            extra["c2_2^26x2^27_zipf0.5"] = c2
    finally:
        for k in ("dR", "dS"):
            if keep.get(k) is not None:
                keep[k].free()
    if not (args.headline_only or args.no_batched):
        extra["batched_4096_mixed"] = batched_run(4096, 20.0, max(5, args.steps // 2), 2)
    cpu = None if args.no_cpu else cpu_baseline()
    out = {
        "metric": "joined tuples/s (|R|+|S|)",
        "value": round(value, 1),
        "unit": "tuples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "uint32",
        "data": "synthetic (seeded generator: R = permutation of 1..2^26, S uniform over R's keys)",
        "config": {"workload": workload, "nR": nR, "nS": nS, "zipf": zipf, "algorithm": "radix-partitioned hash join"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
