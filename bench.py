"""Triangle-count benchmarks on one MI355X.

Default workload (static): R-MAT scale 24, edge factor 32,
(a,b,c,d)=(0.57,0.19,0.19,0.05), seed 4, symmetrised / deduplicated /
loop-free, generated on the device with the project's counter-based
generator (synthetic data; nothing is read from disk).  A "step" is one full
static count of that graph -- orientation, job planning and the intersection
kernels -- producing the uint32 per-vertex counts (kept in HBM) and the uint64
global count.  Three legs run on the same graph, each K timed steps:

  fixed    built-in schedule, fixed degree-ratio strategy heuristic;
  tuned    autotuned bin thresholds / job grains, fixed heuristic;
  learned  autotuned schedule with the learned strategy table (MLP trained on
           the device from in-situ count times of R-MAT graphs of other
           seeds, custingeralg_amd/selector_model.json).

value = nnz / step time of the learned leg; all three must give the same
global count.  The roofline object describes the dominant count kernel by
event time in that leg: the small or mid bin's kernel, or the largest bin's
hash-task kernel on bitmap or hash-table owner tiles.

Other workloads (``--workload``): ``stream`` (R-MAT scale 22 ef 16 seed 3 base,
20 update batches of 10^3..10^6 pairs, updates/s, running total checked against
a recount after every batch), ``batched`` (4096 R-MAT graphs of 1k-64k
vertices counted in one call, edges/s), ``copy`` (a plain device copy of
known size, the calibration kernel for the HBM traffic counters) and ``cpu``
(the reference CPU path counting ER-16 and R-MAT 20 in full, on one and on
all host cores; no GPU).

The CPU baseline times the oracle's restatement of the reference CPU path
(hostCountTriangles' per-entry merge, tests/cct.cpp:32-89) on a uniform
sample of the same graph's entries, on this host's cores.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload static|stream|batched|copy|cpu]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "TC edges/sec + HBM GB/s at RMAT-24; learned vs fixed schedule, exact counts"
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_count_long_rmat24_traffic.md")
GOLDEN_FILE = os.path.join(ROOT, "tests", "golden", "big_fixtures.json")


def golden(section: str):
    """Expected results of a full-size workload, computed on the CPU by the
    oracle (tests/golden/make_big_fixtures.py), or None."""
    try:
        with open(GOLDEN_FILE) as f:
            return json.load(f).get(section)
    except OSError:
        return None


def sha256(a) -> str:
    import hashlib
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(off, adj, budget_s: float, seed: int = 7) -> dict:
    """hostCountTriangles' per-entry intersection on uniformly drawn entries."""
    import oracle as O

    threads = O.nthreads()
    rng = np.random.default_rng(seed)
    nnz = len(adj)
    n = 1 << 10
    spent, done, t_all = 0.0, 0, 0.0
    while spent < budget_s and n <= nnz:
        e = np.sort(rng.integers(0, nnz, size=n, dtype=np.int64))
        t0 = time.perf_counter()
        O.host_count_entries(off, adj, e, threads)
        dt = time.perf_counter() - t0
        spent += dt
        done += n
        t_all += dt
        if dt < budget_s / 4:
            n *= 2
    return {"value": done / t_all, "unit": "edges/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(),
            "sample": f"sampled, not a full count: {done} directed entries drawn uniformly from "
                      f"the same graph ({done / nnz:.2%} of its {nnz}); each runs the per-entry "
                      f"merge of hostCountTriangles (tests/cct.cpp:32-89); {t_all:.1f} s on "
                      f"{threads} threads. Full counts on ER-16 and R-MAT 20: --workload cpu"}


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def timed_host_count(off, adj, threads: int, label: str):
    """The oracle's hostCountTriangles (tests/cct.cpp:71-89) over the whole
    graph, in 32 row slices of equal entry count so that a long single-thread
    run prints its progress; returns (int64 total, seconds)."""
    import oracle as O

    nv = len(off) - 1
    cuts = np.searchsorted(off, np.linspace(0, off[-1], 33).astype(np.int64), side="left")
    cuts[0], cuts[-1] = 0, nv
    total, spent, t_last = 0, 0.0, time.perf_counter()
    for k in range(32):
        v0, v1 = int(cuts[k]), int(cuts[k + 1])
        if v1 <= v0:
            continue
        t0 = time.perf_counter()
        total += O.host_count_rows(off, adj, v0, v1, threads)
        spent += time.perf_counter() - t0
        if time.perf_counter() - t_last > 20:
            log(f"[bench] cpu {label}: rows < {v1} of {nv}, {spent:.0f} s")
            t_last = time.perf_counter()
    return total, spent


def measured_traffic():
    """HBM bytes per launch of the dominant count kernel, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (copy-kernel calibrated)."""
    try:
        with open(TRAFFIC_FILE) as f:
            m = re.search(r"^traffic_bytes_per_launch:\s*([0-9.eE+]+)", f.read(), re.M)
        return float(m.group(1)) if m else None
    except OSError:
        return None


def timed_steps(T, step, steps, warmup):
    """Run warmup + steps calls of step(); returns (ms per step, [stats])."""
    for _ in range(warmup):
        step()
    T.synchronize()
    stats = []
    t0 = time.perf_counter()
    for k in range(steps):
        stats.append(step())
    T.synchronize()
    return (time.perf_counter() - t0) / steps * 1e3, stats


def bench_static(T, args) -> dict:
    from custingeralg_amd import learned as Lm

    t0 = time.time()
    g = T.rmat_graph_device(args.scale, args.ef, args.seed)
    nnz = g.info()["nnz"]
    log(f"[bench] R-MAT scale {args.scale} ef {args.ef} seed {args.seed}: nnz={nnz} "
        f"max_degree={g.info()['max_degree']} built in {time.time() - t0:.1f}s")
    out = T.DeviceArray(g.nv, np.uint32)
    ref = {}

    def leg(name, policy):
        sched = T.default_schedule(policy=policy)

        def step():
            _, glob, st = T.static_count(g, sched, per_vertex=True, out=out)
            if ref.setdefault("glob", glob) != glob:
                raise RuntimeError(f"{name}: global count {glob} != {ref['glob']}")
            return st

        ms, stats = timed_steps(T, step, args.steps, args.warmup)
        best = min(stats, key=lambda s: s["ms_total"])
        log(f"[bench] {name}: {ms:.1f} ms/step (events best {best['ms_total']:.1f} ms, "
            f"bins {[round(x, 1) for x in best['ms_bin']]})")
        return {"ms_per_step": ms, "stats": best}

    Lm.set_default_schedule(None)
    legs = {"fixed": leg("fixed", T.POLICY_FIXED)}

    # learned selector and tuned schedule: saved by custingeralg_amd.train_selector
    # (trained on the GPU from in-situ count times), else trained here
    t0 = time.time()
    if os.path.exists(args.selector):
        sel = Lm.load_selector(args.selector)
        table, tuned, meta = sel["table"], sel["schedule"], sel["meta"]
        how = (f"{meta['timed_tasks']} timed tasks" if "timed_tasks" in meta
               else meta.get("costs", "?"))
        source = (f"{os.path.relpath(args.selector, ROOT)}: trained on R-MAT "
                  f"{meta.get('train_scales')} ({how}), tuned on {meta.get('tune_graph')}")
    else:
        train = [T.rmat_graph_device(s, 16, 100 + s) for s in (18, 20, 22)]
        _, table, losses, (feat, _) = Lm.train_selector(train, samples_per_graph=4096, steps=400)
        tuned, _ = Lm.autotune(train[-1])
        for tg in train:
            tg.close()
        source = f"trained in this run on R-MAT 18/20/22 ({len(feat)} timed tasks), tuned on R-MAT 22"
    picks = np.bincount(table.ravel(), minlength=3).tolist()
    log(f"[bench] selector {source}; table picks merge/bsearch/hash {picks}; schedule "
        f"small={tuned.small_max} mid={tuned.mid_max} flush={tuned.flush_items} "
        f"job={tuned.job_items} ({time.time() - t0:.1f}s)")
    legs["tuned"] = leg("tuned", T.POLICY_FIXED)
    legs["learned"] = leg("learned", T.POLICY_LEARNED)
    Lm.unload()
    Lm.set_default_schedule(None)

    # outside the timed region: the learned leg's output (the last one written
    # to out) against the oracle's golden for this graph
    check = "no golden for this graph"
    gd = golden("rmat24")
    if gd and (args.scale, args.ef, args.seed) == (gd["scale"], gd["edge_factor"], gd["seed"]):
        pv = out.to_host()
        if ref["glob"] != gd["global"] or sha256(pv) != gd["pv_sha256"]:
            raise RuntimeError(f"count differs from the oracle's golden: global {ref['glob']} "
                               f"!= {gd['global']} or per-vertex SHA-256 differs")
        check = (f"global and per-vertex SHA-256 equal the oracle's golden "
                 f"(T={gd['T']}, {gd['pv_wrapped']} per-vertex values wrapped past 2^32)")
        del pv
    log(f"[bench] parity: {check}")

    main = legs["learned"]
    st = main["stats"]
    # the dominant count kernel by event time: the small and mid bins' kernels,
    # or the largest bin's hash-task kernel on bitmap / hash-table tiles (the
    # rest of the largest bin is a short pass over its other tasks)
    kernels = [("k_count<64,1,true> (bin 0)", st["ms_bin"][0], st["stream_bytes_bin"][0]),
               ("k_count<512,2,true> (bin 1)", st["ms_bin"][1], st["stream_bytes_bin"][1]),
               ("k_count_long<4096,8,true> (bin 2, bitmap tiles)", st["ms_long"][0],
                st["stream_bytes_long"][0]),
               ("k_count_long<4096,8,false> (bin 2, hash-table tiles)", st["ms_long"][1],
                st["stream_bytes_long"][1])]
    kname, ms_b, bytes_b = max(kernels, key=lambda k: k[1])
    # bytes the kernel must read: each task's list suffix, task entry and
    # vertex record, each job's owner tile (the owner list stays in LDS)
    achieved = bytes_b / (ms_b * 1e-3) / 1e9 if ms_b > 0 else 0.0

    cpu = None
    if not args.no_cpu:
        off, adj = g.to_csr()
        cpu = cpu_baseline(off, adj, args.cpu_budget)
        del off, adj
    out.free()
    g.close()
    glob = ref["glob"]
    return {
        "metric": METRIC, "value": nnz / (main["ms_per_step"] * 1e-3), "unit": "edges/s",
        "ms_per_step": main["ms_per_step"], "higher_is_better": True, "dtype": "int32",
        "config": {"workload": f"static TC, R-MAT scale {args.scale} ef {args.ef} "
                               f"(0.57,0.19,0.19,0.05) seed {args.seed}",
                   "nnz": nnz, "triangles": glob // 6, "global_count": glob,
                   "schedule": "autotuned + learned strategy table", "selector": source,
                   "parity": check,
                   "legs_ms_per_step": {k: v["ms_per_step"] for k, v in legs.items()},
                   "learned_vs_fixed": legs["fixed"]["ms_per_step"] / main["ms_per_step"],
                   "tuned": {"small_max": tuned.small_max, "mid_max": tuned.mid_max,
                             "flush_items": tuned.flush_items, "job_items": tuned.job_items},
                   "table_picks": picks},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(),
                     "kernel": kname, "kernel_ms": ms_b, "alg_bytes": bytes_b},
        "cpu_baseline": cpu,
        "bins": {"ms": st["ms_bin"], "tasks": st["tasks_bin"], "jobs": st["jobs_bin"],
                 "items": st["items_bin"], "stream_bytes": st["stream_bytes_bin"],
                 "merge_equiv_bytes": st["bytes_alg_bin"], "strategy_tasks": st["strategy_tasks"],
                 "long_ms": st["ms_long"], "long_tasks": st["tasks_long"],
                 "long_items": st["items_long"], "long_stream_bytes": st["stream_bytes_long"],
                 "ms_prepare": st["ms_prepare"], "ms_count": st["ms_count"],
                 "build": build_roofline(nnz, st["ms_prepare"])},
    }


def build_roofline(nnz: int, ms: float) -> dict:
    """The orientation build against the bytes any build must move: per store
    entry its id and its neighbour's rank (4 + 4 B), per plus entry (half of
    them) its N+ rank and its task record (4 + 8 B), 14 B per store entry."""
    alg = 14 * nnz
    gbs = alg / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    return {"ms": ms, "alg_bytes_per_entry": 14, "alg_bytes": alg, "achieved": gbs,
            "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS}


def _device_pairs(T):
    """The device pair generators in the form the workload recipe takes."""
    def er(n, m, seed):
        du, dv = T.DeviceArray(m, np.int32), T.DeviceArray(m, np.int32)
        T.check(T.lib().tc_gen_er_pairs(n, m, seed, du.ptr, dv.ptr, None))
        return du.to_host(), dv.to_host()

    def rmat(scale, m, seed):
        du, dv = T.DeviceArray(m, np.int32), T.DeviceArray(m, np.int32)
        T.check(T.lib().tc_gen_rmat_pairs(scale, m, 0.57, 0.19, 0.19, seed, du.ptr, dv.ptr, None))
        return du.to_host(), dv.to_host()
    return er, rmat


def bench_stream(T, args) -> dict:
    from custingeralg_amd import workloads as W
    scale, ef, seed = (W.STREAM_BASE[k] for k in ("scale", "edge_factor", "seed"))
    g = T.rmat_graph_device(scale, ef, seed)
    nv = g.nv
    log(f"[bench] stream base R-MAT {scale} ef {ef}: nnz={g.info()['nnz']}")
    _, total, _ = T.static_count(g, per_vertex=False)
    delta = T.DeviceArray(nv, np.uint32).zero()
    rng = np.random.default_rng(W.STREAM_RNG_SEED)
    sizes = W.STREAM_SIZES
    er, rmat = _device_pairs(T)
    gd = golden("stream")
    # batches padded to size classes and replayed as hipGraphs (the first
    # batch of a class runs plainly, the second records)
    sched = T.default_schedule(capture=0 if args.no_capture else 1)
    per_size = {}
    t_all, n_all = 0.0, 0
    for i, B in enumerate(sizes):
        off, adj = g.to_csr()
        u, v, dl = W.stream_batch(er, rmat, off, adj, nv, B, scale, W.stream_batch_seed(i), rng)
        del off, adj
        du, dv, dd = (T.DeviceArray.from_host(x) for x in (u, v, dl))
        st = T.BatchStats()
        dg = C.c_int64(0)
        T.synchronize()
        t0 = time.perf_counter()
        T.check(T.lib().tc_batch_apply(g.h, du.ptr, dv.ptr, dd.ptr, len(u), delta.ptr,
                                       C.byref(dg), C.byref(sched), C.byref(st)))
        T.synchronize()
        dt = time.perf_counter() - t0
        total += dg.value
        _, recount, _ = T.static_count(g, per_vertex=False)
        if recount != total:
            raise RuntimeError(f"batch {i}: running total {total} != recount {recount}")
        if gd and total != 6 * gd["batches"][i]["T"]:
            raise RuntimeError(f"batch {i}: running total {total} != the oracle's "
                               f"{6 * gd['batches'][i]['T']}")
        s = per_size.setdefault(B, {"pairs": 0, "s": 0.0, "ms_update": 0.0, "ms_delta": 0.0,
                                    "wall_ms": [], "replayed": 0})
        s["pairs"] += len(u)
        s["s"] += dt
        s["wall_ms"].append(dt * 1e3)
        s["replayed"] += st.replayed
        s["ms_update"] += st.ms_update
        s["ms_delta"] += st.ms_delta
        t_all += dt
        n_all += len(u)
        log(f"[bench] batch {i}: B={len(u)} ({int(dl.sum())} deletions) {dt * 1e3:.2f} ms "
            f"(update {st.ms_update:.2f} delta {st.ms_delta:.2f}{' replayed' if st.replayed else ''}) "
            f"delta={dg.value} ok")
    g.close()
    rates = {str(B): {"updates_per_s": s["pairs"] / s["s"],
                      "ms_update": s["ms_update"] / 5, "ms_delta": s["ms_delta"] / 5,
                      "wall_ms_per_batch": [round(x, 3) for x in s["wall_ms"]],
                      "replayed_batches": s["replayed"]}
             for B, s in per_size.items()}
    return {"metric": "streaming TC updates/sec (exact delta, checked by recount)",
            "value": n_all / t_all, "unit": "updates/s", "ms_per_step": t_all / len(sizes) * 1e3,
            "higher_is_better": True, "dtype": "int32",
            "config": {"workload": f"streaming TC, R-MAT scale {scale} ef {ef} seed {seed} base, "
                                   "20 batches (5 each of 10^3, 10^4, 10^5 and 10^6 pairs), "
                                   "80% inserts / 20% deletes",
                       "per_batch_size": rates, "final_global_count": total,
                       "recorded_batches": not args.no_capture,
                       "parity": ("running total equals a GPU recount and the oracle's T "
                                  "after every batch" if gd else
                                  "running total equals a GPU recount after every batch")},
            "roofline": None, "cpu_baseline": None}


def bench_batched(T, args) -> dict:
    n, ef, seeds = T.batched_spec()
    t0 = time.time()
    dvoff, doff, dadj = T.batched_graph_device(n, ef, seeds)
    nnz = dadj.n
    log(f"[bench] batched: {len(n)} graphs, {int(n.sum())} vertices, nnz={nnz} "
        f"built in {time.time() - t0:.1f}s")
    pg = T.DeviceArray(len(n), np.uint64)
    bg = T.BatchedGraph(dvoff, doff, dadj)  # validated once, like the static graph
    ref = {}

    def step():
        st = bg.count(pg)
        if ref.setdefault("tri", st["triangles"]) != st["triangles"]:
            raise RuntimeError("batched count changed between steps")
        return st

    ms, stats = timed_steps(T, step, args.steps, args.warmup)
    per_graph = pg.to_host()
    bg.close()
    gd = golden("batched")
    if gd and not np.array_equal(per_graph, np.array(gd["T"], np.uint64) * np.uint64(6)):
        raise RuntimeError("per-graph counts differ from the oracle's golden")
    st = min(stats, key=lambda s: s["ms_total"])
    return {"metric": "batched TC edges/sec over 4096 graphs", "value": nnz / (ms * 1e-3),
            "unit": "edges/s", "ms_per_step": ms, "higher_is_better": True, "dtype": "int32",
            "config": {"workload": "batched static TC, 4096 R-MAT graphs, n log-uniform "
                                   "[1k,64k], ef U{4..32}, seeds 1000+i",
                       "nnz": nnz, "sum_per_graph": int(per_graph.sum()),
                       "parity": ("every per-graph count equals the oracle's golden" if gd
                                  else "no golden"),
                       "bins_ms": st["ms_bin"], "ms_prepare": st["ms_prepare"],
                       "bins": {"tasks": st["tasks_bin"], "jobs": st["jobs_bin"],
                                "items": st["items_bin"], "stream_bytes": st["stream_bytes_bin"],
                                "strategy_tasks": st["strategy_tasks"]}},
            "roofline": None, "cpu_baseline": None}


def bench_copy(T, args) -> dict:
    """A plain device-to-device copy of known size (1 GiB, the runtime's copy
    kernel); its FETCH_SIZE / WRITE_SIZE readings calibrate the traffic
    counters of the count kernel."""
    nbytes = 1 << 30
    src = T.DeviceArray(nbytes // 4, np.int32)
    dst = T.DeviceArray(nbytes // 4, np.int32)
    T.check(T.lib().tc_memset(src.ptr, 1, nbytes))

    def step():
        T.check(T.lib().tc_memcpy(dst.ptr, src.ptr, nbytes, 2))

    ms, _ = timed_steps(T, step, args.steps, args.warmup)
    src.free()
    dst.free()
    return {"metric": "device copy GB/s (traffic calibration)", "value": 2 * nbytes / ms / 1e6,
            "unit": "GB/s", "ms_per_step": ms, "higher_is_better": True, "dtype": "int32",
            "config": {"workload": f"device-to-device copy of {nbytes} bytes"}, "roofline": None,
            "cpu_baseline": None}


def bench_cpu(T, args) -> dict:
    """The reference CPU path (hostCountTriangles, tests/cct.cpp:71-89, as the
    oracle restates it) run in full on ER-16 and R-MAT 20, single-threaded and
    on every host core; each total is checked against 6 x the oracle's
    degree-oriented count."""
    import oracle as O

    nth = O.nthreads()
    runs = {}
    for name, make in (("ER-16 (n=2^16, avg degree 16, seed 1)", lambda: O.er_graph(1 << 16, 16, 1)),
                       ("R-MAT 20 (ef 16, seed 2)", lambda: O.rmat_graph(20, 16, 2))):
        off, adj = make()
        want = 6 * O.forward_count(off, adj)[0]
        for th in (nth, 1):
            total, dt = timed_host_count(off, adj, th, f"{name} {th}T")
            if total != want:
                raise RuntimeError(f"{name}: CPU total {total} != 6T = {want}")
            runs[f"{name}, {th} threads"] = {"edges_per_s": len(adj) / dt, "seconds": dt,
                                             "nnz": len(adj), "total_6T": total}
            log(f"[bench] cpu {name} {th} threads: {dt:.2f} s, {len(adj) / dt:.3e} edges/s")
    key = f"R-MAT 20 (ef 16, seed 2), {nth} threads"
    return {"metric": "reference CPU TC edges/sec (hostCountTriangles, full count)",
            "value": runs[key]["edges_per_s"], "unit": "edges/s",
            "ms_per_step": runs[key]["seconds"] * 1e3, "higher_is_better": True,
            "dtype": "int32",
            "config": {"workload": "CPU reference path, full counts of ER-16 and R-MAT 20",
                       "cpu_model": cpu_model(), "threads": nth, "runs": runs},
            "roofline": None, "cpu_baseline": None}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="static", choices=["static", "stream", "batched", "copy", "cpu"])
    ap.add_argument("--scale", type=int, default=24)
    ap.add_argument("--ef", type=int, default=32)
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-capture", action="store_true",
                    help="stream workload: no size classes, no recorded batches")
    ap.add_argument("--selector", default=os.path.join(ROOT, "custingeralg_amd",
                                                       "selector_model.json"))
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU "
              "sharding of this path is a design, not a run", file=sys.stderr)
        return 2

    if args.workload == "cpu":
        T = None
    else:
        import custingeralg_amd as T
        log(f"[bench] device {T.device_name()}, workload {args.workload}")
    fn = {"static": bench_static, "stream": bench_stream, "batched": bench_batched,
          "copy": bench_copy, "cpu": bench_cpu}[args.workload]
    rec = fn(T, args)
    line = {"metric": rec.pop("metric"), "value": rec.pop("value"), "unit": rec.pop("unit"),
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": rec.pop("ms_per_step"), "higher_is_better": rec.pop("higher_is_better"),
            "scaling": "weak", "vs_baseline": None, "dtype": rec.pop("dtype"),
            "data": "synthetic"}
    line.update(rec)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
