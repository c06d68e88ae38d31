"""Train the learned strategy selector on the GPU and save it.

    python -m custingeralg_amd.train_selector [--out PATH] [--tune-scale 24] [--timings]

By default the costs are measured in situ: whole counts of R-MAT graphs of
scales 20, 22 and 24 (edge factor 16, seeds 100 + scale) with one group of
strategy-table cells switched at a time (learned.train_insitu); a group leaves
the fixed heuristic's strategy only for one that saved at least --min-gain of
the count on average.  With --timings, the costs come from one-wavefront
timings of sampled tasks instead.  The cost model is trained by the library's
own FP32 kernels, and the schedule is autotuned on an R-MAT graph of the tune
scale (edge factor 32, seed 5: a different graph from the benchmark's seed
4).  The JSON file holds the FP32 weights (rounded to BF16 when loaded), the
tuned schedule and how both were made.
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

from . import device_name, rmat_graph_device
from . import learned as Lm


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=Lm.default_model_path())
    ap.add_argument("--scales", default="20,22,24")
    ap.add_argument("--min-gain", type=float, default=0.015)
    ap.add_argument("--samples", type=int, default=4096)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--tune-scale", type=int, default=24)
    ap.add_argument("--tune-ef", type=int, default=32)
    ap.add_argument("--tune-seed", type=int, default=5)
    ap.add_argument("--timings", action="store_true", help="one-wavefront task timings")
    args = ap.parse_args(argv)
    scales = tuple(int(x) for x in args.scales.split(","))
    t0 = time.time()
    if args.timings:
        model, table, losses, ntimed = Lm.train_default(scales, 16, args.samples, args.steps)
        source = {"costs": "one-wavefront timings of sampled tasks", "timed_tasks": ntimed}
        print(f"[train] {ntimed} timed tasks from R-MAT {scales}: loss {losses[0]:.4f} -> "
              f"{losses[-1]:.4f} ({time.time() - t0:.1f}s)", flush=True)
    else:
        graphs = [rmat_graph_device(s, 16, 100 + s) for s in scales]
        try:
            model, table, losses, sweeps = Lm.train_insitu(graphs, steps=args.steps,
                                                            min_gain=args.min_gain)
        finally:
            for g in graphs:
                g.close()
        ncounts = sum(len(sw) * 3 for sw in sweeps)
        # per cell group: mean relative count time of each strategy over the graphs
        rel = {}
        for key in sweeps[0]:
            rel[f"rows {key[0][0]}-{key[0][1]} cols {key[1][0]}-{key[1][1]}"] = {
                str(st): round(float(np.mean([sw[key][st] for sw in sweeps])), 4)
                for st in sweeps[0][key]}
        for k, v in rel.items():
            print(f"[train]   {k}: {v}", flush=True)
        source = {"costs": "in situ: whole counts with one cell group switched",
                  "sweep_points": ncounts, "min_gain": args.min_gain,
                  "sweep_relative_time": rel}
        print(f"[train] in-situ sweeps of R-MAT {scales}: loss {losses[0]:.4f} -> "
              f"{losses[-1]:.4f} ({time.time() - t0:.1f}s)", flush=True)
    t0 = time.time()
    g = rmat_graph_device(args.tune_scale, args.tune_ef, args.tune_seed)
    Lm.set_default_schedule(None)
    tuned, ms = Lm.autotune(g)
    g.close()
    print(f"[train] autotune on R-MAT {args.tune_scale} ef {args.tune_ef} seed {args.tune_seed}: "
          f"{ms:.1f} ms, small={tuned.small_max} mid={tuned.mid_max} flush={tuned.flush_items} "
          f"job={tuned.job_items} ({time.time() - t0:.1f}s)", flush=True)
    meta = {"device": device_name(), "train_scales": list(scales), "edge_factor": 16,
            **source, "steps": args.steps, "loss_first": losses[0],
            "loss_last": losses[-1],
            "table_picks": np.bincount(table.ravel(), minlength=3).tolist(),
            "tune_graph": f"R-MAT {args.tune_scale} ef {args.tune_ef} seed {args.tune_seed}",
            "tune_ms": ms}
    Lm.save_selector(args.out, model, tuned, meta)
    Lm.unload()
    Lm.set_default_schedule(None)
    print(f"[train] wrote {os.path.relpath(args.out)}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
