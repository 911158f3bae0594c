"""Benchmark: exact range-sum queries/s of the GPU batch path on one MI355X.

Workload: the plain GPU drop-in at paper scale (benchmark config 2): N = 2^20 keys, 300,000
branching states (parent = previous w.p. 0.9, else uniform among the prior 1,024), 128 Zipf(1.1)
point updates per state with int64 deltas in [-10^6, 10^6], 32 log-uniform range queries per
state -> 9,600,000 exact int64 answers per pass.  A "step" is one complete pass (block replay,
carry chunks, carry scan, chunk replay, combine) over inputs already resident in HBM; the worker
times every pass with HIP events on its stream, bracketed by stream synchronisation.

Prints one JSON line.  `roofline` is the dominant kernel's algorithmic bytes per launch over its
average HIP-event duration against the MI355X HBM peak (8 TB/s); `cpu_baseline` is the oracle's
rollback-Fenwick solver timed on this host (1 thread) on the same workload.
"""
import argparse
import json
import os
import subprocess
import sys

REPO = os.path.dirname(os.path.abspath(__file__))
BW_TREE = os.path.join(REPO, "biworlddb-artifact_amd", "build", "bw_tree")
ORACLE = os.path.join(REPO, "oracle", "bw_oracle")
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(REPO, "profiles", "traffic_cfg2.json")


def last_json(text):
    return json.loads([ln for ln in text.splitlines() if ln.startswith("{")][-1])


def cpu_baseline(cfg):
    """Oracle (rollback-Fenwick DFS, 1 thread) on the full workload; generation time excluded."""
    p = subprocess.run([ORACLE, "tree", "--cfg", str(cfg)], capture_output=True, text=True, check=True, timeout=600)
    r = last_json(p.stdout)
    return r


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cfg", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build measures one GPU; the path's sharding is described in DESIGN.md "
              "(no multi-GPU run is made).", file=sys.stderr)
        sys.exit(2)
    if not os.path.exists(BW_TREE):
        sys.exit("bench.py: native build missing (make -C biworlddb-artifact_amd)")
    print(f"bench: bw_tree cfg {a.cfg} warmup {a.warmup} steps {a.steps}", file=sys.stderr, flush=True)
    p = subprocess.run([BW_TREE, "--cfg", str(a.cfg), "--warmup", str(a.warmup), "--repeat", str(a.steps)],
                       capture_output=True, text=True, timeout=1200)
    if p.returncode != 0:
        sys.exit(f"bench.py: bw_tree failed ({p.returncode}): {p.stderr[-2000:]}")
    r = last_json(p.stdout)
    stages = {k: v for k, v in r["stage_ms"].items() if v > 0}
    dom = max(stages, key=stages.get)
    dom_ms = stages[dom]
    dom_bytes = r["stage_bytes"][dom]
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            traffic = json.load(f).get(dom)
    out = {
        "metric": "exact range-sum queries/s at 300k states (plain GPU drop-in, one launch sequence per query batch)",
        "value": round(r["queries_per_s"], 1),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(r["pass_ms"]["avg"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic (seeded counter-based generator, config 2 shape)",
        "config": {"workload": f"config {a.cfg}: {r['S']} states x {r['U']} updates, {r['Q']} queries/state, "
                               f"N={r['N']}", "states": r["S"], "updates_per_state": r["U"],
                   "queries_per_state": r["Q"], "keys": r["N"], "answers_per_step": r["n_answers"],
                   "answers_sha256": r["answers_sha256"]},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel_ms": round(dom_ms, 4), "kernel_bytes_alg": dom_bytes,
                     "pass_bytes_alg": r["bytes_alg"],
                     "pass_effective_GBs": round(r["bytes_alg"] / (r["pass_ms"]["avg"] * 1e-3) / 1e9, 1)},
        "stage_ms": r["stage_ms"],
        "setup_ms": r["phases_ms"]["setup"],
        "device": r["device"],
    }
    if not a.no_cpu:
        c = cpu_baseline(a.cfg)
        out["cpu_baseline"] = {"value": round(r["n_answers"] / c["solve_s"], 1), "unit": "queries/s", "cores": 1,
                               "kind": "port",
                               "sample": "the full config-2 workload (9.6M queries), oracle rollback-Fenwick DFS "
                                         "on 1 host thread, generation excluded; answers SHA "
                                         + ("equal" if c["answers_sha256"] == r["answers_sha256"] else "DIFFERENT"),
                               "seconds": round(c["solve_s"], 3)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
