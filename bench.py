"""Benchmark: exact range-sum queries/s at 300,000 states -- the batched form against the plain GPU
drop-in, with the served form's p50/p99 latency -- on one MI355X.

Headline (`value`): the batched form, planning included.  A request stream of 100,000 query
batches (sizes uniform in [1, 2048], 102,546,012 queries over uniform states, log-uniform ranges)
over the 300,000-state tree with uneven chains (geometric runs, mean 64, max 4,096; 128 Zipf(1.1)
point updates per state, N = 2^20 keys).  A "step" is one pass of the hot path over the whole
stream, its queries already in HBM: the stream is planned on the device (a two-level bucket sort
of the queries by state, one hipGraph) and answered by ONE launch of a persistent block-replay
kernel whose workgroups take work items (a block's queries cut into runs, a block's runs queued
together) from a device work queue, over the dense-carry index resident
in HBM (one carry row of 2^20+1 int64 per block of up to 4,096 keys, built once; its build time is
reported beside); the replay stores the answers in sorted order and two streaming passes (a
partition by query-id window, a die-local scatter) put them in query-id order, all in a second
hipGraph.  The block size, the replay's workgroup shape, the work item and the answer store are the autotuner's
(the "dense" entry of tuned.json, written by bw_tune --form dense).  bw_batch, in a process of its
own, times each step with one pair of HIP events on its stream and checks every timed step's
answers (device FNV-1a); the answers' SHA-256 must equal the two-level form's (which the test
suite pins against the oracle).

Beside it, measured in the same run (every figure planning-inclusive unless named otherwise):
  two_level -- the same stream over the two-level carry index (blocks of 1,024 keys, level-C rows,
               level-B chunk chains): device plan + one hipGraph replay of block replay, chunk
               replay and combine, per step; its replay alone as a sub-field;
  plain     -- the generalised plain pass at paper scale (config 2: 300,000 states, 9.6M queries,
               one launch sequence that plans the batch on the device, rebuilds the index and
               answers), the same for config 3 (4,194,304 distinct 32-bit keys through the
               rank-LUT, on its own tuned knobs), and on the batched stream's first 64 batches the
               plain pass once per batch (host plan, outside the timed region);
  dropin    -- the reference-ABI kernels (gpu_v7b.h / gpu_session.h arms of bw_panel) on the
               reference's own 300,000-state workload R1 (chain, factor period 37, q = 2 per
               state: 599,998 answers), kernel time and execute phase, SHA-256 = the reference
               CPU engine's;
  reference_cpu -- that CPU engine itself (src/panel_batch_v7fix.cpp built from the reference
               sources into oracle/_ref/, OpenMP, B = threads) on R1, on the host's cores;
  served    -- the served form (bw_serve): 20,000 Poisson requests (log-uniform sizes in [1, 4096],
               half point / half range queries, 5% appended branch states) by continuous batching,
               each batch answered by the two-level form, or by the dense form from 2^20 queries on
               (--form hybrid: both indexes resident), p50/p99 request latency, at several offered
               rates up to saturation; and the same request law over the config-3 tree with its
               rank-LUT resident (served_rank_lut, two-level);
  roofline  -- the compulsory bytes of a step (20 B per query: state, two endpoints, answer; 12 B
               per update: key, delta) over the persistent block replay's average HIP-event time
               (bwd_replay_timed: events around that kernel on its stream), against the MI355X HBM
               peak (8 TB/s); the whole step's figure beside it; `traffic` is the rocprofv3
               FETCH_SIZE + WRITE_SIZE of that kernel per launch (profiles/), scaled by the same
               counters' reading for a copy of known size, when that summary exists; and the
               kernel's random-access floor: its 2 carry-row reads per query at the rate bw_floor
               measures on the device for random 8-byte gathers over 8 MiB rows, spread over every
               row and in the replay's own pattern (its workgroups, one row per work item, a
               block's items together), against its time;
  cpu_baseline -- the oracle's rollback-Fenwick stream solver (a port, not the reference engine:
               that engine cannot take branching multi-update trees) on the host's cores over the
               same whole stream; its answers' SHA-256 must equal the GPU's.
"""
import argparse
import importlib.util
import json
import os
import subprocess
import sys
import tempfile

REPO = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(REPO, "biworlddb-artifact_amd", "build")
BW_BATCH = os.path.join(BUILD, "bw_batch")
BW_TREE = os.path.join(BUILD, "bw_tree")
BW_SERVE = os.path.join(BUILD, "bw_serve")
BW_PANEL = os.path.join(BUILD, "bw_panel")
BW_FLOOR = os.path.join(BUILD, "bw_floor")
TUNED = os.path.join(REPO, "biworlddb-artifact_amd", "tuned.json")
ORACLE = os.path.join(REPO, "oracle", "bw_oracle")
REF_ENGINE = os.path.join(REPO, "oracle", "_ref", "compiled_300000")
REF_CASES = os.path.join(REPO, "tests", "golden", "ref_cases.json")
MAKE_FIXTURE = os.path.join(REPO, "tests", "golden", "make_fixture.py")
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(REPO, "profiles", "traffic_dense.json")
SERVED_RATES = (20000.0, 80000.0, 320000.0)  # offered request rates (req/s), then saturation


def last_json(text):
    return json.loads([ln for ln in text.splitlines() if ln.startswith("{")][-1])


def run(cmd, timeout, env=None):
    print("bench: " + " ".join(os.path.basename(c) if i == 0 else c for i, c in enumerate(cmd)), file=sys.stderr,
          flush=True)
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout, env=env)
    if p.returncode != 0:
        sys.exit(f"bench.py: {os.path.basename(cmd[0])} failed ({p.returncode}): {p.stderr[-2000:]}")
    return last_json(p.stdout)


def host_threads():
    """Host cores this process may use: its CPU affinity, bounded by OMP_NUM_THREADS when set."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def r1_case(tmp):
    """The reference's own 300,000-state workload R1 and its fixture (regenerated from parameters)."""
    with open(REF_CASES) as f:
        g = json.load(f)["R1_chain_300k_fp37_q2"]
    spec = importlib.util.spec_from_file_location("make_fixture", MAKE_FIXTURE)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    path = os.path.join(tmp, "r1_fixture.bin")
    mod.make_fixture(path, *g["fixture_params"])
    args = ["--fixture", path, "--topology", str(g["topology"]), "--seed", str(g["seed"]), "--factor-period",
            str(g["factor_period"]), "--q", str(g["q"])]
    return g, args


def dropin_legs(tmp, threads):
    g, args = r1_case(tmp)
    out = {"workload": "R1: chain of 300,000 states, one edit each, factor period 37, q = 2 per-state ranges "
                       "(599,998 answers), reference-native"}
    for arm in ("v7b", "sess"):
        j = os.path.join(tmp, f"p_{arm}.json")
        r = run([BW_PANEL] + args + ["--states", str(g["states"]), "--arm", arm, "--carry", "gpu", "--blocks", "64",
                                     "--threads", "256", "--per-state", "--answers", os.path.join(tmp, "p.bin"),
                                     "--json", j], 600)
        if r["answers_sha256"] != g["compiled"]:
            sys.exit(f"bench.py: drop-in {arm} answers {r['answers_sha256']} != reference {g['compiled']}")
        ph, gp = r["phases_ns"], r["gpu"]
        out[arm] = {"kernel_ms": round(gp["kernel_ns"] / 1e6, 4), "execute_ms": round(ph["execute"] / 1e6, 4),
                    "upload_ms": round(gp["upload_ns"] / 1e6, 4),
                    # outside the execute phase, with the host plan: the device arena, the pinned
                    # staging buffer and the copy engines' and host threads' start
                    "setup_alloc_ms": round(gp["alloc_ns"] / 1e6, 4),
                    "kernel_queries_per_s": round(r["n_answers"] / (gp["kernel_ns"] * 1e-9), 1),
                    "execute_queries_per_s": round(r["n_answers"] / (ph["execute"] * 1e-9), 1),
                    "full_ms": round(r["full_ns"] / 1e6, 3), "blocks": 64, "answers_sha256": r["answers_sha256"]}
    ref = None
    if os.path.exists(REF_ENGINE):
        j = os.path.join(tmp, "ref.json")
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), BWH_RANGE_SCOPE="per_state")
        print(f"bench: compiled_300000 (OMP_NUM_THREADS={threads})", file=sys.stderr, flush=True)
        p = subprocess.run([REF_ENGINE] + args + ["--answers", os.path.join(tmp, "ref.bin"), "--json", j],
                           capture_output=True, text=True, timeout=600, env=env)
        if p.returncode != 0:
            sys.exit(f"bench.py: reference engine failed: {p.stderr[-2000:]}")
        with open(j) as f:
            c = json.load(f)
        if c["answers_sha256"] != g["compiled"]:
            sys.exit("bench.py: reference engine answers differ from its golden hash")
        ph = c["phases_ns"]
        ref = {"kind": "reference", "binary": "oracle/_ref/compiled_300000 (src/panel_batch_v7fix.cpp)",
               "mode": c["mode"], "threads": threads, "execute_ms": round(ph["execute"] / 1e6, 3),
               "setup_ms": round(ph["setup"] / 1e6, 3), "full_ms": round(c["full_ns"] / 1e6, 3),
               "execute_queries_per_s": round(c["n_answers"] / (ph["execute"] * 1e-9), 1),
               "full_queries_per_s": round(c["n_answers"] / (c["full_ns"] * 1e-9), 1),
               "answers_sha256": c["answers_sha256"]}
    return out, ref


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-served", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: measures one GPU only (no multi-GPU run is made); the path would shard by contiguous "
              "block ranges of the Euler walk with one exclusive scan of per-GPU carry totals.", file=sys.stderr)
        sys.exit(2)
    for exe in (BW_BATCH, BW_TREE, BW_SERVE, BW_PANEL, BW_FLOOR):
        if not os.path.exists(exe):
            sys.exit("bench.py: native build missing (make -C biworlddb-artifact_amd)")
    knobs = ["--tuned", TUNED] if os.path.exists(TUNED) else []
    threads = host_threads()
    # the headline step in a process of its own, whose only large allocations are the two indexes
    # and the stream; the two-level form, the device-built index, the device generator and the
    # all-on-device leg in a second process over the same stream
    rd = run([BW_BATCH, "--warmup", str(a.warmup), "--repeat", str(a.steps), "--plain-sample", "0", "--dense"]
             + knobs, 900)
    r = run([BW_BATCH, "--warmup", str(a.warmup), "--repeat", str(a.steps), "--plain-sample", "64", "--stages",
             "--device-gen", "--device-index", "--device-all"] + knobs, 1200)
    d = rd["dense"]
    if d["mismatch"] != 0 or not d["answers_sha256"] == rd["answers_sha256"] == r["answers_sha256"]:
        sys.exit("bench.py: dense-carry answers differ from the two-level form's")
    plain = run([BW_TREE, "--cfg", "2", "--mode", "pass", "--device-plan", "--warmup", "2", "--repeat", "5",
                 "--extract", "both"] + knobs, 600)
    # config 3 on its own tuned knobs (the "plain3" entry: key tiles of 2^10 over its 2^22 ranks)
    plain3 = run([BW_TREE, "--cfg", "3", "--mode", "pass", "--device-plan", "--warmup", "2", "--repeat", "5"] + knobs,
                 600)
    # the rank remap's HBM fraction: two 4-byte endpoints read and two 4-byte ranks written per
    # query, and the CSR LUT (65,537 row offsets, one 16-bit low half per distinct key) read once
    m3, d3 = plain3["n_answers"], plain3["N"]
    remap_ms = plain3["stage_ms"]["rank_remap"]
    remap_bytes = 16 * m3 + 4 * 65537 + 2 * d3
    remap_gbs = remap_bytes / (remap_ms * 1e-3) / 1e9
    M, S, U = r["n_answers"], r["S"], r["U"]
    alg_bytes = 20.0 * M + 12.0 * S * U  # compulsory bytes of one step
    step_ms, replay_ms = d["step_ms"]["avg"], d["replay_ms"]
    phases = d["replay_phases_ms"]
    block_ms = phases["block"]  # the dominant kernel, timed alone by HIP events on its stream
    achieved = alg_bytes / (block_ms * 1e-3) / 1e9
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            traffic = json.load(f).get("bwd_block_kernel")
    # the block replay's own floor: its random stream (two carry-row reads per query) at the device's
    # measured random-gather rate, (a) spread over all rows and (b) in the replay's own pattern: its
    # workgroups each reading one row per work item, the items of a block (on one row) taken
    # together; the answers leave it in sorted order (coalesced) and reach their ids through the
    # partition and the die-local scatter, timed beside it
    items = d["work_items"] if d["chunk"] > 0 else d["nblk"]
    share = max(1, round(items / d["nblk"]))
    per_visit = max(1, round(2 * M / items))
    fl = run([BW_FLOOR, "--gathers", str(2 * M), "--scatters", str(M), "--workgroups", str(d["workgroups"]),
              "--visits", str(items), "--per-visit", str(per_visit), "--share", str(share)], 300)
    floor_ms = 2 * M / fl["gather"]["per_s"] * 1e3
    replay_floor_ms = 2 * M / fl["replay"]["per_s"] * 1e3
    tl = r["device_plan"]
    value = d["queries_per_s"]
    out = {
        "metric": "exact range-sum queries/s at 300k states, batched (device plan + one persistent-kernel launch per "
                  "stream of 100,000 query batches) vs the plain GPU drop-in, served p99",
        "value": round(value, 1),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(step_ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic (seeded counter-based generators: config-4 tree and its 100,000-batch request stream)",
        "config": {"workload": f"batched: {r['batches']} query batches (sizes U[1,{r['max_batch']}], "
                               f"{M} queries) over {S} states x {U} updates, N={r['N']}; device plan + persistent "
                               "replay over the dense-carry index",
                   "states": S, "updates_per_state": U, "keys": r["N"], "batches": r["batches"],
                   "answers_per_step": M, "answers_sha256": d["answers_sha256"],
                   "dense_index": {"rmax": d["rmax"], "blocks": d["nblk"], "carry_bytes": d["carry_bytes"],
                                   "carry_build_ms": d["carry_build_ms"], "threads": d["threads"],
                                   "wg_per_cu": d["wg_resident"], "store": d["store"], "chunk": d["chunk"],
                                   "work_items": d["work_items"], "lds_bytes": d["lds_bytes"],
                                   "knobs_from": os.path.basename(d["knobs_from"])}},
        "step": {"plan_ms": d["plan_ms"], "replay_ms": d["replay_ms"],
                 "replay_phases_ms": phases, "min_ms": d["step_ms"]["min"], "max_ms": d["step_ms"]["max"],
                 "every_step_checked": d["step_mismatch"] == 0},
        "two_level": {"step_queries_per_s": tl["step_queries_per_s"], "step_ms": tl["step_ms"],
                      "plan_ms": tl["ms"], "replay_ms": r["graph_ms"]["avg"],
                      "replay_queries_per_s": round(r["queries_per_s"], 1), "stage_ms": r["stage_ms"],
                      "plan_equals_host_plan": tl["mismatch"] == 0, "index_build_ms": r["phases_ms"]["index_build"],
                      "knobs": r["plan"]},
        "plain": {
            "form": "the generalised plain pass (bwt_pass_launch: one launch sequence per batch that plans the "
                    "batch on the device, rebuilds the index and answers); the reference-ABI kernels are `dropin`",
            "config2_queries_per_s": round(plain["queries_per_s"], 1),
            "config2_pass_ms": plain["pass_ms"]["avg"],
            "config2_stage_ms": plain["stage_ms"],
            "config2_answers_sha256": plain["answers_sha256"],
            "config3_queries_per_s": round(plain3["queries_per_s"], 1),
            "config3_pass_ms": plain3["pass_ms"]["avg"],
            "config3_stage_ms": plain3["stage_ms"],
            "config3_answers_sha256": plain3["answers_sha256"],
            "config3_plan": plain3["plan"],
            "config3_rank_remap": {"ms": remap_ms, "bytes_alg": remap_bytes, "achieved_GBps": round(remap_gbs, 1),
                                   "frac_of_hbm_peak": round(remap_gbs / HBM_PEAK_GBS, 4),
                                   "rule": "2 x 4 B endpoints read + 2 x 4 B ranks written per query + the CSR LUT "
                                           "(row offsets and 16-bit low halves) once"},
            "per_batch_pass_queries_per_s": round(r["plain"]["pass_queries_per_s"], 1),
            "per_batch_query_only_queries_per_s": round(r["plain"]["query_queries_per_s"], 1),
            "per_batch_sample_batches": r["plain"]["batches"],
            "per_batch_note": "each of the first batches alone: index rebuild + answers (host plan, outside the clock)",
        },
        "batched_vs_plain": round(value / plain["queries_per_s"], 3),
        "batched_vs_plain_per_batch": round(value / max(1.0, r["plain"]["pass_queries_per_s"]), 1),
        "two_level_vs_plain": round(tl["step_queries_per_s"] / plain["queries_per_s"], 3),
        "extract_config2": {"answers": plain["n_answers"], "serial_ms": plain["extract"]["serial"],
                            "overlap_ms": plain["extract"]["overlap"], "same": plain["extract"]["same"]},
        "roofline": {"bound": "hbm", "kernel": "bwd_block_kernel", "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "kernel_ms": round(block_ms, 4), "kernel_bytes_alg": alg_bytes,
                     "bytes_alg_rule": "20 B per query (state, lo, hi, answer) + 12 B per update (key, delta)",
                     "step_achieved": round(alg_bytes / (step_ms * 1e-3) / 1e9, 1),
                     "step_frac": round(alg_bytes / (step_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic_over_alg": round(traffic / alg_bytes, 2) if traffic else None,
                     # the measured HBM traffic of a launch (profiles/) over the kernel's time: how busy
                     # the DRAM is, whatever share of the bytes the algorithm needs
                     "traffic_GBps": round(traffic / (block_ms * 1e-3) / 1e9, 1) if traffic else None,
                     "traffic_frac": round(traffic / (block_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if traffic else None,
                     "random_access": {
                         "rule": "2 random 8-byte carry-row reads per query; on gfx950 each such read fetches a "
                                 "128-byte line",
                         "gathers": 2 * M, "gathers_per_s": round(fl["gather"]["per_s"], 1),
                         "floor_ms": round(floor_ms, 4), "kernel_ms": round(block_ms, 4),
                         "floor_over_kernel": round(floor_ms / block_ms, 3),
                         "replay_pattern": {
                             "rule": "the same loads by the replay's workgroups, one row per work item, the items "
                                     "of a block read together (bw_floor --workgroups/--visits/--per-visit/--share)",
                             "workgroups": d["workgroups"], "work_items": items, "items_per_row": share,
                             "rows_at_once": fl["replay"]["rows_at_once"],
                             "gathers_per_s": round(fl["replay"]["per_s"], 1),
                             "floor_ms": round(replay_floor_ms, 4),
                             "floor_over_kernel": round(replay_floor_ms / block_ms, 3)},
                         "random_scatters_per_s": round(fl["scatter"]["per_s"], 1),
                         "copy_GBps": fl["copy"]["GBps"]}},
        "device_index": {"total_ms": r["device_index"]["total_ms"], "device_ms": r["device_index"]["gpu_ms"],
                         "host_checks_ms": r["device_index"]["host_ms"],
                         "host_index_ms": r["device_index"]["host_plan_ms"],
                         "equals_host_index": r["device_index"]["mismatch"] == 0},
        "all_on_device": {k: r["device_all"][k] for k in ("generate_ms", "index_ms", "alloc_ms", "plan_answer_ms",
                                                          "extract_ms", "total_ms")}
        | {"same_answers": r["device_all"]["mismatch"] == 0},
        "device_gen": {"gpu_ms": r["device_gen"]["gpu_ms"], "wall_ms": r["device_gen"]["wall_ms"],
                       "host_generator_ms": r["device_gen"]["host_ms"],
                       "equals_host_generator": r["device_gen"]["mismatch"] == 0, "ties": r["device_gen"]["ties"]},
        "device": r["device"],
    }
    with tempfile.TemporaryDirectory() as tmp:
        out["dropin"], ref = dropin_legs(tmp, threads)
        if ref:
            out["reference_cpu"] = ref
    if not a.no_served:
        curve = []
        for rate in SERVED_RATES + (1e9,):
            s = run([BW_SERVE, "--form", "hybrid", "--check-alone", "16" if rate == SERVED_RATES[0] else "0", "--rate",
                     str(rate)] + knobs, 600)
            runs = None
            if rate >= 1e9:
                # offered at once the stream is two batches and ~10 ms of wall time: three runs, the
                # median one reported, all three rates listed
                tries = [s] + [run([BW_SERVE, "--form", "hybrid", "--check-alone", "0", "--rate", str(rate)] + knobs, 600)
                               for _ in range(2)]
                tries.sort(key=lambda t: t["queries_per_s"])
                s = tries[1]
                runs = [round(t["queries_per_s"], 1) for t in tries]
            curve.append({"offered_requests_per_s": rate if rate < 1e9 else "saturated",
                          "offered_queries_per_s": round(s["offered_queries_per_s"], 1),
                          "queries_per_s": round(s["queries_per_s"], 1), "latency_ms": s["latency_ms"],
                          "batches": s["batches"], "mean_batch_queries": s["mean_batch_queries"],
                          "batch_gpu_ms": s["batch_gpu_ms"], "answers_sha256": s["answers_sha256"],
                          "alone_mismatch": s["alone"]["mismatch"], "answer_form": s["answer_form"],
                          "dense_from": s["dense_from"], "dense_batches": s["dense_batches"]}
                         | ({"runs_queries_per_s": runs} if runs else {}))
        if len({c["answers_sha256"] for c in curve}) != 1:
            sys.exit("bench.py: served answers depend on the offered rate")
        out["served"] = curve[0] | {"requests": s["requests"], "appends": s["appends"], "planner": s["planner"]}
        out["served_curve"] = curve
        u = run([BW_SERVE, "--cfg", "3", "--check-alone", "16"] + knobs, 600)
        out["served_rank_lut"] = {"queries_per_s": round(u["queries_per_s"], 1),
                                  "offered_queries_per_s": round(u["offered_queries_per_s"], 1),
                                  "latency_ms": u["latency_ms"], "requests": u["requests"], "appends": u["appends"],
                                  "answers_sha256": u["answers_sha256"], "alone_mismatch": u["alone"]["mismatch"]}
    if not a.no_cpu:
        p = subprocess.run([ORACLE, "stream", "--form", "4", "--threads", str(threads)],
                           capture_output=True, text=True, check=True, timeout=900)
        c = last_json(p.stdout)
        if c["answers_sha256"] != d["answers_sha256"]:
            sys.exit(f"bench.py: CPU baseline answers {c['answers_sha256']} != GPU {d['answers_sha256']}")
        out["cpu_baseline"] = {"value": round(c["n_answers"] / c["solve_s"], 1), "unit": "queries/s",
                               "cores": threads, "kind": "port",
                               "sample": f"the whole batched stream ({c['n_answers']} queries, {r['batches']} batches) "
                                         "over the same 300,000-state tree: the oracle's rollback-Fenwick DFS (tree "
                                         f"replay + queries) cut into {threads} pieces on {threads} host threads (the "
                                         "host's cores), generation excluded; answers' SHA-256 equal to the GPU's",
                               "seconds": round(c["solve_s"], 3)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
