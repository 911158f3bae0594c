"""Regenerates the golden answer hashes under tests/golden/ (run in the build container).

tree_cases.json: branching multi-update workloads (the benchmark configs' shapes).
  * "oracle"  SHA-256 of the answers of oracle/bw_oracle tree (rollback-Fenwick);
  * "brute"   the same workload solved by the oracle's independent ancestor-walk (small cases);
  * "ref"     the reference's own Solver (src/oracle.cpp:225-451) run by oracle/_ref/tree_harness_*
              on the same workload (small cases; needs the reference sources to build).
stream_cases.json: request streams of the batched (form 4) and served (form 5) forms
  (oracle/bw_stream.h): "oracle" = bw_oracle stream (rollback-Fenwick over the tree extended by
  the appended states), "brute" = its independent ancestor walk (small cases), "stream" = the
  SHA-256 of the generated stream (the product's generator must reproduce it).
ref_cases.json: reference-native workloads (one edit per state, src/common.hpp:279-374) on
  fixtures written by make_fixture.py; "oracle" = oracle/_ref/oracle_K, "compiled" =
  oracle/_ref/compiled_K (the reference's CPU v7 batch engine, BWH_BATCH_FORCE=1), "mine" = the
  oracle restatement.  All three must agree; the GPU tests compare against them.
Only answer hashes and workload parameters are stored -- no reference source or output vectors.
"""
import json
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
ORACLE = os.path.join(REPO, "oracle", "bw_oracle")
REF = os.path.join(REPO, "oracle", "_ref")
# scratch files (solver dumps, answer files) and the regenerated fixtures live in the system temp dir
WORK = os.path.join(tempfile.gettempdir(), "bwgpu_golden")
FIXTURES = os.path.join(tempfile.gettempdir(), "bwgpu_fixtures")

TREE_CASES = [
    # name, cfg, S, U, Q, N, seed, solvers
    ("t1_cfg1_small", 1, 250, 16, 8, 65536, 99, ("brute", "ref")),
    ("t2_cfg2_small", 2, 500, 8, 8, 1 << 20, 99, ("brute", "ref")),
    ("t3_cfg3_small", 3, 250, 16, 8, 65536, 99, ("brute", "ref")),
    ("t4_cfg4_small", 4, 500, 8, 8, 1 << 20, 99, ("brute", "ref")),
    ("t5_cfg2_dense", 2, 250, 16, 8, 4096, 5, ("brute", "ref")),
    ("t6_cfg1_single", 1, 1, 1, 1, 2, 3, ("brute",)),
    ("t7_cfg2_deep", 2, 20000, 16, 4, 1 << 20, 11, ("brute",)),
    ("c1_cfg1", 1, 1000, 64, 16, 65536, 20261204, ("brute",)),
    ("c2_cfg2", 2, 300000, 128, 32, 1 << 20, 20261205, ()),
    ("c3_cfg3", 3, 300000, 256, 32, 4194304, 20261206, ()),
    ("c4_cfg4", 4, 300000, 128, 32, 1 << 20, 20261207, ()),
]

STREAM_CASES = [
    # name, form, tree (S, U, N, seed or None for the form's default), stream args, solvers
    ("s1_batched_small", 4, (300, 8, 4096, None), ["--requests", "50", "--max-batch", "64"], ("brute",)),
    ("s2_batched_mid", 4, (20000, 16, 1 << 20, None), ["--requests", "2000", "--max-batch", "512"], ()),
    ("s3_served_small", 5, (300, 8, 4096, None),
     ["--requests", "200", "--max-batch", "64", "--append-U", "8", "--rate", "1000"], ("brute",)),
    ("s4_served_mid", 5, (20000, 16, 1 << 20, None),
     ["--requests", "2000", "--max-batch", "1024", "--append-U", "16", "--rate", "5000"], ()),
    ("S4_batched_full", 4, None, [], ()),
    ("S5_served_full", 5, None, [], ()),
    # the served form over a config-3 tree: raw 32-bit keys and endpoints through the rank-LUT
    ("s6_served_lut_small", 5, (300, 8, 4096, None),
     ["--cfg", "3", "--requests", "200", "--max-batch", "64", "--append-U", "8", "--rate", "1000"], ("brute",)),
    ("s7_served_lut_mid", 5, (20000, 16, 65536, None),
     ["--cfg", "3", "--requests", "2000", "--max-batch", "1024", "--append-U", "16", "--rate", "5000"], ()),
    ("S6_served_lut_full", 5, None, ["--cfg", "3"], ()),
]

REF_CASES = [
    # name, fixture (rows, groups, span, seed), states, topology, seed, factor period, q, per_state
    ("r1_chain_fp37_q2", (2000, 4, 86_400_000_000, 20261131), 2049, 0, 7, 37, 2, True),
    ("r2_chain_fp4_q8", (2000, 4, 86_400_000_000, 20261131), 2049, 0, 7, 4, 8, True),
    ("r3_chain_shared_q8", (2000, 4, 86_400_000_000, 20261131), 2049, 0, 11, 37, 8, False),
    ("r4_chain_fp37_q2_g64", (20000, 64, 86_400_000_000, 7), 2049, 0, 7, 37, 2, True),
    ("r5_chain_9001_q2", (20000, 64, 86_400_000_000, 7), 9001, 0, 5, 37, 2, True),
    ("r6_chain_9001_q8_fp4", (5000, 16, 3_600_000_000, 9), 9001, 0, 5, 4, 8, True),
]

# paper-scale reference-native cases (300,000 states, the reference's own workload shape): the
# O(states x rows) oracles are infeasible here, so the pin is the reference's compiled CPU v7
# engine, run with 1 and with 8 blocks (it agreed with both oracles on every case above).
REF_BIG = [
    ("R1_chain_300k_fp37_q2", (100000, 64, 86_400_000_000, 7), 300000, 0, 7, 37, 2, True),
    ("R2_chain_300k_fp4_q8", (100000, 64, 86_400_000_000, 7), 300000, 0, 7, 4, 8, True),
]


def sha_of(cmd, env=None):
    p = subprocess.run(cmd, capture_output=True, text=True, check=True, env=env)
    for line in p.stdout.splitlines()[::-1]:
        if line.startswith("{"):
            return json.loads(line)["answers_sha256"]
    raise RuntimeError(p.stdout + p.stderr)


def tree(only=None):
    os.makedirs(WORK, exist_ok=True)
    out = {}
    for name, cfg, S, U, Q, N, seed, solvers in TREE_CASES:
        if only and name not in only:
            continue
        base = [ORACLE, "tree", "--cfg", str(cfg), "--S", str(S), "--U", str(U), "--Q", str(Q),
                "--N", str(N), "--seed", str(seed)]
        dump = os.path.join(WORK, name + ".dump")
        rec = {"cfg": cfg, "S": S, "U": U, "Q": Q, "N": N, "seed": seed}
        rec["oracle"] = sha_of(base + (["--dump", dump] if "ref" in solvers else []))
        if "brute" in solvers:
            rec["brute"] = sha_of(base + ["--algo", "brute"])
        if "ref" in solvers:
            k = 1 + S * U
            exe = os.path.join(REF, f"tree_harness_{k}")
            subprocess.run(["make", "-C", os.path.join(REPO, "oracle"), f"_ref/tree_harness_{k}"], check=True,
                           stdout=subprocess.DEVNULL)
            rec["ref"] = sha_of([exe, dump, os.path.join(WORK, name + ".ref.bin")])
        shas = {v for k2, v in rec.items() if k2 in ("oracle", "brute", "ref")}
        assert len(shas) == 1, (name, rec)
        out[name] = rec
        print(name, rec, flush=True)
    return out


def stream_args(form, tree_, args):
    cmd = ["stream", "--form", str(form)]
    if tree_:
        S, U, N, seed = tree_
        cmd += ["--S", str(S), "--U", str(U), "--N", str(N)] + (["--seed", str(seed)] if seed else [])
    return cmd + list(args)


def streams(only=None):
    out = {}
    for name, form, tree_, args, solvers in STREAM_CASES:
        if only and name not in only:
            continue
        base = [ORACLE] + stream_args(form, tree_, args)
        rec = {"form": form, "tree": list(tree_) if tree_ else None, "args": list(args)}
        p = subprocess.run(base + ["--stream-sha"], capture_output=True, text=True, check=True)
        st = json.loads(p.stdout.strip().splitlines()[-1])
        rec["stream"] = st["stream_sha256"]
        rec["nq"] = st["nq"]
        rec["napp"] = st["napp"]
        rec["oracle"] = sha_of(base + ["--out", os.path.join(WORK, "s.bin")])
        if "brute" in solvers:
            rec["brute"] = sha_of(base + ["--algo", "brute", "--out", os.path.join(WORK, "sb.bin")])
            assert rec["brute"] == rec["oracle"], (name, rec)
        out[name] = rec
        print(name, rec, flush=True)
    return out


def fixture_path(fx):
    """Fixtures are regenerated from their parameters (make_fixture.py) and cached, not committed."""
    sys.path.insert(0, HERE)
    from make_fixture import make_fixture

    d = FIXTURES
    os.makedirs(d, exist_ok=True)
    p = os.path.join(d, "fx_%d_%d_%d_%d.bin" % tuple(fx))
    if not os.path.exists(p):
        make_fixture(p + ".tmp", *fx)
        os.replace(p + ".tmp", p)
    return p


def refbig():
    os.makedirs(WORK, exist_ok=True)
    out = {}
    for name, fx, K, topo, seed, fp, q, per_state in REF_BIG:
        fxpath = fixture_path(fx)
        common = ["--fixture", fxpath, "--topology", str(topo), "--seed", str(seed), "--factor-period", str(fp),
                  "--q", str(q)]
        rec = {"fixture_params": list(fx), "states": K, "topology": topo, "seed": seed, "factor_period": fp, "q": q,
               "per_state": per_state}
        for threads in (1, 8):
            env = dict(os.environ, OMP_NUM_THREADS=str(threads))
            if per_state:
                env["BWH_RANGE_SCOPE"] = "per_state"
            jp = os.path.join(WORK, "big.json")
            subprocess.run([os.path.join(REF, f"compiled_{K}")] + common + ["--answers", os.path.join(WORK, "b.bin"),
                           "--json", jp], capture_output=True, text=True, check=True, env=env)
            with open(jp) as f:
                rec[f"compiled_b{threads}"] = json.load(f)["answers_sha256"]
        assert rec["compiled_b1"] == rec["compiled_b8"], rec
        rec["compiled"] = rec["compiled_b1"]
        out[name] = rec
        print(name, rec, flush=True)
    return out


def refnative():
    os.makedirs(WORK, exist_ok=True)
    out = {}
    for name, fx, K, topo, seed, fp, q, per_state in REF_CASES:
        fxpath = fixture_path(fx)
        env = dict(os.environ)
        if per_state:
            env["BWH_RANGE_SCOPE"] = "per_state"
        common = ["--fixture", fxpath, "--topology", str(topo), "--seed", str(seed), "--factor-period", str(fp),
                  "--q", str(q)]
        rec = {"fixture_params": list(fx), "states": K, "topology": topo,
               "seed": seed, "factor_period": fp, "q": q, "per_state": per_state}
        rec["oracle"] = sha_of([os.path.join(REF, f"oracle_{K}")] + common + ["--out", os.path.join(WORK, "o.bin")],
                               env)
        cenv = dict(env, BWH_BATCH_FORCE="1", OMP_NUM_THREADS="8")
        p = subprocess.run([os.path.join(REF, f"compiled_{K}")] + common +
                           ["--answers", os.path.join(WORK, "c.bin"), "--json", os.path.join(WORK, "c.json")],
                           capture_output=True, text=True, check=True, env=cenv)
        with open(os.path.join(WORK, "c.json")) as f:
            rec["compiled"] = json.load(f)["answers_sha256"]
        rec["mine"] = sha_of([ORACLE, "ref", "--states", str(K)] + common + (["--per-state"] if per_state else []) +
                             ["--out", os.path.join(WORK, "m.bin")])
        assert rec["oracle"] == rec["compiled"] == rec["mine"], (name, rec)
        out[name] = rec
        print(name, rec, flush=True)
    return out


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "all"
    if what in ("all", "tree"):
        with open(os.path.join(HERE, "tree_cases.json"), "w") as f:
            json.dump(tree(), f, indent=1, sort_keys=True)
    if what in ("all", "stream"):
        os.makedirs(WORK, exist_ok=True)
        with open(os.path.join(HERE, "stream_cases.json"), "w") as f:
            json.dump(streams(), f, indent=1, sort_keys=True)
    if what in ("all", "ref"):
        cases = refnative()
        cases.update(refbig())
        with open(os.path.join(HERE, "ref_cases.json"), "w") as f:
            json.dump(cases, f, indent=1, sort_keys=True)
