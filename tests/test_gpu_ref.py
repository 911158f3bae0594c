"""GPU parity of the reference drop-in path: bw_panel drives only the drop-in C-ABI (gpu_v7.h,
gpu_v7b.h, gpu_session.h, bwgpu_carry.h) over the reference's own declared workloads and must
reproduce the complete-answer SHA-256 of the reference oracle (src/oracle.cpp) and of the
reference's compiled CPU v7 engine (src/panel_batch_v7fix.cpp), bit for bit."""
import json
import os
import subprocess

import pytest

from conftest import GOLDEN, fixture_file, golden, run_json, spec_files

REF = golden("ref_cases.json")
SMALL = [k for k, v in REF.items() if v["states"] <= 9001]
# R2 (300,000 states, factor period 4, q=8 per state) has 3.18M classes; the reference's own v7
# engine rejects it with its correction-table gate (src/batch_engine_v7.hpp:611-616) and answers it
# with a different CPU arm, so it has no GPU-path counterpart and is only a CPU golden.
BIG = ["R1_chain_300k_fp37_q2"]


def _cmd(pkg, g, tmp_path, arm, carry="gpu", blocks=256, threads=256):
    cmd = [pkg.BW_PANEL, "--fixture", fixture_file(g["fixture_params"]), "--topology", str(g["topology"]),
           "--seed", str(g["seed"]), "--factor-period", str(g["factor_period"]), "--q", str(g["q"]),
           "--states", str(g["states"]), "--arm", arm, "--carry", carry, "--blocks", str(blocks),
           "--threads", str(threads), "--answers", str(tmp_path / "a.bin"), "--json", str(tmp_path / "a.json")]
    if g["per_state"]:
        cmd.append("--per-state")
    return cmd


LDS_LIMIT = 160 * 1024


@pytest.mark.gpu
@pytest.mark.parametrize("case", SMALL)
@pytest.mark.parametrize("arm", ["v7", "v7b", "sess", "sess_global"])
def test_dropin_matches_reference(pkg, tmp_path, case, arm):
    g = REF[case]
    cmd = _cmd(pkg, g, tmp_path, arm)
    print("RUN", " ".join(cmd), flush=True)
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    if p.returncode != 0 and arm in ("v7", "v7b") and "of LDS" in p.stderr:
        # the shared-memory kernels refuse class domains beyond one CU's LDS with a clear error
        # (the reference's own kernels stop at 96 KiB, src/gpu_v7.cu:102-103); the session arms
        # of the same case must then be exact
        assert p.returncode == 5, p.stderr
        r = run_json(_cmd(pkg, g, tmp_path, "sess"), timeout=300)
        assert r["answers_sha256"] == g["compiled"], (case, "sess")
        return
    assert p.returncode == 0, p.stderr[-2000:]
    r = json.loads([ln for ln in p.stdout.splitlines() if ln.startswith("{")][-1])
    assert r["answers_sha256"] == g["compiled"], (case, arm)


@pytest.mark.gpu
@pytest.mark.parametrize("blocks,threads,carry", [(1, 64, "cpu"), (7, 128, "gpu"), (32, 1024, "cpu"),
                                                  (2048, 256, "gpu")])
@pytest.mark.parametrize("arm", ["v7", "v7b", "sess_global"])
def test_dropin_block_counts(pkg, tmp_path, arm, blocks, threads, carry):
    """B = 1 .. states (every carry boundary), odd workgroup sizes, host and device carry."""
    g = REF["r5_chain_9001_q2"]
    r = run_json(_cmd(pkg, g, tmp_path, arm, carry, blocks, threads), timeout=300)
    assert r["answers_sha256"] == g["compiled"], (arm, blocks, threads, carry)


@pytest.mark.gpu
def test_session_repeated_batches_stay_exact(pkg, tmp_path):
    """Resident session: every batch restores the snapshots and replays; the answers never drift."""
    g = REF["r1_chain_fp37_q2"]
    r = run_json(_cmd(pkg, g, tmp_path, "sess", blocks=64) + ["--repeat", "3"], timeout=300)
    assert r["answers_sha256"] == g["compiled"]


@pytest.mark.gpu
@pytest.mark.parametrize("case", BIG)
@pytest.mark.parametrize("arm", ["sess", "v7b"])
def test_dropin_paper_scale(pkg, tmp_path, case, arm):
    """300,000 states: the reference's CPU v7 engine hash (identical with 1 and 8 blocks)."""
    g = REF[case]
    r = run_json(_cmd(pkg, g, tmp_path, arm, blocks=512), timeout=600)
    assert r["answers_sha256"] == g["compiled"], (case, arm)


SPECS = golden("spec_cases.json")


@pytest.mark.gpu
@pytest.mark.parametrize("case", sorted(SPECS))
@pytest.mark.parametrize("arm", ["v7", "v7b", "sess", "sess_global"])
def test_dropin_spec_matches_reference(pkg, tmp_path, case, arm):
    """BWHHIST1 declared histories written by the reference's own builders, through the drop-in
    C-ABI: chain histories as they are, the branching history (trunk, two branches with declared
    factor nodes, merged replay) as the chain of its depth-first walk.  Every version's answers
    must hash to the reference's brute-force spec oracle (src/panel_real_history_oracle.cpp,
    src/panel_real_branch_oracle.cpp)."""
    base, spec = spec_files(case, tmp_path)
    cmd = [pkg.BW_PANEL, "--fixture", base, "--spec", spec, "--arm", arm, "--blocks", "64",
           "--answers", str(tmp_path / "a.bin"), "--json", str(tmp_path / "a.json")]
    plan = run_json(cmd + ["--plan-only"], timeout=120)
    print("RUN", " ".join(cmd), flush=True)
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    if arm in ("v7", "v7b") and p.returncode != 0 and "of LDS" in p.stderr:
        # the branching case's 10,380 classes (two int64 rows of them) exceed one CU's LDS for
        # the shared-memory kernels: a clear refusal, only for a class domain that large (the
        # session arms answer it)
        assert p.returncode == 5 and plan["ncls"] * 16 > LDS_LIMIT * 3 // 4, (plan, p.stderr[-500:])
        return
    assert p.returncode == 0, p.stderr[-2000:]
    r = json.loads([ln for ln in p.stdout.splitlines() if ln.startswith("{")][-1])
    assert r["n_answers"] == (SPECS[case]["states"] - 1) * SPECS[case]["q"]
    assert r["answers_sha256"] == SPECS[case]["oracle"], (case, arm)


@pytest.mark.gpu
@pytest.mark.parametrize("arm", ["v7", "v7b", "sess", "sess_global"])
@pytest.mark.parametrize("seed,blocks,K,q", [(77, 64, 1500, 3), (78, 1, 1500, 3), (79, 256, 1500, 3),
                                         (80, 1, 2, 1), (81, 4, 7, 5)])
def test_dropin_synthetic_spec_all_ops(pkg, oracle_bin, tmp_path, arm, seed, blocks, K, q):
    """A branching history with every reference op (inserts, deletes, replaces, moves, factor
    edits) through the depth-first walk: the undo states must restore every row's hidden values
    exactly.  Checked against the oracle's spec mode on the same files."""
    import sys

    sys.path.insert(0, GOLDEN)
    from make_spec import write_synthetic_spec

    base, spec = str(tmp_path / "syn.base"), str(tmp_path / "syn.spec")
    write_synthetic_spec(base, spec, K, q, seed)  # K = 2: one version, the smallest history
    ref = run_json([oracle_bin, "spec", "--fixture", base, "--spec", spec, "--out", str(tmp_path / "o.bin")])
    r = run_json([pkg.BW_PANEL, "--fixture", base, "--spec", spec, "--arm", arm, "--blocks", str(blocks),
                  "--answers", str(tmp_path / "a.bin"), "--json", str(tmp_path / "a.json")], timeout=300)
    assert r["n_answers"] == ref["n_answers"]
    assert r["spec"]["branching"] or K <= 2  # a one-version history is a chain
    assert r["answers_sha256"] == ref["answers_sha256"], (seed, arm, blocks)
