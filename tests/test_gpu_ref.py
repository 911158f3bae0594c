"""GPU parity of the reference drop-in path: bw_panel drives only the drop-in C-ABI (gpu_v7.h,
gpu_v7b.h, gpu_session.h, bwgpu_carry.h) over the reference's own declared workloads and must
reproduce the complete-answer SHA-256 of the reference oracle (src/oracle.cpp) and of the
reference's compiled CPU v7 engine (src/panel_batch_v7fix.cpp), bit for bit."""
import os

import pytest

from conftest import fixture_file, golden, run_json

REF = golden("ref_cases.json")
SMALL = [k for k, v in REF.items() if v["states"] <= 9001]
BIG = [k for k, v in REF.items() if v["states"] > 9001]


def _cmd(pkg, g, tmp_path, arm, carry="gpu", blocks=256, threads=256):
    cmd = [pkg.BW_PANEL, "--fixture", fixture_file(g["fixture_params"]), "--topology", str(g["topology"]),
           "--seed", str(g["seed"]), "--factor-period", str(g["factor_period"]), "--q", str(g["q"]),
           "--states", str(g["states"]), "--arm", arm, "--carry", carry, "--blocks", str(blocks),
           "--threads", str(threads), "--answers", str(tmp_path / "a.bin"), "--json", str(tmp_path / "a.json")]
    if g["per_state"]:
        cmd.append("--per-state")
    return cmd


@pytest.mark.gpu
@pytest.mark.parametrize("case", SMALL)
@pytest.mark.parametrize("arm", ["v7", "v7b", "sess", "sess_global"])
def test_dropin_matches_reference(pkg, tmp_path, case, arm):
    g = REF[case]
    r = run_json(_cmd(pkg, g, tmp_path, arm), timeout=300)
    assert r["answers_sha256"] == g["compiled"], (case, arm)


@pytest.mark.gpu
@pytest.mark.parametrize("blocks,threads,carry", [(1, 64, "cpu"), (7, 128, "gpu"), (32, 1024, "cpu"),
                                                  (2048, 256, "gpu")])
@pytest.mark.parametrize("arm", ["v7", "v7b", "sess_global"])
def test_dropin_block_counts(pkg, tmp_path, arm, blocks, threads, carry):
    """B = 1 .. states (every carry boundary), odd workgroup sizes, host and device carry."""
    g = REF["r2_chain_fp4_q8"]
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
@pytest.mark.parametrize("arm", ["v7b", "sess"])
def test_dropin_paper_scale(pkg, tmp_path, case, arm):
    """300,000 states: the reference's CPU v7 engine hash (identical with 1 and 8 blocks)."""
    g = REF[case]
    r = run_json(_cmd(pkg, g, tmp_path, arm, blocks=512), timeout=600)
    assert r["answers_sha256"] == g["compiled"], (case, arm)
