"""The allocation-free entry point (bwd_answer_noalloc_launch, library-held scratch) gives the
answers of the workspace entry point (bwd_answer_launch) on two batches in a row, the second
smaller, both called through the C-ABI by bw_batch; both equal the whole stream's answers (the
oracle's), and a batch beyond the scratch is refused with an error."""
import pytest

from conftest import golden, run_json

STREAMS = golden("stream_cases.json")


def _cmd(pkg, g, extra):
    cmd = [pkg.BW_BATCH]
    if g["tree"]:
        S, U, N, seed = g["tree"]
        cmd += ["--S", str(S), "--U", str(U), "--N", str(N)] + (["--seed", str(seed)] if seed else [])
    return cmd + g["args"] + ["--warmup", "0", "--repeat", "1", "--plain-sample", "0"] + extra


@pytest.mark.gpu
@pytest.mark.parametrize("case,knobs", [
    ("s2_batched_mid", ["--dense-rmax", "1024"]),
    ("S4_batched_full", ["--tuned", None]),
])
def test_noalloc_matches_workspace_entry(pkg, case, knobs):
    g = STREAMS[case]
    knobs = [pkg.TUNED if k is None else k for k in knobs]
    r = run_json(_cmd(pkg, g, knobs + ["--noalloc-check"]), timeout=900)
    n = r["noalloc"]
    assert n["mismatch"] == 0, n
    assert n["m2"] < n["m1"] and n["m1_needs_bytes"] <= n["scratch_bytes"], n
    assert r["dense"]["answers_sha256"] == g["oracle"]
    if case == "S4_batched_full":
        assert n["refuses_oversize"], n
