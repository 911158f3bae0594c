"""GPU parity of the generalised tree path: bw_tree (HIP kernels behind bwt_* C-ABI) against the
oracle's golden hashes, which independent solvers (brute ancestor walk, the reference's own
Solver) pinned.  Integer work: the bar is bit-exact (SHA-256 equality of the answer vector)."""
import hashlib

import pytest

from conftest import golden, run_json

TREE = golden("tree_cases.json")

# (case, knobs): small blocks / chunks / segments stress every block, chunk and segment boundary
CASES = [
    ("t1_cfg1_small", []),
    ("t1_cfg1_small", ["--rmax", "64", "--tile-shift", "6", "--chunk", "256", "--seg", "2", "--threads", "64",
                       "--waves", "1"]),
    ("t2_cfg2_small", []),
    ("t2_cfg2_small", ["--rmax", "32", "--chunk", "128", "--seg", "3", "--threads", "128", "--waves", "2"]),
    ("t3_cfg3_small", []),
    ("t3_cfg3_small", ["--rmax", "48", "--tile-shift", "5", "--chunk", "200", "--seg", "1", "--waves", "3"]),
    ("t4_cfg4_small", ["--rmax", "16", "--tile-shift", "9", "--chunk", "64"]),
    ("t5_cfg2_dense", []),
    ("t5_cfg2_dense", ["--rmax", "17", "--tile-shift", "3", "--chunk", "100", "--seg", "4", "--threads", "1024"]),
    ("t6_cfg1_single", []),
    ("t7_cfg2_deep", ["--rmax", "512", "--chunk", "4096"]),
    ("c1_cfg1", []),
    # wide key tiles (many slots with both endpoints on one level-B chain: paired results), work
    # items cut at odd sizes and at small slot ranges (range-ordered work)
    ("c1_cfg1", ["--tile-shift", "12", "--ep-per-work", "65", "--range-slots", "100"]),
    ("t2_cfg2_small", ["--tile-shift", "11", "--chunk", "64", "--ep-per-work", "67", "--range-slots", "333"]),
    # 65,536 keys in tiles of 8 and 16: 8,193 and 4,097 key-tile classes (beyond 12 bits)
    ("c1_cfg1", ["--tile-shift", "3"]),
    ("c1_cfg1", ["--tile-shift", "4", "--chunk", "128"]),
]


def _cmd(pkg, g, knobs):
    return [pkg.BW_TREE, "--cfg", str(g["cfg"]), "--S", str(g["S"]), "--U", str(g["U"]), "--Q", str(g["Q"]),
            "--N", str(g["N"]), "--seed", str(g["seed"]), "--warmup", "0", "--repeat", "2"] + knobs


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["pass", "query"])
@pytest.mark.parametrize("case,knobs", CASES)
def test_tree_gpu_matches_golden(pkg, case, knobs, mode):
    """pass: index rebuilt inside every pass; query: resident index, only touched blocks/chunks."""
    g = TREE[case]
    r = run_json(_cmd(pkg, g, knobs + ["--mode", mode]), timeout=300)
    assert r["answers_sha256"] == g["oracle"], (case, knobs, mode)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["pass", "query"])
@pytest.mark.parametrize("case", ["c2_cfg2", "c4_cfg4", "c3_cfg3"])
def test_tree_gpu_full_size(pkg, case, mode, tmp_path):
    """The benchmark configs at full size (300,000 states): same complete-answer SHA as the oracle.
    In pass mode the answers are extracted both ways (serial copy / write / hashes, and the
    overlapped bw_extract_answers) into the answers file; checksums and file must agree."""
    g = TREE[case]
    path = str(tmp_path / "answers.bin")
    extra = ["--extract", "both", "--answers", path] if mode == "pass" else []
    r = run_json(_cmd(pkg, g, ["--mode", mode] + extra), timeout=600)
    assert r["answers_sha256"] == g["oracle"], case
    assert r["n_answers"] == g["S"] * g["Q"]
    if mode == "pass":
        assert r["extract"]["same"], r["extract"]
        with open(path, "rb") as f:
            assert hashlib.sha256(f.read()).hexdigest() == g["oracle"]


@pytest.mark.gpu
@pytest.mark.parametrize("ts", ["9", "10"])
def test_tree_gpu_cfg3_fine_tiles(pkg, ts):
    """Config 3 at full size over 4,194,304 ranks in key tiles of 512 or 1,024 (8,193 or 4,097
    key-tile classes), planned on the device: the oracle's complete-answer SHA."""
    g = TREE["c3_cfg3"]
    r = run_json(_cmd(pkg, g, ["--mode", "pass", "--device-plan", "--tile-shift", ts]), timeout=600)
    assert r["answers_sha256"] == g["oracle"], ts


@pytest.mark.gpu
def test_rank_remap_matches_host_ranks(pkg):
    """The device CSR rank-LUT remap of raw 32-bit endpoints equals the host ranks exactly."""
    g = TREE["c3_cfg3"]
    r = run_json(_cmd(pkg, g, ["--verify-remap"]), timeout=600)
    assert r["remap_mismatch"] == 0
    assert r["answers_sha256"] == g["oracle"]
