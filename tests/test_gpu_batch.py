"""GPU parity of the batched form (bw_batch: every batch of a request stream answered by one
hipGraph replay over a resident index) against the stream oracle's golden hashes.  Integer work:
bit-exact SHA-256 of the whole answer vector in (batch, query) order; the first batches are also
answered one by one the plain way and must match the batched answers element for element."""
import pytest

from conftest import golden, run_json

STREAMS = golden("stream_cases.json")


def _cmd(pkg, g, extra):
    cmd = [pkg.BW_BATCH]
    if g["tree"]:
        S, U, N, seed = g["tree"]
        cmd += ["--S", str(S), "--U", str(U), "--N", str(N)] + (["--seed", str(seed)] if seed else [])
    return cmd + g["args"] + ["--warmup", "0", "--repeat", "2"] + extra


@pytest.mark.gpu
@pytest.mark.parametrize("case,knobs", [
    ("s1_batched_small", []),
    ("s1_batched_small", ["--rmax", "64", "--tile-shift", "6", "--chunk", "128", "--threads", "64", "--waves", "1"]),
    ("s2_batched_mid", []),
    ("s2_batched_mid", ["--rmax", "1024", "--tile-shift", "9", "--chunk", "2048", "--threads", "128"]),
])
def test_batched_matches_oracle(pkg, case, knobs):
    g = STREAMS[case]
    r = run_json(_cmd(pkg, g, knobs + ["--plain-sample", "6"]), timeout=300)
    assert r["answers_sha256"] == g["oracle"], (case, knobs)
    assert r["n_answers"] == g["nq"]
    assert r["plain"]["mismatch"] == 0


@pytest.mark.gpu
def test_batched_full_stream(pkg):
    """100,000 batches (102.5M queries) over the 300,000-state tree, tuned knobs."""
    g = STREAMS["S4_batched_full"]
    r = run_json(_cmd(pkg, g, ["--plain-sample", "3", "--tuned", pkg.TUNED]), timeout=900)
    assert r["answers_sha256"] == g["oracle"]
    assert r["plain"]["mismatch"] == 0


@pytest.mark.gpu
@pytest.mark.parametrize("case,knobs", [
    ("s1_batched_small", []),
    ("s1_batched_small", ["--rmax", "64", "--tile-shift", "6", "--chunk", "128", "--ep-per-work", "65",
                          "--range-slots", "100", "--stage-shift", "8"]),
    ("s2_batched_mid", ["--range-slots", "65536", "--ep-per-work", "129"]),
    ("s2_batched_mid", ["--bt-table-mb", "0", "--device-gen"]),  # positions by binary search, not the table
    ("S4_batched_full", ["--tuned", None, "--device-gen"]),
])
def test_device_plan_equals_host_plan(pkg, case, knobs):
    """The device planner (sorts, scans and searches on the GPU) builds the host planner's plan
    array for array; the queries replayed through it give the oracle's answers.  With
    --device-gen the queries planned and replayed are the device generator's (equal to the
    host's element for element), so the stream never crosses PCIe."""
    g = STREAMS[case]
    knobs = [pkg.TUNED if k is None else k for k in knobs]
    r = run_json(_cmd(pkg, g, knobs + ["--plain-sample", "0", "--device-plan"]), timeout=900)
    assert r["device_plan"]["used"] and r["device_plan"]["mismatch"] == 0, (case, knobs)
    assert r["answers_sha256"] == g["oracle"], (case, knobs)
    if "--device-gen" in knobs:
        assert r["device_gen"]["used"] and r["device_gen"]["mismatch"] == 0, r["device_gen"]


@pytest.mark.gpu
@pytest.mark.parametrize("case,knobs", [
    ("s1_batched_small", []),
    ("s1_batched_small", ["--rmax", "64", "--tile-shift", "6", "--chunk", "128", "--ep-per-work", "65",
                          "--range-slots", "100", "--stage-shift", "8"]),
    ("s2_batched_mid", ["--bt-table-mb", "0"]),
    ("S4_batched_full", ["--tuned", None]),
])
def test_device_index_equals_host_index(pkg, case, knobs):
    """The index built on the device (the greedy block cut by per-step block ends and pointer
    doubling, per-block keys and ranks by LDS hash set and bitonic sort, chains by a stable radix
    sort of the tile slots) equals the host-built index array for array; the stream planned and
    answered over it gives the oracle's answers."""
    g = STREAMS[case]
    knobs = [pkg.TUNED if k is None else k for k in knobs]
    r = run_json(_cmd(pkg, g, knobs + ["--plain-sample", "0", "--device-index"]), timeout=900)
    d = r["device_index"]
    assert d["mismatch"] == 0 and d["answers_sha256"] == g["oracle"], (case, knobs, d)
    assert r["answers_sha256"] == g["oracle"]


@pytest.mark.gpu
@pytest.mark.parametrize("case,knobs", [
    ("s1_batched_small", []),
    ("s2_batched_mid", ["--range-slots", "65536"]),
    ("S4_batched_full", ["--tuned", None]),
])
def test_all_device_pipeline(pkg, case, knobs):
    """From the seed to the answers on the device: the tree's updates and the stream's queries
    generated, the index built (walk, cut, tables), the stream planned and answered, the answers
    hashed -- the oracle's answers."""
    g = STREAMS[case]
    knobs = [pkg.TUNED if k is None else k for k in knobs]
    r = run_json(_cmd(pkg, g, knobs + ["--plain-sample", "0", "--device-all"]), timeout=900)
    d = r["device_all"]
    assert d["mismatch"] == 0 and d["answers_sha256"] == g["oracle"], (case, d)



@pytest.mark.gpu
@pytest.mark.parametrize("case,knobs", [
    ("s1_batched_small", []),
    ("s1_batched_small", ["--dense-rmax", "64", "--dense-threads", "256"]),
    ("s1_batched_small", ["--dense-store", "2"]),
    ("s2_batched_mid", []),
    ("s2_batched_mid", ["--dense-store", "1"]),
    ("s2_batched_mid", ["--dense-rmax", "1024", "--dense-threads", "512"]),
    ("s2_batched_mid", ["--dense-rmax", "2048", "--dense-threads", "256", "--device-gen"]),
    ("s2_batched_mid", ["--dense-threads", "512", "--dense-wg", "2", "--dense-store", "2"]),
    # work items of a block's queries, each its own workgroup's (the one-state edge case of 2^20
    # queries is split over many)
    ("s1_batched_small", ["--dense-chunk", "256"]),
    ("s2_batched_mid", ["--dense-chunk", "300", "--dense-threads", "512"]),
    ("s2_batched_mid", ["--dense-chunk", "4096", "--dense-threads", "1024", "--dense-store", "1"]),
    ("S4_batched_full", ["--tuned", None]),
    ("S4_batched_full", ["--tuned", None, "--dense-chunk", "2048"]),
])
def test_dense_form_matches_oracle(pkg, case, knobs):
    """The dense-carry form (include/bwgpu_dense.h): carry rows built from the blocks' weights, the
    stream sorted by state rank on the device, one persistent work-queue replay launch (replayed
    from a hipGraph) -- the oracle's answers at every block size and workgroup shape, small trees
    (one block, or blocks of 64 keys) to the full 102.5M-query stream, with the answers stored at
    their ids directly or staged (sorted order, partition by id window, die-local scatter; the
    default once the ids span more than one window, so from s2 on).  Each run also checks the plan's
    edge cases against root-path sums computed on the host: the stream's first queries all on one
    state (one bucket, one group), on four states, with out-of-range queries mixed in (flag set,
    answer 0), and batches of 1 and 7 queries (`dense.edge`, counted in `dense.mismatch`)."""
    g = STREAMS[case]
    knobs = [pkg.TUNED if k is None else k for k in knobs]
    r = run_json(_cmd(pkg, g, knobs + ["--plain-sample", "0", "--dense"]), timeout=900)
    d = r["dense"]
    assert d["mismatch"] == 0 and d["answers_sha256"] == g["oracle"], (case, knobs, d)
    assert d["edge"]["mismatch"] == 0 and d["edge"]["queries"] == min(g["nq"], 1 << 20), d["edge"]
    assert r["answers_sha256"] == g["oracle"]
    if pkg.TUNED in knobs:  # the tuned file's dense entry is the one the run used
        assert d["knobs_from"] == pkg.TUNED, d
