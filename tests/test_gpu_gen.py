"""GPU parity of the device workload generator (include/bwgpu_gen.h): the request streams it
materialises hash to the golden stream SHA-256s (the oracle's generator's), and the tree workloads
equal the host generator's (itself pinned to the oracle's generator) at the benchmark sizes.
Integer work: bit-exact."""
import pytest

from conftest import golden, run_json

STREAMS = golden("stream_cases.json")


def _stream_cmd(g):
    cmd = ["--form", str(g["form"])]
    if g["tree"]:
        S, U, N, seed = g["tree"]
        cmd += ["--S", str(S), "--U", str(U), "--N", str(N)] + (["--seed", str(seed)] if seed else [])
    return cmd + g["args"]


@pytest.mark.gpu
@pytest.mark.parametrize("case", list(STREAMS))
def test_device_generated_stream_hashes_to_golden(pkg, case):
    g = STREAMS[case]
    r = run_json([pkg.BW_BATCH] + _stream_cmd(g) + ["--stream-sha", "--device-gen"], timeout=300)
    assert r["device_gen"], r
    assert r["stream_sha256"] == g["stream"] and r["nq"] == g["nq"] and r["napp"] == g["napp"], (case, r)


@pytest.mark.gpu
@pytest.mark.parametrize("cfg,S,U,Q,N", [
    (1, 1000, 64, 16, 65536),
    (2, 300000, 128, 32, 1 << 20),
    (3, 300000, 256, 32, 4194304),
    (4, 300000, 128, 32, 1 << 20),
    (5, 800, 16, 8, 4096),
])
def test_device_generated_tree_equals_host(pkg, cfg, S, U, Q, N):
    args = [pkg.BW_TREE, "--cfg", str(cfg), "--S", str(S), "--U", str(U), "--Q", str(Q), "--N", str(N),
            "--workload-sha"]
    host = run_json(args, timeout=300)
    dev = run_json(args + ["--device-gen"], timeout=300)
    assert dev["device_gen"] and not host["device_gen"]
    assert dev["workload_sha256"] == host["workload_sha256"], (cfg, dev, host)
