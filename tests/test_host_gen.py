"""Device workload generator (csrc/gen_kernels.hip) checked on the CPU: its per-element functions
(csrc/host/gen_draws.hpp, the kernels' bodies) run under a CPU emulation with a stand-in for the
device's exp (tests/host/gen_emulator.cpp, test infrastructure), and every array must equal the
host generator's (tree_generate / stream_generate, itself pinned to the oracle's generator by
SHA).  The Zipf keys go through the guide-bracketed search, the log-uniform ranges through the tie
test and the host settlement; a libm that differs from the host's by a few ulps, or rounds
near-integers differently, must give identical arrays once ties are settled, and leaving the
settlement out (or a libm far outside the tie window) must be caught."""
import json
import os
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def emulator(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("gen") / "gen_emulator")
    host = os.path.join(REPO, "biworlddb-artifact_amd", "csrc", "host")
    subprocess.run(["g++", "-O2", "-std=c++17", "-fopenmp", "-Wno-deprecated-declarations",
                    "-I" + os.path.join(REPO, "include"), "-I" + host, "-o", out,
                    os.path.join(REPO, "tests", "host", "gen_emulator.cpp"), os.path.join(host, "tree_workload.cpp"),
                    "-lcrypto"], check=True, capture_output=True, timeout=300)
    return out


def _run(emu, args):
    p = subprocess.run([emu] + [str(a) for a in args], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr[-2000:]
    return json.loads(p.stdout.strip().splitlines()[-1])


@pytest.mark.parametrize("args", [
    ["tree", 1, 500, 64, 16, 65536, 20261204],
    ["tree", 2, 2000, 128, 32, 1 << 20, 20261205, "--ulps", 4],
    ["tree", 3, 1500, 256, 32, 65536, 99, "--ulps", 2],
    ["tree", 4, 3000, 32, 8, 1 << 20, 5],
    ["tree", 5, 800, 16, 8, 4096, 7, "--ulps", 8],
    ["stream", 4, 300, 4096, 0, 50, 64, 0, 128, 20000],
    ["stream", 5, 3000, 1 << 20, 0, 200, 64, 0.05, 8, 1000, "--ulps", 3],
])
def test_emulated_generator_equals_host_generator(emulator, args):
    r = _run(emulator, args)
    assert r["mismatch"] == 0 and r["elements"] > 0, r


@pytest.mark.parametrize("kind", ["tree", "stream"])
def test_ties_are_settled_by_the_host(emulator, kind):
    """A stand-in exp that snaps values within 1e-9 of an integer to it: many draws tie, the
    settlement restores every one; without it, floors cross and the arrays differ."""
    base = (["tree", 2, 300000, 8, 32, 1 << 20, 20261205] if kind == "tree"
            else ["stream", 4, 20000, 1 << 20, 0, 5000, 2048, 0, 128, 20000])
    good = _run(emulator, base + ["--snap", "1e-9"])
    assert good["ties"] > 100 and good["mismatch"] == 0, good
    bad = _run(emulator, base + ["--snap", "1e-9", "--no-settle"])
    assert bad["mismatch"] > 0, bad


def test_libm_far_outside_tie_window_is_caught(emulator):
    """An exp off by a relative 1e-12 (about 4,500 ulps, beyond the 64-ulp tie window) moves some
    floors without flagging them: the comparison must see it."""
    r = _run(emulator, ["tree", 2, 300000, 8, 32, 1 << 20, 20261205, "--rel", "1e-12"])
    assert r["mismatch"] > 0 and r["ties"] == 0, r
