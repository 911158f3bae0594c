"""Overlapped extract (bw_extract_answers, csrc/host/extract.cpp) checked on the CPU: built against a
stand-in for its two HIP runtime calls (tests/host/extract_check.cpp), its chunked, threaded
copy / file write / SHA-256 / FNV-1a must produce the bytes and checksums that one-shot hashing of
the same answers gives; the checksums are pinned independently by Python's hashlib and a Python
FNV-1a 64 over the written file, and the empty input by the empty-message values (SHA-256's
published one; FNV-1a's start value as the reference defines it, src/common.hpp:396, which is
not the published FNV offset basis)."""
import hashlib
import json
import os
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def extract_check(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("extract") / "extract_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wno-deprecated-declarations", "-D__HIP_PLATFORM_AMD__",
                    "-I/opt/rocm/include", "-I" + os.path.join(REPO, "include"), "-o", out,
                    os.path.join(REPO, "tests", "host", "extract_check.cpp"),
                    os.path.join(REPO, "biworlddb-artifact_amd", "csrc", "host", "extract.cpp"), "-lcrypto",
                    "-lpthread"], check=True, capture_output=True, timeout=300)
    return out


def _fnv1a64(data):
    h = 1469598103934665603
    for b in data:
        h = ((h ^ b) * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
    return h


@pytest.mark.parametrize("n", [0, 1, 1000, 524288, 524289, 3 * 524288 + 17])
def test_overlapped_extract_equals_one_shot(extract_check, tmp_path, n):
    path = str(tmp_path / "a.bin")
    p = subprocess.run([extract_check, str(n), path], capture_output=True, text=True, timeout=120, check=True)
    r = json.loads(p.stdout)
    assert r["rc"] == 0 and r["copied"], r
    assert r["sha"] == r["sha_one_shot"] and r["fnv"] == r["fnv_one_shot"], r
    with open(path, "rb") as f:
        data = f.read()
    assert len(data) == 8 * n
    assert hashlib.sha256(data).hexdigest() == r["sha"]
    if n <= 1000:
        assert "%016x" % _fnv1a64(data) == r["fnv"]
    if n == 0:
        assert r["sha"] == "e3b0c44298fc1c149afbf4c8996fb92427ae41e4649b934ca495991b7852b855"
        assert r["fnv"] == "14650fb0739d0383"


def test_unwritable_path_is_reported(extract_check, tmp_path):
    p = subprocess.run([extract_check, "10", str(tmp_path / "missing_dir" / "a.bin")], capture_output=True, text=True,
                       timeout=60, check=True)
    assert json.loads(p.stdout)["rc"] == -1


@pytest.fixture(scope="module")
def fnv_check(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("fnv") / "fnv_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I" + os.path.join(REPO, "biworlddb-artifact_amd", "csrc", "host"),
                    "-o", out, os.path.join(REPO, "tests", "host", "fnv_check.cpp")], check=True, capture_output=True,
                   timeout=300)
    return out


@pytest.mark.parametrize("n", [0, 1, 3, 4095, 4096, 4097, 100000, 300001, 17000003])
def test_fnv_block_split_equals_sequential(fnv_check, n):
    """The device FNV-1a's block split (low-byte maps for all 256 entries, composed 64 at a time
    over up to three levels, block polynomials in the prime) restated step by step on the CPU
    equals the sequential hash, at sizes that end inside a word, a block and a group."""
    r = json.loads(subprocess.run([fnv_check, str(n), "5"], capture_output=True, text=True, check=True,
                                  timeout=300).stdout)
    assert r["split"] == r["sequential"], r
    assert r["levels"] == (1 if n <= 64 * 4096 else 2 if n <= 64 * 64 * 4096 else 3)


def test_fnv_file_mode_pins_the_start_value(fnv_check, tmp_path):
    path = tmp_path / "b.bin"
    data = bytes(range(256)) * 3 + b"xyz"
    path.write_bytes(data)
    r = json.loads(subprocess.run([fnv_check, "--file", str(path)], capture_output=True, text=True, check=True,
                                  timeout=60).stdout)
    assert r["sequential"] == "%016x" % _fnv1a64(data)

