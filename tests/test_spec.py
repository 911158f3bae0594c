"""CPU tests of BWHHIST1 declared-history specs (src/real_history_spec.hpp,
src/real_branch_spec.hpp): the oracle's spec mode pinned to the SHA-256 of the reference's own
brute-force spec oracles (src/panel_real_history_oracle.cpp, src/panel_real_branch_oracle.cpp)
on specs written by the reference's own builders (tests/golden/make_spec.py); the drop-in
driver's spec loader and depth-first walk checked against the builders' receipts (no GPU)."""
import json
import os
import struct
import subprocess
import sys

import pytest

from conftest import GOLDEN, fixture_file, golden, spec_files

sys.path.insert(0, GOLDEN)
from make_spec import write_synthetic_spec  # noqa: E402

SPECS = golden("spec_cases.json")


@pytest.mark.parametrize("case", sorted(SPECS))
def test_oracle_spec_matches_reference(oracle_bin, tmp_path, case):
    base, spec = spec_files(case, tmp_path)
    out = subprocess.run([oracle_bin, "spec", "--fixture", base, "--spec", spec, "--out", str(tmp_path / "a.bin")],
                         capture_output=True, text=True, check=True).stdout
    r = json.loads(out.strip().splitlines()[-1])
    assert r["n_answers"] == (SPECS[case]["states"] - 1) * SPECS[case]["q"]
    assert r["answers_sha256"] == SPECS[case]["oracle"]


def _edit_offset(spec_path, v):
    """Byte offset of edit v in a spec file (either header layout)."""
    size = os.path.getsize(spec_path)
    with open(spec_path, "rb") as f:
        h = f.read(124)
    for nf in (4, 5):
        K, q = struct.unpack_from("<ii", h, 8)
        nins, nep = struct.unpack_from("<q", h, 8 + 4 * nf + 8)[0], struct.unpack_from("<q", h, 8 + 4 * nf + 32)[0]
        hdr = 8 + 4 * nf + 96
        if hdr + K * 4 + K * 40 + K * q * 16 + nins * 12 + nep * 8 == size:
            return hdr + K * 4 + 40 * v
    raise AssertionError("layout")


def test_oracle_spec_planted_mistake(oracle_bin, tmp_path):
    """One factor edit's value changed in the oracle's input must change the hash: the golden
    hash is sensitive to the factor path, not only to row edits."""
    case = "h2_chain_fp4"
    base, spec = spec_files(case, tmp_path)
    raw = bytearray(open(spec, "rb").read())
    off = _edit_offset(spec, 4)  # fp 4: version 4 is a factor edit
    assert raw[off] == 4
    val = struct.unpack_from("<q", raw, off + 24)[0]
    struct.pack_into("<q", raw, off + 24, val + 1)
    bad = tmp_path / "bad.spec"
    bad.write_bytes(bytes(raw))
    out = subprocess.run([oracle_bin, "spec", "--fixture", base, "--spec", str(bad), "--out", str(tmp_path / "a.bin")],
                         capture_output=True, text=True, check=True).stdout
    assert json.loads(out.strip().splitlines()[-1])["answers_sha256"] != SPECS[case]["oracle"]


@pytest.mark.parametrize("case", sorted(SPECS))
def test_panel_spec_plan_only(pkg, tmp_path, case):
    """The driver reads both header layouts; a chain walks onto itself, a branching history onto
    one chain state per version entry plus one per exit before the last entry."""
    g = SPECS[case]
    base, spec = spec_files(case, tmp_path)
    p = subprocess.run([pkg.BW_PANEL, "--fixture", base, "--spec", spec, "--answers", str(tmp_path / "a.bin"),
                        "--json", str(tmp_path / "a.json"), "--plan-only"], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    r = json.loads(p.stdout.strip().splitlines()[-1])
    s, rc = r["spec"], g["receipt"]
    assert s["versions"] == g["states"] and s["q"] == g["q"] and s["n_ins"] == rc["n_ins"]
    if g["kind"] == "chain":
        assert s["header_fields"] == 4 and not s["branching"] and s["walk_states"] == g["states"]
        assert (s["inserts"], s["deletes"], s["credits"], s["factor_edits"]) == \
            (rc["inserts"], rc["deletes"], rc["credits"], rc["factor_edits"])
    else:
        assert s["header_fields"] == 5 and s["branching"]
        # exits before the last entry: the first branch (fork + 1 .. its tip) and the second
        # branch (its factor node .. its tip) are left before the merged replay is entered
        exits = (rc["branch_a_tip"] - rc["fork_parent"]) + (rc["branch_b_tip"] - rc["branch_a_tip"])
        assert s["walk_states"] == g["states"] + exits
        assert s["deletes"] == rc["real_reversals"] and s["factor_edits"] == 3


def test_panel_spec_rejects_bad_input(pkg, tmp_path):
    base, spec = spec_files("h1_chain_fp37", tmp_path)
    raw = open(spec, "rb").read()
    cut = tmp_path / "cut.spec"
    cut.write_bytes(raw[:-8])
    other = fixture_file([2000, 4, 86400000000, 20261131])  # another relation (2,000 rows, 4 groups)
    for fx, sp, msg in ((base, str(cut), "neither header layout"), (other, spec, "disagree")):
        p = subprocess.run([pkg.BW_PANEL, "--fixture", fx, "--spec", sp, "--answers", str(tmp_path / "a.bin"),
                            "--json", str(tmp_path / "a.json"), "--plan-only"], capture_output=True, text=True,
                           timeout=120)
        assert p.returncode == 3 and msg in p.stderr, (p.returncode, p.stderr)


def synthetic_case(tmp_path, K=1500, q=3, seed=77):
    base, spec = str(tmp_path / "syn.base"), str(tmp_path / "syn.spec")
    write_synthetic_spec(base, spec, K, q, seed)
    return base, spec


def test_synthetic_spec_all_ops(oracle_bin, pkg, tmp_path):
    """A branching spec with every op (moves and replaces whose undo must restore hidden row
    values): the oracle's spec mode answers it, the driver walks it (entries + exits)."""
    base, spec = synthetic_case(tmp_path)
    out = subprocess.run([oracle_bin, "spec", "--fixture", base, "--spec", spec, "--out", str(tmp_path / "o.bin")],
                         capture_output=True, text=True, check=True).stdout
    assert json.loads(out.strip().splitlines()[-1])["n_answers"] == 1499 * 3
    p = subprocess.run([pkg.BW_PANEL, "--fixture", base, "--spec", spec, "--answers", str(tmp_path / "a.bin"),
                        "--json", str(tmp_path / "a.json"), "--plan-only"], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    s = json.loads(p.stdout.strip().splitlines()[-1])["spec"]
    assert s["branching"] and s["header_fields"] == 5 and 1500 < s["walk_states"] < 3000


@pytest.mark.parametrize("K,q", [(2, 1), (7, 5)])
def test_tiny_specs(oracle_bin, pkg, tmp_path, K, q):
    """The smallest history (one version) and a short ragged one: the oracle answers them and the
    driver plans them."""
    base, spec = synthetic_case(tmp_path, K=K, q=q, seed=80 + K)
    out = subprocess.run([oracle_bin, "spec", "--fixture", base, "--spec", spec, "--out", str(tmp_path / "o.bin")],
                         capture_output=True, text=True, check=True).stdout
    assert json.loads(out.strip().splitlines()[-1])["n_answers"] == (K - 1) * q
    p = subprocess.run([pkg.BW_PANEL, "--fixture", base, "--spec", spec, "--answers", str(tmp_path / "a.bin"),
                        "--json", str(tmp_path / "a.json"), "--plan-only"], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    s = json.loads(p.stdout.strip().splitlines()[-1])["spec"]
    assert s["versions"] == K and s["q"] == q and 1 <= s["walk_states"] - 1 <= 2 * (K - 1)
