"""Writes the BWHHIST1 history-spec fixtures and their golden answers (run in a tree that has
the reference sources and `make -C oracle`).

Source relation: a BWBASE02 fixture from this project's own generator (splitmix64 draws of its
own definition): uniform timestamps and groups, positive amounts; after some lines a
cancellation line -- the same group and magnitude, negated, a little later -- and some credit
lines whose magnitude matches no positive line.  The reference's own builders, compiled from its
sources by oracle/Makefile, turn it into declared histories:
  * chain specs: src/real_history_build.cpp (a version-0 snapshot of the rows before
    --period-lo, one edit per version from the later rows in time order -- appends, reversals
    of a present line, credit appends -- and a declared group-factor edit every fp-th version;
    two ranges per version);
  * branch specs: src/real_branch_build.cpp (a trunk, two branches with declared factor nodes,
    and a merged descendant that replays branch B on top of branch A's tip; 5-field header).
The reference's brute-force spec oracles (src/panel_real_history_oracle.cpp,
src/panel_real_branch_oracle.cpp, built with kStates = the spec's state count) give the SHA-256
of every version's answers in (version, range) order.  Each case's version-0 base fixture and
spec are committed gzipped under tests/golden/specs/; spec_cases.json holds the SHA-256s, the
builder's receipt and the oracle's own checks (the branch oracle compares the tree replay of the
two tips and the merged tip with independent sequential replays).

    python tests/golden/make_spec.py
"""
import gzip
import json
import os
import shutil
import struct
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(REPO, "oracle", "_ref")
OUT = os.path.join(HERE, "specs")
sys.path.insert(0, HERE)
from make_fixture import _splitmix  # noqa: E402

SPAN = 1_000_000_000
PERIOD_LO = 400_000_000


def make_source(path, n_pos, groups, seed, cancel_every=6, credit_every=11):
    """n_pos positive lines, a cancellation after every cancel_every-th one and an unmatched
    credit after every credit_every-th one (source order = generation order)."""
    d = _splitmix(seed, 5 * n_pos).reshape(n_pos, 5)
    t = (d[:, 0] % np.uint64(SPAN)).astype(np.int64)
    g = (d[:, 1] % np.uint64(groups)).astype(np.int32)
    a = (d[:, 2] % np.uint64(100000)).astype(np.int64) + 1
    lag = (d[:, 3] % np.uint64(SPAN // 20)).astype(np.int64) + 1
    cg = (d[:, 4] % np.uint64(groups)).astype(np.int32)
    rows = []
    for i in range(n_pos):
        rows.append((int(t[i]), int(a[i]), int(g[i])))
        if i % cancel_every == cancel_every - 1 and t[i] + lag[i] < SPAN:
            rows.append((int(t[i] + lag[i]), -int(a[i]), int(g[i])))
        if i % credit_every == credit_every - 1:
            # magnitudes above every positive line's: never matched, so a credit append
            rows.append((int((t[i] * 7 + lag[i]) % SPAN), -(100000 + int(a[i])), int(cg[i])))
    rec = np.zeros(len(rows), dtype=[("t", "<i8"), ("a", "<i8"), ("g", "<i4"), ("p", "<i4")])
    arr = np.array(rows, dtype=np.int64)
    rec["t"], rec["a"], rec["g"] = arr[:, 0], arr[:, 1], arr[:, 2].astype(np.int32)
    hdr = struct.pack("<8sqiiqqqqq", b"BWBASE02", len(rows), groups, 0, len(rows), 0, int(rec["t"].min()),
                      int(rec["t"].max()), 2)
    with open(path, "wb") as f:
        f.write(hdr)
        f.write(rec.tobytes())
        f.write(np.full(groups, 1000, dtype="<i8").tobytes())
        f.write(np.arange(groups, dtype="<i4").tobytes())


def write_synthetic_spec(base_path, spec_path, K, q, seed, n_base=600, groups=5, window=24):
    """A branching BWHHIST1 spec (5-field layout) of this project's own definition, with every
    reference op: inserts of new rows, deletes, replaces and moves of base rows, group-factor
    edits; parent[v] uniform over the `window` versions before v.  Replaces and moves touch base
    rows only (a replace keeps the row's hidden time and group, so it must not meet a row whose
    hidden values differ between readers' initialisations).  The base fixture is written too."""
    d = _splitmix(seed, 3 * n_base + K * (8 + 2 * q)).astype(np.uint64)
    it = iter(d.tolist())
    nxt = lambda: next(it)  # noqa: E731
    span = 1_000_000
    base = [(nxt() % span, nxt() % 90000 + 1, nxt() % groups) for _ in range(n_base)]
    rec = np.zeros(n_base, dtype=[("t", "<i8"), ("a", "<i8"), ("g", "<i4"), ("p", "<i4")])
    rec["t"], rec["a"], rec["g"] = [b[0] for b in base], [b[1] for b in base], [b[2] for b in base]
    hdr = struct.pack("<8sqiiqqqqq", b"BWBASE02", n_base, groups, 0, n_base, 0, 0, span - 1, 2)
    with open(base_path, "wb") as f:
        f.write(hdr + rec.tobytes() + np.full(groups, 1000, dtype="<i8").tobytes() +
                np.arange(groups, dtype="<i4").tobytes())
    parent = [-1] + [max(0, v - 1 - int(nxt() % window)) for v in range(1, K)]
    edits = [(0, 0, 0, 0, 0)]  # (op, g, id, val, t)
    ins_t, ins_g = [], []
    for v in range(1, K):
        roll = nxt() % 100
        if roll < 30:
            t, g = nxt() % span, nxt() % groups
            edits.append((1, g, n_base + len(ins_t), int(nxt() % 200001) - 100000, t))
            ins_t.append(t)
            ins_g.append(g)
        elif roll < 50:
            # delete a base row or a row inserted earlier (it may be absent on this path)
            nrow = n_base + len(ins_t)
            edits.append((2, 0, int(nxt() % nrow), 0, 0))
        elif roll < 70:
            edits.append((0, 0, int(nxt() % n_base), int(nxt() % 200001) - 100000, 0))
        elif roll < 88:
            edits.append((3, int(nxt() % groups), int(nxt() % n_base), int(nxt() % 200001) - 100000, int(nxt() % span)))
        else:
            edits.append((4, int(nxt() % groups), 0, 900 + int(nxt() % 201), 0))
    qs = []
    for v in range(K):
        for _ in range(q):
            # endpoints on a 1,000-wide grid: about a thousand classes, so the shared-memory
            # arms (v7, v7b) take the case too
            a, b = sorted((int(nxt() % (span + 2000)) // 1000 * 1000 - 1000, int(nxt() % (span + 2000)) // 1000 * 1000 - 1000))
            qs.append((a, b))
    eps = sorted({x for ab in qs[q:] for x in ab})
    with open(spec_path, "wb") as f:
        f.write(b"BWHHIST1" + struct.pack("<5i", K, q, 0, groups, 1))
        n_ins = len(ins_t)
        nins = sum(1 for e in edits[1:] if e[0] == 1)
        ndel = sum(1 for e in edits[1:] if e[0] == 2)
        nfac = sum(1 for e in edits[1:] if e[0] == 4)
        f.write(struct.pack("<12q", n_base, n_ins, 0, span - 1, len(eps), 0, 0, 0, nins, ndel, 0, nfac))
        f.write(struct.pack(f"<{K}i", *parent))
        for op, g, i, val, t in edits:
            f.write(struct.pack("<B3xii4xqqq", op, g, 0, i, val, t))  # 40 bytes, id at 16
        for a, b in qs:
            f.write(struct.pack("<qq", a, b))
        f.write(struct.pack(f"<{n_ins}q", *ins_t) + struct.pack(f"<{n_ins}i", *ins_g))
        f.write(struct.pack(f"<{len(eps)}q", *eps))


def _run(cmd):
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(" ".join(cmd) + "\n" + p.stderr)
    return p.stdout


def _gz(src, dst):
    with open(src, "rb") as fi, gzip.GzipFile(dst, "wb", mtime=0) as fo:
        shutil.copyfileobj(fi, fo)


def main():
    os.makedirs(OUT, exist_ok=True)
    cases = {}
    with tempfile.TemporaryDirectory() as tmp:
        src = os.path.join(tmp, "source.bin")
        make_source(src, n_pos=4000, groups=6, seed=20261203)
        chain = [("h1_chain_fp37", 37), ("h2_chain_fp4", 4)]
        for name, fp in chain:
            base, spec = os.path.join(tmp, name + ".base"), os.path.join(tmp, name + ".spec")
            receipt = json.loads(_run([os.path.join(REF, "real_history_build"), "--source", src, "--period-lo",
                                       str(PERIOD_LO), "--period-hi", str(SPAN), "--bucket", "10000000",
                                       "--factor-period", str(fp), "--versions", "2049", "--out-base", base,
                                       "--out-spec", spec]).splitlines()[-1])
            aj = os.path.join(tmp, name + ".json")
            _run([os.path.join(REF, "history_oracle_2049"), "--fixture", base, "--spec", spec, "--stride", "1",
                  "--answers", os.path.join(tmp, name + ".ans"), "--json", aj])
            oj = json.load(open(aj))
            _gz(base, os.path.join(OUT, name + ".base.gz"))
            _gz(spec, os.path.join(OUT, name + ".spec.gz"))
            cases[name] = {"kind": "chain", "states": 2049, "q": 2, "factor_period": fp, "receipt": receipt,
                           "oracle": oj["answers_sha256"]}
        # branch: 1005 trunk + 500 (A) + 2*270 (B, merged replay) + 4 = 2049 versions
        K, NA, NB, q = 1005, 500, 270, 4
        name = "b1_branch_merge"
        base, spec = os.path.join(tmp, name + ".base"), os.path.join(tmp, name + ".spec")
        receipt = json.loads(_run([os.path.join(REF, "real_branch_build"), "--source", src, "--period-lo",
                                   str(PERIOD_LO), "--trunk-events", str(K), "--branch-a-events", str(NA),
                                   "--branch-b-events", str(NB), "--q", str(q), "--seed", "7", "--out-base", base,
                                   "--out-spec", spec]).splitlines()[-1])
        vFA = K + 1
        vAtip = vFA + NA
        vFB = vAtip + 1
        vBtip = vFB + NB
        vFM = vBtip + 1
        aj = os.path.join(tmp, name + ".json")
        _run([os.path.join(REF, "branch_oracle_2049"), "--fixture", base, "--spec", spec, "--stride", "1",
              "--fork", str(K), "--fa", str(vFA), "--a-start", str(vFA + 1), "--a-tip", str(vAtip), "--fb", str(vFB),
              "--b-start", str(vFB + 1), "--b-tip", str(vBtip), "--fm", str(vFM), "--merge", str(vFM + NB),
              "--answers", os.path.join(tmp, name + ".ans"), "--json", aj])
        oj = json.load(open(aj))
        # with stride 1 every version is sampled, so the sampled SHA-256 covers all answers
        assert oj["sampled_answers"] == (receipt["versions"] - 1) * q, oj
        _gz(base, os.path.join(OUT, name + ".base.gz"))
        _gz(spec, os.path.join(OUT, name + ".spec.gz"))
        cases[name] = {"kind": "branch", "states": 2049, "q": q, "receipt": receipt,
                       "oracle": oj["sampled_sha256"],
                       "oracle_checks": {k: v for k, v in oj.items() if "merge" in k or "tip" in k or "bad" in k}}
    with open(os.path.join(HERE, "spec_cases.json"), "w") as f:
        json.dump(cases, f, indent=1, sort_keys=True)
        f.write("\n")
    print(json.dumps({k: v["oracle"] for k, v in cases.items()}, indent=1))


if __name__ == "__main__":
    main()
