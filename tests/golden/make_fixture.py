"""Writes synthetic BWBASE02 fixtures for the reference-native parity cases.

The format is the reference's (src/common.hpp:57-75,112-155): a 64-byte header
("BWBASE02", n, groups, unit, original_lines, missing_values, t_min, t_max, amount_scale),
n 24-byte rows (t i64, amount i64, g i32, pad i32), G i64 base factors, G i32 labels.
Row values come from a splitmix64 stream of this project's own definition (uniform t over
[0, span), uniform group, amount uniform in [1, 100000]); base factors are 1000.
"""
import argparse
import struct

import numpy as np

MASK = (1 << 64) - 1


def _splitmix(seed, n):
    idx = np.arange(1, n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed & MASK) + idx * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def make_fixture(path, rows, groups, span, seed):
    d = _splitmix(seed, 3 * rows).reshape(rows, 3)
    t = (d[:, 0] % np.uint64(span)).astype(np.int64)
    g = (d[:, 1] % np.uint64(groups)).astype(np.int32)
    a = (d[:, 2] % np.uint64(100000)).astype(np.int64) + 1
    rec = np.zeros(rows, dtype=[("t", "<i8"), ("a", "<i8"), ("g", "<i4"), ("p", "<i4")])
    rec["t"], rec["a"], rec["g"] = t, a, g
    hdr = struct.pack("<8sqiiqqqqq", b"BWBASE02", rows, groups, 0, rows, 0, int(t.min()), int(t.max()), 2)
    with open(path, "wb") as f:
        f.write(hdr)
        f.write(rec.tobytes())
        f.write(np.full(groups, 1000, dtype="<i8").tobytes())
        f.write(np.arange(groups, dtype="<i4").tobytes())


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--rows", type=int, default=2000)
    ap.add_argument("--groups", type=int, default=4)
    ap.add_argument("--span", type=int, default=86_400_000_000)
    ap.add_argument("--seed", type=int, default=20261131)
    a = ap.parse_args()
    make_fixture(a.out, a.rows, a.groups, a.span, a.seed)
