"""The generated workloads follow the laws the benchmark configs state (CPU only).

The product's generator and the oracle's are byte-identical (tests/test_oracle.py); this checks
what neither restates: that their output has the stated statistics.  Each case dumps one workload
from the oracle (`bw_oracle tree --dump`: parent[1..S], keys, deltas, lo, hi) and measures it:
  config 1  parent uniform over the 32 most recent states; keys uniform in [0, N); deltas uniform
            in [-1000, 1000]; queries [min(a,b), max(a,b)+1) of two uniform draws
  config 2  parent = previous with probability 0.9, else uniform over the prior 1,024; Zipf(1.1)
            keys; deltas uniform in [-10^6, 10^6]; range lengths log-uniform in [1, N]
  config 3  Zipf(0.9) keys over the distinct set; deltas in [-10^4, 10^4]
  config 4  chains of geometric length (mean 64, at most 4,096), else config 2's parent law
Seeds are fixed, so the outcomes are deterministic; tolerances are several standard errors of the
sample sizes used.  Zipf exponents come from the rank-frequency slope of the observed keys (no
assumption about how ranks map to keys).
"""
import subprocess

import numpy as np
from scipy import stats


def dump(oracle_bin, tmp_path, cfg, S, U, Q, N, seed):
    path = str(tmp_path / f"w{cfg}.bin")
    subprocess.run([oracle_bin, "tree", "--cfg", str(cfg), "--S", str(S), "--U", str(U), "--Q", str(Q), "--N", str(N),
                    "--seed", str(seed), "--dump", path, "--workload-sha"], check=True, capture_output=True,
                   timeout=600)
    raw = open(path, "rb").read()
    hdr = np.frombuffer(raw[:24], dtype=np.int64)
    assert tuple(hdr) == (S, U, Q)
    o = 24
    parts = {}
    for name, dt, n in (("parent", np.int32, S), ("key", np.uint32, S * U), ("delta", np.int64, S * U),
                        ("lo", np.uint32, S * Q), ("hi", np.uint32, S * Q)):
        nb = n * np.dtype(dt).itemsize
        parts[name] = np.frombuffer(raw[o:o + nb], dtype=dt)
        o += nb
    assert o == len(raw)
    return parts


def zipf_slope(keys, ranks=200):
    """Least-squares slope of log(count) on log(rank) over the most frequent keys."""
    counts = np.sort(np.unique(keys, return_counts=True)[1])[::-1][:ranks].astype(np.float64)
    r = np.arange(1, len(counts) + 1, dtype=np.float64)
    return np.polyfit(np.log(r), np.log(counts), 1)[0]


def uniform_p(x, lo, hi, seed=7):
    """p-value of "x is uniform over the integers lo..hi": x - lo plus an independent U[0, 1)
    jitter, divided by the count, is exactly U[0, 1) under that law (KS test)."""
    assert x.min() >= lo and x.max() <= hi
    jit = np.random.default_rng(seed).random(len(x))
    u = (x.astype(np.float64) - lo + jit) / (float(hi) - lo + 1)
    return stats.kstest(u, "uniform").pvalue


def test_config1_laws(oracle_bin, tmp_path):
    S, U, Q, N = 1000, 64, 16, 65536
    w = dump(oracle_bin, tmp_path, 1, S, U, Q, N, 20261204)
    v = np.arange(1, S + 1)
    back = v - w["parent"]
    assert back.min() >= 1 and back.max() <= 32
    # beyond the first 32 states the distance back is uniform over 1..32
    hist = np.bincount(back[32:], minlength=33)[1:]
    assert stats.chisquare(hist).pvalue > 1e-4
    assert uniform_p(w["key"], 0, N - 1) > 1e-4
    assert uniform_p(w["delta"], -1000, 1000) > 1e-4
    assert (w["lo"] < w["hi"]).all() and w["hi"].max() <= N
    # hi - lo - 1 = |a - b| of two uniform draws: mean (N^2 - 1) / 3N, sd ~ 0.236 N
    gap = w["hi"].astype(np.int64) - w["lo"].astype(np.int64) - 1
    assert abs(gap.mean() - (N * N - 1) / (3 * N)) < 0.01 * N


def test_config2_laws(oracle_bin, tmp_path):
    S, U, Q, N = 300000, 8, 8, 1 << 20
    w = dump(oracle_bin, tmp_path, 2, S, U, Q, N, 20261205)
    v = np.arange(1, S + 1)
    back = v - w["parent"]
    assert back.min() >= 1 and back.max() <= 1024
    p_prev = np.mean(back == 1)
    # P(parent = v-1) = 0.9 + 0.1 / 1024; standard error of the mean ~ 0.00055
    assert abs(p_prev - (0.9 + 0.1 / 1024)) < 0.003, p_prev
    far = back[(back > 1) & (v > 1024)]
    hist = np.bincount(far, minlength=1025)[2:]
    assert stats.chisquare(hist).pvalue > 1e-4  # the jumps are uniform over the prior 1,024
    assert abs(zipf_slope(w["key"]) + 1.1) < 0.05
    assert uniform_p(w["delta"], -10**6, 10**6) > 1e-4
    length = w["hi"].astype(np.int64) - w["lo"].astype(np.int64)
    assert length.min() >= 1 and length.max() <= N and w["hi"].max() <= N
    # log-uniform lengths: len = floor(X), X = N^u, so P(len <= L) = log(L + 1) / log(N) exactly;
    # 2.4M samples give a pointwise standard error <= 0.00033
    grid = np.unique(np.logspace(0, np.log10(N - 2), 40).astype(np.int64))
    emp = np.searchsorted(np.sort(length), grid, side="right") / len(length)
    assert np.max(np.abs(emp - np.log(grid + 1) / np.log(N))) < 0.002
    # the range's start is uniform over the N - len + 1 positions that fit it
    jit = np.random.default_rng(11).random(len(length))
    pos = (w["lo"].astype(np.float64) + jit) / (N - length + 1)
    assert stats.kstest(pos, "uniform").pvalue > 1e-4


def test_config3_laws(oracle_bin, tmp_path):
    S, U, Q, D = 20000, 64, 8, 65536
    w = dump(oracle_bin, tmp_path, 3, S, U, Q, D, 99)
    assert abs(zipf_slope(w["key"]) + 0.9) < 0.05
    assert len(np.unique(w["key"])) <= D
    assert uniform_p(w["delta"], -10**4, 10**4) > 1e-4


def test_config4_laws(oracle_bin, tmp_path):
    S, U, Q, N = 300000, 8, 4, 1 << 20
    w = dump(oracle_bin, tmp_path, 4, S, U, Q, N, 20261207)
    v = np.arange(1, S + 1)
    chain = w["parent"] == v - 1
    # runs: maximal stretches of states whose parent is the previous state, plus their head
    starts = np.flatnonzero(~chain)
    ends = np.append(starts[1:], S)
    runs = ends - starts
    assert runs.max() <= 4096
    # geometric with continue probability 63/64: mean 64 (the 4,096 cap removes ~e^-64 of it);
    # ~4,700 runs give a standard error of the mean ~ 0.93
    assert abs(runs.mean() - 64.0) < 4.0, runs.mean()
    assert abs(np.median(runs) - np.log(2) / -np.log(63 / 64)) < 5.0
    # a jump's distance back is uniform over the prior 1,024 (as config 2)
    back = (v - w["parent"])[~chain & (v > 1024)]
    assert back.min() >= 1 and back.max() <= 1024
    # (a jump of distance 1 reads as a continued chain, so compare over 2..1024)
    assert stats.chisquare(np.bincount(back, minlength=1025)[2:]).pvalue > 1e-4
    assert abs(zipf_slope(w["key"]) + 1.1) < 0.05
