"""GPU parity of the CSR rank-LUT remap (bwt_rank_remap_launch, csrc/tree_kernels.hip) called through
the C-ABI: out[i] = #keys < in[i] over a sorted set of distinct 32-bit keys, the rank convention of
the reference's sparse engine (src/sparse_engine.hpp:62-97).  Three paths: large remaps with the
LUT's row offsets staged in LDS, the same when 64 consecutive rows hold more keys than a 16-bit
offset reaches (the workgroup falls back to global row offsets), and small remaps (global).  The
expected ranks are numpy's searchsorted on the same keys.  Integer work: bit-exact."""
import ctypes

import numpy as np
import pytest


class RankArgs(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("inp", ctypes.c_void_p), ("out", ctypes.c_void_p),
                ("row_ptr", ctypes.c_void_p), ("low16", ctypes.c_void_p)]


@pytest.fixture(scope="module")
def hip(pkg):
    lib = pkg.load_gpu_lib()
    rt = ctypes.CDLL("libamdhip64.so")
    lib.bwt_rank_remap_launch.argtypes = [ctypes.POINTER(RankArgs), ctypes.c_void_p]
    rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    rt.hipFree.argtypes = [ctypes.c_void_p]
    rt.hipDeviceSynchronize.argtypes = []
    return lib, rt


def _csr(keys):
    hi = (keys >> 16).astype(np.int64)
    row_ptr = np.searchsorted(hi, np.arange(65537, dtype=np.int64), side="left").astype(np.uint32)
    return row_ptr, (keys & 0xFFFF).astype(np.uint16)


def _remap(hip, keys, vals):
    lib, rt = hip
    row_ptr, low16 = _csr(keys)
    bufs = []

    def up(a):
        p = ctypes.c_void_p()
        assert rt.hipMalloc(ctypes.byref(p), max(a.nbytes, 4)) == 0
        bufs.append(p)
        assert rt.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0
        return p

    try:
        d_in, d_rp, d_lo = up(vals), up(row_ptr), up(low16)
        d_out = ctypes.c_void_p()
        assert rt.hipMalloc(ctypes.byref(d_out), vals.nbytes) == 0
        bufs.append(d_out)
        args = RankArgs(len(vals), d_in, d_out, d_rp, d_lo)
        print("remap", len(keys), "keys", len(vals), "endpoints", flush=True)
        assert lib.bwt_rank_remap_launch(ctypes.byref(args), None) == 0
        assert rt.hipDeviceSynchronize() == 0
        out = np.empty(len(vals), dtype=np.uint32)
        assert rt.hipMemcpy(out.ctypes.data, d_out, out.nbytes, 2) == 0
        return out
    finally:
        for p in bufs:
            rt.hipFree(p)


def _endpoints(rng, keys, n):
    """Uniform 32-bit values, the keys themselves, their neighbours, and the extremes."""
    v = rng.integers(0, 1 << 32, size=n, dtype=np.uint64)
    k = keys[rng.integers(0, len(keys), size=n // 4)].astype(np.uint64)
    v[: n // 4] = k
    v[n // 4: n // 2] = np.minimum(k + 1, (1 << 32) - 1)
    v[:4] = [0, (1 << 32) - 1, keys[0], keys[-1]]
    return v.astype(np.uint32)


@pytest.mark.gpu
@pytest.mark.parametrize("layout,n", [("uniform", 1 << 21), ("clustered", 1 << 21), ("uniform", 1000)])
def test_rank_remap_matches_searchsorted(hip, layout, n):
    rng = np.random.default_rng(20261204 + n)
    if layout == "uniform":  # config 3's key set: 4,194,304 distinct uniform 32-bit keys
        keys = np.unique(rng.integers(0, 1 << 32, size=4_300_000, dtype=np.uint64))[:4_194_304]
    else:  # 200,000 keys under the first 64 rows: their offsets do not fit 16 bits
        dense = rng.choice(1 << 22, size=200_000, replace=False).astype(np.uint64)
        rest = rng.integers(1 << 22, 1 << 32, size=100_000, dtype=np.uint64)
        keys = np.unique(np.concatenate([dense, rest]))
    keys = keys.astype(np.uint32)
    vals = _endpoints(rng, keys, n)
    got = _remap(hip, keys, vals)
    want = np.searchsorted(keys, vals, side="left").astype(np.uint32)
    bad = np.flatnonzero(got != want)
    assert bad.size == 0, (layout, n, bad[:5], got[bad[:5]], want[bad[:5]])
