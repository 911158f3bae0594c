"""GPU parity of the device FNV-1a (bw_fnv1a_device, csrc/hash_kernels.hip): for byte strings that
end inside a word, a block and a group of blocks, over one to three composition levels, on an
aligned and a misaligned (4-byte) start, the device value equals the sequential FNV-1a of the same
bytes (tests/host/fnv_check.cpp --file, compiled here).  Integer work: bit-exact."""
import ctypes
import json
import os
import subprocess

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
START = 1469598103934665603  # the reference's FNV-1a start value


@pytest.fixture(scope="module")
def fnv_check(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("fnvg") / "fnv_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I" + os.path.join(REPO, "biworlddb-artifact_amd", "csrc", "host"),
                    "-o", out, os.path.join(REPO, "tests", "host", "fnv_check.cpp")], check=True, capture_output=True,
                   timeout=300)
    return out


@pytest.fixture(scope="module")
def hip(pkg):
    lib = pkg.load_gpu_lib()
    rt = ctypes.CDLL("libamdhip64.so")
    lib.bw_fnv1a_workspace.argtypes = [ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
    lib.bw_fnv1a_device.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint64, ctypes.c_void_p,
                                    ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p]
    rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    rt.hipFree.argtypes = [ctypes.c_void_p]
    rt.hipDeviceSynchronize.argtypes = []
    return lib, rt


def _device_fnv(hip, data, offset):
    lib, rt = hip
    n = len(data)
    buf = ctypes.c_void_p()
    assert rt.hipMalloc(ctypes.byref(buf), n + 16) == 0
    work = ctypes.c_void_p()
    res = ctypes.c_void_p()
    try:
        src = np.frombuffer(data, dtype=np.uint8)
        if n:
            assert rt.hipMemcpy(ctypes.c_void_p(buf.value + offset), src.ctypes.data, n, 1) == 0
        wb = ctypes.c_size_t()
        assert lib.bw_fnv1a_workspace(n, ctypes.byref(wb)) == 0
        assert rt.hipMalloc(ctypes.byref(work), wb.value) == 0
        assert rt.hipMalloc(ctypes.byref(res), 8) == 0
        rc = lib.bw_fnv1a_device(ctypes.c_void_p(buf.value + offset), n, START, work, wb.value, res, None)
        assert rc == 0, rc
        out = np.zeros(1, dtype=np.uint64)
        assert rt.hipMemcpy(out.ctypes.data, res, 8, 2) == 0
        return "%016x" % int(out[0])
    finally:
        for p in (buf, work, res):
            if p.value:
                rt.hipFree(p)


@pytest.mark.gpu
@pytest.mark.parametrize("n,offset", [(0, 0), (1, 0), (5, 4), (4095, 0), (4097, 4), (300001, 0), (300001, 4),
                                      (17000003, 0), (76800000, 0)])
def test_device_fnv_equals_sequential(hip, fnv_check, tmp_path, n, offset):
    data = np.random.default_rng(n + offset).integers(0, 256, size=n, dtype=np.uint8).tobytes()
    path = tmp_path / "b.bin"
    path.write_bytes(data)
    want = json.loads(subprocess.run([fnv_check, "--file", str(path)], capture_output=True, text=True, check=True,
                                     timeout=300).stdout)["sequential"]
    assert _device_fnv(hip, data, offset) == want, (n, offset)
