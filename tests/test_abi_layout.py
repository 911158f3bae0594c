"""The drop-in structs keep the reference's byte layout (CPU only).

include/gpu_v7.h, gpu_v7b.h and gpu_session.h must lay out every struct exactly as the reference's
src/gpu_v7.h:7-47, src/gpu_v7b.h:6-48 and src/gpu_session.h:4-20 do, since bw_panel and the
reference's drivers (src/panel_batch_gpu*.cpp, src/panel_gpu_session.cpp) pass them by pointer
across the C-ABI.  tests/golden/abi_layout.json holds the layout (sizeof and every field's
offsetof, x86-64 g++); it is checked against our headers always and against the reference's own
headers where the reference sources are present.
"""
import json
import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(HERE, "host"))
from abi_layout import layout, struct_fields  # noqa: E402

GOLDEN = os.path.join(HERE, "golden", "abi_layout.json")
REF_SRC = "/root/reference/src"


def golden():
    with open(GOLDEN) as f:
        return json.load(f)


def test_field_lists_parsed():
    fields = {}
    for h in ("gpu_v7.h", "gpu_v7b.h", "gpu_session.h"):
        fields.update(struct_fields(os.path.join(REPO, "include", h)))
    assert len(fields) == 12
    assert fields["GpuV7Node"][:4] == ["kind", "g_old", "c_old", "ti_old"]
    assert fields["GpuSessRunArgs"][-1] == "threads"


def test_our_headers_match_golden():
    assert layout(os.path.join(REPO, "include")) == golden()


def test_known_sizes():
    # sizes stated by the reference's own comments and its worker's packing (src/gpu_v7b.h:6
    # "40 bytes"; GpuV7Node packed at src/panel_batch_gpu.cpp:86-97), independent of the printer
    g = golden()
    assert g["GpuV7bNode"]["size"] == 40 and g["GpuSessNode"]["size"] == 40
    assert g["GpuV7Node"]["size"] == 72
    assert g["GpuV7bNode"]["fields"]["g_old"] == 2  # int16
    assert g["GpuV7Args"]["size"] == 152 and g["GpuSessRunArgs"]["size"] == 168


@pytest.mark.skipif(not os.path.isdir(REF_SRC), reason="reference sources not present")
def test_reference_headers_match_golden():
    assert layout(REF_SRC) == golden()
