"""Prints the layout (sizeof, every field's offsetof) of the drop-in POD structs of a header set.

The field lists are read from this repository's include/gpu_v7.h, gpu_v7b.h and gpu_session.h;
a C++ program is generated that includes the headers of the directory given and prints the
layout as JSON.  Run against the reference's own headers (src/gpu_v7.h:7-47, src/gpu_v7b.h:6-48,
src/gpu_session.h:4-20) it must print the same layout: the drop-in contract is byte identity.

  python abi_layout.py INCLUDE_DIR OUT_JSON
"""
import json
import os
import re
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
HEADERS = ("gpu_v7.h", "gpu_v7b.h", "gpu_session.h")


def struct_fields(path):
    """{struct name: [field names]} from the struct bodies of one of our headers."""
    txt = open(path).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    txt = re.sub(r"//[^\n]*", "", txt)
    out = {}
    for name, body in re.findall(r"struct\s+(\w+)\s*\{(.*?)\};", txt, flags=re.S):
        fields = []
        for decl in body.split(";"):
            decl = decl.strip()
            if not decl:
                continue
            # drop the type: everything up to the first declarator; declarators are comma separated
            m = re.match(r"^(?:const\s+)?(?:struct\s+)?\w+\s*\**\s*(.*)$", decl, flags=re.S)
            for d in m.group(1).split(","):
                d = d.strip().lstrip("*").strip()
                fields.append(re.sub(r"\[.*\]", "", d))
        out[name] = fields
    return out


def layout(include_dir, cxx="g++"):
    structs = {}
    for h in HEADERS:
        structs.update(struct_fields(os.path.join(REPO, "include", h)))
    lines = ["#include <cstddef>", "#include <cstdint>", "#include <cstdio>"]
    lines += [f'#include "{h}"' for h in HEADERS]
    lines.append("int main() {")
    lines.append('  std::printf("{");')
    first = True
    for s, fields in structs.items():
        sep = "" if first else ","
        first = False
        lines.append(f'  std::printf("{sep}\\"{s}\\":{{\\"size\\":%zu,\\"fields\\":{{", sizeof({s}));')
        for i, f in enumerate(fields):
            fs = "" if i == 0 else ","
            lines.append(f'  std::printf("{fs}\\"{f}\\":%zu", offsetof({s}, {f}));')
        lines.append('  std::printf("}}");')
    lines.append('  std::printf("}\\n");')
    lines.append("  return 0;\n}")
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "abi.cpp"), os.path.join(d, "abi")
        with open(src, "w") as f:
            f.write("\n".join(lines))
        subprocess.run([cxx, "-std=c++17", "-w", "-I", include_dir, src, "-o", exe], check=True, capture_output=True)
        return json.loads(subprocess.run([exe], check=True, capture_output=True, text=True).stdout)


if __name__ == "__main__":
    res = layout(sys.argv[1])
    with open(sys.argv[2], "w") as f:
        json.dump(res, f, indent=1, sort_keys=True)
        f.write("\n")
