import importlib.util
import json
import os
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(REPO, "oracle", "bw_oracle")
REF_DIR = os.path.join(REPO, "oracle", "_ref")
GOLDEN = os.path.join(REPO, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def load_pkg():
    spec = importlib.util.spec_from_file_location(
        "bwgpu_pkg", os.path.join(REPO, "biworlddb-artifact_amd", "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def pkg():
    return load_pkg()


@pytest.fixture(scope="session")
def oracle_bin():
    if not os.path.exists(ORACLE):
        subprocess.run(["make", "-C", os.path.join(REPO, "oracle"), "bw_oracle"], check=True,
                       stdout=subprocess.DEVNULL)
    return ORACLE


def run_json(cmd, timeout=600, env=None):
    """Runs a worker, prints the command first (so a failing GPU step names its case)."""
    print("RUN", " ".join(cmd), flush=True)
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout, env=env)
    if p.returncode != 0:
        raise AssertionError(f"{cmd[0]} exited {p.returncode}: {p.stderr[-2000:]}")
    lines = [ln for ln in p.stdout.splitlines() if ln.startswith("{")]
    return json.loads(lines[-1])


def golden(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def fixture_file(params):
    """BWBASE02 fixture regenerated from its parameters (deterministic; cached, not committed)."""
    sys.path.insert(0, GOLDEN)
    from make_golden import fixture_path

    return fixture_path(params)


def spec_files(case, tmp_path):
    """(base fixture, spec) of a committed BWHHIST1 case, decompressed into tmp_path."""
    import gzip
    import shutil

    out = []
    for part in ("base", "spec"):
        dst = os.path.join(str(tmp_path), f"{case}.{part}")
        with gzip.open(os.path.join(GOLDEN, "specs", f"{case}.{part}.gz"), "rb") as fi, open(dst, "wb") as fo:
            shutil.copyfileobj(fi, fo)
        out.append(dst)
    return out
