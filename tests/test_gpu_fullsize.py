"""The benchmark workloads at full size: the product racon CLI (GPU hot path)
must print byte-for-byte the FASTA that the CPU oracle CLI printed for the
same seeded inputs. The oracle's sha256 per case is committed in
tests/golden/fullsize_sha256.json (tests/golden/make_fullsize_hashes.py,
a CPU run); the inputs are regenerated here from their seeds.
"""
import hashlib
import json
import os
import subprocess
import time

import pytest

from fullsize_cases import CASES, inputs
from racon_amd import _lib

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SIMGEN = os.path.join(ROOT, "racon_amd", "bin", "simgen")
GOLDEN = os.path.join(ROOT, "tests", "golden", "fullsize_sha256.json")


def _golden():
    with open(GOLDEN) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def workdir(tmp_path_factory):
    return str(tmp_path_factory.mktemp("fullsize"))


@pytest.mark.timeout(900)
@pytest.mark.parametrize("name", sorted(CASES))
def test_fullsize_matches_oracle_hash(name, workdir):
    golden = _golden()
    if name not in golden:
        pytest.skip(f"no committed oracle hash for {name}")
    t0 = time.time()
    files = inputs(name, workdir, SIMGEN)
    t1 = time.time()
    cmd = [_lib.RACON_BIN, "-t", "16"] + CASES[name]["args"] + files
    p = subprocess.run(cmd, capture_output=True, timeout=880)
    t2 = time.time()
    assert p.returncode == 0, p.stderr.decode()[-2000:]
    digest = hashlib.sha256(p.stdout).hexdigest()
    print(f"{name}: inputs {t1 - t0:.1f} s, racon {t2 - t1:.1f} s, {len(p.stdout)} B, "
          f"{p.stdout.count(b'>')} sequences, sha256 {digest}", flush=True)
    assert len(p.stdout) == golden[name]["bytes"]
    assert p.stdout.count(b">") == golden[name]["sequences"]
    assert digest == golden[name]["sha256"]
