"""The autotuner's choices across processes (SPH_TUNE_CACHE): the first process
that creates a solver of a given (precision, n) times the variants and appends
one line to the cache file; the next process finds it there, times nothing,
and reports the same choice and timings; a line whose variants do not exist
in this build is ignored (the next process times again and appends a valid
line).
"""
import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r"""
import json, sys
sys.path.insert(0, sys.argv[1])
import spectral_solver_amd as sph
s = sph.TaylorGreenSolver(64, 1600.0, "fp32")
print(json.dumps({"tuning": s.tuning(), "layout": s.layout()}))
s.close()
"""


def child(env):
    r = subprocess.run([sys.executable, "-c", CHILD, ROOT], capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout.strip().splitlines()[-1])


def test_choices_persist_across_processes(tmp_path):
    path = str(tmp_path / "tune.txt")
    env = dict(os.environ, SPH_TUNE_CACHE=path)
    env.pop("SPH_TILE_VARIANT", None)
    env.pop("SPH_AUTOTUNE", None)
    print("case tune cache: first process", flush=True)
    first = child(env)
    lines = open(path).read().splitlines()
    assert len(lines) == 1 and lines[0].startswith("sph-tune-2 ")
    key = " ".join(lines[0].split()[:7])
    print("case tune cache: second process", flush=True)
    second = child(env)
    assert open(path).read().splitlines() == lines  # found, nothing timed or appended
    assert second["layout"][0] == first["layout"][0]
    for k, (v, l, ms) in first["tuning"].items():
        v2, l2, ms2 = second["tuning"][k]
        assert (v2, l2) == (v, l)
        assert all(abs(a - b) <= 1e-5 * abs(a) for a, b in zip(ms, ms2))  # stored with 6 digits
    # an entry with a variant this build does not have is ignored
    fields = lines[0].split()
    fields[8] = "7"  # variant of the inverse y pass
    with open(path, "a") as f:
        f.write(" ".join(fields) + "\n")
    print("case tune cache: invalid entry", flush=True)
    third = child(env)
    lines3 = open(path).read().splitlines()
    assert len(lines3) == 3 and lines3[2].startswith(key + " ")
    assert third["tuning"]["inverse_y"][0] < 3
