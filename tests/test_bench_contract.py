"""bench.py's output contract: one JSON line with the metric, the throughput
and the roofline / CPU-baseline objects; more than one GPU is refused with a
message (this build has no multi-GPU path)."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BENCH = os.path.join(ROOT, "bench.py")


def test_more_than_one_gpu_is_refused():
    # exits before the library or any GPU is touched
    r = subprocess.run([sys.executable, BENCH, "--gpus", "2"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


@pytest.mark.gpu
def test_json_line_small_grid():
    print("case bench.py --n 64", flush=True)
    r = subprocess.run([sys.executable, BENCH, "--n", "64", "--steps", "2", "--warmup", "1", "--no-cpu",
                        "--no-ensemble", "--no-fft-compare", "--reps", "1"],
                       capture_output=True, text=True, timeout=240)
    print(r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    lines = [x for x in r.stdout.splitlines() if x.strip()]
    assert len(lines) == 1
    d = json.loads(lines[0])
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better",
              "scaling", "vs_baseline", "dtype", "data", "config", "roofline", "cpu_baseline"):
        assert k in d, k
    assert d["n_gpus"] == 1 and d["steps"] == 2 and d["warmup"] == 1 and d["unit"] == "Mpts/s"
    assert d["higher_is_better"] is True and d["dtype"] == "f32"
    # Mpts/s = n^3 / seconds per step
    assert abs(d["value"] - 64 ** 3 / (d["ms_per_step"] * 1e-3) / 1e6) <= 1e-3 * d["value"] + 1e-3
    assert d["vs_baseline"] is None  # a published figure exists for 512^3 only
    assert "workload" in d["config"]
    rf = d["roofline"]
    for k in ("bound", "achieved", "peak", "unit", "frac", "traffic"):
        assert k in rf, k
    assert rf["bound"] == "hbm" and rf["unit"] == "GB/s" and rf["peak"] == 8000.0
    assert abs(rf["frac"] - rf["achieved"] / rf["peak"]) < 1e-3
    assert d["cpu_baseline"] is None  # --no-cpu
    # the FP64 leg ran and its error against FP32 is small at this size
    assert d["fp64"]["ms_per_step"] > 0 and d["dissipation_rel_err_fp32_vs_fp64"] < 1e-4
    # the GPU times of the timed steps (events around each graph replay) sit
    # below the host clock over them, and the layout report names a layout
    assert 0 < d["ms_per_step_gpu_median"] <= 1.5 * d["ms_per_step"]
    assert 0 < d["fp64"]["ms_per_step_gpu_median"] <= 1.5 * d["fp64"]["ms_per_step"]
    assert rf["work_layout"]["layout"] in ("line", "transposed", "ky-blocked")
