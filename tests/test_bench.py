"""bench.py's output contract: one JSON line with the fields the harness reads
(metric, value, unit, steps, roofline, cpu_baseline, ...), and the refusal of
runs on more than one GPU."""
import json
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def run_bench(*args, timeout=600):
    return subprocess.run([sys.executable, str(ROOT / "bench.py"), *args], capture_output=True, text=True,
                          timeout=timeout, cwd=ROOT)


def test_bench_refuses_more_than_one_gpu():
    p = run_bench("--gpus", "2")
    assert p.returncode != 0
    msg = json.loads(p.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]


def test_traffic_table_keys_and_profiles():
    """profiles/traffic.json: every key is a kernel description reduced by
    bench.traffic_key (so a description that differs only in wave shape, item
    count or projection form finds it), every entry's bytes follow from its
    FETCH_SIZE / WRITE_SIZE and the copy calibration, and its profile exists."""
    sys.path.insert(0, str(ROOT))
    import bench
    t = json.loads((ROOT / "profiles" / "traffic.json").read_text())
    assert abs(t["fetch_scale"] - 2.0) < 1e-3 and t["write_scale"] == 1.0
    for key, e in t["kernels"].items():
        assert bench.traffic_key(key) == key, key
        want = 1024 * (e["fetch_kib"] * t["fetch_scale"] + e["write_kib"] * t["write_scale"])
        assert abs(e["hbm_bytes"] - want) <= 1e-6 * want, key
        assert (ROOT / "profiles" / e["profile"]).exists(), e["profile"]
    desc = ("model=DNKG n=1024 copies=16384 kernel=wave lanes_per_chain=64 sites_per_lane=16 exact=1 "
            "force=cubic_unit")
    assert bench.traffic_bytes(desc) == t["kernels"][bench.traffic_key(desc)]["hbm_bytes"]
    assert bench.traffic_key(desc + " projection=mfma_wide_f64") == bench.traffic_key(desc)


@pytest.mark.gpu
def test_bench_json_contract(gpu):
    p = run_bench("--steps", "3", "--warmup", "1", "--cpu-seconds", "1")
    assert p.returncode == 0, p.stderr[-2000:]
    line = p.stdout.strip().splitlines()[-1]
    d = json.loads(line)
    for key in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better", "scaling",
                "vs_baseline", "dtype", "data", "config", "roofline", "cpu_baseline"):
        assert key in d, key
    assert d["n_gpus"] == 1 and d["steps"] == 3 and d["warmup"] == 1 and d["higher_is_better"] is True
    assert d["unit"] == "site-steps/s" and d["value"] > 1e11 and d["dtype"] == "f64"
    assert "workload" in d["config"] and "model" not in d["config"]
    r = d["roofline"]
    assert r["bound"] in ("hbm", "mfma") and r["unit"] in ("GB/s", "TFLOP/s")
    assert 0 < r["frac"] <= 1 and abs(r["frac"] - r["achieved"] / r["peak"]) < 1e-9
    assert r["traffic"] is None or r["traffic"] > 0
    c = d["cpu_baseline"]
    assert c["kind"] in ("reference", "port") and c["value"] > 0 and c["sample"]
    assert 1 <= c["cores"] <= c["host_cores"]
    bt = d["batched"]  # the batched form of the same workload, same process
    assert bt["value"] > 1e11 and bt["vs_plain"] > 0 and "wave shapes" in bt["autotune"]
    assert bt["energy_drift"]["over_steps"] == 4000 and bt["energy_drift"]["nonfinite_chains"] == 0


@pytest.mark.gpu
def test_bench_dnls_fp32c_line(gpu):
    """--precision fp32c: the FP32-compensated DNLS rate, its roofline against
    the FP32 peak, and its norm / energy error next to an FP64 run."""
    p = run_bench("--workload", "dnls", "--precision", "fp32c", "--steps", "2", "--warmup", "1", "--batch", "50",
                  "--cpu-seconds", "0")
    assert p.returncode == 0, p.stderr[-2000:]
    d = json.loads(p.stdout.strip().splitlines()[-1])
    assert d["dtype"] == "f32" and "fp32c" in d["config"]["kernel"] and d["roofline"]["peak"] == 157.3
    pr = d["precision"]
    assert pr["over_steps"] == 150 and pr["fp64"]["value"] > 0 and d["vs_fp64"] > 0
    assert pr["norm_drift"]["max"] < 1e-6 and pr["fp64"]["norm_drift"]["max"] < 1e-12
    assert pr["linenergies_vs_fp64"] < 1e-3


@pytest.mark.gpu
def test_bench_toda_ragged_fp32c_line(gpu):
    """--precision fp32c on the ragged Toda workload: the FP32-compensated
    persistent launch, its rate next to the FP64 one and its energy drift."""
    p = run_bench("--workload", "toda_ragged", "--precision", "fp32c", "--steps", "2", "--warmup", "1", "--batch",
                  "20", "--cpu-seconds", "0")
    assert p.returncode == 0, p.stderr[-2000:]
    d = json.loads(p.stdout.strip().splitlines()[-1])
    assert d["dtype"] == "f32" and "precision=fp32c" in d["config"]["kernel"]
    assert "ragged_persistent" in d["config"]["kernel"] or "per_class" in d["config"]["kernel"]
    pr = d["precision"]
    assert pr["over_steps"] == 60 and pr["fp64"]["value"] > 0 and d["vs_fp64"] > 0 and pr["norm_drift"] is None
    assert pr["energy_drift"]["mean"] < 1e-5 and pr["fp64"]["energy_drift"]["mean"] < 1e-5


@pytest.mark.gpu
def test_bench_c1_pipelined(gpu):
    p = run_bench("--workload", "c1", "--steps", "20", "--warmup", "2", "--cpu-seconds", "1")
    assert p.returncode == 0, p.stderr[-2000:]
    d = json.loads(p.stdout.strip().splitlines()[-1])
    assert d["us_per_batch"] > 0 and d["config"]["mode"] == "batched" and "queued" in d["config"]["workload"]
    assert d["cpu_baseline"]["value"] > 0 and d["energy_drift"]["over_steps"] == 2200
