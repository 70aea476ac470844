"""The CPU side of bench.py: the --gpus guard, the CPU-baseline record (all
cores and one core, both from the oracle) and the measured-traffic lookup.
No GPU is touched."""
import json
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))

import bench  # noqa: E402
from gartus_amd import scenes  # noqa: E402


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr) or "1 GPU" in (r.stdout + r.stderr)


def test_cpu_baseline_record_has_the_contract_fields():
    data, targets, cam = scenes.cornell()
    cam = dict(cam, width=24, grid=2)
    frames = [dict(data=data, lights=None, env=None, env_rgb=None, cam=cam, targets=targets)]
    rec = bench.cpu_baseline(frames, 0.5)
    for k in ("value", "unit", "cores", "kind", "sample", "cpu_model", "one_core"):
        assert k in rec
    assert rec["unit"] == "Mrays/s" and rec["kind"] == "port" and rec["cores"] >= 1
    assert rec["value"] > 0 and rec["one_core"]["value"] > 0
    assert rec["one_core"]["unit"] == "Mrays/s" and rec["one_core"]["samples_per_s"] > 0
    assert isinstance(rec["cpu_model"], str) and rec["cpu_model"]
    json.dumps(rec)  # one JSON line


def test_traffic_record_of_the_default_workload():
    rec = bench.traffic_for("ggx_env_1080p")
    assert rec["bytes_per_launch"] > 0
    assert 0.0 < rec["l2_hit_rate"] < 1.0
    assert bench.traffic_for("no_such_workload") == {}


@pytest.mark.parametrize("name", bench.WORKLOADS[:3])
def test_single_frame_workloads_build(name):
    frames, desc = bench.build_workload(name)
    assert len(frames) == 1 and desc
    assert frames[0]["cam"]["width"] in (256, 1280, 1920)
