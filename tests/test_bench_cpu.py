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


def test_traffic_record_applies_only_to_the_kernels_it_was_measured_on(tmp_path, monkeypatch):
    f = tmp_path / "traffic.json"
    cur = bench.kernel_sources_sha()
    f.write_text(json.dumps({"a": {"kernel_sources_sha16": cur, "bytes_per_launch": 5, "l2_hit_rate": 0.5},
                             "b": {"kernel_sources_sha16": "0" * 16, "bytes_per_launch": 7}}))
    monkeypatch.setattr(bench, "TRAFFIC_FILE", f)
    assert bench.traffic_for("a")["bytes_per_launch"] == 5
    assert bench.traffic_for("b") == {"stale": "0" * 16}
    assert bench.traffic_for("no_such_workload") == {}
    stale = bench.roofline_record(bench.traffic_for("b"), None, None, 9000.0, 1e9, 0.5)
    assert stale["traffic"] is None and "no traffic record" in stale["basis"]
    ok = bench.roofline_record(bench.traffic_for("a"), 4e9, 4000.0, 9000.0, 1e9, 0.5)
    assert ok["frac"] == 0.5 and ok["achieved"] == 4000.0 and ok["algorithmic_gbs"] == 9000.0


def test_the_committed_traffic_record_is_complete():
    rec = json.loads(bench.TRAFFIC_FILE.read_text())
    for name, r in rec.items():
        assert len(r["kernel_sources_sha16"]) == 16
        assert r["bytes_per_launch"] > 0 and 0.0 < r["l2_hit_rate"] < 1.0


@pytest.mark.parametrize("name", bench.WORKLOADS[:3])
def test_single_frame_workloads_build(name):
    frames, desc = bench.build_workload(name)
    assert len(frames) == 1 and desc
    assert frames[0]["cam"]["width"] in (256, 1280, 1920)


def _write_csv(path, header, rows):
    import csv
    path.parent.mkdir(parents=True, exist_ok=True)
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(header)
        w.writerows(rows)


def test_record_profile_turns_counter_passes_into_a_stamped_record(tmp_path):
    """Synthetic rocprofv3 output: two extend dispatches per counter run with
    known counters, a copy calibration with factors 2 (fetch) and 1 (write),
    and a trace run whose bench line names the current kernel sources."""
    out = tmp_path / "prof"
    hdr = ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"]
    ext = "void gtk::(anonymous namespace)::k_extend<false, false, false>(gtk::Frame, gtk::Batch, gtk::Queues, unsigned int)"
    cnt = "void gtk::(anonymous namespace)::k_extend<true, false, false>(gtk::Frame, gtk::Batch, gtk::Queues, unsigned int)"
    sq = {"SQ_WAVE_CYCLES": 1000, "SQ_WAIT_ANY": 600, "SQ_ACTIVE_INST_VALU": 100, "SQ_THREAD_CYCLES_VALU": 3200}
    _write_csv(out / "sq_w" / "run_counter_collection.csv", hdr,
               [[i, k, c, v] for i, k in enumerate([ext, ext, cnt]) for c, v in sq.items()])
    _write_csv(out / "fetch_w" / "run_counter_collection.csv", hdr,
               [[0, ext, "FETCH_SIZE", 1000], [1, ext, "FETCH_SIZE", 3000], [2, cnt, "FETCH_SIZE", 99999]])
    _write_csv(out / "write_w" / "run_counter_collection.csv", hdr,
               [[i, k, c, v] for i, k in enumerate([ext, ext, cnt])
                for c, v in (("WRITE_SIZE", 500), ("TCC_HIT_sum", 30), ("TCC_MISS_sum", 70))])
    neg = "void at::native::vectorized_elementwise_kernel<4, at::native::neg_kernel_cuda>(int)"
    _write_csv(out / "cal_fetch" / "run_counter_collection.csv", hdr, [[0, neg, "FETCH_SIZE", 512 * 1024]])
    _write_csv(out / "cal_write" / "run_counter_collection.csv", hdr, [[0, neg, "WRITE_SIZE", 1024 * 1024]])
    # trace: 3 renders (statistics, warm-up, one timed step) of 2 extend launches, then a counting render
    trace = [[i, ext if i < 6 else cnt, 1000 * i, 1000 * i + (100 if i < 4 else 300)] for i in range(8)]
    _write_csv(out / "trace_w" / "run_kernel_trace.csv", ["Dispatch_Id", "Kernel_Name", "Start_Timestamp",
                                                          "End_Timestamp"], trace)
    _write_csv(out / "trace_w" / "run_kernel_stats.csv", ["Name", "Calls", "TotalDurationNs", "AverageNs",
                                                          "Percentage"], [[ext, 6, 1000, 166.7, 50.0]])
    line = {"metric": "Mrays/s", "value": 1.0, "ms_per_step": 1.0, "steps": 1, "warmup": 1, "config": {"workload": "w"},
            "roofline": {"avg_launch_ms": 0.0003, "kernel_sources_sha16": bench.kernel_sources_sha()}}
    (out / "trace_w.log").write_text("step 0\n" + json.dumps(line) + "\n")
    dest = tmp_path / "profiles"
    dest.mkdir()
    assert bench.record_profile(out, "w", dest=dest) == 0
    r = json.loads((dest / "extend_hbm_traffic.json").read_text())["w"]
    assert r["kernel_sources_sha16"] == bench.kernel_sources_sha()
    assert (r["fetch_factor"], r["write_factor"]) == (2.0, 1.0)
    # mean FETCH 2000 KiB x 2 + WRITE 500 KiB x 1, the counting dispatch left out
    assert r["bytes_per_launch"] == (2000 * 2 + 500) * 1024
    assert r["l2_hit_rate"] == 0.3 and r["wave_wait_frac"] == 0.6 and r["valu_busy_frac"] == 0.1
    assert r["valu_lane_util"] == 0.5
    assert r["rocprof_avg_launch_us"] == 0.3  # the timed step's two launches (300 ns each)
    assert "k_extend<false, false, false>" in (dest / "w_counters.md").read_text()
    # a set measured on other kernel sources is refused
    line["roofline"]["kernel_sources_sha16"] = "0" * 16
    (out / "trace_w.log").write_text(json.dumps(line) + "\n")
    with pytest.raises(SystemExit):
        bench.record_profile(out, "w", dest=dest)


def test_limiter_reads_the_counters():
    base = {"traffic": 1, "frac": 0.37, "wave_wait_frac": 0.63, "valu_busy_frac": 0.08}
    assert bench.limiter(base).startswith("latency of dependent memory fetches")
    assert bench.limiter(dict(base, frac=0.8)) == "HBM bandwidth"
    assert bench.limiter(dict(base, traffic=None)) is None
    assert bench.limiter(dict(base, wave_wait_frac=0.2)).startswith("mixed")

