"""Host logic of bench.py: the kernel-source hash that ties the committed
counter records under profiles/ to the device code they were measured on."""
import json
import os
import shutil
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def _copy_sources(tmp_path):
    d = tmp_path / "csrc"
    d.mkdir()
    for f in bench.KERNEL_SOURCES:
        shutil.copy(os.path.join(ROOT, "trace_amd", "csrc", f), d / f)
    return str(d)


def test_hash_ignores_comments_and_whitespace(tmp_path):
    d = _copy_sources(tmp_path)
    h0 = bench.kernel_source_hash(d)
    p = os.path.join(d, "kernels.hip")
    text = open(p, encoding="utf-8").read()
    edited = ("// a new line comment\n/* and a block\n   comment */\n"
              + text.replace("\n", "\n\n", 5) + "\n// trailing comment\n")
    open(p, "w", encoding="utf-8").write(edited)
    assert bench.kernel_source_hash(d) == h0


def test_hash_follows_code(tmp_path):
    d = _copy_sources(tmp_path)
    h0 = bench.kernel_source_hash(d)
    p = os.path.join(d, "traverse.h")
    text = open(p, encoding="utf-8").read()
    assert "constexpr int kWindowMargin = 1;" in text
    open(p, "w", encoding="utf-8").write(
        text.replace("constexpr int kWindowMargin = 1;", "constexpr int kWindowMargin = 2;"))
    assert bench.kernel_source_hash(d) != h0


def test_host_sources_do_not_enter_the_hash(tmp_path):
    d = _copy_sources(tmp_path)
    h0 = bench.kernel_source_hash(d)
    (tmp_path / "csrc" / "capi.hip").write_text("int host_only_change = 1;\n")
    assert bench.kernel_source_hash(d) == h0


def test_committed_records_match_the_current_kernels():
    """Every traffic record names the current device code, so bench.py
    reports the measured HBM traffic instead of null."""
    recs = json.load(open(os.path.join(ROOT, "profiles", "traffic.json")))["records"]
    assert recs
    h = bench.kernel_source_hash()
    assert all(r["source_sha"] == h for r in recs)


def test_counter_records_match_the_current_kernels():
    recs = json.load(open(os.path.join(ROOT, "profiles", "counters.json")))["records"]
    assert {r["kernel"] for r in recs} >= {"k_forward_staged", "k_bp_tile", "k_update_em"}
    h = bench.kernel_source_hash()
    assert all(r["source_sha"] == h and r["limiter"] for r in recs)


def test_traffic_records_are_consistent():
    """The size-resolved request counts reproduce the copy kernel's bytes,
    and every record's bytes agree with FETCH_SIZE and WRITE_SIZE scaled by
    the copy's factors (all of these kernels read whole 128-byte lines)."""
    t = json.load(open(os.path.join(ROOT, "profiles", "traffic.json")))
    cal = t["calibration"]
    assert abs(cal["request_sized_read_error"]) < 1e-3
    assert abs(cal["request_sized_write_error"]) < 1e-3
    for r in t["records"]:
        assert r["bytes_per_launch"] == r["read_bytes"] + r["write_bytes"]
        assert abs(r["bytes_per_launch"] / r["fetch_write_calibrated_bytes"] - 1) < 0.02, r["kernel"]


def test_more_than_one_gpu_is_refused():
    """bench.py measures one GPU: --gpus N > 1 exits with a message (before
    anything touches a device)."""
    import subprocess
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_host_description():
    threads, info = bench.host_cpus()[0], bench.cpu_model()
    assert threads >= 1 and isinstance(info, str) and info
