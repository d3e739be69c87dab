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
