"""bench.py's host logic (no GPU): the one-GPU guard, the warm-up length and
the staleness rule for the HBM traffic read from a committed profile."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused_before_any_gpu_work():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr and r.stdout == ""


@pytest.mark.parametrize("requested, expect", [(0, 16), (5, 16), (16, 16), (17, 32), (40, 48)])
def test_warmup_covers_whole_captured_graphs(requested, expect):
    assert bench.warmup_steps(requested) == expect


def test_traffic_read_only_from_a_summary_of_the_same_build(tmp_path, monkeypatch):
    prof = tmp_path / "profiles"
    prof.mkdir()
    (prof / "k.md").write_text("text\nhbm_bytes_per_launch: 12345\nkernel_ms: 2.0\n")
    (prof / "nodur.md").write_text("hbm_bytes_per_launch: 12345\n")
    monkeypatch.setattr(bench, "ROOT", str(tmp_path))
    assert bench.traffic_from_profile("k.md", 2.1) == 12345      # within 10 %
    assert bench.traffic_from_profile("k.md", 1.7) is None       # another build of the kernel
    assert bench.traffic_from_profile("nodur.md", 2.0) is None   # no recorded duration
    assert bench.traffic_from_profile("missing.md", 2.0) is None


def test_committed_summaries_carry_traffic_and_duration():
    for name in ("step_kernel_khi512_fp32_fast_hbm.md", "step_kernel_khi512_fp64_fast_hbm.md",
                 "ens_step_kernel_ensemble256_fp32_fast_hbm.md"):
        txt = open(os.path.join(ROOT, "profiles", name)).read()
        assert "hbm_bytes_per_launch:" in txt and "kernel_ms:" in txt, name
