"""Host check of the tree kernel's Hines round schedule: tests/sched_check.cpp replays the
schedule block exactly as the kernel decodes it and compares, bit for bit, with the serial
depth-first elimination (eden/Eden.cpp:4416-4486, 6114-6225) on random trees of several
shapes, for every workgroup size, in both fold-task formats (two-word tasks, and the
four-word merged tasks of constant trees replayed as the kernel runs them: diagonal
pre-sweep, then right-hand sides with fixed ratios)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_schedule_replay_matches_serial_elimination(tmp_path):
    exe = tmp_path / "sched_check"
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-I", ROOT,
                    os.path.join(ROOT, "tests", "sched_check.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe), "200"], check=True, capture_output=True, text=True).stdout
    assert out.strip() == "ok 2400", out
