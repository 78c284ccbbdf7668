"""The CPU oracle and the host producer built with AddressSanitizer and
UndefinedBehaviorSanitizer (g++, host code only) and run over seeded frames
(six option variants, all three tile edges, the learned-cull outputs, the
depth target), the PLY fixture, every 7th truncation of it, byte
corruptions of its header and the slot packer."""
from __future__ import annotations

import os
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="g++ is missing")


def test_oracle_and_producer_under_asan_ubsan(tmp_path):
    exe = tmp_path / "driver"
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-pthread", "-ffp-contract=off", "-fno-fast-math",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
           "-fno-omit-frame-pointer", f"-I{ROOT / 'oracle'}", f"-I{ROOT / 'include'}",
           f"-I{ROOT / '3dgs_tile_webgpu_amd' / 'csrc'}",
           str(ROOT / "tests" / "sanitize" / "driver.cpp"), str(ROOT / "oracle" / "gs_oracle.cpp"),
           str(ROOT / "3dgs_tile_webgpu_amd" / "csrc" / "gs_producer.cpp"), "-o", str(exe)]
    build = subprocess.run(cmd, capture_output=True, text=True)
    assert build.returncode == 0, build.stderr[-4000:]
    env = dict(os.environ, ASAN_OPTIONS="halt_on_error=1:detect_leaks=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    env.pop("LD_PRELOAD", None)
    run = subprocess.run([str(exe), str(ROOT / "tests" / "golden" / "scene_64.ply")],
                         capture_output=True, text=True, env=env, timeout=600)
    assert run.returncode == 0, (run.stdout + run.stderr)[-6000:]
    assert "0 failures" in run.stdout
