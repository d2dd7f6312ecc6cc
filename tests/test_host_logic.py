"""Host-side logic of the library that needs no GPU, compiled with g++ from
the library's own headers: the split-and-retry rule of the batched form and
the request server (csrc/split_retry.h)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "quantum-harness-peps-naive_amd", "csrc")


def test_memory_shortage_splits_and_other_failures_do_not(tmp_path):
    exe = tmp_path / "split_retry_check"
    src = os.path.join(os.path.dirname(__file__), "native", "split_retry_check.cpp")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", CSRC, "-o", str(exe), src],
                   check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "split_retry: ok" in r.stdout
