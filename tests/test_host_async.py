"""The host's persistent async threads (breseq_amd/csrc/host_util.cpp run_async), which run the
coverage fit and the DP candidate regions beside the device passes: results and exceptions reach the
futures, tasks that wait on each other all run, and WaitOnExit waits before a scope's locals go."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_async_runner(tmp_path):
    exe = str(tmp_path / "async_check")
    csrc = os.path.join(ROOT, "breseq_amd", "csrc")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-o", exe, os.path.join(ROOT, "tests", "native", "async_check.cpp"),
                           os.path.join(csrc, "host_util.cpp"), "-lpthread"])
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
