"""The exact decimal formatter shared by the host and the device GD writer
(breseq_amd/csrc/dec_format.h) against std::to_chars (printf's %.*e / %.*f, correctly rounded),
compiled for the CPU from the same header: boundary values, exact ties at the printed precision,
subnormals, the largest double and random bit patterns."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_formatter_matches_to_chars(tmp_path):
    exe = str(tmp_path / "dec_format_check")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-o", exe, os.path.join(ROOT, "tests", "native", "dec_format_check.cpp")])
    for seed, count, prec in ((1, 400000, 20), (2, 100000, 60)):
        r = subprocess.run([exe, str(count), str(seed), str(prec)], capture_output=True, text=True)
        assert r.returncode == 0, r.stdout + r.stderr
