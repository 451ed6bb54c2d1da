"""The error-rate table's values without their text: g6_round_trip (breseq_amd/csrc/host_util.cpp)
against strtod of printf's %g text, which is what the reference's error_rates.tab round trip gives
(error_count.cpp:1005-1026 writes the table with six significant digits, identify_mutations reads it
back).  Covers exact ties at the sixth digit, the neighbours of powers of ten, the table's own
log-ratio values, random bit patterns and the values outside the direct route's range."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_round_trip_matches_text(tmp_path):
    exe = str(tmp_path / "g6_round_trip_check")
    csrc = os.path.join(ROOT, "breseq_amd", "csrc")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-o", exe, os.path.join(ROOT, "tests", "native", "g6_round_trip_check.cpp"),
                           os.path.join(csrc, "host_util.cpp"), "-lpthread"])
    for seed in (1, 2):
        r = subprocess.run([exe, "200000", str(seed)], capture_output=True, text=True)
        assert r.returncode == 0, r.stdout + r.stderr
