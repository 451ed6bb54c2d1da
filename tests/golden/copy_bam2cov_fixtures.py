"""Copies the BAM2COV golden tables of breseq's own tests (tests/bam2cov, tests/bam2cov_csv,
tests/bam2cov_per_read_group) and the reference sequence they print bases from
(tests/data/bull/bull_1.fasta) into tests/golden/bam2cov/.  Their input BAMs are not in the
reference tree, so the tests rebuild alignments that reproduce each table's coverage columns.

    python tests/golden/copy_bam2cov_fixtures.py /root/reference
"""
import json
import os
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "bam2cov")

FILES = [
    ("tests/bam2cov/expected.tab", "expected.tab"),
    ("tests/bam2cov_csv/expected.csv", "expected.csv"),
    ("tests/bam2cov_per_read_group/expected.no_read_groups.tab", "expected.no_read_groups.tab"),
    ("tests/bam2cov_per_read_group/expected.multiple_read_groups.tab", "expected.multiple_read_groups.tab"),
    ("tests/data/bull/bull_1.fasta", "bull_1.fasta"),
]


def main(ref_root):
    os.makedirs(OUT, exist_ok=True)
    for src, dst in FILES:
        shutil.copyfile(os.path.join(ref_root, src), os.path.join(OUT, dst))
    # --show-average prints references.reference[seq_id].coverage_average of summary.json
    with open(os.path.join(ref_root, "tests/bam2cov/summary.json")) as f:
        summary = json.load(f)
    avg = summary["references"]["reference"]["rachael"]["coverage_average"]
    with open(os.path.join(OUT, "coverage_average.json"), "w") as f:
        json.dump({"rachael": avg}, f)
        f.write("\n")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
