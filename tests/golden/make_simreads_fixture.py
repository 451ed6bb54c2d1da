"""Writes tests/golden/simreads_lambda.json.gz from the reference's own SIMULATE-READS test
(tests/simulate_reads_paired: lambda.gbk, paired-end, 25 pairs of 100 bp, fragments N(350, 40),
seed 1): the first sequence of the GenBank file, its topology, the command's options, and the two
expected FASTQ files.  Run in a container that holds the reference tree:
    python tests/golden/make_simreads_fixture.py /root/reference"""
import gzip
import json
import os
import sys


def genbank_first_sequence(path):
    """(sequence, circular) of the first record: the ORIGIN lines without digits and spaces,
    upper-cased; circular when the LOCUS line says so."""
    seq, circular, in_origin = [], False, False
    with open(path) as f:
        for line in f:
            if line.startswith("LOCUS"):
                circular = " circular " in line.lower()
            elif line.startswith("ORIGIN"):
                in_origin = True
            elif line.startswith("//"):
                if in_origin:
                    break
            elif in_origin:
                seq.append("".join(c for c in line if c.isalpha()))
    return "".join(seq).upper(), circular


def main(root):
    t = os.path.join(root, "tests")
    seq, circular = genbank_first_sequence(os.path.join(t, "data", "lambda", "lambda.gbk"))
    out = {
        "reference": seq,
        "circular": circular,
        "options": {"mode": "paired-end", "n_reads": 25, "read_length": 100, "fragment_mean": 350,
                    "fragment_stdev": 40, "seed": 1},
        "expected_1": open(os.path.join(t, "simulate_reads_paired", "expected.reads_1.fastq")).read(),
        "expected_2": open(os.path.join(t, "simulate_reads_paired", "expected.reads_2.fastq")).read(),
    }
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "simreads_lambda.json.gz")
    with gzip.open(dst, "wt") as f:
        json.dump(out, f)
    print(dst, len(seq), circular)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
