"""Extract test vectors from the reference's own golden files (tests/<name>/expected.gd).

Writes tests/golden/reference_ra_entries.json: RA evidence lines, reduced to the fields that
test_RA_evidence and the Fisher strand test read, with the mode of the run that produced them.
Each line is data from the reference's test fixtures (inputs and expected outputs), not code.

Run in a container where the reference is mounted:
    python tests/golden/make_reference_vectors.py /root/reference
"""
import glob
import json
import os
import random
import sys

KEEP = {"score", "frequency", "frequency_lower", "frequency_upper", "major_base", "minor_base", "major_cov",
        "minor_cov", "new_cov", "ref_cov", "total_cov", "major_frequency", "allele_frequencies",
        "fisher_strand_p_value", "ks_quality_p_value", "prediction", "reject", "consensus_reject",
        "polymorphism_reject", "user_defined"}

# options that change test_RA_evidence's settings or the evidence set; runs using them are skipped
SKIP_OPTIONS = ("--polymorphism-frequency-cutoff", "--polymorphism-reject", "--consensus-", "--polymorphism-minimum",
                "--user-evidence-gd", "--mask-gd", "--polymorphism-no-indels", "--polymorphism-bias")


def main(ref_root, out_path, per_file_cap=120, seed=7):
    rng = random.Random(seed)
    entries = []
    for path in sorted(glob.glob(os.path.join(ref_root, "tests", "*", "expected.gd"))):
        name = os.path.basename(os.path.dirname(path))
        lines = open(path, encoding="utf-8", errors="replace").read().splitlines()
        cmd = next((l for l in lines if l.startswith("#=COMMAND")), None)
        if cmd is None:
            continue
        if any(opt in cmd for opt in SKIP_OPTIONS):
            continue
        toks = cmd.split()
        poly = ("--polymorphism-prediction" in toks) or ("-p" in toks)
        ras = [l for l in lines if l.startswith("RA\t")]
        if len(ras) > per_file_cap:
            ras = rng.sample(ras, per_file_cap)
        for l in ras:
            f = l.split("\t")
            pos = [x for x in f[3:] if "=" not in x]
            kv = dict(x.split("=", 1) for x in f[3:] if "=" in x)
            kv = {k: v for k, v in kv.items() if k in KEEP}
            entries.append({"test": name, "polymorphism": poly, "id": f[1], "spec": pos[:5], "fields": kv})
    with open(out_path, "w") as fh:
        json.dump(entries, fh, indent=0, sort_keys=True)
    print("wrote %d RA entries to %s" % (len(entries), out_path))


if __name__ == "__main__":
    root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    here = os.path.dirname(os.path.abspath(__file__))
    main(root, os.path.join(here, "reference_ra_entries.json"))
