"""Checks against vectors taken from the reference's own golden files (tests/<name>/expected.gd),
committed as tests/golden/reference_ra_entries.json by tests/golden/make_reference_vectors.py.

* fisher_strand_p_value of every RA entry is a function of its printed major/minor coverage:
  pins the oracle's Fisher test to the reference's printed digits.
* prediction / reject / consensus_reject / polymorphism_reject are what test_RA_evidence wrote from
  the printed score, bounds and coverages: pins the product's classifier (host code of libbsq.so).
* every printed floating point field re-formats to the same text: pins the number formatting.
"""
import ctypes as C
import json
import os

import pytest

import breseq_amd as B
import oracle_lib as O
from breseq_amd import _lib as L

HERE = os.path.dirname(os.path.abspath(__file__))
ENTRIES = json.load(open(os.path.join(HERE, "golden", "reference_ra_entries.json")))
CLASS_FIELDS = ("prediction", "reject", "consensus_reject", "polymorphism_reject")


def test_fixture_is_populated():
    assert len(ENTRIES) > 500
    assert any(e["polymorphism"] for e in ENTRIES) and any(not e["polymorphism"] for e in ENTRIES)


def test_fisher_strand_p_values_match_reference_digits():
    checked = exact = 0
    for e in ENTRIES:
        f = e["fields"]
        if "fisher_strand_p_value" not in f or "major_cov" not in f or "minor_cov" not in f:
            continue
        mj = [int(x) for x in f["major_cov"].split("/")]
        mn = [int(x) for x in f["minor_cov"].split("/")]
        p = O.fisher(mn[0], mn[1], mj[0], mj[1])
        text = "%.5e" % p
        want = f["fisher_strand_p_value"]
        checked += 1
        if text == want:
            exact += 1
        else:
            # libm ulps may move the last printed digit only
            assert abs(p - float(want)) <= 1.01e-5 * abs(float(want)), (e["test"], e["spec"], text, want)
    assert checked > 500
    assert exact >= 0.99 * checked, (exact, checked)


def _gd_line(e, drop):
    f = {k: v for k, v in e["fields"].items() if k not in drop}
    parts = ["RA", e["id"], "."] + e["spec"] + ["%s=%s" % (k, f[k]) for k in sorted(f)]
    return "\t".join(parts)


@pytest.mark.parametrize("poly", [False, True])
def test_ra_classification_matches_reference(poly):
    entries = [e for e in ENTRIES if e["polymorphism"] == poly and "user_defined" not in e["fields"]]
    assert entries
    got = {}
    # evidence IDs are only unique within one reference test, so each test is its own GD file
    for test in sorted({e["test"] for e in entries}):
        mine = [e for e in entries if e["test"] == test]
        text = "#=GENOME_DIFF\t1.0\n" + "\n".join(_gd_line(e, CLASS_FIELDS) for e in mine) + "\n"
        out = B.test_RA_evidence(text, polymorphism_prediction=poly)
        for ln in out.splitlines():
            if not ln.startswith("RA\t"):
                continue
            f = ln.split("\t")
            got[(test, f[1])] = dict(x.split("=", 1) for x in f[3:] if "=" in x)
    mismatches = []
    for e in entries:
        g = got.get((e["test"], e["id"]))
        if g is None:
            mismatches.append((e["test"], e["id"], "deleted"))
            continue
        for k in CLASS_FIELDS:
            if e["fields"].get(k) != g.get(k):
                mismatches.append((e["test"], e["id"], k, e["fields"].get(k), g.get(k)))
    assert not mismatches, mismatches[:20]


def test_number_formatting_round_trips_reference_fields():
    lib = L.lib()
    buf = C.create_string_buffer(128)
    n = 0
    for e in ENTRIES:
        for k, v in e["fields"].items():
            if k in ("frequency", "frequency_lower", "frequency_upper", "major_frequency"):
                places, sci = len(v.split("e")[0].split(".")[1]), 1
            elif k in ("fisher_strand_p_value", "ks_quality_p_value"):
                places, sci = 5, 1
            elif k == "score":
                places, sci = 1, 0
            else:
                continue
            val = float("nan") if v == "NA" else float(v)
            assert lib.bsq_format_double(val, places, sci, buf, 128) > 0
            assert buf.value.decode() == v, (k, v, buf.value)
            n += 1
    assert n > 1000
