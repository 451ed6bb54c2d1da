"""The oracle fixtures of the full-size workloads (tests/golden/oracle_*.json.gz) describe the
workloads they claim to, and exercise the evidence kinds the full-size parity tests require."""
import os

import pytest

from breseq_amd import workloads
from full_size import GOLDEN, fixture, required_cases


@pytest.mark.parametrize("name", ["C2", "C3"])
def test_fixture_matches_its_workload(name):
    if not os.path.exists(os.path.join(GOLDEN, "oracle_%s.json.gz" % name)):
        pytest.fail("fixture of %s missing: run tests/golden/make_oracle_fixtures.py %s" % (name, name))
    fx = fixture(name)
    assert fx["config"] == getattr(workloads, name)
    assert fx["mode"] == workloads.MODES[name]
    req = required_cases(fx["gd"], polymorphism=(fx["mode"] == "polymorphism"))
    assert req["MC_with_ranges"] >= 1 and req["RA_insert_slot"] >= 1
    if fx["mode"] == "polymorphism":
        assert req["RA_interior_bounds"] >= 1
    # the cutoff the oracle identified with is its own fit's: MC runs close above it
    cut = fx["fit"]["deletion_coverage_propagation_cutoff"]
    assert cut >= 1
    assert fx["dp"].startswith("seq_id,start,end,strand,orientation")
