"""breseq_amd -- MI355X-native pileup-scoring engine for breseq (error_count, coverage
distribution, identify_mutations, test_RA_evidence), behind a C ABI (include/bsq.h)."""
from .api import (  # noqa: F401
    CONSENSUS,
    POLYMORPHISM,
    BsqError,
    ErrorCountResult,
    Sample,
    Session,
    calculate_trims,
    error_count,
    error_rates_text_to_table,
    fit_coverage_distribution,
    identify_mutations,
    mode_settings,
    ra_test_settings,
    test_RA_evidence,
)
