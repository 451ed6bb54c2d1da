"""Shared helpers of the parity tests: synthetic samples, oracle runs, GD comparison."""
import numpy as np

import breseq_amd as B
import oracle_lib as O

# Small versions of the benchmark configurations (sizes the oracle finishes in seconds).
SMALL_CONFIGS = {
    # consensus mode, single-end 100 bp, 20 SNPs + indels, uniform reference (C1-shaped)
    "se100": dict(seed=11, ref_length=48502, gc=0.5, read_length=100, coverage=60.0, paired=False,
                  qual_mean=32, qual_sd=6, qual_min=2, qual_max=41, n_snps=20, n_indels=6, indel_max=5,
                  redundant_fraction=0.02, mapq=42),
    # paired 150 bp (two read files, R1/R2 error models), GC 50.8% (C2-shaped, smaller genome)
    "pe150": dict(seed=12, ref_length=60000, gc=0.508, read_length=150, coverage=80.0, paired=True,
                  insert_mean=350, insert_sd=50, qual_mean=34, qual_sd=5, qual_min=2, qual_max=41,
                  n_snps=30, n_indels=10, indel_max=5, redundant_fraction=0.01, mapq=42),
    # polymorphism mode, mixed-frequency variants (C3-shaped)
    "poly": dict(seed=13, ref_length=30000, gc=0.508, read_length=150, coverage=200.0, paired=True,
                 insert_mean=350, insert_sd=50, qual_mean=33, qual_sd=6, qual_min=2, qual_max=41,
                 n_snps=40, n_indels=10, indel_max=5, variant_freq_min=0.05, variant_freq_max=0.5,
                 redundant_fraction=0.01, mapq=42),
}


def make_sample(name):
    cfg = dict(SMALL_CONFIGS[name])
    return B.Sample.simulate(**cfg)


def oracle_error_count(sample, **kw):
    args = B.api.make_error_count_args(sample, **kw)
    return args, O.error_count(args)


def oracle_identify(sample, table, S, Q, prop, seed=None, mode=B.CONSENSUS, **kw):
    args = B.api.make_identify_args(sample, table, S, Q, prop, seed, mode=mode, **kw)
    return O.identify_mutations(args)


def parse_gd(text):
    """GD text -> list of (type, positional fields tuple, dict of key=value)."""
    out = []
    for ln in text.splitlines():
        if not ln or ln.startswith("#"):
            continue
        f = ln.split("\t")
        pos = []
        kv = {}
        for x in f[3:]:
            if "=" in x:
                k, v = x.split("=", 1)
                kv[k] = v
            else:
                pos.append(x)
        out.append((f[0], f[1], tuple(pos), kv))
    return out


FLOAT_FIELDS = ("score", "frequency", "major_frequency", "frequency_lower", "frequency_upper",
                "ks_quality_p_value", "fisher_strand_p_value")


def last_place(txt):
    """Value of one unit in the last printed digit of a %.Nf / %.Ne string."""
    m = txt.lower()
    if "e" in m:
        mant, ex = m.split("e")
        dec = len(mant.split(".")[1]) if "." in mant else 0
        return 10.0 ** (int(ex) - dec)
    dec = len(m.split(".")[1]) if "." in m else 0
    return 10.0 ** (-dec)


def compare_gd(gpu_text, ref_text):
    """Compare two GD texts: identical entry sets, ids, positional fields and integer fields; a
    printed floating point field may differ by at most one unit in its last printed digit (libm
    ulp differences landing on a rounding boundary).
    Returns (n_entries, n_identical_lines, max_abs_err per field)."""
    g = parse_gd(gpu_text)
    r = parse_gd(ref_text)
    assert len(g) == len(r), "entry counts differ: gpu %d oracle %d\nGPU:\n%s\nORACLE:\n%s" % (
        len(g), len(r), gpu_text[:4000], ref_text[:4000])
    errs = {}
    identical = 0
    for (gt, gid, gp, gk), (rt, rid, rp, rk) in zip(g, r):
        assert (gt, gid, gp) == (rt, rid, rp), "entry differs: %s %s %s vs %s %s %s" % (gt, gid, gp, rt, rid, rp)
        assert set(gk) == set(rk), "field sets differ at %s %s: %s vs %s" % (gt, gp, sorted(gk), sorted(rk))
        same = True
        for k in gk:
            if gk[k] == rk[k]:
                continue
            same = False
            if k in FLOAT_FIELDS:
                a, b = float(gk[k]) if gk[k] != "NA" else np.nan, float(rk[k]) if rk[k] != "NA" else np.nan
                assert np.isnan(a) == np.isnan(b), (k, gk[k], rk[k])
                if not np.isnan(a):
                    err = abs(a - b)
                    assert err <= 1.01 * last_place(rk[k]), (
                        "field %s differs beyond one printed unit at %s %s: %s vs %s" % (k, gt, gp, gk[k], rk[k]))
                    errs[k] = max(errs.get(k, 0.0), err)
            elif k == "allele_frequencies":
                ga = dict(x.split(":") for x in gk[k].split(","))
                ra = dict(x.split(":") for x in rk[k].split(","))
                assert set(ga) == set(ra), (gk[k], rk[k])
                for base in ga:
                    assert abs(float(ga[base]) - float(ra[base])) <= 1.01 * last_place(ra[base]), (gk[k], rk[k])
            else:
                raise AssertionError("field %s differs at %s %s: %s vs %s" % (k, gt, gp, gk[k], rk[k]))
        identical += same
    return len(g), identical, errs
