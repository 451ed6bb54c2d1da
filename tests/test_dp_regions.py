"""Discordant-pair (DP) candidate regions, DP_candidate_regions.csv (identify_mutations.cpp:883-908,
1395-1457, 2358-2422, 2550-2573).

The oracle restates breseq's per-column deques; the product (bsq_dp_candidate_regions, also run
inside identify_mutations) visits only the columns where a window count changes.  Both are pinned
by hand-computed regions: window exits at exactly start + width (the >= test), the open-time
snapshot, reads appended while a region is open (also on the column that closes it), the flush
at L+1, names without ':', negative insert sizes for R2, redundancy majority, excluded flags and
the always-open regions of seed 0.  Randomized alignment sets then compare product and oracle."""
import numpy as np
import pytest

import breseq_amd as B
import oracle_lib as O
from aln_builder import Built, BuiltSample

FLAG_R1, FLAG_R2 = 0x41, 0x81          # paired + first / second mate, not a proper pair
REV = 0x10


def _read(ref, pos, flag, xp, name, tlen, cigar="30M", x1=-1, rg=-1, tid=0):
    seq, rp = [], pos
    for ln, op in __import__("re").findall(r"(\d+)([MD])", cigar):
        ln = int(ln)
        if op == "M":
            seq.append(ref[rp:rp + ln])
        rp += ln
    return dict(tid=tid, pos=pos, cigar=cigar, seq="".join(seq), flag=flag, xp=xp, name=name, tlen=tlen, x1=x1,
                rg=rg)


def _refs(n=2, length=1000, seed=5):
    rng = np.random.default_rng(seed)
    return [("t%d" % t, "".join(rng.choice(list("ACGT"), size=length))) for t in range(n)]


def kat_sample():
    refs = _refs()
    r0, r1 = refs[0][1], refs[1][1]
    reads = [
        # t0, bin F/FR: regions open at 151 (snapshot A, B), C appended at 301, closes at 351
        _read(r0, 100, FLAG_R1, "FR", "1:7", 5000),                       # A
        _read(r0, 150, FLAG_R1, "FR", "1:8", 5100),                       # B
        _read(r0, 150, FLAG_R1 | 0x100, "FR", "1:90", 1),                 # secondary: never counted
        _read(r0, 160, FLAG_R1 | 0x2, "FR", "1:91", 1),                   # proper pair
        _read(r0, 170, FLAG_R1 | 0x8, "FR", "1:92", 1),                   # mate unmapped
        _read(r0, 175, 0x40, "FR", "1:93", 1),                            # not paired
        _read(r0, 300, FLAG_R2, "FR", "2:9", -7000),                      # C (R2: insert size negated)
        # t0, bin R/RF: D and E enter on the opening column, F right after; 2 of 3 redundant
        _read(r0, 600, FLAG_R1 | REV, "RF", "1:20", -300, x1=2),          # D
        _read(r0, 600, FLAG_R1 | REV, "RF", "1:21", 400),                 # E
        _read(r0, 601, FLAG_R1 | REV, "RF", "1:22", 0, x1=3),             # F
        # t0, bin F/FF: names without ':'; still open at the end of the target (flush at L+1)
        _read(r0, 900, FLAG_R1, "FF", "q5", 10),                          # G
        _read(r0, 949, FLAG_R2, "FF", "q6", 20),                          # H
        # t1, bin R/FR: K enters on the column that closes the region and is still listed
        _read(r1, 100, FLAG_R1 | REV, "FR", "1:30", 1, tid=1),            # I
        _read(r1, 100, FLAG_R1 | REV, "FR", "1:31", 2, tid=1),            # J
        _read(r1, 300, FLAG_R1 | REV, "FR", "1:32", 3, cigar="10M5D20M", tid=1),   # K
    ]
    return refs, reads


KAT_SEED2 = """seq_id,start,end,strand,orientation,length,max_discordant_count,redundant,discordant_pairs
t0,151,350,F,FR,200,2,0,1:7__2:7__5000__101__130;1:8__2:8__5100__151__180;1:9__2:9__7000__301__330
t0,601,800,R,RF,200,3,1,1:20__2:20__-300__601__630;1:21__2:21__400__601__630;1:22__2:22__0__602__631
t0,950,1000,F,FF,51,2,0,q5____10__901__930;__q6__-20__950__979
t1,101,300,R,FR,200,2,0,1:30__2:30__1__101__130;1:31__2:31__2__101__130;1:32__2:32__3__301__335
"""

# seed 0: every bin of a target is one region 1..L; bins without reads list nothing
KAT_SEED0_ONE_READ = """seq_id,start,end,strand,orientation,length,max_discordant_count,redundant,discordant_pairs
t0,1,1000,F,FF,1000,0,0,
t0,1,1000,F,FR,1000,1,0,1:7__2:7__5000__101__130
t0,1,1000,F,RF,1000,0,0,
t0,1,1000,R,FF,1000,0,0,
t0,1,1000,R,FR,1000,0,0,
t0,1,1000,R,RF,1000,0,0,
"""


def _oracle_dp(s, dp, prop=None, call_targets=None):
    args = B.api.make_error_count_args(s)
    _, _, l_o, texts = O.error_count(args)
    table = O.read_log10_prob_table(texts["error_rates"], len(l_o))
    prop = prop if prop is not None else [3.0] * s.n_targets
    iargs = B.api.make_identify_args(s, table, args.read_set_max, args.quality_max, prop, dp=dp,
                                     call_targets=call_targets)
    return O.identify_with_dp(iargs)[1]


def test_kat_regions_oracle_and_product():
    refs, reads = kat_sample()
    s = BuiltSample(Built(refs, reads))
    dp = B.api.dp_settings([dict(median=200.0, mad=0.0, r1_prefix=1, r2_prefix=2)], seed=2)
    assert _oracle_dp(s, dp) == KAT_SEED2
    assert B.api.dp_candidate_regions(s, dp) == KAT_SEED2


def test_kat_seed_zero_always_open():
    refs = _refs(n=1)
    s = BuiltSample(Built(refs, [_read(refs[0][1], 100, FLAG_R1, "FR", "1:7", 5000)]))
    dp = B.api.dp_settings([dict(median=200.0, mad=0.0, r1_prefix=1, r2_prefix=2)], seed=0)
    assert _oracle_dp(s, dp) == KAT_SEED0_ONE_READ
    assert B.api.dp_candidate_regions(s, dp) == KAT_SEED0_ONE_READ


def test_no_paired_group_writes_header_only():
    refs, reads = kat_sample()
    s = BuiltSample(Built(refs, reads))
    header = KAT_SEED2.splitlines()[0] + "\n"
    for dp in (None, B.api.dp_settings([dict(median=200.0, mad=0.0, paired=False)]),
               B.api.dp_settings([dict(median=0.0, mad=5.0)])):
        assert B.api.dp_candidate_regions(s, dp) == header
        assert _oracle_dp(s, dp) == header


def test_failed_coverage_fit_and_uncalled_targets_have_no_regions():
    refs, reads = kat_sample()
    s = BuiltSample(Built(refs, reads))
    dp = B.api.dp_settings([dict(median=200.0, mad=0.0, r1_prefix=1, r2_prefix=2)], seed=2)
    t0_only = "".join(ln + "\n" for ln in KAT_SEED2.splitlines() if not ln.startswith("t1,"))
    t1_only = "".join(ln + "\n" for ln in KAT_SEED2.splitlines() if not ln.startswith("t0,"))
    assert B.api.dp_candidate_regions(s, dp, propagation_cutoffs=[3.0, -1.0]) == t0_only
    assert _oracle_dp(s, dp, prop=[3.0, -1.0]) == t0_only
    assert B.api.dp_candidate_regions(s, dp, call_targets={1}) == t1_only
    assert _oracle_dp(s, dp, call_targets={1}) == t1_only


def test_missing_orientation_is_refused():
    """breseq drops such a read from its start column's tallies (identify_mutations.cpp:1411);
    the product reports it instead of changing the columns."""
    refs, reads = kat_sample()
    reads[0]["xp"] = None
    s = BuiltSample(Built(refs, reads))
    dp = B.api.dp_settings([dict(median=200.0, mad=0.0)], seed=2)
    with pytest.raises(B.BsqError) as e:
        B.api.dp_candidate_regions(s, dp)
    assert e.value.code == 6
    # without DP groups the same input is fine
    assert B.api.dp_candidate_regions(s, None).count("\n") == 1


def random_pair_reads(seed, n_targets=2, target_len=4000, n_clusters=14, read_len=40):
    """Discordant reads clustered in every bin, in three read file sets (two paired with different
    window widths, one unpaired), with excluded flags, absent RG tags, names with and without
    ':', redundant reads, equal start columns, deletions and reads at target ends."""
    rng = np.random.default_rng(seed)
    refs = _refs(n_targets, target_len, seed + 100)
    reads = []
    counter = [0]

    def one(t, pos, strand, orient, rg):
        counter[0] += 1
        mate2 = rng.random() < 0.5
        flag = (FLAG_R2 if mate2 else FLAG_R1) | (REV if strand else 0)
        u = rng.random()
        if u < 0.06:
            flag |= int(rng.choice([0x2, 0x8, 0x100, 0x800]))
        elif u < 0.09:
            flag &= ~0x1
        name = ("%d:%d" % (2 if mate2 else 1, counter[0])) if rng.random() < 0.9 else "n%d" % counter[0]
        cig = "%dM" % read_len if rng.random() < 0.8 else "15M%dD%dM" % (int(rng.integers(1, 6)), read_len - 15)
        span = read_len + (int(cig.split("D")[0].split("M")[1]) if "D" in cig else 0)
        pos = min(pos, target_len - span)
        tlen = int(rng.integers(-20000, 20000))
        x1 = int(rng.integers(2, 4)) if rng.random() < 0.25 else -1
        reads.append(_read(refs[t][1], pos, flag, ["FR", "RF", "FF"][orient], name, tlen, cigar=cig, x1=x1,
                           rg=rg, tid=t))

    for _ in range(n_clusters):
        t = int(rng.integers(n_targets))
        center = int(rng.integers(0, target_len - read_len))
        strand, orient = int(rng.integers(2)), int(rng.integers(3))
        rg = int(rng.choice([-1, 0, 1, 2]))
        for _ in range(int(rng.integers(1, 8))):
            jitter = 0 if rng.random() < 0.2 else int(rng.integers(0, 260))
            one(t, center + jitter, strand, orient, rg)
    for _ in range(60):   # scattered singles
        one(int(rng.integers(n_targets)), int(rng.integers(0, target_len)), int(rng.integers(2)),
            int(rng.integers(3)), int(rng.choice([-1, 0, 1, 2])))
    reads.sort(key=lambda r: (r["tid"], r["pos"]))
    return refs, reads


SETS = [dict(median=180.0, mad=20.0, r1_prefix=1, r2_prefix=2),      # width 228.4
        dict(median=100.0, mad=0.0, r1_prefix=3, r2_prefix=4),       # width 100 (integer)
        dict(median=150.0, mad=10.0, paired=False)]


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
@pytest.mark.parametrize("dp_seed", [1, 2, 3])
def test_random_product_matches_oracle(seed, dp_seed):
    refs, reads = random_pair_reads(seed)
    s = BuiltSample(Built(refs, reads))
    dp = B.api.dp_settings(SETS, seed=dp_seed)
    want = _oracle_dp(s, dp)
    got = B.api.dp_candidate_regions(s, dp)
    assert got == want
    if dp_seed <= 2:
        assert want.count("\n") > 3, want


def test_simulated_discordant_pairs_round_trip_and_regions(tmp_path):
    s = B.Sample.simulate(seed=21, ref_length=20000, read_length=100, coverage=20.0, paired=True, insert_mean=300,
                          insert_sd=30, n_snps=5, n_indels=2, discordant_fraction=0.03, n_junctions=4)
    a = s.arrays()
    disc = ((a["flag"] & 0x2) == 0) & ((a["flag"] & 0x1) != 0)
    assert disc.sum() > 20
    assert set(np.unique(a["xp"][disc])) <= {0, 1, 2}
    bam, fa = str(tmp_path / "d.bam"), str(tmp_path / "d.fasta")
    s.write_bam(bam)
    s.write_fasta(fa)
    r = B.Sample.read(bam, fa, read_file_groups=s.read_file_groups)
    ra = r.arrays()
    for k in a:
        assert np.array_equal(a[k], ra[k]), k
    assert r.read_names() == s.read_names()
    dp = B.api.dp_settings([dict(median=300.0, mad=20.0, r1_prefix=1, r2_prefix=2)], seed=3)
    want = _oracle_dp(s, dp)
    assert B.api.dp_candidate_regions(s, dp) == want
    assert B.api.dp_candidate_regions(r, dp) == want
    assert want.count("\n") > 2
