"""Synthetic workloads of the benchmark configurations, as generator settings (bsq_simulate).

Each is seeded and deterministic: the same settings give the same sample on any machine and thread
count.  Every workload also carries a few missing-coverage regions (deletions in the sampled genome,
reached into from both edges by repeat-mapped reads), so that the missing-coverage evidence, with
its redundant-coverage sub-ranges, is exercised at full size.
"""

# consensus mode, 48,502 bp uniform reference, 100,000 single-end 100-bp reads, qualities
# N(32, 6) clipped to [2, 41], 20 SNPs + 2 small indels
C1 = dict(seed=101, ref_length=48502, gc=0.5, read_length=100, coverage=100000 * 100 / 48502, paired=False,
          qual_mean=32, qual_sd=6, qual_min=2, qual_max=41, n_snps=20, n_indels=2, indel_max=5,
          redundant_fraction=0.01, mapq=42)

# C1 with one 400-bp deletion reached by 6 repeat-mapped reads from each edge
C1_DEL = dict(C1, seed=102, n_missing=1, missing_length=400, missing_edge_reads=6)

# consensus mode, 4.6 Mb reference at GC 50.8%, 100x of 2x150 bp pairs, insert N(350, 50), qualities
# N(34, 5), 500 SNPs + 50 indels of 1-5 bp, 0.2% discordant pairs at 40 junctions, three deletions
C2 = dict(seed=2, ref_length=4_600_000, gc=0.508, read_length=150, coverage=100.0, paired=True, insert_mean=350,
          insert_sd=50, qual_mean=34, qual_sd=5, qual_min=2, qual_max=41, n_snps=500, n_indels=50, indel_max=5,
          redundant_fraction=0.01, mapq=42, discordant_fraction=0.002, n_junctions=40, n_missing=3,
          missing_length=400, missing_edge_reads=6)

# polymorphism mode, 4.6 Mb at 500x of 2x150 bp pairs, 2,000 variants at log-uniform frequencies in
# [1%, 50%], qualities N(33, 6), three deletions
C3 = dict(seed=3, ref_length=4_600_000, gc=0.508, read_length=150, coverage=500.0, paired=True, insert_mean=350,
          insert_sd=50, qual_mean=33, qual_sd=6, qual_min=2, qual_max=41, n_snps=1800, n_indels=200, indel_max=5,
          variant_freq_min=0.01, variant_freq_max=0.5, redundant_fraction=0.01, mapq=42, discordant_fraction=0.002,
          n_junctions=40, n_missing=3, missing_length=400, missing_edge_reads=12)

# the paired-mapping-distance fit of the one read file set of the paired workloads: median and
# MAD of N(350, 50); read names carry the file as their first field (1: R1, 2: R2)
DP_SET = dict(median=350.0, mad=50.0 * 0.6745, r1_prefix=1, r2_prefix=2)
DP_SEED = 3

MODES = {"C1": "consensus", "C1_DEL": "consensus", "C2": "consensus", "C3": "polymorphism"}
