"""BAM2COV coverage tables (coverage_output.cpp:190-503): the oracle against tables worked out by
hand from the reference's traversal, for downsampling, the positions after the last covered column,
a region no read overlaps, read groups and the total-only / CSV forms."""
import ctypes as C

import pytest

import bam2cov_golden
import oracle_lib as O
from aln_builder import Built, BuiltSample
from breseq_amd import _lib as L
from breseq_amd import api

REF = "ACGTACGTAC"
# 1-based coverage: A 2-4 top (begins at 2); B 3-4,6-7 bottom with a deletion at 5 (its first base
# in read direction is at 7); C 4-5 top X1=2; D 4-5 bottom X1=3
READS = [
    dict(pos=1, cigar="3M", seq="CGT", rg=0),
    dict(pos=2, cigar="2M1D2M", seq="GTCG", flag=16, rg=1),
    dict(pos=3, cigar="2M", seq="TA", x1=2, rg=1),
    dict(pos=3, cigar="2M", seq="TA", flag=16, x1=3, rg=-1),
]


def table(resolution=0, start=0, end=0, n_rg=0, **kw):
    b = Built([("chr", REF)], READS)
    s = BuiltSample(b, n_read_groups=n_rg)
    args = api.coverage_table_args(s, 0, start, end, resolution, **kw)
    return O.coverage_table(C.pointer(b.ref), C.pointer(b.aln), args)


def rows(*r):
    return "".join("\t".join(str(x) for x in row) + "\n" for row in r)


HEAD8 = ["position", "ref_base", "unique_top_cov", "unique_bot_cov", "redundant_top_cov", "redundant_bot_cov",
         "raw_redundant_top_cov", "raw_redundant_bot_cov", "unique_top_begin", "unique_bot_begin"]
Z = [0] * 8
FULL = {
    1: ["A"] + Z,
    2: ["C", 1, 0, 0, 0, 0, 0, 1, 0],
    3: ["G", 1, 1, 0, 0, 0, 0, 0, 0],
    4: ["T", 1, 1, 0.5, "0.333333", 1, 1, 0, 0],
    5: ["A", 0, 0, 0.5, "0.333333", 1, 1, 0, 0],
    6: ["C", 0, 1, 0, 0, 0, 0, 0, 0],
    7: ["G", 0, 1, 0, 0, 0, 0, 0, 1],
    8: ["T"] + Z,
    9: ["A"] + Z,
    10: ["C"] + Z,
}


def foot(u, r, t, n):
    return rows(["#", "region_unique_average_cov", u], ["#", "region_repeat_average_cov", r],
                ["#", "region_average_cov", t], ["#", "number_of_positions", n])


def test_every_position():
    want = rows(HEAD8, *[[p] + FULL[p] for p in range(1, 11)]) + foot(0.7, "0.166667", "0.866667", 10)
    assert table() == want


def test_downsampled_tail_prints_skipped_positions():
    # resolution 5 over 10 positions: step 2, positions with (p + 1) even.  Up to the last covered
    # column (7) each handled position has one row; after it, position 9 also prints position 8
    # (unhandled) as a zero row, since only handled positions advance the traversal there
    want = rows(HEAD8, *[[p] + FULL[p] for p in (1, 3, 5, 7)], [8] + FULL[8], [9] + FULL[9])
    want += foot(0.5, "0.138889", "0.638889", 6)
    assert table(resolution=5, start=1, end=10) == want


def test_region_without_reads_starts_from_position_one():
    # no read overlaps 8-10 (B ends at 7), so the traversal starts at position 1: the first handled
    # position (8) prints positions 1-7 as zero rows even though reads cover 2-7
    want = rows(HEAD8, *[[p, REF[p - 1]] + Z for p in range(1, 11)]) + foot(0, 0, 0, 10)
    assert table(resolution=0, start=8, end=10) == want


def test_region_clips_and_counts_columns_past_its_end():
    # region 3-4: reads run to 7, so the traversal ends past the region; no tail rows
    want = rows(HEAD8, [3] + FULL[3], [4] + FULL[4]) + foot(2, "0.416667", "2.41667", 2)
    assert table(start=3, end=4) == want


def test_total_only_csv_with_average():
    got = table(start=4, end=5, total_only=True, csv=True, reference_average=12.5)
    want = ("position,ref_base,unique_cov,redundant_cov,total_cov\n"
            "4,T,2,0.833333,2.83333\n"
            "5,A,0,0.833333,0.833333\n"
            "#,reference_unique_average_cov,12.5\n"
            "#,region_unique_average_cov,1\n"
            "#,region_repeat_average_cov,0.833333\n"
            "#,region_average_cov,1.83333\n"
            "#,number_of_positions,2\n")
    assert got == want


def test_per_read_group_columns():
    # groups: A -> 0, B -> 1, C -> 1, D has no tag -> group 0
    got = table(start=4, end=5, per_read_group=True, n_rg=2)
    h = HEAD8 + ["RG-%d_%s" % (g, c) for g in range(2) for c in HEAD8[2:]]
    r4 = [4] + FULL[4] + [1, 0, 0, "0.333333", 0, 1, 0, 0] + [0, 1, 0.5, 0, 1, 0, 0, 0]
    r5 = [5] + FULL[5] + [0, 0, 0, "0.333333", 0, 1, 0, 0] + [0, 0, 0.5, 0, 1, 0, 0, 0]
    want = rows(h, r4, r5) + foot(1, "0.833333", "1.83333", 2)
    want += rows(["#", "RG-0_region_unique_average_cov", 0.5], ["#", "RG-0_region_repeat_average_cov", "0.333333"],
                 ["#", "RG-0_region_average_cov", "0.833333"], ["#", "RG-1_region_unique_average_cov", 0.5],
                 ["#", "RG-1_region_repeat_average_cov", 0.5], ["#", "RG-1_region_average_cov", 1])
    assert got == want


def test_one_read_group_maps_every_read_to_group_zero():
    # with a single @RG line every read is group 0 (read_group_index_map::index)
    got = table(start=4, end=4, per_read_group=True, n_rg=1, total_only=True)
    assert got.splitlines()[1] == "4\tT\t2\t0.833333\t2.83333\t2\t0.833333\t2.83333"


def test_region_checks():
    with pytest.raises(L.BsqError):
        table(start=0, end=11)
    with pytest.raises(L.BsqError):
        table(start=5, end=4)


@pytest.mark.parametrize("case", sorted(bam2cov_golden.CASES))
def test_reference_golden_tables(case):
    # breseq's own BAM2COV tables, over alignments rebuilt from their coverage columns
    refs, reads, n_rg, kw, want = bam2cov_golden.CASES[case]()
    b = Built(refs, reads)
    s = BuiltSample(b, n_read_groups=n_rg)
    got = O.coverage_table(C.pointer(b.ref), C.pointer(b.aln), api.coverage_table_args(s, **kw))
    assert got == want
