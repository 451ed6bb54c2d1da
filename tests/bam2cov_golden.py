"""breseq's own BAM2COV golden tables (tests/golden/bam2cov, copied from the reference's
tests/bam2cov*) and alignment sets rebuilt from them.

The tables' input BAMs are not in the reference tree.  Every row of these tables has unique reads
only, so a table fixes, at each printed position and per strand, the unique coverage and the
number of reads whose first base (in read direction) lies there.  `rebuild_reads` makes unique,
ungapped reads that give exactly those numbers at the printed positions: a top-strand read begins
at its leftmost aligned base unless it carries a left soft clip, a bottom-strand read at its
rightmost aligned base unless it carries a right soft clip.  Running the table over the rebuilt
reads must then print the golden text byte for byte: the header, which positions get rows under
the default resolution, the reference bases, the per-group columns and the averages."""
import json
import os

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "bam2cov")
CLIP = 3


def path(name):
    return os.path.join(HERE, name)


def read_text(name):
    with open(path(name)) as f:
        return f.read()


def read_fasta(name):
    seqs, cur = [], None
    for line in read_text(name).splitlines():
        if line.startswith(">"):
            cur = [line[1:].split()[0], []]
            seqs.append(cur)
        elif cur is not None:
            cur[1].append(line.strip())
    return [(n, "".join(s)) for n, s in seqs]


def coverage_average(seq_id):
    with open(path("coverage_average.json")) as f:
        return json.load(f)[seq_id]


def parse_table(text, delim="\t"):
    """(header columns, rows as lists of strings) of a table, without the '#' footer lines."""
    lines = text.splitlines()
    head = lines[0].split(delim)
    rows = [ln.split(delim) for ln in lines[1:] if not ln.startswith("#")]
    return head, rows


def _strand_reads(positions, cov, begin, reversed_, ref):
    """Ungapped reads of one strand with coverage cov[i] and begin[i] read starts at positions[i]."""
    out = []
    active = []   # [start, clipped_left] of open reads, oldest first
    n = len(positions)

    def emit(start, end, clip_left, clip_right):
        m = end - start + 1
        seq = ("A" * CLIP if clip_left else "") + ref[start - 1:end] + ("A" * CLIP if clip_right else "")
        cig = ("%dS" % CLIP if clip_left else "") + "%dM" % m + ("%dS" % CLIP if clip_right else "")
        out.append(dict(pos=start - 1, cigar=cig, seq=seq, flag=16 if reversed_ else 0))

    if not reversed_:
        prev = None
        for i, p in enumerate(positions):
            c, b = cov[i], begin[i]
            assert b <= c
            k = max(0, len(active) + b - c)   # reads whose last base is the previous printed position
            for _ in range(k):
                s, cl = active.pop(0)
                emit(s, prev, cl, False)
            s_extra = c - len(active) - b
            assert s_extra >= 0
            active += [[p, False]] * b + [[p, True]] * s_extra
            prev = p
        for s, cl in active:
            emit(s, prev, cl, False)
    else:
        for i, p in enumerate(positions):
            c, b = cov[i], begin[i]
            assert b <= c
            opened = c - len(active)
            assert opened >= 0
            active += [p] * opened
            c_next = cov[i + 1] if i + 1 < n else 0
            closed = max(b, c - c_next)
            for j in range(closed):
                s = active.pop(0)
                emit(s, p, False, j >= b)   # the first b end here with no clip: their first base is p
    return out


def rebuild_reads(rows, cols, ref, rg=-1):
    """Reads reproducing columns cols = (unique_top, unique_bot, top_begin, bot_begin) (indexes
    into each row) at the rows' positions."""
    positions = [int(r[0]) for r in rows]
    reads = []
    for strand, (ci, bi) in enumerate(((cols[0], cols[2]), (cols[1], cols[3]))):
        cov = [int(r[ci]) for r in rows]
        beg = [int(r[bi]) for r in rows]
        reads += _strand_reads(positions, cov, beg, strand == 1, ref)
    for r in reads:
        r["rg"] = rg
    return reads


def sort_reads(reads):
    return sorted(reads, key=lambda r: r["pos"])


def case_bull(name="expected.tab"):
    """tests/bam2cov and tests/bam2cov_csv: rachael:657-2167 of bull_1, resolution 600,
    --show-average with summary.json's coverage_average."""
    delim = "," if name.endswith(".csv") else "\t"
    text = read_text(name)
    _, rows = parse_table(text, delim)
    refs = read_fasta("bull_1.fasta")
    reads = rebuild_reads(rows, (2, 3, 8, 9), refs[0][1])
    kw = dict(target=0, start=657, end=2167, resolution=600, csv=delim == ",",
              reference_average=coverage_average("rachael"))
    return refs, sort_reads(reads), 0, kw, text


def case_no_read_groups():
    """tests/bam2cov_per_read_group, first table: rachael:657-767 of bull_1 (no @RG lines),
    --per-read-group."""
    text = read_text("expected.no_read_groups.tab")
    _, rows = parse_table(text)
    refs = read_fasta("bull_1.fasta")
    reads = rebuild_reads(rows, (2, 3, 8, 9), refs[0][1])
    kw = dict(target=0, start=657, end=767, resolution=600, per_read_group=True)
    return refs, sort_reads(reads), 0, kw, text


def case_multiple_read_groups():
    """tests/bam2cov_per_read_group, second table: NC_001416-0:8000-8110 of a BAM with four @RG
    lines.  Reads are rebuilt per group from the RG-g columns only, so the aggregate columns test
    the sum over groups.  That BAM's FASTA is made by another test, so the reference bases outside
    the printed rows are filler."""
    text = read_text("expected.multiple_read_groups.tab")
    head, rows = parse_table(text)
    n_rg = sum(1 for h in head if h.endswith("_unique_top_cov") and h.startswith("RG-"))
    end = int(rows[-1][0])
    seq = ["A"] * (end + 50)
    for r in rows:
        seq[int(r[0]) - 1] = r[1]
    ref = "".join(seq)
    reads = []
    for g in range(n_rg):
        i = head.index("RG-%d_unique_top_cov" % g)
        reads += rebuild_reads(rows, (i, i + 1, i + 6, i + 7), ref, rg=g)
    kw = dict(target=0, start=int(rows[0][0]), end=end, resolution=600, per_read_group=True)
    return [("NC_001416-0", ref)], sort_reads(reads), n_rg, kw, text


CASES = {
    "bam2cov": lambda: case_bull("expected.tab"),
    "bam2cov_csv": lambda: case_bull("expected.csv"),
    "per_read_group_none": case_no_read_groups,
    "per_read_group_multiple": case_multiple_read_groups,
}
