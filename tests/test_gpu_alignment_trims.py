"""XL / XR of every alignment through the GPU entry point (bsq_session_alignment_trims) against the
oracle's restatement of get_alignment_trims (calculate_trims.cpp:183-285): the hand-worked set,
alignments on repeat-rich references (homopolymers past the 255 cap, units up to 18 and beyond,
soft and hard clips, indels, N bases, records that are not packed), and a simulated genome."""
import sys

import numpy as np
import pytest

import breseq_amd as B
import oracle_lib as O
from aln_builder import Built, BuiltSample
from test_oracle_pins import TRIM_READS, trim_sample

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def trims_on_gpu(s):
    sess = B.Session().load(s)
    try:
        return sess.alignment_trims()
    finally:
        sess.close()


def test_hand_worked_trims():
    say("[gpu] alignment trims: hand-worked set")
    xl, xr = trims_on_gpu(trim_sample())
    assert list(xl) == [r[4] for r in TRIM_READS]
    assert list(xr) == [r[5] for r in TRIM_READS]


def repeat_rich(seed, n_targets=2, target_len=3000, n_reads=4000):
    rng = np.random.default_rng(seed)
    refs = []
    for t in range(n_targets):
        parts, n = [], 0
        while n < target_len:
            u = rng.random()
            if u < 0.4:
                piece = "".join(rng.choice(list("ACGT"), int(rng.integers(5, 40))))
            elif u < 0.6:
                piece = str(rng.choice(list("ACGT"))) * int(rng.integers(2, 30))
            elif u < 0.95:
                unit = "".join(rng.choice(list("ACGT"), int(rng.integers(2, 22))))
                piece = unit * int(rng.integers(2, 6))
            else:
                piece = str(rng.choice(list("ACGT"))) * int(rng.integers(250, 320))   # past the 255 cap
            parts.append(piece)
            n += len(piece)
        refs.append(("t%d" % t, "".join(parts)[:target_len]))
    reads = []
    for i in range(n_reads):
        t = int(rng.integers(n_targets))
        ref = refs[t][1]
        ops = []
        if rng.random() < 0.1:
            ops.append(("H", int(rng.integers(1, 5))))
        if rng.random() < 0.3:
            ops.append(("S", int(rng.integers(1, 12))))
        for k in range(int(rng.integers(1, 4))):
            ops.append(("M", int(rng.integers(8, 60))))
            if k < 2 and rng.random() < 0.4:
                ops.append((str(rng.choice(list("DI"))), int(rng.integers(1, 4))))
        if ops[-1][0] in "DI":
            ops.append(("M", int(rng.integers(8, 30))))
        if rng.random() < 0.3:
            ops.append(("S", int(rng.integers(1, 12))))
        if rng.random() < 0.1:
            ops.append(("H", int(rng.integers(1, 5))))
        span = sum(l for o, l in ops if o in "MD")
        pos = int(rng.integers(0, target_len - span + 1))
        seq, rp = [], pos
        for o, l in ops:
            if o == "M":
                seq.extend(ref[rp:rp + l])
                rp += l
            elif o == "D":
                rp += l
            elif o in "IS":
                # clipped and inserted bases often continue the repeat they sit next to
                src = ref[max(0, rp - l):rp] if rng.random() < 0.5 and rp >= l else "".join(rng.choice(list("ACGT"), l))
                seq.extend(src)
        seq = ["N" if rng.random() < 0.005 else b for b in seq]
        flag = int(rng.choice([0, 16]))
        if rng.random() < 0.01:
            flag |= 4
        reads.append(dict(tid=t, pos=pos, cigar="".join("%d%s" % (l, o) for o, l in ops), seq="".join(seq), flag=flag))
    reads.sort(key=lambda r: (r["tid"], r["pos"]))
    return refs, reads


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_repeat_rich_alignments(seed):
    say("[gpu] alignment trims: repeat-rich seed", seed)
    refs, reads = repeat_rich(seed)
    s = BuiltSample(Built(refs, reads))
    want_l, want_r = O.alignment_trims(s.ref, s.aln)
    got_l, got_r = trims_on_gpu(s)
    assert np.array_equal(got_l, want_l) and np.array_equal(got_r, want_r)
    assert (want_l > 20).any() and (want_r > 255).any() and (want_l == -1).any()


def test_simulated_genome():
    say("[gpu] alignment trims: simulated genome")
    s = B.Sample.simulate(seed=5, ref_length=60000, read_length=150, coverage=40.0, paired=True, n_snps=20,
                          n_indels=4)
    want_l, want_r = O.alignment_trims(s.ref, s.aln)
    got_l, got_r = trims_on_gpu(s)
    assert np.array_equal(got_l, want_l) and np.array_equal(got_r, want_r)
