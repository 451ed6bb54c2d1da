"""Test-side binding of the oracle (oracle/liboracle.so): the CPU restatement of breseq's
pileup-scoring path, used only as the checker."""
import ctypes as C
import os
import subprocess

import numpy as np

from breseq_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "liboracle.so")

_O = None


def build():
    subprocess.check_call(["make", "-s", "-C", ORACLE_DIR])


def lib():
    global _O
    if _O is not None:
        return _O
    if not os.path.exists(ORACLE_LIB):
        build()
    O = C.CDLL(ORACLE_LIB)
    P = C.POINTER
    vp = C.c_void_p
    st = P(L.Status)
    O.orc_error_count.argtypes = [P(L.ErrorCountArgs), P(vp), st]
    O.orc_ec_histogram.argtypes = [vp, C.c_int32, P(L.c_u32p), L.c_i64p]
    O.orc_ec_counts.argtypes = [vp, P(P(C.c_uint64)), L.c_i64p]
    O.orc_ec_log10_prob.argtypes = [vp, P(L.c_f64p), L.c_i64p]
    O.orc_ec_text.argtypes = [vp, C.c_int32, C.c_int32, P(C.c_char_p), L.c_i64p]
    O.orc_ec_free.argtypes = [vp]
    O.orc_ec_no_pos_hash_pr.argtypes = [vp, P(L.c_f64p), L.c_i64p]
    O.orc_read_log10_prob_table.argtypes = [C.c_char_p, L.c_f64p, C.c_int64]
    O.orc_identify_mutations.argtypes = [P(L.IdentifyArgs), P(vp), st]
    O.orc_im_gd_text.argtypes = [vp, P(C.c_char_p), L.c_i64p]
    O.orc_im_dp_text.argtypes = [vp, P(C.c_char_p), L.c_i64p]
    O.orc_im_position_scores.argtypes = [vp, P(L.c_f64p), L.c_i64p]
    O.orc_im_free.argtypes = [vp]
    O.orc_sequence_trims.argtypes = [C.c_char_p, C.c_int64, L.c_u8p]
    O.orc_alignment_trims.argtypes = [P(L.Reference), P(L.Alignments), L.c_i32p, L.c_i32p]
    O.orc_coverage_table.argtypes = [P(L.Reference), P(L.Alignments), P(L.CoverageTableArgs), P(vp), L.c_i64p, st]
    O.orc_free_text.argtypes = [vp]
    O.orc_fisher_exact_2x2.restype = C.c_double
    O.orc_fisher_exact_2x2.argtypes = [C.c_uint32] * 4
    O.orc_ks_test_less.restype = C.c_double
    O.orc_ks_test_less.argtypes = [L.c_f64p, C.c_int64, L.c_f64p, C.c_int64]
    O.orc_fit_alleles.restype = C.c_double
    O.orc_fit_alleles.argtypes = [L.c_f64p, L.c_f64p, L.c_i32p, C.c_int64, L.c_i32p, C.c_double, L.c_f64p, L.c_i32p]
    O.orc_pure_genotype.restype = C.c_double
    O.orc_pure_genotype.argtypes = [L.c_f64p, C.c_uint32, L.c_i32p]
    O.orc_pileup_column.restype = C.c_int64
    O.orc_pileup_column.argtypes = [P(L.Alignments), C.c_int32, C.c_int64, L.c_i32p, C.c_int64]
    _O = O
    return O


def error_count(args):
    O = lib()
    h = C.c_void_p()
    st = L.Status()
    rc = O.orc_error_count(C.byref(args), C.byref(h), C.byref(st))
    L.check(rc, st)
    try:
        hists = []
        for g in range(args.n_coverage_groups):
            p = L.c_u32p()
            n = C.c_int64()
            O.orc_ec_histogram(h, g, C.byref(p), C.byref(n))
            hists.append(np.ctypeslib.as_array(p, shape=(n.value,)).copy() if n.value else np.zeros(0, np.uint32))
        p64 = C.POINTER(C.c_uint64)()
        n = C.c_int64()
        O.orc_ec_counts(h, C.byref(p64), C.byref(n))
        counts = np.ctypeslib.as_array(p64, shape=(n.value,)).copy()
        pd = L.c_f64p()
        O.orc_ec_log10_prob(h, C.byref(pd), C.byref(n))
        log10p = np.ctypeslib.as_array(pd, shape=(n.value,)).copy()
        texts = {}
        t = C.c_char_p()
        O.orc_ec_text(h, 0, 0, C.byref(t), C.byref(n))
        texts["error_rates"] = C.string_at(t, n.value).decode()
        for g in range(args.n_coverage_groups):
            O.orc_ec_text(h, 1, g, C.byref(t), C.byref(n))
            texts["coverage_%d" % g] = C.string_at(t, n.value).decode()
        for r in range(args.read_set_max):
            if O.orc_ec_text(h, 2, r, C.byref(t), C.byref(n)) == 0:
                texts["base_qual_%d" % r] = C.string_at(t, n.value).decode()
        O.orc_ec_no_pos_hash_pr(h, C.byref(pd), C.byref(n))
        texts["no_pos_hash_pr"] = np.ctypeslib.as_array(pd, shape=(n.value,)).copy()
    finally:
        O.orc_ec_free(h)
    return hists, counts, log10p, texts


def identify_mutations(args):
    O = lib()
    h = C.c_void_p()
    st = L.Status()
    rc = O.orc_identify_mutations(C.byref(args), C.byref(h), C.byref(st))
    L.check(rc, st)
    try:
        t = C.c_char_p()
        n = C.c_int64()
        O.orc_im_gd_text(h, C.byref(t), C.byref(n))
        text = C.string_at(t, n.value).decode()
        ps = L.c_f64p()
        O.orc_im_position_scores(h, C.byref(ps), C.byref(n))
        scores = np.ctypeslib.as_array(ps, shape=(n.value,)).copy() if n.value else np.zeros(0)
    finally:
        O.orc_im_free(h)
    return text, scores


def identify_with_dp(args):
    """(ra_mc_evidence.gd text, DP_candidate_regions.csv text) of the oracle."""
    O = lib()
    h = C.c_void_p()
    st = L.Status()
    L.check(O.orc_identify_mutations(C.byref(args), C.byref(h), C.byref(st)), st)
    try:
        t = C.c_char_p()
        n = C.c_int64()
        O.orc_im_gd_text(h, C.byref(t), C.byref(n))
        gd = C.string_at(t, n.value).decode()
        O.orc_im_dp_text(h, C.byref(t), C.byref(n))
        dp = C.string_at(t, n.value).decode()
    finally:
        O.orc_im_free(h)
    return gd, dp


def read_log10_prob_table(text, n_bins):
    out = np.zeros(n_bins, dtype=np.float64)
    rc = lib().orc_read_log10_prob_table(text.encode(), out.ctypes.data_as(L.c_f64p), n_bins)
    assert rc == 0
    return out


def sequence_trims(seq):
    b = seq.encode() if isinstance(seq, str) else bytes(seq)
    out = np.zeros(2 * len(b), dtype=np.uint8)
    lib().orc_sequence_trims(b, len(b), out.ctypes.data_as(L.c_u8p))
    return out[: len(b)], out[len(b):]


def fisher(a, b, c, d):
    return lib().orc_fisher_exact_2x2(a, b, c, d)


def ks_less(x, y):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    return lib().orc_ks_test_less(x.ctypes.data_as(L.c_f64p), x.size, y.ctypes.data_as(L.c_f64p), y.size)


def alignment_trims(ref_ptr, aln_ptr):
    """XL / XR of every alignment (get_alignment_trims, calculate_trims.cpp:183-285); -1 where the
    record is unmapped, has no target or covers no reference base."""
    n = int(aln_ptr.contents.n_reads)
    xl = np.zeros(n, dtype=np.int32)
    xr = np.zeros(n, dtype=np.int32)
    rc = lib().orc_alignment_trims(ref_ptr, aln_ptr, xl.ctypes.data_as(L.c_i32p), xr.ctypes.data_as(L.c_i32p))
    assert rc == 0, rc
    return xl, xr


def pileup_column(aln_ptr, tid, pos0, cap=4096):
    out = np.zeros(4 * cap, dtype=np.int32)
    n = lib().orc_pileup_column(aln_ptr, tid, pos0, out.ctypes.data_as(L.c_i32p), cap)
    assert n >= 0
    return out[: 4 * min(n, cap)].reshape(-1, 4)


def fit_alleles(r, lpmax, obs, allowed=(1, 1, 1, 1, 1), tol=1e-6):
    r = np.ascontiguousarray(r, dtype=np.float64)
    lpmax = np.ascontiguousarray(lpmax, dtype=np.float64)
    obs = np.ascontiguousarray(obs, dtype=np.int32)
    al = np.ascontiguousarray(allowed, dtype=np.int32)
    f = np.zeros(5)
    it = C.c_int32()
    ll = lib().orc_fit_alleles(r.ctypes.data_as(L.c_f64p), lpmax.ctypes.data_as(L.c_f64p), obs.ctypes.data_as(L.c_i32p),
                               len(obs), al.ctypes.data_as(L.c_i32p), tol, f.ctypes.data_as(L.c_f64p), C.byref(it))
    return f, ll, it.value


def pure_genotype(lps, n):
    lps = np.ascontiguousarray(lps, dtype=np.float64)
    best = C.c_int32()
    score = lib().orc_pure_genotype(lps.ctypes.data_as(L.c_f64p), n, C.byref(best))
    return best.value, score


def coverage_table(ref, aln, args):
    """BAM2COV table text of one region (coverage_output.cpp:190-503)."""
    O = lib()
    out = C.c_void_p()
    n = C.c_int64()
    st = L.Status()
    L.check(O.orc_coverage_table(ref, aln, C.byref(args), C.byref(out), C.byref(n), C.byref(st)), st)
    try:
        return C.string_at(out, n.value).decode()
    finally:
        O.orc_free_text(out)


def _bind_extra(O):
    if getattr(O, "_extra_bound", False):
        return O
    O.orc_read_base_likelihoods.restype = C.c_int
    O.orc_read_base_likelihoods.argtypes = [L.c_f64p, C.c_int64, C.c_int32, C.c_int32, C.c_uint32, C.c_uint32,
                                            C.c_uint32, C.c_int32, C.c_int32, L.c_f64p, L.c_f64p, L.c_f64p]
    O.orc_presence_and_bounds.restype = C.c_int
    O.orc_presence_and_bounds.argtypes = [L.c_f64p, L.c_f64p, L.c_i32p, C.c_int64, C.c_int32, C.c_double, C.c_double,
                                          L.c_f64p, L.c_f64p]
    O._extra_bound = True
    return O


def read_base_likelihoods(table, S, Q, rs, q, obs, strand, mapq):
    """fill_read_base_likelihoods of one read base: (lp[5], r[5], lpmax)."""
    O = _bind_extra(lib())
    t = np.ascontiguousarray(table, dtype=np.float64)
    lp, r, mx = np.zeros(5), np.zeros(5), C.c_double()
    rc = O.orc_read_base_likelihoods(t.ctypes.data_as(L.c_f64p), t.size, S, Q, rs, q, obs, strand, mapq,
                                     lp.ctypes.data_as(L.c_f64p), r.ctypes.data_as(L.c_f64p), C.byref(mx))
    assert rc == 0, rc
    return lp, r, mx.value


def presence_and_bounds(r, lpmax, obs, v, tol=1e-6, log10_ref_length=0.0):
    """(f[5], full-fit log10 likelihood, presence score, lower bound, upper bound) of one column."""
    O = _bind_extra(lib())
    r = np.ascontiguousarray(r, dtype=np.float64)
    lpmax = np.ascontiguousarray(lpmax, dtype=np.float64)
    obs = np.ascontiguousarray(obs, dtype=np.int32)
    f, out = np.zeros(5), np.zeros(4)
    O.orc_presence_and_bounds(r.ctypes.data_as(L.c_f64p), lpmax.ctypes.data_as(L.c_f64p), obs.ctypes.data_as(L.c_i32p),
                              len(obs), v, tol, log10_ref_length, f.ctypes.data_as(L.c_f64p),
                              out.ctypes.data_as(L.c_f64p))
    return f, out[0], out[1], out[2], out[3]


def _bind_fit(O):
    if getattr(O, "_fit_bound", False):
        return O
    O.orc_fit_coverage_distribution.restype = C.c_int
    O.orc_fit_coverage_distribution.argtypes = [L.c_u32p, C.c_int64, C.c_uint32, L.c_f64p]
    O.orc_special.restype = C.c_double
    O.orc_special.argtypes = [C.c_int32, C.c_double, C.c_double, C.c_double]
    O.orc_nelder_mead_quadratic.restype = C.c_int
    O.orc_nelder_mead_quadratic.argtypes = [C.c_double] * 5 + [L.c_f64p]
    O._fit_bound = True
    return O


FIT_FIELDS = ("nb_fit_size", "nb_fit_mu", "average", "variance", "relative_variance",
              "deletion_coverage_propagation_cutoff", "censor_start", "censor_end", "nb_fit_scale",
              "restarts_converged")


def fit_coverage_distribution(hist, sequence_length):
    """fit_coverage_distribution of a unique-only coverage histogram (hist[j], rows j >= 1)."""
    O = _bind_fit(lib())
    h = np.ascontiguousarray(hist, dtype=np.uint32)
    out = np.zeros(len(FIT_FIELDS))
    O.orc_fit_coverage_distribution(h.ctypes.data_as(L.c_u32p), h.size, int(sequence_length),
                                    out.ctypes.data_as(L.c_f64p))
    return dict(zip(FIT_FIELDS, out.tolist()))


SPECIAL = {"lngamma": 0, "gamma": 1, "lgam": 2, "incbet": 3, "nbdtr": 4, "dnbinom_mu": 5, "qnbinom_mu": 6}


def special(name, a, b=0.0, x=0.0):
    return _bind_fit(lib()).orc_special(SPECIAL[name], a, b, x)


def nelder_mead_quadratic(c0, c1, w, x0, x1):
    O = _bind_fit(lib())
    est = np.zeros(2)
    conv = O.orc_nelder_mead_quadratic(c0, c1, w, x0, x1, est.ctypes.data_as(L.c_f64p))
    return est, bool(conv)
