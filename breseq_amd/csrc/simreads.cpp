// simreads.cpp -- SIMULATE-READS: breseq's read simulator, restated so that its FASTQ output is
// byte-identical for a given seed (do_simulate_reads, breseq_cmdline.cpp:1150-1330;
// cSimFastqSequence, fastq.cpp:1254-1583; cPortableRandom, portable_random.cpp:22-88).
//
// Every read pair (or single read) draws from its own Mersenne-twister stream seeded through
// std::seed_seq{seed, index}: the standard fixes both, so the streams are the reference's.  All
// arithmetic on the stream's outputs is integer: unbiased draws below n by rejection of the top
// partial bucket, normal deviates as the Irwin-Hall sum of twelve 16-bit uniforms, per-base error
// events against Phred error rates held as exact integers per billion.
//
// Besides the FASTQ text, each read carries its alignment by construction (the reference span it
// was copied from, with the simulated insertions and deletions as CIGAR operations), which
// bsq_simulate_reads_sample packs into a sample for the pileup path.
#include "simreads.h"

#include <algorithm>
#include <cctype>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "host_util.h"

namespace bsq {

namespace {

class PortableRandom {
 public:
  PortableRandom(uint32_t seed, uint32_t stream) {
    std::seed_seq seq{seed, stream};
    eng_.seed(seq);
  }
  // uniform in [0, n): draws at or above the largest multiple of n below 2^32 are rejected
  uint32_t below(uint32_t n) {
    const uint64_t limit = (4294967296ull / n) * n;
    uint64_t x;
    do {
      x = eng_();
    } while (x >= limit);
    return (uint32_t)(x % n);
  }
  uint32_t in_range_1(uint32_t lo, uint32_t hi) { return lo + below(hi - lo + 1); }
  bool coin() { return below(2) == 1; }
  bool event_per_billion(uint32_t rate) { return below(1000000000u) < rate; }
  // mean + stdev * (sum of 12 uniforms on [0, 65536) - 393210) / 65536, rounded half up
  int32_t gaussian(int32_t mean, int32_t stdev) {
    if (stdev == 0) return mean;
    int64_t sum = -393210;
    for (int i = 0; i < 12; i++) sum += below(65536u);
    // floor((sum * stdev + 32768) / 65536), computed on a non-negative numerator
    const int64_t bias = (int64_t)1 << 40;
    const int64_t num = sum * stdev + 32768 + bias;
    return mean + (int32_t)(num / 65536 - bias / 65536);
  }

 private:
  std::mt19937 eng_;
};

// Phred error probability in parts per billion: the ten mantissas 1e9 * 10^(-k/10), then one
// division by ten per decade (integer, so the same on every platform)
uint32_t error_per_billion(uint32_t phred) {
  static const uint32_t mant[10] = {1000000000u, 794328235u, 630957344u, 501187234u, 398107171u,
                                    316227766u,  251188643u, 199526231u, 158489319u, 125892541u};
  uint32_t r = mant[phred % 10];
  for (uint32_t d = phred / 10; d > 0; d--) r /= 10;
  return r;
}

char complement(char c) {
  switch (c) {
    case 'A': return 'T';
    case 'T': return 'A';
    case 'C': return 'G';
    case 'G': return 'C';
    default: return 'N';
  }
}

// substitution choices for a miscalled base (every base but the true one; any base for N)
const char* error_choices(char b) {
  switch (b) {
    case 'A': return "TCG";
    case 'T': return "ACG";
    case 'C': return "ATG";
    case 'G': return "ATC";
    case 'N': return "ACTG";
    default: return nullptr;
  }
}

struct Model {
  uint32_t q_start, q_end, q_stdev, q_min, q_max;
  uint32_t indel_per_million;
};

// Phred score of one cycle: a quadratic decline from q_start to q_end across the read, a normal
// scatter of q_stdev, clipped to [q_min, q_max]
uint32_t draw_quality(PortableRandom& rng, const Model& m, uint32_t cycle, uint32_t read_size) {
  const uint64_t n = read_size > 1 ? read_size - 1 : 1;
  const uint64_t drop = m.q_start > m.q_end ? m.q_start - m.q_end : 0;
  const uint32_t shift = (uint32_t)(drop * cycle * cycle / (n * n));
  int32_t q = rng.gaussian((int32_t)m.q_start - (int32_t)shift, (int32_t)m.q_stdev);
  q = std::max(q, (int32_t)m.q_min);
  q = std::min(q, (int32_t)m.q_max);
  return (uint32_t)q;
}

// One read starting at left_1 (1-based) on `strand`, built along a buffer of the reference from the
// read's 5' end (the minus strand's buffer reverse complemented).  ops: the alignment in read order
// (M / I / D runs); buf_origin: the reference index of the buffer's first base, whose later bases
// run forward on the plus strand and backward on the minus strand.
struct SimRead {
  std::string name, name_plus, seq, qual;
  std::vector<std::pair<char, uint32_t>> ops;
  int64_t buf_origin;   // reference index (0-based, may lie outside [0, len) when circular) of buffer[0]
  int strand;
};

void add_op(std::vector<std::pair<char, uint32_t>>& ops, char op) {
  if (!ops.empty() && ops.back().first == op) ops.back().second++;
  else ops.emplace_back(op, 1u);
}

SimRead simulate_read(const std::string& ref, bool circular, int32_t left_1, uint32_t read_size, int strand,
                      const Model& m, PortableRandom& rng, const std::string& name) {
  SimRead r;
  r.name = name;
  r.name_plus = "[strand]:" + std::to_string(strand) + "\t[left_1]:" + std::to_string(left_1);
  r.strand = strand;
  const int32_t len = (int32_t)ref.size();
  // the buffer: the read's footprint plus read_size more bases in its direction of travel
  const int32_t fetch = strand == -1 ? left_1 - (int32_t)read_size : left_1;
  std::string buf;
  if (circular) {
    int64_t s = fetch;
    while (s < 1) s += len;
    s %= len;
    if (s == 0) s = len;
    buf.resize(2 * (size_t)read_size);
    for (size_t k = 0; k < buf.size(); k++) buf[k] = ref[(size_t)((s - 1 + (int64_t)k) % len)];
    r.buf_origin = s - 1;
  } else {
    const int32_t lo = std::max<int32_t>(1, fetch), hi = std::min<int32_t>(len, fetch + 2 * (int32_t)read_size - 1);
    if (lo <= hi) buf = ref.substr((size_t)lo - 1, (size_t)(hi - lo + 1));
    r.buf_origin = lo - 1;
  }
  if (strand == -1) {
    std::string rc(buf.size(), 'N');
    for (size_t k = 0; k < buf.size(); k++) rc[k] = complement(buf[buf.size() - 1 - k]);
    buf.swap(rc);
    r.buf_origin += (int64_t)buf.size() - 1;   // buffer[0] is now the last fetched base
  }
  if (buf.size() < read_size)
    throw BsqError(BSQ_ERR_ARG, "simulated read at " + std::to_string(left_1) + " would extend past the sequence end");
  r.seq.resize(read_size);
  r.qual.resize(read_size);
  const uint32_t indel_rate = m.indel_per_million * 1000u;
  size_t out = 0, at = 0;
  while (out < read_size) {
    if (at + 1 < buf.size() && rng.event_per_billion(indel_rate)) {
      at++;   // a deletion: a reference base consumed, none emitted
      add_op(r.ops, 'D');
      continue;
    }
    const uint32_t q = draw_quality(rng, m, (uint32_t)out, read_size);
    const char qc = (char)(q + 33);
    if (rng.event_per_billion(indel_rate)) {
      static const char kBases[4] = {'A', 'C', 'G', 'T'};
      r.seq[out] = kBases[rng.below(4)];   // an insertion: a base emitted, none consumed
      r.qual[out] = qc;
      out++;
      add_op(r.ops, 'I');
      continue;
    }
    const char b = buf[at];
    if (rng.event_per_billion(error_per_billion(q))) {
      const char* ch = error_choices(b);
      if (!ch) throw BsqError(BSQ_ERR_DATA, std::string("unexpected reference base ") + b);
      r.seq[out] = ch[rng.below((uint32_t)std::strlen(ch))];
    } else {
      r.seq[out] = b;
    }
    r.qual[out] = qc;
    out++;
    at++;
    add_op(r.ops, 'M');
  }
  return r;
}

void append_fastq(std::string& t, const SimRead& r) {
  t += '@';
  t += r.name;
  t += '\n';
  t += r.seq;
  t += "\n+";
  t += r.name_plus;
  t += '\n';
  t += r.qual;
  t += '\n';
}

}  // namespace

void simulate_reads(const std::string& ref_in, const bsq_simreads_args& a, std::vector<SimReadOut>* reads,
                    std::string* fastq_1, std::string* fastq_2) {
  std::string ref(ref_in);
  for (char& c : ref) c = (char)std::toupper((unsigned char)c);
  const int32_t len = (int32_t)ref.size();
  if (len <= 0) throw BsqError(BSQ_ERR_ARG, "empty reference sequence");
  if (a.read_length == 0) throw BsqError(BSQ_ERR_ARG, "read length 0");
  if (a.quality_min > a.quality_max) throw BsqError(BSQ_ERR_ARG, "quality_min above quality_max");
  const bool paired = a.mode == 1;
  if (a.mode != 0 && a.mode != 1) throw BsqError(BSQ_ERR_ARG, "mode must be 0 (single) or 1 (paired-end)");
  if (!a.circular && (int64_t)a.read_length > len) throw BsqError(BSQ_ERR_ARG, "read longer than the sequence");
  uint32_t n = a.n_reads;
  if (n == 0) n = (uint32_t)((uint64_t)len * a.coverage / ((uint64_t)a.read_length * (paired ? 2 : 1)));
  const Model m{a.quality_start, a.quality_end, a.quality_stdev, a.quality_min, a.quality_max,
                a.indel_error_rate_per_million};
  const uint32_t rs = a.read_length;
  auto keep = [&](const SimRead& r, int mate) {
    if (!reads) return;
    SimReadOut o;
    o.seq = r.seq;
    o.qual = r.qual;
    o.name = r.name;
    o.strand = r.strand;
    o.mate = mate;
    // alignment by construction on the forward strand: minus-strand reads are stored reverse
    // complemented with their operations reversed; the leftmost aligned base is found by walking
    // the buffer
    int64_t ref_consumed = 0;
    for (const auto& op : r.ops)
      if (op.first != 'I') ref_consumed += op.second;
    if (r.strand == 1) {
      o.left0 = r.buf_origin;
      o.ops = r.ops;
    } else {
      o.left0 = r.buf_origin - ref_consumed + 1;
      o.ops.assign(r.ops.rbegin(), r.ops.rend());
      std::string s(r.seq.rbegin(), r.seq.rend()), q(r.qual.rbegin(), r.qual.rend());
      for (char& c : s) c = complement(c);
      o.seq = s;
      o.qual = q;
    }
    // a deletion drawn before the first or after the last emitted base is no part of the alignment
    while (!o.ops.empty() && o.ops.front().first == 'D') {
      o.left0 += o.ops.front().second;
      ref_consumed -= o.ops.front().second;
      o.ops.erase(o.ops.begin());
    }
    while (!o.ops.empty() && o.ops.back().first == 'D') {
      ref_consumed -= o.ops.back().second;
      o.ops.pop_back();
    }
    o.ref_span = ref_consumed;
    reads->push_back(std::move(o));
  };
  for (uint32_t i = 0; i < n; i++) {
    PortableRandom rng(a.seed, i);
    const std::string name = "READ-" + std::to_string(i + 1);
    if (!paired) {
      const int strand = rng.coin() ? 1 : -1;
      const int32_t left = a.circular ? (int32_t)rng.in_range_1(1, (uint32_t)len)
                                      : (int32_t)rng.in_range_1(1, (uint32_t)(len - (int32_t)rs + 1));
      SimRead r = simulate_read(ref, a.circular != 0, left, rs, strand, m, rng, name);
      if (fastq_1) append_fastq(*fastq_1, r);
      keep(r, 0);
      continue;
    }
    // fragment length: normal, at least one read long, at most the sequence; FR pair with a coin
    // deciding which end read 1 comes from
    int32_t L = rng.gaussian((int32_t)a.fragment_mean, (int32_t)a.fragment_stdev);
    L = std::max<int32_t>(L, (int32_t)rs);
    L = std::min<int32_t>(L, len);
    const int32_t left = a.circular ? (int32_t)rng.in_range_1(1, (uint32_t)len)
                                    : (int32_t)rng.in_range_1(1, (uint32_t)(len - L + 1));
    const bool r1_left = rng.coin();
    const int32_t lpos = left, rpos = left + L - (int32_t)rs;
    SimRead r1 = simulate_read(ref, a.circular != 0, r1_left ? lpos : rpos, rs, r1_left ? 1 : -1, m, rng, name + "/1");
    SimRead r2 = simulate_read(ref, a.circular != 0, r1_left ? rpos : lpos, rs, r1_left ? -1 : 1, m, rng, name + "/2");
    if (fastq_1) append_fastq(*fastq_1, r1);
    if (fastq_2) append_fastq(*fastq_2, r2);
    keep(r1, 1);
    keep(r2, 2);
  }
}

}  // namespace bsq

using bsq::BsqError;

extern "C" int bsq_simreads_defaults(bsq_simreads_args* a) {
  if (!a) return BSQ_ERR_ARG;
  *a = bsq_simreads_args{};
  a->mode = 0;
  a->coverage = 80;
  a->read_length = 50;
  a->fragment_mean = 200;
  a->fragment_stdev = 20;
  a->quality_start = 38;
  a->quality_end = 28;
  a->quality_stdev = 4;
  a->quality_min = 2;
  a->quality_max = 41;
  a->indel_error_rate_per_million = 10;
  return BSQ_OK;
}

extern "C" int bsq_simulate_reads(const char* seq, int64_t len, const bsq_simreads_args* args, char** fastq_1,
                                  int64_t* len_1, char** fastq_2, int64_t* len_2, bsq_status* status) {
  try {
    if (!seq || len <= 0 || !args || !fastq_1 || !len_1) throw BsqError(BSQ_ERR_ARG, "bad arguments");
    if (args->mode == 1 && (!fastq_2 || !len_2)) throw BsqError(BSQ_ERR_ARG, "paired-end output needs two files");
    if (len > 0x7fffffffLL) throw BsqError(BSQ_ERR_LIMIT, "sequence longer than 2^31 - 1 bases");
    std::string t1, t2;
    bsq::simulate_reads(std::string(seq, (size_t)len), *args, nullptr, &t1, args->mode == 1 ? &t2 : nullptr);
    auto give = [](const std::string& t, char** out, int64_t* n) {
      char* p = (char*)std::malloc(t.size() + 1);
      if (!p) throw BsqError(BSQ_ERR_LIMIT, "out of host memory");
      std::memcpy(p, t.data(), t.size());
      p[t.size()] = 0;
      *out = p;
      *n = (int64_t)t.size();
    };
    give(t1, fastq_1, len_1);
    if (args->mode == 1) give(t2, fastq_2, len_2);
    bsq::set_status(status, BSQ_OK, "");
    return BSQ_OK;
  } catch (const BsqError& e) {
    bsq::set_status(status, e.code, e.what());
    return e.code;
  }
}

#include "sample.h"

namespace {

// The genome the reads come from: `ref` with the mutations applied (sorted, non-overlapping), and
// for each of its bases the reference base it stands for (-1 for an inserted base).  The first
// min(ref_len, alt length) bases of a mutation replace reference bases one for one.
void apply_mutations(const std::string& ref, const bsq_sim_mutation* m, int32_t n, std::string& genome,
                     std::vector<int64_t>& orig) {
  genome.clear();
  orig.clear();
  int64_t at = 0;
  for (int32_t k = 0; k < n; k++) {
    const int64_t p = m[k].pos;
    const int64_t rl = m[k].ref_len;
    const std::string alt = m[k].alt ? m[k].alt : "";
    if (p < at || rl < 0 || p + rl > (int64_t)ref.size())
      throw BsqError(BSQ_ERR_ARG, "mutations must be sorted, non-overlapping and inside the sequence");
    for (; at < p; at++) {
      genome += ref[(size_t)at];
      orig.push_back(at);
    }
    for (size_t j = 0; j < alt.size(); j++) {
      genome += (char)std::toupper((unsigned char)alt[j]);
      orig.push_back((int64_t)j < rl ? p + (int64_t)j : -1);
    }
    at = p + rl;
  }
  for (; at < (int64_t)ref.size(); at++) {
    genome += ref[(size_t)at];
    orig.push_back(at);
  }
}

void push_cigar(std::vector<uint32_t>& cig, uint32_t op, uint32_t n) {
  if (n == 0) return;
  if (!cig.empty() && (cig.back() & 15u) == op) cig.back() += n << 4;
  else cig.push_back((n << 4) | op);
}

}  // namespace

extern "C" int bsq_simulate_reads_sample(const char* name, const char* seq, int64_t len, const bsq_sim_mutation* muts,
                                         int32_t n_muts, const bsq_simreads_args* args, int32_t mapq,
                                         bsq_sample** out, bsq_status* status) {
  try {
    if (!name || !seq || len <= 0 || !args || !out || n_muts < 0 || (n_muts > 0 && !muts))
      throw BsqError(BSQ_ERR_ARG, "bad arguments");
    if (args->circular) throw BsqError(BSQ_ERR_ARG, "aligned samples are built from linear sequences only");
    if (mapq < 0 || mapq > 255) throw BsqError(BSQ_ERR_ARG, "mapq outside 0-255");
    std::string ref(seq, (size_t)len);
    for (char& c : ref) c = (char)std::toupper((unsigned char)c);
    std::string genome;
    std::vector<int64_t> orig;
    apply_mutations(ref, muts, n_muts, genome, orig);
    std::vector<bsq::SimReadOut> reads;
    bsq::simulate_reads(genome, *args, &reads, nullptr, nullptr);
    const bool paired = args->mode == 1;
    // each read's alignment to the original sequence, walked through the genome's base map:
    // inserted genome bases become I, reference bases the genome skips become D
    struct Rec {
      int64_t pos = -1, end = -1;   // 0-based first and last aligned reference base
      std::vector<uint32_t> cig;
    };
    std::vector<Rec> rec(reads.size());
    bsq::parallel_ranges((int64_t)reads.size(), bsq::host_threads(), [&](int, int64_t i0, int64_t i1) {
      for (int64_t i = i0; i < i1; i++) {
        const bsq::SimReadOut& r = reads[(size_t)i];
        Rec& R = rec[(size_t)i];
        int64_t g = r.left0, prev = -1;
        for (const auto& op : r.ops) {
          if (op.first == 'I') {
            push_cigar(R.cig, 1, op.second);
            continue;
          }
          for (uint32_t k = 0; k < op.second; k++, g++) {
            const int64_t o = orig[(size_t)g];
            if (op.first == 'D') continue;   // a simulated deletion: the reference gap shows at the next base
            if (o < 0) {
              push_cigar(R.cig, 1, 1);
              continue;
            }
            if (prev >= 0 && o > prev + 1) push_cigar(R.cig, 2, (uint32_t)(o - prev - 1));
            if (R.pos < 0) R.pos = o;
            push_cigar(R.cig, 0, 1);
            prev = o;
          }
        }
        R.end = prev;
      }
    });
    bsq_sample* smp = new bsq_sample();
    bsq::Sample& S = smp->s;
    S.names.push_back(name);
    S.seqs.push_back(ref);
    // file order: by first aligned base, then generation order; reads without an aligned base dropped
    std::vector<size_t> order;
    for (size_t i = 0; i < reads.size(); i++)
      if (rec[i].pos >= 0) order.push_back(i);
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return rec[a].pos < rec[b].pos; });
    std::vector<uint8_t> gtrims(2 * (size_t)len);
    bsq::host_sequence_trims(ref.data(), len, gtrims.data());
    // mate of each read of a pair: the other read generated from the same fragment
    std::vector<int64_t> mate(reads.size(), -1);
    if (paired)
      for (size_t i = 0; i + 1 < reads.size(); i += 2) {
        mate[i] = (int64_t)i + 1;
        mate[i + 1] = (int64_t)i;
      }
    const size_t nr = order.size();
    S.tid.assign(nr, 0);
    S.rg.assign(nr, 0);
    S.x1.assign(nr, 1);
    S.cigar_off.assign(nr + 1, 0);
    S.seq_off.assign(nr + 1, 0);
    static const uint8_t kCode[5] = {1, 2, 4, 8, 15};
    for (size_t k = 0; k < nr; k++) {
      const size_t i = order[k];
      const bsq::SimReadOut& r = reads[i];
      const Rec& R = rec[i];
      S.pos.push_back((int32_t)R.pos);
      uint16_t flag = r.strand == -1 ? 0x10 : 0;
      int32_t mpos = -1, tl = 0;
      int8_t xp = -1;
      if (paired) {
        const int64_t j = mate[i];
        flag |= 0x1 | (r.mate == 1 ? 0x40 : 0x80);
        if (j >= 0 && rec[(size_t)j].pos >= 0) {
          flag |= 0x2;
          if (reads[(size_t)j].strand == -1) flag |= 0x20;
          mpos = (int32_t)rec[(size_t)j].pos;
          const int64_t lo = std::min(R.pos, rec[(size_t)j].pos), hi = std::max(R.end, rec[(size_t)j].end);
          tl = (int32_t)(hi - lo + 1) * (R.pos <= rec[(size_t)j].pos ? 1 : -1);
          xp = 0;   // FR
        } else {
          flag |= 0x8;
        }
      }
      S.flag.push_back(flag);
      S.mapq.push_back((uint8_t)mapq);
      S.mate_pos.push_back(mpos);
      S.tlen.push_back(tl);
      S.xp.push_back(xp);
      S.mate_tid.push_back(paired ? 0 : -1);
      S.read_names.push_back(r.name);
      for (uint32_t c : R.cig) S.cigar.push_back(c);
      S.cigar_off[k + 1] = (int64_t)S.cigar.size();
      for (size_t q = 0; q < r.seq.size(); q++) {
        const char c = r.seq[q];
        S.seq.push_back(kCode[c == 'A' ? 0 : c == 'C' ? 1 : c == 'G' ? 2 : c == 'T' ? 3 : 4]);
        S.qual.push_back((uint8_t)(r.qual[q] - 33));
      }
      S.seq_off[k + 1] = (int64_t)S.seq.size();
      // XL / XR as breseq writes them (get_alignment_trims): no soft clips here
      uint32_t el = 1, er = 1;
      bsq::host_edge_trims(r.seq, el, er);
      S.xl.push_back((int32_t)std::max<uint32_t>(gtrims[(size_t)R.pos], el));
      S.xr.push_back((int32_t)std::max<uint32_t>(gtrims[(size_t)len + (size_t)R.end], er));
    }
    S.rg_ids.push_back("0");
    S.rg_libs.push_back(paired ? "reads" : "reads_1");
    S.rf_base.push_back(0);
    S.rf_count.push_back(paired ? 2 : 1);
    S.rf_names.push_back("reads_1");
    if (paired) S.rf_names.push_back("reads_2");
    S.finalize();
    *out = smp;
    bsq::set_status(status, BSQ_OK, "");
    return BSQ_OK;
  } catch (const BsqError& e) {
    bsq::set_status(status, e.code, e.what());
    return e.code;
  }
}
