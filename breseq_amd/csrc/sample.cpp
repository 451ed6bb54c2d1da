// sample.cpp -- bsq_sample: container, deterministic synthetic sample generator, and BAM/FASTA
// I/O on zlib (the parts of htslib breseq's pileup path relies on: sam_read1 record decoding,
// BGZF blocks, fai_fetch whole-sequence loads).
#include "sample.h"

#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <numeric>
#include <sstream>
#include <stdexcept>
#include <thread>

#include "host_util.h"

namespace bsq {

void Sample::finalize() {
  name_ptrs.clear();
  seq_ptrs.clear();
  for (size_t i = 0; i < names.size(); i++) {
    name_ptrs.push_back(names[i].c_str());
    seq_ptrs.push_back(seqs[i].c_str());
  }
  lengths.resize(seqs.size());
  for (size_t i = 0; i < seqs.size(); i++) lengths[i] = (int64_t)seqs[i].size();
  ref_view.n_targets = (int32_t)names.size();
  ref_view.names = name_ptrs.data();
  ref_view.lengths = lengths.data();
  ref_view.seqs = seq_ptrs.data();
  aln_view.n_reads = (int64_t)pos.size();
  aln_view.tid = tid.data();
  aln_view.pos = pos.data();
  aln_view.flag = flag.data();
  aln_view.mapq = mapq.data();
  aln_view.cigar_off = cigar_off.data();
  aln_view.cigar = cigar.data();
  aln_view.seq_off = seq_off.data();
  aln_view.seq = seq.data();
  aln_view.qual = qual.data();
  aln_view.x1 = x1.data();
  aln_view.xl = xl.data();
  aln_view.xr = xr.data();
  aln_view.rg = rg.data();
  const size_t nr = pos.size();
  if (tlen.size() != nr) tlen.assign(nr, 0);
  if (xp.size() != nr) xp.assign(nr, (int8_t)-1);
  name_off.assign(nr + 1, 0);
  for (size_t i = 0; i < nr; i++) name_off[i + 1] = name_off[i] + (int64_t)(i < read_names.size() ? read_names[i].size() : 0);
  name_chars.resize((size_t)name_off[nr]);
  for (size_t i = 0; i < nr && i < read_names.size(); i++)
    if (!read_names[i].empty()) memcpy(&name_chars[(size_t)name_off[i]], read_names[i].data(), read_names[i].size());
  aln_view.tlen = tlen.data();
  aln_view.xp = xp.data();
  aln_view.name_off = name_off.data();
  aln_view.names = name_chars.data();
  rf_name_ptrs.clear();
  for (auto& n : rf_names) rf_name_ptrs.push_back(n.c_str());
}

/* ------------------------------------------------------------------------------------------ */
/* Deterministic RNG: xoshiro256** seeded by splitmix64; one stream per read index so the      */
/* output does not depend on thread count.                                                    */

struct Rng {
  uint64_t s[4];
  static uint64_t splitmix(uint64_t& x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  }
  explicit Rng(uint64_t seed) {
    uint64_t x = seed;
    for (int i = 0; i < 4; i++) s[i] = splitmix(x);
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  uint64_t below(uint64_t n) { return n ? next() % n : 0; }
  double normal() {   // Box-Muller
    double u1 = uniform(), u2 = uniform();
    if (u1 < 1e-300) u1 = 1e-300;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
};

static const uint8_t kBam[4] = {1, 2, 4, 8};
static const char kChr[4] = {'A', 'C', 'G', 'T'};

static inline int chr_index(char c) {
  switch (c) {
    case 'A': return 0;
    case 'C': return 1;
    case 'G': return 2;
    case 'T': return 3;
  }
  return -1;
}

struct Variant {
  int64_t pos;       // 0-based reference position (SNP: base; DEL: first deleted base; INS: base after which)
  int type;          // 0 SNP, 1 INS, 2 DEL
  std::string alt;   // SNP base / inserted bases
  int32_t del_len;
  double freq;
};

// read one fragment's aligned read of query length `len` starting at reference `start`,
// applying the carried variants; returns false if the read would run off the reference.
struct Built {
  std::vector<uint32_t> cigar;
  std::string bases;
  int64_t pos;
};

static void push_op(std::vector<uint32_t>& cig, uint32_t op, uint32_t n) {
  if (n == 0) return;
  if (!cig.empty() && (cig.back() & 0xf) == op) cig.back() += n << 4;
  else cig.push_back((n << 4) | op);
}

static bool build_read(const std::string& ref, const std::vector<Variant>& vars, const std::vector<char>& carry,
                       size_t vfirst, int64_t start, int32_t len, int32_t edge_guard, Built& out) {
  out.cigar.clear();
  out.bases.clear();
  out.pos = start;
  const int64_t L = (int64_t)ref.size();
  int64_t p = start;
  size_t vi = vfirst;
  while ((int32_t)out.bases.size() < len) {
    if (p >= L) return false;
    while (vi < vars.size() && vars[vi].pos < p) vi++;
    const Variant* v = (vi < vars.size() && vars[vi].pos == p && carry[vi]) ? &vars[vi] : nullptr;
    int32_t q = (int32_t)out.bases.size();
    bool near_edge = (q < edge_guard) || (q > len - 1 - edge_guard - 8);
    if (v && v->type == 2 && !near_edge && q > 0) {            // deletion starting at p
      push_op(out.cigar, 2, (uint32_t)v->del_len);
      p += v->del_len;
      vi++;
      continue;
    }
    char b = ref[p];
    if (v && v->type == 0) b = v->alt[0];
    push_op(out.cigar, 0, 1);
    out.bases.push_back(b);
    if (v && v->type == 1 && !near_edge) {                      // insertion after p
      int32_t room = len - (int32_t)out.bases.size();
      int32_t n = std::min<int32_t>((int32_t)v->alt.size(), room);
      if (n == (int32_t)v->alt.size() && room - n > edge_guard) {
        push_op(out.cigar, 1, (uint32_t)n);
        out.bases += v->alt;
      }
    }
    p++;
  }
  return true;
}

}  // namespace bsq

using namespace bsq;

/* ------------------------------------------------------------------------------------------ */

extern "C" int bsq_simulate(const bsq_sim_config* cfg, bsq_sample** out, bsq_status* status) {
  try {
    if (!cfg || !out) throw BsqError(BSQ_ERR_ARG, "null argument");
    if (cfg->ref_length < 1000) throw BsqError(BSQ_ERR_ARG, "ref_length must be >= 1000");
    if (cfg->read_length < 20 || cfg->read_length > 1000) throw BsqError(BSQ_ERR_ARG, "read_length out of range");
    if (cfg->qual_min < 0 || cfg->qual_max > 93 || cfg->qual_min > cfg->qual_max) throw BsqError(BSQ_ERR_ARG, "bad quality range");
    bsq_sample* smp = new bsq_sample();
    Sample& S = smp->s;
    Rng rng(cfg->seed * 0x100000001B3ULL + 17);
    // reference
    const int64_t L = cfg->ref_length;
    std::string ref(L, 'A');
    for (int64_t i = 0; i < L; i++) {
      double u = rng.uniform();
      bool gc = u < cfg->gc;
      uint64_t k = rng.next() & 1;
      ref[i] = gc ? (k ? 'G' : 'C') : (k ? 'T' : 'A');
    }
    S.names.push_back("synthetic_1");
    // planted variants: distinct positions, >= 20 bp apart, away from the ends
    std::vector<Variant> vars;
    {
      int32_t nv = cfg->n_snps + cfg->n_indels;
      std::vector<int64_t> sites;
      int guard = 0;
      while ((int32_t)sites.size() < nv && guard < nv * 100) {
        guard++;
        int64_t p = 200 + (int64_t)rng.below((uint64_t)std::max<int64_t>(1, L - 400));
        bool ok = true;
        for (int64_t q : sites) if (std::llabs(q - p) < 20) { ok = false; break; }
        if (ok) sites.push_back(p);
      }
      for (size_t i = 0; i < sites.size(); i++) {
        Variant v;
        v.pos = sites[i];
        v.del_len = 0;
        double fmin = cfg->variant_freq_min, fmax = cfg->variant_freq_max;
        if (fmin <= 0 || fmax <= 0) { fmin = fmax = 1.0; }
        v.freq = (fmin == fmax) ? fmin : std::exp(std::log(fmin) + rng.uniform() * (std::log(fmax) - std::log(fmin)));
        if ((int32_t)i < cfg->n_snps) {
          v.type = 0;
          int r0 = chr_index(ref[v.pos]);
          int alt = (r0 + 1 + (int)rng.below(3)) & 3;
          v.alt = std::string(1, kChr[alt]);
        } else {
          int32_t len = 1 + (int32_t)rng.below((uint64_t)std::max(1, cfg->indel_max));
          if (rng.below(2)) {
            v.type = 1;
            for (int32_t k = 0; k < len; k++) v.alt.push_back(kChr[rng.below(4)]);
          } else {
            v.type = 2;
            v.del_len = len;
          }
        }
        vars.push_back(v);
      }
      std::sort(vars.begin(), vars.end(), [](const Variant& a, const Variant& b) { return a.pos < b.pos; });
    }
    // quality distribution: clipped normal, discretized; cumulative table over [qmin, qmax]
    const int qmin = cfg->qual_min, qmax = cfg->qual_max;
    std::vector<double> qcdf;
    {
      std::vector<double> w;
      for (int q = qmin; q <= qmax; q++) {
        double lo = (q == qmin) ? -1e300 : (q - 0.5 - cfg->qual_mean) / cfg->qual_sd;
        double hi = (q == qmax) ? 1e300 : (q + 0.5 - cfg->qual_mean) / cfg->qual_sd;
        w.push_back(0.5 * (std::erfc(-hi / std::sqrt(2.0)) - std::erfc(-lo / std::sqrt(2.0))));
      }
      double tot = std::accumulate(w.begin(), w.end(), 0.0);
      double c = 0;
      for (double x : w) { c += x / tot; qcdf.push_back(c); }
      qcdf.back() = 1.0;
    }
    std::vector<uint64_t> err_thr(94, 0);
    for (int q = 0; q <= 93; q++) err_thr[q] = (uint64_t)(std::pow(10.0, -q / 10.0) * 4294967296.0);

    const int32_t RL = cfg->read_length;
    const bool paired = cfg->paired != 0;
    // discordant-pair junctions: two far sites and the orientation of the pairs spanning them
    struct Junction {
      int64_t a, b;
      int8_t orient;
    };
    std::vector<Junction> junctions;
    if (paired && cfg->discordant_fraction > 0 && cfg->n_junctions > 0) {
      const uint64_t span = (uint64_t)std::max<int64_t>(1, L - 2 * (int64_t)RL - 1000);
      for (int32_t j = 0; j < cfg->n_junctions; j++) {
        Junction J;
        J.a = 500 + (int64_t)rng.below(span);
        J.b = 500 + (int64_t)rng.below(span);
        J.orient = (int8_t)rng.below(3);
        junctions.push_back(J);
      }
    }
    // missing-coverage regions: no fragment overlapping one is sequenced (a deletion in the sample
    // genome); a few repeat-mapped reads (X1 = 2) reach into each region from both edges
    struct Missing {
      int64_t a, b;   // [a, b)
    };
    std::vector<Missing> missing;
    if (cfg->n_missing > 0) {
      if (cfg->missing_length < 1 || cfg->missing_length * (int64_t)cfg->n_missing > L / 4)
        throw BsqError(BSQ_ERR_ARG, "missing-coverage regions must be non-empty and cover at most a quarter of the reference");
      int guard = 0;
      while ((int32_t)missing.size() < cfg->n_missing && guard < cfg->n_missing * 100) {
        guard++;
        const int64_t a = 1000 + (int64_t)rng.below((uint64_t)std::max<int64_t>(1, L - 2000 - cfg->missing_length));
        const int64_t b = a + cfg->missing_length;
        bool ok = true;
        for (const Missing& m : missing)
          if (a < m.b + 2000 && m.a < b + 2000) ok = false;
        if (ok) missing.push_back({a, b});
      }
      std::sort(missing.begin(), missing.end(), [](const Missing& x, const Missing& y) { return x.a < y.a; });
    }
    auto in_missing = [&](int64_t a, int64_t b) {   // [a, b) overlaps a missing region
      for (const Missing& m : missing)
        if (a < m.b && m.a < b) return true;
      return false;
    };
    const double cov = cfg->coverage;
    int64_t n_frag = (int64_t)std::llround(cov * (double)L / (double)(RL * (paired ? 2 : 1)));
    if (n_frag < 1) n_frag = 1;

    // Reads of a range of fragments, in fragment order, in flat buffers.  Each fragment draws from
    // its own generator, so the reads do not depend on how the fragments are split over threads.
    struct Reads {
      std::vector<int64_t> pos, mate_pos;
      std::vector<uint32_t> cig_off{0}, cig;
      std::vector<char> bases;
      std::vector<uint8_t> q;
      std::vector<uint16_t> flag;
      std::vector<uint8_t> mapq;
      std::vector<int32_t> x1, tlen, file;
      std::vector<int8_t> xp;
      void add(const Built& B, int32_t RL_, uint16_t fl, uint8_t mq, int32_t x1v, int64_t mpos, int32_t tl, int32_t f,
               int8_t x) {
        pos.push_back(B.pos);
        cig.insert(cig.end(), B.cigar.begin(), B.cigar.end());
        cig_off.push_back((uint32_t)cig.size());
        bases.insert(bases.end(), B.bases.begin(), B.bases.begin() + RL_);
        q.resize(q.size() + (size_t)RL_);
        flag.push_back(fl);
        mapq.push_back(mq);
        x1.push_back(x1v);
        mate_pos.push_back(mpos);
        tlen.push_back(tl);
        file.push_back(f);
        xp.push_back(x);
      }
      size_t size() const { return pos.size(); }
    };
    // per-base qualities and sequencing errors of the last added read, from generator fr
    auto finish_read = [&](Reads& R, Rng& fr) {
      char* bs = R.bases.data() + R.bases.size() - (size_t)RL;
      uint8_t* qs = R.q.data() + R.q.size() - (size_t)RL;
      for (int32_t i = 0; i < RL; i++) {
        uint64_t u = fr.next();
        double uq = (double)(u >> 40) * (1.0 / 16777216.0);
        int q = qmin + (int)(std::lower_bound(qcdf.begin(), qcdf.end(), uq) - qcdf.begin());
        if (q > qmax) q = qmax;
        qs[i] = (uint8_t)q;
        uint64_t ue = u & 0xFFFFFFFFULL;
        if (ue < err_thr[q]) {
          int c = chr_index(bs[i]);
          int alt = (c + 1 + (int)((u >> 32) % 3)) & 3;
          bs[i] = kChr[alt];
        }
        if ((u >> 24 & 0xFFFF) == 7) {
          bs[i] = 'N';
          qs[i] = (uint8_t)std::max(qmin, 2);
        }
      }
    };
    auto first_var = [&](int64_t p) {
      return (size_t)(std::lower_bound(vars.begin(), vars.end(), p, [](const Variant& v, int64_t x) { return v.pos < x; }) - vars.begin());
    };
    auto ref_end = [](const Built& B) {
      int64_t e = B.pos;
      for (uint32_t c : B.cigar)
        if ((c & 0xf) == 0 || (c & 0xf) == 2) e += c >> 4;
      return e;
    };
    const int n_chunks = (int)std::max<int64_t>(1, std::min<int64_t>(host_threads() * 4, n_frag / 1024 + 1));
    std::vector<Reads> chunks((size_t)n_chunks);
    parallel_ranges(n_chunks, host_threads(), [&](int, int64_t c0, int64_t c1) {
      std::vector<char> carry(vars.size(), 0);
      Built b1, b2;
      for (int64_t ch = c0; ch < c1; ch++) {
        Reads& R = chunks[(size_t)ch];
        const int64_t f0 = n_frag * ch / n_chunks, f1 = n_frag * (ch + 1) / n_chunks;
        for (int64_t f = f0; f < f1; f++) {
          Rng fr(cfg->seed ^ (0xA0761D6478BD642FULL * (uint64_t)(f + 1)));
          int64_t flen = RL;
          if (paired) {
            double x = cfg->insert_mean + cfg->insert_sd * fr.normal();
            flen = (int64_t)std::llround(x);
            if (flen < RL) flen = RL;
          }
          if (flen + 16 >= L) continue;
          int64_t start = (int64_t)fr.below((uint64_t)(L - flen - 8));
          bool frag_rev = (fr.next() & 1) != 0;
          const int32_t guard = 4;
          int n_mates = paired ? 2 : 1;
          // discordant pair: R1 just before junction site a, R2 just after site b
          const Junction* J = nullptr;
          if (!junctions.empty() && fr.uniform() < cfg->discordant_fraction) J = &junctions[fr.below(junctions.size())];
          bool ok;
          bool rev0 = frag_rev, rev1 = !frag_rev;
          if (J) {
            const uint64_t spread = (uint64_t)std::max<int64_t>(1, (int64_t)cfg->insert_mean - RL);
            const int64_t hi = L - RL - 64;
            const int64_t s1 = std::min<int64_t>(hi, std::max<int64_t>(0, J->a - RL - (int64_t)fr.below(spread)));
            const int64_t s2 = std::min<int64_t>(hi, std::max<int64_t>(0, J->b + (int64_t)fr.below(spread)));
            if (J->orient == 0) { rev0 = false; rev1 = true; }
            else if (J->orient == 1) { rev0 = true; rev1 = false; }
            else { rev0 = rev1 = (fr.next() & 1) != 0; }
            const size_t v1 = first_var(s1), v2 = first_var(s2);
            for (size_t vi = v1; vi < vars.size() && vars[vi].pos < s1 + RL + 16; vi++) carry[vi] = fr.uniform() < vars[vi].freq;
            ok = build_read(ref, vars, carry, v1, s1, RL, guard, b1);
            for (size_t vi = v1; vi < vars.size() && vars[vi].pos < s1 + RL + 16; vi++) carry[vi] = 0;
            for (size_t vi = v2; vi < vars.size() && vars[vi].pos < s2 + RL + 16; vi++) carry[vi] = fr.uniform() < vars[vi].freq;
            ok = ok && build_read(ref, vars, carry, v2, s2, RL, guard, b2);
            for (size_t vi = v2; vi < vars.size() && vars[vi].pos < s2 + RL + 16; vi++) carry[vi] = 0;
          } else {
            // haplotype: each variant in the fragment carried with its frequency
            size_t vfirst = first_var(start);
            for (size_t vi = vfirst; vi < vars.size() && vars[vi].pos < start + flen + 16; vi++) carry[vi] = fr.uniform() < vars[vi].freq;
            ok = build_read(ref, vars, carry, vfirst, start, RL, guard, b1);
            if (ok && paired) {
              // mate 2 ends at start + flen - 1 on the reference (approximately, before indels)
              int64_t s2 = start + flen - RL;
              ok = build_read(ref, vars, carry, first_var(s2), s2, RL, guard, b2);
            }
            for (size_t vi = vfirst; vi < vars.size() && vars[vi].pos < start + flen + 16; vi++) carry[vi] = 0;
          }
          if (!ok) continue;
          if (!missing.empty()) {
            if (in_missing(b1.pos, ref_end(b1)) || (paired && in_missing(b2.pos, ref_end(b2)))) continue;
            if (!J && in_missing(start, start + flen)) continue;
          }
          // TLEN as mark_pair_info sets it: +distance on the leftmost mate, -distance on the other
          const int64_t dist = paired ? (J ? std::max(ref_end(b1), ref_end(b2)) - std::min(b1.pos, b2.pos) : flen) : 0;
          const bool m0_left = !paired || b1.pos <= b2.pos;
          bool redundant = fr.uniform() < cfg->redundant_fraction;
          int32_t x1v = redundant ? (2 + (int32_t)fr.below(2)) : 1;
          for (int m = 0; m < n_mates; m++) {
            const Built& B = (m == 0) ? b1 : b2;
            bool rev = paired ? ((m == 0) ? rev0 : rev1) : frag_rev;
            uint16_t flag = rev ? 0x10 : 0;
            if (paired) {
              flag |= 0x1 | (J ? 0 : 0x2) | ((m == 0) ? 0x40 : 0x80);
              bool mrev = (m == 0) ? rev1 : rev0;
              if (mrev) flag |= 0x20;
            }
            R.add(B, RL, flag, redundant ? (uint8_t)1 : (uint8_t)cfg->mapq, x1v,
                  paired ? ((m == 0) ? b2.pos : b1.pos) : -1, paired ? (int32_t)(((m == 0) == m0_left) ? dist : -dist) : 0,
                  paired ? m : 0, paired ? (J ? J->orient : (int8_t)0) : (int8_t)-1);
            finish_read(R, fr);
          }
        }
      }
    });
    // the repeat-mapped reads at the missing regions' edges: unpaired, X1 = 2, reference bases
    Reads edge;
    for (size_t k = 0; k < missing.size(); k++) {
      Rng er(cfg->seed ^ (0x9FB21C651E98DF25ULL * (uint64_t)(k + 1)));
      for (int32_t j = 0; j < 2 * cfg->missing_edge_reads; j++) {
        const bool left = j < cfg->missing_edge_reads;
        const int64_t reach = 10 + (int64_t)er.below(31);   // bases inside the region
        int64_t s = left ? missing[k].a + reach - RL : missing[k].b - reach;
        s = std::max<int64_t>(0, std::min<int64_t>(s, L - RL));
        Built B;
        B.pos = s;
        B.cigar.assign(1, ((uint32_t)RL << 4) | 0u);
        B.bases.assign(ref, (size_t)s, (size_t)RL);
        edge.add(B, RL, (er.next() & 1) ? 0x10 : 0, 1, 2, -1, 0, 0, (int8_t)-1);
        finish_read(edge, er);
      }
    }
    chunks.push_back(std::move(edge));
    // concatenation in fragment order, then a stable counting sort by position
    std::vector<int64_t> chunk_base(chunks.size() + 1, 0);
    for (size_t c = 0; c < chunks.size(); c++) chunk_base[c + 1] = chunk_base[c] + (int64_t)chunks[c].size();
    const size_t nr = (size_t)chunk_base.back();
    std::vector<int32_t> src_chunk(nr), src_idx(nr);
    std::vector<int64_t> bucket((size_t)L + 1, 0);
    for (size_t c = 0; c < chunks.size(); c++)
      for (int64_t p : chunks[c].pos) bucket[(size_t)p + 1]++;
    for (int64_t p = 0; p < L; p++) bucket[(size_t)p + 1] += bucket[(size_t)p];
    std::vector<int64_t> order_of(nr);   // file-order read index -> generation order
    for (size_t c = 0; c < chunks.size(); c++)
      for (size_t i = 0; i < chunks[c].size(); i++) {
        const int64_t dst = bucket[(size_t)chunks[c].pos[i]]++;
        src_chunk[(size_t)dst] = (int32_t)c;
        src_idx[(size_t)dst] = (int32_t)i;
        order_of[(size_t)dst] = chunk_base[c] + (int64_t)i;
      }
    // trims for XL/XR (calculate_trims.cpp:239-285 + alignment.cpp:920-946)
    std::vector<uint8_t> gtrims(2 * (size_t)L);
    host_sequence_trims(ref.data(), L, gtrims.data());
    S.seqs.push_back(ref);
    S.tid.assign(nr, 0);
    S.pos.resize(nr);
    S.flag.resize(nr);
    S.mapq.resize(nr);
    S.x1.resize(nr);
    S.xl.resize(nr);
    S.xr.resize(nr);
    S.rg.assign(nr, 0);
    S.mate_tid.assign(nr, paired ? 0 : -1);
    S.mate_pos.resize(nr);
    S.tlen.resize(nr);
    S.xp.resize(nr);
    S.read_names.resize(nr);
    S.cigar_off.assign(nr + 1, 0);
    S.seq_off.assign(nr + 1, 0);
    for (size_t i = 0; i < nr; i++) {
      const Reads& R = chunks[(size_t)src_chunk[i]];
      const size_t k = (size_t)src_idx[i];
      S.cigar_off[i + 1] = S.cigar_off[i] + (int64_t)(R.cig_off[k + 1] - R.cig_off[k]);
      S.seq_off[i + 1] = S.seq_off[i] + RL;
    }
    S.cigar.resize((size_t)S.cigar_off[nr]);
    S.seq.resize((size_t)S.seq_off[nr]);
    S.qual.resize((size_t)S.seq_off[nr]);
    parallel_ranges((int64_t)nr, host_threads(), [&](int, int64_t i0, int64_t i1) {
      std::string bases;
      for (int64_t i = i0; i < i1; i++) {
        const Reads& R = chunks[(size_t)src_chunk[(size_t)i]];
        const size_t k = (size_t)src_idx[(size_t)i];
        S.pos[i] = (int32_t)R.pos[k];
        S.flag[i] = R.flag[k];
        S.mapq[i] = R.mapq[k];
        S.x1[i] = R.x1[k];
        S.mate_pos[i] = (int32_t)R.mate_pos[k];
        S.tlen[i] = R.tlen[k];
        S.xp[i] = R.xp[k];
        S.read_names[i] = std::to_string(R.file[k] + 1) + ":" + std::to_string(order_of[(size_t)i] / (paired ? 2 : 1) + 1);
        std::copy(R.cig.begin() + R.cig_off[k], R.cig.begin() + R.cig_off[k + 1], S.cigar.begin() + S.cigar_off[i]);
        bases.assign(R.bases.data() + k * (size_t)RL, (size_t)RL);
        for (int32_t j = 0; j < RL; j++) {
          int c = chr_index(bases[j]);
          S.seq[(size_t)S.seq_off[i] + j] = c < 0 ? 15 : kBam[c];
          S.qual[(size_t)S.seq_off[i] + j] = R.q[k * (size_t)RL + j];
        }
        // alignment trims: max(genomic, read-edge) + query offset
        int64_t rend = R.pos[k];
        for (uint32_t c = R.cig_off[k]; c < R.cig_off[k + 1]; c++) {
          uint32_t op = R.cig[c] & 0xf;
          if (op == 0 || op == 2 || op == 3 || op == 7 || op == 8) rend += R.cig[c] >> 4;
        }
        uint32_t etl = 1, etr = 1;
        host_edge_trims(bases, etl, etr);
        uint32_t gl = gtrims[R.pos[k]], gr = gtrims[L + (rend - 1)];
        S.xl[i] = (int32_t)std::max(gl, etl);      // query_start_0 == 0 (no soft clips)
        S.xr[i] = (int32_t)std::max(gr, etr);      // read_length - query_end_1 == 0
      }
    });
    S.rg_ids.push_back("0");
    S.rg_libs.push_back(paired ? "reads" : "reads_1");
    S.rf_base.push_back(0);
    S.rf_count.push_back(paired ? 2 : 1);
    S.rf_names.push_back("reads_1");
    if (paired) S.rf_names.push_back("reads_2");
    S.finalize();
    *out = smp;
    set_status(status, BSQ_OK, "");
    return BSQ_OK;
  } catch (const BsqError& e) {
    set_status(status, e.code, e.what());
    return e.code;
  } catch (const std::exception& e) {
    set_status(status, BSQ_ERR_ARG, e.what());
    return BSQ_ERR_ARG;
  }
}

/* ------------------------------------------------------------------------------------------ */
/* FASTA                                                                                       */

static void read_fasta(const std::string& path, Sample& S) {
  std::ifstream in(path);
  if (!in) throw BsqError(BSQ_ERR_IO, "could not open FASTA " + path);
  std::string line;
  while (std::getline(in, line)) {
    if (!line.empty() && line.back() == '\r') line.pop_back();
    if (line.empty()) continue;
    if (line[0] == '>') {
      std::string name = line.substr(1);
      size_t sp = name.find_first_of(" \t");
      if (sp != std::string::npos) name = name.substr(0, sp);
      S.names.push_back(name);
      S.seqs.emplace_back();
    } else {
      if (S.seqs.empty()) throw BsqError(BSQ_ERR_FORMAT, "FASTA sequence before header");
      S.seqs.back() += line;
    }
  }
}

extern "C" int bsq_sample_write_fasta(const bsq_sample* s, const char* path, bsq_status* status) {
  try {
    std::ofstream out(path);
    if (!out) throw BsqError(BSQ_ERR_IO, std::string("could not write ") + path);
    for (size_t i = 0; i < s->s.names.size(); i++) {
      out << ">" << s->s.names[i] << "\n";
      const std::string& q = s->s.seqs[i];
      for (size_t p = 0; p < q.size(); p += 60) out << q.substr(p, 60) << "\n";
    }
    set_status(status, BSQ_OK, "");
    return BSQ_OK;
  } catch (const BsqError& e) {
    set_status(status, e.code, e.what());
    return e.code;
  }
}

/* ------------------------------------------------------------------------------------------ */
/* BAM (SAMv1 section 4): BGZF blocks of raw-deflate data with a BC extra field.              */

namespace {

struct BgzfWriter {
  FILE* f = nullptr;
  std::string buf;
  explicit BgzfWriter(const char* path) {
    f = fopen(path, "wb");
    if (!f) throw BsqError(BSQ_ERR_IO, std::string("could not write ") + path);
  }
  void block(const char* data, size_t n) {
    std::vector<unsigned char> comp(n + 1024);
    z_stream zs;
    memset(&zs, 0, sizeof(zs));
    if (deflateInit2(&zs, 6, Z_DEFLATED, -15, 8, Z_DEFAULT_STRATEGY) != Z_OK) throw BsqError(BSQ_ERR_IO, "deflateInit2");
    zs.next_in = (Bytef*)data;
    zs.avail_in = (uInt)n;
    zs.next_out = comp.data();
    zs.avail_out = (uInt)comp.size();
    if (deflate(&zs, Z_FINISH) != Z_STREAM_END) throw BsqError(BSQ_ERR_IO, "deflate");
    size_t clen = zs.total_out;
    deflateEnd(&zs);
    uint32_t crc = (uint32_t)crc32(0L, (const Bytef*)data, (uInt)n);
    size_t bsize = 18 + clen + 8;
    unsigned char hdr[18] = {31, 139, 8, 4, 0, 0, 0, 0, 0, 255, 6, 0, 'B', 'C', 2, 0, 0, 0};
    hdr[16] = (unsigned char)((bsize - 1) & 0xff);
    hdr[17] = (unsigned char)(((bsize - 1) >> 8) & 0xff);
    fwrite(hdr, 1, 18, f);
    fwrite(comp.data(), 1, clen, f);
    unsigned char tail[8];
    for (int i = 0; i < 4; i++) tail[i] = (unsigned char)(crc >> (8 * i));
    for (int i = 0; i < 4; i++) tail[4 + i] = (unsigned char)(n >> (8 * i));
    fwrite(tail, 1, 8, f);
  }
  void write(const void* p, size_t n) {
    buf.append((const char*)p, n);
    while (buf.size() >= 65280) {
      block(buf.data(), 65280);
      buf.erase(0, 65280);
    }
  }
  void close() {
    if (!buf.empty()) block(buf.data(), buf.size());
    buf.clear();
    block("", 0);   // EOF marker block
    fclose(f);
    f = nullptr;
  }
};

template <class T>
void put(std::string& s, T v) {
  s.append((const char*)&v, sizeof(T));
}

int reg2bin(int64_t beg, int64_t end) {
  --end;
  if (beg >> 14 == end >> 14) return (int)(((1 << 15) - 1) / 7 + (beg >> 14));
  if (beg >> 17 == end >> 17) return (int)(((1 << 12) - 1) / 7 + (beg >> 17));
  if (beg >> 20 == end >> 20) return (int)(((1 << 9) - 1) / 7 + (beg >> 20));
  if (beg >> 23 == end >> 23) return (int)(((1 << 6) - 1) / 7 + (beg >> 23));
  if (beg >> 26 == end >> 26) return (int)(((1 << 3) - 1) / 7 + (beg >> 26));
  return 0;
}

void put_int_tag(std::string& s, const char* tag, int32_t v) {
  s.append(tag, 2);
  if (v >= 0 && v <= 255) { s.push_back('C'); put<uint8_t>(s, (uint8_t)v); }
  else if (v >= 0 && v <= 65535) { s.push_back('S'); put<uint16_t>(s, (uint16_t)v); }
  else { s.push_back('i'); put<int32_t>(s, v); }
}

}  // namespace

extern "C" int bsq_sample_write_bam(const bsq_sample* smp, const char* path, bsq_status* status) {
  try {
    const Sample& S = smp->s;
    BgzfWriter w(path);
    std::string text = "@HD\tVN:1.6\tSO:coordinate\n";
    for (size_t i = 0; i < S.names.size(); i++) text += "@SQ\tSN:" + S.names[i] + "\tLN:" + std::to_string(S.seqs[i].size()) + "\n";
    for (size_t i = 0; i < S.rg_ids.size(); i++) text += "@RG\tID:" + S.rg_ids[i] + "\tLB:" + S.rg_libs[i] + "\n";
    std::string h = "BAM\1";
    put<int32_t>(h, (int32_t)text.size());
    h += text;
    put<int32_t>(h, (int32_t)S.names.size());
    for (size_t i = 0; i < S.names.size(); i++) {
      put<int32_t>(h, (int32_t)S.names[i].size() + 1);
      h += S.names[i];
      h.push_back('\0');
      put<int32_t>(h, (int32_t)S.seqs[i].size());
    }
    w.write(h.data(), h.size());
    std::string rec;
    for (int64_t r = 0; r < S.n_reads(); r++) {
      rec.clear();
      std::string name = (r < (int64_t)S.read_names.size() && !S.read_names[r].empty()) ? S.read_names[r] : ("r" + std::to_string(r));
      int64_t c0 = S.cigar_off[r], c1 = S.cigar_off[r + 1];
      int64_t s0 = S.seq_off[r], s1 = S.seq_off[r + 1];
      int32_t lseq = (int32_t)(s1 - s0);
      int64_t end = S.pos[r];
      for (int64_t k = c0; k < c1; k++) {
        uint32_t op = S.cigar[k] & 0xf;
        if (op == 0 || op == 2 || op == 3 || op == 7 || op == 8) end += S.cigar[k] >> 4;
      }
      if (end == S.pos[r]) end++;
      put<int32_t>(rec, 0);   // block_size placeholder
      put<int32_t>(rec, S.tid[r]);
      put<int32_t>(rec, S.pos[r]);
      put<uint8_t>(rec, (uint8_t)(name.size() + 1));
      put<uint8_t>(rec, S.mapq[r]);
      put<uint16_t>(rec, (uint16_t)reg2bin(S.pos[r], end));
      put<uint16_t>(rec, (uint16_t)(c1 - c0));
      put<uint16_t>(rec, S.flag[r]);
      put<int32_t>(rec, lseq);
      put<int32_t>(rec, r < (int64_t)S.mate_tid.size() ? S.mate_tid[r] : -1);
      put<int32_t>(rec, r < (int64_t)S.mate_pos.size() ? S.mate_pos[r] : -1);
      put<int32_t>(rec, r < (int64_t)S.tlen.size() ? S.tlen[r] : 0);
      rec += name;
      rec.push_back('\0');
      for (int64_t k = c0; k < c1; k++) put<uint32_t>(rec, S.cigar[k]);
      for (int32_t i = 0; i < lseq; i += 2) {
        uint8_t hi = S.seq[s0 + i] & 0xf;
        uint8_t lo = (i + 1 < lseq) ? (S.seq[s0 + i + 1] & 0xf) : 0;
        rec.push_back((char)((hi << 4) | lo));
      }
      rec.append((const char*)&S.qual[s0], (size_t)lseq);
      if (S.x1[r] >= 0) put_int_tag(rec, "X1", S.x1[r]);
      if (S.xl[r] >= 0) put_int_tag(rec, "XL", S.xl[r]);
      if (S.xr[r] >= 0) put_int_tag(rec, "XR", S.xr[r]);
      if (r < (int64_t)S.xp.size() && S.xp[r] >= 0 && S.xp[r] <= 2) {
        static const char* kXp[3] = {"FR", "RF", "FF"};
        rec += "XPZ";
        rec += kXp[S.xp[r]];
        rec.push_back('\0');
      }
      if (S.rg[r] >= 0 && S.rg[r] < (int32_t)S.rg_ids.size()) {
        rec += "RGZ";
        rec += S.rg_ids[S.rg[r]];
        rec.push_back('\0');
      }
      int32_t bs = (int32_t)rec.size() - 4;
      memcpy(&rec[0], &bs, 4);
      w.write(rec.data(), rec.size());
    }
    w.close();
    set_status(status, BSQ_OK, "");
    return BSQ_OK;
  } catch (const BsqError& e) {
    set_status(status, e.code, e.what());
    return e.code;
  }
}

namespace {

struct Reader {
  const unsigned char* p;
  const unsigned char* e;
  template <class T>
  T get() {
    if ((size_t)(e - p) < sizeof(T)) throw BsqError(BSQ_ERR_FORMAT, "truncated BAM");
    T v;
    memcpy(&v, p, sizeof(T));
    p += sizeof(T);
    return v;
  }
  void need(size_t n) const {
    if ((size_t)(e - p) < n) throw BsqError(BSQ_ERR_FORMAT, "truncated BAM");
  }
};

// integer value of an aux tag payload of the given type; advances past it
bool aux_int(char t, const unsigned char* v, int64_t& out) {
  switch (t) {
    case 'c': out = (int8_t)v[0]; return true;
    case 'C': out = v[0]; return true;
    case 's': { int16_t x; memcpy(&x, v, 2); out = x; return true; }
    case 'S': { uint16_t x; memcpy(&x, v, 2); out = x; return true; }
    case 'i': { int32_t x; memcpy(&x, v, 4); out = x; return true; }
    case 'I': { uint32_t x; memcpy(&x, v, 4); out = x; return true; }
  }
  return false;
}

size_t aux_size(char t, const unsigned char* v, const unsigned char* end) {
  switch (t) {
    case 'A': case 'c': case 'C': return 1;
    case 's': case 'S': return 2;
    case 'i': case 'I': case 'f': return 4;
    case 'Z': case 'H': {
      const unsigned char* q = v;
      while (q < end && *q) q++;
      if (q >= end) throw BsqError(BSQ_ERR_FORMAT, "unterminated aux string");
      return (size_t)(q - v) + 1;
    }
    case 'B': {
      if (end - v < 5) throw BsqError(BSQ_ERR_FORMAT, "truncated aux array");
      char st = (char)v[0];
      uint32_t n;
      memcpy(&n, v + 1, 4);
      size_t es = (st == 'c' || st == 'C') ? 1 : (st == 's' || st == 'S') ? 2 : 4;
      return 5 + es * n;
    }
  }
  throw BsqError(BSQ_ERR_FORMAT, std::string("unknown aux type ") + t);
}

std::vector<unsigned char> read_file_bytes(const char* path) {
  FILE* f = fopen(path, "rb");
  if (!f) throw BsqError(BSQ_ERR_IO, std::string("could not open ") + path);
  std::vector<unsigned char> buf;
  if (fseek(f, 0, SEEK_END) == 0) {
    const long sz = ftell(f);
    if (sz > 0) buf.resize((size_t)sz);
    fseek(f, 0, SEEK_SET);
  }
  const size_t got = buf.empty() ? 0 : fread(buf.data(), 1, buf.size(), f);
  fclose(f);
  if (got != buf.size()) throw BsqError(BSQ_ERR_IO, std::string("short read of ") + path);
  return buf;
}

// BGZF (SAMv1 section 4.1): a series of gzip members, each with a BC extra subfield holding the
// member size.  Members are independent deflate streams, so they are inflated in parallel into
// their offsets in the output (prefix sum of the ISIZE trailers), each one's CRC32 checked.
// Returns false when the file is not BGZF-framed (the caller then reads it as one gzip stream).
bool bgzf_inflate_parallel(const std::vector<unsigned char>& raw, std::vector<unsigned char>& out) {
  struct Block {
    size_t cdata, clen, uoff;
    uint32_t usize, crc;
  };
  std::vector<Block> blocks;
  size_t p = 0, uoff = 0;
  while (p < raw.size()) {
    if (raw.size() - p < 18 || raw[p] != 31 || raw[p + 1] != 139 || raw[p + 2] != 8 || !(raw[p + 3] & 4)) return false;
    const uint16_t xlen = (uint16_t)(raw[p + 10] | (raw[p + 11] << 8));
    size_t x = p + 12;
    const size_t xe = x + xlen;
    if (xe > raw.size()) return false;
    int bsize = -1;
    while (x + 4 <= xe) {
      const uint16_t slen = (uint16_t)(raw[x + 2] | (raw[x + 3] << 8));
      if (raw[x] == 'B' && raw[x + 1] == 'C' && slen == 2 && x + 6 <= xe) bsize = raw[x + 4] | (raw[x + 5] << 8);
      x += 4 + (size_t)slen;
    }
    if (bsize < 0) return false;
    const size_t end = p + (size_t)bsize + 1;
    if (end > raw.size() || end < xe + 8) return false;
    Block b;
    b.cdata = xe;
    b.clen = end - 8 - xe;
    memcpy(&b.crc, &raw[end - 8], 4);
    memcpy(&b.usize, &raw[end - 4], 4);
    b.uoff = uoff;
    uoff += b.usize;
    blocks.push_back(b);
    p = end;
  }
  out.resize(uoff);
  parallel_ranges((int64_t)blocks.size(), host_threads(), [&](int, int64_t b0, int64_t b1) {
    z_stream zs;
    memset(&zs, 0, sizeof(zs));
    if (inflateInit2(&zs, -15) != Z_OK) throw BsqError(BSQ_ERR_FORMAT, "zlib initialisation failed");
    for (int64_t i = b0; i < b1; i++) {
      const Block& b = blocks[(size_t)i];
      inflateReset(&zs);
      zs.next_in = const_cast<unsigned char*>(&raw[b.cdata]);
      zs.avail_in = (uInt)b.clen;
      zs.next_out = out.data() + b.uoff;
      zs.avail_out = b.usize;
      const int rc = inflate(&zs, Z_FINISH);
      if (rc != Z_STREAM_END || zs.total_out != b.usize) {
        inflateEnd(&zs);
        throw BsqError(BSQ_ERR_FORMAT, "BGZF block does not inflate to its recorded size");
      }
      if ((uint32_t)crc32(0L, out.data() + b.uoff, b.usize) != b.crc) {
        inflateEnd(&zs);
        throw BsqError(BSQ_ERR_FORMAT, "BGZF block CRC mismatch");
      }
    }
    inflateEnd(&zs);
  });
  return true;
}

std::vector<unsigned char> read_gzip_stream(const char* path) {
  gzFile gz = gzopen(path, "rb");
  if (!gz) throw BsqError(BSQ_ERR_IO, std::string("could not open BAM ") + path);
  std::vector<unsigned char> data;
  std::vector<unsigned char> chunk(1 << 22);
  for (;;) {
    const int n = gzread(gz, chunk.data(), (unsigned)chunk.size());
    if (n < 0) {
      gzclose(gz);
      throw BsqError(BSQ_ERR_FORMAT, "BGZF decompression failed");
    }
    if (n == 0) break;
    data.insert(data.end(), chunk.begin(), chunk.begin() + n);
  }
  gzclose(gz);
  return data;
}

}  // namespace

extern "C" int bsq_sample_read(const char* bam_path, const char* fasta_path, bsq_sample** out, bsq_status* status) {
  bsq_sample* smp = nullptr;
  try {
    smp = new bsq_sample();
    Sample& S = smp->s;
    read_fasta(fasta_path, S);
    std::vector<unsigned char> data;
    {
      const std::vector<unsigned char> raw = read_file_bytes(bam_path);
      if (!bgzf_inflate_parallel(raw, data)) data = read_gzip_stream(bam_path);
    }
    Reader R{data.data(), data.data() + data.size()};
    R.need(4);
    if (memcmp(R.p, "BAM\1", 4) != 0) throw BsqError(BSQ_ERR_FORMAT, "not a BAM file");
    R.p += 4;
    int32_t l_text = R.get<int32_t>();
    R.need((size_t)l_text);
    std::string text((const char*)R.p, (size_t)l_text);
    R.p += l_text;
    // @RG lines
    {
      std::istringstream ts(text);
      std::string line;
      while (std::getline(ts, line)) {
        if (line.compare(0, 3, "@RG") != 0) continue;
        std::string id, lb;
        std::istringstream ls(line);
        std::string f;
        while (std::getline(ls, f, '\t')) {
          if (f.compare(0, 3, "ID:") == 0) id = f.substr(3);
          if (f.compare(0, 3, "LB:") == 0) lb = f.substr(3);
        }
        S.rg_ids.push_back(id);
        S.rg_libs.push_back(lb);
      }
    }
    int32_t n_ref = R.get<int32_t>();
    std::vector<std::string> bam_names;
    std::vector<int64_t> bam_lens;
    for (int32_t i = 0; i < n_ref; i++) {
      int32_t ln = R.get<int32_t>();
      R.need((size_t)ln);
      bam_names.emplace_back((const char*)R.p, strnlen((const char*)R.p, (size_t)ln));
      R.p += ln;
      bam_lens.push_back(R.get<int32_t>());
    }
    // order FASTA sequences as the BAM header targets (pileup_base.cpp:80-86)
    {
      std::map<std::string, size_t> idx;
      for (size_t i = 0; i < S.names.size(); i++) idx[S.names[i]] = i;
      std::vector<std::string> seqs;
      for (int32_t i = 0; i < n_ref; i++) {
        auto it = idx.find(bam_names[i]);
        if (it == idx.end()) throw BsqError(BSQ_ERR_FORMAT, "FASTA lacks target " + bam_names[i]);
        if ((int64_t)S.seqs[it->second].size() != bam_lens[i]) throw BsqError(BSQ_ERR_FORMAT, "FASTA length differs from BAM header for " + bam_names[i]);
        seqs.push_back(S.seqs[it->second]);
      }
      S.names = bam_names;
      S.seqs = seqs;
    }
    // record boundaries and the CIGAR / sequence offsets of every record (sequential, cheap)
    std::vector<size_t> rec_at;
    S.cigar_off.assign(1, 0);
    S.seq_off.assign(1, 0);
    while (R.p < R.e) {
      const int32_t bs = R.get<int32_t>();
      if (bs < 32) throw BsqError(BSQ_ERR_FORMAT, "BAM record shorter than its fixed fields");
      R.need((size_t)bs);
      uint16_t nc;
      int32_t lseq;
      memcpy(&nc, R.p + 12, 2);
      memcpy(&lseq, R.p + 16, 4);
      if (lseq < 0) throw BsqError(BSQ_ERR_FORMAT, "negative BAM sequence length");
      rec_at.push_back((size_t)(R.p - data.data()));
      S.cigar_off.push_back(S.cigar_off.back() + nc);
      S.seq_off.push_back(S.seq_off.back() + lseq);
      R.p += bs;
    }
    const size_t nr = rec_at.size();
    S.tid.resize(nr);
    S.pos.resize(nr);
    S.flag.resize(nr);
    S.mapq.resize(nr);
    S.mate_tid.resize(nr);
    S.mate_pos.resize(nr);
    S.tlen.resize(nr);
    S.xp.resize(nr);
    S.x1.resize(nr);
    S.xl.resize(nr);
    S.xr.resize(nr);
    S.rg.resize(nr);
    S.read_names.resize(nr);
    S.cigar.resize((size_t)S.cigar_off.back());
    S.seq.resize((size_t)S.seq_off.back());
    S.qual.resize((size_t)S.seq_off.back());
    // decode the records in parallel, each into its own slots
    parallel_ranges((int64_t)nr, host_threads(), [&](int, int64_t r0, int64_t r1) {
      for (int64_t r = r0; r < r1; r++) {
        const size_t ri = (size_t)r;
        const unsigned char* rec = data.data() + rec_at[ri];
        int32_t bs;
        memcpy(&bs, rec - 4, 4);
        const unsigned char* rec_end = rec + bs;
        Reader Q{rec, rec_end};
        const int32_t tid = Q.get<int32_t>();
        const int32_t pos = Q.get<int32_t>();
        const uint8_t lrn = Q.get<uint8_t>();
        const uint8_t mq = Q.get<uint8_t>();
        (void)Q.get<uint16_t>();
        const uint16_t nc = Q.get<uint16_t>();
        const uint16_t flag = Q.get<uint16_t>();
        const int32_t lseq = Q.get<int32_t>();
        const int32_t mtid = Q.get<int32_t>();
        const int32_t mpos = Q.get<int32_t>();
        const int32_t tlen = Q.get<int32_t>();
        Q.need(lrn);
        S.read_names[ri].assign((const char*)Q.p, strnlen((const char*)Q.p, lrn));
        Q.p += lrn;
        Q.need((size_t)nc * 4);
        if (nc) memcpy(&S.cigar[(size_t)S.cigar_off[ri]], Q.p, (size_t)nc * 4);
        Q.p += (size_t)nc * 4;
        Q.need((size_t)(lseq + 1) / 2 + (size_t)lseq);
        if (lseq) {
          uint8_t* sq = &S.seq[(size_t)S.seq_off[ri]];
          for (int32_t i = 0; i < lseq; i++) {
            const uint8_t byte = Q.p[i >> 1];
            sq[i] = (i & 1) ? (byte & 0xf) : (byte >> 4);
          }
        }
        Q.p += (lseq + 1) / 2;
        if (lseq) memcpy(&S.qual[(size_t)S.seq_off[ri]], Q.p, (size_t)lseq);
        Q.p += lseq;
        int32_t x1 = -1, xl = -1, xr = -1, rgi = -1;
        int8_t xpc = -1;
        while (Q.p < rec_end) {
          Q.need(3);
          const char t0 = (char)Q.p[0], t1 = (char)Q.p[1], ty = (char)Q.p[2];
          const unsigned char* v = Q.p + 3;
          const size_t sz = aux_size(ty, v, rec_end);
          if ((size_t)(rec_end - v) < sz) throw BsqError(BSQ_ERR_FORMAT, "truncated aux field");
          int64_t iv;
          if (t0 == 'X' && t1 == '1' && aux_int(ty, v, iv)) x1 = (int32_t)iv;
          else if (t0 == 'X' && t1 == 'L' && aux_int(ty, v, iv)) xl = (int32_t)iv;
          else if (t0 == 'X' && t1 == 'R' && aux_int(ty, v, iv)) xr = (int32_t)iv;
          else if (t0 == 'X' && t1 == 'P' && ty == 'Z') {
            const char* o = (const char*)v;
            if (sz == 3 && o[0] == 'F' && o[1] == 'R') xpc = 0;
            else if (sz == 3 && o[0] == 'R' && o[1] == 'F') xpc = 1;
            else if (sz == 3 && o[0] == 'F' && o[1] == 'F') xpc = 2;
          } else if (t0 == 'R' && t1 == 'G' && ty == 'Z') {
            const char* id = (const char*)v;
            for (size_t g = 0; g < S.rg_ids.size(); g++)
              if (S.rg_ids[g] == id) {
                rgi = (int32_t)g;
                break;
              }
          }
          Q.p = v + sz;
        }
        S.tid[ri] = tid;
        S.pos[ri] = pos;
        S.flag[ri] = flag;
        S.mapq[ri] = mq;
        S.mate_tid[ri] = mtid;
        S.mate_pos[ri] = mpos;
        S.tlen[ri] = tlen;
        S.xp[ri] = xpc;
        S.x1[ri] = x1;
        S.xl[ri] = xl;
        S.xr[ri] = xr;
        S.rg[ri] = rgi;
      }
    });
    S.finalize();
    *out = smp;
    set_status(status, BSQ_OK, "");
    return BSQ_OK;
  } catch (const BsqError& e) {
    delete smp;
    set_status(status, e.code, e.what());
    return e.code;
  } catch (const std::exception& e) {
    delete smp;
    set_status(status, BSQ_ERR_FORMAT, e.what());
    return BSQ_ERR_FORMAT;
  }
}

extern "C" const bsq_reference* bsq_sample_reference(const bsq_sample* s) { return &s->s.ref_view; }
extern "C" const bsq_alignments* bsq_sample_alignments(const bsq_sample* s) { return &s->s.aln_view; }
extern "C" int32_t bsq_sample_n_read_groups(const bsq_sample* s) { return (int32_t)s->s.rg_ids.size(); }
extern "C" const char* bsq_sample_read_group(const bsq_sample* s, int32_t i) {
  if (i < 0 || i >= (int32_t)s->s.rg_ids.size()) return nullptr;
  return s->s.rg_ids[i].c_str();
}
extern "C" void bsq_sample_free(bsq_sample* s) { delete s; }
