// evidence.cpp -- see evidence.h.
#include "evidence.h"

#include <charconv>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <sstream>

#include "host_util.h"

namespace bsq {

static const char kIdx[6] = {'A', 'C', 'G', 'T', '.', 'N'};

void counts_to_log10_prob_range(int S, int Q, const std::vector<uint64_t>& counts, double* out, int q0, int q1) {
  // the table is quality-major: the bins of one quality are the 25 * S consecutive entries from q * 25 * S
  (void)Q;
  const size_t blk = (size_t)25 * S;
  for (int q = q0; q < q1; q++) {
    const uint64_t* c = counts.data() + (size_t)q * blk;
    double* o = out + (size_t)q * blk;
    // summed over obs_base for each (read_set, ref_base); the denominator's logarithm once each
    double den[5 * 256];
    for (size_t j = 0; j < (size_t)5 * S; j++) den[j] = 0.0;
    for (size_t j = 0; j < blk; j++) den[j % ((size_t)5 * S)] += (double)c[j];
    for (size_t j = 0; j < (size_t)5 * S; j++) den[j] = std::log10(den[j] + 5.0);
    for (size_t j = 0; j < blk; j++) o[j] = std::log10((double)c[j] + 1.0) - den[j % ((size_t)5 * S)];
  }
}

void counts_to_log10_prob(int S, int Q, const std::vector<uint64_t>& counts, std::vector<double>& out) {
  out.resize((size_t)S * 25 * Q);
  counts_to_log10_prob_range(S, Q, counts, out.data(), 0, Q);
}

std::string error_rates_text(int S, int Q, const std::vector<double>& log10p) {
  std::string s;
  s.reserve(log10p.size() * 24 + 128);
  s += "read_set=" + std::to_string(S) + ",ref_base,obs_base,quality=" + std::to_string(Q) + "\n";
  s += "read_set\tref_base\tobs_base\tquality\tlog10_probability\n";
  char num[32];
  auto put_u = [&](size_t x) {
    const std::to_chars_result r = std::to_chars(num, num + sizeof(num), x);
    s.append(num, (size_t)(r.ptr - num));
  };
  for (size_t i = 0; i < log10p.size(); i++) {
    const size_t rs = i % S, rb = (i / S) % 5, ob = (i / (5 * (size_t)S)) % 5, q = i / (25 * (size_t)S);
    // "%zu\t%c\t%c\t%zu\t%g\n"
    put_u(rs);
    s += '\t';
    s += kIdx[rb];
    s += '\t';
    s += kIdx[ob];
    s += '\t';
    put_u(q);
    s += '\t';
    s += fmt_g(log10p[i]);
    s += '\n';
  }
  return s;
}

std::string coverage_text(const std::vector<uint32_t>& hist) {
  std::string s = "coverage\tn\n";
  char buf[48];
  for (size_t j = 1; j < hist.size(); j++) {
    snprintf(buf, sizeof(buf), "%zu\t%u\n", j, hist[j]);
    s += buf;
  }
  return s;
}

std::vector<std::string> base_qual_texts(int S, int Q, const std::vector<uint64_t>& counts) {
  const size_t blk = (size_t)25 * Q;
  std::vector<double> ot((size_t)S * blk, 0.0);
  for (size_t i = 0; i < counts.size(); i++) {
    const size_t rs = i % S, rb = (i / S) % 5, ob = (i / (5 * (size_t)S)) % 5, q = i / (25 * (size_t)S);
    ot[rs * blk + rb + 5 * ob + 25 * q] += (double)counts[i];
  }
  std::vector<std::string> texts;
  double running = 0;
  for (int r = 0; r < S; r++) {
    for (size_t k = 0; k < blk; k++) {
      const size_t i = (size_t)r * blk + k;
      running += ot[i];
      if (i % 5 == 4) {
        for (size_t j = i - 4; j <= i; j++) {
          if (running > 0) ot[j] /= running;
          if (ot[j] == 0) ot[j] = NAN;
        }
        running = 0;
      }
    }
    std::string s = "quality";
    for (int b1 = 0; b1 < 5; b1++)
      for (int b2 = 0; b2 < 5; b2++) {
        s += "\t";
        s += kIdx[b1];
        s += kIdx[b2];
      }
    s += "\n";
    for (int q = 0; q < Q; q++) {
      s += std::to_string(q);
      for (int b1 = 0; b1 < 5; b1++)
        for (int b2 = 0; b2 < 5; b2++) {
          const double v = ot[(size_t)r * blk + (size_t)q * 25 + b1 * 5 + b2];
          s += "\t";
          s += std::isnan(v) ? std::string("NA") : fmt_g(v);
        }
      s += "\n";
    }
    texts.push_back(s);
  }
  return texts;
}

bool parse_error_rates(const std::string& text, std::vector<double>& out, int& S, int& Q) {
  std::istringstream in(text);
  std::string line;
  if (!std::getline(in, line)) return false;
  S = Q = 0;
  {
    std::stringstream cs(line);
    std::string part;
    bool rs = false, rb = false, ob = false, q = false;
    while (std::getline(cs, part, ',')) {
      size_t eq = part.find('=');
      std::string name = part.substr(0, eq);
      if (name == "read_set") { rs = true; S = atoi(part.substr(eq + 1).c_str()); }
      else if (name == "ref_base") rb = true;
      else if (name == "obs_base") ob = true;
      else if (name == "quality") { q = true; Q = atoi(part.substr(eq + 1).c_str()); }
      else return false;   // only the covariates of breseq's main pipeline are supported
    }
    if (!(rs && rb && ob && q) || S <= 0 || Q <= 0) return false;
  }
  if (!std::getline(in, line)) return false;
  const size_t nb = (size_t)S * 25 * Q;
  out.assign(nb, 0.0);
  for (size_t i = 0; i < nb; i++) {
    if (!std::getline(in, line)) return false;
    size_t tab = line.rfind('\t');
    std::string last = (tab == std::string::npos) ? line : line.substr(tab + 1);
    out[i] = strtod(last.c_str(), NULL);
  }
  return true;
}

/* ---------------------------------------------------------------- genome diff */

static double reported(const RaOut& r, int b) {
  if (r.n == 0 || b >= 5) return 0.0;
  return (r.f[b] < 0.5 / (double)r.n) ? 0.0 : r.f[b];
}

void ra_entry_from_result(const RaOut& r, const std::string& seq_id, uint32_t position, uint32_t places, GdEntry& e) {
  e.type = 0;
  e.seq_id = seq_id;
  e.pos = position;
  e.ins = r.slot;
  e.ref_base = std::string(1, kIdx[r.ref_idx]);
  e.new_base = std::string(1, kIdx[r.variant]);
  // the fields in the order of their names (the entry's std::map order when written), written
  // straight into one string; a value that formats empty is left out, as the writer does
  std::string& t = e.fields_text;
  t.clear();
  t.reserve(400);
  char buf[64];
  auto key = [&](const char* k) {
    t += '\t';
    t += k;
    t += '=';
  };
  auto put_i = [&](int64_t x) {
    const std::to_chars_result c = std::to_chars(buf, buf + sizeof(buf), x);
    t.append(buf, (size_t)(c.ptr - buf));
  };
  auto put_d = [&](const char* k, double v, unsigned prec, bool sci) {
    const int n = fmt_double_buf(v, prec, sci, buf, sizeof(buf));
    if (n <= 0) return;
    key(k);
    t.append(buf, (size_t)n);
  };
  auto put_cov = [&](const char* k, int b) {
    key(k);
    if (b >= 5) {
      t += "0/0";
      return;
    }
    put_i((int32_t)r.cov[b][0]);
    t += '/';
    put_i((int32_t)r.cov[b][1]);
  };
  bool any = false;
  for (int b = 0; b < 5; b++) {
    const double fr = reported(r, b);
    if (fr <= 0.0) continue;
    const int n = fmt_double_buf(fr, places, true, buf, sizeof(buf));
    if (n <= 0) continue;
    if (!any) key("allele_frequencies");
    else t += ',';
    any = true;
    t += kIdx[b];
    t += ':';
    t.append(buf, (size_t)n);
  }
  put_d("fisher_strand_p_value", r.fisher_p, 5, true);
  put_d("frequency", reported(r, r.variant), places, true);
  put_d("frequency_lower", r.lower, places, true);
  put_d("frequency_upper", r.upper, places, true);
  put_d("ks_quality_p_value", r.ks_p, 5, true);
  key("major_base");
  t += kIdx[r.major];
  put_cov("major_cov", r.major);
  put_d("major_frequency", reported(r, r.major), places, true);
  key("minor_base");
  t += kIdx[r.minor];
  put_cov("minor_cov", r.minor);
  put_cov("new_cov", r.variant);
  put_cov("ref_cov", r.ref_idx);
  put_d("score", r.variant_score, 1, false);
  key("total_cov");
  put_i(r.total_top);
  t += '/';
  put_i(r.total_bot);
  e.fields.clear();
}

static int type_group(int t) { return t == 0 ? 3 : (t == 1 ? 4 : 7); }

std::string write_gd(std::vector<GdEntry>& entries) {
  // unique ids in add() order
  std::vector<size_t> idx(entries.size());
  for (size_t i = 0; i < idx.size(); i++) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) {
    const GdEntry &x = entries[a], &y = entries[b];
    if (x.order_target != y.order_target) return x.order_target < y.order_target;
    if (x.order_pos != y.order_pos) return x.order_pos < y.order_pos;
    return x.order_sub < y.order_sub;
  });
  for (size_t k = 0; k < idx.size(); k++) entries[idx[k]].id = (uint32_t)(k + 1);
  // cGenomeDiff::sort (genome_diff_entry.cpp:566-700)
  std::vector<const GdEntry*> v;
  for (const GdEntry& e : entries) v.push_back(&e);
  std::stable_sort(v.begin(), v.end(), [](const GdEntry* a, const GdEntry* b) {
    if (type_group(a->type) != type_group(b->type)) return type_group(a->type) < type_group(b->type);
    if (a->seq_id != b->seq_id) return a->seq_id < b->seq_id;
    if (a->pos != b->pos) return a->pos < b->pos;
    if (a->type == 0) {
      if (a->ins != b->ins) return a->ins < b->ins;
      if (a->ref_base != b->ref_base) return a->ref_base < b->ref_base;
      if (a->new_base != b->new_base) return a->new_base < b->new_base;
    } else {
      if (a->end != b->end) return a->end < b->end;
      if (a->type == 1) {
        if (a->start_range != b->start_range) return a->start_range < b->start_range;
        if (a->end_range != b->end_range) return a->end_range < b->end_range;
      }
    }
    return a->id < b->id;
  });
  std::string out;
  out.reserve(64 + v.size() * 400);
  out += "#=GENOME_DIFF\t1.0\n";
  static const char* tn[3] = {"RA", "MC", "UN"};
  char num[16];
  auto put_u = [&](int64_t x) {
    const std::to_chars_result r = std::to_chars(num, num + sizeof(num), x);
    out.append(num, (size_t)(r.ptr - num));
  };
  for (const GdEntry* e : v) {
    out += tn[e->type];
    out += '\t';
    put_u(e->id);
    out += "\t.\t";
    out += e->seq_id;
    out += '\t';
    put_u(e->pos);
    out += '\t';
    if (e->type == 0) {
      put_u(e->ins);
      out += '\t';
      out += e->ref_base;
      out += '\t';
      out += e->new_base;
    } else if (e->type == 1) {
      put_u(e->end);
      out += '\t';
      put_u(e->start_range);
      out += '\t';
      put_u(e->end_range);
    } else {
      put_u(e->end);
    }
    out += e->fields_text;
    for (const auto& kv : e->fields) {
      if (kv.second.empty() || kv.first[0] == '_') continue;
      out += '\t';
      out += kv.first;
      out += '=';
      out += kv.second;
    }
    out += '\n';
  }
  return out;
}

/* ---------------------------------------------------------------- test_RA_evidence */

namespace {

// The fields of one entry, sorted by name as the std::map of a cGenomeDiff entry keeps them (its
// writer prints them in that order); a repeated name keeps its last value.  A flat sorted vector:
// an entry has about twenty fields.
struct Fields {
  std::vector<std::pair<std::string, std::string>> kv;

  std::vector<std::pair<std::string, std::string>>::iterator lower(const std::string& k) {
    return std::lower_bound(kv.begin(), kv.end(), k,
                            [](const std::pair<std::string, std::string>& a, const std::string& b) { return a.first < b; });
  }
  int find(const std::string& k) const {
    auto it = std::lower_bound(kv.begin(), kv.end(), k,
                               [](const std::pair<std::string, std::string>& a, const std::string& b) { return a.first < b; });
    return (it != kv.end() && it->first == k) ? (int)(it - kv.begin()) : -1;
  }
  bool has(const std::string& k) const { return find(k) >= 0; }
  const std::string& get(const std::string& k) const {
    static const std::string empty;
    const int i = find(k);
    return i < 0 ? empty : kv[(size_t)i].second;
  }
  std::string& at(const std::string& k) {   // inserted empty when absent, as std::map::operator[]
    auto it = lower(k);
    if (it == kv.end() || it->first != k) it = kv.insert(it, std::make_pair(k, std::string()));
    return it->second;
  }
  void erase(const std::string& k) {
    auto it = lower(k);
    if (it != kv.end() && it->first == k) kv.erase(it);
  }
  // from name=value fields in line order: sorted by name, the last of a repeated name kept
  void assign(std::vector<std::pair<std::string, std::string>>&& fields) {
    kv = std::move(fields);
    std::stable_sort(kv.begin(), kv.end(),
                     [](const std::pair<std::string, std::string>& a, const std::pair<std::string, std::string>& b) {
                       return a.first < b.first;
                     });
    size_t w = 0;
    for (size_t r = 0; r < kv.size(); r++) {
      if (w > 0 && kv[w - 1].first == kv[r].first) kv[w - 1].second = std::move(kv[r].second);
      else if (w != r) kv[w++] = std::move(kv[r]);
      else w++;
    }
    kv.resize(w);
  }
};

bool is_digit(char c) { return c >= '0' && c <= '9'; }

// from_string<double> (common.h:892-899): the value a stream extraction leaves, whether or not it
// fails; that is 0 when the text holds no number and +-DBL_MAX when it is out of range.  (The
// reference asserts on an empty string; here it reads as 0.)  A plain decimal number ([+-] digits
// [. digits] [e [+-] digits]) is read with strtod, which rounds it as the stream does; anything
// else, and a value out of range, takes the stream itself.
double from_string_d(const std::string& s) {
  const char* p = s.c_str();
  const size_t n = s.size();
  size_t i = 0, nd = 0;
  if (i < n && (p[i] == '+' || p[i] == '-')) i++;
  while (i < n && is_digit(p[i])) i++, nd++;
  if (i < n && p[i] == '.') {
    i++;
    while (i < n && is_digit(p[i])) i++, nd++;
  }
  bool plain = nd > 0;
  if (plain && i < n && (p[i] == 'e' || p[i] == 'E')) {
    size_t e = i + 1;
    if (e < n && (p[e] == '+' || p[e] == '-')) e++;
    const size_t e0 = e;
    while (e < n && is_digit(p[e])) e++;
    plain = e > e0;
    i = e;
  }
  if (plain && i == n) {
    errno = 0;
    const double v = strtod(p, nullptr);
    if (errno != ERANGE) return v;
  }
  std::istringstream is(s);
  double t = 0.0;
  is >> t;
  return t;
}

double double_from_string(const std::string& in) {   // common.h:939-950
  std::string u = in;
  for (char& c : u) c = (char)toupper((unsigned char)c);
  if (u == "NA" || u == "#NA" || u == "NAN") return NAN;
  if (u == "INF") return INFINITY;
  if (u == "-INF") return -INFINITY;
  return from_string_d(in);
}

// the two numbers of an "a/b" field (b is 0 when absent)
void pair_of(const std::string& v, double& a, double& b) {
  const size_t k = v.find('/');
  if (k == std::string::npos) {
    a = from_string_d(v);
    b = 0.0;
  } else {
    a = from_string_d(v.substr(0, k));
    b = from_string_d(v.substr(k + 1, v.find('/', k + 1) == std::string::npos ? std::string::npos : v.find('/', k + 1) - k - 1));
  }
}

struct Ra {
  Fields& m;
  bool has(const std::string& k) const { return m.has(k); }
  const std::string& get(const std::string& k) const { return m.get(k); }
  void add_reject(const std::string& reason, const std::string& field = "reject") {
    if (!has(field)) {
      m.at(field) = reason;
      return;
    }
    std::string& cur = m.at(field);
    size_t b = 0;
    for (;;) {   // a reason already in the comma-separated list is not added again
      const size_t e = cur.find(',', b);
      if (cur.compare(b, (e == std::string::npos ? cur.size() : e) - b, reason) == 0) return;
      if (e == std::string::npos) break;
      b = e + 1;
    }
    cur += ",";
    cur += reason;
  }
};

struct RefAccess {
  const bsq_reference* ref;
  int target(const std::string& seq_id) const {
    for (int i = 0; ref && i < ref->n_targets; i++)
      if (seq_id == ref->names[i]) return i;
    return -1;
  }
  int64_t length(int t) const { return t < 0 ? 0 : ref->lengths[t]; }
  // get_sequence_1(pos, pos) == b, where a position outside the target reads as ""
  bool base_is(int t, int64_t pos1, const std::string& b) const {
    if (t < 0 || pos1 < 1 || pos1 > ref->lengths[t]) return b.empty();
    return b.size() == 1 && ref->seqs[t][pos1 - 1] == b[0];
  }
};

double score_of(Ra& ra) {
  if (ra.has("score")) return double_from_string(ra.get("score"));
  double legacy = NAN;
  if (ra.has("consensus_score")) legacy = double_from_string(ra.get("consensus_score"));
  if (ra.has("polymorphism_score")) {
    double p = double_from_string(ra.get("polymorphism_score"));
    if (std::isnan(legacy) || p > legacy) legacy = p;
  }
  return legacy;
}

double bd_lower(double k, double n);
double bd_upper(double k, double n);

bool freq_bounds(Ra& ra, double& lo, double& hi) {
  if (ra.has("frequency_lower") && ra.has("frequency_upper")) {
    lo = from_string_d(ra.get("frequency_lower"));
    hi = from_string_d(ra.get("frequency_upper"));
    return true;
  }
  if (!ra.has("total_cov") || !ra.has("frequency")) return false;
  const std::string& tc = ra.get("total_cov");
  if (tc.find('/') == std::string::npos) return false;
  double a, b;
  pair_of(tc, a, b);
  double n = a + b;
  if (!(n > 0.0)) return false;
  double k = from_string_d(ra.get("frequency")) * n;
  lo = bd_lower(k, n);
  hi = bd_upper(k, n);
  return true;
}

void bias(Ra& ra, const bsq_ra_test_settings& s) {   // rejected_RA_polymorphism_bias
  if (s.polymorphism_ks_quality_p_value_cutoff != 0 && ra.has("ks_quality_p_value"))
    if (from_string_d(ra.get("ks_quality_p_value")) < s.polymorphism_ks_quality_p_value_cutoff) ra.add_reject("KS_BASE_QUALITY");
  if (s.polymorphism_fisher_strand_p_value_cutoff != 0 && ra.has("fisher_strand_p_value"))
    if (from_string_d(ra.get("fisher_strand_p_value")) < s.polymorphism_fisher_strand_p_value_cutoff) ra.add_reject("FISHER_STRAND");
}

void poly_coverage(Ra& ra, const bsq_ra_test_settings& s) {   // rejected_RA_polymorphism_coverage
  double a, b;
  if (s.polymorphism_minimum_variant_coverage_each_strand > 0) {
    bool passed = true;
    const std::string& ref_base = ra.get("ref_base");
    bool major_var = ra.has("major_base") && ra.get("major_base") != ref_base;
    bool minor_var = ra.has("minor_base") && ra.get("minor_base") != ref_base;
    if (!major_var && !minor_var) major_var = minor_var = true;
    const double c = s.polymorphism_minimum_variant_coverage_each_strand;
    if (major_var && ra.has("major_cov")) {
      pair_of(ra.get("major_cov"), a, b);
      passed = passed && a >= c && b >= c;
    }
    if (minor_var && ra.has("minor_cov")) {
      pair_of(ra.get("minor_cov"), a, b);
      passed = passed && a >= c && b >= c;
    }
    if (!passed) ra.add_reject("VARIANT_STRAND_COVERAGE");
  }
  if (s.polymorphism_minimum_total_coverage_each_strand > 0) {
    pair_of(ra.get("total_cov"), a, b);
    if (!(a >= s.polymorphism_minimum_total_coverage_each_strand && b >= s.polymorphism_minimum_total_coverage_each_strand))
      ra.add_reject("TOTAL_STRAND_COVERAGE");
  }
  if (s.polymorphism_minimum_variant_coverage > 0) {
    bool passed = true;
    pair_of(ra.get("major_cov"), a, b);
    passed = passed && (a + b >= s.polymorphism_minimum_variant_coverage);
    pair_of(ra.get("minor_cov"), a, b);
    passed = passed && (a + b >= s.polymorphism_minimum_variant_coverage);
    if (!passed) ra.add_reject("VARIANT_COVERAGE");
  }
  if (s.polymorphism_minimum_total_coverage > 0) {
    pair_of(ra.get("total_cov"), a, b);
    if (a + b < s.polymorphism_minimum_total_coverage) ra.add_reject("TOTAL_COVERAGE");
  }
}

void cons_coverage(Ra& ra, const bsq_ra_test_settings& s) {   // rejected_RA_consensus_coverage
  double a, b;
  if (s.consensus_minimum_variant_coverage_each_strand > 0) {
    pair_of(ra.get("major_cov"), a, b);
    if (a < s.consensus_minimum_variant_coverage_each_strand || b < s.consensus_minimum_variant_coverage_each_strand) ra.add_reject("VARIANT_STRAND_COVERAGE");
  }
  if (s.consensus_minimum_total_coverage_each_strand > 0) {
    pair_of(ra.get("total_cov"), a, b);
    if (a < s.consensus_minimum_total_coverage_each_strand || b < s.consensus_minimum_total_coverage_each_strand) ra.add_reject("TOTAL_STRAND_COVERAGE");
  }
  if (s.consensus_minimum_variant_coverage > 0) {
    pair_of(ra.get("major_cov"), a, b);
    if (a + b < s.consensus_minimum_variant_coverage) ra.add_reject("VARIANT_COVERAGE");
  }
  if (s.consensus_minimum_total_coverage > 0) {
    pair_of(ra.get("total_cov"), a, b);
    if (a + b < s.consensus_minimum_total_coverage) ra.add_reject("TOTAL_COVERAGE");
  }
}

void homopolymer(Ra& ra, const RefAccess& R, int32_t indel_len, int32_t surround_len, bool no_indels) {   // :400-487
  const std::string rb = ra.get("ref_base"), nb = ra.get("new_base");
  if (indel_len && (rb == "." || nb == ".")) {
    const int t = R.target(ra.get("seq_id"));
    int32_t mut_pos = (int32_t)from_string_d(ra.get("position"));
    bool is_ins = from_string_d(ra.get("insert_position")) > 0;
    const std::string& mb = (rb == ".") ? nb : rb;
    int32_t hl = 0;
    bool no_match = false;
    if (is_ins) {
      if (!R.base_is(t, mut_pos, mb)) {
        mut_pos++;
        if (mut_pos > (int32_t)R.length(t)) no_match = true;
        else if (!R.base_is(t, mut_pos, mb)) no_match = true;
      }
    }
    if (!no_match) {
      int32_t sp = mut_pos - 1;
      while (R.base_is(t, sp, mb) && sp > 0) sp--;
      sp++;
      int32_t ep = mut_pos + 1;
      while (R.base_is(t, ep, mb) && ep <= (int32_t)R.length(t)) ep++;
      ep--;
      hl = ep - sp + 1;
    }
    if (hl >= indel_len) ra.add_reject("INDEL_HOMOPOLYMER");
  }
  if (surround_len && rb != "." && nb != ".") {
    const int t = R.target(ra.get("seq_id"));
    int32_t mut_pos = (int32_t)from_string_d(ra.get("position"));
    int32_t sp = mut_pos - 1;
    while (sp >= 1) {
      if (!R.base_is(t, sp, nb)) break;
      sp--;
    }
    sp++;
    int32_t ep = mut_pos + 1;
    while (ep <= (int32_t)R.length(t)) {
      if (!R.base_is(t, ep, nb)) break;
      ep++;
    }
    ep--;
    if (sp < mut_pos && ep > mut_pos && ep - sp + 1 >= surround_len) ra.add_reject("SURROUNDING_HOMOPOLYMER");
  }
  if (no_indels && (rb == "." || nb == ".")) ra.add_reject("POLYMORPHIC_INDEL");
}

bool test_mode(Ra& ra, const RefAccess& R, const bsq_ra_test_settings& s, bool poly_mode) {   // :522-685
  const double score = score_of(ra);
  double lower, upper;
  if (!freq_bounds(ra, lower, upper)) lower = upper = from_string_d(ra.get("frequency"));
  if (score < s.mutation_log10_e_value_cutoff) ra.add_reject("SCORE_CUTOFF");
  const double fcons = poly_mode ? upper : lower;
  if (s.consensus_frequency_cutoff > 0.0 && fcons < s.consensus_frequency_cutoff) ra.add_reject("FREQUENCY_CUTOFF");
  cons_coverage(ra, s);
  homopolymer(ra, R, s.consensus_reject_indel_homopolymer_length, s.consensus_reject_surrounding_homopolymer_length, false);
  if (!ra.has("reject")) {
    ra.m.at("prediction") = "consensus";
    return ra.get("ref_base") == ra.get("major_base");
  }
  std::string moved = ra.get("reject");   // a copy: at() may insert and move the fields
  ra.m.at("consensus_reject") = moved;
  ra.m.erase("reject");
  if (score < s.polymorphism_log10_e_value_cutoff) ra.add_reject("SCORE_CUTOFF");
  if (s.polymorphism_frequency_cutoff > 0.0 && lower < s.polymorphism_frequency_cutoff) ra.add_reject("FREQUENCY_CUTOFF");
  bias(ra, s);
  poly_coverage(ra, s);
  homopolymer(ra, R, s.polymorphism_reject_indel_homopolymer_length, s.polymorphism_reject_surrounding_homopolymer_length,
              s.polymorphism_no_indels != 0);
  if (!ra.has("reject")) {
    ra.m.at("prediction") = "polymorphism";
    return false;
  }
  moved = ra.get("reject");
  ra.m.at("polymorphism_reject") = moved;
  ra.m.erase("reject");
  if (!poly_mode) return true;
  ra.m.at("prediction") = "polymorphism";
  moved = ra.get("polymorphism_reject");
  ra.m.at("reject") = moved;
  ra.m.erase("polymorphism_reject");
  return false;
}

// Clopper-Pearson bounds for evidence written without recorded bounds (stats.cpp:2395-2415):
// lower = incbi(k, n-k+1, 0.05), upper = incbi(k+1, n-k, 0.95) via bisection on the regularized
// incomplete beta function.
double ibeta(double a, double b, double x) {   // continued fraction (Numerical Recipes betacf form)
  if (x <= 0.0) return 0.0;
  if (x >= 1.0) return 1.0;
  const double lbt = std::lgamma(a + b) - std::lgamma(a) - std::lgamma(b) + a * std::log(x) + b * std::log(1.0 - x);
  auto cf = [](double aa, double bb, double xx) {
    const double tiny = 1e-300;
    double qab = aa + bb, qap = aa + 1.0, qam = aa - 1.0, c = 1.0, d = 1.0 - qab * xx / qap;
    if (std::fabs(d) < tiny) d = tiny;
    d = 1.0 / d;
    double h = d;
    for (int m = 1; m < 400; m++) {
      int m2 = 2 * m;
      double an = m * (bb - m) * xx / ((qam + m2) * (aa + m2));
      d = 1.0 + an * d; if (std::fabs(d) < tiny) d = tiny;
      c = 1.0 + an / c; if (std::fabs(c) < tiny) c = tiny;
      d = 1.0 / d; h *= d * c;
      an = -(aa + m) * (qab + m) * xx / ((aa + m2) * (qap + m2));
      d = 1.0 + an * d; if (std::fabs(d) < tiny) d = tiny;
      c = 1.0 + an / c; if (std::fabs(c) < tiny) c = tiny;
      d = 1.0 / d;
      double del = d * c;
      h *= del;
      if (std::fabs(del - 1.0) < 1e-15) break;
    }
    return h;
  };
  if (x < (a + 1.0) / (a + b + 2.0)) return std::exp(lbt) * cf(a, b, x) / a;
  return 1.0 - std::exp(lbt) * cf(b, a, 1.0 - x) / b;
}

double ibeta_inv(double a, double b, double y) {
  double lo = 0.0, hi = 1.0;
  for (int i = 0; i < 200; i++) {
    double mid = 0.5 * (lo + hi);
    if (ibeta(a, b, mid) < y) lo = mid;
    else hi = mid;
  }
  return 0.5 * (lo + hi);
}

double bd_lower(double k, double n) { return (k <= 0) ? 0.0 : ibeta_inv(k, n - k + 1, 0.05); }
double bd_upper(double k, double n) { return (k >= n) ? 1.0 : ibeta_inv(k + 1, n - k, 0.95); }

}  // namespace

// Each line is read once: entries are split at tabs, their name=value fields sorted (the order the
// reference's writer prints), RA entries classified, and every entry printed back.
std::string test_ra_evidence_text(const std::string& text, const bsq_reference* ref, const bsq_ra_test_settings& s) {
  RefAccess R{ref};
  std::string out;
  out.reserve(text.size() + text.size() / 8);
  std::vector<std::string> spec;
  std::vector<std::pair<std::string, std::string>> fields;
  Fields m;
  // a malformed line anywhere fails the whole text before any entry is tested
  for (size_t b = 0; b < text.size();) {
    size_t e = text.find('\n', b);
    if (e == std::string::npos) e = text.size();
    const char* l = text.data() + b;
    const size_t ln = e - b;
    b = e + 1;
    if (ln == 0 || l[0] == '#') continue;
    const void* t1 = memchr(l, '\t', ln);
    const void* t2 = t1 ? memchr((const char*)t1 + 1, '\t', ln - (size_t)((const char*)t1 + 1 - l)) : nullptr;
    if (!t2) throw BsqError(BSQ_ERR_FORMAT, "malformed genome diff line: " + std::string(l, ln));
  }
  size_t b = 0;
  while (b < text.size()) {
    size_t e = text.find('\n', b);
    if (e == std::string::npos) e = text.size();
    const char* l = text.data() + b;
    const size_t ln = e - b;
    b = e + 1;
    if (ln == 0 || l[0] == '#') {
      out.append(l, ln);
      out += '\n';
      continue;
    }
    // split at tabs, keeping empty fields
    spec.clear();
    fields.clear();
    std::string type, id, evidence;
    int nf = 0;
    size_t f0 = 0;
    for (size_t k = 0; k <= ln; k++) {
      if (k < ln && l[k] != '\t') continue;
      const char* f = l + f0;
      const size_t fn = k - f0;
      if (nf == 0) type.assign(f, fn);
      else if (nf == 1) id.assign(f, fn);
      else if (nf == 2) evidence.assign(f, fn);
      else {
        const void* eq = memchr(f, '=', fn);
        if (!eq) spec.emplace_back(f, fn);
        else {
          const size_t kl = (size_t)((const char*)eq - f);
          fields.emplace_back(std::string(f, kl), std::string(f + kl + 1, fn - kl - 1));
        }
      }
      nf++;
      f0 = k + 1;
    }
    m.assign(std::move(fields));
    fields = std::vector<std::pair<std::string, std::string>>();
    if (type == "RA") {
      if (spec.size() < 5) throw BsqError(BSQ_ERR_FORMAT, "RA entry without its 5 positional fields");
      m.at("seq_id") = spec[0];
      m.at("position") = spec[1];
      m.at("insert_position") = spec[2];
      m.at("ref_base") = spec[3];
      m.at("new_base") = spec[4];
      Ra ra{m};
      if (!ra.has("score") && !ra.has("consensus_score") && !ra.has("polymorphism_score"))
        throw BsqError(BSQ_ERR_FORMAT, "Expected field 'score' in evidence item");
      if (!ra.has("frequency")) throw BsqError(BSQ_ERR_FORMAT, "Expected field 'frequency' in evidence item");
      bool del = test_mode(ra, R, s, s.polymorphism_prediction != 0);
      if (ra.has("user_defined")) del = false;
      if (del) continue;
      for (const char* k : {"seq_id", "position", "insert_position", "ref_base", "new_base"}) m.erase(k);
    }
    out += type;
    out += '\t';
    out += id;
    out += '\t';
    out += evidence;
    for (const auto& sp : spec) {
      out += '\t';
      out += sp;
    }
    for (const auto& kv : m.kv) {
      if (kv.second.empty() || kv.first[0] == '_') continue;
      out += '\t';
      out += kv.first;
      out += '=';
      out += kv.second;
    }
    out += '\n';
  }
  return out;
}

}  // namespace bsq

extern "C" int32_t bsq_format_double(double v, uint32_t precision, int32_t scientific, char* buf, int32_t buf_len) {
  const std::string s = (scientific == 2) ? bsq::fmt_g(v, (int)precision) : bsq::fmt_double(v, precision, scientific != 0);
  if (!buf || buf_len <= (int32_t)s.size()) return -1;
  memcpy(buf, s.c_str(), s.size() + 1);
  return (int32_t)s.size();
}
