// C ABI (include/dmr.h): drop-in entry points, blastp/blastx drivers,
// frame merging and outfmt-6 rendering.
#include "../../include/dmr.h"

#include "host_pool.hpp"

#include <emmintrin.h>
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "api_internal.hpp"

namespace dmr_api {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

// Index of the first letter code outside the 25-letter alphabet
// (src/alphabet.rs:15-16), n when there is none.
uint64_t first_bad_letter(const uint8_t* p, uint64_t n) {
  constexpr uint64_t kBlock = 4096;
  for (uint64_t b = 0; b < n; b += kBlock) {
    const uint64_t e = std::min(n, b + kBlock);
    uint8_t m = 0;
    for (uint64_t i = b; i < e; ++i) m = std::max(m, p[i]);
    if (m >= dmr::kAlphabet)
      for (uint64_t i = b; i < e; ++i)
        if (p[i] >= dmr::kAlphabet) return i;
  }
  return n;
}

void check_letters(const uint8_t* p, uint64_t n, const char* what) {
  const uint64_t i = n && p ? first_bad_letter(p, n) : n;
  if (i < n)
    throw dmr::Error(DMR_E_ARG, std::string(what) + ": letter code " + std::to_string(p[i]) +
                                    " at position " + std::to_string(i) +
                                    " is outside the 25-letter alphabet");
}

using dmr::HspRec;

bool hsp_less(const HspRec& a, const HspRec& b) {
  if (a.score != b.score) return a.score > b.score;
  if (a.evalue < b.evalue) return true;
  if (b.evalue < a.evalue) return false;
  return a.subject < b.subject;
}

// search::merge_frames (src/search.rs:357-396): per query, all frames'
// HSPs sorted by cmp_hsp, at most max_hsps per subject, subjects in order of
// first appearance up to max_targets.
std::vector<std::vector<HspRec>> merge_frames(std::vector<std::vector<HspRec>>& per_target,
                                              const dmr::TargetSet& t, size_t nqueries,
                                              size_t max_hsps, size_t max_targets) {
  // One target per query (protein queries): the search already returns each
  // target's HSPs in this order (subjects by their best HSP, each subject's
  // HSPs sorted, both limits applied per target), so the lists move as they
  // are.
  bool single = per_target.size() == nqueries && t.query_of.size() == nqueries;
  for (size_t ti = 0; single && ti < nqueries; ++ti) single = t.query_of[ti] == ti && t.frame_of[ti] == 0;
  if (single) {
    for (size_t ti = 0; ti < nqueries; ++ti)
      for (auto& h : per_target[ti]) {
        h.frame = 0;
        h.query = (uint32_t)ti;
      }
    return std::move(per_target);
  }
  std::vector<std::vector<HspRec>> per(nqueries);
  for (size_t ti = 0; ti < per_target.size(); ++ti) {
    size_t q = t.query_of[ti];
    for (auto& h : per_target[ti]) {
      h.frame = t.frame_of[ti];
      h.query = (uint32_t)q;
      per[q].push_back(std::move(h));
    }
  }
  for (auto& hs : per) {
    if (hs.empty() || (hs.size() == 1 && max_hsps >= 1 && max_targets >= 1)) continue;
    std::stable_sort(hs.begin(), hs.end(), hsp_less);
    // Subjects in order of first appearance (few per query: a linear scan),
    // each keeping its first max_hsps HSPs.
    std::vector<uint32_t> order;
    std::vector<std::vector<HspRec>> by;
    for (auto& h : hs) {
      size_t k = 0;
      while (k < order.size() && order[k] != h.subject) ++k;
      if (k < order.size()) {
        if (by[k].size() < max_hsps) by[k].push_back(std::move(h));
      } else if (order.size() < max_targets) {
        order.push_back(h.subject);
        by.emplace_back();
        by.back().push_back(std::move(h));
      }
    }
    hs.clear();
    for (auto& b : by)
      for (auto& h : b) hs.push_back(std::move(h));
  }
  return per;
}

// translate::min_orf_len (src/translate.rs:124-137) of every read: the
// user's --min-orf, else 1 below 30 bases, 20 below 100, then 40.
std::vector<uint32_t> min_orf_lengths(const uint64_t* r_offsets, uint32_t n, int32_t user) {
  std::vector<uint32_t> mo(n);
  for (uint32_t i = 0; i < n; ++i) {
    const uint64_t len = r_offsets[i + 1] - r_offsets[i];
    mo[i] = user >= 0 ? (uint32_t)user : (len < 30 ? 1u : (len < 100 ? 20u : 40u));
  }
  return mo;
}

dmr::SearchOptions options(const dmr_search_config* c) {
  dmr::SearchOptions o;
  if (!c) return o;
  o.sensitivity = c->sensitivity ? c->sensitivity : "default";
  o.evalue = c->evalue;
  o.max_target_seqs = c->max_target_seqs;
  o.max_hsps = c->max_hsps;
  o.index_chunks = c->index_chunks;
  o.seed_cut = c->seed_cut;
  o.gap_open = c->gap_open == -1 ? 11 : c->gap_open;
  o.gap_extend = c->gap_extend == -1 ? 1 : c->gap_extend;
  return o;
}

void assemble(dmr_ctx* ctx, dmr_result* r, std::vector<std::vector<HspRec>>& per_target,
              const dmr::SearchOptions& o, size_t nqueries) {
  auto per = merge_frames(per_target, r->targets, nqueries, o.max_hsps, o.max_target_seqs);
  // Target of each (query, frame): frames are 0..5.
  std::vector<int> tix(nqueries * 6, -1);
  for (size_t t = 0; t < r->targets.query_of.size(); ++t) {
    const size_t q = r->targets.query_of[t], f = r->targets.frame_of[t];
    if (q < nqueries && f < 6) tix[q * 6 + f] = (int)t;
  }
  size_t total = 0;
  for (auto& hs : per) total += hs.size();
  r->target_of.reserve(total);
  r->hsps.reserve(total);
  for (auto& hs : per)
    for (auto& h : hs) {
      r->target_of.push_back(h.query < nqueries && h.frame < 6 ? tix[(size_t)h.query * 6 + h.frame] : -1);
      r->hsps.push_back(std::move(h));
    }
  r->db_offs = ctx->eng->db_offsets_shared();
  r->db_res = ctx->eng->db_letters_shared();
}

void finish(dmr_ctx* ctx, dmr_result* r, const dmr::SearchOptions& o, size_t nqueries,
            bool upload = true) {
  if (upload) ctx->eng->load_targets(r->targets);
  auto per_target = ctx->eng->search_targets(o, &r->stats, &r->ops_stores);
  assemble(ctx, r, per_target, o, nqueries);
}

// Closes the target just appended to t.letters: its end offset is 32-bit in
// the device layout (with one padding letter per target and two more), so
// batches past that are refused before the offset is narrowed.
void end_target(dmr::TargetSet& t) {
  if ((uint64_t)t.letters.size() + t.offs.size() + 2 >= (1ull << 32))
    throw dmr::Error(DMR_E_UNSUPPORTED, "query batch larger than 4G letters");
  t.offs.push_back((uint32_t)t.letters.size());
}

dmr::TargetSet protein_targets(const uint8_t* queries, const uint64_t* q_offsets, uint32_t nq) {
  dmr::TargetSet t;
  t.offs.push_back(0);
  for (uint32_t i = 0; i < nq; ++i) {
    uint64_t a = q_offsets[i], b = q_offsets[i + 1];
    if (b < a) throw dmr::Error(DMR_E_ARG, "query offsets are not monotone");
    if (b == a) continue;  // empty queries have no target (src/workflow.rs:129-132)
    if (!queries) throw dmr::Error(DMR_E_ARG, "NULL argument");
    check_letters(queries + a, b - a, ("query " + std::to_string(i)).c_str());
    t.letters.insert(t.letters.end(), queries + a, queries + b);
    end_target(t);
    t.query_of.push_back(i);
    t.frame_of.push_back(0);
  }
  return t;
}

// {:.1} of 0 <= x < 100 by exact integer rounding of x * 10 (x = m / 2^s
// exactly; round half to even on the exact value, as the exact decimal
// formatting does). False outside that range.
bool fixed1_fast(double x, char* b, char** end) {
  if (!(x >= 0.0 && x < 100.0)) return false;
  uint64_t q = 0;
  if (x != 0.0) {
    int e2;
    const uint64_t m = (uint64_t)std::ldexp(std::frexp(x, &e2), 53);  // exact
    const int sh = 53 - e2;
    if (sh <= 0) {
      q = (m << -sh) * 10;
    } else if (sh <= 60) {  // otherwise x < 2^-7, below 0.05: q = 0
      const uint64_t t = m * 10;  // < 2^57
      q = t >> sh;
      const uint64_t rem = t & ((1ull << sh) - 1), half = 1ull << (sh - 1);
      if (rem > half || (rem == half && (q & 1))) ++q;
    }
  }
  char* p = std::to_chars(b, b + 8, q / 10).ptr;
  *p++ = '.';
  *p++ = (char)('0' + q % 10);
  *end = p;
  return true;
}

// Correctly rounded 10^n, n in [0, 308].
const double* pow10_table() {
  static const std::vector<double> t = [] {
    std::vector<double> v(309);
    for (int n = 0; n <= 308; ++n) v[n] = std::strtod(("1e" + std::to_string(n)).c_str(), nullptr);
    return v;
  }();
  return t.data();
}

// {:.2e} of e when its three significant digits are clear of a rounding tie
// (and of the 100 / 1000 bounds) by far more than the error of e * 10^n
// (two roundings); false otherwise, and the caller formats exactly.
bool sci2_fast(double e, char* b, char** end) {
  if (!(e >= 1e-290 && e < 1e290)) return false;
  const double* p10 = pow10_table();
  int e2;
  std::frexp(e, &e2);  // e in [2^(e2-1), 2^e2)
  int k = (int)std::floor((e2 - 1) * 0.30102999566398120);  // floor(log10 e) or one less
  auto scaled = [&](int kk) { return 2 - kk >= 0 ? e * p10[2 - kk] : e / p10[kk - 2]; };
  double s = scaled(k);
  if (s >= 1000.0) s = scaled(++k);
  constexpr double kMargin = 1e-6;
  if (s < 100.0 + kMargin || s >= 1000.0 - kMargin) return false;
  const double q = std::floor(s), frac = s - q;
  if (std::fabs(frac - 0.5) < kMargin) return false;
  unsigned d = (unsigned)q + (frac > 0.5 ? 1u : 0u);
  if (d == 1000) {
    d = 100;
    ++k;
  }
  char* p = b;
  *p++ = (char)('0' + d / 100);
  *p++ = '.';
  *p++ = (char)('0' + d / 10 % 10);
  *p++ = (char)('0' + d % 10);
  *p++ = 'e';
  p = std::to_chars(p, p + 8, k).ptr;
  *end = p;
  return true;
}

// src/output.rs:235-242: values >= 100 print as their floor, others with one
// decimal ({:.1}, the exact binary value rounded half to even, as printf's
// %.1f; std::to_chars with a precision is specified to match printf).
void put_double(std::string& out, double x) {
  char b[64];
  char* end;
  if (fixed1_fast(x, b, &end)) {
    out.append(b, end);
    return;
  }
  std::to_chars_result r;
  if (x >= 100.0) r = std::to_chars(b, b + sizeof b, (long long)std::floor(x));
  else r = std::to_chars(b, b + sizeof b, x, std::chars_format::fixed, 1);
  out.append(b, r.ptr);
}

// src/output.rs:244-250 (Rust {:.2e}: no exponent sign padding)
void put_evalue(std::string& out, double e) {
  if (!std::isfinite(e) || e <= 0.0) {
    out += "0.0";
    return;
  }
  char b[64];
  char* fe;
  if (sci2_fast(e, b, &fe)) {
    out.append(b, fe);
    return;
  }
  const auto r = std::to_chars(b, b + sizeof b, e, std::chars_format::scientific, 2);
  const char* end = r.ptr;
  const char* p = std::find(static_cast<const char*>(b), end, 'e');
  out.append(b, (size_t)(p - b));
  out += 'e';
  const char* x = p + 1;
  const bool neg = *x == '-';
  if (*x == '-' || *x == '+') ++x;
  while (x + 1 < end && *x == '0') ++x;
  if (neg) out += '-';
  out.append(x, (size_t)(end - x));
}

std::vector<std::string> split(const char* s, char sep) {
  std::vector<std::string> out;
  if (!s) return out;
  std::string cur;
  for (const char* p = s; *p; ++p) {
    if (*p == sep) {
      if (sep != ' ' || !cur.empty()) out.push_back(cur);
      cur.clear();
    } else {
      cur += *p;
    }
  }
  if (!cur.empty()) out.push_back(cur);
  return out;
}

void set_aligned(const dmr::SwOut& o, dmr_aligned* a, int64_t q_shift) {
  std::memset(a, 0, sizeof *a);
  a->found = o.score > 0 ? 1 : 0;
  a->score = o.score;
  if (!a->found) return;
  a->q_begin = (uint64_t)(q_shift + o.q_begin);
  a->q_end = (uint64_t)(q_shift + o.best_i);
  a->s_begin = o.s_begin;
  a->s_end = o.best_j;
  a->stats = {o.identities, o.positives, o.mismatches, o.gap_openings, o.gaps, o.length};
  a->n_ops = o.n_ops;
}

}  // namespace dmr_api

using namespace dmr_api;

namespace dmr {
void set_last_error(const std::string& msg) { g_err = msg; }
}  // namespace dmr

extern "C" {

const char* dmr_last_error(void) { return g_err.c_str(); }
const char* dmr_version(void) { return "dmr 0.1 (gfx950)"; }

void dmr_search_config_default(dmr_search_config* c) {
  if (!c) return;
  c->sensitivity = "default";
  c->evalue = 0.001;
  c->max_target_seqs = 25;
  c->max_hsps = 1;
  c->index_chunks = -1;
  c->gap_open = -1;
  c->gap_extend = -1;
  c->min_orf = -1;
  c->seed_cut = -1.0;
}

int dmr_device_count(int* n) {
  return guard([&]() -> int {
    DMR_HIP(hipGetDeviceCount(n));
    return DMR_OK;
  });
}

int dmr_ctx_create(int device, dmr_ctx** out) {
  if (!out) return fail(DMR_E_ARG, "out is NULL");
  return guard([&]() -> int {
    auto c = std::make_unique<dmr_ctx>();
    c->eng = std::make_unique<dmr::Engine>(device);
    *out = c.release();
    return DMR_OK;
  });
}

void dmr_ctx_destroy(dmr_ctx* ctx) { delete ctx; }

int dmr_ctx_share_db(dmr_ctx* ctx, const dmr_ctx* from) {
  if (!ctx || !from) return fail(DMR_E_ARG, "NULL context");
  if (ctx == from) return DMR_OK;
  return guard([&]() -> int {
    ctx->eng->share_db(*from->eng);
    ctx->db_ids = from->db_ids;
    ctx->db_headers = from->db_headers;
    ctx->db_has_ids = from->db_has_ids;
    ctx->db_header_nl = from->db_header_nl;
    return DMR_OK;
  });
}

int dmr_ctx_set_tuning(dmr_ctx* ctx, const char* key, double value) {
  if (!ctx || !key) return fail(DMR_E_ARG, "NULL argument");
  return guard([&] {
    const std::string k = key;
    dmr::Engine& e = *ctx->eng;
    if (k == "slice_hits") {
      if (!(value >= 1)) return fail(DMR_E_ARG, "slice_hits must be at least 1");
      e.set_slice_hits((uint64_t)value);
    } else if (k == "profile") {
      e.set_profile(value != 0);
    } else {
      e.set_layout_key(k, value);
    }
    return DMR_OK;
  });
}

int dmr_ctx_get_tuning(const dmr_ctx* ctx, const char* key, double* value) {
  if (!ctx || !key || !value) return fail(DMR_E_ARG, "NULL argument");
  return guard([&] {
    const std::string k = key;
    const dmr::Engine& e = *ctx->eng;
    if (k == "slice_hits") *value = (double)e.slice_hits();
    else if (k == "last_slices") *value = e.last_slices();
    else if (k == "tuned_presets") *value = (double)e.tuned_presets();
    else if (k == "graph_launches") *value = (double)e.graph_launches();
    else if (k == "graph_builds") *value = (double)e.graph_builds();
    else *value = e.layout_key(k);
    return DMR_OK;
  });
}

int dmr_ctx_load_tuning(dmr_ctx* ctx, const char* path) {
  if (!ctx) return fail(DMR_E_ARG, "NULL argument");
  return guard([&] {
    ctx->eng->load_tuning(path ? path : "");
    return DMR_OK;
  });
}

int dmr_sw16_variants(int32_t* out, uint32_t cap, uint32_t* n) {
  if (!n) return fail(DMR_E_ARG, "NULL argument");
  const auto& v = dmr::sw16_variants();
  *n = (uint32_t)v.size();
  if (!out) return DMR_OK;
  if (cap < v.size()) return fail(DMR_E_CAPACITY, "variant buffer too small");
  for (size_t k = 0; k < v.size(); ++k) {
    out[4 * k] = v[k].cls;
    out[4 * k + 1] = v[k].width;
    out[4 * k + 2] = v[k].lanes;
    out[4 * k + 3] = v[k].threads;
  }
  return DMR_OK;
}

uint32_t dmr_band_for(uint64_t len) { return dmr::band_for(len); }

int dmr_evalue(int32_t go, int32_t ge, int32_t raw, uint64_t qlen, uint64_t slen, uint64_t db,
               double* evalue, double* bits) {
  dmr::GapRow r;
  if (go == -1) go = 11;
  if (ge == -1) ge = 1;
  if (!dmr::blosum62_gap_row(go, ge, &r)) return fail(DMR_E_ARG, "unsupported gap costs");
  dmr::AlpParams p = dmr::alp_make(r.lambda, r.k, r.alpha, r.alpha_v, r.sigma, go, ge,
                                   dmr::kB62UngappedAlpha, dmr::kB62UngappedAlphaV);
  if (evalue) *evalue = dmr::alp_evalue(p, raw, qlen, slen, db);
  if (bits) *bits = dmr::bit_score(p, raw);
  return DMR_OK;
}

int dmr_seed_build(dmr_ctx* ctx, const uint8_t* seqs, const uint64_t* offsets, uint32_t nseq,
                   const char* shape_code, const char* reduction_def, uint32_t base,
                   double complexity_cut, dmr_seed_entry* out, uint64_t cap, uint64_t* n_out) {
  if (!ctx || !offsets || !shape_code || !reduction_def || !n_out || (nseq && offsets[nseq] && !seqs))
    return fail(DMR_E_ARG, "NULL argument");
  dmr::ShapeDef sh;
  if (!dmr::parse_shape(shape_code, &sh)) return fail(DMR_E_ARG, "invalid seed shape");
  dmr::ReductionDef red = dmr::parse_reduction(reduction_def);
  if (red.size == 0) return fail(DMR_E_ARG, "empty reduction definition");
  return guard([&]() -> int {
    std::vector<uint64_t> k;
    std::vector<uint32_t> s, p;
    if (nseq) check_letters(seqs, offsets[nseq], "sequences");
    ctx->eng->seed_build(seqs, offsets, nseq, sh, red, base, complexity_cut, &k, &s, &p);
    *n_out = k.size();
    if (k.size() > cap) return fail(DMR_E_CAPACITY, "output capacity too small");
    for (size_t i = 0; i < k.size(); ++i) out[i] = dmr_seed_entry{k[i], s[i], p[i]};
    return DMR_OK;
  });
}

int dmr_seed_join(dmr_ctx* ctx, const dmr_seed_entry* q, uint64_t nq, const dmr_seed_entry* r,
                  uint64_t nr, uint64_t max_expand, dmr_seed_hit* out, uint64_t cap,
                  uint64_t* n_out) {
  if (!ctx || !n_out || (nq && !q) || (nr && !r)) return fail(DMR_E_ARG, "NULL argument");
  return guard([&]() -> int {
    std::vector<uint64_t> qk(nq), rk(nr);
    std::vector<uint32_t> qs(nq), qp(nq), rs(nr), rp(nr);
    for (uint64_t i = 0; i < nq; ++i) { qk[i] = q[i].key; qs[i] = q[i].seq; qp[i] = q[i].pos; }
    for (uint64_t i = 0; i < nr; ++i) { rk[i] = r[i].key; rs[i] = r[i].seq; rp[i] = r[i].pos; }
    std::vector<uint32_t> h;
    ctx->eng->seed_join(qk.data(), qs.data(), qp.data(), nq, rk.data(), rs.data(), rp.data(), nr,
                        max_expand, &h);
    uint64_t n = h.size() / 4;
    *n_out = n;
    if (n > cap) return fail(DMR_E_CAPACITY, "output capacity too small");
    for (uint64_t i = 0; i < n; ++i) out[i] = dmr_seed_hit{h[4 * i], h[4 * i + 1], h[4 * i + 2], h[4 * i + 3]};
    return DMR_OK;
  });
}

int dmr_xdrop_batch(dmr_ctx* ctx, const uint8_t* query, uint64_t qlen, const uint8_t* subject,
                    uint64_t slen, const uint32_t* qa, const uint32_t* sa, uint64_t n, int32_t xdrop,
                    dmr_segment* out) {
  if (!ctx || (n && (!qa || !sa || !out)) || (qlen && !query) || (slen && !subject))
    return fail(DMR_E_ARG, "NULL argument");
  return guard([&]() -> int {
    std::vector<dmr::SegRec> s(n);
    check_letters(query, qlen, "query");
    check_letters(subject, slen, "subject");
    ctx->eng->xdrop_batch(query, qlen, subject, slen, qa, sa, n, xdrop, s.data());
    for (uint64_t i = 0; i < n; ++i) out[i] = dmr_segment{s[i].q_start, s[i].s_start, s[i].len, s[i].score, 0};
    return DMR_OK;
  });
}

int dmr_xdrop(dmr_ctx* ctx, const uint8_t* query, uint64_t qlen, const uint8_t* subject,
              uint64_t slen, uint64_t qa, uint64_t sa, int32_t xdrop, dmr_segment* out) {
  uint32_t a = (uint32_t)qa, b = (uint32_t)sa;
  return dmr_xdrop_batch(ctx, query, qlen, subject, slen, &a, &b, 1, xdrop, out);
}

int dmr_extend_span_batch(dmr_ctx* ctx, const uint8_t* qbuf, uint64_t qbuf_len, const uint8_t* sbuf,
                          uint64_t sbuf_len, const dmr_span_job* jobs, uint64_t njobs, int32_t go,
                          int32_t ge, int traceback, dmr_aligned* out, uint32_t* ops_pool,
                          uint64_t ops_cap, uint64_t* ops_offsets, uint64_t* ops_needed) {
  if (!ctx || (njobs && (!jobs || !out)) || (qbuf_len && !qbuf) || (sbuf_len && !sbuf))
    return fail(DMR_E_ARG, "NULL argument");
  dmr::GapRow gr;
  if (!dmr::blosum62_gap_row(go, ge, &gr)) return fail(DMR_E_ARG, "unsupported gap costs");
  return guard([&]() -> int {
    check_letters(qbuf, qbuf_len, "query buffer");
    check_letters(sbuf, sbuf_len, "subject buffer");
    std::vector<dmr::Engine::SpanJob> sj(njobs);
    for (uint64_t i = 0; i < njobs; ++i)
      sj[i] = {jobs[i].q_off, jobs[i].q_len, jobs[i].s_off, jobs[i].s_len, jobs[i].d0, jobs[i].d1};
    std::vector<dmr::SwOut> o;
    std::vector<std::vector<uint32_t>> ops;
    ctx->eng->extend_span_batch(qbuf, qbuf_len, sbuf, sbuf_len, sj.data(), njobs, go, ge,
                                traceback != 0, &o, &ops);
    uint64_t need = 0;
    for (uint64_t i = 0; i < njobs; ++i) {
      set_aligned(o[i], &out[i], 0);
      if (!traceback) {
        std::memset(&out[i].q_begin, 0, 4 * sizeof(uint64_t) + sizeof(dmr_hsp_stats) + 8);
      }
      if (ops_offsets) ops_offsets[i] = need;
      need += traceback ? ops[i].size() : 0;
    }
    if (ops_offsets) ops_offsets[njobs] = need;
    if (ops_needed) *ops_needed = need;
    if (traceback) {
      if (need > ops_cap || (need && !ops_pool)) return fail(DMR_E_CAPACITY, "ops pool too small");
      uint64_t k = 0;
      for (auto& v : ops)
        for (uint32_t x : v) ops_pool[k++] = x;
    }
    return DMR_OK;
  });
}

int dmr_extend_span(dmr_ctx* ctx, const uint8_t* query, uint64_t qlen, const uint8_t* subject,
                    uint64_t slen, int64_t d0, int64_t d1, int32_t go, int32_t ge, dmr_aligned* out,
                    uint32_t* ops, uint32_t ops_cap) {
  if (!ctx || !out) return fail(DMR_E_ARG, "NULL argument");
  dmr_span_job j{0, qlen, 0, slen, d0, d1};
  uint64_t need = 0, offs[2];
  std::vector<uint32_t> tmp(qlen + slen + 1);
  int rc = dmr_extend_span_batch(ctx, query, qlen, subject, slen, &j, 1, go, ge, 1, out, tmp.data(),
                                 tmp.size(), offs, &need);
  if (rc != DMR_OK) return rc;
  if (need > ops_cap) return fail(DMR_E_CAPACITY, "ops capacity too small");
  if (need) std::memcpy(ops, tmp.data(), need * 4);
  return DMR_OK;
}

int dmr_extend_span_noalloc(dmr_ctx* ctx, const uint8_t* query, uint64_t qlen, const uint8_t* subject,
                            uint64_t slen, int64_t d0, int64_t d1, int32_t go, int32_t ge, dmr_aligned* out,
                            uint32_t* ops, uint32_t ops_cap) {
  if (!ctx || !out || (qlen && !query) || (slen && !subject)) return fail(DMR_E_ARG, "NULL argument");
  dmr::GapRow gr;
  if (!dmr::blosum62_gap_row(go, ge, &gr)) return fail(DMR_E_ARG, "unsupported gap costs");
  return guard([&]() -> int {
    check_letters(query, qlen, "query");
    check_letters(subject, slen, "subject");
    dmr::SwOut o;
    bool fit = false;
    ctx->eng->extend_span_noalloc(query, qlen, subject, slen, d0, d1, go, ge, &o, ops, ops_cap, &fit);
    set_aligned(o, out, 0);
    if (!fit) return fail(DMR_E_CAPACITY, "ops capacity too small");
    return DMR_OK;
  });
}

int dmr_extend(dmr_ctx* ctx, const uint8_t* query, uint64_t qlen, const uint8_t* subject,
               uint64_t slen, uint64_t aq, uint64_t as, uint64_t band, int32_t go, int32_t ge,
               dmr_aligned* out, uint32_t* ops, uint32_t ops_cap) {
  int64_t diag = (int64_t)aq - (int64_t)as;
  int64_t b = (int64_t)std::max<uint64_t>(band, 1);
  return dmr_extend_span(ctx, query, qlen, subject, slen, diag - b, diag + b, go, ge, out, ops, ops_cap);
}

int dmr_db_load(dmr_ctx* ctx, const uint8_t* letters, const uint64_t* offsets, uint32_t nseq) {
  if (!ctx || !offsets || (offsets[nseq] && !letters)) return fail(DMR_E_ARG, "NULL argument");
  if (offsets[0] != 0) return fail(DMR_E_ARG, "offsets[0] must be 0");
  for (uint32_t i = 0; i < nseq; ++i)
    if (offsets[i + 1] < offsets[i]) return fail(DMR_E_ARG, "offsets are not monotone");
  return guard([&]() -> int {
    check_letters(letters, offsets[nseq], "database");
    ctx->eng->load_db(letters, offsets, nseq);
    ctx->db_ids.clear();
    ctx->db_has_ids = false;
    return DMR_OK;
  });
}

namespace {

// DRDMND01 layout (src/db.rs:24-29, 117-131): a 72-byte header (magic,
// version u32, flags u32, then seq_count, total_letters, defs_len,
// seq_off_start, def_off_start, letters_start, defs_start as u64), then
// seq_count + 1 u64 letter offsets, seq_count + 1 u64 header offsets, the
// encoded letters and the newline-free FASTA headers. Little-endian.
constexpr size_t kDbHeaderLen = 72;
constexpr uint32_t kDbVersion = 1;

uint64_t le64(const uint8_t* p) {
  uint64_t v;
  std::memcpy(&v, p, 8);
  return v;
}

bool path_exists(const std::string& p) {
  struct stat st;
  return ::stat(p.c_str(), &st) == 0;
}

// db::resolve (src/db.rs:39-60): the path as given, else with ".dmnd"
// appended when its file name has no extension.
bool resolve_db_path(const std::string& path, std::string* out) {
  if (path_exists(path)) {
    *out = path;
    return true;
  }
  std::string name = path;
  while (name.size() > 1 && name.back() == '/') name.pop_back();
  const size_t slash = name.rfind('/');
  if (slash != std::string::npos) name = name.substr(slash + 1);
  const size_t dot = name.rfind('.');
  const bool has_ext = name != ".." && dot != std::string::npos && dot > 0;
  if (has_ext) return false;
  *out = path + ".dmnd";
  return path_exists(*out);
}

// std::str::from_utf8 acceptance: well-formed UTF-8, no overlong forms,
// surrogates or code points above U+10FFFF.
bool valid_utf8(const uint8_t* s, size_t n) {
  size_t i = 0;
  while (i < n) {
    const uint8_t c = s[i];
    if (c < 0x80) {
      ++i;
      continue;
    }
    size_t len;
    uint32_t cp;
    if (c >= 0xC2 && c <= 0xDF) {
      len = 2;
      cp = c & 0x1F;
    } else if (c >= 0xE0 && c <= 0xEF) {
      len = 3;
      cp = c & 0x0F;
    } else if (c >= 0xF0 && c <= 0xF4) {
      len = 4;
      cp = c & 0x07;
    } else {
      return false;
    }
    if (n - i < len) return false;
    for (size_t k = 1; k < len; ++k) {
      if ((s[i + k] & 0xC0) != 0x80) return false;
      cp = cp << 6 | (s[i + k] & 0x3F);
    }
    if (len == 3 && (cp < 0x800 || (cp >= 0xD800 && cp <= 0xDFFF))) return false;
    if (len == 4 && (cp < 0x10000 || cp > 0x10FFFF)) return false;
    i += len;
  }
  return true;
}

// Database::id (src/db.rs:262-265) over alphabet::seq_id (src/alphabet.rs:87-93):
// the header (empty when it is not UTF-8) without leading '>' and ' ', up to
// the first ASCII whitespace.
void append_seq_id(const uint8_t* h, size_t n, std::string* out) {
  if (valid_utf8(h, n)) {
    size_t a = 0;
    while (a < n && (h[a] == '>' || h[a] == ' ')) ++a;
    size_t b = a;
    while (b < n && h[b] != ' ' && h[b] != '\t' && h[b] != '\n' && h[b] != '\f' && h[b] != '\r') ++b;
    out->append(reinterpret_cast<const char*>(h + a), b - a);
  }
  out->push_back('\n');
}

// Read-only mapping of a whole file, unmapped on scope exit.
struct Mapped {
  const uint8_t* p = nullptr;
  size_t n = 0;
  ~Mapped() {
    if (p && n) ::munmap(const_cast<uint8_t*>(p), n);
  }
};

}  // namespace

int dmr_db_load_file(dmr_ctx* ctx, const char* path, uint32_t* nseq) {
  if (!ctx || !path) return fail(DMR_E_ARG, "NULL argument");
  std::string file;
  if (!resolve_db_path(path, &file)) return fail(DMR_E_ARG, std::string("Database file not found: ") + path);
  return guard([&]() -> int {
    Mapped m;
    {
      const int fd = ::open(file.c_str(), O_RDONLY);
      struct stat st;
      if (fd < 0 || ::fstat(fd, &st) != 0) {
        const std::string why = std::strerror(errno);
        if (fd >= 0) ::close(fd);
        return fail(DMR_E_ARG, "cannot read " + file + ": " + why);
      }
      m.n = (size_t)st.st_size;
      if (m.n) {
        void* a = ::mmap(nullptr, m.n, PROT_READ, MAP_PRIVATE, fd, 0);
        if (a == MAP_FAILED) {
          const std::string why = std::strerror(errno);
          ::close(fd);
          m.n = 0;
          return fail(DMR_E_ARG, "cannot read " + file + ": " + why);
        }
        m.p = static_cast<const uint8_t*>(a);
      }
      ::close(fd);
    }
    const uint8_t* d = m.p;
    const uint64_t len = m.n;
    // Database::open (src/db.rs:172-222).
    if (len < kDbHeaderLen) return fail(DMR_E_ARG, file + ": not a diamond-rust database (file too short)");
    if (std::memcmp(d, "DRDMND01", 8) != 0)
      return fail(DMR_E_ARG, file + ": not a diamond-rust database (bad magic). Build one with `diamond-rust makedb`.");
    uint32_t version;
    std::memcpy(&version, d + 8, 4);
    if (version != kDbVersion)
      return fail(DMR_E_ARG, file + ": database version " + std::to_string(version) +
                                 " is not supported (expected " + std::to_string(kDbVersion) + ")");
    const uint64_t n = le64(d + 16), total_letters = le64(d + 24), defs_len = le64(d + 32);
    const uint64_t seq_off_start = le64(d + 40), def_off_start = le64(d + 48);
    const uint64_t letters_start = le64(d + 56), defs_start = le64(d + 64);
    if (defs_start > len || len - defs_start != defs_len)
      return fail(DMR_E_ARG, file + ": truncated or corrupt database");
    auto table_fits = [&](uint64_t start) { return n < len / 8 && start <= len && (n + 1) * 8 <= len - start; };
    if (!table_fits(seq_off_start) || !table_fits(def_off_start))
      return fail(DMR_E_ARG, file + ": corrupt offset table");
    // Offsets the reference reads only when a sequence or header is used
    // (Database::seq / header, src/db.rs:243-259) are checked here once:
    // monotone and inside the file.
    if (n > UINT32_MAX) return fail(DMR_E_UNSUPPORTED, file + ": more than 2^32 - 1 sequences");
    std::vector<uint64_t> so(n + 1);
    std::memcpy(so.data(), d + seq_off_start, (n + 1) * 8);
    const uint8_t* dofs = d + def_off_start;
    for (uint64_t i = 0; i < n; ++i) {
      if (so[i + 1] < so[i] || le64(dofs + 8 * (i + 1)) < le64(dofs + 8 * i))
        return fail(DMR_E_ARG, file + ": corrupt offset table");
    }
    if (letters_start > len || so[n] > len - letters_start || le64(dofs + 8 * n) > len - defs_start)
      return fail(DMR_E_ARG, file + ": corrupt offset table");
    // Sequence i is letters[so[i]..so[i + 1]] (Database::seq, src/db.rs:256-260);
    // letters before so[0] belong to no sequence, so the offsets are rebased.
    const uint8_t* letters = d + letters_start + so[0];
    const uint64_t base = so[0];
    for (auto& o : so) o -= base;
    const uint64_t bad = first_bad_letter(letters, so[n]);
    if (bad < so[n])
      return fail(DMR_E_ARG, file + ": corrupt letters section (code " + std::to_string(letters[bad]) +
                                 " at letter " + std::to_string(base + bad) + " is outside the 25-letter alphabet)");
    std::string ids, heads;
    ids.reserve(n * 16);
    heads.reserve(defs_len + n);
    bool nl = false;
    for (uint64_t i = 0; i < n; ++i) {
      const uint64_t a = le64(dofs + 8 * i), b = le64(dofs + 8 * (i + 1));
      const uint8_t* h = d + defs_start + a;
      append_seq_id(h, b - a, &ids);
      // Database::header (src/db.rs:262-266): the bytes as text, "" when not UTF-8.
      if (valid_utf8(h, b - a)) {
        nl |= std::memchr(h, '\n', b - a) != nullptr;
        heads.append(reinterpret_cast<const char*>(h), b - a);
      }
      heads += '\n';
    }
    // E-values use the header's letter count (Database::total_letters,
    // src/db.rs:199,236-238 and src/search.rs:158), which a file may state
    // differently from its offset table.
    ctx->eng->load_db(letters, so.data(), (uint32_t)n, &total_letters);
    ctx->db_ids = std::move(ids);
    ctx->db_headers = std::move(heads);
    ctx->db_header_nl = nl;
    ctx->db_has_ids = true;
    if (nseq) *nseq = (uint32_t)n;
    return DMR_OK;
  });
}

int dmr_db_ids(const dmr_ctx* ctx, char* buf, uint64_t cap, uint64_t* len) {
  if (!ctx || !len) return fail(DMR_E_ARG, "NULL argument");
  if (!ctx->db_has_ids) return fail(DMR_E_ARG, "the resident database was not loaded from a file");
  *len = ctx->db_ids.size();
  if (cap < ctx->db_ids.size() || (ctx->db_ids.size() && !buf))
    return fail(DMR_E_CAPACITY, "output capacity too small");
  if (!ctx->db_ids.empty()) std::memcpy(buf, ctx->db_ids.data(), ctx->db_ids.size());
  return DMR_OK;
}

int dmr_db_headers(const dmr_ctx* ctx, char* buf, uint64_t cap, uint64_t* len) {
  if (!ctx || !len) return fail(DMR_E_ARG, "NULL argument");
  if (!ctx->db_has_ids) return fail(DMR_E_ARG, "the resident database was not loaded from a file");
  if (ctx->db_header_nl)
    return fail(DMR_E_UNSUPPORTED, "a database header holds a newline; the header list cannot be newline-separated");
  *len = ctx->db_headers.size();
  if (cap < ctx->db_headers.size() || (ctx->db_headers.size() && !buf))
    return fail(DMR_E_CAPACITY, "output capacity too small");
  if (!ctx->db_headers.empty()) std::memcpy(buf, ctx->db_headers.data(), ctx->db_headers.size());
  return DMR_OK;
}

int dmr_db_build_index(dmr_ctx* ctx, const char* sens, int32_t index_chunks, double seed_cut,
                       double* build_ms) {
  if (!ctx) return fail(DMR_E_ARG, "NULL context");
  return guard([&]() -> int {
    ctx->eng->build_index(sens ? sens : "default", index_chunks, seed_cut);
    if (build_ms) *build_ms = ctx->eng->last_index_ms();
    return DMR_OK;
  });
}

int dmr_blastp(dmr_ctx* ctx, const uint8_t* queries, const uint64_t* q_offsets, uint32_t nq,
               const dmr_search_config* cfg, dmr_result** out) {
  if (!ctx || !q_offsets || !out) return fail(DMR_E_ARG, "NULL argument");
  dmr::SearchOptions o = options(cfg);
  dmr::GapRow gr;
  if (!dmr::blosum62_gap_row(o.gap_open, o.gap_extend, &gr))
    return fail(DMR_E_ARG, "Gap penalty settings are outside the supported range for BLOSUM62");
  return guard([&]() -> int {
    auto r = std::make_unique<dmr_result>();
    r->targets = protein_targets(queries, q_offsets, nq);
    r->dna_len.resize(nq, 0);
    finish(ctx, r.get(), o, nq);
    *out = r.release();
    return DMR_OK;
  });
}

int dmr_queries_load(dmr_ctx* ctx, const uint8_t* queries, const uint64_t* q_offsets, uint32_t nq) {
  if (!ctx || !q_offsets) return fail(DMR_E_ARG, "NULL argument");
  return guard([&]() -> int {
    ctx->loaded = protein_targets(queries, q_offsets, nq);
    ctx->loaded_nq = nq;
    ctx->eng->load_targets(ctx->loaded);
    ctx->has_loaded = true;
    return DMR_OK;
  });
}

int dmr_search_loaded(dmr_ctx* ctx, const dmr_search_config* cfg, dmr_result** out) {
  if (!ctx || !out) return fail(DMR_E_ARG, "NULL argument");
  if (!ctx->has_loaded) return fail(DMR_E_ARG, "no query batch loaded");
  dmr::SearchOptions o = options(cfg);
  dmr::GapRow gr;
  if (!dmr::blosum62_gap_row(o.gap_open, o.gap_extend, &gr))
    return fail(DMR_E_ARG, "Gap penalty settings are outside the supported range for BLOSUM62");
  return guard([&]() -> int {
    auto r = std::make_unique<dmr_result>();
    r->targets = ctx->loaded;
    r->dna_len.resize(ctx->loaded_nq, 0);
    finish(ctx, r.get(), o, ctx->loaded_nq, false);
    *out = r.release();
    return DMR_OK;
  });
}

int dmr_blastp_plain(dmr_ctx* ctx, const uint8_t* queries, const uint64_t* q_offsets, uint32_t nq,
                     const dmr_search_config* cfg, dmr_result** out) {
  if (!ctx || !q_offsets || !out) return fail(DMR_E_ARG, "NULL argument");
  dmr::SearchOptions o = options(cfg);
  dmr::GapRow gr;
  if (!dmr::blosum62_gap_row(o.gap_open, o.gap_extend, &gr))
    return fail(DMR_E_ARG, "Gap penalty settings are outside the supported range for BLOSUM62");
  return guard([&]() -> int {
    auto r = std::make_unique<dmr_result>();
    r->targets = protein_targets(queries, q_offsets, nq);
    r->dna_len.resize(nq, 0);
    std::shared_ptr<const dmr::OpsStore> store;
    auto per_target = dmr::search_plain(*ctx->eng, r->targets, o, &r->stats, &store);
    r->ops_stores.push_back(std::move(store));
    assemble(ctx, r.get(), per_target, o, nq);
    *out = r.release();
    return DMR_OK;
  });
}

int dmr_blastx(dmr_ctx* ctx, const char* reads, const uint64_t* r_offsets, uint32_t nreads,
               const dmr_search_config* cfg, dmr_result** out) {
  if (!ctx || !r_offsets || !out) return fail(DMR_E_ARG, "NULL argument");
  dmr::SearchOptions o = options(cfg);
  dmr::GapRow gr;
  if (!dmr::blosum62_gap_row(o.gap_open, o.gap_extend, &gr))
    return fail(DMR_E_ARG, "Gap penalty settings are outside the supported range for BLOSUM62");
  if (nreads && !reads && r_offsets[nreads] > r_offsets[0]) return fail(DMR_E_ARG, "NULL argument");
  for (uint32_t i = 0; i < nreads; ++i)
    if (r_offsets[i + 1] < r_offsets[i]) return fail(DMR_E_ARG, "read offsets are not monotone");
  return guard([&]() -> int {
    auto r = std::make_unique<dmr_result>();
    r->translated = true;
    r->dna_len.resize(nreads);
    const auto mo = min_orf_lengths(r_offsets, nreads, cfg ? cfg->min_orf : -1);
    for (uint32_t i = 0; i < nreads; ++i) r->dna_len[i] = r_offsets[i + 1] - r_offsets[i];
    // Six frames, masking and the target set on the device
    // (src/workflow.rs:112-127); the result keeps the host copy.
    ctx->eng->load_reads(reads, r_offsets, nreads, mo.data(), &r->targets);
    finish(ctx, r.get(), o, nreads, false);
    r->stats.ms_translate = ctx->eng->last_translate_ms();
    *out = r.release();
    return DMR_OK;
  });
}

int dmr_six_frames(dmr_ctx* ctx, const char* reads, const uint64_t* r_offsets, uint32_t nreads, int32_t min_orf,
                   uint8_t* letters, uint64_t cap, uint64_t* frame_offsets, uint64_t* n_letters) {
  if (!ctx || !r_offsets || !n_letters) return fail(DMR_E_ARG, "NULL argument");
  if (nreads && !reads && r_offsets[nreads] > r_offsets[0]) return fail(DMR_E_ARG, "NULL argument");
  for (uint32_t i = 0; i < nreads; ++i)
    if (r_offsets[i + 1] < r_offsets[i]) return fail(DMR_E_ARG, "read offsets are not monotone");
  return guard([&]() -> int {
    const auto mo = min_orf_lengths(r_offsets, nreads, min_orf);
    std::vector<uint8_t> v;
    std::vector<uint64_t> off;
    ctx->eng->six_frames(reads, r_offsets, nreads, mo.data(), &v, &off);
    *n_letters = v.size();
    if (frame_offsets) std::copy(off.begin(), off.end(), frame_offsets);
    if (v.size() > cap || (!v.empty() && !letters)) return fail(DMR_E_CAPACITY, "letter capacity too small");
    if (!v.empty()) std::memcpy(letters, v.data(), v.size());
    return DMR_OK;
  });
}

int64_t dmr_result_count(const dmr_result* r) { return r ? (int64_t)r->hsps.size() : -1; }

int dmr_result_hsps(const dmr_result* r, int64_t* o, double* d) {
  if (!r) return fail(DMR_E_ARG, "NULL result");
  for (size_t i = 0; i < r->hsps.size(); ++i) {
    const HspRec& h = r->hsps[i];
    int64_t* x = o + 16 * i;
    x[0] = h.query; x[1] = h.subject; x[2] = h.frame; x[3] = h.q_begin; x[4] = h.q_end;
    x[5] = h.s_begin; x[6] = h.s_end; x[7] = h.score; x[8] = h.identities; x[9] = h.positives;
    x[10] = h.mismatches; x[11] = h.gap_openings; x[12] = h.gaps; x[13] = h.length;
    x[14] = (int64_t)h.ops.size(); x[15] = 0;
    d[2 * i] = h.bits;
    d[2 * i + 1] = h.evalue;
  }
  return DMR_OK;
}

int dmr_result_ops(const dmr_result* r, int64_t i, uint32_t* ops, uint32_t cap, uint32_t* n) {
  if (!r || i < 0 || (size_t)i >= r->hsps.size()) return fail(DMR_E_ARG, "bad HSP index");
  const auto& v = r->hsps[i].ops;
  if (n) *n = (uint32_t)v.size();
  if (v.size() > cap) return fail(DMR_E_CAPACITY, "ops capacity too small");
  std::copy(v.begin(), v.end(), ops);
  return DMR_OK;
}

int dmr_result_stats(const dmr_result* r, double* o) {
  if (!r) return fail(DMR_E_ARG, "NULL result");
  const dmr::StageStats& s = r->stats;
  double v[21] = {(double)s.query_seeds, (double)s.runs, (double)s.hits, (double)s.hit_runs,
                  (double)s.segments, (double)s.groups, (double)s.jobs, (double)s.maybe,
                  (double)s.selected, (double)s.traceback_cells, s.ms_seed, s.ms_join,
                  s.ms_sort_hits, s.ms_ungapped, s.ms_bands, s.ms_sw, s.ms_select, s.ms_traceback,
                  s.ms_total, s.ms_host, s.ms_translate};
  std::memcpy(o, v, sizeof v);
  return DMR_OK;
}

int dmr_last_dp_cells(dmr_ctx* ctx, uint64_t* cells) {
  if (!ctx || !cells) return fail(DMR_E_ARG, "NULL argument");
  return guard([&]() -> int {
    *cells = ctx->eng->count_last_dp_cells();
    return DMR_OK;
  });
}

}  // extern "C"

// Tabular::write (src/output.rs:150-223).
namespace {

enum Field {
  kQseqid, kQlen, kSseqid, kSlen, kQstart, kQend, kSstart, kSend, kQseq, kSseq, kEvalue,
  kBitscore, kScore, kLength, kPident, kNident, kMismatch, kPositive, kGapopen, kGaps, kPpos,
  kQframe, kQcovhsp, kScovhsp, kCigar, kBtop, kQtitle, kStitle, kFullSseq, kQseqTranslated,
  kQstrand, kFieldCount
};
const char* const kFieldNames[kFieldCount] = {
    "qseqid", "qlen", "sseqid", "slen", "qstart", "qend", "sstart", "send", "qseq", "sseq",
    "evalue", "bitscore", "score", "length", "pident", "nident", "mismatch", "positive",
    "gapopen", "gaps", "ppos", "qframe", "qcovhsp", "scovhsp", "cigar", "btop", "qtitle",
    "stitle", "full_sseq", "qseq_translated", "qstrand"};

// The lines of a newline-separated list, in place: line k is text[start[k],
// start[k + 1] - 1) (a last line without its newline ends at the text's end).
// The id lists of a large database run to millions of lines per format call,
// so the newlines are found 16 bytes at a time (SSE2, the x86-64 baseline)
// and each line costs one 32-bit offset.
struct LineIndex {
  const char* text = nullptr;
  std::vector<uint32_t> start;
  size_t size() const { return start.empty() ? 0 : start.size() - 1; }
  std::pair<const char*, size_t> operator[](size_t k) const {
    return {text + start[k], (size_t)(start[k + 1] - 1 - start[k])};
  }
};

LineIndex lines_of(const char* s) {
  LineIndex out;
  if (!s) return out;
  const size_t len = std::strlen(s);
  if (len >= 0xFFFFFFFEull) throw dmr::Error(DMR_E_ARG, "id or header list longer than 4 GB");
  out.text = s;
  out.start.reserve(len / 12 + 2);
  out.start.push_back(0);
  const __m128i nl = _mm_set1_epi8('\n');
  size_t i = 0;
  for (; i + 16 <= len; i += 16) {
    unsigned m = (unsigned)_mm_movemask_epi8(
        _mm_cmpeq_epi8(_mm_loadu_si128(reinterpret_cast<const __m128i*>(s + i)), nl));
    while (m) {
      out.start.push_back((uint32_t)(i + (size_t)__builtin_ctz(m) + 1));
      m &= m - 1;
    }
  }
  for (; i < len; ++i)
    if (s[i] == '\n') out.start.push_back((uint32_t)(i + 1));
  if (out.start.back() < len) out.start.push_back((uint32_t)(len + 1));
  return out;
}

using Lines = LineIndex;

void render_lines(const dmr_result* r, const Lines& qid, const Lines& sid, const Lines& qtl,
                  const Lines& stl, const std::vector<int>& f, size_t i0, size_t i1,
                  std::string* text);

int render(const dmr_result* r, const char* qids, const char* sids, const char* qtitles,
           const char* stitles, const char* fields, std::vector<std::string>* parts_out) {
  const auto qid = lines_of(qids), sid = lines_of(sids);
  const auto qtl = lines_of(qtitles), stl = lines_of(stitles);
  std::vector<int> f;
  for (const std::string& n : split(fields, ' ')) {
    int code = -1;
    for (int k = 0; k < kFieldCount; ++k)
      if (n == kFieldNames[k]) code = k;
    if (code < 0) return fail(DMR_E_ARG, "Invalid output field: " + n);
    f.push_back(code);
  }
  if (f.empty())
    f = {kQseqid, kSseqid, kPident, kLength, kMismatch, kGapopen,
         kQstart, kQend, kSstart, kSend, kEvalue, kBitscore};
  bool want_qt = false, want_st = false;
  for (int code : f) {
    want_qt |= code == kQtitle;
    want_st |= code == kStitle;
  }
  // qtitle / stitle print the full headers (src/output.rs:201-202).
  if (want_qt && !qtitles) return fail(DMR_E_ARG, "qtitle needs the query headers");
  if (want_st && !stitles) return fail(DMR_E_ARG, "stitle needs the database headers");
  for (size_t i = 0; i < r->hsps.size(); ++i) {
    if (r->target_of[i] < 0) continue;
    const HspRec& h = r->hsps[i];
    if (h.query >= qid.size() || h.subject >= sid.size())
      return fail(DMR_E_ARG, "id list shorter than the query or database");
    if ((want_qt && h.query >= qtl.size()) || (want_st && h.subject >= stl.size()))
      return fail(DMR_E_ARG, "header list shorter than the query or database");
  }
  // Lines are independent: blocks of at least 64 lines (up to four per host
  // thread, claimed as threads come free) render into their own strings; the
  // blocks are joined in order.
  const size_t nh = r->hsps.size();
  const size_t nblocks = std::max<size_t>(1, std::min<size_t>(4 * dmr::host_threads(), (nh + 63) / 64));
  const size_t block = (nh + nblocks - 1) / nblocks;
  std::vector<std::string>& parts = *parts_out;
  parts.assign(nblocks, std::string());
  dmr::host_parallel_for(nblocks, 1, [&](uint64_t b0, uint64_t b1) {
    for (uint64_t blk = b0; blk < b1; ++blk)
      render_lines(r, qid, sid, qtl, stl, f, std::min(nh, blk * block),
                   std::min(nh, (size_t)(blk + 1) * block), &parts[blk]);
  });
  return DMR_OK;
}

// outfmt-6 lines of HSPs [i0, i1) (ids already checked) appended to out.
void render_lines(const dmr_result* r, const Lines& qid, const Lines& sid, const Lines& qtl,
                  const Lines& stl, const std::vector<int>& f, size_t i0, size_t i1,
                  std::string* text) {
  std::string& out = *text;
  out.reserve((i1 - i0) * 96);
  char nb[32];
  auto num = [&](long long x) { out.append(nb, std::to_chars(nb, nb + sizeof nb, x).ptr); };
  for (size_t i = i0; i < i1; ++i) {
    const HspRec& h = r->hsps[i];
    int ti = r->target_of[i];
    if (ti < 0) continue;
    const uint8_t* letters = r->targets.letters.data() + r->targets.offs[ti];
    size_t frame_len = r->targets.offs[ti + 1] - r->targets.offs[ti];
    const uint8_t* subj = r->db_res->data() + (*r->db_offs)[h.subject];
    size_t slen = (*r->db_offs)[h.subject + 1] - (*r->db_offs)[h.subject];
    size_t dna = r->translated ? r->dna_len[h.query] : frame_len;
    size_t qb = h.q_begin, qe = h.q_end;
    if (r->translated) {
      size_t fr = h.frame;
      if (fr < 3) {
        qb = fr + 3 * h.q_begin;
        qe = fr + 3 * h.q_end;
      } else {
        size_t o = fr - 3;
        qb = dna - (o + 3 * h.q_end);
        qe = dna - (o + 3 * h.q_begin);
      }
    }
    size_t qlen = r->translated ? dna : frame_len;
    auto aligned = [&](bool query_side) {
      size_t qi = h.q_begin, si = h.s_begin;
      for (uint32_t op : h.ops)
        for (uint32_t k = 0; k < (op >> 2); ++k) {
          uint32_t c = op & 3;
          if (c == 0) {
            out += dmr::decode_letter(query_side ? letters[qi] : subj[si]);
            ++qi;
            ++si;
          } else if (c == 1) {
            out += query_side ? dmr::decode_letter(letters[qi]) : '-';
            ++qi;
          } else {
            out += query_side ? '-' : dmr::decode_letter(subj[si]);
            ++si;
          }
        }
    };
    for (size_t k = 0; k < f.size(); ++k) {
      if (k) out += '\t';
      switch (f[k]) {
        case kQseqid: out.append(qid[h.query].first, qid[h.query].second); break;
        case kQtitle: out.append(qtl[h.query].first, qtl[h.query].second); break;
        case kQlen: num((long long)qlen); break;
        case kSseqid: out.append(sid[h.subject].first, sid[h.subject].second); break;
        case kStitle: out.append(stl[h.subject].first, stl[h.subject].second); break;
        case kSlen: num((long long)slen); break;
        case kQstart: num((long long)qb + 1); break;
        case kQend: num((long long)qe); break;
        case kSstart: num((long long)h.s_begin + 1); break;
        case kSend: num((long long)h.s_end); break;
        case kEvalue: put_evalue(out, h.evalue); break;
        case kBitscore: put_double(out, h.bits); break;
        case kScore: num(h.score); break;
        case kLength: num(h.length); break;
        case kPident: put_double(out, h.length ? h.identities * 100.0 / h.length : 0.0); break;
        case kNident: num(h.identities); break;
        case kMismatch: num(h.mismatches); break;
        case kPositive: num(h.positives); break;
        case kGapopen: num(h.gap_openings); break;
        case kGaps: num(h.gaps); break;
        case kPpos: put_double(out, h.length ? h.positives * 100.0 / h.length : 0.0); break;
        case kQframe: num(h.frame < 3 ? (int)h.frame + 1 : 2 - (int)h.frame); break;
        case kQcovhsp:
          if (qlen == 0) out += "0.0";
          else put_double(out, (double)(qe - qb) * 100.0 / (double)qlen);
          break;
        case kScovhsp:
          if (slen == 0) out += "0.0";
          else put_double(out, (double)(h.s_end - h.s_begin) * 100.0 / (double)slen);
          break;
        case kCigar:
          for (uint32_t op : h.ops) {
            num(op >> 2);
            out += "MID"[op & 3];
          }
          break;
        case kBtop: {
          uint32_t run = 0;
          size_t qi = h.q_begin, si = h.s_begin;
          auto flush = [&]() {
            if (run) num(run);
            run = 0;
          };
          for (uint32_t op : h.ops)
            for (uint32_t t = 0; t < (op >> 2); ++t) {
              uint32_t c = op & 3;
              if (c == 0) {
                char a = dmr::decode_letter(letters[qi++]), b = dmr::decode_letter(subj[si++]);
                if (a == b) {
                  ++run;
                } else {
                  flush();
                  out += a;
                  out += b;
                }
              } else if (c == 1) {
                flush();
                out += dmr::decode_letter(letters[qi++]);
                out += '-';
              } else {
                flush();
                out += '-';
                out += dmr::decode_letter(subj[si++]);
              }
            }
          flush();
          break;
        }
        case kQseq: aligned(true); break;
        case kSseq: aligned(false); break;
        case kFullSseq:
          for (size_t t = 0; t < slen; ++t) out += dmr::decode_letter(subj[t]);
          break;
        case kQseqTranslated: {
          size_t e = std::min<size_t>(h.q_end, frame_len), b = std::min<size_t>(h.q_begin, e);
          for (size_t t = b; t < e; ++t) out += dmr::decode_letter(letters[t]);
          break;
        }
        case kQstrand: out += h.frame >= 3 ? "-" : "+"; break;
        default: break;
      }
    }
    out += '\n';
  }
}

}  // namespace

extern "C" {

int dmr_result_format(const dmr_result* r, const char* qids, const char* sids, const char* qtitles,
                      const char* stitles, const char* fields, char* buf, uint64_t cap, uint64_t* len) {
  if (!r || !len) return fail(DMR_E_ARG, "NULL argument");
  return guard([&]() -> int {
    // The text is rendered once per distinct request (the size query and
    // the fill of one request render once) and kept as its blocks. A NULL
    // title list is kept as "\x01", which no list of lines equals.
    auto kept_as = [](const char* s) { return std::string(s ? s : "\x01"); };
    auto same = [&](const std::string& kept, const char* s) { return kept == kept_as(s); };
    bool fresh = false;
    if (!r->fmt_valid || !same(r->fmt_q, qids) || !same(r->fmt_s, sids) || !same(r->fmt_qt, qtitles) ||
        !same(r->fmt_st, stitles) || !same(r->fmt_f, fields)) {
      r->fmt_valid = false;
      const int rc = render(r, qids, sids, qtitles, stitles, fields, &r->fmt_parts);
      if (rc != DMR_OK) return rc;
      r->fmt_len = 0;
      for (const auto& pt : r->fmt_parts) r->fmt_len += pt.size();
      fresh = true;
    }
    *len = r->fmt_len;
    const bool fits = buf && cap >= r->fmt_len;
    if (fits) {
      char* o = buf;
      for (const auto& pt : r->fmt_parts) {
        std::memcpy(o, pt.data(), pt.size());
        o += pt.size();
      }
    }
    // A request that did not fit (or asked for the size) keeps the text and
    // the lists it came from, for the fill that follows; one served in a
    // single call keeps nothing (the lists of a large database run to
    // megabytes).
    if (fresh && !fits) {
      r->fmt_q = kept_as(qids);
      r->fmt_s = kept_as(sids);
      r->fmt_qt = kept_as(qtitles);
      r->fmt_st = kept_as(stitles);
      r->fmt_f = kept_as(fields);
      r->fmt_valid = true;
    }
    if (buf && !fits) return fail(DMR_E_CAPACITY, "buffer too small");
    return DMR_OK;
  });
}

int dmr_last_class_profile(dmr_ctx* ctx, uint64_t* jobs, uint64_t* cells, double* ms) {
  static_assert(DMR_SCORE_CLASSES == dmr::Engine::kClasses, "class count");
  if (!ctx || !jobs || !cells || !ms) return fail(DMR_E_ARG, "NULL argument");
  return guard([&]() -> int {
    ctx->eng->last_class_profile(jobs, cells, ms);
    return DMR_OK;
  });
}

int dmr_score_layout(dmr_ctx* ctx, uint32_t* split_mask, uint32_t* f16) {
  if (!ctx) return fail(DMR_E_ARG, "NULL argument");
  const dmr::ScoreLayout& L = ctx->eng->layout();
  if (split_mask) {
    *split_mask = 0;
    for (int c = 0; c < dmr::kExactClasses; ++c)
      if (L.lanes[c] > 1) *split_mask |= 1u << c;
  }
  if (f16) *f16 = L.f16 ? 1u : 0u;
  return DMR_OK;
}

void dmr_result_destroy(dmr_result* r) { delete r; }

}  // extern "C"
