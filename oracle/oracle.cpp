// Oracle: CPU restatement of diamond-rust's search core (test infrastructure only).
//
// Each function restates the behaviour of the reference function cited above
// it (paths are relative to the reference repository). The product never
// links this file; tests, the smoke check and the CPU-baseline leg of the
// benchmark load it through the C API at the bottom, to compare results with.
#include "oracle.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <thread>
#include <tuple>
#include <unordered_map>

#include "oracle_data.hpp"

namespace orc {

static const char kSymbols[] = "ARNDCQEGHILKMFPSTWYVBJZX*";

// src/alphabet.rs:24-46 — ASCII -> NCBI code, unknown -> X (23), U -> C, O -> K.
struct EncTable {
  uint8_t t[256];
  EncTable() {
    for (int i = 0; i < 256; ++i) t[i] = 23;
    for (int i = 0; i < 25; ++i) {
      unsigned char c = (unsigned char)kSymbols[i];
      t[c] = (uint8_t)i;
      if (c >= 'A' && c <= 'Z') t[c - 'A' + 'a'] = (uint8_t)i;
    }
    t['U'] = t['u'] = 4;
    t['O'] = t['o'] = 11;
  }
};
static const EncTable kEnc;

uint8_t encode_byte(uint8_t c) { return kEnc.t[c]; }

// Rust u8::is_ascii_whitespace: space, \t, \n, \x0C, \r (not \x0B).
static bool ascii_ws(uint8_t c) { return c == ' ' || c == '\t' || c == '\n' || c == 0x0C || c == '\r'; }

// src/alphabet.rs:54-63
std::vector<uint8_t> encode(const std::string& raw) {
  std::vector<uint8_t> out;
  out.reserve(raw.size());
  for (unsigned char c : raw) {
    if (ascii_ws(c) || c == '-' || c == '.') continue;
    out.push_back(kEnc.t[c]);
  }
  return out;
}

// src/alphabet.rs:66-68
char decode_byte(uint8_t l) { return kSymbols[l < 24 ? l : 24]; }

// src/reduction.rs:19-36
Reduction reduction_parse(const std::string& def) {
  Reduction r;
  std::memset(r.map, 0xFF, sizeof(r.map));
  uint8_t bucket = 0;
  size_t i = 0;
  while (i < def.size()) {
    while (i < def.size() && ascii_ws((uint8_t)def[i])) ++i;
    if (i >= def.size()) break;
    while (i < def.size() && !ascii_ws((uint8_t)def[i])) {
      uint8_t l = encode_byte((uint8_t)def[i]);
      if (l < 20) r.map[l] = bucket;
      ++i;
    }
    ++bucket;
  }
  r.size = bucket;
  return r;
}

// src/reduction.rs:48-79
Reduction reduction_for(const std::string& sensitivity) {
  if (sensitivity == "very-sensitive") return reduction_parse("A G P CH ST NDE RQK FWY ILMV");
  return reduction_parse("A KR EDNQ C G H ILVM FYW P ST");
}

// src/shape.rs:23-43
Shape shape_parse(const std::string& code) {
  Shape s;
  s.code = code;
  for (size_t i = 0; i < code.size(); ++i) {
    if (code[i] == '1') {
      s.positions.push_back((uint32_t)i);
      s.span = (uint32_t)i + 1;
    } else if (code[i] == '0') {
      s.span = (uint32_t)i + 1;
    } else {
      std::fprintf(stderr, "invalid seed shape %s\n", code.c_str());
      std::abort();
    }
  }
  s.weight = (uint32_t)s.positions.size();
  return s;
}

// src/shape.rs:60-74
bool Shape::packed(const uint8_t* seq, size_t len, size_t pos, const Reduction& r, uint64_t* key) const {
  if (pos + span > len) return false;
  uint64_t k = 0;
  for (uint32_t off : positions) {
    uint8_t l = seq[pos + off];
    if (l >= 20) return false;
    k = k * r.size + r.bucket(l);
  }
  *key = k;
  return true;
}

// src/shape.rs:78-220 (preset shape codes)
std::vector<Shape> shapes_for(const std::string& sens, size_t limit) {
  static const std::vector<std::string> fast = {"1101110101101111"};
  static const std::vector<std::string> def = {"111101110111", "111011010010111"};
  static const std::vector<std::string> mid = {
      "11110110111",    "1101100111101",    "1110010101111",    "11010101100111",
      "11101110001011", "1110100100010111", "1101000011010111", "1110011000011011"};
  static const std::vector<std::string> sensitive = {
      "1011110111",       "110100100010111",  "11001011111",      "101110001111",
      "11011101100001",   "1111010010101",    "111001001001011",  "10101001101011",
      "111101010011",     "1111000010000111", "1100011011011",    "1101010000011011",
      "1110001010101001", "110011000110011",  "11011010001101",   "1101001100010011"};
  static const std::vector<std::string> very = {
      "11110111",          "11100100111",         "110010101011",        "11010001001011",
      "10101100001101",    "110100100100011",     "1010010100010011",    "1100101000001011",
      "11100000100010101", "11000100010010011",   "11010000001000111",   "110001001000010011",
      "1010001000100001011", "1100010100000010011", "1100100000101000011", "1101000010000001011"};
  static const std::vector<std::string> ultra = {
      "1111111",          "11101111",         "110011111",        "110110111",
      "111111001",        "1010111011",       "1011110101",       "1111000111",
      "10011110011",      "10101101101",      "10111010101",      "11001010111",
      "11001100111",      "11010101101",      "11110001011",      "100111010011",
      "101100110101",     "101110000111",     "110100101011",     "110110001101",
      "111000110011",     "1010001011011",    "1010101000111",    "1010110100011",
      "1100100110011",    "1100101001011",    "1101001100101",    "1101010101001",
      "1110001010101",    "1110010010011",    "10100001101101",   "11000100010111",
      "11010000100111",   "11010100110001",   "11101000011001",   "11110000001101",
      "11110100000011",   "101001000001111",  "110000100101011",  "110010010000111",
      "110101100001001",  "110110000010011",  "111001000100011",  "111100000100101",
      "1000110010010101", "1001000100101101", "1001000110011001", "1010001001001011",
      "1010001010010011", "1010010001010101", "1010010100010011", "1010010101001001",
      "1010100000101011", "1010100011000101", "1011000010001011", "1100010000111001",
      "1100010010001011", "1100100001001011", "1100100100100011", "1100110000001101",
      "1101000100010011", "1101000110000101", "1110000001010011", "1110100000010101"};
  const std::vector<std::string>* codes = &def;
  if (sens == "faster" || sens == "fast") codes = &fast;
  else if (sens == "mid-sensitive") codes = &mid;
  else if (sens == "sensitive" || sens == "more-sensitive") codes = &sensitive;
  else if (sens == "very-sensitive") codes = &very;
  else if (sens == "ultra-sensitive") codes = &ultra;
  size_t n = limit == 0 ? codes->size() : std::min(limit, codes->size());
  std::vector<Shape> out;
  for (size_t i = 0; i < n; ++i) out.push_back(shape_parse((*codes)[i]));
  return out;
}

// src/seed.rs:29-38 — t[n] = t[n-1] + ln n, summed in increasing n.
static const double* lnfact_table() {
  static double t[33];
  static bool init = false;
  if (!init) {
    t[0] = t[1] = 0.0;
    for (int n = 2; n < 33; ++n) t[n] = t[n - 1] + std::log((double)n);
    init = true;
  }
  return t;
}
double ln_factorial(int n) { return lnfact_table()[n]; }

// src/seed.rs:46-70 — counts per bucket, entropy subtracted in bucket order.
bool seed_is_complex(const uint8_t* seq, const Shape& s, const Reduction& r, double cut) {
  if (cut <= 0.0) return true;
  uint32_t count[32] = {0};
  for (uint32_t off : s.positions) {
    uint8_t l = seq[off];
    if (l >= 20) return false;
    uint8_t b = r.bucket(l);
    if (b >= 32) return false;
    count[b] += 1;
  }
  const double* lf = lnfact_table();
  double e = lf[s.weight];
  for (int g = 0; g < 32; ++g)
    if (count[g] > 0) e -= lf[count[g]];
  return e >= cut;
}

// src/seed.rs:76-104 — enumerate then sort by key (order among equal keys is
// unspecified in the reference; downstream consumers re-sort, H13).
std::vector<SeedEntry> seed_build(const std::vector<const uint8_t*>& seqs,
                                  const std::vector<uint32_t>& lens, const Shape& s,
                                  const Reduction& r, uint32_t base, double cut) {
  std::vector<SeedEntry> out;
  for (size_t i = 0; i < seqs.size(); ++i) {
    size_t len = lens[i];
    if (len < s.span) continue;
    for (size_t pos = 0; pos + s.span <= len; ++pos) {
      uint64_t key;
      if (!s.packed(seqs[i], len, pos, r, &key)) continue;
      if (!seed_is_complex(seqs[i] + pos, s, r, cut)) continue;
      out.push_back({key, base + (uint32_t)i, (uint32_t)pos});
    }
  }
  std::sort(out.begin(), out.end(), [](const SeedEntry& a, const SeedEntry& b) {
    if (a.key != b.key) return a.key < b.key;
    if (a.seq != b.seq) return a.seq < b.seq;
    return a.pos < b.pos;
  });
  return out;
}

// Per-target map subject -> hits in emission order (HashMap<u32, Vec<SeedHit>>).
struct HitMap {
  std::unordered_map<uint32_t, std::vector<SeedHit>> m;
};

// src/seed.rs:110-143 — equal-key runs; a run is dropped when nq*nr > max_expand.
static void join_into(const std::vector<SeedEntry>& q, const std::vector<SeedEntry>& r,
                      std::vector<HitMap>& out, size_t max_expand, uint64_t* nhits) {
  size_t i = 0, j = 0;
  while (i < q.size() && j < r.size()) {
    if (q[i].key < r[j].key) {
      ++i;
    } else if (q[i].key > r[j].key) {
      ++j;
    } else {
      uint64_t key = q[i].key;
      size_t qi = i, rj = j;
      while (qi < q.size() && q[qi].key == key) ++qi;
      while (rj < r.size() && r[rj].key == key) ++rj;
      if ((qi - i) * (rj - j) <= max_expand) {
        for (size_t a = i; a < qi; ++a) {
          auto& bucket = out[q[a].seq];
          for (size_t b = j; b < rj; ++b) {
            bucket.m[r[b].seq].push_back({q[a].seq, q[a].pos, r[b].seq, r[b].pos});
            ++*nhits;
          }
        }
      }
      i = qi;
      j = rj;
    }
  }
}

// src/seed.rs:146-164
std::vector<std::pair<size_t, size_t>> chunks(size_t n, size_t parts) {
  std::vector<std::pair<size_t, size_t>> out;
  if (parts < 1) parts = 1;
  if (n == 0) return out;
  parts = std::min(parts, n);
  size_t start = 0;
  for (size_t p = 0; p < parts; ++p) {
    size_t end = (p + 1 == parts) ? n : start + n / parts + (p < n % parts ? 1 : 0);
    out.push_back({start, end});
    start = end;
  }
  return out;
}

// ---------------------------------------------------------------------------
// ALP E-values (src/alp.rs:60-237), GSL log_erfc restated.
// ---------------------------------------------------------------------------
static const double kLn2 = 0.6931471805599453;  // std::f64::consts::LN_2
static const double kNatCutOff = 2.0;
static const double kRoot6Eps = 2.4607833005759251e-3;
static const double kSqrtPi = 1.772453850905516;
static const double kInvSqrt2Pi = 0.3989422804014327;
static const double kPi = 3.141592653589793;

static double cheb_eval(const double* c, int len, double x) {
  double y2 = 2.0 * x, d = 0.0, dd = 0.0;
  for (int j = len - 1; j >= 1; --j) {
    double tmp = d;
    d = y2 * d - dd + c[j];
    dd = tmp;
  }
  return x * d - dd + 0.5 * c[0];
}

static double erfc8_sum(double x) {
  double num = kErfc8P[5];
  for (int i = 4; i >= 0; --i) num = x * num + kErfc8P[i];
  double den = kErfc8Q[6];
  for (int i = 5; i >= 0; --i) den = x * den + kErfc8Q[i];
  return num / den;
}

static double erfc_gsl(double x) {
  double ax = std::fabs(x), v;
  if (ax <= 1.0) v = cheb_eval(kErfcXlt1, 20, 2.0 * ax - 1.0);
  else if (ax <= 5.0) v = std::exp(-x * x) * cheb_eval(kErfcX15, 25, 0.5 * (ax - 3.0));
  else if (ax < 10.0) v = std::exp(-x * x) / ax * cheb_eval(kErfcX510, 20, (2.0 * ax - 15.0) / 5.0);
  else v = erfc8_sum(ax) * std::exp(-x * x);
  return x < 0.0 ? 2.0 - v : v;
}

double log_erfc(double x) {
  if (x < -8.0) return kLn2;
  if (x * x < 10.0 * kRoot6Eps) {
    double y = x / kSqrtPi;
    const double c3 = (4.0 - kPi) / 3.0;
    const double c4 = 2.0 * (1.0 - kPi / 3.0);
    const double c5 = -0.001829764677455021, c6 = 0.02629651521057465, c7 = -0.01621575378835404;
    const double c8 = 0.00125993961762116, c9 = 0.00556964649138, c10 = -0.0045563339802;
    const double c11 = 0.0009461589032, c12 = 0.0013200243174, c13 = -0.00142906, c14 = 0.00048204;
    double series = c8 + y * (c9 + y * (c10 + y * (c11 + y * (c12 + y * (c13 + c14 * y)))));
    series = y * (1.0 + y * (1.0 + y * (c3 + y * (c4 + y * (c5 + y * (c6 + y * (c7 + y * series)))))));
    return -2.0 * series;
  }
  if (x > 8.0) return std::log(erfc8_sum(x)) - x * x;
  return std::log(erfc_gsl(x));
}

static double ln_or_neg_inf(double x) { return x > 0.0 ? std::log(x) : -INFINITY; }

static double log_sum(double a, double b) {
  double hi = a >= b ? a : b, lo = a >= b ? b : a;
  if (!std::isfinite(hi)) return hi;
  if (lo == -INFINITY) return hi;
  return hi + std::log(1.0 + std::exp(lo - hi));
}

static double log_diff(double a, double b) {
  if (a < b) return -INFINITY;
  if (!std::isfinite(a)) return a;
  if (b == -INFINITY) return a;
  double t = 1.0 - std::exp(b - a);
  return t <= 0.0 ? -INFINITY : a + std::log(t);
}

struct Alp {
  double lambda, k, a, b, alpha, beta, a_j, b_j, alpha_j, beta_j, sigma, tau, vi_thr, vj_thr, c_thr;
};

// src/alp.rs:60-92
static Alp alp_new(const Matrix& m) {
  Alp p;
  double g = (double)(m.gap_open + m.gap_extend);
  p.lambda = m.lambda;
  p.k = m.k;
  p.a = m.alpha;
  p.b = 2.0 * g * (m.u_alpha - m.alpha);
  p.alpha = m.alpha_v;
  p.beta = 2.0 * g * (m.u_alpha_v - m.alpha_v);
  p.a_j = p.a;
  p.b_j = p.b;
  p.alpha_j = p.alpha;
  p.beta_j = p.beta;
  p.sigma = m.sigma;
  p.tau = 2.0 * g * (m.u_alpha_v - m.sigma);
  p.vi_thr = std::max(kNatCutOff * m.alpha_v / m.lambda, 0.0);
  p.vj_thr = std::max(kNatCutOff * m.alpha_v / m.lambda, 0.0);
  p.c_thr = std::max(kNatCutOff * m.sigma / m.lambda, 0.0);
  return p;
}

// src/alp.rs:125-163
static double log_area(const Alp& p, double y, double m, double n) {
  double m_li_y = m - (p.a * y + p.b);
  double vi_y = std::max(p.vi_thr, p.alpha * y + p.beta);
  double sqrt_vi = std::sqrt(vi_y);
  double m_f = sqrt_vi == 0.0 ? 1e100 : m_li_y / sqrt_vi;
  double log_p_m = std::log(0.5) + log_erfc(-std::sqrt(0.5) * m_f);
  double log_e_m = std::log(kInvSqrt2Pi) - 0.5 * m_f * m_f;
  double log_sqrt_e_m = ln_or_neg_inf(sqrt_vi) + log_e_m;
  double log_p1 = m_li_y < 0.0 ? log_diff(log_sqrt_e_m, ln_or_neg_inf(-m_li_y) + log_p_m)
                               : log_sum(log_sqrt_e_m, ln_or_neg_inf(m_li_y) + log_p_m);
  double n_lj_y = n - (p.a_j * y + p.b_j);
  double vj_y = std::max(p.vj_thr, p.alpha_j * y + p.beta_j);
  double sqrt_vj = std::sqrt(vj_y);
  double n_f = sqrt_vj == 0.0 ? 1e100 : n_lj_y / sqrt_vj;
  double log_p_n = std::log(0.5) + log_erfc(-std::sqrt(0.5) * n_f);
  double log_e_n = std::log(kInvSqrt2Pi) - 0.5 * n_f * n_f;
  double log_sqrt_e_n = ln_or_neg_inf(sqrt_vj) + log_e_n;
  double log_p2 = n_lj_y < 0.0 ? log_diff(log_sqrt_e_n, ln_or_neg_inf(-n_lj_y) + log_p_n)
                               : log_sum(log_sqrt_e_n, ln_or_neg_inf(n_lj_y) + log_p_n);
  double c_y = std::max(p.c_thr, p.sigma * y + p.tau);
  double log_c_y = ln_or_neg_inf(c_y);
  return log_sum(log_p1 + log_p2, log_c_y + log_p_m + log_p_n);
}

// src/alp.rs:95-122
static double alp_evalue(const Alp& p, int raw, size_t qlen, size_t slen, uint64_t db) {
  double y = (double)raw;
  double m = (double)std::max<size_t>(slen, 1);
  double n = (double)std::max<size_t>(qlen, 1);
  double la = log_area(p, y, m, n);
  if (!std::isfinite(la)) {
    double ex = -p.lambda * y + std::log(p.k);
    return ex > 700.0 ? INFINITY : (double)qlen * (double)db * std::exp(ex);
  }
  double log_e = la + std::log(p.k) - p.lambda * y;
  return log_e > 700.0 ? INFINITY : std::exp(log_e) * (double)db / m;
}

// src/matrix.rs:42-98 (BLOSUM62 only; other built-in matrices are out of scope)
bool matrix_blosum62(int gap_open, int gap_extend, Matrix* out) {
  if (gap_open == -1) gap_open = 11;
  if (gap_extend == -1) gap_extend = 1;
  if (gap_open <= 0 || gap_extend <= 0) return false;
  const GapParams* gp = nullptr;
  for (const auto& g : kBlosum62Gaps)
    if (g.gap_open == gap_open && g.gap_extend == gap_extend) gp = &g;
  if (!gp) return false;
  std::memcpy(out->s, kBlosum62, sizeof(kBlosum62));
  out->gap_open = gap_open;
  out->gap_extend = gap_extend;
  out->lambda = gp->lambda;
  out->k = gp->k;
  out->ln_k = std::log(gp->k);
  out->alpha = gp->alpha;
  out->alpha_v = gp->alpha_v;
  out->sigma = gp->sigma;
  out->u_alpha = kBlosum62UngappedAlpha;
  out->u_alpha_v = kBlosum62UngappedAlphaV;
  return true;
}

// src/matrix.rs:182-200
int Matrix::raw_score(double bits) const { return (int)std::ceil((bits * kLn2 + ln_k) / lambda); }
double Matrix::bit_score(int raw) const { return (lambda * (double)raw - ln_k) / kLn2; }
double Matrix::evalue(int raw, size_t qlen, size_t slen, uint64_t db) const {
  Alp p = alp_new(*this);
  return alp_evalue(p, raw, qlen, slen, db);
}

// src/ungapped.rs:38-88 — left extension from (qa-1, sa-1), then right from
// (qa, sa) continuing from the left-hand best score.
Segment xdrop(const uint8_t* q, size_t qlen, const uint8_t* s, size_t slen, size_t qa, size_t sa,
              const Matrix& m, int xd) {
  int32_t score = 0, st = 0;
  size_t n = 1, delta = 0, len = 0;
  long qi = (long)qa - 1, si = (long)sa - 1;
  while (score - st < xd && qi >= 0 && si >= 0) {
    st += m.score(q[qi], s[si]);
    if (st > score) {
      score = st;
      delta = n;
    }
    --qi;
    --si;
    ++n;
  }
  qi = (long)qa;
  si = (long)sa;
  st = score;
  n = 1;
  while (score - st < xd && (size_t)qi < qlen && (size_t)si < slen) {
    st += m.score(q[qi], s[si]);
    if (st > score) {
      score = st;
      len = n;
    }
    ++qi;
    ++si;
    ++n;
  }
  return Segment{qa - delta, sa - delta, len + delta, score};
}

// src/gapped.rs:27-32
size_t band_for(size_t len) {
  static const size_t thr[7] = {50, 100, 150, 200, 250, 350, 500};
  static const size_t band[8] = {12, 16, 30, 30, 30, 40, 64, 64};
  size_t b = 0;
  for (size_t t : thr)
    if (len >= t) ++b;
  return band[b];
}

static const int32_t kNeg = INT32_MIN / 4;  // src/gapped.rs:12

// src/gapped.rs:67-280 — row-major banded Gotoh with 4-bit trace codes
// (choice | e_ext<<2 | f_ext<<3), first strict maximum, then traceback.
bool extend_span(const uint8_t* q, size_t n, const uint8_t* s, size_t m, long d0, long d1,
                 const Matrix& mat, Aligned* out) {
  if (n == 0 || m == 0) return false;
  if (d0 > d1) std::swap(d0, d1);
  std::vector<size_t> j0(n + 1), j1(n + 1), off(n + 2, 0);
  size_t max_w = 1;
  for (size_t i = 0; i <= n; ++i) {
    long lo = std::clamp((long)i - d1, 0L, (long)m);
    long hi = std::clamp((long)i - d0, 0L, (long)m);
    j0[i] = (size_t)lo;
    j1[i] = (size_t)std::max(hi, lo);
    max_w = std::max(max_w, j1[i] - j0[i] + 1);
  }
  for (size_t i = 0; i <= n; ++i) off[i + 1] = off[i] + (j1[i] - j0[i] + 1);
  std::vector<uint8_t> trace(off[n + 1], 0);
  const int32_t oc = mat.gap_open + mat.gap_extend, ge = mat.gap_extend;
  std::vector<int32_t> ph(max_w, 0), pe(max_w, kNeg), ch(max_w, 0), ce(max_w, kNeg);
  int32_t best = 0;
  size_t bi = 0, bj = 0;
  uint64_t cells = 0;
  for (size_t i = 1; i <= n; ++i) {
    uint8_t qi = q[i - 1];
    size_t a = j0[i], b = j1[i], pa = j0[i - 1], pb = j1[i - 1];
    int32_t f_left = kNeg;
    for (size_t j = a; j <= b; ++j) {
      ++cells;
      int32_t d = kNeg;
      if (j >= 1) {
        int32_t hd = (j > pa && j - 1 <= pb) ? ph[j - 1 - pa] : 0;
        d = hd + mat.score(qi, s[j - 1]);
      }
      int32_t e = kNeg;
      uint8_t e_ext = 0;
      if (j >= pa && j <= pb) {
        int32_t hp = ph[j - pa];
        if (hp > kNeg / 2 && hp - oc > e) e = hp - oc;
        int32_t pev = pe[j - pa];
        if (pev > kNeg / 2 && pev - ge > e) {
          e = pev - ge;
          e_ext = 1;
        }
      }
      int32_t f = kNeg;
      uint8_t f_ext = 0;
      if (j > a) {
        int32_t hl = ch[j - 1 - a];
        if (hl > kNeg / 2 && hl - oc > f) f = hl - oc;
        if (f_left > kNeg / 2 && f_left - ge > f) {
          f = f_left - ge;
          f_ext = 1;
        }
      }
      int32_t hv = 0;
      uint8_t choice = 0;
      if (d > hv) { hv = d; choice = 1; }
      if (e > hv) { hv = e; choice = 2; }
      if (f > hv) { hv = f; choice = 3; }
      ch[j - a] = hv;
      ce[j - a] = e;
      trace[off[i] + j - a] = (uint8_t)(choice | (e_ext << 2) | (f_ext << 3));
      f_left = f;
      if (hv > best) {
        best = hv;
        bi = i;
        bj = j;
      }
    }
    std::swap(ph, ch);
    std::swap(pe, ce);
  }
  out->cells = cells;
  if (best <= 0) return false;
  // Traceback (src/gapped.rs:205-266).
  std::vector<Op> rev;
  Stats st;
  int state = 0;
  size_t i = bi, j = bj;
  bool in_gap = false;
  while (i > 0 && j > 0) {
    if (j < j0[i] || j > j1[i]) break;
    uint8_t t = trace[off[i] + j - j0[i]];
    if (state == 0) {
      uint8_t c = t & 3;
      if (c == 0) break;
      if (c == 1) {
        uint8_t x = q[i - 1], y = s[j - 1];
        st.length++;
        if (x == y) st.identities++;
        else st.mismatches++;
        if (mat.score(x, y) > 0) st.positives++;
        rev.push_back(OP_M);
        in_gap = false;
        --i;
        --j;
      } else if (c == 2) {
        state = 1;
      } else {
        state = 2;
      }
    } else if (state == 1) {
      st.length++;
      st.gaps++;
      if (!in_gap) {
        st.gap_openings++;
        in_gap = true;
      }
      rev.push_back(OP_I);
      if (((t >> 2) & 1) == 0) state = 0;
      --i;
    } else {
      st.length++;
      st.gaps++;
      if (!in_gap) {
        st.gap_openings++;
        in_gap = true;
      }
      rev.push_back(OP_D);
      if (((t >> 3) & 1) == 0) state = 0;
      --j;
    }
  }
  std::reverse(rev.begin(), rev.end());
  out->ops.clear();
  for (Op o : rev) {
    if (!out->ops.empty() && out->ops.back().first == o) out->ops.back().second++;
    else out->ops.push_back({o, 1});
  }
  if (out->ops.empty()) return false;
  out->score = best;
  out->q_begin = i;
  out->q_end = bi;
  out->s_begin = j;
  out->s_end = bj;
  out->stats = st;
  return true;
}

// src/search.rs:55-79
double SearchConfig::seed_cut_for() const {
  if (seed_cut >= 0) return seed_cut;
  if (sensitivity == "faster" || sensitivity == "fast") return 0.9;
  if (sensitivity == "mid-sensitive" || sensitivity == "sensitive" ||
      sensitivity == "more-sensitive" || sensitivity == "very-sensitive" ||
      sensitivity == "ultra-sensitive")
    return 1.0;
  return 0.8;
}
size_t SearchConfig::index_chunks_for() const {
  if (index_chunks >= 0) return (size_t)std::max(index_chunks, 1L);
  if (sensitivity == "very-sensitive" || sensitivity == "ultra-sensitive") return 1;
  return 4;
}

// src/search.rs:273-282 — score desc, E-value asc, subject asc.
static bool hsp_less(const Hsp& a, const Hsp& b) {
  if (a.score != b.score) return a.score > b.score;
  if (a.evalue < b.evalue) return true;
  if (b.evalue < a.evalue) return false;
  return a.subject < b.subject;
}

// Reference-stable insertion sort (what Rust's sort_unstable does for
// slices of at most 20 elements). Longer slices with ties are counted as
// events whose permutation the reference leaves unspecified.
template <class T, class Less>
static void small_stable_sort(std::vector<T>& v, Less less) {
  std::stable_sort(v.begin(), v.end(), less);
}

// src/search.rs:287-354
static bool gapped_hsp(const uint8_t* q, size_t qlen, const uint8_t* s, size_t slen,
                       std::pair<size_t, size_t> anchor, long d0, long d1, size_t band,
                       const SearchConfig& cfg, uint64_t db_letters, const Target& t,
                       uint32_t sid, Hsp* out, SearchCounters* c) {
  size_t row_cells = (size_t)std::max(d1 - d0, 0L) + 1;
  size_t budget = 16000000;
  size_t max_rows = std::max(budget / row_cells, 2 * band + 16);
  size_t q0 = 0, q1 = qlen;
  if (q1 - q0 > max_rows) {
    q0 = anchor.first > max_rows / 2 ? anchor.first - max_rows / 2 : 0;
    q1 = std::min(q0 + max_rows, qlen);
    q0 = q1 > max_rows ? q1 - max_rows : 0;
    c->clipped_dps++;
  }
  long shift = (long)q0;
  Aligned a;
  c->dps++;
  bool ok = extend_span(q + q0, q1 - q0, s, slen, d0 - shift, d1 - shift, cfg.matrix, &a);
  c->dp_cells += a.cells;
  if (!ok) return false;
  out->query = t.query;
  out->subject = sid;
  out->frame = t.frame;
  out->q_begin = q0 + a.q_begin;
  out->q_end = q0 + a.q_end;
  out->s_begin = a.s_begin;
  out->s_end = a.s_end;
  out->score = a.score;
  out->bits = cfg.matrix.bit_score(a.score);
  out->evalue = cfg.matrix.evalue(a.score, qlen, slen, db_letters);
  out->stats = a.stats;
  out->ops = std::move(a.ops);
  return true;
}

// src/search.rs:148-271
static std::vector<Hsp> align_target(const Database& db, const Target& t, const HitMap& hits,
                                     const SearchConfig& cfg, SearchCounters* c) {
  std::vector<Hsp> outv;
  const auto& query = t.letters;
  if (query.empty() || hits.m.empty()) return outv;
  const uint64_t db_letters = db.total_letters();
  const Matrix& mat = cfg.matrix;
  const int cutoff = 19;
  const int xd = cfg.ungapped_xdrop();
  std::vector<uint32_t> subjects;
  for (auto& kv : hits.m) subjects.push_back(kv.first);
  std::sort(subjects.begin(), subjects.end());
  std::vector<std::vector<Hsp>> found;
  for (uint32_t sid : subjects) {
    const uint8_t* subj = db.seq(sid);
    size_t slen = db.seq_len(sid);
    if (slen == 0) continue;
    std::vector<SeedHit> seeds = hits.m.at(sid);
    std::sort(seeds.begin(), seeds.end(), [](const SeedHit& a, const SeedHit& b) {
      long da = (long)a.qpos - (long)a.spos, dbv = (long)b.qpos - (long)b.spos;
      if (da != dbv) return da < dbv;
      return a.spos < b.spos;
    });
    std::vector<Segment> segs;
    for (const SeedHit& h : seeds) {
      long diag = (long)h.qpos - (long)h.spos;
      if (!segs.empty() && segs.back().diag() == diag && segs.back().s_end() >= h.spos) continue;
      c->xdrops++;
      Segment sg = xdrop(query.data(), query.size(), subj, slen, h.qpos, h.spos, mat, xd);
      if (sg.score >= cutoff) segs.push_back(sg);
    }
    if (segs.empty()) continue;
    c->segments += segs.size();
    // segments.sort_unstable_by_key(Reverse(score)) then order by diag: the
    // band (d0, d1) and its best score do not depend on tie order; only the
    // anchor of a band whose maximum is tied does (used by window clipping).
    // This restatement takes the first maximum in (diag, spos) order.
    size_t band = band_for(query.size());
    std::vector<size_t> order(segs.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) {
      if (segs[x].diag() != segs[y].diag()) return segs[x].diag() < segs[y].diag();
      return segs[x].s_start < segs[y].s_start;
    });
    struct Band { long b0, b1; size_t best; bool tie; };
    std::vector<Band> bands;
    for (size_t idx : order) {
      long diag = segs[idx].diag();
      long b0 = diag - (long)band, b1 = diag + (long)band;
      if (!bands.empty() && b0 <= bands.back().b1) {
        Band& l = bands.back();
        l.b1 = std::max(l.b1, b1);
        if (segs[idx].score > segs[l.best].score) {
          l.best = idx;
          l.tie = false;
        } else if (segs[idx].score == segs[l.best].score) {
          l.tie = true;
        }
      } else {
        bands.push_back({b0, b1, idx, false});
      }
    }
    if (segs.size() > 20) {
      for (auto& b : bands)
        if (b.tie) c->anchor_tie_events++;
    }
    // bands.sort_unstable_by(best score desc): stable for <= 20 bands.
    if (bands.size() > 20) {
      std::vector<int32_t> sc;
      for (auto& b : bands) sc.push_back(segs[b.best].score);
      std::sort(sc.begin(), sc.end());
      if (std::adjacent_find(sc.begin(), sc.end()) != sc.end()) c->band_tie_events++;
    }
    std::stable_sort(bands.begin(), bands.end(), [&](const Band& x, const Band& y) {
      return segs[x.best].score > segs[y.best].score;
    });
    c->bands += bands.size();
    std::vector<Hsp> hsps;
    for (const Band& b : bands) {
      if (hsps.size() >= cfg.max_hsps) break;
      const Segment& sg = segs[b.best];
      Hsp h;
      if (!gapped_hsp(query.data(), query.size(), subj, slen, {sg.q_start, sg.s_start}, b.b0, b.b1,
                      band, cfg, db_letters, t, sid, &h, c))
        continue;
      if (h.evalue <= cfg.evalue) hsps.push_back(std::move(h));
    }
    if (!hsps.empty()) {
      std::stable_sort(hsps.begin(), hsps.end(), hsp_less);
      found.push_back(std::move(hsps));
    }
  }
  std::stable_sort(found.begin(), found.end(),
                   [](const std::vector<Hsp>& a, const std::vector<Hsp>& b) { return hsp_less(a[0], b[0]); });
  if (found.size() > cfg.max_target_seqs) found.resize(cfg.max_target_seqs);
  for (auto& hs : found) {
    if (hs.size() > cfg.max_hsps) hs.resize(cfg.max_hsps);
    for (auto& h : hs) outv.push_back(std::move(h));
  }
  return outv;
}

static double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// src/search.rs:95-146 — query seeds once; per DB chunk and shape: build the
// reference seeds and join (serial); then a thread pool over targets.
std::vector<std::vector<Hsp>> search_run(const Database& db, const std::vector<Target>& targets,
                                         const SearchConfig& cfg, SearchCounters* c) {
  SearchCounters local;
  if (!c) c = &local;
  std::vector<std::vector<Hsp>> results(targets.size());
  if (targets.empty() || db.size() == 0) return results;
  double t0 = now_s();
  auto shapes = shapes_for(cfg.sensitivity, 0);
  Reduction red = reduction_for(cfg.sensitivity);
  double cut = cfg.seed_cut_for() * kLn2 * (double)shapes[0].weight;
  size_t nchunks = cfg.index_chunks_for();
  std::vector<const uint8_t*> qs;
  std::vector<uint32_t> ql;
  for (auto& t : targets) {
    qs.push_back(t.letters.data());
    ql.push_back((uint32_t)t.letters.size());
  }
  std::vector<std::vector<SeedEntry>> qseeds;
  for (auto& sh : shapes) {
    qseeds.push_back(seed_build(qs, ql, sh, red, 0, cut));
    c->query_seeds += qseeds.back().size();
  }
  std::vector<HitMap> hits(targets.size());
  double t1 = now_s();
  for (auto [c0, c1] : chunks(db.size(), nchunks)) {
    std::vector<const uint8_t*> rs;
    std::vector<uint32_t> rl;
    for (size_t i = c0; i < c1; ++i) {
      rs.push_back(db.seq(i));
      rl.push_back((uint32_t)db.seq_len(i));
    }
    for (size_t si = 0; si < shapes.size(); ++si) {
      auto rseeds = seed_build(rs, rl, shapes[si], red, (uint32_t)c0, cut);
      c->ref_seeds += rseeds.size();
      join_into(qseeds[si], rseeds, hits, 4096, &c->hits);
    }
  }
  double t2 = now_s();
  c->t_seed = t1 - t0;
  c->t_join = t2 - t1;
  int nthreads = std::max(cfg.threads, 1);
  std::atomic<size_t> next{0};
  std::vector<SearchCounters> tc(nthreads);
  auto worker = [&](int w) {
    for (;;) {
      size_t ti = next.fetch_add(1);
      if (ti >= targets.size()) break;
      results[ti] = align_target(db, targets[ti], hits[ti], cfg, &tc[w]);
    }
  };
  std::vector<std::thread> pool;
  for (int w = 1; w < nthreads; ++w) pool.emplace_back(worker, w);
  worker(0);
  for (auto& th : pool) th.join();
  for (auto& x : tc) {
    c->xdrops += x.xdrops;
    c->segments += x.segments;
    c->bands += x.bands;
    c->dps += x.dps;
    c->dp_cells += x.dp_cells;
    c->clipped_dps += x.clipped_dps;
    c->band_tie_events += x.band_tie_events;
    c->anchor_tie_events += x.anchor_tie_events;
  }
  c->t_align = now_s() - t2;
  return results;
}

// src/search.rs:357-396
std::vector<std::vector<Hsp>> merge_frames(std::vector<std::vector<Hsp>> results,
                                           const std::vector<Target>& targets, size_t queries,
                                           size_t max_hsps, size_t max_targets) {
  std::vector<std::vector<Hsp>> per(queries);
  for (size_t ti = 0; ti < results.size(); ++ti) {
    size_t q = targets[ti].query;
    if (q < queries)
      for (auto& h : results[ti]) per[q].push_back(std::move(h));
  }
  for (auto& hs : per) {
    std::stable_sort(hs.begin(), hs.end(), hsp_less);
    std::vector<size_t> order;
    std::map<size_t, std::vector<Hsp>> by;
    for (auto& h : hs) {
      auto it = by.find(h.subject);
      if (it != by.end()) {
        if (it->second.size() < max_hsps) it->second.push_back(std::move(h));
      } else if (order.size() < max_targets) {
        order.push_back(h.subject);
        by[h.subject].push_back(std::move(h));
      }
    }
    hs.clear();
    for (size_t sid : order)
      for (auto& h : by[sid]) hs.push_back(std::move(h));
  }
  return per;
}

// src/translate.rs:9-82
static uint8_t nt_base(uint8_t b) {
  switch (b) {
    case 'T': case 't': case 'U': case 'u': return 0;
    case 'C': case 'c': return 1;
    case 'A': case 'a': return 2;
    case 'G': case 'g': return 3;
    default: return 4;
  }
}
static uint8_t nt_complement(uint8_t b) {
  switch (b) {
    case 'A': return 'T'; case 'T': return 'A'; case 'C': return 'G'; case 'G': return 'C';
    case 'U': return 'A'; case 'R': return 'Y'; case 'Y': return 'R'; case 'S': return 'S';
    case 'W': return 'W'; case 'K': return 'M'; case 'M': return 'K'; case 'B': return 'V';
    case 'V': return 'B'; case 'D': return 'H'; case 'H': return 'D'; case 'N': return 'N';
    default: return b;
  }
}
static std::vector<uint8_t> translate_frame(const std::string& nt, size_t offset) {
  static const char codon[] = "FFLLSSSSYY**CC*WLLLLPPPPHHQQRRRRIIIMTTTTNNKKSSRRVVVVAAAADDEEGGGG";
  std::vector<uint8_t> out;
  for (size_t i = offset; i + 3 <= nt.size(); i += 3) {
    uint8_t a = nt_base(nt[i]), b = nt_base(nt[i + 1]), c = nt_base(nt[i + 2]);
    if (a > 3 || b > 3 || c > 3) out.push_back(23);
    else out.push_back(encode_byte((uint8_t)codon[a * 16 + b * 4 + c]));
  }
  return out;
}
std::vector<std::vector<uint8_t>> six_frames(const std::string& nt) {
  std::string rc(nt.rbegin(), nt.rend());
  for (auto& ch : rc) ch = (char)nt_complement((uint8_t)ch);
  std::vector<std::vector<uint8_t>> f;
  for (size_t o = 0; o < 3; ++o) f.push_back(translate_frame(nt, o));
  for (size_t o = 0; o < 3; ++o) f.push_back(translate_frame(rc, o));
  return f;
}
// src/translate.rs:101-121
void mask_short_orfs(std::vector<uint8_t>& frame, size_t min_len) {
  if (min_len <= 1) return;
  size_t begin = 0;
  for (size_t i = 0; i < frame.size(); ++i) {
    if (frame[i] == 24) {
      if (i - begin < min_len)
        for (size_t k = begin; k < i; ++k) frame[k] = 23;
      begin = i + 1;
    }
  }
  if (frame.size() - begin < min_len)
    for (size_t k = begin; k < frame.size(); ++k) frame[k] = 23;
}
// src/translate.rs:124-137
size_t min_orf_len(size_t dna_len, long user) {
  if (user >= 0) return (size_t)user;
  if (dna_len < 30) return 1;
  if (dna_len < 100) return 20;
  return 40;
}

// src/output.rs:235-241 — >= 100 prints floor as an integer, else one decimal
// rounded half-to-even on the exact binary value (glibc %.1f does the same).
std::string fmt_double(double x) {
  char buf[64];
  if (x >= 100.0) std::snprintf(buf, sizeof buf, "%lld", (long long)std::floor(x));
  else std::snprintf(buf, sizeof buf, "%.1f", x);
  return buf;
}

// src/output.rs:244-250 — Rust {:.2e}: mantissa with two decimals, exponent
// without sign padding ("1.00e-3", "3.21e-105", "1.50e10").
std::string fmt_evalue(double e) {
  if (!std::isfinite(e) || e <= 0.0) return "0.0";
  char buf[64];
  std::snprintf(buf, sizeof buf, "%.2e", e);
  std::string s(buf);
  size_t p = s.find('e');
  std::string mant = s.substr(0, p);
  int ex = std::atoi(s.c_str() + p + 1);
  return mant + "e" + std::to_string(ex);
}

// src/hsp.rs:84-90 — qframe column: +1..+3 forward, -1..-3 reverse.
int q_frame(uint8_t frame) { return frame < 3 ? (int)frame + 1 : 2 - (int)frame; }

// src/hsp.rs:97-110 — nucleotide range on the forward strand (0-based,
// half-open) of a frame's amino-acid range.
std::pair<size_t, size_t> nt_range(const Hsp& h, size_t dna_len) {
  const size_t f = h.frame;
  if (f < 3) return {f + 3 * h.q_begin, f + 3 * h.q_end};
  const size_t o = f - 3;
  return {dna_len - (o + 3 * h.q_end), dna_len - (o + 3 * h.q_begin)};
}

// src/hsp.rs:164-192 — gapped query and subject strings of the aligned region.
std::pair<std::string, std::string> aligned_strings(const Hsp& h, const uint8_t* q, const uint8_t* s) {
  std::string qs, ss;
  size_t qi = h.q_begin, si = h.s_begin;
  for (auto& [op, n] : h.ops)
    for (uint32_t t = 0; t < n; ++t) {
      if (op == OP_M) { qs += decode_byte(q[qi++]); ss += decode_byte(s[si++]); }
      else if (op == OP_I) { qs += decode_byte(q[qi++]); ss += '-'; }
      else { qs += '-'; ss += decode_byte(s[si++]); }
    }
  return {qs, ss};
}

// src/hsp.rs:152-159
std::string cigar(const Hsp& h) {
  std::string s;
  for (auto& [op, n] : h.ops) s += std::to_string(n) + (op == OP_M ? 'M' : op == OP_I ? 'I' : 'D');
  return s;
}

// src/hsp.rs:195-241
std::string btop(const Hsp& h, const uint8_t* q, const uint8_t* s) {
  std::string out;
  uint32_t matches = 0;
  size_t qi = h.q_begin, si = h.s_begin;
  auto flush = [&]() {
    if (matches > 0) {
      out += std::to_string(matches);
      matches = 0;
    }
  };
  for (auto& [op, n] : h.ops) {
    for (uint32_t k = 0; k < n; ++k) {
      if (op == OP_M) {
        char a = decode_byte(q[qi]), b = decode_byte(s[si]);
        if (a == b) matches++;
        else { flush(); out += a; out += b; }
        ++qi; ++si;
      } else if (op == OP_I) {
        flush(); out += decode_byte(q[qi]); out += '-'; ++qi;
      } else {
        flush(); out += '-'; out += decode_byte(s[si]); ++si;
      }
    }
  }
  flush();
  return out;
}

}  // namespace orc

// ===========================================================================
// C API for the tests / benchmark (ctypes).
// ===========================================================================
using namespace orc;

namespace {
std::vector<std::string> split_lines(const char* s) {
  std::vector<std::string> out;
  if (!s) return out;
  std::string cur;
  for (const char* p = s; *p; ++p) {
    if (*p == '\n') { out.push_back(cur); cur.clear(); }
    else cur += *p;
  }
  if (!cur.empty()) out.push_back(cur);
  return out;
}

struct Result {
  std::vector<Target> targets;
  std::vector<std::string> raw_nt;  // blastx reads
  std::vector<std::vector<Hsp>> per_query;
  std::vector<const Hsp*> flat;
  std::vector<int> flat_target;  // index into targets of each flat hsp
  SearchCounters counters;
  bool translated = false;
  const Database* db = nullptr;
  Database db_copy;
  size_t nqueries = 0;
};

std::string aa_string(const std::vector<uint8_t>& v, size_t b, size_t e) {
  std::string s;
  for (size_t i = b; i < e && i < v.size(); ++i) s += decode_byte(v[i]);
  return s;
}
}  // namespace

extern "C" {

int orc_encode(const char* raw, uint8_t* out, int cap) {
  auto v = encode(raw);
  int n = (int)std::min<size_t>(v.size(), (size_t)cap);
  std::memcpy(out, v.data(), n);
  return (int)v.size();
}

int orc_shape_key(const char* code, const uint8_t* seq, int len, int pos, const char* reduction,
                  uint64_t* key) {
  Shape s = shape_parse(code);
  Reduction r = reduction_parse(reduction);
  return s.packed(seq, (size_t)len, (size_t)pos, r, key) ? 1 : 0;
}

int orc_seed_is_complex(const uint8_t* seq, const char* code, const char* reduction, double cut) {
  return seed_is_complex(seq, shape_parse(code), reduction_parse(reduction), cut) ? 1 : 0;
}

int orc_shapes_for(const char* sens, int limit, char* buf, int cap) {
  std::string s;
  for (auto& sh : shapes_for(sens, (size_t)limit)) s += sh.code + "\n";
  std::snprintf(buf, cap, "%s", s.c_str());
  return (int)s.size();
}

int orc_reduction_map(const char* sens, uint8_t* map25) {
  Reduction r = reduction_for(sens);
  std::memcpy(map25, r.map, 25);
  return (int)r.size;
}

double orc_seed_cut(const char* sens, double override_cut) {
  SearchConfig c;
  c.sensitivity = sens;
  c.seed_cut = override_cut;
  return c.seed_cut_for();
}

long orc_index_chunks(const char* sens, long override_n) {
  SearchConfig c;
  c.sensitivity = sens;
  c.index_chunks = override_n;
  return (long)c.index_chunks_for();
}

// Seeds of seqs[offs[i]..offs[i+1]] sorted by (key, seq, pos).
int64_t orc_seed_build(const uint8_t* seqs, const uint64_t* offs, uint32_t nseq, const char* code,
                       const char* reduction, uint32_t base, double cut, uint64_t* keys,
                       uint32_t* seq, uint32_t* pos, int64_t cap) {
  std::vector<const uint8_t*> p;
  std::vector<uint32_t> l;
  for (uint32_t i = 0; i < nseq; ++i) {
    p.push_back(seqs + offs[i]);
    l.push_back((uint32_t)(offs[i + 1] - offs[i]));
  }
  auto v = seed_build(p, l, shape_parse(code), reduction_parse(reduction), base, cut);
  for (int64_t i = 0; i < (int64_t)v.size() && i < cap; ++i) {
    keys[i] = v[i].key;
    seq[i] = v[i].seq;
    pos[i] = v[i].pos;
  }
  return (int64_t)v.size();
}

// Join of two sorted seed arrays; hits flattened as (query, qpos, subject, spos)
// sorted by (query, subject, diag, spos).
int64_t orc_seed_join(const uint64_t* qk, const uint32_t* qs, const uint32_t* qp, int64_t nq,
                      const uint64_t* rk, const uint32_t* rs, const uint32_t* rp, int64_t nr,
                      uint32_t ntargets, int64_t max_expand, uint32_t* out4, int64_t cap) {
  std::vector<SeedEntry> q(nq), r(nr);
  for (int64_t i = 0; i < nq; ++i) q[i] = {qk[i], qs[i], qp[i]};
  for (int64_t i = 0; i < nr; ++i) r[i] = {rk[i], rs[i], rp[i]};
  std::vector<HitMap> hits(ntargets);
  uint64_t n = 0;
  join_into(q, r, hits, (size_t)max_expand, &n);
  std::vector<SeedHit> flat;
  for (auto& hm : hits)
    for (auto& kv : hm.m)
      for (auto& h : kv.second) flat.push_back(h);
  std::sort(flat.begin(), flat.end(), [](const SeedHit& a, const SeedHit& b) {
    if (a.query != b.query) return a.query < b.query;
    if (a.subject != b.subject) return a.subject < b.subject;
    long da = (long)a.qpos - (long)a.spos, dbv = (long)b.qpos - (long)b.spos;
    if (da != dbv) return da < dbv;
    return a.spos < b.spos;
  });
  for (int64_t i = 0; i < (int64_t)flat.size() && i < cap; ++i) {
    out4[4 * i] = flat[i].query;
    out4[4 * i + 1] = flat[i].qpos;
    out4[4 * i + 2] = flat[i].subject;
    out4[4 * i + 3] = flat[i].spos;
  }
  return (int64_t)flat.size();
}

int orc_chunks(int64_t n, int64_t parts, int64_t* out) {
  auto v = chunks((size_t)n, (size_t)parts);
  for (size_t i = 0; i < v.size(); ++i) {
    out[2 * i] = (int64_t)v[i].first;
    out[2 * i + 1] = (int64_t)v[i].second;
  }
  return (int)v.size();
}

static Matrix blosum62_default() {
  Matrix m;
  matrix_blosum62(11, 1, &m);
  return m;
}

int orc_matrix_score(int a, int b) { return blosum62_default().score((uint8_t)a, (uint8_t)b); }

int orc_matrix_ok(int go, int ge) {
  Matrix m;
  return matrix_blosum62(go, ge, &m) ? 1 : 0;
}

void orc_xdrop(const uint8_t* q, int64_t qlen, const uint8_t* s, int64_t slen, int64_t qa,
               int64_t sa, int xd, int64_t* out4) {
  Segment g = xdrop(q, (size_t)qlen, s, (size_t)slen, (size_t)qa, (size_t)sa, blosum62_default(), xd);
  out4[0] = (int64_t)g.q_start;
  out4[1] = (int64_t)g.s_start;
  out4[2] = (int64_t)g.len;
  out4[3] = g.score;
}

// out: score, q_begin, q_end, s_begin, s_end, identities, positives,
// mismatches, gap_openings, gaps, length, cells. ops: (op | count << 2).
int orc_extend_span(const uint8_t* q, int64_t n, const uint8_t* s, int64_t m, int64_t d0, int64_t d1,
                    int go, int ge, int64_t* out12, uint32_t* ops, int ops_cap, int* nops) {
  Matrix mat;
  if (!matrix_blosum62(go, ge, &mat)) return -1;
  Aligned a;
  bool ok = extend_span(q, (size_t)n, s, (size_t)m, (long)d0, (long)d1, mat, &a);
  std::memset(out12, 0, 12 * sizeof(int64_t));
  out12[11] = (int64_t)a.cells;
  if (!ok) {
    *nops = 0;
    return 0;
  }
  out12[0] = a.score;
  out12[1] = (int64_t)a.q_begin;
  out12[2] = (int64_t)a.q_end;
  out12[3] = (int64_t)a.s_begin;
  out12[4] = (int64_t)a.s_end;
  out12[5] = a.stats.identities;
  out12[6] = a.stats.positives;
  out12[7] = a.stats.mismatches;
  out12[8] = a.stats.gap_openings;
  out12[9] = a.stats.gaps;
  out12[10] = a.stats.length;
  *nops = (int)a.ops.size();
  for (int i = 0; i < (int)a.ops.size() && i < ops_cap; ++i)
    ops[i] = (uint32_t)a.ops[i].first | (a.ops[i].second << 2);
  return 1;
}

int orc_band_for(int64_t len) { return (int)band_for((size_t)len); }

double orc_evalue(int raw, int64_t qlen, int64_t slen, uint64_t db) {
  return blosum62_default().evalue(raw, (size_t)qlen, (size_t)slen, db);
}
double orc_bit_score(int raw) { return blosum62_default().bit_score(raw); }
int orc_raw_score(double bits) { return blosum62_default().raw_score(bits); }
double orc_log_erfc(double x) { return log_erfc(x); }

int orc_fmt_double(double x, char* buf, int cap) {
  return std::snprintf(buf, cap, "%s", fmt_double(x).c_str());
}
int orc_fmt_evalue(double x, char* buf, int cap) {
  return std::snprintf(buf, cap, "%s", fmt_evalue(x).c_str());
}

// Letters written up to cap; returns the count of all six frames' letters
// (more than cap means the output was cut).
int orc_six_frames(const char* nt, uint8_t* out, int64_t cap, int64_t* lens6, long min_orf) {
  std::string s(nt);
  auto f = six_frames(s);
  size_t mo = min_orf_len(s.size(), min_orf);
  int64_t o = 0;
  for (int k = 0; k < 6; ++k) {
    mask_short_orfs(f[k], mo);
    lens6[k] = (int64_t)f[k].size();
    for (auto c : f[k]) {
      if (o < cap) out[o] = c;
      ++o;
    }
  }
  return (int)o;
}

int orc_q_frame(int frame) { return q_frame((uint8_t)frame); }

void orc_nt_range(int frame, int64_t q_begin, int64_t q_end, int64_t dna_len, int64_t* out2) {
  Hsp h;
  h.frame = (uint8_t)frame;
  h.q_begin = (size_t)q_begin;
  h.q_end = (size_t)q_end;
  auto r = nt_range(h, (size_t)dna_len);
  out2[0] = (int64_t)r.first;
  out2[1] = (int64_t)r.second;
}

// Text of one alignment given by its ops (op | count << 2) starting at
// (q_begin, s_begin) of the encoded query q and subject s: which = 0 CIGAR,
// 1 BTOP, 2 qseq, 3 sseq. Returns the byte length; writes up to cap.
int64_t orc_alignment_text(int which, const uint32_t* ops, int64_t nops, const uint8_t* q,
                           int64_t q_begin, const uint8_t* s, int64_t s_begin, char* buf, int64_t cap) {
  Hsp h;
  h.q_begin = (size_t)q_begin;
  h.s_begin = (size_t)s_begin;
  for (int64_t i = 0; i < nops; ++i) h.ops.push_back({(Op)(ops[i] & 3), ops[i] >> 2});
  std::string t;
  if (which == 0) t = cigar(h);
  else if (which == 1) t = btop(h, q, s);
  else {
    auto al = aligned_strings(h, q, s);
    t = which == 2 ? al.first : al.second;
  }
  if (buf && cap > 0) std::memcpy(buf, t.data(), (size_t)std::min<int64_t>(cap, (int64_t)t.size()));
  return (int64_t)t.size();
}

// Full search. Queries: encoded residues (blastp) or nucleotide ASCII (blastx).
// header_letters: the database's header letter count (0 = its letters).
void* orc_search(const uint8_t* dbl, const uint64_t* db_off, uint32_t ndb, const uint8_t* ql,
                 const uint64_t* q_off, uint32_t nq, const char* sens, int threads, double evalue,
                 int max_target_seqs, int max_hsps, long index_chunks, double seed_cut,
                 int translated, long min_orf, uint64_t header_letters) {
  Result* r = new Result();
  r->db_copy.letters.assign(dbl, dbl + db_off[ndb]);
  r->db_copy.offsets.assign(db_off, db_off + ndb + 1);
  r->db_copy.header_letters = header_letters;
  r->db = &r->db_copy;
  r->translated = translated != 0;
  r->nqueries = nq;
  r->raw_nt.resize(nq);
  SearchConfig cfg;
  matrix_blosum62(11, 1, &cfg.matrix);
  cfg.evalue = evalue;
  cfg.max_target_seqs = (size_t)max_target_seqs;
  cfg.max_hsps = (size_t)max_hsps;
  cfg.sensitivity = sens;
  cfg.threads = threads;
  cfg.index_chunks = index_chunks;
  cfg.seed_cut = seed_cut;
  for (uint32_t i = 0; i < nq; ++i) {
    if (translated) {
      std::string nt((const char*)ql + q_off[i], (const char*)ql + q_off[i + 1]);
      r->raw_nt[i] = nt;
      size_t mo = min_orf_len(nt.size(), min_orf);
      auto frames = six_frames(nt);
      for (int k = 0; k < 6; ++k) {
        mask_short_orfs(frames[k], mo);
        size_t good = 0;
        for (auto c : frames[k]) good += c < 23;
        if (good == 0) continue;
        r->targets.push_back({i, (uint8_t)k, frames[k]});
      }
    } else {
      std::vector<uint8_t> v(ql + q_off[i], ql + q_off[i + 1]);
      if (v.empty()) continue;
      r->targets.push_back({i, 0, v});
    }
  }
  auto res = search_run(*r->db, r->targets, cfg, &r->counters);
  r->per_query = merge_frames(std::move(res), r->targets, nq, cfg.max_hsps, cfg.max_target_seqs);
  for (auto& hs : r->per_query)
    for (auto& h : hs) {
      r->flat.push_back(&h);
      int ti = -1;
      for (size_t t = 0; t < r->targets.size(); ++t)
        if (r->targets[t].query == h.query && r->targets[t].frame == h.frame) ti = (int)t;
      r->flat_target.push_back(ti);
    }
  return r;
}

int64_t orc_result_count(void* h) { return (int64_t)((Result*)h)->flat.size(); }

// 16 int64 per HSP: query, subject, frame, q_begin, q_end, s_begin, s_end,
// score, identities, positives, mismatches, gap_openings, gaps, length,
// n_ops, 0. Doubles: bits, evalue.
void orc_result_hsps(void* h, int64_t* out, double* outd) {
  Result* r = (Result*)h;
  for (size_t i = 0; i < r->flat.size(); ++i) {
    const Hsp& x = *r->flat[i];
    int64_t* o = out + 16 * i;
    o[0] = (int64_t)x.query; o[1] = (int64_t)x.subject; o[2] = x.frame;
    o[3] = (int64_t)x.q_begin; o[4] = (int64_t)x.q_end; o[5] = (int64_t)x.s_begin; o[6] = (int64_t)x.s_end;
    o[7] = x.score; o[8] = x.stats.identities; o[9] = x.stats.positives; o[10] = x.stats.mismatches;
    o[11] = x.stats.gap_openings; o[12] = x.stats.gaps; o[13] = x.stats.length;
    o[14] = (int64_t)x.ops.size(); o[15] = 0;
    outd[2 * i] = x.bits;
    outd[2 * i + 1] = x.evalue;
  }
}

// Counters: query_seeds, ref_seeds, hits, xdrops, segments, bands, dps,
// dp_cells, clipped_dps, band_tie_events, anchor_tie_events, t_seed, t_join, t_align.
void orc_result_counters(void* h, double* out) {
  const SearchCounters& c = ((Result*)h)->counters;
  double v[14] = {(double)c.query_seeds, (double)c.ref_seeds, (double)c.hits, (double)c.xdrops,
                  (double)c.segments, (double)c.bands, (double)c.dps, (double)c.dp_cells,
                  (double)c.clipped_dps, (double)c.band_tie_events, (double)c.anchor_tie_events,
                  c.t_seed, c.t_join, c.t_align};
  std::memcpy(out, v, sizeof v);
}

// outfmt-6 text for the given field names (space separated, empty = defaults)
// following src/output.rs:150-223. Ids and titles (the FASTA headers of the
// queries, the database headers) are newline-separated lists; titles may be
// NULL when no title field is asked for. Returns the byte length; writes up
// to cap.
int64_t orc_result_outfmt6(void* h, const char* qids, const char* sids, const char* qtitles,
                           const char* stitles, const char* fields, char* buf, int64_t cap) {
  Result* r = (Result*)h;
  auto qid = split_lines(qids), sid = split_lines(sids);
  auto qtl = split_lines(qtitles), stl = split_lines(stitles);
  std::vector<std::string> f;
  {
    std::string cur;
    for (const char* p = fields; p && *p; ++p) {
      if (*p == ' ') { if (!cur.empty()) f.push_back(cur); cur.clear(); }
      else cur += *p;
    }
    if (!cur.empty()) f.push_back(cur);
    if (f.empty())
      f = {"qseqid", "sseqid", "pident", "length", "mismatch", "gapopen",
           "qstart", "qend", "sstart", "send", "evalue", "bitscore"};
  }
  std::string out;
  for (size_t i = 0; i < r->flat.size(); ++i) {
    const Hsp& x = *r->flat[i];
    int ti = r->flat_target[i];
    if (ti < 0) continue;
    const auto& letters = r->targets[ti].letters;
    const uint8_t* subj = r->db->seq(x.subject);
    size_t slen = r->db->seq_len(x.subject);
    size_t dna_len = r->translated ? r->raw_nt[x.query].size() : letters.size();
    size_t qb = x.q_begin, qe = x.q_end;
    if (r->translated) std::tie(qb, qe) = nt_range(x, dna_len);
    size_t query_len = r->translated ? dna_len : letters.size();
    std::string row;
    for (size_t k = 0; k < f.size(); ++k) {
      const std::string& n = f[k];
      std::string v;
      if (n == "qseqid") v = qid[x.query];
      else if (n == "qtitle") v = x.query < qtl.size() ? qtl[x.query] : "?";
      else if (n == "qlen") v = std::to_string(query_len);
      else if (n == "sseqid") v = sid[x.subject];
      else if (n == "stitle") v = x.subject < stl.size() ? stl[x.subject] : "?";
      else if (n == "slen") v = std::to_string(slen);
      else if (n == "qstart") v = std::to_string(qb + 1);
      else if (n == "qend") v = std::to_string(qe);
      else if (n == "sstart") v = std::to_string(x.s_begin + 1);
      else if (n == "send") v = std::to_string(x.s_end);
      else if (n == "evalue") v = fmt_evalue(x.evalue);
      else if (n == "bitscore") v = fmt_double(x.bits);
      else if (n == "score") v = std::to_string(x.score);
      else if (n == "length") v = std::to_string(x.stats.length);
      else if (n == "pident") v = fmt_double(x.stats.length ? x.stats.identities * 100.0 / x.stats.length : 0.0);
      else if (n == "nident") v = std::to_string(x.stats.identities);
      else if (n == "mismatch") v = std::to_string(x.stats.mismatches);
      else if (n == "positive") v = std::to_string(x.stats.positives);
      else if (n == "gapopen") v = std::to_string(x.stats.gap_openings);
      else if (n == "gaps") v = std::to_string(x.stats.gaps);
      else if (n == "ppos") v = fmt_double(x.stats.length ? x.stats.positives * 100.0 / x.stats.length : 0.0);
      else if (n == "qframe") v = std::to_string(q_frame(x.frame));
      else if (n == "qcovhsp") v = query_len == 0 ? "0.0" : fmt_double((double)(qe - qb) * 100.0 / (double)query_len);
      else if (n == "scovhsp") v = slen == 0 ? "0.0" : fmt_double((double)(x.s_end - x.s_begin) * 100.0 / (double)slen);
      else if (n == "cigar") v = cigar(x);
      else if (n == "btop") v = btop(x, letters.data(), subj);
      else if (n == "full_sseq") v = aa_string(std::vector<uint8_t>(subj, subj + slen), 0, slen);
      else if (n == "qseq_translated") {
        size_t e = std::min(x.q_end, letters.size()), b = std::min(x.q_begin, e);
        v = aa_string(letters, b, e);
      } else if (n == "qstrand") v = x.frame >= 3 ? "-" : "+";
      else if (n == "qseq" || n == "sseq") {
        auto al = aligned_strings(x, letters.data(), subj);
        v = n == "qseq" ? al.first : al.second;
      } else v = "?";
      if (k) row += '\t';
      row += v;
    }
    out += row;
    out += '\n';
  }
  if (buf && cap > 0) std::memcpy(buf, out.data(), (size_t)std::min<int64_t>(cap, (int64_t)out.size()));
  return (int64_t)out.size();
}

void orc_result_free(void* h) { delete (Result*)h; }

}  // extern "C"
