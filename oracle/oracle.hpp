// Oracle: a CPU restatement of diamond-rust's search core.
//
// TEST INFRASTRUCTURE ONLY. Nothing in the product links, includes or calls
// this code; only tests/, the smoke check and bench.py's CPU-baseline leg do,
// and only to compare results with. Every function cites the reference file:line whose
// behaviour it restates (paths relative to the reference repository).
//
// Parity unpinned (nothing in the reference's sources or tests fixes them):
//  - the order Rust's sort_unstable gives tied elements in slices longer than
//    20 (src/search.rs:206,214,230). The oracle sorts stably and counts such
//    events (band_tie_events, anchor_tie_events in the search counters);
//  - compile-time folding of 0.5f64.ln() and INV_SQRT_2PI.ln()
//    (src/alp.rs:134-135), computed here with glibc log at run time.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace orc {

// ---- alphabet (src/alphabet.rs:15-63) ------------------------------------
uint8_t encode_byte(uint8_t c);
std::vector<uint8_t> encode(const std::string& raw);
char decode_byte(uint8_t l);

// ---- reduction (src/reduction.rs:19-79) ----------------------------------
struct Reduction {
  uint8_t map[25];
  uint64_t size;
  uint8_t bucket(uint8_t l) const { return map[l < 24 ? l : 24]; }
};
Reduction reduction_parse(const std::string& def);
Reduction reduction_for(const std::string& sensitivity);

// ---- shape (src/shape.rs:23-220) -----------------------------------------
struct Shape {
  std::string code;
  uint32_t weight = 0, span = 0;
  std::vector<uint32_t> positions;
  bool packed(const uint8_t* seq, size_t len, size_t pos, const Reduction& r, uint64_t* key) const;
};
Shape shape_parse(const std::string& code);
std::vector<Shape> shapes_for(const std::string& sensitivity, size_t limit);

// ---- seeds (src/seed.rs:14-164) ------------------------------------------
struct SeedEntry { uint64_t key; uint32_t seq, pos; };
struct SeedHit { uint32_t query, qpos, subject, spos; };
double ln_factorial(int n);
bool seed_is_complex(const uint8_t* seq, const Shape& s, const Reduction& r, double cut);
std::vector<SeedEntry> seed_build(const std::vector<const uint8_t*>& seqs,
                                  const std::vector<uint32_t>& lens, const Shape& s,
                                  const Reduction& r, uint32_t base, double cut);
// Hits are appended to out[q.seq] keyed by subject; `out` holds one vector of
// (subject, hit) per target in emission order.
struct HitMap;
std::vector<std::pair<size_t, size_t>> chunks(size_t n, size_t parts);

// ---- scoring (src/matrix.rs:42-200, src/alp.rs) ---------------------------
struct Matrix {
  int8_t s[25][25];
  int gap_open = 11, gap_extend = 1;
  double lambda = 0, k = 0, ln_k = 0;
  double alpha = 0, alpha_v = 0, sigma = 0, u_alpha = 0, u_alpha_v = 0;
  int score(uint8_t a, uint8_t b) const { return s[a < 24 ? a : 24][b < 24 ? b : 24]; }
  int raw_score(double bits) const;
  double bit_score(int raw) const;
  double evalue(int raw, size_t qlen, size_t slen, uint64_t db_letters) const;
};
bool matrix_blosum62(int gap_open, int gap_extend, Matrix* out);
double log_erfc(double x);

// ---- ungapped (src/ungapped.rs:10-88) ------------------------------------
struct Segment {
  size_t q_start, s_start, len;
  int32_t score;
  long diag() const { return (long)q_start - (long)s_start; }
  size_t s_end() const { return s_start + len; }
};
Segment xdrop(const uint8_t* q, size_t qlen, const uint8_t* s, size_t slen, size_t qa, size_t sa,
              const Matrix& m, int xd);

// ---- gapped (src/gapped.rs:12-280, src/hsp.rs:8-57) -----------------------
enum Op : uint8_t { OP_M = 0, OP_I = 1, OP_D = 2 };
struct Stats { uint32_t identities = 0, positives = 0, mismatches = 0, gap_openings = 0, gaps = 0, length = 0; };
struct Aligned {
  int32_t score = 0;
  size_t q_begin = 0, q_end = 0, s_begin = 0, s_end = 0;
  Stats stats;
  std::vector<std::pair<Op, uint32_t>> ops;
  uint64_t cells = 0;  // DP cells visited (one iteration of the inner loop)
};
size_t band_for(size_t len);
bool extend_span(const uint8_t* q, size_t n, const uint8_t* s, size_t m, long d0, long d1,
                 const Matrix& mat, Aligned* out);

// ---- HSP + search (src/hsp.rs, src/search.rs) -----------------------------
struct Hsp {
  size_t query = 0, subject = 0;
  uint8_t frame = 0;
  size_t q_begin = 0, q_end = 0, s_begin = 0, s_end = 0;
  int32_t score = 0;
  double bits = 0, evalue = 0;
  Stats stats;
  std::vector<std::pair<Op, uint32_t>> ops;
};

struct Target {
  size_t query;
  uint8_t frame;
  std::vector<uint8_t> letters;
};

struct SearchConfig {
  Matrix matrix;
  double evalue = 0.001;
  size_t max_target_seqs = 25, max_hsps = 1;
  std::string sensitivity = "default";
  int threads = 1;
  long index_chunks = -1;  // <0: preset
  double seed_cut = -1;    // <0: preset
  double seed_cut_for() const;
  size_t index_chunks_for() const;
  int ungapped_xdrop() const { return matrix.raw_score(12.3); }
};

struct SearchCounters {
  uint64_t query_seeds = 0, ref_seeds = 0, hits = 0, xdrops = 0, segments = 0, bands = 0;
  uint64_t dps = 0, dp_cells = 0, clipped_dps = 0;
  // Events where the reference's unstable sorts could order tied elements
  // differently from this restatement (only slices longer than 20 qualify).
  uint64_t band_tie_events = 0, anchor_tie_events = 0;
  double t_seed = 0, t_join = 0, t_align = 0;
};

struct Database {
  std::vector<uint8_t> letters;
  std::vector<uint64_t> offsets;  // n+1
  size_t size() const { return offsets.empty() ? 0 : offsets.size() - 1; }
  const uint8_t* seq(size_t i) const { return letters.data() + offsets[i]; }
  size_t seq_len(size_t i) const { return offsets[i + 1] - offsets[i]; }
  // src/db.rs:199,236-238 — the letter count of the file header (offset 24),
  // which a file may state differently from its offset table; 0 = letters.
  uint64_t header_letters = 0;
  uint64_t total_letters() const { return header_letters ? header_letters : letters.size(); }
};

std::vector<std::vector<Hsp>> search_run(const Database& db, const std::vector<Target>& targets,
                                         const SearchConfig& cfg, SearchCounters* counters);
std::vector<std::vector<Hsp>> merge_frames(std::vector<std::vector<Hsp>> results,
                                           const std::vector<Target>& targets, size_t queries,
                                           size_t max_hsps, size_t max_targets);

// ---- translation (src/translate.rs:9-137) --------------------------------
std::vector<std::vector<uint8_t>> six_frames(const std::string& nt);
void mask_short_orfs(std::vector<uint8_t>& frame, size_t min_len);
size_t min_orf_len(size_t dna_len, long user);

// ---- output (src/output.rs:150-250, src/hsp.rs:113-241) -------------------
std::string fmt_double(double x);
std::string fmt_evalue(double e);
int q_frame(uint8_t frame);
std::pair<size_t, size_t> nt_range(const Hsp& h, size_t dna_len);
std::string cigar(const Hsp& h);
std::string btop(const Hsp& h, const uint8_t* q, const uint8_t* s);
std::pair<std::string, std::string> aligned_strings(const Hsp& h, const uint8_t* q, const uint8_t* s);

}  // namespace orc
