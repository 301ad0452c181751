// Scoring, reduction and seed-shape tables of the search core.
//
// Values equal the reference's: BLOSUM62 in NCBI letter order
// (src/matrix_data.rs:265-384) with its gap-cost rows (src/matrix_data.rs:370-382),
// reductions (src/reduction.rs:48-79), shape presets (src/shape.rs:78-220),
// seed-cut / index-chunk presets (src/search.rs:55-79) and the band table
// (src/gapped.rs:27-32).
#pragma once
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

namespace dmr {

constexpr int kAlphabet = 25;
constexpr int kTrueAA = 20;
constexpr int kMaxShapeWeight = 16;
constexpr int kMaxShapeSpan = 32;

// NCBI order: A R N D C Q E G H I L K M F P S T W Y V B J Z X *
inline const int8_t (&blosum62())[25][25] {
  static const int8_t t[25][25] = {
      {4, -1, -2, -2, 0, -1, -1, 0, -2, -1, -1, -1, -1, -2, -1, 1, 0, -3, -2, 0, -2, -1, -1, -1, -4},
      {-1, 5, 0, -2, -3, 1, 0, -2, 0, -3, -2, 2, -1, -3, -2, -1, -1, -3, -2, -3, -1, -2, 0, -1, -4},
      {-2, 0, 6, 1, -3, 0, 0, 0, 1, -3, -3, 0, -2, -3, -2, 1, 0, -4, -2, -3, 4, -3, 0, -1, -4},
      {-2, -2, 1, 6, -3, 0, 2, -1, -1, -3, -4, -1, -3, -3, -1, 0, -1, -4, -3, -3, 4, -3, 1, -1, -4},
      {0, -3, -3, -3, 9, -3, -4, -3, -3, -1, -1, -3, -1, -2, -3, -1, -1, -2, -2, -1, -3, -1, -3, -1, -4},
      {-1, 1, 0, 0, -3, 5, 2, -2, 0, -3, -2, 1, 0, -3, -1, 0, -1, -2, -1, -2, 0, -2, 4, -1, -4},
      {-1, 0, 0, 2, -4, 2, 5, -2, 0, -3, -3, 1, -2, -3, -1, 0, -1, -3, -2, -2, 1, -3, 4, -1, -4},
      {0, -2, 0, -1, -3, -2, -2, 6, -2, -4, -4, -2, -3, -3, -2, 0, -2, -2, -3, -3, -1, -4, -2, -1, -4},
      {-2, 0, 1, -1, -3, 0, 0, -2, 8, -3, -3, -1, -2, -1, -2, -1, -2, -2, 2, -3, 0, -3, 0, -1, -4},
      {-1, -3, -3, -3, -1, -3, -3, -4, -3, 4, 2, -3, 1, 0, -3, -2, -1, -3, -1, 3, -3, 3, -3, -1, -4},
      {-1, -2, -3, -4, -1, -2, -3, -4, -3, 2, 4, -2, 2, 0, -3, -2, -1, -2, -1, 1, -4, 3, -3, -1, -4},
      {-1, 2, 0, -1, -3, 1, 1, -2, -1, -3, -2, 5, -1, -3, -1, 0, -1, -3, -2, -2, 0, -3, 1, -1, -4},
      {-1, -1, -2, -3, -1, 0, -2, -3, -2, 1, 2, -1, 5, 0, -2, -1, -1, -1, -1, 1, -3, 2, -1, -1, -4},
      {-2, -3, -3, -3, -2, -3, -3, -3, -1, 0, 0, -3, 0, 6, -4, -2, -2, 1, 3, -1, -3, 0, -3, -1, -4},
      {-1, -2, -2, -1, -3, -1, -1, -2, -2, -3, -3, -1, -2, -4, 7, -1, -1, -4, -3, -2, -2, -3, -1, -1, -4},
      {1, -1, 1, 0, -1, 0, 0, 0, -1, -2, -2, 0, -1, -2, -1, 4, 1, -3, -2, -2, 0, -2, 0, -1, -4},
      {0, -1, 0, -1, -1, -1, -1, -2, -2, -1, -1, -1, -1, -2, -1, 1, 5, -2, -2, 0, -1, -1, -1, -1, -4},
      {-3, -3, -4, -4, -2, -2, -3, -2, -2, -3, -2, -3, -1, 1, -4, -3, -2, 11, 2, -3, -4, -2, -2, -1, -4},
      {-2, -2, -2, -3, -2, -1, -2, -3, 2, -1, -1, -2, -1, 3, -3, -2, -2, 2, 7, -1, -3, -1, -2, -1, -4},
      {0, -3, -3, -3, -1, -2, -2, -3, -3, 3, 1, -2, 1, -1, -2, -2, 0, -3, -1, 4, -3, 2, -2, -1, -4},
      {-2, -1, 4, 4, -3, 0, 1, -1, 0, -3, -4, 0, -3, -3, -2, 0, -1, -4, -3, -3, 4, -3, 0, -1, -4},
      {-1, -2, -3, -3, -1, -2, -3, -4, -3, 3, 3, -3, 2, 0, -3, -2, -1, -2, -1, 2, -3, 3, -3, -1, -4},
      {-1, 0, 0, 1, -3, 4, 4, -2, 0, -3, -3, 1, -1, -3, -1, 0, -1, -2, -2, -2, 0, -3, 4, -1, -4},
      {-1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -4},
      {-4, -4, -4, -4, -4, -4, -4, -4, -4, -4, -4, -4, -4, -4, -4, -4, -4, -4, -4, -4, -4, -4, -4, -4, 1},
  };
  return t;
}

struct GapRow {
  int gap_open, gap_extend;
  double lambda, k, alpha, alpha_v, sigma;
};

// BLOSUM62 (open, extend, lambda, K, alpha, alpha_v, sigma) and the
// ungapped (alpha, alpha_v) of the matrix.
inline bool blosum62_gap_row(int go, int ge, GapRow* out) {
  static const GapRow rows[] = {
      {11, 2, 0.297, 0.082, 1.1, 12.6738, 12.7576}, {10, 2, 0.291, 0.075, 1.3, 16.474, 16.6026},
      {9, 2, 0.279, 0.058, 1.5, 22.7519, 22.95},     {8, 2, 0.264, 0.045, 1.8, 35.4838, 35.8213},
      {7, 2, 0.239, 0.027, 2.5, 61.2383, 61.886},    {6, 2, 0.201, 0.012, 3.3, 140.417, 141.882},
      {13, 1, 0.292, 0.071, 1.2, 19.5063, 19.8931}, {12, 1, 0.283, 0.059, 1.5, 27.8562, 28.4699},
      {11, 1, 0.267, 0.041, 1.9, 42.6028, 43.6362}, {10, 1, 0.243, 0.024, 2.5, 83.1787, 85.0656},
      {9, 1, 0.206, 0.01, 4.0, 210.333, 214.842},
  };
  for (const auto& r : rows)
    if (r.gap_open == go && r.gap_extend == ge) {
      *out = r;
      return true;
    }
  return false;
}
constexpr double kB62UngappedAlpha = 0.7916;
constexpr double kB62UngappedAlphaV = 4.96466;

// Reduced alphabets as DIAMOND-style group strings.
inline const char* reduction_def(const std::string& sensitivity) {
  return sensitivity == "very-sensitive" ? "A G P CH ST NDE RQK FWY ILMV"
                                         : "A KR EDNQ C G H ILVM FYW P ST";
}

inline const std::vector<std::string>& shape_codes(const std::string& sens) {
  static const std::vector<std::string> fast = {"1101110101101111"};
  static const std::vector<std::string> def = {"111101110111", "111011010010111"};
  static const std::vector<std::string> mid = {"11110110111",    "1101100111101",    "1110010101111",
                                               "11010101100111", "11101110001011", "1110100100010111",
                                               "1101000011010111", "1110011000011011"};
  static const std::vector<std::string> sensitive = {
      "1011110111",       "110100100010111", "11001011111",      "101110001111",
      "11011101100001",   "1111010010101",   "111001001001011",  "10101001101011",
      "111101010011",     "1111000010000111", "1100011011011",   "1101010000011011",
      "1110001010101001", "110011000110011", "11011010001101",   "1101001100010011"};
  static const std::vector<std::string> very = {
      "11110111",          "11100100111",         "110010101011",        "11010001001011",
      "10101100001101",    "110100100100011",     "1010010100010011",    "1100101000001011",
      "11100000100010101", "11000100010010011",   "11010000001000111",   "110001001000010011",
      "1010001000100001011", "1100010100000010011", "1100100000101000011", "1101000010000001011"};
  static const std::vector<std::string> ultra = {
      "1111111",          "11101111",         "110011111",        "110110111",        "111111001",
      "1010111011",       "1011110101",       "1111000111",       "10011110011",      "10101101101",
      "10111010101",      "11001010111",      "11001100111",      "11010101101",      "11110001011",
      "100111010011",     "101100110101",     "101110000111",     "110100101011",     "110110001101",
      "111000110011",     "1010001011011",    "1010101000111",    "1010110100011",    "1100100110011",
      "1100101001011",    "1101001100101",    "1101010101001",    "1110001010101",    "1110010010011",
      "10100001101101",   "11000100010111",   "11010000100111",   "11010100110001",   "11101000011001",
      "11110000001101",   "11110100000011",   "101001000001111",  "110000100101011",  "110010010000111",
      "110101100001001",  "110110000010011",  "111001000100011",  "111100000100101",  "1000110010010101",
      "1001000100101101", "1001000110011001", "1010001001001011", "1010001010010011", "1010010001010101",
      "1010010100010011", "1010010101001001", "1010100000101011", "1010100011000101", "1011000010001011",
      "1100010000111001", "1100010010001011", "1100100001001011", "1100100100100011", "1100110000001101",
      "1101000100010011", "1101000110000101", "1110000001010011", "1110100000010101"};
  if (sens == "faster" || sens == "fast") return fast;
  if (sens == "mid-sensitive") return mid;
  if (sens == "sensitive" || sens == "more-sensitive") return sensitive;
  if (sens == "very-sensitive") return very;
  if (sens == "ultra-sensitive") return ultra;
  return def;
}

inline double preset_seed_cut(const std::string& s) {
  if (s == "faster" || s == "fast") return 0.9;
  if (s == "mid-sensitive" || s == "sensitive" || s == "more-sensitive" || s == "very-sensitive" ||
      s == "ultra-sensitive")
    return 1.0;
  return 0.8;
}

inline int preset_index_chunks(const std::string& s) {
  return (s == "very-sensitive" || s == "ultra-sensitive") ? 1 : 4;
}

// Half-width of the diagonal band by query length.
inline uint32_t band_for(uint64_t len) {
  static const uint64_t thr[7] = {50, 100, 150, 200, 250, 350, 500};
  static const uint32_t band[8] = {12, 16, 30, 30, 30, 40, 64, 64};
  int b = 0;
  for (uint64_t t : thr) b += len >= t;
  return band[b];
}

// A parsed shape: offsets of the '1' positions, weight and span.
struct ShapeDef {
  uint32_t weight = 0, span = 0;
  uint8_t pos[kMaxShapeWeight] = {0};
};

inline bool parse_shape(const std::string& code, ShapeDef* out) {
  ShapeDef s;
  for (size_t i = 0; i < code.size(); ++i) {
    if (code[i] == '1') {
      if (s.weight >= (uint32_t)kMaxShapeWeight) return false;
      s.pos[s.weight++] = (uint8_t)i;
      s.span = (uint32_t)i + 1;
    } else if (code[i] == '0') {
      s.span = (uint32_t)i + 1;
    } else {
      return false;
    }
  }
  if (s.weight == 0 || s.span > (uint32_t)kMaxShapeSpan) return false;
  *out = s;
  return true;
}

// ASCII -> NCBI code as the reference encodes (unknown -> X, U -> C, O -> K).
inline uint8_t encode_letter(uint8_t c) {
  static const char sym[] = "ARNDCQEGHILKMFPSTWYVBJZX*";
  struct T {
    uint8_t t[256];
    T() {
      for (int i = 0; i < 256; ++i) t[i] = 23;
      for (int i = 0; i < 25; ++i) {
        unsigned char ch = (unsigned char)sym[i];
        t[ch] = (uint8_t)i;
        if (ch >= 'A' && ch <= 'Z') t[ch - 'A' + 'a'] = (uint8_t)i;
      }
      t['U'] = t['u'] = 4;
      t['O'] = t['o'] = 11;
    }
  };
  static const T tab;
  return tab.t[c];
}

inline char decode_letter(uint8_t l) { return "ARNDCQEGHILKMFPSTWYVBJZX*"[l < 24 ? l : 24]; }

// Reduction: letter -> bucket (255 for non-canonical letters) and its size.
struct ReductionDef {
  uint8_t map[25];
  uint32_t size = 0;
};

inline ReductionDef parse_reduction(const char* def) {
  ReductionDef r;
  std::memset(r.map, 0xFF, sizeof r.map);
  uint32_t bucket = 0;
  const char* p = def;
  auto ws = [](char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == 0x0C; };
  while (*p) {
    while (*p && ws(*p)) ++p;
    if (!*p) break;
    while (*p && !ws(*p)) {
      uint8_t l = encode_letter((uint8_t)*p);
      if (l < kTrueAA) r.map[l] = (uint8_t)bucket;
      ++p;
    }
    ++bucket;
  }
  r.size = bucket;
  return r;
}

}  // namespace dmr
