// Score-pass kernel settings: which compiled packed kernel runs each class,
// which merged-band kernels are in use, the arithmetic and the stream plan.
// None of them changes a result. They come from built-in defaults, a tuning
// file written by the autotuner (diamond-rust_amd/autotune.py: settings per
// search preset, measured on the device) and dmr_ctx_set_tuning, in that
// order of precedence (last wins).
#pragma once

#include <map>
#include <string>

#include "dmr_kernels.hpp"

namespace dmr {

struct ScoreLayout {
  int lanes[kPackedClasses];    // lanes per DP pair (1 = k_sw16, classes 0-4 only)
  int threads[kPackedClasses];  // threads per block
  unsigned merged_mask;         // merged kernels in use: bit i = class kExactClasses + i
  bool f16;                     // packed pairs in f16 (else saturating int16)
  int plan;                     // how the class kernels share the streams (0, 1, 2)
  bool graph;                   // ungapped .. score-pass preprocessing replayed as one hipGraph
  int transpose[kPackedClasses];  // row sequence of a packed DP: 0 query, 1 fewer rows, 2 subject
  bool fuse;                    // merged classes, w81, w61 in one launch (k_sw16p_fused)
};
// Row-sequence modes of the packed classes, 2 bits per class (k_classify).
uint32_t transpose_modes(const ScoreLayout& L);

// Built-in defaults.
ScoreLayout default_layout();
// Name of packed class c as used in keys: w25 w33 w61 w81 w129 (band_for
// widths), m97 m129 m161 m193 m225 m257 (merged-band width bounds).
std::string packed_class_name(int c);

using Settings = std::map<std::string, double>;

// Applies settings to L: "lanes_<class>" (a compiled lane count; also sets
// the class's default block size for it), then "threads_<class>" (a compiled
// block size for the class's lanes), "merged_mask" (8 bits, m97 .. m513; m257
// is always in use), "f16" (0/1), "score_plan" (0-2), "graph" (0/1),
// "transpose_<class>" (0-2: which sequence runs along the rows of the
// class's DPs), "transpose" (every class's; reads -1 when they differ),
// "fuse" (0/1).
// Throws Error(-1) naming the key on an unknown key or a value with no
// compiled kernel.
void apply_settings(ScoreLayout* L, const Settings& s);
// Value of a layout key in L (throws Error(-1) on an unknown key).
double layout_value(const ScoreLayout& L, const std::string& key);
bool is_layout_key(const std::string& key);

// Settings per preset of a tuning file: {"presets": {"<preset>": {"settings":
// {"<key>": number, ...}, ...}, ...}, ...}; other members are ignored. A
// preset named "*" applies to every preset before the preset's own entry.
// Throws Error(-1) with the position on malformed JSON.
std::map<std::string, Settings> parse_tuning(const std::string& text);

}  // namespace dmr
