// Score-pass kernel settings: defaults, keys, and the tuning file reader.
#include "tuning.hpp"

#include <cmath>
#include <cstdlib>
#include <cstring>

#include "engine.hpp"

namespace dmr {

namespace {

const char* const kClassNames[kPackedClasses] = {"w25", "w33", "w61", "w81", "w129", "m97", "m129",
                                                  "m161", "m193", "m225", "m257", "m385", "m513"};

int class_of(const std::string& name) {
  for (int c = 0; c < kPackedClasses; ++c)
    if (name == kClassNames[c]) return c;
  return -1;
}

[[noreturn]] void bad(const std::string& m) { throw Error(-1, m); }

// First compiled block size for (cls, lanes), or 0.
int default_threads(int cls, int lanes) {
  for (const Sw16Variant& v : sw16_variants())
    if (v.cls == cls && v.lanes == lanes) return v.threads;
  return 0;
}

int as_int(const std::string& key, double v) {
  if (!(v == std::floor(v)) || std::fabs(v) > 1e9) bad("tuning key " + key + ": not an integer");
  return (int)v;
}

// Minimal JSON reader: the values the tuning file holds (objects, arrays,
// strings, numbers, true/false/null), keeping only numbers and objects.
struct Json {
  enum Kind { kNull, kNum, kStr, kObj, kArr } kind = kNull;
  double num = 0;
  std::string str;
  std::map<std::string, Json> obj;
};

class Reader {
 public:
  explicit Reader(const std::string& t) : t_(t) {}
  Json value() {
    ws();
    if (p_ >= t_.size()) fail("unexpected end");
    const char c = t_[p_];
    Json j;
    if (c == '{') {
      j.kind = Json::kObj;
      ++p_;
      ws();
      if (peek('}')) return j;
      for (;;) {
        ws();
        std::string k = string();
        ws();
        expect(':');
        j.obj[k] = value();
        ws();
        if (peek('}')) return j;
        expect(',');
      }
    }
    if (c == '[') {
      j.kind = Json::kArr;
      ++p_;
      ws();
      if (peek(']')) return j;
      for (;;) {
        value();
        ws();
        if (peek(']')) return j;
        expect(',');
      }
    }
    if (c == '"') {
      j.kind = Json::kStr;
      j.str = string();
      return j;
    }
    for (const char* w : {"true", "false", "null"})
      if (t_.compare(p_, std::strlen(w), w) == 0) {
        p_ += std::strlen(w);
        if (w[0] != 'n') {
          j.kind = Json::kNum;
          j.num = w[0] == 't' ? 1 : 0;
        }
        return j;
      }
    const char* b = t_.c_str() + p_;
    char* e = nullptr;
    j.num = std::strtod(b, &e);
    if (e == b) fail("unexpected character");
    j.kind = Json::kNum;
    p_ += (size_t)(e - b);
    return j;
  }
  void end() {
    ws();
    if (p_ != t_.size()) fail("trailing characters");
  }

 private:
  [[noreturn]] void fail(const char* what) { bad("tuning file: " + std::string(what) + " at byte " + std::to_string(p_)); }
  void ws() {
    while (p_ < t_.size() && (t_[p_] == ' ' || t_[p_] == '\n' || t_[p_] == '\r' || t_[p_] == '\t')) ++p_;
  }
  bool peek(char c) {
    if (p_ < t_.size() && t_[p_] == c) {
      ++p_;
      return true;
    }
    return false;
  }
  void expect(char c) {
    if (!peek(c)) fail("malformed JSON");
  }
  std::string string() {
    expect('"');
    std::string s;
    while (p_ < t_.size() && t_[p_] != '"') {
      if (t_[p_] == '\\') {
        if (++p_ >= t_.size()) break;
        const char e = t_[p_];
        if (e == 'u') {
          p_ += 4;  // keys and names here are ASCII; escaped code points are skipped
          s += '?';
        } else {
          s += e == 'n' ? '\n' : e == 't' ? '\t' : e;
        }
        ++p_;
        continue;
      }
      s += t_[p_++];
    }
    expect('"');
    return s;
  }
  const std::string& t_;
  size_t p_ = 0;
};

}  // namespace

std::string packed_class_name(int c) { return kClassNames[c]; }

ScoreLayout default_layout() {
  ScoreLayout L{};
  // Band_for widths: 81 and 129 split (two and four lanes per pair), the
  // narrower widths whole in one lane; merged bands over eight lanes, those
  // wider than 256 over sixteen.
  for (int c = 0; c < kPackedClasses; ++c) {
    L.lanes[c] = c < kExactClasses ? 1 : (sw16_width(c) > 257 ? 16 : 8);
    L.threads[c] = c < kExactClasses ? 256 : 512;
  }
  L.lanes[3] = 2;
  L.threads[3] = 512;
  L.lanes[4] = 4;
  L.threads[4] = 768;
  L.merged_mask = 0x28u;  // width bounds 193 and 257 (wider bands: int32 wave kernels)
  L.f16 = true;
  L.plan = 0;
  L.graph = false;
  // Row sequence: the subject for the widest band_for class (its long
  // queries meet shorter subjects), the query elsewhere.
  for (int c = 0; c < kPackedClasses; ++c) L.transpose[c] = 0;
  L.transpose[4] = 2;
  L.fuse = false;
  return L;
}

uint32_t transpose_modes(const ScoreLayout& L) {
  uint32_t m = 0;
  for (int c = 0; c < kPackedClasses; ++c) m |= (uint32_t)L.transpose[c] << (2 * c);
  return m;
}

bool is_layout_key(const std::string& k) {
  if (k == "merged_mask" || k == "f16" || k == "score_plan" || k == "graph" || k == "transpose" || k == "fuse")
    return true;
  for (const char* pre : {"lanes_", "threads_", "transpose_"}) {
    const size_t n = std::strlen(pre);
    if (k.compare(0, n, pre) == 0 && class_of(k.substr(n)) >= 0) return true;
  }
  return false;
}

void apply_settings(ScoreLayout* L, const Settings& s) {
  // Lane counts first: each resets its class's block size to the first
  // compiled one, which a threads_ key then overrides.
  for (const auto& [k, v] : s) {
    if (k.compare(0, 6, "lanes_") != 0) continue;
    const int c = class_of(k.substr(6));
    if (c < 0) bad("unknown tuning key: " + k);
    const int p = as_int(k, v), t = default_threads(c, p);
    if (!t) bad("tuning key " + k + ": no compiled kernel with " + std::to_string(p) + " lanes per DP pair");
    L->lanes[c] = p;
    L->threads[c] = t;
  }
  for (const auto& [k, v] : s) {
    if (k.compare(0, 6, "lanes_") == 0) continue;
    if (k.compare(0, 8, "threads_") == 0) {
      const int c = class_of(k.substr(8));
      if (c < 0) bad("unknown tuning key: " + k);
      const int t = as_int(k, v);
      if (!sw16_has_variant(c, L->lanes[c], t))
        bad("tuning key " + k + ": no compiled kernel with " + std::to_string(L->lanes[c]) +
            " lanes per DP pair and " + std::to_string(t) + " threads");
      L->threads[c] = t;
    } else if (k == "merged_mask") {
      const int m = as_int(k, v);
      if (m < 0 || m >= (1 << kMergedClasses)) bad("merged_mask is an 8-bit mask");
      L->merged_mask = (unsigned)m | (1u << kMergedAlways);
    } else if (k == "f16") {
      L->f16 = v != 0;
    } else if (k == "graph") {
      L->graph = v != 0;
    } else if (k == "fuse") {
      L->fuse = v != 0;
    } else if (k == "transpose" || k.compare(0, 10, "transpose_") == 0) {
      const int c = k == "transpose" ? -1 : class_of(k.substr(10));
      if (c < 0 && k != "transpose") bad("unknown tuning key: " + k);
      const int t = as_int(k, v);
      if (t < 0 || t > 2) bad(k + " is 0, 1 or 2");
      for (int w = 0; w < kPackedClasses; ++w)
        if (c < 0 || w == c) L->transpose[w] = t;
    } else if (k == "score_plan") {
      const int p = as_int(k, v);
      if (p < 0 || p > 2) bad("score_plan is 0, 1 or 2");
      L->plan = p;
    } else {
      bad("unknown tuning key: " + k);
    }
  }
}

double layout_value(const ScoreLayout& L, const std::string& k) {
  if (k == "merged_mask") return L.merged_mask;
  if (k == "f16") return L.f16 ? 1 : 0;
  if (k == "score_plan") return L.plan;
  if (k == "graph") return L.graph ? 1 : 0;
  if (k == "fuse") return L.fuse ? 1 : 0;
  if (k == "transpose") {
    for (int c = 1; c < kPackedClasses; ++c)
      if (L.transpose[c] != L.transpose[0]) return -1;
    return L.transpose[0];
  }
  if (k.compare(0, 10, "transpose_") == 0 && class_of(k.substr(10)) >= 0) return L.transpose[class_of(k.substr(10))];
  if (k.compare(0, 6, "lanes_") == 0 && class_of(k.substr(6)) >= 0) return L.lanes[class_of(k.substr(6))];
  if (k.compare(0, 8, "threads_") == 0 && class_of(k.substr(8)) >= 0) return L.threads[class_of(k.substr(8))];
  bad("unknown tuning key: " + k);
}

std::map<std::string, Settings> parse_tuning(const std::string& text) {
  Reader r(text);
  const Json root = r.value();
  r.end();
  std::map<std::string, Settings> out;
  if (root.kind != Json::kObj) bad("tuning file: not a JSON object");
  auto pr = root.obj.find("presets");
  if (pr == root.obj.end()) return out;
  if (pr->second.kind != Json::kObj) bad("tuning file: \"presets\" is not an object");
  for (const auto& [name, entry] : pr->second.obj) {
    if (entry.kind != Json::kObj) bad("tuning file: preset " + name + " is not an object");
    auto st = entry.obj.find("settings");
    if (st == entry.obj.end() || st->second.kind != Json::kObj)
      bad("tuning file: preset " + name + " has no \"settings\" object");
    Settings s;
    for (const auto& [k, v] : st->second.obj) {
      if (v.kind != Json::kNum) bad("tuning file: preset " + name + " key " + k + " is not a number");
      s[k] = v.num;
    }
    ScoreLayout check = default_layout();
    apply_settings(&check, s);  // every key known and compiled
    out[name] = s;
  }
  return out;
}

}  // namespace dmr
