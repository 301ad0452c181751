// FASTA ingress: io::parse_fasta (src/io.rs:40-102), alphabet::encode
// (src/alphabet.rs:54-63) and io::looks_like_dna (src/io.rs:104-121), on the
// host. Large inputs are cut at lines that start with '>' and the parts are
// parsed on the host worker threads; a part's records are the same as the
// whole input's because a record never spans a '>' line.
#include "../../include/dmr.h"

#include <algorithm>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "dmr_tables.hpp"
#include "host_pool.hpp"

struct dmr_fasta {
  std::unique_ptr<uint8_t[]> raw;  // upper-cased residues, layout characters dropped
  uint64_t nraw = 0;
  std::vector<uint64_t> offs;  // nrec + 1
  std::string ids;             // Record::id of each record, each followed by '\n'
  std::string headers;         // Record::header of each record, each followed by '\n'
  bool dna = false;
};

namespace {

// Bytes u8::is_ascii_whitespace accepts: space, tab, LF, form feed, CR.
inline bool ascii_ws(uint8_t b) { return b == ' ' || b == '\t' || b == '\n' || b == '\f' || b == '\r'; }

struct ByteClass {
  uint8_t keep[256];   // residue bytes kept by parse_fasta
  uint8_t upper[256];  // to_ascii_uppercase
  ByteClass() {
    for (int b = 0; b < 256; ++b) {
      keep[b] = !(ascii_ws((uint8_t)b) || b == '-' || b == '.');
      upper[b] = (uint8_t)(b >= 'a' && b <= 'z' ? b - 32 : b);
    }
  }
};
const ByteClass kByte;

void put_utf8(uint32_t cp, std::string* out) {
  if (cp < 0x80) {
    out->push_back((char)cp);
  } else if (cp < 0x800) {
    out->push_back((char)(0xC0 | cp >> 6));
    out->push_back((char)(0x80 | (cp & 0x3F)));
  } else if (cp < 0x10000) {
    out->push_back((char)(0xE0 | cp >> 12));
    out->push_back((char)(0x80 | (cp >> 6 & 0x3F)));
    out->push_back((char)(0x80 | (cp & 0x3F)));
  } else {
    out->push_back((char)(0xF0 | cp >> 18));
    out->push_back((char)(0x80 | (cp >> 12 & 0x3F)));
    out->push_back((char)(0x80 | (cp >> 6 & 0x3F)));
    out->push_back((char)(0x80 | (cp & 0x3F)));
  }
}

// String::from_utf8_lossy: valid sequences copied, each maximal invalid
// subpart (a lead byte and the continuation bytes that could still complete
// it) replaced by one U+FFFD, as the Unicode standard's "substitution of
// maximal subparts" practice.
void utf8_lossy(const uint8_t* s, size_t n, std::string* out) {
  size_t i = 0;
  while (i < n) {
    const uint8_t c = s[i];
    if (c < 0x80) {
      out->push_back((char)c);
      ++i;
      continue;
    }
    size_t len = 0;
    uint8_t lo = 0x80, hi = 0xBF;  // range of the second byte
    if (c >= 0xC2 && c <= 0xDF) {
      len = 2;
    } else if (c >= 0xE0 && c <= 0xEF) {
      len = 3;
      if (c == 0xE0) lo = 0xA0;
      if (c == 0xED) hi = 0x9F;
    } else if (c >= 0xF0 && c <= 0xF4) {
      len = 4;
      if (c == 0xF0) lo = 0x90;
      if (c == 0xF4) hi = 0x8F;
    }
    size_t k = 1;
    if (len) {
      while (k < len && i + k < n) {
        const uint8_t b = s[i + k];
        const bool ok = k == 1 ? (b >= lo && b <= hi) : (b >= 0x80 && b <= 0xBF);
        if (!ok) break;
        ++k;
      }
    }
    if (len && k == len) {
      out->append(reinterpret_cast<const char*>(s + i), len);
    } else {
      put_utf8(0xFFFD, out);
    }
    i += k;
  }
}

// Decodes the code point starting at s[i] of valid UTF-8 (length in *len).
uint32_t cp_at(const std::string& s, size_t i, size_t* len) {
  const uint8_t c = (uint8_t)s[i];
  if (c < 0x80) {
    *len = 1;
    return c;
  }
  const size_t l = c >= 0xF0 ? 4 : c >= 0xE0 ? 3 : 2;
  uint32_t cp = c & (l == 2 ? 0x1F : l == 3 ? 0x0F : 0x07);
  for (size_t k = 1; k < l; ++k) cp = cp << 6 | ((uint8_t)s[i + k] & 0x3F);
  *len = l;
  return cp;
}

// char::is_whitespace: the Unicode White_Space property.
bool unicode_ws(uint32_t cp) {
  return (cp >= 0x09 && cp <= 0x0D) || cp == 0x20 || cp == 0x85 || cp == 0xA0 || cp == 0x1680 ||
         (cp >= 0x2000 && cp <= 0x200A) || cp == 0x2028 || cp == 0x2029 || cp == 0x202F || cp == 0x205F ||
         cp == 0x3000;
}

// str::trim of valid UTF-8.
std::string trim_unicode(const std::string& s) {
  size_t a = 0;
  while (a < s.size()) {
    size_t l;
    if (!unicode_ws(cp_at(s, a, &l))) break;
    a += l;
  }
  size_t b = a, i = a;
  while (i < s.size()) {
    size_t l;
    const bool ws = unicode_ws(cp_at(s, i, &l));
    i += l;
    if (!ws) b = i;
  }
  return s.substr(a, b - a);
}

struct Part {
  std::unique_ptr<uint8_t[]> raw;  // capacity: the part's byte count
  uint64_t nraw = 0;
  std::vector<uint64_t> lens;
  std::string ids, headers;
  bool seen_header = false;
  int error = 0;  // 1: residues before the first header
};

void finish(const std::string& header, uint64_t len, Part* p) {
  size_t e = 0;
  while (e < header.size() && !ascii_ws((uint8_t)header[e])) ++e;
  if (e) {
    p->ids.append(header, 0, e);
  } else {
    p->ids += "seq" + std::to_string(std::max<uint64_t>(len, 1));
  }
  p->ids.push_back('\n');
  p->headers += header;
  p->headers.push_back('\n');
  p->lens.push_back(len);
}

// parse_fasta over d[0, n) (src/io.rs:40-74); lines split at '\n' with one
// trailing '\r' dropped (split_lines, src/io.rs:96-102).
void parse_part(const uint8_t* d, size_t n, Part* p) {
  p->raw.reset(new uint8_t[n ? n : 1]);
  uint8_t* const dst = p->raw.get();
  uint64_t j = 0;
  bool in_record = false;
  std::string header;
  uint64_t start = 0;
  size_t pos = 0;
  while (true) {
    const uint8_t* nl = static_cast<const uint8_t*>(std::memchr(d + pos, '\n', n - pos));
    const size_t end = nl ? (size_t)(nl - d) : n;
    size_t lend = end;
    if (lend > pos && d[lend - 1] == '\r') --lend;
    const uint8_t* line = d + pos;
    const size_t ll = lend - pos;
    if (ll && line[0] == '>') {
      if (in_record) finish(header, j - start, p);
      in_record = p->seen_header = true;
      std::string lossy;
      utf8_lossy(line + 1, ll - 1, &lossy);
      header = trim_unicode(lossy);
      start = j;
    } else if (in_record) {
      // Every byte is written; j advances over the kept ones only.
      for (size_t k = 0; k < ll; ++k) {
        const uint8_t b = line[k];
        dst[j] = kByte.upper[b];
        j += kByte.keep[b];
      }
    } else if (ll && line[0] != ';' && line[0] != '#') {
      p->error = 1;
      return;
    }
    if (!nl) break;
    pos = end + 1;
  }
  if (in_record) finish(header, j - start, p);
  p->nraw = j;
}

// looks_like_dna (src/io.rs:104-121): the first records' residues, at most
// 20,000 per record, until 20,000 are sampled; DNA when >= 85% are ACGTUN.
bool looks_like_dna(const dmr_fasta& f) {
  uint64_t canonical = 0, total = 0;
  const uint64_t nrec = f.offs.size() - 1;
  for (uint64_t r = 0; r < nrec; ++r) {
    const uint64_t a = f.offs[r], b = std::min(f.offs[r + 1], a + 20000);
    for (uint64_t i = a; i < b; ++i) {
      const uint8_t c = f.raw.get()[i];
      ++total;
      canonical += c == 'A' || c == 'C' || c == 'G' || c == 'T' || c == 'U' || c == 'N';
    }
    if (total >= 20000) break;
  }
  return total > 0 && canonical * 100 >= total * 85;
}

// Part boundaries: starts of lines that begin with '>', near equal splits.
std::vector<size_t> cut_points(const uint8_t* d, size_t n, unsigned parts) {
  std::vector<size_t> cuts{0};
  for (unsigned k = 1; k < parts; ++k) {
    size_t at = std::max(cuts.back(), (size_t)((unsigned __int128)n * k / parts));
    while (at < n) {
      const uint8_t* q = static_cast<const uint8_t*>(std::memchr(d + at, '\n', n - at));
      if (!q || (size_t)(q - d) + 1 >= n) {
        at = n;
        break;
      }
      at = (size_t)(q - d) + 1;
      if (d[at] == '>') break;
    }
    if (at >= n) break;
    if (at > cuts.back()) cuts.push_back(at);
  }
  cuts.push_back(n);
  return cuts;
}

}  // namespace

namespace dmr {
// The calling thread's dmr_last_error text (api.cpp).
void set_last_error(const std::string& msg);
}  // namespace dmr

int dmr_fasta_parse(const char* data, uint64_t len, dmr_fasta** out) {
  if (!out || (len && !data)) {
    dmr::set_last_error("NULL argument");
    return DMR_E_ARG;
  }
  *out = nullptr;
  try {
    const uint8_t* d = reinterpret_cast<const uint8_t*>(len ? data : "");
    constexpr size_t kMinPart = 1 << 20;
    const unsigned want = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(dmr::host_threads(), len / kMinPart));
    const std::vector<size_t> cuts = cut_points(d, (size_t)len, want);
    const size_t np = cuts.size() - 1;
    std::vector<Part> parts(np);
    dmr::host_parallel_for(np, 1, [&](uint64_t a, uint64_t b) {
      for (uint64_t k = a; k < b; ++k) parse_part(d + cuts[k], cuts[k + 1] - cuts[k], &parts[k]);
    });
    // Only the first part can hold lines before a header: the others start
    // with a '>' line.
    if (np && parts[0].error) {
      dmr::set_last_error("malformed FASTA: sequence data found before the first '>' header");
      return DMR_E_ARG;
    }
    bool seen = false;
    for (const Part& p : parts) seen = seen || p.seen_header;
    if (!seen && len) {
      dmr::set_last_error("malformed FASTA: no '>' header found");
      return DMR_E_ARG;
    }
    auto f = std::make_unique<dmr_fasta>();
    uint64_t nres = 0, nrec = 0, nid = 0, nhd = 0;
    for (const Part& p : parts) {
      nres += p.nraw;
      nrec += p.lens.size();
      nid += p.ids.size();
      nhd += p.headers.size();
    }
    // Each part's residues, ids and headers go to their final places in
    // parallel; the offsets are a running sum.
    std::vector<uint64_t> raw_at(np + 1, 0);
    for (size_t k = 0; k < np; ++k) raw_at[k + 1] = raw_at[k] + parts[k].nraw;
    f->nraw = nres;
    if (np == 1) {
      f->raw = std::move(parts[0].raw);
    } else {
      f->raw.reset(new uint8_t[nres ? nres : 1]);
      dmr::host_parallel_for(np, 1, [&](uint64_t a, uint64_t b) {
        for (uint64_t k = a; k < b; ++k)
          if (parts[k].nraw) std::memcpy(f->raw.get() + raw_at[k], parts[k].raw.get(), parts[k].nraw);
      });
    }
    if (!f->raw) f->raw.reset(new uint8_t[1]);
    f->offs.reserve(nrec + 1);
    f->ids.reserve(nid);
    f->headers.reserve(nhd);
    f->offs.push_back(0);
    for (Part& p : parts) {
      for (uint64_t l : p.lens) f->offs.push_back(f->offs.back() + l);
      f->ids += p.ids;
      f->headers += p.headers;
      p.raw.reset();
    }
    f->dna = looks_like_dna(*f);
    *out = f.release();
    return DMR_OK;
  } catch (const std::bad_alloc&) {
    dmr::set_last_error("host allocation failed");
    return DMR_E_OOM;
  }
}

int dmr_fasta_sizes(const dmr_fasta* f, uint64_t* nrec, uint64_t* nres, uint64_t* ids_len,
                    uint64_t* headers_len, int* looks_like_dna) {
  if (!f) {
    dmr::set_last_error("NULL argument");
    return DMR_E_ARG;
  }
  if (nrec) *nrec = f->offs.size() - 1;
  if (nres) *nres = f->nraw;
  if (ids_len) *ids_len = f->ids.size();
  if (headers_len) *headers_len = f->headers.size();
  if (looks_like_dna) *looks_like_dna = f->dna ? 1 : 0;
  return DMR_OK;
}

int dmr_fasta_residues(const dmr_fasta* f, char* raw, uint8_t* letters, uint64_t* offsets) {
  if (!f) {
    dmr::set_last_error("NULL argument");
    return DMR_E_ARG;
  }
  const uint64_t n = f->nraw;
  if (raw && n) std::memcpy(raw, f->raw.get(), n);
  if (letters) {
    uint8_t tab[256];
    for (int b = 0; b < 256; ++b) tab[b] = dmr::encode_letter((uint8_t)b);
    const uint8_t* src = f->raw.get();
    dmr::host_parallel_for(n, 1 << 20, [&](uint64_t a, uint64_t b) {
      for (uint64_t i = a; i < b; ++i) letters[i] = tab[src[i]];
    });
  }
  if (offsets) std::memcpy(offsets, f->offs.data(), f->offs.size() * sizeof(uint64_t));
  return DMR_OK;
}

int dmr_fasta_text(const dmr_fasta* f, char* ids, char* headers) {
  if (!f) {
    dmr::set_last_error("NULL argument");
    return DMR_E_ARG;
  }
  if (ids && !f->ids.empty()) std::memcpy(ids, f->ids.data(), f->ids.size());
  if (headers && !f->headers.empty()) std::memcpy(headers, f->headers.data(), f->headers.size());
  return DMR_OK;
}

void dmr_fasta_destroy(dmr_fasta* f) { delete f; }
