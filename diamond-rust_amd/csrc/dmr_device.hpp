// Device-side data layout and kernel launchers of the search core.
//
// HBM layout (one resident database, one batch of query targets):
//   residues   uint8 codes 0..24, sequences concatenated without separators
//   offsets    uint32 start of each sequence (n+1 entries)
//   seq_of     uint32 per residue: the sequence it belongs to
//   seeds      (key u64, loc u32) sorted by key; loc = global residue index
//   hits       u64 packed (target | subject | diag+bias | spos), sorted
//   segments   SegRec per kept ungapped segment, in hit order
//   jobs       SwJob per banded DP, grouped by (target, subject), band rank order
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace dmr {

constexpr int kScoreStride = 32;  // row stride of the 25x25 matrix in memory tables
constexpr uint8_t kPadLetter = 25;  // subject padding code (scores -64 against everything)

struct ShapeDev {
  uint32_t weight, span;
  uint8_t pos[16];
};

struct SeedParams {
  uint32_t nshapes;
  uint32_t red_size;
  uint8_t red_map[25];
  double cut;           // complexity cut (<= 0 disables)
  double lnfact[33];    // t[n] = t[n-1] + ln n, built on the host
  ShapeDev shapes[64];
};

// Packed hit-key geometry (bit widths of target, subject, diag, spos).
struct HitGeom {
  uint32_t bt, bs, bd, bp;
  int64_t diag_bias;  // added to diag = qpos - spos so it is non-negative
};

struct SegRec {
  uint32_t q_start, s_start, len;
  int32_t score;
};

// One banded DP: query window [q0, q1) of target `target` against subject
// `subject`, diagonals [d0, d1] relative to the window start.
struct SwJob {
  uint32_t target, subject;
  uint32_t q0, q1;
  int32_t d0, d1;
  uint32_t group;   // (target, subject) group index
  uint32_t rank;    // position of the band in the group's score order
};

struct SwOut {
  int32_t score;
  uint32_t best_i, best_j;  // first strict maximum (row-major), window-relative
  uint32_t q_begin, s_begin;
  uint32_t identities, positives, mismatches, gap_openings, gaps, length;
  uint32_t n_ops;
  uint32_t flags;  // bit0: saturated (int16 pass)
};

// ---- launchers (stream-ordered; no host synchronisation inside) ----------
void scan_exclusive_u64(const uint64_t* in, uint64_t* out, uint64_t n, uint64_t* tmp,
                        hipStream_t st);
uint64_t scan_tmp_elems(uint64_t n);
void scan_exclusive_u32_to_u64(const uint32_t* in, uint64_t* out, uint64_t n, uint64_t* tmp,
                               hipStream_t st);

// LSD radix sort of (key, value) pairs on the low `bits` bits of key.
// Sorted result lands in (k0, v0); (k1, v1) is scratch of the same size.
// n_dev: sort the first *n_dev of a buffer of n (a bound; the grid), the
// count read on the device.
void radix_sort_pairs(uint64_t* k0, uint32_t* v0, uint64_t* k1, uint32_t* v1, uint64_t n, int bits,
                      uint32_t* hist, uint64_t* scan_tmp, hipStream_t st, const uint64_t* n_dev = nullptr);
void radix_sort_keys(uint64_t* k0, uint64_t* k1, uint64_t n, int bits, uint32_t* hist,
                     uint64_t* scan_tmp, hipStream_t st, const uint64_t* n_dev = nullptr);
uint64_t radix_hist_elems(uint64_t n);
uint64_t radix_scan_tmp_elems(uint64_t n);

}  // namespace dmr
