// Packed score pass: job classes, lane pairing and the launches of the
// band_for width classes (kernels in sw16_kernels.hpp).
#include "sw16_kernels.hpp"

namespace dmr {

namespace {

// Merged-width bucket of a width in 62..512: 62..97, then 32 widths each
// (98..129, ..., 226..256), then 64 each (257..320, ..., 449..512).
__host__ __device__ __forceinline__ int merged_bucket(int w) {
  return w <= 97 ? 0 : w <= 256 ? (w - 66) >> 5 : 6 + ((w - 257) >> 6);
}
constexpr int kMergedBuckets = 10;
// Largest width of each bucket.
__host__ __device__ __forceinline__ int merged_bucket_top(int b) {
  return b < 5 ? 97 + 32 * b : b == 5 ? 256 : 256 + 64 * (b - 5);
}

// Packed class of a band width: 0-4 the widths 2 * band + 1 of the band_for
// values (25/33/61/81/129); a merged band (widths 62..512 other than 81 and
// 129) by its bucket b: class (merged >> 4b) & 15, the narrowest merged
// kernel in use that holds it (sw16_merged_map; 15 for none: widths past 256
// when neither m385 nor m513 is in use); -1 for any other width.
__device__ __forceinline__ int packed_class(int w, uint64_t merged) {
  switch (w) {
    case 25: return 0;
    case 33: return 1;
    case 61: return 2;
    case 81: return 3;
    case 129: return 4;
    default:
      if (w <= 61 || w > 512) return -1;
      const int c = (int)((merged >> (4 * merged_bucket(w))) & 15u);
      return c == 15 ? -1 : c;
  }
}

// Sort key per job: class | orientation | row sequence | rows. Classes below
// kPackedClasses: packed 16-bit kernels (packed_class); then kWaveKinds
// wave-per-DP int32 classes with K = 1..32 offsets per lane (widths up to
// 64 * K); the last: wider bands (block-per-DP and row-exact int32 kernels).
// A packed job runs with the query or the subject on the rows (subject_rows
// with its class's mode, bits 2c..2c+1 of transpose; the other classes
// always with the query), and the
// packed kernels pair two jobs with the same row sequence in a lane, so the
// key groups the jobs of one class and orientation by their row sequence
// (target or subject, seq_bits bits). Within a run they are ordered by row
// span and first row (32-row buckets): a lane runs the union of its two
// DPs' rows, and a wave as many rows as its longest lane, so neighbours
// should start and end alike.
__global__ void k_classify(const SwJob* __restrict__ jobs, uint64_t njobs, SeqSet q, SeqSet db,
                           int seq_bits, uint64_t merged, uint32_t transpose, uint64_t* __restrict__ keys,
                           uint32_t* __restrict__ idx, const uint64_t* __restrict__ njobs_dev) {
  uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (njobs_dev ? *njobs_dev : njobs)) return;
  const SwJob j = jobs[k];
  const int64_t qlen = q.offs[j.target + 1] - q.offs[j.target];
  const int64_t m = db.offs[j.subject + 1] - db.offs[j.subject];
  const int64_t n = (int64_t)j.q1 - j.q0;
  const int64_t w = (int64_t)j.d1 - j.d0 + 1;
  int cls = packed_class((int)(w > 1000 ? 1000 : w), merged);
  const int64_t rlo = j.d0 + 1 > 1 ? j.d0 + 1 : 1;
  const int64_t istar = m + j.d1;
  const int64_t rhi = istar < n ? istar : n;
  if (j.q0 != 0 || n != qlen || rlo > rhi || istar < 1 || m < 1 || n > 16383 || m > 16383) cls = -1;
  if (cls < 0) {
    cls = kPackedClasses + kWaveKinds;
    for (int c = 0; c < kWaveKinds; ++c)
      if (w <= (64ll << c)) {
        cls = kPackedClasses + c;
        break;
      }
  }
  const bool tr = cls < kPackedClasses && subject_rows(n, m, j.d0, j.d1, (transpose >> (2 * cls)) & 3u);
  // Rows in the job's orientation.
  const int64_t lo = tr ? max((int64_t)1, (int64_t)(1 - j.d1)) : rlo;
  const int64_t hi = tr ? min(m, n - (int64_t)j.d0) : rhi;
  const int64_t span = hi - lo + 1;
  const uint64_t sb = (uint64_t)(span < 0 ? 0 : (span > 16383 ? 16383 : span)) >> 5;
  const uint64_t lb = (uint64_t)(lo > 16383 ? 16383 : lo) >> 5;
  const uint64_t seq = tr ? j.subject : j.target;
  // With the subject on the rows, two DPs of one subject (from different
  // queries) may cover distant parts of it: those runs are ordered by first
  // row, then span, so a lane's two row ranges overlap.
  const uint64_t rows = tr ? (lb << 9) | sb : (sb << 9) | lb;
  keys[k] = ((uint64_t)cls << (seq_bits + 1 + kRowKeyBits)) | ((uint64_t)tr << (seq_bits + kRowKeyBits)) |
            (seq << kRowKeyBits) | rows;
  idx[k] = (uint32_t)k;
}

// One thread per sorted job. A job heads a lane when it sits at an even
// position of its (class, orientation, row sequence) run (run r of job i is
// pos[i] + flags[i] - 1, starting at run_start[r]) in a packed class (class <
// npacked); the class ranges of the sorted order go to the summary:
// summary[c] = first job of class c, summary[nclasses + c] = one past its
// last (absent classes keep 0), and summary[tfirst + c] = 1 + the first job
// of class c with the subject on the rows (the orientation bit lies right
// below the class; 0: none).
__global__ void k_lane_heads(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ flags,
                             const uint64_t* __restrict__ pos, const uint32_t* __restrict__ run_start,
                             uint64_t n, int cls_shift, int npacked, int nclasses, int tfirst,
                             uint32_t* __restrict__ heads, uint32_t* __restrict__ summary,
                             const uint64_t* __restrict__ n_dev) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (n_dev) {  // a device count within the bound n: no heads past it
    n = *n_dev;
    if (i >= n) {
      heads[i] = 0;
      return;
    }
  }
  auto cls_of = [&](uint64_t k) { return (int)min<uint64_t>(keys[k] >> cls_shift, (uint64_t)(nclasses - 1)); };
  auto tr_of = [&](uint64_t k) { return (int)((keys[k] >> (cls_shift - 1)) & 1u); };
  const int c = cls_of(i);
  const uint64_t r = pos[i] + flags[i] - 1;
  heads[i] = (c < npacked && ((i - run_start[r]) & 1) == 0) ? 1u : 0u;
  if (i == 0 || cls_of(i - 1) != c) summary[c] = (uint32_t)i;
  if (i + 1 == n || cls_of(i + 1) != c) summary[nclasses + c] = (uint32_t)(i + 1);
  if (c < npacked && tr_of(i) && (i == 0 || cls_of(i - 1) != c || !tr_of(i - 1)))
    summary[tfirst + c] = (uint32_t)(i + 1);
}

// One thread per sorted job: a lane head pairs itself with the next job of
// its run (none at the run's end) into lane hpos[i].
__global__ void k_make_lanes(const uint32_t* __restrict__ sorted_idx, const uint32_t* __restrict__ flags,
                             const uint32_t* __restrict__ heads, const uint64_t* __restrict__ hpos, uint64_t n,
                             uint2* __restrict__ lanes, const uint64_t* __restrict__ n_dev) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !heads[i]) return;
  if (n_dev) n = *n_dev;  // the job after the last is no partner
  lanes[hpos[i]] = make_uint2(sorted_idx[i], i + 1 < n && flags[i + 1] == 0 ? sorted_idx[i + 1] : ~0u);
}

// Lanes per packed class from the class ranges: summary[2 * nclasses + c] =
// first lane of class c, summary[2 * nclasses + npacked + c] = its lane
// count, summary[2 * (nclasses + npacked)] = all lanes, summary[trow + c] =
// the lanes of class c with the query on the rows (they come first).
__global__ void k_lane_summary(const uint64_t* __restrict__ hpos, uint64_t n, int npacked, int nclasses,
                               int tfirst, int trow, uint32_t* __restrict__ summary) {
  const int c = threadIdx.x;
  if (c < npacked) {
    const uint32_t b = summary[c], e = summary[nclasses + c];
    const bool any = e > b;
    const uint32_t t = summary[tfirst + c];
    summary[2 * nclasses + c] = any ? (uint32_t)hpos[b] : 0u;
    summary[2 * nclasses + npacked + c] = any ? (uint32_t)(hpos[e] - hpos[b]) : 0u;
    summary[trow + c] = any ? (uint32_t)(hpos[t ? t - 1 : e] - hpos[b]) : 0u;
  }
  if (c == 0) summary[2 * (nclasses + npacked)] = (uint32_t)hpos[n];
}

// Sort key per lane: class | orientation | (row span, longest first). The
// rows a lane runs are the union of its two DPs' row ranges in their
// orientation (subject_rows with the classification's modes), as in the
// packed kernels; the lanes of a class with the query on the rows stay first.
__global__ void k_lane_keys(const uint2* __restrict__ lanes, uint32_t nlanes, const SwJob* __restrict__ jobs,
                            SeqSet q, SeqSet db, uint64_t merged, uint32_t transpose, uint64_t* __restrict__ keys,
                            uint32_t* __restrict__ idx) {
  const uint32_t L = blockIdx.x * blockDim.x + threadIdx.x;
  if (L >= nlanes) return;
  const uint2 ln = lanes[L];
  int lo = INT32_MAX, hi = 0, cls = 0;
  bool tr = false;
  for (int h = 0; h < 2; ++h) {
    const uint32_t k = h ? ln.y : ln.x;
    if (k == ~0u) continue;
    const SwJob j = jobs[k];
    const int n = (int)(q.offs[j.target + 1] - q.offs[j.target]);
    const int m = (int)(db.offs[j.subject + 1] - db.offs[j.subject]);
    cls = packed_class(j.d1 - j.d0 + 1, merged);
    tr = subject_rows(n, m, j.d0, j.d1, (transpose >> (2 * cls)) & 3u);
    lo = min(lo, tr ? max(1, 1 - j.d1) : max(1, j.d0 + 1));
    hi = max(hi, tr ? min(m, n - j.d0) : min(n, m + j.d1));
  }
  const int span = min(max(hi - lo + 1, 0), (int)kLaneSpanMax);
  keys[L] = ((uint64_t)cls << (kLaneSpanBits + 1)) | ((uint64_t)tr << kLaneSpanBits) |
            (uint64_t)(kLaneSpanMax - span);
  idx[L] = L;
}

__global__ void k_gather_lanes(const uint2* __restrict__ in, const uint32_t* __restrict__ idx, uint32_t n,
                               uint2* __restrict__ out) {
  const uint32_t L = blockIdx.x * blockDim.x + threadIdx.x;
  if (L < n) out[L] = in[idx[L]];
}

// One block per sequence: its letters times 4 at offs[k] + k + 2 and the
// padding code before it (and after the last).
__global__ void k_pad_letters(SeqSet s, uint32_t n, uint8_t* __restrict__ pad, uint32_t* __restrict__ pad_offs) {
  const uint32_t k = blockIdx.x;
  const uint32_t b = s.offs[k], len = s.offs[k + 1] - b, base = b + k + 2;
  for (uint32_t p = threadIdx.x; p < len; p += blockDim.x) pad[base + p] = (uint8_t)(4 * s.res[b + p]);
  if (threadIdx.x == 0) {
    pad_offs[k] = base;
    pad[base - 1] = 4 * 25;
    if (k == 0) pad[0] = pad[1] = 4 * 25;
    if (k + 1 == n) pad[base + len] = 4 * 25;
  }
}

inline unsigned nblk(uint64_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

}  // namespace

int pair_table_entries() { return kPairTab; }

// Score table for the packed kernel: entry (qa, sA, sB) holds
// (score(qa, sA) + o, score(qa, sB) + o) as two int16; letter 25 is padding.
// With the subject on the rows the kernels read entry (s, qA, qB) for
// (score(qA, s), score(qB, s)): BLOSUM62 is symmetric, so one table serves
// both orientations.
// IEEE half-precision bits of a small integer (|v| < 2048, exact).
static uint16_t half_bits(int v) {
  if (v == 0) return 0;
  const uint16_t sign = v < 0 ? 0x8000u : 0u;
  unsigned a = (unsigned)(v < 0 ? -v : v);
  int e = 0;
  while ((a >> e) > 1) ++e;                           // a = 1.f * 2^e
  const unsigned frac = (a << (10 - e)) & 0x3FFu;     // e <= 10: exact
  return (uint16_t)(sign | (unsigned)((e + 15) << 10) | frac);
}

void fill_pair_table(const int8_t (&m)[25][25], bool f16, std::vector<int>* out) {
  out->assign(kPairTab, 0);
  auto sc = [&](int a, int b) { return (a < 25 && b < 25) ? (int)m[a][b] : -64; };
  auto enc = [&](int v) { return f16 ? (uint32_t)half_bits(v) : (uint32_t)(uint16_t)(int16_t)v; };
  for (int qa = 0; qa < kCodes; ++qa)
    for (int a = 0; a < kCodes; ++a)
      for (int b = 0; b < kCodes; ++b) {
        const int lo = sc(qa, a) + kOpen, hi = sc(qa, b) + kOpen;
        (*out)[qa * kPairRow + a * kCodes + b] = (int)(enc(lo) | (enc(hi) << 16));
      }
}

void launch_classify(const SwJob* jobs, uint64_t njobs, SeqSet q, SeqSet db, int seq_bits, uint64_t merged_map,
                     uint32_t transpose, uint64_t* keys, uint32_t* idx, hipStream_t st, const uint64_t* njobs_dev) {
  if (!njobs) return;
  hipLaunchKernelGGL(k_classify, dim3(nblk(njobs, 256)), dim3(256), 0, st, jobs, njobs, q, db,
                     seq_bits, merged_map, transpose, keys, idx, njobs_dev);
}

void launch_lane_heads(const uint64_t* keys, const uint32_t* flags, const uint64_t* pos, const uint32_t* run_start,
                       uint64_t n, int cls_shift, int npacked, int nclasses, int tfirst, uint32_t* heads,
                       uint32_t* summary, hipStream_t st, const uint64_t* n_dev) {
  if (!n) return;
  hipLaunchKernelGGL(k_lane_heads, dim3(nblk(n, 256)), dim3(256), 0, st, keys, flags, pos, run_start, n,
                     cls_shift, npacked, nclasses, tfirst, heads, summary, n_dev);
}

void launch_make_lanes(const uint32_t* sorted_idx, const uint32_t* flags, const uint32_t* heads,
                       const uint64_t* hpos, uint64_t n, uint2* lanes, hipStream_t st,
                       const uint64_t* n_dev) {
  if (!n) return;
  hipLaunchKernelGGL(k_make_lanes, dim3(nblk(n, 256)), dim3(256), 0, st, sorted_idx, flags, heads, hpos, n,
                     lanes, n_dev);
}

void launch_lane_summary(const uint64_t* hpos, uint64_t n, int npacked, int nclasses, int tfirst, int trow,
                         uint32_t* summary, hipStream_t st) {
  hipLaunchKernelGGL(k_lane_summary, dim3(1), dim3(64), 0, st, hpos, n, npacked, nclasses, tfirst, trow, summary);
}

void launch_lane_keys(const uint2* lanes, uint32_t nlanes, const SwJob* jobs, SeqSet q, SeqSet db,
                      uint64_t merged_map, uint32_t transpose, uint64_t* keys, uint32_t* idx, hipStream_t st) {
  if (!nlanes) return;
  hipLaunchKernelGGL(k_lane_keys, dim3(nblk(nlanes, 256)), dim3(256), 0, st, lanes, nlanes, jobs, q, db,
                     merged_map, transpose, keys, idx);
}

void launch_gather_lanes(const uint2* in, const uint32_t* idx, uint32_t n, uint2* out, hipStream_t st) {
  if (!n) return;
  hipLaunchKernelGGL(k_gather_lanes, dim3(nblk(n, 256)), dim3(256), 0, st, in, idx, n, out);
}

void launch_pad_letters(SeqSet s, uint32_t n, uint8_t* pad, uint32_t* pad_offs, hipStream_t st) {
  if (!n) return;
  hipLaunchKernelGGL(k_pad_letters, dim3(n), dim3(256), 0, st, s, n, pad, pad_offs);
}

static const int kWidths[kPackedClasses] = {25, 33, 61, 81, 129, 97, 129, 161, 193, 225, 257, 385, 513};
int sw16_width(int cls) { return kWidths[cls]; }

uint64_t sw16_merged_map(unsigned mask) {
  // Widest width each merged kernel takes (m257 stops at 256).
  static const int kHolds[kMergedClasses] = {97, 129, 161, 193, 225, 256, 384, 512};
  mask |= 1u << kMergedAlways;
  uint64_t map = 0;
  for (int b = 0; b < kMergedBuckets; ++b) {
    int i = 0;
    while (i < kMergedClasses && (!((mask >> i) & 1u) || kHolds[i] < merged_bucket_top(b))) ++i;
    map |= (uint64_t)(i < kMergedClasses ? kExactClasses + i : 15) << (4 * b);
  }
  return map;
}

const std::vector<Sw16Variant>& sw16_variants() {
  static const std::vector<Sw16Variant> v = [] {
    std::vector<Sw16Variant> r;
    for (int c = 0; c < kExactClasses; ++c) r.push_back({c, kWidths[c], 1, 256});
#define DMR_SW16_LIST(c, w, p, t) r.push_back({c, w, p, t});
    DMR_SW16_EXACT_SPLIT(DMR_SW16_LIST)
    DMR_SW16_MERGED(DMR_SW16_LIST)
#undef DMR_SW16_LIST
    return r;
  }();
  return v;
}

bool sw16_has_variant(int cls, int lanes, int threads) {
  for (const Sw16Variant& v : sw16_variants())
    if (v.cls == cls && v.lanes == lanes && v.threads == threads) return true;
  return false;
}

// Band_for width classes: k_sw16 (one lane per DP pair, 256 threads) or a
// split kernel of DMR_SW16_EXACT_SPLIT.
template <class Tr>
bool launch_exact(int cls, int P, int T, const uint2* lanes, uint32_t nlanes, uint32_t nrow, const SwJob* jobs,
                  SeqSet q, SeqSet db,
                  const int* pair_tab, const int8_t* m32, uint32_t* counter, int32_t* scores, uint32_t* redo,
                  uint32_t* nredo, int cu_count, hipStream_t st) {
#define DMR_SW16_SPLIT_CASE(c, w, p, t)                                                                       \
  if (cls == c && P == p && T == t) {                                                                          \
    launch_split<w, p, t, Tr, false>(lanes, nlanes, nrow, jobs, q, db, pair_tab, counter, scores, redo, nredo, \
                                     cu_count, st);                                                          \
    return true;                                                                                             \
  }
  DMR_SW16_EXACT_SPLIT(DMR_SW16_SPLIT_CASE)
#undef DMR_SW16_SPLIT_CASE
  if (P != 1 || T != 256) return false;
#define DMR_SW16_FULL(c, w)                                                                                   \
  case c: {                                                                                                   \
    static const unsigned resident = resident_blocks(k_sw16<w, Tr>, 256, cu_count);                        \
    const unsigned tasks = nblk(nrow, 64) + nblk(nlanes - nrow, 64);                                          \
    const unsigned blocks = std::min<unsigned>(resident, nblk(tasks, 4));                                     \
    hipLaunchKernelGGL((k_sw16<w, Tr>), dim3(blocks), dim3(256), 0, st, lanes, nlanes, nrow, jobs, q, db,     \
                       pair_tab, m32, counter, scores, redo, nredo);                                            \
    return true;                                                                                              \
  }
  switch (cls) {
    DMR_SW16_FULL(0, 25)
    DMR_SW16_FULL(1, 33)
    DMR_SW16_FULL(2, 61)
    DMR_SW16_FULL(3, 81)
    DMR_SW16_FULL(4, 129)
    default:
      return false;
  }
#undef DMR_SW16_FULL
}

bool launch_sw16(int cls, int lanes_per_pair, int threads, bool f16, const uint2* lanes, uint32_t nlanes,
                 uint32_t nrow, const SwJob* jobs, SeqSet q, SeqSet db, const int* pair_tab, const int8_t* m32,
                 uint32_t* counter, int32_t* scores, uint32_t* redo, uint32_t* nredo, int cu_count, hipStream_t st) {
  if (!sw16_has_variant(cls, lanes_per_pair, threads)) return false;
  if (!nlanes) return true;
  if (nrow > nlanes) return false;
  if (cls >= kExactClasses)
    return launch_sw16_merged(cls, lanes_per_pair, threads, f16, lanes, nlanes, nrow, jobs, q, db, pair_tab, counter,
                              scores, redo, nredo, cu_count, st);
  return f16 ? launch_exact<F16>(cls, lanes_per_pair, threads, lanes, nlanes, nrow, jobs, q, db, pair_tab, m32,
                                 counter, scores, redo, nredo, cu_count, st)
             : launch_exact<I16>(cls, lanes_per_pair, threads, lanes, nlanes, nrow, jobs, q, db, pair_tab, m32,
                                 counter, scores, redo, nredo, cu_count, st);
}

}  // namespace dmr
