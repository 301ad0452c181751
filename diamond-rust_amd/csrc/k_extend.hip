// Ungapped X-drop with the adjacency filter, band merging and DP job setup.
//
// Hits arrive sorted by (target, subject, diagonal, spos), the order
// align_target walks them (src/search.rs:167-202). A run is all hits of one
// (target, subject, diagonal); the "last kept segment" of the reference can
// only suppress a hit on its own diagonal, so runs are independent and each
// thread scans one run: skip a hit covered by the last kept segment
// (same diagonal, s_end >= spos), else X-drop it (src/ungapped.rs:38-88) and
// keep the segment if it scores >= 19.
//
// Bands (src/search.rs:208-230, src/gapped.rs:27-32): per (target, subject)
// group, segments in diagonal order widen to [diag-band, diag+band] and merge
// while they touch; each band remembers its first strictly best segment;
// bands are ordered by that score, ties in diagonal order (a stable sort,
// which is what the reference's sort_unstable does for <= 20 bands). Each
// band becomes one DP job with the window clipping of src/search.rs:298-330.
#include "dmr_kernels.hpp"

namespace dmr {

namespace {

__device__ __forceinline__ void load_matrix(const int8_t* __restrict__ m32, int8_t* lds) {
  for (int i = threadIdx.x; i < kScoreStride * kScoreStride; i += blockDim.x) lds[i] = m32[i];
  __syncthreads();
}

// ungapped::xdrop (src/ungapped.rs:38-88), one lane per extension: left
// from (qa-1, sa-1), then right from (qa, sa) starting at the left-hand best
// score; each direction stops once best - running >= xdrop. Steps go in
// batches of 16 whose 32 letter loads are issued together (one memory
// latency per batch instead of one per step); the last steps before a
// sequence end go one at a time, so no load leaves the two sequences.
constexpr int kXBatch = 16;

// One direction: steps k = 0 .. limit-1 read q[dir*k], s[dir*k] (q and s
// point at step 0). (score, st) carry the best and running scores in and
// out; `at` becomes 1 + the step where the final best was first reached
// (unchanged if no step improves it).
template <int kDir>
__device__ __forceinline__ void xdrop_dir(const int8_t* M, const uint8_t* __restrict__ q,
                                          const uint8_t* __restrict__ s, int64_t limit, int xdrop,
                                          int& score, int& st, int64_t& at) {
  if (score - st >= xdrop) return;  // the loop test comes before every step
  int64_t k = 0;
  for (; k + kXBatch <= limit; k += kXBatch) {
    uint32_t a[kXBatch], b[kXBatch];
#pragma unroll
    for (int j = 0; j < kXBatch; ++j) {
      a[j] = q[kDir * (k + j)];
      b[j] = s[kDir * (k + j)];
    }
#pragma unroll
    for (int j = 0; j < kXBatch; ++j) {
      st += M[a[j] * kScoreStride + b[j]];
      if (st > score) {
        score = st;
        at = k + j + 1;
      }
      if (score - st >= xdrop) return;
    }
  }
  for (; k < limit; ++k) {
    st += M[q[kDir * k] * kScoreStride + s[kDir * k]];
    if (st > score) {
      score = st;
      at = k + 1;
    }
    if (score - st >= xdrop) return;
  }
}

__device__ __forceinline__ SegRec xdrop_dev(const int8_t* M, const uint8_t* __restrict__ qs, int64_t qlen,
                                            const uint8_t* __restrict__ ss, int64_t slen, int64_t qa, int64_t sa,
                                            int xdrop) {
  int score = 0, st = 0;
  int64_t delta = 0, len = 0;
  xdrop_dir<-1>(M, qs + qa - 1, ss + sa - 1, min(qa, sa), xdrop, score, st, delta);
  st = score;
  xdrop_dir<1>(M, qs + qa, ss + sa, min(qlen - qa, slen - sa), xdrop, score, st, len);
  SegRec sg;
  sg.q_start = (uint32_t)(qa - delta);
  sg.s_start = (uint32_t)(sa - delta);
  sg.len = (uint32_t)(len + delta);
  sg.score = score;
  return sg;
}

__global__ void k_xdrop_pairs(const uint8_t* __restrict__ q, uint64_t qlen, const uint8_t* __restrict__ s,
                              uint64_t slen, const uint32_t* __restrict__ qa,
                              const uint32_t* __restrict__ sa, uint64_t n,
                              const int8_t* __restrict__ m32, int xdrop, SegRec* __restrict__ out) {
  __shared__ int8_t M[kScoreStride * kScoreStride];
  load_matrix(m32, M);
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = xdrop_dev(M, q, (int64_t)qlen, s, (int64_t)slen, qa[i], sa[i], xdrop);
}

__global__ void k_gather(const SwJob* __restrict__ jobs, const int32_t* __restrict__ scores,
                         const uint32_t* __restrict__ idx, uint64_t n, SwJob* __restrict__ jout,
                         int32_t* __restrict__ sout) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  jout[i] = jobs[idx[i]];
  sout[i] = scores[idx[i]];
}

__global__ void k_gather_u64(const uint64_t* __restrict__ in, const uint32_t* __restrict__ idx,
                             uint64_t n, uint64_t* __restrict__ out) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[idx[i]];
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint64_t n, uint32_t v) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Device counts: a kernel given n_dev runs over its host count as a bound
// (the grid) and takes *n_dev as the count; its flags and counts past the
// count are 0, so a scan over the bound totals the same.
__device__ __forceinline__ uint64_t dcount(uint64_t bound, const uint64_t* n_dev) {
  return n_dev ? *n_dev : bound;
}

__global__ void k_key_flags(const uint64_t* __restrict__ keys, uint64_t n, int shift,
                            uint32_t* __restrict__ flags, const uint64_t* __restrict__ n_dev) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  flags[i] = i < dcount(n, n_dev) && (i == 0 || (keys[i] >> shift) != (keys[i - 1] >> shift)) ? 1u : 0u;
}

// dst[k] = src[k] for k < n (counts kept past the buffers they came from).
__global__ void k_copy_u64(const uint64_t* __restrict__ src, uint64_t* __restrict__ dst, uint32_t n) {
  if (threadIdx.x < n) dst[threadIdx.x] = src[threadIdx.x];
}

// Segments of the sorted hits: one lane per (target, subject, diagonal) run,
// walking the run's hits in order as one thread of the reference does.
__global__ void k_xdrop_runs(const uint64_t* __restrict__ hits, const uint32_t* __restrict__ run_start,
                             uint64_t nruns, uint64_t nhits, HitGeom g, SeqSet q, SeqSet db,
                             const int8_t* __restrict__ m32, int xdrop, int cutoff,
                             SegRec* __restrict__ segs, uint32_t* __restrict__ keep,
                             const uint64_t* __restrict__ nruns_dev) {
  __shared__ int8_t M[kScoreStride * kScoreStride];
  nruns = dcount(nruns, nruns_dev);
  if ((uint64_t)blockIdx.x * blockDim.x >= nruns) return;  // the whole block lies past the runs
  load_matrix(m32, M);
  uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nruns) return;
  uint64_t h0 = run_start[r], h1 = r + 1 < nruns ? run_start[r + 1] : nhits;
  uint64_t k0 = hits[h0];
  const uint64_t pmask = (1ull << g.bp) - 1, dmask = (1ull << g.bd) - 1, smask = (1ull << g.bs) - 1;
  int64_t diag = (int64_t)((k0 >> g.bp) & dmask) - g.diag_bias;
  uint32_t sj = (uint32_t)((k0 >> (g.bp + g.bd)) & smask);
  uint32_t tg = (uint32_t)(k0 >> (g.bp + g.bd + g.bs));
  const uint8_t* qs = q.res + q.offs[tg];
  const int64_t qlen = q.offs[tg + 1] - q.offs[tg];
  const uint8_t* ss = db.res + db.offs[sj];
  const int64_t slen = db.offs[sj + 1] - db.offs[sj];
  bool has_last = false;
  int64_t last_send = 0;
  for (uint64_t h = h0; h < h1; ++h) {
    int64_t spos = (int64_t)(hits[h] & pmask);
    if (has_last && last_send >= spos) {
      keep[h] = 0;
      continue;
    }
    SegRec sg = xdrop_dev(M, qs, qlen, ss, slen, diag + spos, spos, xdrop);
    if (sg.score >= cutoff) {
      segs[h] = sg;
      keep[h] = 1;
      has_last = true;
      last_send = (int64_t)sg.s_start + sg.len;
    } else {
      keep[h] = 0;
    }
  }
}

__global__ void k_compact_segments(const SegRec* __restrict__ segs, const uint64_t* __restrict__ hits,
                                   const uint32_t* __restrict__ keep, const uint64_t* __restrict__ pos,
                                   uint64_t nhits, int group_shift, SegRec* __restrict__ out,
                                   uint64_t* __restrict__ out_gkey) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nhits || !keep[i]) return;
  uint64_t p = pos[i];
  out[p] = segs[i];
  out_gkey[p] = hits[i] >> group_shift;
}

__device__ __forceinline__ uint32_t band_for_dev(uint64_t len) {
  uint32_t b = (len >= 50) + (len >= 100) + (len >= 150) + (len >= 200) + (len >= 250) +
               (len >= 350) + (len >= 500);
  const uint32_t t[8] = {12, 16, 30, 30, 30, 40, 64, 64};
  return t[b];
}

__global__ void k_bands(const SegRec* __restrict__ segs, const uint64_t* __restrict__ gkey,
                        const uint32_t* __restrict__ group_start, uint64_t ngroups, uint64_t nsegs,
                        uint32_t bs, SeqSet q, SeqSet db, int32_t* __restrict__ bd0,
                        int32_t* __restrict__ bd1, uint32_t* __restrict__ bbest,
                        SwJob* __restrict__ tmp_jobs, uint32_t* __restrict__ nbands,
                        const uint64_t* __restrict__ ngroups_dev, const uint64_t* __restrict__ nsegs_dev) {
  uint64_t gi = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gi >= ngroups) return;
  ngroups = dcount(ngroups, ngroups_dev);
  nsegs = dcount(nsegs, nsegs_dev);
  if (gi >= ngroups) {
    nbands[gi] = 0;
    return;
  }
  uint64_t s0 = group_start[gi], s1 = gi + 1 < ngroups ? group_start[gi + 1] : nsegs;
  uint64_t key = gkey[s0];
  uint32_t sj = (uint32_t)(key & ((1ull << bs) - 1));
  uint32_t tg = (uint32_t)(key >> bs);
  int64_t qlen = q.offs[tg + 1] - q.offs[tg];
  int64_t band = band_for_dev((uint64_t)qlen);
  uint32_t nb = 0;
  for (uint64_t k = s0; k < s1; ++k) {
    SegRec sg = segs[k];
    int64_t diag = (int64_t)sg.q_start - (int64_t)sg.s_start;
    int64_t b0 = diag - band, b1 = diag + band;
    if (nb > 0 && b0 <= bd1[s0 + nb - 1]) {
      uint64_t l = s0 + nb - 1;
      if (b1 > bd1[l]) bd1[l] = (int32_t)b1;
      if (sg.score > segs[bbest[l]].score) bbest[l] = (uint32_t)k;
    } else {
      bd0[s0 + nb] = (int32_t)b0;
      bd1[s0 + nb] = (int32_t)b1;
      bbest[s0 + nb] = (uint32_t)k;
      ++nb;
    }
  }
  // Stable insertion sort by best segment score, descending.
  for (uint32_t a = 1; a < nb; ++a) {
    int32_t x0 = bd0[s0 + a], x1 = bd1[s0 + a];
    uint32_t xb = bbest[s0 + a];
    int32_t xs = segs[xb].score;
    uint32_t b = a;
    while (b > 0 && segs[bbest[s0 + b - 1]].score < xs) {
      bd0[s0 + b] = bd0[s0 + b - 1];
      bd1[s0 + b] = bd1[s0 + b - 1];
      bbest[s0 + b] = bbest[s0 + b - 1];
      --b;
    }
    bd0[s0 + b] = x0;
    bd1[s0 + b] = x1;
    bbest[s0 + b] = xb;
  }
  const int64_t budget = 16000000;
  for (uint32_t r = 0; r < nb; ++r) {
    int64_t d0 = bd0[s0 + r], d1 = bd1[s0 + r];
    int64_t row_cells = (d1 - d0 > 0 ? d1 - d0 : 0) + 1;
    int64_t max_rows = budget / row_cells;
    if (max_rows < 2 * band + 16) max_rows = 2 * band + 16;
    int64_t q0 = 0, q1 = qlen;
    if (q1 - q0 > max_rows) {
      int64_t anchor = segs[bbest[s0 + r]].q_start;
      q0 = anchor > max_rows / 2 ? anchor - max_rows / 2 : 0;
      q1 = q0 + max_rows < qlen ? q0 + max_rows : qlen;
      q0 = q1 > max_rows ? q1 - max_rows : 0;
    }
    SwJob j;
    j.target = tg;
    j.subject = sj;
    j.q0 = (uint32_t)q0;
    j.q1 = (uint32_t)q1;
    j.d0 = (int32_t)(d0 - q0);
    j.d1 = (int32_t)(d1 - q0);
    j.group = (uint32_t)gi;
    j.rank = r;
    tmp_jobs[s0 + r] = j;
  }
  nbands[gi] = nb;
}

__global__ void k_jobs_compact(const SwJob* __restrict__ tmp_jobs, const uint32_t* __restrict__ group_start,
                               const uint32_t* __restrict__ nbands, const uint64_t* __restrict__ job_off,
                               uint64_t ngroups, SwJob* __restrict__ jobs, const uint64_t* __restrict__ ngroups_dev) {
  uint64_t gi = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gi >= dcount(ngroups, ngroups_dev)) return;
  uint64_t s0 = group_start[gi], o = job_off[gi];
  for (uint32_t r = 0; r < nbands[gi]; ++r) jobs[o + r] = tmp_jobs[s0 + r];
}

__global__ void k_evalue_filter(const SwJob* __restrict__ jobs, const int32_t* __restrict__ scores,
                                uint64_t njobs, SeqSet q, SeqSet db, AlpParams P, uint64_t db_letters,
                                double threshold, uint32_t* __restrict__ flags) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= njobs) return;
  int32_t s = scores[i];
  uint32_t f = 0;
  if (s > 0) {
    const SwJob j = jobs[i];
    uint64_t qlen = q.offs[j.target + 1] - q.offs[j.target];
    uint64_t slen = db.offs[j.subject + 1] - db.offs[j.subject];
    double e = alp_evalue(P, s, qlen, slen, db_letters);
    // Conservative: only bands whose E-value is far above the threshold are
    // dropped here; the host decides the rest with the exact libm values.
    f = (e > 16.0 * threshold) ? 0u : 1u;
  }
  flags[i] = f;
}

__global__ void k_u32_times(const uint32_t* __restrict__ in, uint64_t n, uint32_t mul,
                            uint64_t* __restrict__ out) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (uint64_t)in[i] * mul;
}

// sums[k] += v[b_k .. e_k) for the ranges of ranges[2k], ranges[2k + 1]:
// grid (blocks over the longest range, k); each block adds its part.
__global__ void k_range_sums(const uint64_t* __restrict__ v, const uint64_t* __restrict__ ranges,
                             unsigned long long* __restrict__ sums) {
  __shared__ unsigned long long part[4];
  const uint32_t k = blockIdx.y;
  const uint64_t b = ranges[2 * k], e = ranges[2 * k + 1];
  unsigned long long x = 0;
  for (uint64_t i = b + (uint64_t)blockIdx.x * 1024 + threadIdx.x; i < e && i < b + (uint64_t)(blockIdx.x + 1) * 1024;
       i += 256)
    x += v[i];
  for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned long long t = part[0] + part[1] + part[2] + part[3];
    if (t) atomicAdd(&sums[k], t);
  }
}

inline unsigned nblk(uint64_t n, unsigned t = 256) { return (unsigned)((n + t - 1) / t); }

}  // namespace

void launch_range_sums(const uint64_t* v, const uint64_t* ranges, uint32_t k, uint64_t longest, uint64_t* sums,
                       hipStream_t st) {
  if (!k || !longest) return;
  hipLaunchKernelGGL(k_range_sums, dim3(nblk(longest, 1024), k), dim3(256), 0, st, v, ranges,
                     reinterpret_cast<unsigned long long*>(sums));
}

void launch_fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t st) {
  if (!n) return;
  hipLaunchKernelGGL(k_fill_u32, dim3(nblk(n)), dim3(256), 0, st, p, n, v);
}

void launch_gather_u64(const uint64_t* in, const uint32_t* idx, uint64_t n, uint64_t* out,
                       hipStream_t st) {
  if (!n) return;
  hipLaunchKernelGGL(k_gather_u64, dim3(nblk(n)), dim3(256), 0, st, in, idx, n, out);
}

void launch_gather(const SwJob* jobs, const int32_t* scores, const uint32_t* idx, uint64_t n,
                   SwJob* jout, int32_t* sout, hipStream_t st) {
  if (!n) return;
  hipLaunchKernelGGL(k_gather, dim3(nblk(n)), dim3(256), 0, st, jobs, scores, idx, n, jout, sout);
}

void launch_xdrop_pairs(const uint8_t* q, uint64_t qlen, const uint8_t* s, uint64_t slen,
                        const uint32_t* qa, const uint32_t* sa, uint64_t n, const int8_t* m32,
                        int xdrop, SegRec* out, hipStream_t st) {
  if (!n) return;
  hipLaunchKernelGGL(k_xdrop_pairs, dim3(nblk(n, 128)), dim3(128), 0, st, q, qlen, s, slen, qa, sa, n,
                     m32, xdrop, out);
}

void launch_key_flags(const uint64_t* keys, uint64_t n, int shift, uint32_t* flags, hipStream_t st,
                      const uint64_t* n_dev) {
  if (!n) return;
  hipLaunchKernelGGL(k_key_flags, dim3(nblk(n)), dim3(256), 0, st, keys, n, shift, flags, n_dev);
}

void launch_copy_u64(const uint64_t* src, uint64_t* dst, uint32_t n, hipStream_t st) {
  if (!n) return;
  hipLaunchKernelGGL(k_copy_u64, dim3(1), dim3(64), 0, st, src, dst, n);
}

void launch_xdrop_runs(const uint64_t* hits, const uint32_t* run_start, uint64_t nruns,
                       uint64_t nhits, HitGeom g, SeqSet q, SeqSet db, const int8_t* matrix32,
                       int xdrop, int cutoff, SegRec* segs, uint32_t* keep, hipStream_t st,
                       const uint64_t* nruns_dev) {
  if (!nruns) return;
  hipLaunchKernelGGL(k_xdrop_runs, dim3(nblk(nruns, 128)), dim3(128), 0, st, hits, run_start, nruns,
                     nhits, g, q, db, matrix32, xdrop, cutoff, segs, keep, nruns_dev);
}

void launch_compact_segments(const SegRec* segs, const uint64_t* hits, const uint32_t* keep,
                             const uint64_t* pos, uint64_t nhits, int group_shift, SegRec* out,
                             uint64_t* out_gkey, hipStream_t st) {
  if (!nhits) return;
  hipLaunchKernelGGL(k_compact_segments, dim3(nblk(nhits)), dim3(256), 0, st, segs, hits, keep, pos,
                     nhits, group_shift, out, out_gkey);
}

void launch_bands(const SegRec* segs, const uint64_t* gkey, const uint32_t* group_start,
                  uint64_t ngroups, uint64_t nsegs, uint32_t bs, SeqSet q, SeqSet db,
                  int32_t* tmp_d0, int32_t* tmp_d1, uint32_t* tmp_best, SwJob* tmp_jobs,
                  uint32_t* nbands, hipStream_t st, const uint64_t* ngroups_dev, const uint64_t* nsegs_dev) {
  if (!ngroups) return;
  hipLaunchKernelGGL(k_bands, dim3(nblk(ngroups, 128)), dim3(128), 0, st, segs, gkey, group_start,
                     ngroups, nsegs, bs, q, db, tmp_d0, tmp_d1, tmp_best, tmp_jobs, nbands, ngroups_dev,
                     nsegs_dev);
}

void launch_jobs_compact(const SwJob* tmp_jobs, const uint32_t* group_start, const uint32_t* nbands,
                         const uint64_t* job_off, uint64_t ngroups, SwJob* jobs, hipStream_t st,
                         const uint64_t* ngroups_dev) {
  if (!ngroups) return;
  hipLaunchKernelGGL(k_jobs_compact, dim3(nblk(ngroups)), dim3(256), 0, st, tmp_jobs, group_start,
                     nbands, job_off, ngroups, jobs, ngroups_dev);
}

void launch_evalue_filter(const SwJob* jobs, const int32_t* scores, uint64_t njobs, SeqSet q,
                          SeqSet db, AlpParams P, uint64_t db_letters, double threshold,
                          uint32_t* flags, hipStream_t st) {
  if (!njobs) return;
  hipLaunchKernelGGL(k_evalue_filter, dim3(nblk(njobs)), dim3(256), 0, st, jobs, scores, njobs, q, db,
                     P, db_letters, threshold, flags);
}

void launch_u32_times(const uint32_t* in, uint64_t n, uint32_t mul, uint64_t* out, hipStream_t st) {
  if (!n) return;
  hipLaunchKernelGGL(k_u32_times, dim3(nblk(n)), dim3(256), 0, st, in, n, mul, out);
}

}  // namespace dmr
