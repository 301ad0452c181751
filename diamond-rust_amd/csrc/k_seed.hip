// Spaced-seed enumeration, reference-index bucket tables and the seed join.
//
// Enumeration restates Shape::packed + seed_is_complex (src/shape.rs:60-74,
// src/seed.rs:46-70): a seed at (sequence, pos) exists when the shape fits in
// the sequence, every sampled letter is canonical (< 20), and the log
// multinomial of the bucket counts reaches the cut. Keys are base-`size`
// numbers with the first sampled letter most significant; the shape index is
// prepended above the key bits so all shapes sort in one pass.
//
// The join restates seed::join_into (src/seed.rs:110-143): for every equal
// key, nq query entries (whole batch) and nr reference entries (one DB
// chunk) expand to nq*nr hits unless nq*nr > max_expand.
#include "dmr_device.hpp"
#include "dmr_kernels.hpp"

namespace dmr {

namespace {

constexpr int kSeedThreads = 256;

// Evaluates the seed at residue `loc` for `shape`; returns true and the
// tagged key if it is emitted.
__device__ __forceinline__ bool seed_at(const uint8_t* __restrict__ res,
                                        const uint32_t* __restrict__ seq_of,
                                        const uint32_t* __restrict__ offs, uint64_t loc,
                                        const SeedParams& P, const ShapeDev& sh, uint32_t s,
                                        int key_bits, uint64_t* key_out) {
  uint32_t sid = seq_of[loc];
  uint64_t start = offs[sid], end = offs[sid + 1];
  if (loc + sh.span > end) return false;
  (void)start;
  uint64_t key = 0;
  // Bucket counts packed as 5-bit fields, 12 buckets per word (counts <= 16).
  uint64_t c_lo = 0, c_hi = 0;
  for (uint32_t k = 0; k < sh.weight; ++k) {
    uint8_t l = res[loc + sh.pos[k]];
    if (l >= 20) return false;
    uint32_t b = P.red_map[l];
    if (b >= 32) return false;
    key = key * P.red_size + b;
    if (b < 12) c_lo += 1ull << (5 * b);
    else c_hi += 1ull << (5 * (b - 12));
  }
  if (P.cut > 0.0) {
    double e = P.lnfact[sh.weight];
    for (uint32_t g = 0; g < P.red_size && g < 24; ++g) {
      uint32_t c = g < 12 ? (uint32_t)((c_lo >> (5 * g)) & 31u) : (uint32_t)((c_hi >> (5 * (g - 12))) & 31u);
      if (c > 0) e -= P.lnfact[c];
    }
    if (!(e >= P.cut)) return false;
  }
  *key_out = ((uint64_t)s << key_bits) | key;
  return true;
}

// grid: (blocks over [r0, r1), nshapes)
__global__ void k_seed_count(const uint8_t* __restrict__ res, const uint32_t* __restrict__ seq_of,
                             const uint32_t* __restrict__ offs, uint64_t r0, uint64_t r1,
                             const SeedParams* __restrict__ Pp, int key_bits,
                             uint32_t* __restrict__ block_counts) {
  __shared__ uint32_t cnt;
  const SeedParams& P = *Pp;
  const uint32_t s = blockIdx.y;
  const ShapeDev sh = P.shapes[s];
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  uint64_t loc = r0 + (uint64_t)blockIdx.x * kSeedThreads + threadIdx.x;
  uint64_t key;
  bool ok = loc < r1 && seed_at(res, seq_of, offs, loc, P, sh, s, key_bits, &key);
  uint64_t m = __ballot(ok);
  if ((threadIdx.x & 63) == 0 && m) atomicAdd(&cnt, (uint32_t)__popcll(m));
  __syncthreads();
  if (threadIdx.x == 0) block_counts[(uint64_t)s * gridDim.x + blockIdx.x] = cnt;
}

__global__ void k_seed_emit(const uint8_t* __restrict__ res, const uint32_t* __restrict__ seq_of,
                            const uint32_t* __restrict__ offs, uint64_t r0, uint64_t r1,
                            const SeedParams* __restrict__ Pp, int key_bits,
                            const uint64_t* __restrict__ block_off, uint64_t* __restrict__ keys,
                            uint32_t* __restrict__ locs) {
  __shared__ uint32_t wave_cnt[kSeedThreads / 64];
  const SeedParams& P = *Pp;
  const uint32_t s = blockIdx.y;
  const ShapeDev sh = P.shapes[s];
  uint64_t loc = r0 + (uint64_t)blockIdx.x * kSeedThreads + threadIdx.x;
  uint64_t key = 0;
  bool ok = loc < r1 && seed_at(res, seq_of, offs, loc, P, sh, s, key_bits, &key);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint64_t m = __ballot(ok);
  if (lane == 0) wave_cnt[wave] = (uint32_t)__popcll(m);
  __syncthreads();
  if (ok) {
    uint64_t pos = block_off[(uint64_t)s * gridDim.x + blockIdx.x];
    for (int w = 0; w < wave; ++w) pos += wave_cnt[w];
    uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    pos += (uint64_t)__popcll(m & lt);
    keys[pos] = key;
    locs[pos] = (uint32_t)loc;
  }
}

// table[b] = first index whose key >= b << shift (lower bound per bucket).
__global__ void k_bucket_table(const uint64_t* __restrict__ keys, uint64_t n, int shift,
                               uint64_t nbuckets, uint64_t* __restrict__ table) {
  uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b > nbuckets) return;
  uint64_t target = b << shift;
  uint64_t lo = 0, hi = n;
  while (lo < hi) {
    uint64_t mid = (lo + hi) >> 1;
    if (keys[mid] < target) lo = mid + 1;
    else hi = mid;
  }
  table[b] = lo;
}

// Device counts: a kernel given n_dev runs over n entries (a bound, the
// grid) and takes *n_dev as the count; flags and counts past it are 0, so a
// scan over the bound totals the same.
__device__ __forceinline__ uint64_t dcount(uint64_t bound, const uint64_t* n_dev) {
  return n_dev ? *n_dev : bound;
}

__global__ void k_run_flags(const uint64_t* __restrict__ keys, uint64_t n, uint32_t* __restrict__ flags,
                            const uint64_t* __restrict__ n_dev) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  flags[i] = i < dcount(n, n_dev) && (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
}

// Runs of equal keys within one record: a run also ends where the record of
// its entries' target changes (entries of a key are in letter order, and a
// record's targets are contiguous), so each record's seeds meet the cap
// alone, as when the record is searched by itself.
__global__ void k_run_flags_grouped(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ locs,
                                    const uint32_t* __restrict__ seq_of, const uint32_t* __restrict__ group,
                                    uint64_t n, uint32_t* __restrict__ flags, const uint64_t* __restrict__ n_dev) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (i >= dcount(n, n_dev)) {
    flags[i] = 0;
    return;
  }
  flags[i] = (i == 0 || keys[i] != keys[i - 1] || group[seq_of[locs[i]]] != group[seq_of[locs[i - 1]]]) ? 1u
                                                                                                      : 0u;
}

__global__ void k_compact_flags(const uint32_t* __restrict__ flags, const uint64_t* __restrict__ pos,
                                uint64_t n, uint32_t* __restrict__ out) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (flags[i]) out[pos[i]] = (uint32_t)i;
}

// First index of a chunk's sorted keys that is >= key (upper: > key),
// through the bucket table: keys past the last bucket are past every key.
__device__ __forceinline__ uint64_t ref_bound(const RefChunkDev& ch, uint64_t key, bool upper) {
  const uint64_t b = key >> ch.shift;
  if (ch.n == 0 || b >= ch.nbuckets) return ch.n;
  uint64_t a = ch.table[b], z = ch.table[b + 1];
  while (a < z) {
    const uint64_t mid = (a + z) >> 1;
    if (upper ? ch.keys[mid] <= key : ch.keys[mid] < key) a = mid + 1;
    else z = mid;
  }
  return a;
}

// Per (query-key run, chunk): the run's reference entries (a range of the
// chunk's sorted keys) and nq * nr, 0 past the cap. One block takes 256
// consecutive runs, whose keys are ascending. For up to kJoinGroup chunks at
// a time, one thread per chunk and bound finds the reference keys between
// the block's first and last run's key (bucket table plus binary search),
// the block stages them in LDS, and each thread finds its key's range in
// each staged chunk, so the reference keys are read once, coalesced, and the
// dependent searches of a block's chunks run side by side. A chunk whose
// range holds more keys than its share of the stage is searched per thread.
constexpr int kJoinRuns = 256;
constexpr int kJoinStage = 4096;  // reference keys staged per block
constexpr int kJoinGroup = 8;     // chunks staged together
__global__ __launch_bounds__(kJoinRuns) void k_join_count(
    const uint64_t* __restrict__ qkeys, const uint32_t* __restrict__ run_start, uint64_t nruns, uint64_t nq_total,
    const RefChunkDev* __restrict__ chunks, uint32_t nchunks, uint64_t max_expand, uint32_t* __restrict__ counts,
    uint64_t* __restrict__ rlo, const uint64_t* __restrict__ nruns_dev, const uint64_t* __restrict__ nq_dev) {
  __shared__ uint64_t stage[kJoinStage];
  __shared__ uint64_t range[2 * kJoinGroup];
  const uint64_t bound = nruns;
  nruns = dcount(nruns, nruns_dev);
  nq_total = dcount(nq_total, nq_dev);
  const uint64_t r0 = (uint64_t)blockIdx.x * kJoinRuns, r = r0 + threadIdx.x;
  if (r0 >= nruns) {  // past the runs: counts 0 (the scan runs over the bound)
    if (r < bound)
      for (uint32_t c = 0; c < nchunks; ++c) {
        counts[r * nchunks + c] = 0;
        rlo[r * nchunks + c] = 0;
      }
    return;
  }
  const bool live = r < nruns;
  const uint64_t rl = min(r0 + kJoinRuns, nruns) - 1;  // the block's last run
  uint64_t s0 = 0, nq = 0, key = 0;
  if (live) {
    s0 = run_start[r];
    nq = ((r + 1 < nruns) ? run_start[r + 1] : nq_total) - s0;
    key = qkeys[s0];
  }
  for (uint32_t c0 = 0; c0 < nchunks; c0 += kJoinGroup) {
    const uint32_t g = min(nchunks - c0, (uint32_t)kJoinGroup);
    const uint32_t share = kJoinStage / g;  // staged keys per chunk
    if (threadIdx.x < 2 * g) {
      const uint32_t c = c0 + threadIdx.x / 2;
      const bool upper = threadIdx.x & 1;
      range[threadIdx.x] = ref_bound(chunks[c], qkeys[run_start[upper ? rl : r0]], upper);
    }
    __syncthreads();
    for (uint32_t k = 0; k < g; ++k) {
      const uint64_t lo0 = range[2 * k], hi0 = range[2 * k + 1];
      const uint64_t ns = hi0 > lo0 ? hi0 - lo0 : 0;
      if (ns <= share) {
        const uint64_t* src = chunks[c0 + k].keys + lo0;
        for (uint32_t e = threadIdx.x; e < ns; e += kJoinRuns) stage[k * share + e] = src[e];
      }
    }
    __syncthreads();
    for (uint32_t k = 0; k < g; ++k) {
      const uint32_t c = c0 + k;
      if (live) {
        const uint64_t lo0 = range[2 * k], hi0 = range[2 * k + 1];
        const uint64_t ns = hi0 > lo0 ? hi0 - lo0 : 0;
        uint64_t lo, hi;
        if (ns <= share) {  // lower and upper bound of key among the staged keys
          const uint64_t* st = stage + k * share;
          uint32_t a = 0, z = (uint32_t)ns;
          while (a < z) {
            const uint32_t mid = (a + z) >> 1;
            if (st[mid] < key) a = mid + 1;
            else z = mid;
          }
          lo = lo0 + a;
          z = (uint32_t)ns;
          while (a < z) {
            const uint32_t mid = (a + z) >> 1;
            if (st[mid] <= key) a = mid + 1;
            else z = mid;
          }
          hi = lo0 + a;
        } else {
          const RefChunkDev ch = chunks[c];
          lo = ref_bound(ch, key, false);
          hi = ref_bound(ch, key, true);
        }
        const uint64_t nr = hi - lo, cnt = nq * nr;
        counts[r * nchunks + c] = (nr > 0 && cnt <= max_expand) ? (uint32_t)cnt : 0u;
        rlo[r * nchunks + c] = lo;
      } else if (r < bound) {
        counts[r * nchunks + c] = 0;
        rlo[r * nchunks + c] = 0;
      }
    }
    __syncthreads();  // the stage and the ranges are reused by the next group
  }
}

// Hits per target: every entry of a kept (run, chunk) pair contributes its
// nr reference entries to its target.
__global__ void k_target_hits(const uint32_t* __restrict__ qlocs, const uint32_t* __restrict__ run_start,
                              uint64_t nruns, uint64_t nq_total, uint32_t nchunks,
                              const uint32_t* __restrict__ counts, const uint32_t* __restrict__ q_seq_of,
                              unsigned long long* __restrict__ thits, const uint64_t* __restrict__ nruns_dev,
                              const uint64_t* __restrict__ nq_dev) {
  uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nruns * nchunks) return;
  nruns = dcount(nruns, nruns_dev);
  nq_total = dcount(nq_total, nq_dev);
  const uint32_t cnt = counts[t];
  if (cnt == 0) return;
  const uint64_t r = t / nchunks;
  const uint64_t s0 = run_start[r], s1 = (r + 1 < nruns) ? run_start[r + 1] : nq_total;
  const uint32_t nr = cnt / (uint32_t)(s1 - s0);
  for (uint64_t a = s0; a < s1; ++a) atomicAdd(&thits[q_seq_of[qlocs[a]]], (unsigned long long)nr);
}

// The part of each kept (run, chunk) pair whose entries' targets lie in
// [t0, t1): its first entry, entry count and hit count (entries of a run are
// in letter order, so the part is one range, found by binary search).
__global__ void k_join_slice(const uint32_t* __restrict__ qlocs, const uint32_t* __restrict__ run_start,
                             uint64_t nruns, uint64_t nq_total, uint32_t nchunks,
                             const uint32_t* __restrict__ counts, const uint32_t* __restrict__ q_seq_of,
                             uint32_t t0, uint32_t t1, uint32_t* __restrict__ slice_counts,
                             uint32_t* __restrict__ slice_start, uint32_t* __restrict__ slice_nq,
                             const uint64_t* __restrict__ nruns_dev, const uint64_t* __restrict__ nq_dev) {
  uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nruns * nchunks) return;
  nruns = dcount(nruns, nruns_dev);
  nq_total = dcount(nq_total, nq_dev);
  const uint32_t cnt = counts[t];  // 0 past the runs
  const uint64_t r = t / nchunks;
  const uint64_t s0 = run_start[r], s1 = (r + 1 < nruns) ? run_start[r + 1] : nq_total;
  uint32_t out = 0, e0 = 0, ne = 0;
  if (cnt) {
    const uint32_t nr = cnt / (uint32_t)(s1 - s0);
    auto first_at_least = [&](uint32_t tg) {
      uint64_t lo = s0, hi = s1;
      while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (q_seq_of[qlocs[mid]] < tg) lo = mid + 1;
        else hi = mid;
      }
      return lo;
    };
    const uint64_t a = first_at_least(t0), b = first_at_least(t1);
    e0 = (uint32_t)a;
    ne = (uint32_t)(b - a);
    out = ne * nr;
  }
  slice_counts[t] = out;
  slice_start[t] = e0;
  slice_nq[t] = ne;
}

// Join emission, one thread per hit: hit o belongs to the (run, chunk) pair
// t with hit_off[t] <= o < hit_off[t + 1] (an empty pair never holds one),
// and is entry a = l / nr of the query run times entry b = l % nr of the
// reference run, l = o - hit_off[t]: the order a pair-by-pair double loop
// writes, so consecutive threads write consecutive hits. Each block finds
// the pair range of its hits by binary search over hit_off (npairs + 1
// offsets, the last the total), each thread its pair inside that range.
// With slice_start / slice_nq, a pair's entries are the slice's part of it.
__global__ __launch_bounds__(256) void k_join_emit(
    const uint32_t* __restrict__ qlocs, const uint32_t* __restrict__ run_start, uint64_t nruns,
    uint64_t nq_total, const RefChunkDev* __restrict__ chunks, uint32_t nchunks, const uint32_t* __restrict__ counts,
    const uint64_t* __restrict__ rlo, const uint64_t* __restrict__ hit_off, uint64_t npairs, uint64_t nhits,
    const uint32_t* __restrict__ q_seq_of, const uint32_t* __restrict__ q_offs,
    const uint32_t* __restrict__ db_seq_of, const uint32_t* __restrict__ db_offs, HitGeom g,
    uint64_t* __restrict__ hits, const uint32_t* __restrict__ slice_start, const uint32_t* __restrict__ slice_nq,
    const uint64_t* __restrict__ nruns_dev, const uint64_t* __restrict__ nq_dev) {
  __shared__ uint64_t range[2];
  const uint64_t o0 = (uint64_t)blockIdx.x * blockDim.x;
  // Last pair whose offset is <= x (x < nhits = hit_off[npairs]).
  auto last_at_most = [&](uint64_t x, uint64_t lo, uint64_t hi) {  // hit_off[lo] <= x < hit_off[hi]
    while (hi - lo > 1) {
      const uint64_t mid = lo + (hi - lo) / 2;
      if (hit_off[mid] <= x) lo = mid;
      else hi = mid;
    }
    return lo;
  };
  if (threadIdx.x < 2) {
    const uint64_t x = threadIdx.x == 0 ? o0 : min(o0 + blockDim.x, nhits) - 1;
    range[threadIdx.x] = last_at_most(x, 0, npairs);
  }
  __syncthreads();
  const uint64_t o = o0 + threadIdx.x;
  if (o >= nhits) return;
  nruns = dcount(nruns, nruns_dev);
  nq_total = dcount(nq_total, nq_dev);
  const uint64_t t = last_at_most(o, range[0], range[1] + 1);
  const uint64_t r = t / nchunks;
  const uint32_t c = (uint32_t)(t - r * nchunks);
  uint64_t s0 = run_start[r];
  uint32_t nq = (uint32_t)(((r + 1 < nruns) ? run_start[r + 1] : nq_total) - s0);
  if (slice_start) {  // the pair's count covers the slice's entries only
    s0 = slice_start[t];
    nq = slice_nq[t];
  }
  const uint32_t nr = counts[t] / nq;
  const uint32_t l = (uint32_t)(o - hit_off[t]);
  const uint32_t a = l / nr, b = l - a * nr;
  const uint32_t ql = qlocs[s0 + a];
  const uint32_t tg = q_seq_of[ql];
  const uint32_t qpos = ql - q_offs[tg];
  const uint32_t rloc = chunks[c].locs[rlo[t] + b];
  const uint32_t sj = db_seq_of[rloc];
  const uint32_t spos = rloc - db_offs[sj];
  const uint64_t diag = (uint64_t)((int64_t)qpos - (int64_t)spos + g.diag_bias);
  hits[o] = ((((uint64_t)tg << g.bs) | sj) << g.bd | diag) << g.bp | spos;
}

// Join emission for explicit (seq, pos) entry arrays (drop-in seed::join_into).
__global__ void k_join_emit_entries(const uint32_t* __restrict__ qs, const uint32_t* __restrict__ qp,
                                    const uint32_t* __restrict__ run_start, uint64_t nruns,
                                    uint64_t nq_total, const uint32_t* __restrict__ rs,
                                    const uint32_t* __restrict__ rp, const uint32_t* __restrict__ counts,
                                    const uint64_t* __restrict__ rlo, const uint64_t* __restrict__ hit_off,
                                    HitGeom g, uint64_t* __restrict__ hits) {
  uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nruns) return;
  uint32_t cnt = counts[r];
  if (cnt == 0) return;
  uint64_t s0 = run_start[r];
  uint64_t s1 = (r + 1 < nruns) ? run_start[r + 1] : nq_total;
  uint32_t nq = (uint32_t)(s1 - s0), nr = cnt / nq;
  uint64_t o = hit_off[r];
  for (uint32_t a = 0; a < nq; ++a) {
    uint32_t tg = qs[s0 + a], qpos = qp[s0 + a];
    for (uint32_t b = 0; b < nr; ++b) {
      uint32_t sj = rs[rlo[r] + b], spos = rp[rlo[r] + b];
      uint64_t diag = (uint64_t)((int64_t)qpos - (int64_t)spos + g.diag_bias);
      hits[o++] = ((((uint64_t)tg << g.bs) | sj) << g.bd | diag) << g.bp | spos;
    }
  }
}

}  // namespace

void launch_join_emit_entries(const uint32_t* qs, const uint32_t* qp, const uint32_t* run_start,
                              uint64_t nruns, uint64_t nq_total, const uint32_t* rs,
                              const uint32_t* rp, const uint32_t* counts, const uint64_t* rlo,
                              const uint64_t* hit_off, HitGeom g, uint64_t* hits, hipStream_t st) {
  if (!nruns) return;
  hipLaunchKernelGGL(k_join_emit_entries, dim3((unsigned)((nruns + 255) / 256)), dim3(256), 0, st, qs,
                     qp, run_start, nruns, nq_total, rs, rp, counts, rlo, hit_off, g, hits);
}

void launch_seed_count(const uint8_t* res, const uint32_t* seq_of, const uint32_t* offs, uint64_t r0,
                       uint64_t r1, const SeedParams* P, uint32_t nshapes, int key_bits, uint32_t* block_counts,
                       uint64_t* block_off, uint64_t* scan_tmp, hipStream_t st) {
  if (r1 <= r0) return;
  uint32_t nb = (uint32_t)((r1 - r0 + kSeedThreads - 1) / kSeedThreads);
  dim3 grid(nb, nshapes);
  hipLaunchKernelGGL(k_seed_count, grid, dim3(kSeedThreads), 0, st, res, seq_of, offs, r0, r1, P,
                     key_bits, block_counts);
  scan_exclusive_u32_to_u64(block_counts, block_off, (uint64_t)nb * nshapes + 1, scan_tmp, st);
}

void launch_seed_emit(const uint8_t* res, const uint32_t* seq_of, const uint32_t* offs, uint64_t r0,
                      uint64_t r1, const SeedParams* P, uint32_t nshapes, int key_bits, const uint64_t* block_off,
                      uint64_t* keys, uint32_t* locs, hipStream_t st) {
  if (r1 <= r0) return;
  uint32_t nb = (uint32_t)((r1 - r0 + kSeedThreads - 1) / kSeedThreads);
  dim3 grid(nb, nshapes);
  hipLaunchKernelGGL(k_seed_emit, grid, dim3(kSeedThreads), 0, st, res, seq_of, offs, r0, r1, P,
                     key_bits, block_off, keys, locs);
}

void launch_seed_enumerate(const uint8_t* res, const uint32_t* seq_of, const uint32_t* offs,
                           uint64_t r0, uint64_t r1, const SeedParams* P, uint32_t nshapes,
                           int key_bits, uint32_t* block_counts, uint64_t* block_off,
                           uint64_t* scan_tmp, uint64_t* keys, uint32_t* locs, hipStream_t st) {
  launch_seed_count(res, seq_of, offs, r0, r1, P, nshapes, key_bits, block_counts, block_off, scan_tmp, st);
  launch_seed_emit(res, seq_of, offs, r0, r1, P, nshapes, key_bits, block_off, keys, locs, st);
}

uint32_t seed_enum_blocks(uint64_t r0, uint64_t r1) {
  return (uint32_t)((r1 - r0 + kSeedThreads - 1) / kSeedThreads);
}

void launch_bucket_table(const uint64_t* keys, uint64_t n, int shift, uint64_t nbuckets,
                         uint64_t* table, hipStream_t st) {
  uint64_t threads = nbuckets + 1;
  hipLaunchKernelGGL(k_bucket_table, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, st, keys,
                     n, shift, nbuckets, table);
}

void launch_run_flags(const uint64_t* keys, uint64_t n, uint32_t* flags, hipStream_t st, const uint64_t* n_dev) {
  if (!n) return;
  hipLaunchKernelGGL(k_run_flags, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, keys, n, flags, n_dev);
}

void launch_compact_flags(const uint32_t* flags, const uint64_t* pos, uint64_t n, uint32_t* out,
                          hipStream_t st) {
  if (!n) return;
  hipLaunchKernelGGL(k_compact_flags, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, flags, pos,
                     n, out);
}

void launch_join_count(const uint64_t* qkeys, const uint32_t* run_start, uint64_t nruns,
                       uint64_t nq_total, const RefChunkDev* chunks, uint32_t nchunks,
                       uint64_t max_expand, uint32_t* counts, uint64_t* rlo, hipStream_t st,
                       const uint64_t* nruns_dev, const uint64_t* nq_dev) {
  if (!nruns || !nchunks) return;
  hipLaunchKernelGGL(k_join_count, dim3((unsigned)((nruns + kJoinRuns - 1) / kJoinRuns)), dim3(kJoinRuns), 0, st,
                     qkeys, run_start, nruns, nq_total, chunks, nchunks, max_expand, counts, rlo, nruns_dev, nq_dev);
}

void launch_join_emit(const uint32_t* qlocs, const uint32_t* run_start, uint64_t nruns,
                      uint64_t nq_total, const RefChunkDev* chunks, uint32_t nchunks,
                      const uint32_t* counts, const uint64_t* rlo, const uint64_t* hit_off, uint64_t nhits,
                      const uint32_t* q_seq_of, const uint32_t* q_offs, const uint32_t* db_seq_of,
                      const uint32_t* db_offs, HitGeom g, uint64_t* hits, hipStream_t st,
                      const uint32_t* slice_start, const uint32_t* slice_nq, const uint64_t* nruns_dev,
                      const uint64_t* nq_dev) {
  const uint64_t npairs = nruns * nchunks;
  if (!npairs || !nhits) return;
  hipLaunchKernelGGL(k_join_emit, dim3((unsigned)((nhits + 255) / 256)), dim3(256), 0, st, qlocs, run_start,
                     nruns, nq_total, chunks, nchunks, counts, rlo, hit_off, npairs, nhits, q_seq_of, q_offs,
                     db_seq_of, db_offs, g, hits, slice_start, slice_nq, nruns_dev, nq_dev);
}

void launch_run_flags_grouped(const uint64_t* keys, const uint32_t* locs, const uint32_t* seq_of,
                              const uint32_t* group, uint64_t n, uint32_t* flags, hipStream_t st,
                              const uint64_t* n_dev) {
  if (!n) return;
  hipLaunchKernelGGL(k_run_flags_grouped, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, keys, locs,
                     seq_of, group, n, flags, n_dev);
}

void launch_target_hits(const uint32_t* qlocs, const uint32_t* run_start, uint64_t nruns, uint64_t nq_total,
                        uint32_t nchunks, const uint32_t* counts, const uint32_t* q_seq_of, uint64_t* thits,
                        hipStream_t st, const uint64_t* nruns_dev, const uint64_t* nq_dev) {
  uint64_t n = nruns * nchunks;
  if (!n) return;
  hipLaunchKernelGGL(k_target_hits, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, qlocs, run_start,
                     nruns, nq_total, nchunks, counts, q_seq_of, reinterpret_cast<unsigned long long*>(thits),
                     nruns_dev, nq_dev);
}

void launch_join_slice(const uint32_t* qlocs, const uint32_t* run_start, uint64_t nruns, uint64_t nq_total,
                       uint32_t nchunks, const uint32_t* counts, const uint32_t* q_seq_of, uint32_t t0,
                       uint32_t t1, uint32_t* slice_counts, uint32_t* slice_start, uint32_t* slice_nq,
                       hipStream_t st, const uint64_t* nruns_dev, const uint64_t* nq_dev) {
  uint64_t n = nruns * nchunks;
  if (!n) return;
  hipLaunchKernelGGL(k_join_slice, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, qlocs, run_start, nruns,
                     nq_total, nchunks, counts, q_seq_of, t0, t1, slice_counts, slice_start, slice_nq, nruns_dev,
                     nq_dev);
}

}  // namespace dmr
