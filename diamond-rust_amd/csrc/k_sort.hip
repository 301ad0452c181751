// Prefix scans and LSD radix sort for the seed and hit arrays.
//
// Radix sort: 8-bit digits, three kernels per pass: per-tile digit histogram
// (with the digit totals), scatter offsets (one block per digit), scatter.
// The scatter ranks keys stably inside a 256-thread tile with 64-lane ballot
// matching (8 ballots give each lane the mask of lanes sharing its digit), so
// every pass is stable and the final order equals a stable sort by key; the
// tile is staged in LDS by digit so its writes are coalesced.
#include "dmr_device.hpp"

namespace dmr {

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

constexpr int kSortThreads = 256;
constexpr int kSortWaves = kSortThreads / 64;
constexpr int kSortItems = 16;
constexpr int kSortTile = kSortThreads * kSortItems;
constexpr int kTotalStripes = 16;  // copies of the 256 digit totals per radix pass

// Block-wide exclusive scan of one value per thread; returns the exclusive
// prefix and writes the block total to *total.
__device__ uint64_t block_exclusive_scan(uint64_t v, uint64_t* total) {
  __shared__ uint64_t wave_sums[kScanThreads / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint64_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint64_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wave_sums[wave] = x;
  __syncthreads();
  uint64_t wave_off = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kScanThreads / 64; ++w) {
    uint64_t s = wave_sums[w];
    if (w < wave) wave_off += s;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return wave_off + x - v;
}

// LDS slot of tile element e: one padding word after every 8, so a thread's
// 8 consecutive elements (reads at stride 9 words of 8 bytes) and the
// wave's consecutive elements both spread over the banks.
__device__ __forceinline__ uint32_t scan_slot(uint32_t e) { return e + (e >> 3); }

// Exclusive scan of one tile per block; adds block_off[blockIdx] if given.
// The tile moves between global memory and LDS in coalesced order (element
// k * kScanThreads + thread); each thread scans 8 consecutive elements from
// LDS. in may equal out: a block reads its whole tile before writing it.
template <class T>
__global__ __launch_bounds__(kScanThreads) void k_scan_apply(const T* in, uint64_t* out, uint64_t n,
                                                            const uint64_t* block_off) {
  __shared__ uint64_t tile[kScanTile + kScanTile / 8];
  const uint64_t b0 = (uint64_t)blockIdx.x * kScanTile;
  const uint32_t t = threadIdx.x;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const uint32_t e = (uint32_t)k * kScanThreads + t;
    tile[scan_slot(e)] = b0 + e < n ? (uint64_t)in[b0 + e] : 0;
  }
  __syncthreads();
  uint64_t v[kScanItems];
  uint64_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    v[k] = tile[scan_slot(t * kScanItems + k)];
    s += v[k];
  }
  uint64_t tot;
  uint64_t run = block_exclusive_scan(s, &tot);
  if (block_off) run += block_off[blockIdx.x];
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    tile[scan_slot(t * kScanItems + k)] = run;
    run += v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const uint32_t e = (uint32_t)k * kScanThreads + t;
    if (b0 + e < n) out[b0 + e] = tile[scan_slot(e)];
  }
}

// Sum of each tile (order-free, so read in coalesced order).
template <class T>
__global__ __launch_bounds__(kScanThreads) void k_scan_reduce_blocked(const T* in, uint64_t n, uint64_t* sums) {
  const uint64_t b0 = (uint64_t)blockIdx.x * kScanTile;
  uint64_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const uint64_t i = b0 + (uint64_t)k * kScanThreads + threadIdx.x;
    if (i < n) s += (uint64_t)in[i];
  }
  uint64_t tot;
  block_exclusive_scan(s, &tot);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

template <class T>
void scan_impl(const T* in, uint64_t* out, uint64_t n, uint64_t* tmp, hipStream_t st) {
  if (n == 0) return;
  uint64_t nb = (n + kScanTile - 1) / kScanTile;
  if (nb == 1) {
    hipLaunchKernelGGL(k_scan_apply<T>, dim3(1), dim3(kScanThreads), 0, st, in, out, n,
                       (const uint64_t*)nullptr);
    return;
  }
  hipLaunchKernelGGL(k_scan_reduce_blocked<T>, dim3((unsigned)nb), dim3(kScanThreads), 0, st, in, n,
                     tmp);
  scan_impl<uint64_t>(tmp, tmp, nb, tmp + nb, st);
  hipLaunchKernelGGL(k_scan_apply<T>, dim3((unsigned)nb), dim3(kScanThreads), 0, st, in, out, n,
                     (const uint64_t*)tmp);
}

// ---- radix sort -----------------------------------------------------------

__global__ void k_radix_hist(const uint64_t* keys, uint64_t n, int shift, uint32_t* hist,
                             uint32_t nblocks, uint32_t* totals, const uint64_t* n_dev) {
  __shared__ uint32_t h[256];
  if (n_dev) n = *n_dev;  // a device count within the bound n (tiles past it count nothing)
  h[threadIdx.x] = 0;
  __syncthreads();
  uint64_t base = (uint64_t)blockIdx.x * kSortTile;
#pragma unroll 4
  for (int k = 0; k < kSortItems; ++k) {
    uint64_t i = base + (uint64_t)k * kSortThreads + threadIdx.x;
    if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[(uint64_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
  // Digit totals, spread over kTotalStripes copies so that few blocks add
  // to the same word.
  if (h[threadIdx.x]) atomicAdd(&totals[(blockIdx.x % kTotalStripes) * 256 + threadIdx.x], h[threadIdx.x]);
}

__global__ void k_zero_u32(uint32_t* p, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// Scatter offsets of one pass: block d (one per digit) writes
// offs[d * nblocks + b] = (keys of digits below d) + (keys of digit d in
// tiles before b), the exclusive scan of the digit-major histogram.
__global__ __launch_bounds__(kScanThreads) void k_radix_offsets(const uint32_t* hist, uint32_t nblocks,
                                                               const uint32_t* totals, uint64_t* offs) {
  __shared__ uint64_t base_s;
  const uint32_t d = blockIdx.x, t = threadIdx.x;
  uint64_t tot, mine = 0;
#pragma unroll
  for (int k = 0; k < kTotalStripes; ++k) mine += totals[k * 256 + t];
  const uint64_t below = block_exclusive_scan(mine, &tot);
  if (t == d) base_s = below;
  __syncthreads();
  uint64_t carry = base_s;
  const uint32_t* row = hist + (uint64_t)d * nblocks;
  uint64_t* out = offs + (uint64_t)d * nblocks;
  for (uint32_t c = 0; c < nblocks; c += kScanThreads) {
    const uint32_t b = c + t;
    const uint64_t v = b < nblocks ? row[b] : 0;
    const uint64_t ex = block_exclusive_scan(v, &tot);
    if (b < nblocks) out[b] = carry + ex;
    carry += tot;
  }
}

// One tile of kSortTile keys per block. Keys are ranked stably by digit in
// tile order (k, wave, lane), which is the input order: per item, 8 ballots
// give each lane the mask of lanes sharing its digit, and per-wave digit
// counts order the waves. The tile is then laid out in LDS by digit and
// written back in that order, so consecutive threads write consecutive
// positions of one digit's run (coalesced), at the tile's global offset for
// that digit (offs, from the scanned histograms).
template <bool kValues>
__global__ __launch_bounds__(kSortThreads) void k_radix_scatter(const uint64_t* kin, const uint32_t* vin,
                                                                uint64_t* kout, uint32_t* vout, uint64_t n,
                                                                int shift, const uint64_t* offs,
                                                                uint32_t nblocks, const uint64_t* n_dev) {
  __shared__ uint32_t wave_hist[kSortWaves][256];
  __shared__ uint32_t count[256];       // keys of each digit ranked so far
  __shared__ uint32_t local_base[256];  // first LDS slot of each digit
  __shared__ uint64_t run[256];         // global position of the tile's first key of each digit
  __shared__ uint64_t skey[kSortTile];
  __shared__ uint32_t sval[kValues ? kSortTile : 1];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (n_dev) n = *n_dev;
  if ((uint64_t)blockIdx.x * kSortTile >= n) return;  // a tile past the device count
  run[threadIdx.x] = offs[(uint64_t)threadIdx.x * nblocks + blockIdx.x];
  count[threadIdx.x] = 0;
#pragma unroll
  for (int w = 0; w < kSortWaves; ++w) wave_hist[w][threadIdx.x] = 0;
  __syncthreads();
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const uint64_t base = (uint64_t)blockIdx.x * kSortTile;
  uint64_t key[kSortItems];
  uint32_t val[kSortItems], pos[kSortItems];
#pragma unroll
  for (int k = 0; k < kSortItems; ++k) {
    const uint64_t i = base + (uint64_t)k * kSortThreads + threadIdx.x;
    const bool valid = i < n;
    key[k] = valid ? kin[i] : 0;
    val[k] = (kValues && valid) ? vin[i] : 0;
    const uint32_t d = (uint32_t)(key[k] >> shift) & 255u;
    uint64_t m = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const uint64_t bb = __ballot((d >> b) & 1u);
      m &= ((d >> b) & 1u) ? bb : ~bb;
    }
    const uint32_t rank = (uint32_t)__popcll(m & lt_mask);
    if (valid && rank == 0) wave_hist[wave][d] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t p = count[d] + rank;
    for (int w = 0; w < wave; ++w) p += wave_hist[w][d];
    pos[k] = p;
    __syncthreads();
    {
      uint32_t sum = 0;
#pragma unroll
      for (int w = 0; w < kSortWaves; ++w) {
        sum += wave_hist[w][threadIdx.x];
        wave_hist[w][threadIdx.x] = 0;
      }
      count[threadIdx.x] += sum;
    }
    __syncthreads();
  }
  uint64_t tile_n;
  local_base[threadIdx.x] = (uint32_t)block_exclusive_scan(count[threadIdx.x], &tile_n);
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kSortItems; ++k) {
    if (base + (uint64_t)k * kSortThreads + threadIdx.x >= n) continue;
    const uint32_t slot = local_base[(uint32_t)(key[k] >> shift) & 255u] + pos[k];
    skey[slot] = key[k];
    if (kValues) sval[slot] = val[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kSortItems; ++k) {
    const uint32_t idx = (uint32_t)k * kSortThreads + threadIdx.x;
    if (idx >= tile_n) break;
    const uint64_t kv = skey[idx];
    const uint32_t d = (uint32_t)(kv >> shift) & 255u;
    const uint64_t g = run[d] + (idx - local_base[d]);
    kout[g] = kv;
    if (kValues) vout[g] = sval[idx];
  }
}

template <bool kValues>
void radix_impl(uint64_t* k0, uint32_t* v0, uint64_t* k1, uint32_t* v1, uint64_t n, int bits,
                uint32_t* hist, uint64_t* scan_tmp, hipStream_t st, const uint64_t* n_dev) {
  if (n <= 1 || bits <= 0) return;
  uint32_t nb = (uint32_t)((n + kSortTile - 1) / kSortTile);
  uint64_t* offs = scan_tmp;                                        // 256 * nb
  uint32_t* totals = reinterpret_cast<uint32_t*>(scan_tmp + 256ull * nb);  // stripes x 256 per pass
  int passes = (bits + 7) / 8;
  hipLaunchKernelGGL(k_zero_u32, dim3((unsigned)(passes * kTotalStripes)), dim3(256), 0, st, totals,
                     (uint32_t)(256 * kTotalStripes * passes));
  uint64_t *ka = k0, *kb = k1;
  uint32_t *va = v0, *vb = v1;
  for (int p = 0; p < passes; ++p) {
    int shift = 8 * p;
    uint32_t* pt = totals + 256 * kTotalStripes * p;
    hipLaunchKernelGGL(k_radix_hist, dim3(nb), dim3(kSortThreads), 0, st, ka, n, shift, hist, nb, pt, n_dev);
    hipLaunchKernelGGL(k_radix_offsets, dim3(256), dim3(kScanThreads), 0, st, hist, nb, pt, offs);
    hipLaunchKernelGGL(k_radix_scatter<kValues>, dim3(nb), dim3(kSortThreads), 0, st, ka, va, kb,
                       vb, n, shift, offs, nb, n_dev);
    std::swap(ka, kb);
    std::swap(va, vb);
  }
  if (ka != k0) {
    (void)hipMemcpyAsync(k0, ka, n * sizeof(uint64_t), hipMemcpyDeviceToDevice, st);
    if (kValues) (void)hipMemcpyAsync(v0, va, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, st);
  }
}

}  // namespace

uint64_t scan_tmp_elems(uint64_t n) {
  uint64_t total = 0;
  while (n > (uint64_t)kScanTile) {
    n = (n + kScanTile - 1) / kScanTile;
    total += n;
  }
  return total + 8;
}

uint64_t radix_hist_elems(uint64_t n) { return 256ull * ((n + kSortTile - 1) / kSortTile) + 256; }

uint64_t radix_scan_tmp_elems(uint64_t n) {
  // Scatter offsets (256 per tile), then the striped digit totals of up to 8
  // passes (u32).
  const uint64_t h = 256ull * ((n + kSortTile - 1) / kSortTile);
  return h + 8 * kTotalStripes * 256 / 2 + 8;
}

void scan_exclusive_u64(const uint64_t* in, uint64_t* out, uint64_t n, uint64_t* tmp,
                        hipStream_t st) {
  scan_impl<uint64_t>(in, out, n, tmp, st);
}

void scan_exclusive_u32_to_u64(const uint32_t* in, uint64_t* out, uint64_t n, uint64_t* tmp,
                               hipStream_t st) {
  scan_impl<uint32_t>(in, out, n, tmp, st);
}

void radix_sort_pairs(uint64_t* k0, uint32_t* v0, uint64_t* k1, uint32_t* v1, uint64_t n, int bits,
                      uint32_t* hist, uint64_t* scan_tmp, hipStream_t st, const uint64_t* n_dev) {
  radix_impl<true>(k0, v0, k1, v1, n, bits, hist, scan_tmp, st, n_dev);
}

void radix_sort_keys(uint64_t* k0, uint64_t* k1, uint64_t n, int bits, uint32_t* hist,
                     uint64_t* scan_tmp, hipStream_t st, const uint64_t* n_dev) {
  radix_impl<false>(k0, nullptr, k1, nullptr, n, bits, hist, scan_tmp, st, n_dev);
}

}  // namespace dmr
