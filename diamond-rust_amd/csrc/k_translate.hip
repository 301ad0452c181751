// Six-frame translation and --min-orf masking of blastx reads
// (src/translate.rs:9-137, src/workflow.rs:112-127), on the device.
//
// Frames 0..2 read the forward strand from offsets 0..2, frames 3..5 the
// reverse complement from offsets 0..2; frame k of read r is slot 6r + k.
// A codon with a base other than T/U, C, A, G (either case) becomes X; the
// complement maps only upper-case IUPAC letters (lower-case letters stay as
// they are, then read as their own base), as complement() does. Masking
// turns every stop-free stretch shorter than min_orf into X; frames left
// without a letter below X are dropped, the rest become search targets in
// slot order.
//
// Frame lengths are (L - offset) / 3 codons, laid out by the host. Kernels:
// translation (one block per read, threads over the codons of its six
// frames), masking (one block per
// frame: each letter scans for stops at most min_orf letters each way, and
// the block counts the frame's letters below X), and the compaction of the
// kept frames into the target set (one block per frame).
#include "dmr_kernels.hpp"

namespace dmr {

namespace {

constexpr uint8_t kX = 23, kStop = 24;

// NCBI letter code of an amino-acid symbol (src/alphabet.rs:24-46).
constexpr uint8_t aa_code(char c) {
  const char* sym = "ARNDCQEGHILKMFPSTWYVBJZX*";
  for (uint8_t i = 0; i < 25; ++i)
    if (sym[i] == c) return i;
  return kX;
}

// Codon table in TCAG order (index 16*b0 + 4*b1 + b2), as letter codes.
struct CodonTable {
  uint8_t code[64];
  constexpr CodonTable() : code{} {
    const char* t = "FFLLSSSSYY**CC*WLLLLPPPPHHQQRRRRIIIMTTTTNNKKSSRRVVVVAAAADDEEGGGG";
    for (int i = 0; i < 64; ++i) code[i] = aa_code(t[i]);
  }
};
__constant__ CodonTable kCodons = CodonTable();

__device__ __forceinline__ uint8_t nt_base(uint8_t b) {
  switch (b) {
    case 'T': case 't': case 'U': case 'u': return 0;
    case 'C': case 'c': return 1;
    case 'A': case 'a': return 2;
    case 'G': case 'g': return 3;
    default: return 4;
  }
}

__device__ __forceinline__ uint8_t nt_complement(uint8_t b) {
  switch (b) {
    case 'A': return 'T'; case 'T': return 'A'; case 'C': return 'G'; case 'G': return 'C';
    case 'U': return 'A'; case 'R': return 'Y'; case 'Y': return 'R'; case 'S': return 'S';
    case 'W': return 'W'; case 'K': return 'M'; case 'M': return 'K'; case 'B': return 'V';
    case 'V': return 'B'; case 'D': return 'H'; case 'H': return 'D'; case 'N': return 'N';
    default: return b;
  }
}

__global__ void k_translate_frames(const uint8_t* __restrict__ nt, const uint64_t* __restrict__ roff,
                                   const uint64_t* __restrict__ foff, uint8_t* __restrict__ out) {
  const uint32_t r = blockIdx.x;
  const uint8_t* s = nt + roff[r];
  const int64_t L = (int64_t)(roff[r + 1] - roff[r]);
  const uint64_t* fo = foff + 6ull * r;
  const uint64_t f0 = fo[0], total = fo[6] - f0;
  for (uint64_t k = threadIdx.x; k < total; k += blockDim.x) {
    int f = 0;
    while (f < 5 && f0 + k >= fo[f + 1]) ++f;
    const int64_t c = (int64_t)(f0 + k - fo[f]);
    const int64_t o = f % 3;
    uint8_t b0, b1, b2;
    if (f < 3) {
      const int64_t i = o + 3 * c;
      b0 = s[i];
      b1 = s[i + 1];
      b2 = s[i + 2];
    } else {
      // rc[k] = complement(nt[L - 1 - k])
      const int64_t i = L - 1 - (o + 3 * c);
      b0 = nt_complement(s[i]);
      b1 = nt_complement(s[i - 1]);
      b2 = nt_complement(s[i - 2]);
    }
    const uint8_t a = nt_base(b0), b = nt_base(b1), d = nt_base(b2);
    out[f0 + k] = (a > 3 || b > 3 || d > 3) ? kX : kCodons.code[a * 16 + b * 4 + d];
  }
}

// mask_short_orfs (src/translate.rs:101-121): a letter in a stop-free
// stretch [b, e) with e - b < min_len becomes X. For letter i, left counts
// the non-stops just before it and right those from i on; the scans stop
// once left + right reaches min_len (the stretch is then long enough).
__global__ void k_mask_orfs(const uint64_t* __restrict__ foff, const uint32_t* __restrict__ min_orf,
                            const uint8_t* __restrict__ in, uint8_t* __restrict__ out,
                            uint32_t* __restrict__ good) {
  __shared__ uint32_t cnt;
  const uint32_t f = blockIdx.x;
  const uint64_t b = foff[f], len = foff[f + 1] - b;
  const uint32_t min_len = min_orf[f / 6];
  const uint8_t* x = in + b;
  uint8_t* y = out + b;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  uint32_t mine = 0;
  for (uint64_t i = threadIdx.x; i < len; i += blockDim.x) {
    uint8_t v = x[i];
    if (v != kStop && min_len > 1) {
      uint64_t right = 0, left = 0;
      while (right < min_len && i + right < len && x[i + right] != kStop) ++right;
      while (left + right < min_len && left < i && x[i - 1 - left] != kStop) ++left;
      if (left + right < min_len) v = kX;
    }
    y[i] = v;
    mine += v < kX;
  }
  if (mine) atomicAdd(&cnt, mine);
  __syncthreads();
  if (threadIdx.x == 0) good[f] = cnt;
}

// Kept frames (good > 0) into the target set: letters at their target
// offset, the target's record and frame.
__global__ void k_compact_frames(const uint64_t* __restrict__ foff, const uint32_t* __restrict__ good,
                                 const uint64_t* __restrict__ kept_before, const uint64_t* __restrict__ tletters,
                                 const uint8_t* __restrict__ in, uint8_t* __restrict__ letters,
                                 uint32_t* __restrict__ toffs, uint32_t* __restrict__ query_of,
                                 uint8_t* __restrict__ frame_of) {
  const uint32_t f = blockIdx.x;
  if (!good[f]) return;
  const uint64_t b = foff[f], len = foff[f + 1] - b;
  const uint64_t t = kept_before[f], dst = tletters[f];
  for (uint64_t i = threadIdx.x; i < len; i += blockDim.x) letters[dst + i] = in[b + i];
  if (threadIdx.x == 0) {
    toffs[t] = (uint32_t)dst;
    query_of[t] = f / 6;
    frame_of[t] = (uint8_t)(f % 6);
  }
}

// Letters of kept frames only (the compaction's letter offsets).
__global__ void k_kept_lengths(const uint64_t* __restrict__ foff, const uint32_t* __restrict__ good,
                               uint32_t nframes, uint32_t* __restrict__ keep, uint32_t* __restrict__ klen) {
  const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nframes) return;
  const bool k = good[f] != 0;
  keep[f] = k ? 1u : 0u;
  klen[f] = k ? (uint32_t)(foff[f + 1] - foff[f]) : 0u;
}

__global__ void k_seq_of(const uint32_t* __restrict__ offs, uint32_t n, uint32_t* __restrict__ seq_of) {
  const uint32_t k = blockIdx.x;
  for (uint32_t p = offs[k] + threadIdx.x; p < offs[k + 1]; p += blockDim.x) seq_of[p] = k;
}

inline unsigned nblk(uint64_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

}  // namespace

void launch_translate_frames(const uint8_t* nt, const uint64_t* roff, uint32_t n, const uint64_t* foff,
                             uint8_t* out, hipStream_t st) {
  if (!n) return;
  hipLaunchKernelGGL(k_translate_frames, dim3(n), dim3(256), 0, st, nt, roff, foff, out);
}

void launch_mask_orfs(const uint64_t* foff, uint32_t nframes, const uint32_t* min_orf, const uint8_t* in,
                      uint8_t* out, uint32_t* good, hipStream_t st) {
  if (!nframes) return;
  hipLaunchKernelGGL(k_mask_orfs, dim3(nframes), dim3(256), 0, st, foff, min_orf, in, out, good);
}

void launch_kept_lengths(const uint64_t* foff, const uint32_t* good, uint32_t nframes, uint32_t* keep,
                         uint32_t* klen, hipStream_t st) {
  if (!nframes) return;
  hipLaunchKernelGGL(k_kept_lengths, dim3(nblk(nframes, 256)), dim3(256), 0, st, foff, good, nframes, keep, klen);
}

void launch_compact_frames(const uint64_t* foff, const uint32_t* good, const uint64_t* kept_before,
                           const uint64_t* tletters, const uint8_t* in, uint8_t* letters, uint32_t* toffs,
                           uint32_t* query_of, uint8_t* frame_of, uint32_t nframes, hipStream_t st) {
  if (!nframes) return;
  hipLaunchKernelGGL(k_compact_frames, dim3(nframes), dim3(256), 0, st, foff, good, kept_before, tletters, in,
                     letters, toffs, query_of, frame_of);
}

void launch_seq_of(const uint32_t* offs, uint32_t n, uint32_t* seq_of, hipStream_t st) {
  if (!n) return;
  hipLaunchKernelGGL(k_seq_of, dim3(n), dim3(256), 0, st, offs, n, seq_of);
}

}  // namespace dmr
