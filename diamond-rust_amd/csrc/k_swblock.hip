// Banded SW with one DP per block of 16 waves, for bands wider than one wave
// covers (64*32 offsets): exact int32, widths up to 1024*K (K = 4, 8, 16), in
// score mode or with trace bytes and traceback.
//
// The recurrence and its row-parallel form are those of k_swwave.hip, over
// a block instead of a wave: thread g (wave w, lane l, g = 64w + l) owns band
// offsets t = g*K .. g*K+K-1 (column j = i - d1 + t of row i). Per row, D and
// E read the previous row (the E source of a thread's last offset is the
// next thread's first offset: DPP inside a wave, LDS across waves, written at
// the end of the previous row), and the horizontal gap F is an exclusive
// max-plus prefix over the row: a wave prefix (DPP) plus the maxima of the
// earlier waves (LDS), so every row costs two block barriers (three with
// trace bytes, whose first-cell extension bit needs the previous wave's last
// F). Subject letters are read per row from memory (the band's columns are
// contiguous).
//
// Score mode writes the DP's best score by job (as the wave kernel). Trace
// mode needs the DP's best score from the score pass: the first row where a
// cell reaches it holds the first maximum in row-major order (its first such
// offset is the column), the rows after it are not run, and thread 0 walks
// the trace back (gapped::traceback, src/gapped.rs:205-279) reading the
// block's trace rows from memory. DPs whose best may lie in the clamped rows
// below the band (best < 11), or that do not fit, go to the fallback list
// for the row-exact kernel.
#include "dmr_kernels.hpp"

namespace dmr {

namespace {

constexpr int32_t kNeg = INT32_MIN / 4;
constexpr int kPad = 25;
constexpr int kBW = 16;       // waves per block
constexpr int kBT = 64 * kBW;  // threads per block

constexpr int kRowShr1 = 0x111, kRowShr2 = 0x112, kRowShr4 = 0x114, kRowShr8 = 0x118;
constexpr int kWaveShl1 = 0x130, kWaveShr1 = 0x138, kRowBcast15 = 0x142, kRowBcast31 = 0x143;

template <int kCtrl, int kRowMask = 0xf>
__device__ __forceinline__ int dpp(int old, int v) {
  return __builtin_amdgcn_update_dpp(old, v, kCtrl, kRowMask, 0xf, false);
}
__device__ __forceinline__ int from_next(int v, int edge) { return dpp<kWaveShl1>(edge, v); }
__device__ __forceinline__ int from_prev(int v, int edge) { return dpp<kWaveShr1>(edge, v); }
__device__ __forceinline__ int scan_max(int v) {
  constexpr int kId = INT32_MIN;
  v = max(v, dpp<kRowShr1>(kId, v));
  v = max(v, dpp<kRowShr2>(kId, v));
  v = max(v, dpp<kRowShr4>(kId, v));
  v = max(v, dpp<kRowShr8>(kId, v));
  v = max(v, dpp<kRowBcast15, 0xa>(kId, v));
  v = max(v, dpp<kRowBcast31, 0xc>(kId, v));
  return v;
}

struct RowRange {
  int d0, d1, m;
  __device__ __forceinline__ int lo(int i) const { return min(max(i - d1, 0), m); }
  __device__ __forceinline__ int hi(int i) const { return max(min(max(i - d0, 0), m), lo(i)); }
};

template <int K, bool kTrace>
__global__ __launch_bounds__(kBT) void k_sw_block(const SwJob* __restrict__ jobs, const uint32_t* __restrict__ list,
                                                 uint32_t njobs, int wmin, SeqSet q, SeqSet db,
                                                 const int8_t* __restrict__ m32, int gap_open, int gap_extend,
                                                 SwOut* __restrict__ outs, int32_t* __restrict__ scores,
                                                 uint8_t* __restrict__ trace,
                                                 const uint64_t* __restrict__ trace_off, uint32_t* __restrict__ ops,
                                                 const uint64_t* __restrict__ ops_off,
                                                 uint32_t* __restrict__ fallback, uint32_t* __restrict__ nfallback,
                                                 const int32_t* __restrict__ known) {
  __shared__ int8_t M[kScoreStride * kScoreStride];
  __shared__ int nbH[2][kBW + 1], nbE[2][kBW + 1];  // first offset of each wave, previous row
  __shared__ int tot[2][kBW];                        // each wave's F prefix maximum
  __shared__ int lfH[2][kBW], lfF[2][kBW];           // last offset's H' and F (trace mode)
  __shared__ int hit[2];                             // first offset reaching the known best
  __shared__ int red[kBW];
  const uint32_t task = blockIdx.x;
  if (task >= njobs) return;  // the whole block: no barrier is skipped
  const uint32_t jid = list[task];
  const SwJob J = jobs[jid];
  const int W = J.d1 - J.d0 + 1;
  // Widths of this launch: (wmin, 1024K]; others are another launch's, and
  // the widest launch hands bands past its reach to the row-exact kernel.
  if (W <= wmin) return;
  if (W > kBT * K) {
    if (K == kMaxBlockK && threadIdx.x == 0) {
      fallback[atomicAdd(nfallback, 1u)] = kTrace ? task : jid;
      if (kTrace) outs[task] = SwOut{};
    }
    return;
  }
  for (int k = threadIdx.x; k < kScoreStride * kScoreStride; k += kBT) M[k] = m32[k];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int oc = gap_open + gap_extend, ge = gap_extend;
  const int d0 = J.d0, d1 = J.d1;
  const uint8_t* qs = q.res + q.offs[J.target] + J.q0;
  const int n = (int)(J.q1 - J.q0);
  const uint32_t sb = db.offs[J.subject];
  const uint8_t* ss = db.res + sb;
  const int m = (int)(db.offs[J.subject + 1] - sb);
  const int rlo = max(1, d0 + 1), rhi = min(n, m + d1);
  const int istar = m + d1;
  int Hp[K], Ep[K];
  bool valid[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int t = tid * K + k;
    valid[k] = t < W;
    Hp[k] = valid[k] ? 0 : kNeg;
    Ep[k] = kNeg;
  }
  if (lane == 0) {
    nbH[rlo & 1][wave] = Hp[0];
    nbE[rlo & 1][wave] = Ep[0];
  }
  if (tid < 2) {
    nbH[tid][kBW] = kNeg;  // the wall right of the band
    nbE[tid][kBW] = kNeg;
    hit[tid] = INT32_MAX;
  }
  __syncthreads();
  const int stop_at = (kTrace && known) ? known[jid] : INT32_MAX;
  const bool rows = rlo <= rhi && m >= 1;
  int lbest = 0, bi = 0, bj = 0;
  bool found = false;
  uint8_t* tr = kTrace ? trace + trace_off[task] : nullptr;
  const int stepK = ge * K;
  for (int i = rlo; rows && i <= rhi; ++i) {
    const int cur = i & 1, nxt = cur ^ 1;
    const int qrow = (int)qs[i - 1] * kScoreStride;
    const int nH = from_next(Hp[0], nbH[cur][wave + 1]), nE = from_next(Ep[0], nbE[cur][wave + 1]);
    int Hq[K], E[K], D[K];
    uint32_t eext = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int j = i - d1 + tid * K + k;
      const int sl = (j >= 1 && j <= m) ? (int)ss[j - 1] : kPad;
      D[k] = Hp[k] + M[qrow + sl];
      const int hs = k + 1 < K ? Hp[k + 1] : nH;
      const int es = k + 1 < K ? Ep[k + 1] : nE;
      const int eo = hs - oc, ee = es - ge;
      E[k] = max(eo, ee);
      eext |= (uint32_t)(ee > eo) << k;
      Hq[k] = max(max(D[k], E[k]), 0);
    }
    int L = Hq[0] - oc - ge * (K - 1);
#pragma unroll
    for (int k = 1; k < K; ++k) L = max(L, Hq[k] - oc - ge * (K - 1 - k));
    // Prefix over the block of Y(g) = L(g) + e*K*g: the wave's inclusive
    // prefix, then the maxima of the earlier waves.
    const int inc = scan_max(L + stepK * lane);
    if (lane == 63) tot[cur][wave] = inc + stepK * 64 * wave;
    __syncthreads();
    int carry = kNeg;
    for (int w = 0; w < wave; ++w) carry = max(carry, tot[cur][w]);
    const int exc = max(carry, from_prev(inc, kNeg) + stepK * 64 * wave);
    int F[K];
    F[0] = exc - stepK * (tid - 1);
    uint32_t fext = 0;
#pragma unroll
    for (int k = 1; k < K; ++k) {
      const int fo = Hq[k - 1] - oc, fe = F[k - 1] - ge;
      F[k] = max(fo, fe);
      fext |= (uint32_t)(fe > fo) << k;
    }
    if (kTrace) {
      // Extension bit of the first owned cell: H' and F of the cell to its
      // left (the previous wave's last lane across waves).
      if (lane == 63) {
        lfH[cur][wave] = Hq[K - 1];
        lfF[cur][wave] = F[K - 1];
      }
      __syncthreads();
      const int pH = wave > 0 ? lfH[cur][wave - 1] : (1 << 29);
      const int pF = wave > 0 ? lfF[cur][wave - 1] : kNeg;
      const int fo = from_prev(Hq[K - 1], pH) - oc;
      const int fe = from_prev(F[K - 1], pF) - ge;
      fext |= (uint32_t)(fe > fo);
    }
    int rowbest = 0;
    uint32_t tw[(K + 3) / 4];
#pragma unroll
    for (int w = 0; w < (K + 3) / 4; ++w) tw[w] = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      int hv;
      if (kTrace) {
        int choice = 0;
        hv = 0;
        if (D[k] > hv) {
          hv = D[k];
          choice = 1;
        }
        if (E[k] > hv) {
          hv = E[k];
          choice = 2;
        }
        if (F[k] > hv) {
          hv = F[k];
          choice = 3;
        }
        const uint32_t byte = (uint32_t)choice | (((eext >> k) & 1u) << 2) | (((fext >> k) & 1u) << 3);
        tw[k / 4] |= byte << (8 * (k % 4));
      } else {
        hv = max(Hq[k], F[k]);
      }
      Hp[k] = valid[k] ? hv : kNeg;
      Ep[k] = valid[k] ? E[k] : kNeg;
      rowbest = max(rowbest, Hp[k]);
    }
    if (kTrace) {
      // Row i of the trace: kBT*K bytes, offset t at byte t.
      uint32_t* row = reinterpret_cast<uint32_t*>(tr + (uint64_t)(i - rlo) * (kBT * K) + (uint64_t)tid * K);
#pragma unroll
      for (int w = 0; w < K / 4; ++w) row[w] = tw[w];
      if (rowbest >= stop_at) {
        int tf = INT32_MAX;
#pragma unroll
        for (int k = K - 1; k >= 0; --k)
          if (valid[k] && Hp[k] == stop_at) tf = tid * K + k;
        atomicMin(&hit[cur], tf);
      }
      if (tid == 0) hit[nxt] = INT32_MAX;
    } else {
      lbest = max(lbest, rowbest);
    }
    if (lane == 0) {
      nbH[nxt][wave] = Hp[0];
      nbE[nxt][wave] = Ep[0];
    }
    __syncthreads();
    if (kTrace && hit[cur] != INT32_MAX) {
      found = true;
      bi = i;
      bj = i - d1 + hit[cur];
      break;
    }
  }
  if constexpr (!kTrace) {
    // Block maximum, then the clamped rows below the band as the wave kernel
    // finishes them (H and E of (istar, m) are offset 0's).
    int b = lbest;
    for (int o = 32; o > 0; o >>= 1) b = max(b, __shfl_down(b, o, 64));
    if (lane == 0) red[wave] = b;
    __syncthreads();
    if (tid == 0) {
      int best = 0;
      for (int w = 0; w < kBW; ++w) best = max(best, red[w]);
      bool need_exact = best < 11 && istar < n;
      if (need_exact && m >= 1 && istar >= 1 && rows) {
        int h = Hp[0], e = Ep[0];
        const int sl = ss[m - 1];
        for (int i = istar + 1; i <= n; ++i) {
          int ev = kNeg;
          if (h > kNeg / 2 && h - oc > ev) ev = h - oc;
          if (e > kNeg / 2 && e - ge > ev) ev = e - ge;
          const int hv = max(0, max((int)M[qs[i - 1] * kScoreStride + sl], ev));
          best = max(best, hv);
          h = hv;
          e = ev;
        }
        need_exact = false;
      }
      if (need_exact) {
        fallback[atomicAdd(nfallback, 1u)] = jid;
      } else {
        scores[jid] = best;
      }
    }
  } else {
    if (tid >= 64) return;  // the first wave walks the trace back
    const int lane = tid;
    const int best = found ? stop_at : 0;
    if (!found || (best < 11 && istar < n)) {
      if (lane == 0) {
        fallback[atomicAdd(nfallback, 1u)] = task;
        outs[task] = SwOut{};
      }
      return;
    }
    // Traceback from (bi, bj) through this block's trace rows: the
    // reference's loop (src/gapped.rs:222-265) a run of steps at a time, as
    // in the wave kernel's walk: lane k reads the trace byte of step k in the
    // current state's direction (the diagonal in state H, the column in state
    // E, the row in state F), checks the loop condition and the row range
    // there, and one ballot gives the run's length.
    SwOut o = {};
    o.score = best;
    o.best_i = (uint32_t)bi;
    o.best_j = (uint32_t)bj;
    const RowRange R{d0, d1, m};
    uint32_t* op = ops + ops_off[task];
    uint32_t nr = 0, cur = 3, cnt = 0;
    int state = 0, i = bi, j = bj;
    bool in_gap = false;
    auto push = [&](uint32_t code, uint32_t n) {
      if (code == cur) {
        cnt += n;
      } else {
        if (cnt) {
          if (lane == 0) op[nr] = cur | (cnt << 2);
          ++nr;
        }
        cur = code;
        cnt = n;
      }
    };
    while (i > 0 && j > 0) {
      if (j < R.lo(i) || j > R.hi(i) || i < rlo) break;
      const int di = state == 2 ? 0 : 1, dj = state == 1 ? 0 : 1;
      const int ik = i - lane * di, jk = j - lane * dj;
      const bool ok = ik > 0 && jk > 0 && jk >= R.lo(ik) && jk <= R.hi(ik) && ik >= rlo;
      const int tv = ok ? (int)tr[(uint64_t)(ik - rlo) * (kBT * K) + (uint64_t)(jk - ik + d1)] : 0;
      const uint64_t bad = __ballot(!ok);  // bit k: the loop stops before step k
      if (state == 0) {
        const int c = __builtin_amdgcn_readfirstlane(tv) & 3;  // lane 0's cell
        if (c == 0) break;
        if (c == 2 || c == 3) {
          state = c - 1;
          continue;
        }
        const uint64_t stopm = bad | __ballot(ok && (tv & 3) != 1);
        const int run = stopm ? __builtin_ctzll(stopm) : 64;
        const bool in_run = lane < run;
        int x = 0, y = 0;
        if (in_run) {
          x = qs[ik - 1];
          y = ss[jk - 1];
        }
        const int idn = __builtin_popcountll(__ballot(in_run && x == y));
        o.length += run;
        o.identities += idn;
        o.mismatches += run - idn;
        o.positives += __builtin_popcountll(__ballot(in_run && M[x * kScoreStride + y] > 0));
        in_gap = false;
        push(0, (uint32_t)run);
        i -= run;
        j -= run;
      } else {
        const int bit = state == 1 ? 2 : 3;
        const uint64_t last = __ballot(ok && ((tv >> bit) & 1) == 0);
        const int f = bad ? __builtin_ctzll(bad) : 64;
        const int z = last ? __builtin_ctzll(last) : 64;
        const int g = z < f ? z + 1 : f;
        o.length += g;
        o.gaps += g;
        if (!in_gap) {
          o.gap_openings++;
          in_gap = true;
        }
        push(state == 1 ? 1u : 2u, (uint32_t)g);
        if (state == 1) i -= g;
        else j -= g;
        if (z < f) state = 0;
      }
    }
    if (cnt) {
      if (lane == 0) op[nr] = cur | (cnt << 2);
      ++nr;
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");  // lane 0's ops are read by every lane
    // The walk runs backwards: reverse the ops, one pair per lane.
    for (uint32_t a = lane; a < nr / 2; a += 64) {
      const uint32_t b = nr - 1 - a, x = op[a], y = op[b];
      op[a] = y;
      op[b] = x;
    }
    o.n_ops = nr;
    o.q_begin = (uint32_t)i;
    o.s_begin = (uint32_t)j;
    if (nr == 0) o.score = 0;
    if (lane == 0) outs[task] = o;
  }
}

}  // namespace

int swblock_k_for(int64_t width) {
  for (int K = 4; K <= kMaxBlockK; K <<= 1)
    if (width <= (int64_t)kBT * K) return K;
  return 0;
}

uint64_t swblock_row_bytes(int K) { return (uint64_t)kBT * K; }

void launch_sw_block(bool traceback, const SwJob* jobs, const uint32_t* list, uint32_t njobs, SeqSet q, SeqSet db,
                     const int8_t* m32, int gap_open, int gap_extend, SwOut* outs, int32_t* scores,
                     uint8_t* trace, const uint64_t* trace_off, uint32_t* ops, const uint64_t* ops_off,
                     uint32_t* fallback, uint32_t* nfallback, const int32_t* known, int only_k, hipStream_t st) {
  if (!njobs) return;
  // One launch per K: a block serves its DP only if the width falls in that
  // K's range ((1024 * K/2, 1024 * K], the smallest from 2048), so every DP
  // runs once, with the fewest idle offsets.
#define DMR_SWB(k, tr, lo)                                                                                    \
  hipLaunchKernelGGL((k_sw_block<k, tr>), dim3(njobs), dim3(kBT), 0, st, jobs, list, njobs, lo, q, db, m32,    \
                     gap_open, gap_extend, outs, scores, trace, trace_off, ops, ops_off, fallback, nfallback, known)
  if (traceback) {
    if (only_k == 0 || only_k == 4) DMR_SWB(4, true, 64 * kMaxWaveK);
    if (only_k == 0 || only_k == 8) DMR_SWB(8, true, kBT * 4);
    if (only_k == 0 || only_k == 16) DMR_SWB(16, true, kBT * 8);
  } else {
    if (only_k == 0 || only_k == 4) DMR_SWB(4, false, 64 * kMaxWaveK);
    if (only_k == 0 || only_k == 8) DMR_SWB(8, false, kBT * 4);
    if (only_k == 0 || only_k == 16) DMR_SWB(16, false, kBT * 8);
  }
#undef DMR_SWB
}

}  // namespace dmr
