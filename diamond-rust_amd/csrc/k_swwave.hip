// Banded SW with one DP per wavefront: exact int32, any band width up to
// 64*K (K = 1..32, so up to 2048), optional trace bytes and traceback.
//
// Lane l owns band offsets t = l*K .. l*K+K-1 (column j = i - d1 + t of row
// i). Rows are processed in order; within a row, D and E only read the
// previous row (E's source t+1 comes from the next lane for the last owned
// offset), and the horizontal gap F is a max-plus prefix over the row:
//   F(t) = max_{t' < t} (H'(t') - o - e*(t-1-t')),  H' = max(0, D, E),
// computed as an exclusive wave-wide prefix maximum of per-lane partials.
// Using H' instead of H gives the same F and the same extension bits as the
// cell-by-cell recurrence of gapped::extend_span (src/gapped.rs:151-169): a gap
// opened from a cell whose H came from F is never better than extending F.
// Trace bytes are choice | e_ext << 2 | f_ext << 3 with the reference's
// strict D > E > F priority (src/gapped.rs:171-188); the best cell is the
// first strict maximum in row-major order.
//
// Row ranges follow the padding scheme of the packed kernel: cells with
// j <= 0 read a padding letter (-64) and stay at H = 0; cells with j > m are
// strictly below an earlier cell and never become the first maximum; the
// column after the band is a wall (no E source). Rows past m + d1 (the single
// clamped cell at column m) are left to the row-exact int32 kernel, which is
// only needed when the best score is below the largest matrix entry (11).
#include "dmr_kernels.hpp"

namespace dmr {

namespace {

constexpr int32_t kNeg = INT32_MIN / 4;
constexpr int kPad = 25;

__device__ __forceinline__ int wmax(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ int wmin(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
  return v;
}
// Inclusive prefix maximum over the 64 lanes.
__device__ __forceinline__ int scan_max(int v, int lane) {
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int u = __shfl_up(v, o, 64);
    if (lane >= o) v = max(v, u);
  }
  return v;
}

// Marks a value that is equal in every lane as wave-uniform.
__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint32_t uniform(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}

struct RowRange {
  int d0, d1, m;
  __device__ __forceinline__ int lo(int i) const { return min(max(i - d1, 0), m); }
  __device__ __forceinline__ int hi(int i) const { return max(min(max(i - d0, 0), m), lo(i)); }
};

template <int K, bool kTrace>
__global__ __launch_bounds__(256) void k_sw_wave(const SwJob* __restrict__ jobs,
                                                 const uint32_t* __restrict__ list, uint32_t njobs,
                                                 SeqSet q, SeqSet db, const int8_t* __restrict__ m32,
                                                 int gap_open, int gap_extend,
                                                 uint32_t* __restrict__ counter,
                                                 SwOut* __restrict__ outs, uint8_t* __restrict__ trace,
                                                 const uint64_t* __restrict__ trace_off,
                                                 uint32_t* __restrict__ ops,
                                                 const uint64_t* __restrict__ ops_off,
                                                 uint32_t* __restrict__ fallback,
                                                 uint32_t* __restrict__ nfallback) {
  __shared__ int8_t M[kScoreStride * kScoreStride];
  for (int k = threadIdx.x; k < kScoreStride * kScoreStride; k += blockDim.x) M[k] = m32[k];
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int oc = gap_open + gap_extend, ge = gap_extend;
  for (;;) {
    // Task fetch without a lane-divergent branch: every lane takes part in
    // the atomic and only lane 0 adds 1, so lane 0's old value is this
    // wave's task. The task and everything derived from it are held in
    // scalar registers, which keeps the loop's exit and the row loop uniform.
    // (A `lane == 0` branch here, correlated with the one that writes the
    // result, lets the optimizer thread lanes 1-63 from the end of one
    // iteration into the next while lane 0 is inactive.)
    const uint32_t task = uniform(atomicAdd(counter, lane == 0 ? 1u : 0u));
    if (task >= njobs) break;
    const uint32_t jid = uniform(list[task]);
    const uint32_t tgt = uniform(jobs[jid].target), sub = uniform(jobs[jid].subject);
    const uint32_t q0 = uniform(jobs[jid].q0), q1 = uniform(jobs[jid].q1);
    const int d0 = uniform(jobs[jid].d0), d1 = uniform(jobs[jid].d1);
    const uint8_t* qs = q.res + uniform(q.offs[tgt]) + q0;
    const int n = (int)(q1 - q0);
    const uint32_t sb = uniform(db.offs[sub]);
    const uint8_t* ss = db.res + sb;
    const int m = (int)(uniform(db.offs[sub + 1]) - sb);
    const int W = d1 - d0 + 1;
    const int rlo = max(1, d0 + 1), rhi = min(n, m + d1);
    const int istar = m + d1;

    int Hp[K], Ep[K];
    bool valid[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int t = lane * K + k;
      valid[k] = t < W;
      Hp[k] = valid[k] ? 0 : kNeg;
      Ep[k] = kNeg;
    }
    int best = 0, lbest = 0, bi = 0, bj = 0;
    uint8_t* tr = kTrace ? trace + trace_off[task] : nullptr;
    for (int i = rlo; i <= rhi; ++i) {
      const int qrow = qs[i - 1] * kScoreStride;
      // E sources of the last owned offset: next lane, previous row.
      int nH = __shfl_down(Hp[0], 1, 64), nE = __shfl_down(Ep[0], 1, 64);
      if (lane == 63) {
        nH = kNeg;
        nE = kNeg;
      }
      int Hq[K], E[K], D[K];
      uint32_t eext = 0;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const int t = lane * K + k;
        const int j = i - d1 + t;
        const int s = (j >= 1 && j <= m) ? (int)ss[j - 1] : kPad;
        D[k] = Hp[k] + M[qrow + s];
        const int hs = k + 1 < K ? Hp[k + 1] : nH;
        const int es = k + 1 < K ? Ep[k + 1] : nE;
        // Reference E: defined only when the source lies in the previous row.
        int e = kNeg;
        bool ex = false;
        if (hs > kNeg / 2 && hs - oc > e) e = hs - oc;
        if (es > kNeg / 2 && es - ge > e) {
          e = es - ge;
          ex = true;
        }
        E[k] = e;
        eext |= (uint32_t)ex << k;
        Hq[k] = max(0, max(D[k], e));
      }
      // Horizontal gaps: exclusive prefix over lanes, then the lane's chain.
      int L = kNeg;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        int c = valid[k] ? Hq[k] - oc - ge * (K - 1 - k) : kNeg;
        L = max(L - 0, c);
      }
      const int stepK = ge * K;
      int Y = L > kNeg / 2 ? L + stepK * lane : kNeg;
      int inc = scan_max(Y, lane);
      int exc = __shfl_up(inc, 1, 64);
      int Fin = (lane == 0 || exc <= kNeg / 2) ? kNeg : exc - stepK * (lane - 1);
      // Left neighbour's last H' for the first offset's extension bit.
      int F[K];
      uint32_t fext = 0;
      int prevHq = __shfl_up(Hq[K - 1], 1, 64);
      // F of the left neighbour's last cell is max(prevHq_prev - oc, ...); its
      // value is recovered from Fin: Fin = max(prevHq - oc, Fprev - ge).
#pragma unroll
      for (int k = 0; k < K; ++k) {
        if (k == 0) {
          F[0] = Fin;
        } else {
          int fo = Hq[k - 1] - oc;
          int fe = F[k - 1] > kNeg / 2 ? F[k - 1] - ge : kNeg;
          F[k] = max(fo, fe);
          fext |= (uint32_t)(fe > fo) << k;
        }
      }
      // Extension bit of the first owned cell needs F of the left cell.
      {
        int lastF = __shfl_up(F[K - 1], 1, 64);
        int fo = prevHq - oc;
        int fe = lastF > kNeg / 2 ? lastF - ge : kNeg;
        if (lane > 0 && fe > fo) fext |= 1u;
      }
      int rowbest = 0;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        int hv = 0, choice = 0;
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
        if (!valid[k]) hv = kNeg;
        Hp[k] = hv;
        Ep[k] = valid[k] ? E[k] : kNeg;
        rowbest = max(rowbest, hv);
        if (kTrace && valid[k])
          tr[(uint64_t)(i - rlo) * W + lane * K + k] =
              (uint8_t)(choice | (((eext >> k) & 1u) << 2) | (((fext >> k) & 1u) << 3));
      }
      if (kTrace) {
        const int rb = uniform(wmax(rowbest));
        if (rb > best) {
          int tf = 1 << 30;
#pragma unroll
          for (int k = K - 1; k >= 0; --k)
            if (valid[k] && Hp[k] == rb) tf = lane * K + k;
          tf = uniform(wmin(tf));
          best = rb;
          bi = i;
          bj = i - d1 + tf;
        }
      } else {
        lbest = max(lbest, rowbest);
      }
    }
    // Clamped rows below the band: rows istar+1..n hold the single cell
    // (i, m), whose diagonal predecessor lies outside the previous row's
    // range (it counts as 0) and which has no horizontal source. Their H is
    // at most max(11, H(istar, m) - o), so they can only raise a best score
    // below 11, the largest matrix entry. Score mode finishes them here from
    // H and E of (istar, m), held by offset 0 after row istar; trace mode and
    // bands that never reach row istar go to the row-exact kernel.
    bool need_exact = best < 11 && istar < n;
    if (!kTrace) {
      best = uniform(wmax(lbest));
      need_exact = best < 11 && istar < n;
      if (need_exact && m >= 1 && istar >= 1 && rlo <= rhi) {
        int h = uniform(Hp[0]), e = uniform(Ep[0]);
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
    }
    if (lane == 0) {
      if (need_exact) {
        fallback[atomicAdd(nfallback, 1u)] = kTrace ? task : jid;
      } else if (!kTrace) {
        outs[jid].score = best;
        outs[jid].flags = 0;
      } else {
        SwOut o = {};
        o.score = best;
        o.best_i = (uint32_t)bi;
        o.best_j = (uint32_t)bj;
        if (best > 0) {
          const RowRange R{d0, d1, m};
          uint32_t* op = ops + ops_off[task];
          uint32_t nr = 0, cur = 3, cnt = 0;
          int state = 0, i = bi, j = bj;
          bool in_gap = false;
          while (i > 0 && j > 0) {
            if (j < R.lo(i) || j > R.hi(i) || i < rlo) break;
            const uint8_t tv = tr[(uint64_t)(i - rlo) * W + (j - i + d1)];
            uint32_t code;
            if (state == 0) {
              const uint8_t c = tv & 3;
              if (c == 0) break;
              if (c == 2) {
                state = 1;
                continue;
              }
              if (c == 3) {
                state = 2;
                continue;
              }
              const uint8_t x = qs[i - 1], y = ss[j - 1];
              o.length++;
              if (x == y) o.identities++;
              else o.mismatches++;
              if (M[x * kScoreStride + y] > 0) o.positives++;
              in_gap = false;
              code = 0;
              --i;
              --j;
            } else {
              o.length++;
              o.gaps++;
              if (!in_gap) {
                o.gap_openings++;
                in_gap = true;
              }
              if (state == 1) {
                code = 1;
                if (((tv >> 2) & 1) == 0) state = 0;
                --i;
              } else {
                code = 2;
                if (((tv >> 3) & 1) == 0) state = 0;
                --j;
              }
            }
            if (code == cur) {
              ++cnt;
            } else {
              if (cnt) op[nr++] = cur | (cnt << 2);
              cur = code;
              cnt = 1;
            }
          }
          if (cnt) op[nr++] = cur | (cnt << 2);
          for (uint32_t a = 0, b = nr ? nr - 1 : 0; a < b; ++a, --b) {
            uint32_t x = op[a];
            op[a] = op[b];
            op[b] = x;
          }
          o.n_ops = nr;
          o.q_begin = (uint32_t)i;
          o.s_begin = (uint32_t)j;
          if (nr == 0) o.score = 0;
        }
        outs[task] = o;
      }
    }
  }
}

inline unsigned nblk(uint64_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

}  // namespace

int swwave_k_for(int64_t width) {
  for (int K = 1; K <= kMaxWaveK; K <<= 1)
    if (width <= 64 * K) return K;
  return 0;
}

void launch_sw_wave(int K, bool traceback, const SwJob* jobs, const uint32_t* list, uint32_t njobs,
                    SeqSet q, SeqSet db, const int8_t* m32, int gap_open, int gap_extend,
                    uint32_t* counter, SwOut* outs, uint8_t* trace, const uint64_t* trace_off,
                    uint32_t* ops, const uint64_t* ops_off, uint32_t* fallback, uint32_t* nfallback,
                    int grid, hipStream_t st) {
  if (!njobs) return;
  unsigned blocks = std::min<unsigned>((unsigned)grid, nblk(njobs, 4));
#define DMR_SWW(k, tr)                                                                             \
  hipLaunchKernelGGL((k_sw_wave<k, tr>), dim3(blocks), dim3(256), 0, st, jobs, list, njobs, q, db,  \
                     m32, gap_open, gap_extend, counter, outs, trace, trace_off, ops, ops_off,     \
                     fallback, nfallback)
  if (traceback) {
    switch (K) {
      case 1: DMR_SWW(1, true); break;
      case 2: DMR_SWW(2, true); break;
      case 4: DMR_SWW(4, true); break;
      case 8: DMR_SWW(8, true); break;
      case 16: DMR_SWW(16, true); break;
      case 32: DMR_SWW(32, true); break;
      default: break;
    }
  } else {
    switch (K) {
      case 1: DMR_SWW(1, false); break;
      case 2: DMR_SWW(2, false); break;
      case 4: DMR_SWW(4, false); break;
      case 8: DMR_SWW(8, false); break;
      case 16: DMR_SWW(16, false); break;
      case 32: DMR_SWW(32, false); break;
      default: break;
    }
  }
#undef DMR_SWW
}

}  // namespace dmr
