// Banded SW with one DP per wavefront: exact int32, any band width up to
// 64*K (K = 1..32, so up to 2048), optional trace and traceback.
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
// Trace cells are 4 bits, choice | e_ext << 2 | f_ext << 3, with the
// reference's strict D > E > F priority (src/gapped.rs:171-188), two per
// byte; the best cell is the first strict maximum in row-major order.
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

// Cross-lane moves in the VALU (DPP, GFX9 encodings) instead of the LDS
// crossbar: they sit on every row's critical path.
constexpr int kRowShr1 = 0x111, kRowShr2 = 0x112, kRowShr4 = 0x114, kRowShr8 = 0x118;
constexpr int kWaveShl1 = 0x130, kWaveShr1 = 0x138, kRowBcast15 = 0x142, kRowBcast31 = 0x143;
constexpr int kQuadIdentity = 0xE4;  // quad_perm [0,1,2,3]: a plain VALU copy

template <int kCtrl, int kRowMask = 0xf>
__device__ __forceinline__ int dpp(int old, int v) {
  return __builtin_amdgcn_update_dpp(old, v, kCtrl, kRowMask, 0xf, false);
}
// Lane l receives lane l+1's v; lane 63 receives `edge`.
__device__ __forceinline__ int from_next(int v, int edge) { return dpp<kWaveShl1>(edge, v); }
// Lane l receives lane l-1's v; lane 0 receives `edge`.
__device__ __forceinline__ int from_prev(int v, int edge) { return dpp<kWaveShr1>(edge, v); }
// Inclusive prefix maximum / minimum over the 64 lanes. The identity of the
// operation as the DPP `old` value lets each step fold into one max/min.
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
__device__ __forceinline__ int scan_min(int v) {
  constexpr int kId = INT32_MAX;
  v = min(v, dpp<kRowShr1>(kId, v));
  v = min(v, dpp<kRowShr2>(kId, v));
  v = min(v, dpp<kRowShr4>(kId, v));
  v = min(v, dpp<kRowShr8>(kId, v));
  v = min(v, dpp<kRowBcast15, 0xa>(kId, v));
  v = min(v, dpp<kRowBcast31, 0xc>(kId, v));
  return v;
}
__device__ __forceinline__ int wmax(int v) { return __builtin_amdgcn_readlane(scan_max(v), 63); }
__device__ __forceinline__ int wmin(int v) { return __builtin_amdgcn_readlane(scan_min(v), 63); }

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

__device__ __forceinline__ int letter_at(const uint8_t* ss, int m, int j) {
  return (j >= 1 && j <= m) ? (int)ss[j - 1] : kPad;
}

// Traceback walk of one DP by a whole wave (gapped::traceback,
// src/gapped.rs:205-279): every lane runs the same state machine on
// wave-uniform values, and the trace cells, query letters and subject letters
// it reads come from a wave-private LDS cache filled a block of kRows rows at
// a time by all lanes (coalesced), so a step costs one LDS read instead of a
// dependent global load. Lane 0 writes the run-length ops (op | count << 2,
// 0 = M, 1 = I, 2 = D, in path order) and the caller's statistics in o.
// Blocks have four waves (256 threads).
template <int K>
__device__ void trace_walk(int lane, int wave, const uint8_t* tr, int rlo, int d0, int d1, int m,
                           const uint8_t* qs, const uint8_t* ss, const int8_t* M, int bi, int bj,
                           uint32_t* op, SwOut& o) {
  constexpr int kRowBytes = 32 * K;       // one trace row (4 bits per cell)
  constexpr int kRows = K == 1 ? 64 : 128 / K;  // rows per block: 2-4 KB of trace
  constexpr int kCols = kRows + 64 * K;   // subject letters those rows' bands cover
  __shared__ uint32_t tcache[4][kRows * kRowBytes / 4];
  __shared__ uint8_t qcache[4][kRows];
  __shared__ uint8_t scache[4][kCols];
  const uint8_t* tc = reinterpret_cast<const uint8_t*>(tcache[wave]);
  const RowRange R{d0, d1, m};
  int r0 = INT32_MAX, c0 = 1;  // first cached row and column
  // Rows r0..i of the trace (all written: i <= bi), their query letters and
  // the subject letters of columns max(1, r0 - d1) .. min(m, i - d0).
  // All loads of a block are issued before its LDS writes (clamped, always
  // in-bounds addresses), so a block costs one memory latency. Column
  // bounds: c0 <= m (r0 <= istar = m + d1) and c1 >= c0 (r0 >= d0 + 1).
  constexpr int kTW = kRows * kRowBytes / 4 / 64;  // trace words per lane
  constexpr int kSW = (kCols + 63) / 64;           // subject letters per lane
  auto refill = [&](int i) {
    r0 = max(rlo, i - kRows + 1);
    const int words = (i - r0 + 1) * (kRowBytes / 4);
    const uint32_t* src = reinterpret_cast<const uint32_t*>(tr + (uint64_t)(r0 - rlo) * kRowBytes);
    uint32_t tv[kTW];
#pragma unroll
    for (int k = 0; k < kTW; ++k) tv[k] = src[min(lane + 64 * k, words - 1)];
    const int nq = i - r0 + 1;  // <= kRows <= 64
    const int qv = qs[r0 - 1 + min(lane, nq - 1)];
    c0 = max(1, r0 - d1);
    const int ns = min(m, i - d0) - c0 + 1;
    int sv[kSW];
#pragma unroll
    for (int k = 0; k < kSW; ++k) sv[k] = ss[c0 - 1 + min(lane + 64 * k, ns - 1)];
#pragma unroll
    for (int k = 0; k < kTW; ++k)
      if (lane + 64 * k < words) tcache[wave][lane + 64 * k] = tv[k];
    if (lane < nq) qcache[wave][lane] = (uint8_t)qv;
#pragma unroll
    for (int k = 0; k < kSW; ++k)
      if (lane + 64 * k < ns) scache[wave][lane + 64 * k] = (uint8_t)sv[k];
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");  // the cache is read by every lane
  };
  uint32_t nr = 0, cur = 3, cnt = 0;
  int state = 0, i = bi, j = bj;
  bool in_gap = false;
  auto emit = [&](uint32_t v) {
    if (lane == 0) op[nr] = v;
    ++nr;
  };
  auto push = [&](uint32_t code, uint32_t n) {
    if (code == cur) {
      cnt += n;
    } else {
      if (cnt) emit(cur | (cnt << 2));
      cur = code;
      cnt = n;
    }
  };
  // A cached trace cell (row ik >= r0 of the cache, in the band).
  auto cell = [&](int ik, int jk) {
    const int t = jk - ik + d1;  // band offset: nibble t % 2 of byte t / 2
    return (int)(tc[(ik - r0) * kRowBytes + (t >> 1)] >> (4 * (t & 1))) & 15;
  };
  // The reference's loop (src/gapped.rs:222-265), a run of steps at a time:
  // lane k looks at step k of the current state's direction (a diagonal run
  // of matches in state 0, a vertical or horizontal gap in states 1 and 2),
  // each lane checks the loop condition and the row range at its own cell,
  // and one ballot gives the run's length. Lane 0's cell is the current one,
  // already checked.
  while (i > 0 && j > 0) {
    if (j < R.lo(i) || j > R.hi(i) || i < rlo) break;
    if (i < r0) refill(i);
    const int di = state == 2 ? 0 : 1, dj = state == 1 ? 0 : 1;
    // Steps that stay in the cached rows (a horizontal gap stays in row i).
    const int span = state == 2 ? 64 : min(64, i - r0 + 1);
    const int ik = i - lane * di, jk = j - lane * dj;
    const bool ok = lane < span && ik > 0 && jk > 0 && jk >= R.lo(ik) && jk <= R.hi(ik) && ik >= rlo;
    const int tv = ok ? cell(ik, jk) : 0;
    const uint64_t bad = __ballot(!ok);  // bit k: the loop stops before step k
    if (state == 0) {
      const int c = uniform(tv) & 3;  // lane 0's cell
      if (c == 0) break;
      if (c == 2 || c == 3) {
        state = c - 1;
        continue;
      }
      // Matches while the cells are in range and choose D.
      const uint64_t stopm = bad | __ballot(ok && (tv & 3) != 1);
      const int run = stopm ? __builtin_ctzll(stopm) : 64;
      const bool in_run = lane < run;
      int x = 0, y = 0;
      if (in_run) {
        x = (int)qcache[wave][ik - r0];
        y = (int)scache[wave][jk - c0];
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
      // Gap steps: each runs (its cell in range), and the first whose
      // extension bit is clear is the last of the gap.
      const int bit = state == 1 ? 2 : 3;
      const uint64_t last = __ballot(ok && ((tv >> bit) & 1) == 0);
      const int f = bad ? __builtin_ctzll(bad) : 64;     // steps before the loop stops
      const int z = last ? __builtin_ctzll(last) : 64;   // the gap's last step
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
  if (cnt) emit(cur | (cnt << 2));
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
}

template <int K, bool kTrace>
__global__ __launch_bounds__(256) void k_sw_wave(const SwJob* __restrict__ jobs,
                                                 const uint32_t* __restrict__ list, uint32_t njobs,
                                                 SeqSet q, SeqSet db, const int8_t* __restrict__ m32,
                                                 int gap_open, int gap_extend,
                                                 uint32_t* __restrict__ counter,
                                                 SwOut* __restrict__ outs, int32_t* __restrict__ scores,
                                                 uint8_t* __restrict__ trace,
                                                 const uint64_t* __restrict__ trace_off,
                                                 uint32_t* __restrict__ ops,
                                                 const uint64_t* __restrict__ ops_off,
                                                 uint32_t* __restrict__ fallback,
                                                 uint32_t* __restrict__ nfallback,
                                                 const int32_t* __restrict__ known,
                                                 const uint32_t* __restrict__ njobs_dev, uint32_t list_sel) {
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
    if (task >= njobs || (njobs_dev && task >= uniform(*njobs_dev))) break;
    uint32_t entry = uniform(list[task]);
    if (list_sel) {  // a list of tagged entries (bit 31): take one kind only
      if ((entry >> 31) != (list_sel == 2 ? 1u : 0u)) continue;
      entry &= 0x7FFFFFFFu;
    }
    const uint32_t jid = entry;
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

    // Band offset t = lane*K + k is column j = i - d1 + t of row i. Offsets
    // past the band (t >= W) hold H = E = NEG: the wall right of the band.
    // S[k] is the subject letter of column j (padding letter outside 1..m);
    // it moves one offset left per row, so each row loads one new letter.
    int Hp[K], Ep[K], S[K];
    bool valid[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int t = lane * K + k;
      valid[k] = t < W;
      Hp[k] = valid[k] ? 0 : kNeg;
      Ep[k] = kNeg;
      S[k] = letter_at(ss, m, rlo - d1 + t);
    }
    // lbest: the lane's largest H so far; in trace mode li/lt hold the
    // first cell (row, band offset) where it occurs (rows in order, offsets
    // in order within a row).
    int best = 0, lbest = 0, li = 0, lt = 0, bi = 0, bj = 0;
    uint8_t* tr = kTrace ? trace + trace_off[task] : nullptr;
    // Per-row letters in blocks of 64 rows, one row per lane: the query
    // letter of row i and the subject letter entering the window after row i
    // (column i + 1 - d1 + 64K - 1). The next block is loaded a block ahead,
    // so no row waits on memory.
    // The loads are unconditional (clamped addresses, padding selected
    // afterwards) and each block is moved out of the load registers by a
    // VALU copy at the block start, so the row loop itself never waits on
    // memory. Needs rlo <= rhi and m >= 1, which the row loop implies below.
    auto qletters = [&](int r0) { return (int)qs[min(r0 + lane, rhi) - 1]; };
    auto eletters = [&](int r0) {
      const int j = r0 + lane + 64 * K - d1;
      const int v = ss[min(max(j, 1), m) - 1];
      return (j >= 1 && j <= m) ? v : kPad;
    };
    const bool rows = rlo <= rhi && m >= 1;
    int qn = rows ? qletters(rlo) : 0, en = rows ? eletters(rlo) : 0;
    // Trace mode with the DP's best score known (from the score pass): the
    // first cell that reaches it is the first maximum in row-major order, so
    // the rows after its row cannot change the result and are not run.
    const int stop_at = (kTrace && known) ? uniform(known[jid]) : INT32_MAX;
    bool stop = false;
    for (int blk = rlo; rows && blk <= rhi && !stop; blk += 64) {
     const int qv = dpp<kQuadIdentity>(0, qn), ev = dpp<kQuadIdentity>(0, en);
     qn = qletters(blk + 64);
     en = eletters(blk + 64);
     const int rend = min(rhi, blk + 63);
     for (int i = blk; i <= rend; ++i) {
      const int qrow = __builtin_amdgcn_readlane(qv, i - blk) * kScoreStride;
      const int edge = __builtin_amdgcn_readlane(ev, i - blk);
      // E sources of the last owned offset: the next lane's first offset in
      // the previous row (NEG beyond lane 63).
      const int nH = from_next(Hp[0], kNeg), nE = from_next(Ep[0], kNeg);
      int Hq[K], E[K], D[K];
      uint32_t eext = 0;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        D[k] = Hp[k] + M[qrow + S[k]];
        const int hs = k + 1 < K ? Hp[k + 1] : nH;
        const int es = k + 1 < K ? Ep[k + 1] : nE;
        // E opens from H or extends E of (i-1, j); with no source in the
        // band both terms are below NEG/2 and E stays undefined in effect
        // (it never wins a max against D >= -64 or 0). The extension bit is
        // only read on cells whose E is defined.
        const int eo = hs - oc, ee = es - ge;
        E[k] = max(eo, ee);
        eext |= (uint32_t)(ee > eo) << k;
        Hq[k] = max(max(D[k], E[k]), 0);
      }
      // Horizontal gaps: F(t) = max_{t'<t} H'(t') - o - e*(t-1-t'), with
      // H' = max(0, D, E). Per lane the best source for the cell after its
      // last offset, then an exclusive prefix maximum across lanes.
      int L = Hq[0] - oc - ge * (K - 1);
#pragma unroll
      for (int k = 1; k < K; ++k) L = max(L, Hq[k] - oc - ge * (K - 1 - k));
      const int stepK = ge * K;
      const int inc = scan_max(L + stepK * lane);
      const int exc = from_prev(inc, kNeg);
      int F[K];
      F[0] = exc - stepK * (lane - 1);
      uint32_t fext = 0;
#pragma unroll
      for (int k = 1; k < K; ++k) {
        const int fo = Hq[k - 1] - oc, fe = F[k - 1] - ge;
        F[k] = max(fo, fe);
        fext |= (uint32_t)(fe > fo) << k;
      }
      {
        // Extension bit of the first owned cell: F and H' of the left cell.
        const int fo = from_prev(Hq[K - 1], 1 << 29) - oc;
        const int fe = from_prev(F[K - 1], kNeg) - ge;
        fext |= (uint32_t)(fe > fo);
      }
      int rowbest = 0;
      uint32_t tw[(K + 7) / 8];  // trace nibbles of the lane's offsets (offset k at bit 4k)
#pragma unroll
      for (int w = 0; w < (K + 7) / 8; ++w) tw[w] = 0;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        int hv;
        if (kTrace) {
          // Strict priority D > E > F from 0 (src/gapped.rs:171-184).
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
          const uint32_t nib = (uint32_t)choice | (((eext >> k) & 1u) << 2) | (((fext >> k) & 1u) << 3);
          tw[k / 8] |= nib << (4 * (k % 8));
        } else {
          hv = max(Hq[k], F[k]);
        }
        Hp[k] = valid[k] ? hv : kNeg;
        Ep[k] = valid[k] ? E[k] : kNeg;
        rowbest = max(rowbest, Hp[k]);
      }
      if (kTrace) {
        // Row i of the trace is 32K bytes, 4 bits per cell: offset t in the
        // low (t even) or high nibble of byte t / 2. Every lane stores its
        // K / 2 bytes at once (with one lane, an even lane stores its own
        // nibble and its neighbour's), and a wave writes whole lines.
        uint8_t* row = tr + (uint64_t)(i - rlo) * (32 * K);
        if (K == 1) {
          const uint32_t pair = tw[0] | ((uint32_t)from_next((int)tw[0], 0) << 4);
          if ((lane & 1) == 0) row[lane >> 1] = (uint8_t)pair;
        } else if (K == 2) {
          row[lane] = (uint8_t)tw[0];
        } else if (K == 4) {
          reinterpret_cast<uint16_t*>(row)[lane] = (uint16_t)tw[0];
        } else {
#pragma unroll
          for (int w = 0; w < K / 8; ++w) reinterpret_cast<uint32_t*>(row + lane * (K / 2))[w] = tw[w];
        }
      }
      // Subject letters for the next row (the next lane's first letter is
      // read before the shift overwrites it).
      const int sn = from_next(S[0], edge);
#pragma unroll
      for (int k = 0; k + 1 < K; ++k) S[k] = S[k + 1];
      S[K - 1] = sn;
      if (kTrace) {
        // Each lane keeps its own first maximum (no cross-lane work per
        // row); the wave's first maximum in row-major order is reduced from
        // them after the last row. Once some lane holds the known best, the
        // first maximum lies in this row or an earlier one.
        if (rowbest > lbest) {
          int tf = 0;
#pragma unroll
          for (int k = K - 1; k >= 0; --k)
            if (valid[k] && Hp[k] == rowbest) tf = k;
          lbest = rowbest;
          li = i;
          lt = lane * K + tf;
        }
        if (known && __ballot(lbest >= stop_at) != 0) {
          stop = true;
          break;
        }
      } else {
        lbest = max(lbest, rowbest);
      }
     }
    }
    // Clamped rows below the band: rows istar+1..n hold the single cell
    // (i, m), whose diagonal predecessor lies outside the previous row's
    // range (it counts as 0) and which has no horizontal source. Their H is
    // at most max(11, H(istar, m) - o), so they can only raise a best score
    // below 11, the largest matrix entry. Score mode finishes them here from
    // H and E of (istar, m), held by offset 0 after row istar; trace mode and
    // bands that never reach row istar go to the row-exact kernel.
    best = wmax(lbest);
    if (kTrace && best > 0) {
      // First maximum in row-major order: the smallest (row, offset) among
      // the lanes holding the wave's largest H.
      bi = wmin(lbest == best ? li : INT32_MAX);
      bj = bi - d1 + wmin(lbest == best && li == bi ? lt : INT32_MAX);
    }
    bool need_exact = best < 11 && istar < n;
    if (!kTrace) {
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
    if constexpr (kTrace) {
      if (!need_exact) {  // the whole wave walks the trace back
        SwOut o = {};
        o.score = best;
        o.best_i = (uint32_t)bi;
        o.best_j = (uint32_t)bj;
        if (best > 0)
          trace_walk<K>(lane, threadIdx.x / 64, tr, rlo, d0, d1, m, qs, ss, M, bi, bj, ops + ops_off[task], o);
        if (lane == 0) outs[task] = o;
      }
    }
    if (lane == 0) {
      if (need_exact) {
        fallback[atomicAdd(nfallback, 1u)] = kTrace ? task : jid;
        if (kTrace) outs[task] = SwOut{};  // no ops: the row-exact kernel redoes this DP
      } else if (!kTrace) {
        scores[jid] = best;
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

// Run-length ops of traceback DPs, packed: sizes[t] = n_ops of DP t
// (sizes[n] = 0, so an exclusive scan gives the packed offsets), then each
// DP's ops move from its capacity slot (ops_off, n + 1 offsets) to its packed
// slot, one wave per DP. A count never exceeds the DP's capacity slot.
__device__ __forceinline__ uint32_t ops_count(const SwOut* outs, const uint64_t* ops_off, uint32_t t) {
  const uint64_t cap = ops_off[t + 1] - ops_off[t];
  const uint32_t c = outs[t].n_ops;
  return c < cap ? c : (uint32_t)cap;
}

__global__ void k_ops_sizes(const SwOut* __restrict__ outs, const uint64_t* __restrict__ ops_off, uint32_t n,
                            uint32_t* __restrict__ sizes) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) sizes[t] = ops_count(outs, ops_off, t);
  else if (t == n) sizes[n] = 0;
}

__global__ void k_ops_pack(const SwOut* __restrict__ outs, const uint32_t* __restrict__ ops,
                           const uint64_t* __restrict__ ops_off, const uint64_t* __restrict__ pack_off,
                           uint32_t n, uint32_t* __restrict__ packed) {
  const uint32_t t = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  if (t >= n) return;
  const uint32_t cnt = ops_count(outs, ops_off, t);
  const uint32_t* src = ops + ops_off[t];
  uint32_t* dst = packed + pack_off[t];
  for (uint32_t k = threadIdx.x & 63; k < cnt; k += 64) dst[k] = src[k];
}

void launch_ops_pack(const SwOut* outs, const uint32_t* ops, const uint64_t* ops_off, uint32_t n,
                     uint32_t* sizes, uint64_t* pack_off, uint64_t* scan_tmp, uint32_t* packed,
                     hipStream_t st) {
  if (!n) return;
  hipLaunchKernelGGL(k_ops_sizes, dim3(nblk(n + 1, 256)), dim3(256), 0, st, outs, ops_off, n, sizes);
  scan_exclusive_u32_to_u64(sizes, pack_off, n + 1, scan_tmp, st);
  hipLaunchKernelGGL(k_ops_pack, dim3(nblk(n, 4)), dim3(256), 0, st, outs, ops, ops_off, pack_off, n, packed);
}

void launch_sw_wave(int K, bool traceback, const SwJob* jobs, const uint32_t* list, uint32_t njobs,
                    SeqSet q, SeqSet db, const int8_t* m32, int gap_open, int gap_extend,
                    uint32_t* counter, SwOut* outs, int32_t* scores, uint8_t* trace, const uint64_t* trace_off,
                    uint32_t* ops, const uint64_t* ops_off, uint32_t* fallback, uint32_t* nfallback,
                    int cu_count, hipStream_t st, const int32_t* known, const uint32_t* njobs_dev,
                    uint32_t list_sel) {
  if (!njobs) return;
#define DMR_SWW(k, tr)                                                                             \
  do {                                                                                             \
    static const unsigned resident = resident_blocks(k_sw_wave<k, tr>, 256, cu_count);            \
    const unsigned blocks = std::min<unsigned>(resident, nblk(njobs, 4));                          \
    hipLaunchKernelGGL((k_sw_wave<k, tr>), dim3(blocks), dim3(256), 0, st, jobs, list, njobs, q,   \
                       db, m32, gap_open, gap_extend, counter, outs, scores, trace, trace_off, ops,\
                       ops_off, fallback, nfallback, known, njobs_dev, list_sel);                  \
  } while (0)
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
