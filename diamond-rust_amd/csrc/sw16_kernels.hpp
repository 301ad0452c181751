// Packed 16-bit banded SW score kernels (k_sw16, k_sw16p), shared by the
// translation units that instantiate them: k_sw16.hip (the band_for widths)
// and k_sw16m.hip (merged bands).
#pragma once

// Banded SW score pass in packed int16: two DPs per lane, band in registers.
//
// Each lane carries two DPs with the same row sequence (lo/hi halves of every
// VGPR: two DPs of one query, or, with the subject on the rows, of one
// subject; see subject_rows) and the same band width W (a compile-time
// constant), so every recurrence step is one packed instruction for two
// cells. The band of a row lives in registers indexed by the in-row offset t
// (column j = i - d1 + t), and the dependencies of gapped::extend_span
// (src/gapped.rs:110-199) become
//   D(i,t) = H(i-1,t) + s,  E(i,t) = max(H(i-1,t+1) - o, E(i-1,t+1) - e),
//   F(i,t) = max(H(i,t-1) - o, F(i,t-1) - e),  H = max(0, D, E, F),
// with o = open + extend. G = H - o is stored instead of H (the score table
// holds s + o), which saves one subtraction per cell pair.
//
// Row frames: the values of row step r are held plus r (G plus r + 1), so a
// vertical move gains exactly the gap extension and E needs no decrement:
//   E^(r,t) = max(r, G^(r-1,t+1), E^(r-1,t+1)),  D^ = G^(r-1,t) + s + o,
//   F^(r,t) = max(G^(r,t-1), F^(r,t-1)) - 1,     H^ = max(D^, E^, F^),
//   G^ = H^ - o + 1,
// with the zero of the local alignment at r. The zero rides on E: carrying
// E' = max(E, 0) instead of E changes no H (H = max(0, D, E, F) =
// max(D, E', F)) and no later E' (max(E', 0) - e <= max(E - e, 0)), and
// makes H one three-way max. Every kFrame rows the frame is moved back by
// kFrame (all carried values and r), which bounds the offset.
//
// Exactness against the row-range rules of the reference:
// - Columns j <= 0 and rows above the band read a padding letter scoring -64:
//   those cells stay at H = 0, which is what the reference computes there
//   (column 0 is always 0; E/F values of -o or NEG never change H).
// - Columns j > m (past the subject) and query rows past n read the padding
//   letter too: such cells only feed cells outside the DP and stay strictly
//   below some in-range H, so they never change the maximum.
// - The column after the band (t = W) is a wall: its G and E stay NEG, so the
//   rightmost cell of a row gets no E from outside the band.
// - Rows past m + d1 hold a single cell at column m (the clamped rows of the
//   reference). Their H is at most max(11, H(m + d1, m) - o), so they only
//   matter when the best score is below 11: such DPs are re-run in int32.
// DPs scoring kRedo or more (int16 saturation, f16 exactness) are re-run in
// int32.
#include <type_traits>

#include "dmr_kernels.hpp"

namespace dmr {

namespace {

constexpr int kCodes = 26;                  // 25 letters + padding
constexpr int kPairRow = kCodes * kCodes;   // (sA, sB) pairs for one query letter
constexpr int kPairTab = kCodes * kPairRow;
constexpr short kNeg16 = -16384;
constexpr int kOpen = 12;                   // gap open + extend of the 16-bit tables
constexpr int kFrame = 64;                  // rows between frame moves
// The pair table in LDS, typed so that a task body compiled as a function of
// its own (the fused kernel) still reads it with LDS instructions.
typedef const __attribute__((address_space(3))) int LdsInt;
typedef const __attribute__((address_space(3))) char LdsChar;
// Letters in global memory, typed likewise (global loads, not flat ones,
// whose completion the LDS waits of the row would also have to cover).
typedef const __attribute__((address_space(1))) uint8_t GlobalByte;

// Arithmetic of the packed kernels, two DPs per register. A value in the
// frame of its row is the true value plus at most kFrame, and no value exceeds
// the running best H of its DP in that frame.
// I16: int16 with saturating adds. Saturation (at 32767) leaves the best at
//   32767 - kFrame - 1 or more after the frame is taken off, so DPs from 32000
//   on are re-run in int32.
// F16: half precision. Every value the recurrence can produce is an integer
//   in [-o - 1, best + kFrame] (or below -16000 for the NEG sentinels, where
//   rounding is harmless), and integers up to 2048 are exact in f16. Every
//   value is at most the frame's running best, so the first value that
//   rounds leaves that best at 2048 or more and the final best at
//   2048 - kFrame or more: a DP whose result is below 1900 was scored
//   exactly; the rest are re-run in int32. max and max3 become
//   v_pk_maximum3_f16 (IEEE maximum: unlike maxNum it needs no canonicalized
//   operands; no NaN ever arises).
struct I16 {
  typedef short V __attribute__((ext_vector_type(2)));
  static constexpr int kRedo = 32000;
  static __device__ __forceinline__ V bits(int x) { return __builtin_bit_cast(V, x); }
  static __device__ __forceinline__ int raw(V v) { return __builtin_bit_cast(int, v); }
  static __device__ __forceinline__ V splat(int x) { return V{(short)x, (short)x}; }
  static __device__ __forceinline__ V pair(int a, int b) { return V{(short)a, (short)b}; }
  static __device__ __forceinline__ V add(V a, V b) { return __builtin_elementwise_add_sat(a, b); }
  static __device__ __forceinline__ V sub(V a, V b) { return __builtin_elementwise_sub_sat(a, b); }
  static __device__ __forceinline__ V max(V a, V b) { return __builtin_elementwise_max(a, b); }
  static __device__ __forceinline__ V max3(V a, V b, V c) { return max(max(a, b), c); }
  static __device__ __forceinline__ V min(V a, V b) { return __builtin_elementwise_min(a, b); }
  typedef short S;  // one half
  static __device__ __forceinline__ S hmax(V v) { return v.x > v.y ? v.x : v.y; }
  static __device__ __forceinline__ int lo(V v) { return v.x; }
  static __device__ __forceinline__ int hi(V v) { return v.y; }
};
struct F16 {
  typedef _Float16 V __attribute__((ext_vector_type(2)));
  static constexpr int kRedo = 1900;
  static __device__ __forceinline__ V bits(int x) { return __builtin_bit_cast(V, x); }
  static __device__ __forceinline__ int raw(V v) { return __builtin_bit_cast(int, v); }
  static __device__ __forceinline__ V splat(int x) { return V{(_Float16)x, (_Float16)x}; }
  static __device__ __forceinline__ V pair(int a, int b) { return V{(_Float16)a, (_Float16)b}; }
  static __device__ __forceinline__ V add(V a, V b) { return a + b; }
  static __device__ __forceinline__ V sub(V a, V b) { return a - b; }
  static __device__ __forceinline__ V max(V a, V b) { return __builtin_elementwise_maximum(a, b); }
  static __device__ __forceinline__ V max3(V a, V b, V c) { return max(max(a, b), c); }
  static __device__ __forceinline__ V min(V a, V b) { return __builtin_elementwise_minimum(a, b); }
  typedef _Float16 S;  // one half
  static __device__ __forceinline__ S hmax(V v) { return __builtin_elementwise_maximum(v.x, v.y); }
  static __device__ __forceinline__ int lo(V v) { return (int)v.x; }
  static __device__ __forceinline__ int hi(V v) { return (int)v.y; }
};

__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return v;
}

// Letter (times 4) at position p of a sequence of length m starting at
// pad[base] in a padded letter set: positions outside 0..m-1 clamp to the
// padding code on either side, so one clamp replaces the range tests.
// Sequences are addressed by 32-bit offsets from the kernel-argument base,
// which keeps per-lane state to one register per sequence.
template <class Bytes>
__device__ __forceinline__ int pad_at(Bytes* pad, uint32_t base, int m, int p) {
  return (int)pad[base + (uint32_t)min(max(p, -1), m)];
}
// Byte offset of the score-table row of a query letter (times 4). The empty
// asm makes the value opaque, so each cell's table address stays one add of
// a window half to it (the compiler would otherwise fold the multiply into a
// multiply-add per cell, after extracting the half).
__device__ __forceinline__ uint32_t row_base(int q4) {
  uint32_t b = __umul24((uint32_t)q4, (uint32_t)kPairRow);
  asm("" : "+v"(b));
  return b;
}
// Byte offset of the letter pair (a, b) in a score-table row, from the
// letters times 4: 4 * (a * kCodes + b), one 24-bit multiply-add.
__device__ __forceinline__ uint32_t pair_off4(int a4, int b4) {
  static_assert(kCodes == 26, "the multiplier is written out below");
  uint32_t r;
  asm("v_mad_u32_u24 %0, %1, 26, %2" : "=v"(r) : "v"(a4), "v"(b4));
  return r;
}
// A loaded letter, settled at this point: the load completes here (at the
// end of the row that issued it) instead of being moved to its use in the
// next row, so no row waits on memory.
__device__ __forceinline__ int settle(int v) {
  asm volatile("" : "+v"(v));
  return v;
}

// Orientation of a packed DP. The score pass needs only a band's best score,
// and that does not depend on which of the two sequences runs along the rows:
// the band holds the same cells, and the recurrence and its boundary are
// symmetric in the two sequences (E and F are the two gap kinds with the same
// costs; outside the band or the matrix a D neighbour is 0 and an E or F
// neighbour absent). The packed kernels step rows and keep a row's W band
// offsets in registers, so a DP costs rows x W cells: with the query on the
// rows the rows are max(1, d0 + 1) .. min(n, m + d1), with the subject on the
// rows (diagonals -d1 .. -d0) max(1, 1 - d1) .. min(m, n - d0). Band cells
// past either end of the column sequence are computed as padding, so the
// orientation with fewer rows computes fewer of them. (The first maximum in
// row-major order, where the traceback starts, does depend on it; the
// traceback runs in the reference's orientation.)
// mode: 0 query rows, 1 the orientation with fewer rows (ties: query rows),
// 2 subject rows.
__device__ __forceinline__ bool subject_rows(int64_t n, int64_t m, int64_t d0, int64_t d1, int mode) {
  if (mode != 1) return mode == 2;
  const int64_t rq = min(n, m + d1) - max((int64_t)1, d0 + 1);
  const int64_t rs = min(m, n - d0) - max((int64_t)1, 1 - d1);
  return rs < rq;
}

// A lane's two DPs in the orientation of its task: n the length of the row
// sequence (one for both DPs: the query, or with tr the subject), mA/mB the
// lengths of the column sequences, rp/ap/bp their padded-letter bases (base
// 1: an empty sequence for an absent DP), the bands in that orientation, and
// whether the reference's clamped rows exist (rows past m + d1 with the query
// on the rows; one cell at the subject's end each), which the 16-bit pass
// leaves to the exact re-run.
struct LanePair {
  int n, mA, mB;
  uint32_t rp, ap, bp;
  int d0A, d1A, d0B, d1B;
  bool clampA, clampB;
};
__device__ __forceinline__ LanePair lane_pair(bool tr, bool hasA, bool hasB, const SwJob& ja, const SwJob& jb,
                                              const SeqSet& q, const SeqSet& db) {
  const SeqSet& R = tr ? db : q;
  const SeqSet& C = tr ? q : db;
  const uint32_t rA = tr ? ja.subject : ja.target;
  const uint32_t cA = tr ? ja.target : ja.subject, cB = tr ? jb.target : jb.subject;
  LanePair lp;
  lp.n = hasA ? (int)(R.offs[rA + 1] - R.offs[rA]) : 0;
  lp.mA = hasA ? (int)(C.offs[cA + 1] - C.offs[cA]) : 0;
  lp.mB = hasB ? (int)(C.offs[cB + 1] - C.offs[cB]) : 0;
  lp.rp = hasA ? R.pad_offs[rA] : 1u;
  lp.ap = hasA ? C.pad_offs[cA] : 1u;
  lp.bp = hasB ? C.pad_offs[cB] : 1u;
  lp.d0A = tr ? -ja.d1 : ja.d0;
  lp.d1A = tr ? -ja.d0 : ja.d1;
  lp.d0B = hasB ? (tr ? -jb.d1 : jb.d0) : lp.d0A;
  lp.d1B = hasB ? (tr ? -jb.d0 : jb.d1) : lp.d1A;
  // Query rows: m + d1 < n. Subject rows: the query is the column sequence
  // and d1 = -d0 of this orientation.
  lp.clampA = tr ? lp.n - lp.d0A < lp.mA : lp.mA + lp.d1A < lp.n;
  lp.clampB = tr ? lp.n - lp.d0B < lp.mB : lp.mB + lp.d1B < lp.n;
  return lp;
}

template <int W, class Tr>
__global__ __launch_bounds__(256) void k_sw16(const uint2* __restrict__ lanes, uint32_t nlanes, uint32_t nrow,
                                              const SwJob* __restrict__ jobs, SeqSet q, SeqSet db,
                                              const int* __restrict__ pair_tab,
                                              const int8_t* __restrict__ m32,
                                              uint32_t* __restrict__ counter,
                                              int32_t* __restrict__ scores,
                                              uint32_t* __restrict__ redo, uint32_t* __restrict__ nredo) {
  __shared__ int T[kPairTab];
  for (int k = threadIdx.x; k < kPairTab; k += blockDim.x) T[k] = pair_tab[k];
  __syncthreads();
  const int lane = threadIdx.x & 63;
  using V = typename Tr::V;
  const V NEG2 = Tr::splat(kNeg16);
  const V ONE2 = Tr::splat(1);
  const V GOFF = Tr::splat(1 - kOpen);
  const V FRAME2 = Tr::splat(kFrame);
  const V ZERO2 = Tr::splat(0);
  // Lanes [0, nrow) run with the query on the rows, lanes [nrow, nlanes) with
  // the subject (subject_rows); no task mixes the two.
  const uint32_t trow = (nrow + 63) / 64, ntasks = trow + (nlanes - nrow + 63) / 64;
  for (;;) {
    // Branch-free task fetch (every lane joins the atomic, lane 0 adds 1):
    // the task lives in a scalar register and the loop exit is uniform.
    const uint32_t task =
        (uint32_t)__builtin_amdgcn_readfirstlane((int)atomicAdd(counter, lane == 0 ? 1u : 0u));
    if (task >= ntasks) break;
    const bool tr = task >= trow;
    const uint32_t L = (tr ? nrow + (task - trow) * 64 : task * 64) + lane;
    const bool active = L < (tr ? nlanes : nrow);
    uint2 ln = active ? lanes[L] : make_uint2(~0u, ~0u);
    const bool hasA = ln.x != ~0u, hasB = ln.y != ~0u;
    SwJob ja = {}, jb = {};
    if (hasA) ja = jobs[ln.x];
    if (hasB) jb = jobs[ln.y];
    const LanePair lp = lane_pair(tr, hasA, hasB, ja, jb, q, db);
    const uint8_t* rpad = tr ? db.pad : q.pad;  // row and column letters
    const uint8_t* cpad = tr ? q.pad : db.pad;
    const int n = lp.n, mA = lp.mA, mB = lp.mB;
    const uint32_t qp = lp.rp, ap = lp.ap, bp = lp.bp;
    const int d1A = lp.d1A, d1B = lp.d1B, d0A = lp.d0A, d0B = lp.d0B;
    // Rows of this lane: the union of its two DPs' row ranges, rows 1..m+d1
    // (the clamped rows below, one cell at column m each, are left to the
    // exact re-run when they can matter). Each lane runs its own rows (row
    // i = lo + r at step r); the wave steps as often as its longest lane
    // needs. Before its first row a lane would only see padding cells (H = 0,
    // E = -o), which the initial state reproduces after one row; after its
    // last row it sees padding cells, which stay below an earlier H.
    int lo = 0, hi = -1;
    if (hasA) {
      lo = max(1, d0A + 1);
      hi = min(n, mA + d1A);
    }
    if (hasB) {
      lo = min(lo, max(1, d0B + 1));
      hi = max(hi, min(n, mB + d1B));
    }
    const int steps = __builtin_amdgcn_readfirstlane(wave_max(hi - lo + 1));

    // Window of (sA, sB) letter pairs for the current row, as byte offsets
    // into a score-table row, two 16-bit entries per register: entry t is
    // column j = i - d1 + t (t even: low half, t odd: high half).
    constexpr int NR = (W + 1) / 2;
    V G[W + 1], E[W + 1];
    uint32_t R[NR];
#pragma unroll
    for (int t = 0; t <= W; ++t) {
      G[t] = t < W ? Tr::splat(-kOpen) : NEG2;
      E[t] = NEG2;
    }
#pragma unroll
    for (int k = 0; k < NR; ++k) {
      uint32_t lo16 = pair_off4(pad_at(cpad, ap, mA, lo - d1A + 2 * k - 1),
                                pad_at(cpad, bp, mB, lo - d1B + 2 * k - 1));
      uint32_t hi16 = 2 * k + 1 < W ? pair_off4(pad_at(cpad, ap, mA, lo - d1A + 2 * k),
                                                pad_at(cpad, bp, mB, lo - d1B + 2 * k))
                                    : 0u;
      R[k] = lo16 | (hi16 << 16);
    }
    // Per-row letters: the query letter of row i and the letter pair entering
    // the window at t = W-1 (subject positions i - d0 - 1).
    // Query letter of row i (padding past n), times 4.
    auto qletter = [&](int i) { return (int)rpad[qp + (uint32_t)min(i - 1, n)]; };
    auto pair_at = [&](int i) {
      return pair_off4(pad_at(cpad, ap, mA, i - d0A - 1), pad_at(cpad, bp, mB, i - d0B - 1));
    };
    // Widths up to 81 load them one row ahead (the clamped addresses are
    // always in bounds); the raw bytes are only read by a VALU copy at the
    // end of the row, so the loads are unconditional and no row waits on
    // memory. W = 129 has no registers to spare for this (it would spill to
    // scratch) and loads them at the start of each row.
    constexpr bool kPrefetch = W < 129;
    int qcur = qletter(lo);
    uint32_t pnext = 0;
    V best = ZERO2;  // in the frame of the current row
    V Zr = ZERO2;    // frame offset of the current row
    for (int r = 0; r < steps; ++r) {
      const int i = lo + r;
      if (r > 0) {
        // Shift the window by one entry and append the pair entering at t = W-1.
        static_assert(W % 2 == 1, "band widths are odd");
#pragma unroll
        for (int k = 0; k < NR - 1; ++k) R[k] = __builtin_amdgcn_alignbit(R[k + 1], R[k], 16);
        if (kPrefetch) {
          R[NR - 1] = pnext;
        } else {
          qcur = qletter(i);
          R[NR - 1] = pair_at(i);
        }
      }
      int rq = 0, ra = 0, rb = 0;
      if (kPrefetch) {  // raw bytes of row i + 1
        rq = qletter(i + 1);
        ra = pad_at(cpad, ap, mA, i - d0A);
        rb = pad_at(cpad, bp, mB, i - d0B);
      }
      const char* Tq = reinterpret_cast<const char*>(T) + row_base(qcur);
      auto score = [&](int t) {
        const uint32_t off = (t & 1) ? (R[t >> 1] >> 16) : (R[t >> 1] & 0xFFFFu);
        return *reinterpret_cast<const int*>(Tq + off);
      };
      // Score-table reads run kAhead columns ahead of their use, with a
      // scheduling fence after each so the compiler cannot sink them back to
      // their use: the LDS latency overlaps the recurrence even at one wave
      // per SIMD.
      constexpr int kAhead = 4;
      int sq[kAhead];
#pragma unroll
      for (int t = 0; t < kAhead; ++t) sq[t] = score(t);
      V F = NEG2, Gl = NEG2;
#pragma unroll
      for (int t = 0; t < W; ++t) {
        const V sc = Tr::bits(sq[t % kAhead]);
        if (t + kAhead < W) sq[t % kAhead] = score(t + kAhead);
        __builtin_amdgcn_sched_barrier(0);
        // D and E (with the zero) are off the F chain that runs along the row.
        const V D = Tr::add(G[t], sc);
        const V Ev = Tr::max3(G[t + 1], E[t + 1], Zr);
        const V Fv = Tr::sub(Tr::max(Gl, F), ONE2);
        const V H = Tr::max3(D, Ev, Fv);
        best = Tr::max(best, H);
        Gl = Tr::add(H, GOFF);
        G[t] = Gl;
        E[t] = Ev;
        F = Fv;
      }
      if (kPrefetch) {  // letters of row i + 1
        qcur = settle(rq);
        pnext = pair_off4(settle(ra), settle(rb));
      }
      // Frame of the next row; every kFrame rows the frame moves back.
      best = Tr::add(best, ONE2);
      Zr = Tr::add(Zr, ONE2);
      if ((r & (kFrame - 1)) == kFrame - 1) {
#pragma unroll
        for (int t = 0; t < W; ++t) {
          G[t] = Tr::sub(G[t], FRAME2);
          E[t] = Tr::sub(E[t], FRAME2);
        }
        best = Tr::sub(best, FRAME2);
        Zr = ZERO2;
      }
    }
    best = Tr::sub(best, Zr);
    // Re-run in int32 (exact) when int16 saturated, or when the score is
    // below 11 and the band reaches past the subject's end: the clamped rows
    // there could then hold the maximum.
    const int bA = Tr::lo(best), bB = Tr::hi(best);
    if (hasA) {
      if (bA >= Tr::kRedo || (bA < 11 && lp.clampA)) {
        redo[atomicAdd(nredo, 1u)] = ln.x;
      } else {
        scores[ln.x] = bA;
      }
    }
    if (hasB) {
      if (bB >= Tr::kRedo || (bB < 11 && lp.clampB)) {
        redo[atomicAdd(nredo, 1u)] = ln.y;
      } else {
        scores[ln.y] = bB;
      }
    }
  }
}

// Split-band variant: each DP pair is carried by P lanes (P = 2 or 4, one
// lane group), lane p of the group holding band offsets p*H .. p*H+H-1 with
// H = ceil((W + 1) / P); offsets >= W are the wall right of the band (and
// dead padding past it) and stay far below any score. A lane keeps H instead of W offsets, so
// the state fits far fewer registers and more waves share a SIMD.
//
// Within a row the halves of the band depend on each other only through the
// horizontal gap F entering a lane's first offset. Every lane first computes
// its offsets with no incoming F; then F_in of lane p is
//   F_in(p) = max_{q<p} (F_out(q) - (p-1-q) * H),
// F_out(q) being the F a lane passes to the offset after its last, an
// exclusive prefix maximum over the group. The lane is then corrected with
// F_true(t0+u) = max(F_local(t0+u), F_in - u): by induction H_true = max(H_local,
// F_in - u), and a gap reopened from a cell the chain raised (H - o) never
// beats the chain's own extension (- e). Only G = H - o and the best score
// carry to the next row, so the correction is G = max(G, F_in - o - u) and
// best = max(best, F_in); it is skipped on rows where no lane's F_in is
// positive (it cannot beat H >= 0 then). E crosses lanes from the previous
// row: a lane's last offset reads the next lane's first.
//
// With kWall the kernel serves any band width up to W (merged bands): each
// DP of a lane has its own right wall at its width W'. Cells at offsets
// t >= W' take G = H + kWallOff instead of H - o + 1 (a per-offset addend,
// so the wall costs registers, not instructions): their G stays far below any
// score, E there never rises above it (E only takes G from the row above,
// one offset right), so the last band cell reads no value from past the
// wall, and their H (at most a horizontal gap out of the band) never exceeds
// the best in the band. The F_in correction adds the same per-offset addend.
template <int W, int P>
struct SplitShape {
  static constexpr int H = (W + 1 + P - 1) / P;  // offsets per lane
  static constexpr int NR = H / 2 + 1;           // window registers (entries 0..H)
  static constexpr int kDeadFrom = W - (P - 1) * H;  // first dead offset of the last lane
};

constexpr int kWallOff = -30000;

// Waves per SIMD the register allocation is held to: one block per CU of
// kThreads threads (a block holds the 70 KB score table, so at most two fit
// a CU). 768 threads hold the allocation to 168 registers (three waves per
// SIMD; for the width-129 kernel the few values that then leave registers are
// per-DP setup, outside the row loop); 512 threads allow up to 256.
template <int kThreads>
constexpr int kMinWaves = kThreads / 256;

// One wave task of a split kernel: the kPer = 64 / P lanes (DP pairs) of
// task `task` of a class, lanes [0, nrow) with the query on the rows and
// [nrow, nlanes) with the subject; T is the block's score table in LDS.
template <int W, int P, class Tr, bool kWall>
__device__ __forceinline__ void sw16p_task(uint32_t task, const uint2* __restrict__ lanes, uint32_t nlanes,
                                           uint32_t nrow, const SwJob* __restrict__ jobs, SeqSet q, SeqSet db,
                                           LdsInt* T, int32_t* __restrict__ scores, uint32_t* __restrict__ redo,
                                           uint32_t* __restrict__ nredo) {
  static_assert(W % 2 == 1, "band widths are odd");
  static_assert(P == 2 || P == 4 || P == 8 || P == 16, "lane groups of 2, 4, 8 or 16");
  constexpr int H = SplitShape<W, P>::H;
  constexpr int NR = SplitShape<W, P>::NR;
  constexpr int kDeadFrom = SplitShape<W, P>::kDeadFrom;
  static_assert(kWall || (kDeadFrom >= 1 && kDeadFrom <= H), "the wall lies in the last lane");
  const int lane = threadIdx.x & 63;
  const int p = lane & (P - 1);
  const bool last = p == P - 1;
  const int t0 = p * H;  // band offset of this lane's first entry
  using V = typename Tr::V;
  const V NEG2 = Tr::splat(kNeg16);
  const V ONE2 = Tr::splat(1);
  const V GOFF = Tr::splat(1 - kOpen);
  const V FRAME2 = Tr::splat(kFrame);
  const V ZERO2 = Tr::splat(0);
  // Moves within a lane group: from the next lane and from the previous
  // lane (quad_perm [1,1,3,3] / [0,0,2,2] for pairs, [1,2,3,3] / [0,0,1,2]
  // for groups of 4, row_shl:1 / row_shr:1 for groups of 8 and 16; a lane
  // whose source lies outside its group receives a value it does not use),
  // and the swaps of the final reduction. mv leaves a lane without source
  // undefined (no copy of the old value is needed); mvk keeps its own value.
  constexpr int kFromNext = P == 2 ? 0xF5 : (P == 4 ? 0xF9 : 0x101);
  constexpr int kFromPrev = P == 2 ? 0xA0 : (P == 4 ? 0x90 : 0x111);
  auto dppi = [](int v, auto ctrl) { return __builtin_amdgcn_mov_dpp(v, decltype(ctrl)::value, 0xf, 0xf, false); };
  auto mv = [&](V v, auto ctrl) { return Tr::bits(dppi(Tr::raw(v), ctrl)); };
  auto mvk = [](V v, auto ctrl) {
    return Tr::bits(__builtin_amdgcn_update_dpp(Tr::raw(v), Tr::raw(v), decltype(ctrl)::value, 0xf, 0xf, false));
  };
  using NextC = std::integral_constant<int, kFromNext>;
  using PrevC = std::integral_constant<int, kFromPrev>;
  using Back2C = std::integral_constant<int, 0x40>;   // [0,0,0,1]
  using Shr2C = std::integral_constant<int, 0x112>;   // row_shr:2
  using Shr4C = std::integral_constant<int, 0x114>;   // row_shr:4
  using Shr8C = std::integral_constant<int, 0x118>;   // row_shr:8
  using RowMirrorC = std::integral_constant<int, 0x140>;  // row_mirror
  using SwapC = std::integral_constant<int, 0xB1>;    // [1,0,3,2]
  using Swap2C = std::integral_constant<int, 0x4E>;   // [2,3,0,1]
  using MirrorC = std::integral_constant<int, 0x141>; // row_half_mirror
  // Lanes [0, nrow) run with the query on the rows, lanes [nrow, nlanes) with
  // the subject (subject_rows); no task mixes the two.
  constexpr uint32_t kPer = 64 / P;  // lanes (DP pairs) per wave task
  const uint32_t trow = (nrow + kPer - 1) / kPer;
  const bool tr = task >= trow;
  const uint32_t L = (tr ? nrow + (task - trow) * kPer : task * kPer) + lane / P;
  const bool active = L < (tr ? nlanes : nrow);
  uint2 ln = active ? lanes[L] : make_uint2(~0u, ~0u);
  const bool hasA = ln.x != ~0u, hasB = ln.y != ~0u;
  SwJob ja = {}, jb = {};
  if (hasA) ja = jobs[ln.x];
  if (hasB) jb = jobs[ln.y];
  const LanePair lp = lane_pair(tr, hasA, hasB, ja, jb, q, db);
  GlobalByte* rpad = (GlobalByte*)(tr ? db.pad : q.pad);  // row and column letters
  GlobalByte* cpad = (GlobalByte*)(tr ? q.pad : db.pad);
  const int n = lp.n, mA = lp.mA, mB = lp.mB;
  const uint32_t qp = lp.rp, ap = lp.ap, bp = lp.bp;
  const int d1A = lp.d1A, d1B = lp.d1B, d0A = lp.d0A, d0B = lp.d0B;
  // Rows as in k_sw16 (all lanes of a group have the same range).
  int lo = 0, hi = -1;
  if (hasA) {
    lo = max(1, d0A + 1);
    hi = min(n, mA + d1A);
  }
  if (hasB) {
    lo = min(lo, max(1, d0B + 1));
    hi = max(hi, min(n, mB + d1B));
  }
  const int steps = __builtin_amdgcn_readfirstlane(wave_max(hi - lo + 1));

  // G/E of this lane's offsets; entry H is the E source past the last one
  // (the next lane's first offset, or the wall for the last lane).
  V G[H + 1], E[H + 1];
  uint32_t R[NR];
  // kWall: per offset, the addend that turns H into G (1 - o in the band,
  // kWallOff past the DP's own width), as two halves.
  V GO[kWall ? H : 1];
  const int liveA = kWall ? (d1A - d0A + 1) - t0 : 0, liveB = kWall ? (d1B - d0B + 1) - t0 : 0;
  const V LIVE2 = Tr::pair(liveA, liveB);
  // Without kWall the last lane's offsets from kDeadFrom on lie past the
  // band: their G takes the addend kWallOff instead of 1 - o, which keeps
  // them (and the E they feed) far below any score, as a per-DP wall does;
  // DADJ moves the F_in correction there by the same amount.
  const V GOD = last ? Tr::splat(kWallOff) : GOFF;
  const V DADJ = last ? Tr::splat(kWallOff - (1 - kOpen)) : ZERO2;
#pragma unroll
  for (int u = 0; u <= H; ++u) {
    if (kWall && u < H) {
      GO[u] = Tr::pair(u < liveA ? 1 - kOpen : kWallOff, u < liveB ? 1 - kOpen : kWallOff);
      G[u] = Tr::pair(u < liveA ? -kOpen : kNeg16, u < liveB ? -kOpen : kNeg16);
    } else {
      const bool dead = u == H || (last && u >= kDeadFrom);
      G[u] = dead ? NEG2 : Tr::splat(-kOpen);
    }
    E[u] = NEG2;
  }
#pragma unroll
  for (int k = 0; k < NR; ++k) {
    const int t = t0 + 2 * k;
    uint32_t lo16 = 2 * k < H ? pair_off4(pad_at(cpad, ap, mA, lo - d1A + t - 1),
                                          pad_at(cpad, bp, mB, lo - d1B + t - 1))
                              : 0u;
    uint32_t hi16 = 2 * k + 1 < H ? pair_off4(pad_at(cpad, ap, mA, lo - d1A + t),
                                              pad_at(cpad, bp, mB, lo - d1B + t))
                                  : 0u;
    R[k] = lo16 | (hi16 << 16);
  }
  // Letters of the next row, loaded one row ahead (clamped, unconditional
  // loads, settled at the end of the row): the query letter and the pair
  // entering the last lane at offset P*H-1.
  int qcur = (int)rpad[qp + (uint32_t)min(lo - 1, n)];
  uint32_t pnext = 0;
  V best = ZERO2;  // in the frame of the current row
  V Zr = ZERO2;    // frame offset of the current row
  // One row step. The window holds the entries of an even row (entry e at
  // R[e / 2], half e % 2) plus, at entry H, the one entering the next row;
  // the odd row (S = 1) reads entry u as entry u + 1 of that layout, so the
  // window moves by whole registers once every two rows.
  auto row = [&](int r, auto Sc) {
    constexpr int S = decltype(Sc)::value;
    const int i = lo + r;
    // Raw letters of row i + 1: subject positions i + 1 - d1 + P*H - 2.
    const int pos = i - 1 + P * H;
    const int rq = (int)rpad[qp + (uint32_t)min(i, n)];
    const int ra = pad_at(cpad, ap, mA, pos - d1A), rb = pad_at(cpad, bp, mB, pos - d1B);
    {
      // E source past the lane: the next lane's first offset, previous row.
      const V g0 = mv(G[0], NextC{}), e0 = mv(E[0], NextC{});
      G[H] = last ? NEG2 : g0;
      E[H] = last ? NEG2 : e0;
    }
    LdsChar* Tq = reinterpret_cast<LdsChar*>(T) + row_base(qcur);
    auto score = [&](int u) {
      const int e = u + S;
      const uint32_t off = (e & 1) ? (R[e >> 1] >> 16) : (R[e >> 1] & 0xFFFFu);
      return *reinterpret_cast<LdsInt*>(Tq + off);
    };
    constexpr int kAhead = H < 4 ? H : 4;
    int sq[kAhead];
#pragma unroll
    for (int u = 0; u < kAhead; ++u) sq[u] = score(u);
    V F = NEG2, Gl = NEG2, Hodd = ZERO2;
#pragma unroll
    for (int u = 0; u < H; ++u) {
      const V sc = Tr::bits(sq[u % kAhead]);
      if (u + kAhead < H) sq[u % kAhead] = score(u + kAhead);
      __builtin_amdgcn_sched_barrier(0);
      // D and E (with the zero) do not depend on F: only max(Gl, F), the
      // F decrement, the max3 and the G offset sit on the chain along the row.
      const V D = Tr::add(G[u], sc);
      const V Ev = Tr::max3(G[u + 1], E[u + 1], Zr);
      const V Fv = Tr::sub(Tr::max(Gl, F), ONE2);
      const V Hh = Tr::max3(D, Ev, Fv);
      // The best score takes cells two at a time as they come (one max3
      // per two cells, and no cell value stays live to the end of the row).
      if (u & 1) best = Tr::max3(best, Hodd, Hh);
      else if (u + 1 < H) Hodd = Hh;
      else best = Tr::max(best, Hh);
      Gl = Tr::add(Hh, kWall ? GO[kWall ? u : 0] : (u >= kDeadFrom ? GOD : GOFF));
      G[u] = Gl;
      E[u] = Ev;
      F = Fv;
    }
    // F entering each lane: exclusive prefix maximum of
    // F_out(q) - (P-1-q)*H over the group, plus (P-p)*H. (Every operand
    // stays at or below some F of the row, which the f16 bound needs.)
    const V fout = Tr::sub(Tr::max(Gl, F), ONE2);
    V y = Tr::sub(fout, Tr::splat((P - 1 - p) * H));
    if (P == 2) {  // one source lane: the exclusive prefix is its value
    } else if (P == 4) {  // quad_perm sources stay inside the quad
      y = Tr::max(y, mv(y, PrevC{}));
      y = Tr::max(y, mv(y, Back2C{}));
    } else if (P == 8) {  // row shifts reach into the other half-row: masked
      const V s1 = mv(y, PrevC{});
      y = p >= 1 ? Tr::max(y, s1) : y;
      const V s2 = mv(y, Shr2C{});
      y = p >= 2 ? Tr::max(y, s2) : y;
      const V s4 = mv(y, Shr4C{});
      y = p >= 4 ? Tr::max(y, s4) : y;
    } else {  // the group is a whole row: a lane without source keeps its own
      y = Tr::max(y, mvk(y, PrevC{}));
      y = Tr::max(y, mvk(y, Shr2C{}));
      y = Tr::max(y, mvk(y, Shr4C{}));
      y = Tr::max(y, mvk(y, Shr8C{}));
    }
    V fin = Tr::add(mv(y, PrevC{}), Tr::splat((P - p) * H));
    if (p == 0) fin = NEG2;
    best = Tr::max(best, fin);
    // The correction can only raise offsets u < F_in (F_in taken out of the
    // frame): from there on F_in - o - u is at most -o, the least G. It runs
    // in chunks of kChunk offsets while some lane of the wave still needs it
    // (with walls: inside its DPs' bands).
    {
      const V fu = Tr::sub(fin, Zr);
      using Sc16 = typename Tr::S;
      const Sc16 fl = Tr::hmax(kWall ? Tr::min(fu, LIVE2) : fu);
      V fc = kWall ? fin : Tr::add(fin, GOFF);
      constexpr int kChunk = 8;
#pragma unroll
      for (int c = 0; c < H; c += kChunk) {
        if (__ballot(fl > (Sc16)c) == 0) break;
#pragma unroll
        for (int u = c; u < (c + kChunk < H ? c + kChunk : H); ++u) {
          G[u] = Tr::max(G[u], kWall ? Tr::add(fc, GO[kWall ? u : 0])
                                     : (u >= kDeadFrom ? Tr::add(fc, DADJ) : fc));
          // Settled here: left to itself the compiler applies every chunk's
          // maxima after the last test, which keeps the whole ramp live
          // (about 30 registers more).
          asm volatile("" : "+v"(G[u]));
          fc = Tr::sub(fc, ONE2);
        }
      }
    }
    // Letters of row i + 1.
    qcur = settle(rq);
    pnext = pair_off4(settle(ra), settle(rb));
    // Frame of the next row; every kFrame rows the frame moves back (an
    // even row never ends a frame).
    best = Tr::add(best, ONE2);
    Zr = Tr::add(Zr, ONE2);
    if (S == 1 && (r & (kFrame - 1)) == kFrame - 1) {
#pragma unroll
      for (int u = 0; u < H; ++u) {
        G[u] = Tr::sub(G[u], FRAME2);
        E[u] = Tr::sub(E[u], FRAME2);
      }
      best = Tr::sub(best, FRAME2);
      Zr = ZERO2;
    }
  };
  // Entry e of the window gets the pair offset x.
  auto put = [&](int e, uint32_t x) {
    if (e & 1) R[e >> 1] = (R[e >> 1] & 0xFFFFu) | (x << 16);
    else R[e >> 1] = x;
  };
  static_assert(kFrame % 2 == 0, "frames end on odd rows");
  for (int r = 0; r < steps; r += 2) {
    // The next lane's entries 0 and 1 of row r: the entries entering at
    // H-1 in rows r+1 and r+2 (the last lane takes new pairs instead).
    const uint32_t nx = (uint32_t)dppi((int)R[0], NextC{});
    row(r, std::integral_constant<int, 0>{});
    if (r + 1 >= steps) break;
    put(H, last ? pnext : (nx & 0xFFFFu));
    row(r + 1, std::integral_constant<int, 1>{});
#pragma unroll
    for (int k = 0; k + 1 < NR; ++k) R[k] = R[k + 1];
    put(H - 1, last ? pnext : (nx >> 16));
  }
  best = Tr::sub(best, Zr);
  best = Tr::max(best, mv(best, SwapC{}));
  if (P >= 4) best = Tr::max(best, mv(best, Swap2C{}));
  if (P >= 8) best = Tr::max(best, mv(best, MirrorC{}));
  if (P == 16) best = Tr::max(best, mv(best, RowMirrorC{}));
  const int bA = Tr::lo(best), bB = Tr::hi(best);
  // Redo entries of bands wider than 256 carry bit 31: the int32 re-run
  // takes them with a wider wave kernel than the rest.
  constexpr uint32_t kWide = W > 257 ? 0x80000000u : 0u;
  if (p == 0) {
    if (hasA) {
      if (bA >= Tr::kRedo || (bA < 11 && lp.clampA)) {
        redo[atomicAdd(nredo, 1u)] = ln.x | kWide;
      } else {
        scores[ln.x] = bA;
      }
    }
    if (hasB) {
      if (bB >= Tr::kRedo || (bB < 11 && lp.clampB)) {
        redo[atomicAdd(nredo, 1u)] = ln.y | kWide;
      } else {
        scores[ln.y] = bB;
      }
    }
  }
}

template <int W, int P, int kThreads, class Tr, bool kWall>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(kMinWaves<kThreads>))) void k_sw16p(const uint2* __restrict__ lanes, uint32_t nlanes, uint32_t nrow,
                                                   const SwJob* __restrict__ jobs, SeqSet q, SeqSet db,
                                                   const int* __restrict__ pair_tab,
                                                   uint32_t* __restrict__ counter,
                                                   int32_t* __restrict__ scores,
                                                   uint32_t* __restrict__ redo,
                                                   uint32_t* __restrict__ nredo) {
  __shared__ int T[kPairTab];
  for (int k = threadIdx.x; k < kPairTab; k += blockDim.x) T[k] = pair_tab[k];
  __syncthreads();
  constexpr uint32_t kPer = 64 / P;
  const uint32_t ntasks = (nrow + kPer - 1) / kPer + (nlanes - nrow + kPer - 1) / kPer;
  const int lane = threadIdx.x & 63;
  for (;;) {
    const uint32_t task =
        (uint32_t)__builtin_amdgcn_readfirstlane((int)atomicAdd(counter, lane == 0 ? 1u : 0u));
    if (task >= ntasks) break;
    sw16p_task<W, P, Tr, kWall>(task, lanes, nlanes, nrow, jobs, q, db, (LdsInt*)T, scores, redo, nredo);
  }
}

// A task of the fused kernel as a function of its own: each class then gets
// its own register allocation (inlined into one loop, the eight bodies
// spill).
template <int W, int P, class Tr, bool kWall>
__device__ __attribute__((noinline)) void sw16p_task_call(uint32_t task, const uint2* __restrict__ lanes,
                                                          uint32_t nlanes, uint32_t nrow, const SwJob* __restrict__ jobs,
                                                          SeqSet q, SeqSet db, LdsInt* T, int32_t* __restrict__ scores,
                                                          uint32_t* __restrict__ redo, uint32_t* __restrict__ nredo) {
  sw16p_task<W, P, Tr, kWall>(task, lanes, nlanes, nrow, jobs, q, db, T, scores, redo, nredo);
}

// Classes of the fused launch, as X(kind, class, width, lanes per DP pair,
// wall), in task order: the merged bands (long DPs) first.
#define DMR_SW16_FUSED(X) \
  X(0, 12, 513, 16, true) \
  X(1, 11, 385, 16, true) \
  X(2, 10, 257, 8, true)  \
  X(3, 9, 225, 8, true)   \
  X(4, 8, 193, 8, true)   \
  X(5, 7, 161, 8, true)   \
  X(6, 6, 129, 8, true)   \
  X(7, 5, 97, 8, true)    \
  X(8, 3, 81, 2, false)   \
  X(9, 2, 61, 2, false)
// Fused launch of several packed classes (launch_sw16_fused): one persistent
// grid whose waves take tasks of any of the classes, the merged bands first.
// The classes share one block size and two waves per SIMD (at most 256
// registers), so a class with few long tasks no longer holds whole CUs while
// the others wait for them: its tasks and the others' interleave on the same
// SIMDs, and the others' short tasks fill its tail.
template <class Tr>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(2))) void k_sw16p_fused(
    FusedSections a, const SwJob* __restrict__ jobs, SeqSet q, SeqSet db, const int* __restrict__ pair_tab,
    uint32_t* __restrict__ counter, int32_t* __restrict__ scores, uint32_t* __restrict__ redo,
    uint32_t* __restrict__ nredo) {
  __shared__ int T[kPairTab];
  for (int k = threadIdx.x; k < kPairTab; k += blockDim.x) T[k] = pair_tab[k];
  __syncthreads();
  const int lane = threadIdx.x & 63;
  for (;;) {
    const uint32_t task =
        (uint32_t)__builtin_amdgcn_readfirstlane((int)atomicAdd(counter, lane == 0 ? 1u : 0u));
    if (task >= a.ntasks) break;
    int k = 0;
    while (k + 1 < kFusedKinds && task >= a.task0[k + 1]) ++k;
    const uint32_t t = task - a.task0[k];
    const uint2* ln = a.lanes[k];
    const uint32_t nl = a.nlanes[k], nr = a.nrow[k];
    switch (k) {
#define DMR_SW16_FUSED_CASE(i, c, w, p, wall) \
      case i:                                  \
        sw16p_task_call<w, p, Tr, wall>(t, ln, nl, nr, jobs, q, db, (LdsInt*)T, scores, redo, nredo); \
        break;
      DMR_SW16_FUSED(DMR_SW16_FUSED_CASE)
#undef DMR_SW16_FUSED_CASE
      default:
        break;
    }
  }
}

// Persistent launch of a split kernel: 64 / P lanes per wave task, as many
// resident blocks as the work needs.
template <int W, int P, int T, class Tr, bool kWall>
void launch_split(const uint2* lanes, uint32_t nlanes, uint32_t nrow, const SwJob* jobs, SeqSet q, SeqSet db,
                  const int* pair_tab, uint32_t* counter, int32_t* scores, uint32_t* redo, uint32_t* nredo,
                  int cu_count, hipStream_t st) {
  constexpr unsigned kPer = 64 / P;
  const unsigned tasks = (nrow + kPer - 1) / kPer + (nlanes - nrow + kPer - 1) / kPer;
  static const unsigned resident = resident_blocks(k_sw16p<W, P, T, Tr, kWall>, T, cu_count);
  const unsigned blocks = std::min<unsigned>(resident, (tasks + T / 64 - 1) / (T / 64));
  hipLaunchKernelGGL((k_sw16p<W, P, T, Tr, kWall>), dim3(blocks), dim3(T), 0, st, lanes, nlanes, nrow, jobs, q, db,
                     pair_tab, counter, scores, redo, nredo);
}

}  // namespace

// Compiled split kernels, as X(class, width, lanes per DP pair, threads).
// Band_for widths (no walls; k_sw16.hip): the last lane of a group holds at
// least one band offset, so W - (P - 1) * ceil((W + 1) / P) >= 1.
#define DMR_SW16_EXACT_SPLIT(X) \
  X(0, 25, 4, 512)              \
  X(1, 33, 4, 512)              \
  X(2, 61, 2, 512)              \
  X(2, 61, 4, 512)              \
  X(3, 81, 2, 512)              \
  X(3, 81, 4, 512)              \
  X(4, 129, 4, 768)             \
  X(4, 129, 4, 512)             \
  X(4, 129, 8, 512)
// Merged bands (walls at each DP's own width; k_sw16m.hip).
#define DMR_SW16_MERGED(X) \
  X(5, 97, 4, 512)         \
  X(5, 97, 8, 512)         \
  X(6, 129, 4, 512)        \
  X(6, 129, 8, 512)        \
  X(7, 161, 8, 512)        \
  X(7, 161, 16, 512)       \
  X(8, 193, 8, 512)        \
  X(8, 193, 16, 512)       \
  X(9, 225, 8, 512)        \
  X(9, 225, 16, 512)       \
  X(10, 257, 8, 512)       \
  X(10, 257, 16, 512)      \
  X(11, 385, 16, 512)      \
  X(12, 513, 16, 512)

}  // namespace dmr
