// Banded affine Smith-Waterman: exact int32 path (any band width, optional
// traceback).
//
// Restates gapped::extend_span (src/gapped.rs:67-280) one DP per lane: rows
// i = 1..n of the query window, columns j in [j0(i), j1(i)] with
// j0 = clamp(i-d1, 0, m), j1 = max(clamp(i-d0, 0, m), j0), every row holding
// at least one cell. H takes the diagonal (H outside the previous row's range
// counts as 0), E (vertical) and F (horizontal) with the strict D > E > F
// priority; the best cell is the first strict maximum in row-major order.
// Trace codes are choice | e_ext << 2 | f_ext << 3, one byte per cell, and the
// traceback follows the reference's state machine, emitting run-length ops
// (op | count << 2, op 0 = M, 1 = I, 2 = D) and the alignment statistics.
#include "dmr_kernels.hpp"

namespace dmr {

namespace {

constexpr int32_t kNeg = INT32_MIN / 4;

struct RowRange {
  int64_t d0, d1, m;
  __device__ __forceinline__ int64_t lo(int64_t i) const {
    int64_t v = i - d1;
    return v < 0 ? 0 : (v > m ? m : v);
  }
  __device__ __forceinline__ int64_t hi(int64_t i) const {
    int64_t v = i - d0;
    v = v < 0 ? 0 : (v > m ? m : v);
    int64_t l = lo(i);
    return v > l ? v : l;
  }
};

__global__ void k_job_geometry(const SwJob* __restrict__ jobs, const uint32_t* __restrict__ job_idx,
                               uint64_t njobs, SeqSet db, uint64_t* __restrict__ cells,
                               uint64_t* __restrict__ width0, uint32_t* __restrict__ stride) {
  uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= njobs) return;
  const SwJob J = jobs[job_idx ? job_idx[t] : t];
  int64_t n = (int64_t)J.q1 - J.q0;
  int64_t m = db.offs[J.subject + 1] - db.offs[J.subject];
  RowRange R{J.d0, J.d1, m};
  uint64_t c = 0;
  int64_t w = 1;
  for (int64_t i = 1; i <= n; ++i) {
    int64_t x = R.hi(i) - R.lo(i) + 1;
    c += (uint64_t)x;
    if (x > w) w = x;
  }
  int64_t x0 = R.hi(0) - R.lo(0) + 1;
  if (x0 > w) w = x0;
  if (cells) cells[t] = c;
  if (width0) width0[t] = (uint64_t)x0;
  if (stride) stride[t] = (uint32_t)w;
}

template <bool kTrace>
__global__ void k_sw_generic(const SwJob* __restrict__ jobs, const uint32_t* __restrict__ job_idx,
                             uint64_t njobs, SeqSet q, SeqSet db, const int8_t* __restrict__ m32,
                             int gap_open, int gap_extend, int32_t* __restrict__ scratch,
                             const uint64_t* __restrict__ scratch_off, uint8_t* __restrict__ trace,
                             const uint64_t* __restrict__ trace_off, uint32_t* __restrict__ ops,
                             const uint64_t* __restrict__ ops_off, SwOut* __restrict__ out,
                             int scatter, int32_t* __restrict__ scores) {
  __shared__ int8_t M[kScoreStride * kScoreStride];
  for (int i = threadIdx.x; i < kScoreStride * kScoreStride; i += blockDim.x) M[i] = m32[i];
  __syncthreads();
  uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= njobs) return;
  const SwJob J = jobs[job_idx ? job_idx[t] : t];
  const uint8_t* qs = q.res + q.offs[J.target] + J.q0;
  const int64_t n = (int64_t)J.q1 - J.q0;
  const uint8_t* ss = db.res + db.offs[J.subject];
  const int64_t m = db.offs[J.subject + 1] - db.offs[J.subject];
  SwOut o = {};
  // Score mode with scores: only the score, at the job's index.
  SwOut* dst = scores ? nullptr : out + (scatter ? job_idx[t] : t);
  int32_t* sdst = scores ? scores + (job_idx ? job_idx[t] : t) : nullptr;
  if (n <= 0 || m <= 0) {
    if (sdst) *sdst = 0;
    else *dst = o;
    return;
  }
  const RowRange R{J.d0, J.d1, m};
  const int32_t oc = gap_open + gap_extend, ge = gap_extend;
  const uint64_t base = scratch_off[t];
  const int64_t W = (int64_t)(scratch_off[t + 1] - base) / 4;
  int32_t* ph = scratch + base;
  int32_t* pe = ph + W;
  int32_t* ch = pe + W;
  int32_t* ce = ch + W;
  for (int64_t k = 0; k < W; ++k) {
    ph[k] = 0;
    pe[k] = kNeg;
  }
  int32_t best = 0;
  int64_t bi = 0, bj = 0;
  uint64_t row_off = 0, best_row_off = 0;
  int64_t pa = R.lo(0), pb = R.hi(0);
  uint8_t* tr = kTrace ? trace + trace_off[t] : nullptr;
  row_off = (uint64_t)(pb - pa + 1);  // row 0 occupies the first width(0) trace bytes
  for (int64_t i = 1; i <= n; ++i) {
    const int64_t a = R.lo(i), b = R.hi(i);
    const int qrow = qs[i - 1] * kScoreStride;
    int32_t f_left = kNeg;
    for (int64_t j = a; j <= b; ++j) {
      int32_t d = kNeg;
      if (j >= 1) {
        int32_t hd = (j > pa && j - 1 <= pb) ? ph[j - 1 - pa] : 0;
        d = hd + M[qrow + ss[j - 1]];
      }
      int32_t e = kNeg;
      uint8_t e_ext = 0;
      if (j >= pa && j <= pb) {
        int32_t hp = ph[j - pa];
        if (hp > kNeg / 2 && hp - oc > e) e = hp - oc;
        int32_t pv = pe[j - pa];
        if (pv > kNeg / 2 && pv - ge > e) {
          e = pv - ge;
          e_ext = 1;
        }
      }
      int32_t f = kNeg;
      uint8_t f_ext = 0;
      if (j > a) {
        int32_t hl = ch[j - 1 - a];
        if (hl > kNeg / 2 && hl - oc > f) f = hl - oc;
        if (f_left > kNeg / 2 && f_left - ge > f) {
          f = f_left - ge;
          f_ext = 1;
        }
      }
      int32_t hv = 0;
      uint8_t choice = 0;
      if (d > hv) {
        hv = d;
        choice = 1;
      }
      if (e > hv) {
        hv = e;
        choice = 2;
      }
      if (f > hv) {
        hv = f;
        choice = 3;
      }
      ch[j - a] = hv;
      ce[j - a] = e;
      if (kTrace) tr[row_off + (uint64_t)(j - a)] = (uint8_t)(choice | (e_ext << 2) | (f_ext << 3));
      f_left = f;
      if (hv > best) {
        best = hv;
        bi = i;
        bj = j;
        best_row_off = row_off;
      }
    }
    int32_t* tmp = ph;
    ph = ch;
    ch = tmp;
    tmp = pe;
    pe = ce;
    ce = tmp;
    pa = a;
    pb = b;
    row_off += (uint64_t)(b - a + 1);
  }
  o.score = best;
  o.best_i = (uint32_t)bi;
  o.best_j = (uint32_t)bj;
  if (kTrace && best > 0) {
    // Traceback from the best cell; runs are collected in reverse.
    uint32_t* op = ops + ops_off[t];
    uint32_t nruns = 0;
    int state = 0;
    int64_t i = bi, j = bj;
    uint64_t roff = best_row_off;
    bool in_gap = false;
    uint32_t cur_op = 3, cur_n = 0;
    auto emit = [&](uint32_t code) {
      if (code == cur_op) {
        ++cur_n;
      } else {
        if (cur_n) op[nruns++] = cur_op | (cur_n << 2);
        cur_op = code;
        cur_n = 1;
      }
    };
    while (i > 0 && j > 0) {
      int64_t a = R.lo(i), b = R.hi(i);
      if (j < a || j > b) break;
      uint8_t tv = tr[roff + (uint64_t)(j - a)];
      if (state == 0) {
        uint8_t c = tv & 3;
        if (c == 0) break;
        if (c == 1) {
          uint8_t x = qs[i - 1], y = ss[j - 1];
          o.length++;
          if (x == y) o.identities++;
          else o.mismatches++;
          if (M[x * kScoreStride + y] > 0) o.positives++;
          emit(0);
          in_gap = false;
          --i;
          --j;
          roff -= (uint64_t)(R.hi(i) - R.lo(i) + 1);
        } else if (c == 2) {
          state = 1;
        } else {
          state = 2;
        }
      } else if (state == 1) {
        o.length++;
        o.gaps++;
        if (!in_gap) {
          o.gap_openings++;
          in_gap = true;
        }
        emit(1);
        if (((tv >> 2) & 1) == 0) state = 0;
        --i;
        roff -= (uint64_t)(R.hi(i) - R.lo(i) + 1);
      } else {
        o.length++;
        o.gaps++;
        if (!in_gap) {
          o.gap_openings++;
          in_gap = true;
        }
        emit(2);
        if (((tv >> 3) & 1) == 0) state = 0;
        --j;
      }
    }
    if (cur_n) op[nruns++] = cur_op | (cur_n << 2);
    // Runs were produced back to front.
    for (uint32_t a = 0, b = nruns ? nruns - 1 : 0; a < b; ++a, --b) {
      uint32_t x = op[a];
      op[a] = op[b];
      op[b] = x;
    }
    o.n_ops = nruns;
    o.q_begin = (uint32_t)i;
    o.s_begin = (uint32_t)j;
    if (nruns == 0) o.score = 0;
  }
  if (sdst) *sdst = o.score;
  else *dst = o;
}

inline unsigned nblk(uint64_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

}  // namespace

void launch_job_geometry(const SwJob* jobs, const uint32_t* job_idx, uint64_t njobs, SeqSet db,
                         uint64_t* cells, uint64_t* width0, uint32_t* stride, hipStream_t st) {
  if (!njobs) return;
  hipLaunchKernelGGL(k_job_geometry, dim3(nblk(njobs, 256)), dim3(256), 0, st, jobs, job_idx, njobs,
                     db, cells, width0, stride);
}

void launch_sw_generic(const SwJob* jobs, const uint32_t* job_idx, uint64_t njobs, SeqSet q,
                       SeqSet db, const int8_t* matrix32, int gap_open, int gap_extend,
                       int32_t* scratch, const uint64_t* scratch_off, uint8_t* trace,
                       const uint64_t* trace_off, uint32_t* ops, const uint64_t* ops_off,
                       SwOut* out, bool traceback, hipStream_t st, bool scatter, int32_t* scores) {
  if (!njobs) return;
  const int sc = (scatter && job_idx) ? 1 : 0;
  if (traceback)
    hipLaunchKernelGGL(k_sw_generic<true>, dim3(nblk(njobs, 64)), dim3(64), 0, st, jobs, job_idx,
                       njobs, q, db, matrix32, gap_open, gap_extend, scratch, scratch_off, trace,
                       trace_off, ops, ops_off, out, sc, traceback ? nullptr : scores);
  else
    hipLaunchKernelGGL(k_sw_generic<false>, dim3(nblk(njobs, 64)), dim3(64), 0, st, jobs, job_idx,
                       njobs, q, db, matrix32, gap_open, gap_extend, scratch, scratch_off, trace,
                       trace_off, ops, ops_off, out, sc, traceback ? nullptr : scores);
}

}  // namespace dmr
