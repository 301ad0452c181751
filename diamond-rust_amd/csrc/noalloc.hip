// One banded DP with traceback that allocates nothing: the device memory
// the row-exact kernel needs (its job record, the two letter sequences, four
// score rows, the trace bytes and the ops) is carved from `scratch`, a
// static device array every GPU holds for the life of the process.
//
// The array's name in host code is a stand-in, not a device address: the
// address is taken once per context (its device current) with
// hipGetSymbolAddress and kept in the engine; only that pointer reaches
// kernels and copies. Every call on a device shares the device's array and
// leaves the last call's contents in it; the kernel writes every row, trace
// byte and op before it reads it, so nothing needs clearing between calls.
// Calls that overlap on one device (two contexts of one GPU used from two
// threads at once) race on the array.
#include "engine.hpp"

namespace dmr {

static __device__ __align__(256) uint8_t scratch[Engine::kNoAllocScratchBytes];

namespace {

// Row range of row i of the band (src/gapped.rs:87-93), as in k_sw.hip.
struct HostRows {
  int64_t d0, d1, m;
  int64_t lo(int64_t i) const { return std::min(std::max<int64_t>(i - d1, 0), m); }
  int64_t hi(int64_t i) const { return std::max(std::min(std::max<int64_t>(i - d0, 0), m), lo(i)); }
};

}  // namespace

void* Engine::noalloc_scratch_address() {
  void* p = nullptr;
  DMR_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) throw Error(-2, "hipGetSymbolAddress returned no address for the no-allocation scratch");
  return p;
}

void Engine::extend_span_noalloc(const uint8_t* q, uint64_t qlen, const uint8_t* s, uint64_t slen,
                                 int64_t d0, int64_t d1, int gap_open, int gap_extend, SwOut* out,
                                 uint32_t* ops, uint32_t ops_cap, bool* ops_fit) {
  DMR_HIP(hipSetDevice(device_));
  if (!noalloc_) throw Error(-2, "no-allocation scratch not set up");
  if (qlen >= (1ull << 31) || slen >= (1ull << 31)) throw Error(-5, "sequence longer than 2^31 letters");
  const int64_t lo = std::min(d0, d1), hi = std::max(d0, d1);
  if (lo < INT32_MIN / 2 || hi > INT32_MAX / 2) throw Error(-1, "diagonal out of range");
  // Geometry of the DP (k_job_geometry on the host): cells of rows 1..n, the
  // width of row 0 (the head of the trace) and the widest row.
  const HostRows R{lo, hi, (int64_t)slen};
  const int64_t n = (int64_t)qlen;
  uint64_t cells = 0;
  int64_t w = R.hi(0) - R.lo(0) + 1;
  const uint64_t width0 = (uint64_t)w;
  for (int64_t i = 1; i <= n; ++i) {
    const int64_t x = R.hi(i) - R.lo(i) + 1;
    cells += (uint64_t)x;
    w = std::max(w, x);
  }
  // Regions of the array, each 256-byte aligned.
  uint64_t at = 0;
  auto take = [&](uint64_t bytes) {
    const uint64_t o = at;
    at = (at + bytes + 255) & ~255ull;
    return o;
  };
  const uint64_t o_job = take(sizeof(SwJob)), o_out = take(sizeof(SwOut));
  const uint64_t o_offs = take(4 * sizeof(uint32_t)), o_tabs = take(6 * sizeof(uint64_t));
  const uint64_t o_q = take(qlen + 1), o_s = take(slen + 1);
  const uint64_t o_rows = take(16ull * (uint64_t)w), o_trace = take(width0 + cells);
  const uint64_t n_ops_cap = qlen + slen + 1, o_ops = take(4 * n_ops_cap);
  if (at > kNoAllocScratchBytes)
    throw Error(-5, "this DP needs " + std::to_string(at) + " bytes of device scratch; the no-allocation form holds " +
                        std::to_string(kNoAllocScratchBytes) + " (use dmr_extend_span)");
  uint8_t* base = static_cast<uint8_t*>(noalloc_);
  const SwJob job{0, 0, 0, (uint32_t)qlen, (int32_t)lo, (int32_t)hi, 0, 0};
  const uint32_t offs[4] = {0, (uint32_t)qlen, 0, (uint32_t)slen};
  // scratch_off counts int32 words (four rows of w), trace_off and ops_off
  // bytes and words.
  const uint64_t tabs[6] = {0, 4ull * (uint64_t)w, 0, width0 + cells, 0, n_ops_cap};
  DMR_HIP(hipMemcpyAsync(base + o_job, &job, sizeof job, hipMemcpyHostToDevice, st_));
  DMR_HIP(hipMemcpyAsync(base + o_offs, offs, sizeof offs, hipMemcpyHostToDevice, st_));
  DMR_HIP(hipMemcpyAsync(base + o_tabs, tabs, sizeof tabs, hipMemcpyHostToDevice, st_));
  if (qlen) DMR_HIP(hipMemcpyAsync(base + o_q, q, qlen, hipMemcpyHostToDevice, st_));
  if (slen) DMR_HIP(hipMemcpyAsync(base + o_s, s, slen, hipMemcpyHostToDevice, st_));
  const uint32_t* d_offs = reinterpret_cast<const uint32_t*>(base + o_offs);
  const uint64_t* d_tabs = reinterpret_cast<const uint64_t*>(base + o_tabs);
  const SeqSet qs{base + o_q, d_offs}, ss{base + o_s, d_offs + 2};
  launch_sw_generic(reinterpret_cast<const SwJob*>(base + o_job), nullptr, 1, qs, ss, matrix32_.as<int8_t>(),
                    gap_open, gap_extend, reinterpret_cast<int32_t*>(base + o_rows), d_tabs, base + o_trace,
                    d_tabs + 2, reinterpret_cast<uint32_t*>(base + o_ops), d_tabs + 4,
                    reinterpret_cast<SwOut*>(base + o_out), true, st_);
  DMR_HIP(hipGetLastError());
  DMR_HIP(hipMemcpyAsync(out, base + o_out, sizeof(SwOut), hipMemcpyDeviceToHost, st_));
  DMR_HIP(hipStreamSynchronize(st_));
  *ops_fit = out->n_ops <= ops_cap;
  if (out->n_ops && *ops_fit) {
    DMR_HIP(hipMemcpyAsync(ops, base + o_ops, 4ull * out->n_ops, hipMemcpyDeviceToHost, st_));
    DMR_HIP(hipStreamSynchronize(st_));
  }
}

}  // namespace dmr
