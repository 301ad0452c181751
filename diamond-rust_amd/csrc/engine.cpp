// Host orchestration of the search core (see engine.hpp).
#include "engine.hpp"

#include "host_pool.hpp"

#include <dlfcn.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <numeric>

namespace dmr {

PinPool::~PinPool() {
  for (auto& b : free_) (void)hipHostFree(b.first);
}

void* PinPool::take(size_t bytes, size_t* cap) {
  {
    std::lock_guard<std::mutex> g(m_);
    size_t best = free_.size();
    for (size_t k = 0; k < free_.size(); ++k)
      if (free_[k].second >= bytes && (best == free_.size() || free_[k].second < free_[best].second)) best = k;
    if (best < free_.size()) {
      auto b = free_[best];
      free_.erase(free_.begin() + (long)best);
      *cap = b.second;
      return b.first;
    }
  }
  const size_t c = bytes + bytes / 4 + 4096;
  void* p = nullptr;
  DMR_HIP(hipHostMalloc(&p, c, hipHostMallocDefault));
  *cap = c;
  return p;
}

void PinPool::give(void* p, size_t cap) {
  std::lock_guard<std::mutex> g(m_);
  free_.push_back({p, cap});
}

OpsStore::OpsStore(std::shared_ptr<PinPool> pool, size_t words) : pool_(std::move(pool)) {
  blk_ = pool_->take(words * sizeof(uint32_t), &cap_);
}

OpsStore::~OpsStore() {
  if (blk_) pool_->give(blk_, cap_);
}

void launch_fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t st);
void launch_gather(const SwJob* jobs, const int32_t* scores, const uint32_t* idx, uint64_t n,
                   SwJob* jout, int32_t* sout, hipStream_t st);
void launch_xdrop_pairs(const uint8_t* q, uint64_t qlen, const uint8_t* s, uint64_t slen,
                        const uint32_t* qa, const uint32_t* sa, uint64_t n, const int8_t* m32,
                        int xdrop, SegRec* out, hipStream_t st);
void launch_join_emit_entries(const uint32_t* qs, const uint32_t* qp, const uint32_t* run_start,
                              uint64_t nruns, uint64_t nq_total, const uint32_t* rs,
                              const uint32_t* rp, const uint32_t* counts, const uint64_t* rlo,
                              const uint64_t* hit_off, HitGeom g, uint64_t* hits, hipStream_t st);

namespace {

int bits_for(uint64_t x) { return x == 0 ? 1 : 64 - __builtin_clzll(x); }

double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

// src/search.rs:273-282 — score desc, E-value asc, subject asc.
bool hsp_less(const HspRec& a, const HspRec& b) {
  if (a.score != b.score) return a.score > b.score;
  if (a.evalue < b.evalue) return true;
  if (b.evalue < a.evalue) return false;
  return a.subject < b.subject;
}

}  // namespace

Engine::Engine(int device) : device_(device) {
  int n = 0;
  DMR_HIP(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) throw Error(-1, "no such HIP device");
  DMR_HIP(hipSetDevice(device));
  DMR_HIP(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
  for (auto& e : ev_) DMR_HIP(hipEventCreate(&e));
  for (auto& s : aux_) DMR_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  DMR_HIP(hipEventCreateWithFlags(&fork_ev_, hipEventDisableTiming));
  for (auto& e : join_ev_) DMR_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  // 32x32 score table; letters >= 25 (padding) score -64 against everything.
  int8_t m[kScoreStride * kScoreStride];
  for (int a = 0; a < kScoreStride; ++a)
    for (int b = 0; b < kScoreStride; ++b)
      m[a * kScoreStride + b] = (a < 25 && b < 25) ? blosum62()[a][b] : (int8_t)-64;
  DMR_HIP(hipMemcpy(matrix32_.ensure<int8_t>(sizeof m), m, sizeof m, hipMemcpyHostToDevice));
  std::vector<int> pt;
  fill_pair_table(blosum62(), false, &pt);
  DMR_HIP(hipMemcpy(pair_tab_.ensure<int>(pt.size()), pt.data(), pt.size() * 4, hipMemcpyHostToDevice));
  fill_pair_table(blosum62(), true, &pt);
  DMR_HIP(hipMemcpy(pair_tab16_.ensure<int>(pt.size()), pt.data(), pt.size() * 4, hipMemcpyHostToDevice));
  hipDeviceProp_t prop;
  DMR_HIP(hipGetDeviceProperties(&prop, device));
  cu_count_ = prop.multiProcessorCount;
  // The no-allocation entry point's array: its address on this device.
  noalloc_ = noalloc_scratch_address();
  const char* fg = std::getenv("DMR_FORCE_GENERIC");
  force_generic_ = fg && fg[0] == '1';
  const char* lr = std::getenv("DMR_LANE_REPORT");
  lane_report_ = lr && lr[0] == '1';
  // DMR_SCORE_SERIAL=1 runs the score classes one after another on the engine
  // stream (per-class kernel times without sharing the chip).
  const char* ss = std::getenv("DMR_SCORE_SERIAL");
  score_serial_ = ss && ss[0] == '1';
  // Score-pass kernel settings: the tuning file DMR_TUNING names (empty:
  // none), else tuning_<arch>.json beside this library when it exists.
  const char* tf = std::getenv("DMR_TUNING");
  if (tf) {
    load_tuning(tf);
  } else {
    Dl_info info{};
    std::string arch = prop.gcnArchName;
    arch = arch.substr(0, arch.find(':'));
    if (dladdr(reinterpret_cast<void*>(&default_layout), &info) && info.dli_fname) {
      std::string lib = info.dli_fname;
      const std::string path = lib.substr(0, lib.rfind('/') + 1) + "tuning_" + arch + ".json";
      if (FILE* f = std::fopen(path.c_str(), "rb")) {
        std::fclose(f);
        load_tuning(path);
      }
    }
  }
  // Hits per slice of targets (device memory of the per-hit and per-DP
  // buffers grows with it: ~52 bytes per hit and ~150 per DP job).
  const char* sh = std::getenv("DMR_SLICE_HITS");
  slice_hits_ = sh ? std::strtoull(sh, nullptr, 0) : (300ull << 20);
}

void Engine::load_tuning(const std::string& path) {
  std::map<std::string, Settings> t;
  if (!path.empty()) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) throw Error(-1, "cannot open tuning file " + path);
    std::string text;
    char buf[4096];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) text.append(buf, n);
    std::fclose(f);
    try {
      t = parse_tuning(text);
    } catch (const Error& e) {
      throw Error(-1, path + ": " + e.what());
    }
  }
  std::swap(tuned_, t);
  try {
    layout_ = resolve_layout(layout_preset_);
  } catch (...) {
    std::swap(tuned_, t);
    throw;
  }
  tuning_file_ = path;
}

ScoreLayout Engine::resolve_layout(const std::string& preset, const Settings* pinned) const {
  ScoreLayout L = default_layout();
  auto any = tuned_.find("*");
  if (any != tuned_.end()) apply_settings(&L, any->second);
  auto own = tuned_.find(preset);
  if (own != tuned_.end() && preset != "*") apply_settings(&L, own->second);
  apply_settings(&L, pinned ? *pinned : pinned_);
  return L;
}

void Engine::set_layout_key(const std::string& key, double value) {
  if (!is_layout_key(key)) throw Error(-1, "unknown tuning key: " + key);
  Settings p = pinned_;
  p[key] = value;
  // A lane count sets its class's block size too: a block size pinned
  // before it goes.
  if (key.compare(0, 6, "lanes_") == 0) p.erase("threads_" + key.substr(6));
  layout_ = resolve_layout(layout_preset_, &p);
  pinned_ = std::move(p);
}

void Engine::copy_tuning(const Engine& o) {
  tuned_ = o.tuned_;
  pinned_ = o.pinned_;
  tuning_file_ = o.tuning_file_;
  layout_ = o.layout_;
  layout_preset_ = o.layout_preset_;
}

Engine::~Engine() {
  (void)hipSetDevice(device_);
  if (st_) (void)hipStreamSynchronize(st_);
  for (auto& s : aux_)
    if (s) (void)hipStreamSynchronize(s);
  for (auto& e : ev_)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : join_ev_)
    if (e) (void)hipEventDestroy(e);
  if (fork_ev_) (void)hipEventDestroy(fork_ev_);
  for (auto& s : aux_)
    if (s) (void)hipStreamDestroy(s);
  if (gexec_) (void)hipGraphExecDestroy(gexec_);
  if (st_) (void)hipStreamDestroy(st_);
}

void Engine::fork() {
  DMR_HIP(hipEventRecord(fork_ev_, st_));
  for (auto& s : aux_) DMR_HIP(hipStreamWaitEvent(s, fork_ev_, 0));
}

void Engine::join() {
  for (int k = 0; k < kAux; ++k) {
    DMR_HIP(hipEventRecord(join_ev_[k], aux_[k]));
    DMR_HIP(hipStreamWaitEvent(st_, join_ev_[k], 0));
  }
}

template <class T>
T Engine::read1(const T* dptr) {
  T v;
  DMR_HIP(hipMemcpyAsync(&v, dptr, sizeof(T), hipMemcpyDeviceToHost, st_));
  DMR_HIP(hipStreamSynchronize(st_));
  return v;
}

void Engine::stage_event(int k) {
  if (capturing_) return;  // recorded after the graph's launch instead
  DMR_HIP(hipEventRecord(ev_[k], st_));
}

thread_local bool g_capture_active = false;

void Engine::end_capture_and_launch() {
  g_capture_active = false;
  capturing_ = false;
  hipGraph_t gr = nullptr;
  DMR_HIP(hipStreamEndCapture(st_, &gr));
  bool updated = false;
  if (gexec_) {
    hipGraphNode_t err = nullptr;
    hipGraphExecUpdateResult res = hipGraphExecUpdateError;
    updated = hipGraphExecUpdate(gexec_, gr, &err, &res) == hipSuccess && res == hipGraphExecUpdateSuccess;
    if (!updated) {
      (void)hipGetLastError();
      (void)hipGraphExecDestroy(gexec_);
      gexec_ = nullptr;
    }
  }
  hipError_t e = hipSuccess;
  if (!gexec_) {
    e = hipGraphInstantiate(&gexec_, gr, nullptr, nullptr, 0);
    ++graph_builds_;
  }
  (void)hipGraphDestroy(gr);
  DMR_HIP(e);
  DMR_HIP(hipGraphLaunch(gexec_, st_));
  ++graph_launches_;
  // The chain's stage events: one span for the whole graph.
  for (int k : {4, 5, 6}) DMR_HIP(hipEventRecord(ev_[k], st_));
}
float Engine::stage_ms(int a, int b) {
  float ms = 0;
  DMR_HIP(hipEventElapsedTime(&ms, ev_[a], ev_[b]));
  return ms;
}

AlpParams Engine::alp(int go, int ge) const {
  GapRow r;
  if (!blosum62_gap_row(go, ge, &r)) throw Error(-1, "unsupported gap costs for BLOSUM62");
  return alp_make(r.lambda, r.k, r.alpha, r.alpha_v, r.sigma, go, ge, kB62UngappedAlpha,
                  kB62UngappedAlphaV);
}

void Engine::share_db(const Engine& other) {
  if (other.device_ != device_) throw Error(-1, "engines of different devices cannot share a database");
  db_ = other.db_;
  idx_.reset();
}

void Engine::load_db(const uint8_t* letters, const uint64_t* offsets, uint32_t n,
                     const uint64_t* stat_letters) {
  DMR_HIP(hipSetDevice(device_));
  uint64_t total = offsets[n];
  if (total + n + 2 >= (1ull << 32)) throw Error(-5, "database larger than 4G letters");
  // A new resident database: engines sharing the previous one keep it.
  auto d = std::make_shared<DbState>();
  d->h_offs.resize(n + 1);
  uint32_t maxl = 0;
  std::vector<uint32_t> seq_of(total);
  for (uint32_t i = 0; i < n; ++i) {
    d->h_offs[i] = (uint32_t)offsets[i];
    uint32_t l = (uint32_t)(offsets[i + 1] - offsets[i]);
    maxl = std::max(maxl, l);
    std::fill(seq_of.begin() + (long)offsets[i], seq_of.begin() + (long)offsets[i + 1], i);
  }
  d->h_offs[n] = (uint32_t)total;
  d->h_offs_sp = std::make_shared<const std::vector<uint32_t>>(d->h_offs);
  d->h_res_sp = std::make_shared<const std::vector<uint8_t>>(letters, letters + total);
  DMR_HIP(hipMemcpy(d->res.ensure<uint8_t>(total + 64), letters, total, hipMemcpyHostToDevice));
  DMR_HIP(hipMemcpy(d->offs.ensure<uint32_t>(n + 1), d->h_offs.data(), (n + 1) * 4, hipMemcpyHostToDevice));
  DMR_HIP(hipMemcpy(d->seq_of.ensure<uint32_t>(total + 1), seq_of.data(), total * 4, hipMemcpyHostToDevice));
  launch_pad_letters(SeqSet{d->res.as<uint8_t>(), d->offs.as<uint32_t>()}, n,
                     d->pad.ensure<uint8_t>(pad_letters_bytes(n, total)), d->pad_offs.ensure<uint32_t>(n + 1), st_);
  DMR_HIP(hipStreamSynchronize(st_));
  d->n = n;
  d->letters = total;
  d->stat_letters = stat_letters ? *stat_letters : total;
  d->max_len = maxl;
  db_ = std::move(d);
  idx_.reset();
}

void Engine::upload_seed_params(const std::string& sens, double cut, DevBuf* params, uint32_t* nshapes,
                                int* key_bits, int* shape_bits) {
  const auto& codes = shape_codes(sens);
  ReductionDef red = parse_reduction(reduction_def(sens));
  SeedParams P;
  std::memset(&P, 0, sizeof P);
  P.nshapes = (uint32_t)codes.size();
  P.red_size = red.size;
  std::memcpy(P.red_map, red.map, 25);
  P.cut = cut;
  P.lnfact[0] = P.lnfact[1] = 0.0;
  for (int k = 2; k < 33; ++k) P.lnfact[k] = P.lnfact[k - 1] + std::log((double)k);
  int kb = 1;
  for (size_t s = 0; s < codes.size(); ++s) {
    ShapeDef sd;
    if (!parse_shape(codes[s], &sd)) throw Error(-1, "invalid shape " + codes[s]);
    P.shapes[s].weight = sd.weight;
    P.shapes[s].span = sd.span;
    std::memcpy(P.shapes[s].pos, sd.pos, sizeof sd.pos);
    // max key = size^weight - 1
    long double mx = std::pow((long double)red.size, (long double)sd.weight) - 1;
    if (mx >= 18446744073709551615.0L) throw Error(-5, "seed key does not fit 64 bits");
    kb = std::max(kb, bits_for((uint64_t)mx));
  }
  *nshapes = P.nshapes;
  *key_bits = kb;
  *shape_bits = P.nshapes > 1 ? bits_for(P.nshapes - 1) : 0;
  if (kb + *shape_bits > 64) throw Error(-5, "tagged seed key does not fit 64 bits");
  DMR_HIP(hipMemcpyAsync(params->ensure<SeedParams>(1), &P, sizeof P, hipMemcpyHostToDevice, st_));
}

uint64_t Engine::index_entries() const {
  uint64_t n = 0;
  if (idx_)
    for (auto& c : idx_->chunks) n += c.n;
  return n;
}

// Resident index: per DB chunk (src/seed.rs:146-164), all shapes enumerated
// and sorted by tagged key, plus a bucket table over the top key bits. One
// index per preset is kept with the database; the first engine to ask for a
// preset builds it (the others wait on the database's lock).
void Engine::build_index(const std::string& sens, int index_chunks, double seed_cut) {
  DMR_HIP(hipSetDevice(device_));
  double cutv = (seed_cut >= 0 ? seed_cut : preset_seed_cut(sens));
  int nch = index_chunks >= 0 ? std::max(index_chunks, 1) : preset_index_chunks(sens);
  const auto& codes = shape_codes(sens);
  ShapeDef s0;
  parse_shape(codes[0], &s0);
  // cfg.seed_complexity_cut = seed_cut * ln 2 * shapes[0].weight (src/search.rs:102-105)
  double cut = cutv * 0.6931471805599453 * (double)s0.weight;
  char keybuf[128];
  std::snprintf(keybuf, sizeof keybuf, "%s|%d|%.17g", sens.c_str(), nch, cut);
  const std::string key = keybuf;
  DbState& D = *db_;
  std::lock_guard<std::mutex> lock(D.m);
  for (auto& kv : D.indexes)
    if (kv.first == key) {
      idx_ = kv.second;
      return;
    }
  double t0 = now_ms();
  auto I = std::make_shared<IndexState>();
  upload_seed_params(sens, cut, &I->seed_params, &I->nshapes, &I->key_bits, &I->shape_bits);
  const int total_bits = I->key_bits + I->shape_bits;
  // chunks(n, parts)
  std::vector<std::pair<uint64_t, uint64_t>> ranges;
  {
    uint64_t n = D.n, parts = std::min<uint64_t>(std::max(nch, 1), std::max<uint64_t>(n, 1));
    uint64_t start = 0;
    for (uint64_t p = 0; p < parts && n > 0; ++p) {
      uint64_t end = (p + 1 == parts) ? n : start + n / parts + (p < n % parts ? 1 : 0);
      ranges.push_back({start, end});
      start = end;
    }
  }
  I->chunks.resize(ranges.size());
  std::vector<RefChunkDev> desc(ranges.size());
  for (size_t c = 0; c < ranges.size(); ++c) {
    uint64_t r0 = D.h_offs[ranges[c].first], r1 = D.h_offs[ranges[c].second];
    uint32_t nb = seed_enum_blocks(r0, r1);
    uint64_t nbk = (uint64_t)nb * I->nshapes + 1;
    uint32_t* bc = b_counts_.ensure<uint32_t>(nbk);
    launch_fill_u32(bc + nbk - 1, 1, 0, st_);
    uint64_t* bo = b_off_.ensure<uint64_t>(nbk);
    uint64_t* tmp = b_scan_.ensure<uint64_t>(scan_tmp_elems(nbk));
    // Count first, then allocate exactly (an index can be tens of GB).
    IndexChunk& ch = I->chunks[c];
    launch_seed_count(D.res.as<uint8_t>(), D.seq_of.as<uint32_t>(), D.offs.as<uint32_t>(), r0, r1,
                      I->seed_params.as<SeedParams>(), I->nshapes, I->key_bits, bc, bo, tmp, st_);
    uint64_t n = read1(bo + nbk - 1);
    ch.n = n;
    uint64_t* keys = ch.keys.ensure<uint64_t>(n + 1);
    uint32_t* locs = ch.locs.ensure<uint32_t>(n + 1);
    launch_seed_emit(D.res.as<uint8_t>(), D.seq_of.as<uint32_t>(), D.offs.as<uint32_t>(), r0, r1,
                     I->seed_params.as<SeedParams>(), I->nshapes, I->key_bits, bo, keys, locs, st_);
    uint64_t* k2 = qkeys2_.ensure<uint64_t>(n + 1);
    uint32_t* l2 = qlocs2_.ensure<uint32_t>(n + 1);
    uint32_t* hist = b_hist_.ensure<uint32_t>(radix_hist_elems(n));
    uint64_t* stmp = b_scan_.ensure<uint64_t>(radix_scan_tmp_elems(n));
    radix_sort_pairs(keys, locs, k2, l2, n, total_bits, hist, stmp, st_);
    int tb = std::min(total_bits, 22);
    ch.shift = total_bits - tb;
    uint64_t maxkey = (((uint64_t)(I->nshapes - 1)) << I->key_bits) |
                      ((I->key_bits >= 64) ? ~0ull : ((1ull << I->key_bits) - 1));
    ch.nbuckets = (maxkey >> ch.shift) + 1;
    launch_bucket_table(keys, n, ch.shift, ch.nbuckets, ch.table.ensure<uint64_t>(ch.nbuckets + 1), st_);
    desc[c] = RefChunkDev{keys, locs, ch.table.as<uint64_t>(), n, ch.nbuckets, ch.shift, 0};
  }
  DMR_HIP(hipMemcpyAsync(I->chunk_desc.ensure<RefChunkDev>(desc.size() + 1), desc.data(),
                         desc.size() * sizeof(RefChunkDev), hipMemcpyHostToDevice, st_));
  DMR_HIP(hipStreamSynchronize(st_));
  // The sort's second buffers were sized for a whole chunk; searches size
  // them for their query seeds again.
  qkeys2_.release();
  qlocs2_.release();
  I->build_ms = now_ms() - t0;
  D.indexes.push_back({key, I});
  idx_ = I;
}

void Engine::load_targets(const TargetSet& t) {
  DMR_HIP(hipSetDevice(device_));
  tset_ = t;
  q_n_ = (uint32_t)(t.offs.size() - 1);
  q_total_ = t.offs.back();
  if ((uint64_t)q_total_ + q_n_ + 2 >= (1ull << 32)) throw Error(-5, "queries larger than 4G letters");
  q_max_len_ = 0;
  std::vector<uint32_t> seq_of(q_total_ + 1);
  for (uint32_t i = 0; i < q_n_; ++i) {
    q_max_len_ = std::max(q_max_len_, t.offs[i + 1] - t.offs[i]);
    for (uint32_t k = t.offs[i]; k < t.offs[i + 1]; ++k) seq_of[k] = i;
  }
  DMR_HIP(hipMemcpyAsync(q_res_.ensure<uint8_t>(q_total_ + 64), t.letters.data(), q_total_,
                         hipMemcpyHostToDevice, st_));
  DMR_HIP(hipMemcpyAsync(q_offs_.ensure<uint32_t>(q_n_ + 1), t.offs.data(), (q_n_ + 1) * 4,
                         hipMemcpyHostToDevice, st_));
  DMR_HIP(hipMemcpyAsync(q_seq_of_.ensure<uint32_t>(q_total_ + 1), seq_of.data(), q_total_ * 4,
                         hipMemcpyHostToDevice, st_));
  // The record of each target (frames of one read share it).
  if (t.query_of.size() != q_n_) throw Error(-1, "target set without a record per target");
  DMR_HIP(hipMemcpyAsync(q_group_.ensure<uint32_t>(q_n_ + 1), t.query_of.data(), q_n_ * 4ull,
                         hipMemcpyHostToDevice, st_));
  launch_pad_letters(SeqSet{q_res_.as<uint8_t>(), q_offs_.as<uint32_t>()}, q_n_,
                     q_pad_.ensure<uint8_t>(pad_letters_bytes(q_n_, q_total_)),
                     q_pad_offs_.ensure<uint32_t>(q_n_ + 1), st_);
  DMR_HIP(hipStreamSynchronize(st_));
}

// Reads to the device, their six frames translated (src/translate.rs:54-82)
// and masked (src/translate.rs:101-121) into r_masked_, with each frame's
// count of letters below X in r_good_; *foff receives the 6n + 1 frame
// offsets. Returns the number of frame letters.
uint64_t Engine::frames_dev(const char* reads, const uint64_t* roffs, uint32_t n, const uint32_t* min_orf,
                            std::vector<uint64_t>* foff) {
  const uint64_t base = roffs[0], nt = roffs[n] - base;
  std::vector<uint64_t> ro(n + 1);
  foff->assign(6ull * n + 1, 0);
  for (uint32_t r = 0; r <= n; ++r) ro[r] = roffs[r] - base;
  for (uint32_t r = 0; r < n; ++r) {
    if (ro[r + 1] < ro[r]) throw Error(-1, "read offsets are not monotone");
    const uint64_t L = ro[r + 1] - ro[r];
    for (int k = 0; k < 6; ++k) {
      const uint64_t o = (uint64_t)(k % 3);
      (*foff)[6ull * r + k + 1] = (*foff)[6ull * r + k] + (L >= o ? (L - o) / 3 : 0);
    }
  }
  const uint64_t total = foff->back();
  DMR_HIP(hipMemcpyAsync(r_nt_.ensure<uint8_t>(nt + 1), reads + base, nt, hipMemcpyHostToDevice, st_));
  DMR_HIP(hipMemcpyAsync(r_off_.ensure<uint64_t>(n + 1), ro.data(), (n + 1) * 8, hipMemcpyHostToDevice, st_));
  DMR_HIP(hipMemcpyAsync(r_minorf_.ensure<uint32_t>(n + 1), min_orf, (uint64_t)n * 4, hipMemcpyHostToDevice, st_));
  DMR_HIP(hipMemcpyAsync(r_foff_.ensure<uint64_t>(6ull * n + 1), foff->data(), (6ull * n + 1) * 8,
                         hipMemcpyHostToDevice, st_));
  uint8_t* fr = r_frames_.ensure<uint8_t>(total + 1);
  launch_translate_frames(r_nt_.as<uint8_t>(), r_off_.as<uint64_t>(), n, r_foff_.as<uint64_t>(), fr, st_);
  launch_mask_orfs(r_foff_.as<uint64_t>(), 6 * n, r_minorf_.as<uint32_t>(), fr, r_masked_.ensure<uint8_t>(total + 1),
                   r_good_.ensure<uint32_t>(6ull * n + 1), st_);
  return total;
}

void Engine::six_frames(const char* reads, const uint64_t* roffs, uint32_t n, const uint32_t* min_orf,
                        std::vector<uint8_t>* letters, std::vector<uint64_t>* offsets) {
  DMR_HIP(hipSetDevice(device_));
  letters->clear();
  offsets->assign(1, 0);
  if (!n) return;
  const uint64_t total = frames_dev(reads, roffs, n, min_orf, offsets);
  letters->resize(total);
  if (total)
    DMR_HIP(hipMemcpyAsync(letters->data(), r_masked_.as<uint8_t>(), total, hipMemcpyDeviceToHost, st_));
  DMR_HIP(hipStreamSynchronize(st_));
}

void Engine::load_reads(const char* reads, const uint64_t* roffs, uint32_t n, const uint32_t* min_orf,
                        TargetSet* host) {
  DMR_HIP(hipSetDevice(device_));
  const double t0 = now_ms();
  *host = TargetSet{};
  host->offs.push_back(0);
  std::vector<uint64_t> foff;
  const uint64_t total = n ? frames_dev(reads, roffs, n, min_orf, &foff) : 0;
  const uint32_t nf = 6 * n;
  uint64_t ntargets = 0, nletters = 0;
  if (nf) {
    // Kept frames (a letter below X), their target numbers and letter offsets.
    uint32_t* keep = r_keep_.ensure<uint32_t>(nf + 1);
    uint32_t* klen = r_klen_.ensure<uint32_t>(nf + 1);
    launch_kept_lengths(r_foff_.as<uint64_t>(), r_good_.as<uint32_t>(), nf, keep, klen, st_);
    launch_fill_u32(keep + nf, 1, 0, st_);
    launch_fill_u32(klen + nf, 1, 0, st_);
    uint64_t* kept = r_kept_.ensure<uint64_t>(nf + 1);
    uint64_t* tlet = r_tlet_.ensure<uint64_t>(nf + 1);
    scan_exclusive_u32_to_u64(keep, kept, nf + 1, b_scan_.ensure<uint64_t>(scan_tmp_elems(nf + 1)), st_);
    scan_exclusive_u32_to_u64(klen, tlet, nf + 1, b_scan_.ensure<uint64_t>(scan_tmp_elems(nf + 1)), st_);
    uint64_t tail[2];
    DMR_HIP(hipMemcpyAsync(&tail[0], kept + nf, 8, hipMemcpyDeviceToHost, st_));
    DMR_HIP(hipMemcpyAsync(&tail[1], tlet + nf, 8, hipMemcpyDeviceToHost, st_));
    DMR_HIP(hipStreamSynchronize(st_));
    ntargets = tail[0];
    nletters = tail[1];
    if (nletters + ntargets + 2 >= (1ull << 32)) throw Error(-5, "query batch larger than 4G letters");
    uint32_t* toffs = q_offs_.ensure<uint32_t>(ntargets + 1);
    launch_compact_frames(r_foff_.as<uint64_t>(), r_good_.as<uint32_t>(), kept, tlet, r_masked_.as<uint8_t>(),
                          q_res_.ensure<uint8_t>(nletters + 64), toffs, r_qof_.ensure<uint32_t>(ntargets + 1),
                          r_fof_.ensure<uint8_t>(ntargets + 1), nf, st_);
    const uint32_t end = (uint32_t)nletters;
    DMR_HIP(hipMemcpyAsync(toffs + ntargets, &end, 4, hipMemcpyHostToDevice, st_));
    host->letters.resize(nletters);
    host->offs.resize(ntargets + 1);
    host->query_of.resize(ntargets);
    host->frame_of.resize(ntargets);
    if (nletters)
      DMR_HIP(hipMemcpyAsync(host->letters.data(), q_res_.as<uint8_t>(), nletters, hipMemcpyDeviceToHost, st_));
    DMR_HIP(hipMemcpyAsync(host->offs.data(), toffs, (ntargets + 1) * 4, hipMemcpyDeviceToHost, st_));
    if (ntargets) {
      DMR_HIP(hipMemcpyAsync(host->query_of.data(), r_qof_.as<uint32_t>(), ntargets * 4, hipMemcpyDeviceToHost, st_));
      DMR_HIP(hipMemcpyAsync(host->frame_of.data(), r_fof_.as<uint8_t>(), ntargets, hipMemcpyDeviceToHost, st_));
    }
    DMR_HIP(hipStreamSynchronize(st_));
  }
  (void)total;
  // The device copies load_targets would make, from the device.
  tset_ = *host;
  q_n_ = (uint32_t)ntargets;
  q_total_ = nletters;
  q_max_len_ = 0;
  for (uint64_t t = 0; t < ntargets; ++t) q_max_len_ = std::max(q_max_len_, host->offs[t + 1] - host->offs[t]);
  if (!ntargets) launch_fill_u32(q_offs_.ensure<uint32_t>(1), 1, 0, st_);
  if (ntargets)
    DMR_HIP(hipMemcpyAsync(q_group_.ensure<uint32_t>(ntargets + 1), r_qof_.as<uint32_t>(), ntargets * 4,
                           hipMemcpyDeviceToDevice, st_));
  launch_seq_of(q_offs_.as<uint32_t>(), q_n_, q_seq_of_.ensure<uint32_t>(q_total_ + 1), st_);
  launch_pad_letters(SeqSet{q_res_.ensure<uint8_t>(q_total_ + 64), q_offs_.as<uint32_t>()}, q_n_,
                     q_pad_.ensure<uint8_t>(pad_letters_bytes(q_n_, q_total_)),
                     q_pad_offs_.ensure<uint32_t>(q_n_ + 1), st_);
  DMR_HIP(hipStreamSynchronize(st_));
  translate_ms_ = now_ms() - t0;
}

std::vector<std::vector<HspRec>> Engine::search_targets(
    const SearchOptions& opt, StageStats* stats, std::vector<std::shared_ptr<const OpsStore>>* ops_stores) {
  DMR_HIP(hipSetDevice(device_));
  StageStats S;
  double t_start = now_ms();
  std::vector<std::vector<HspRec>> result(q_n_);
  slices_ = 0;
  last_njobs_ = 0;
  last_slices_ = 0;
  if (q_n_ == 0 || db_->n == 0 || q_total_ == 0) {
    if (stats) *stats = S;
    return result;
  }
  build_index(opt.sensitivity, opt.index_chunks, opt.seed_cut);
  const AlpParams P = alp(opt.gap_open, opt.gap_extend);
  const int total_bits = idx_->key_bits + idx_->shape_bits;
  SeqSet qs{q_res_.as<uint8_t>(), q_offs_.as<uint32_t>(), q_pad_.as<uint8_t>(), q_pad_offs_.as<uint32_t>()};
  SeqSet dbs{db_->res.as<uint8_t>(), db_->offs.as<uint32_t>(), db_->pad.as<uint8_t>(),
             db_->pad_offs.as<uint32_t>()};
  stage_event(0);

  // ---- query seeds (all shapes, tagged, sorted) ---------------------------
  // Counts stay on the device up to the join's hit total, the one host read
  // before the hits (it sizes their buffer): the seeds and runs are sized by
  // their bound (every letter of every shape seeds) and the kernels read the
  // counts where they need them.
  uint64_t* dc = counts_dev_.ensure<uint64_t>(kCountSlots);
  uint32_t nb = seed_enum_blocks(0, q_total_);
  uint64_t nbk = (uint64_t)nb * idx_->nshapes + 1;
  uint32_t* bc = b_counts_.ensure<uint32_t>(nbk);
  launch_fill_u32(bc + nbk - 1, 1, 0, st_);
  uint64_t* bo = b_off_.ensure<uint64_t>(nbk);
  const uint64_t nQb = q_total_ * idx_->nshapes;  // bound of the query seeds
  uint64_t* qk = qkeys_.ensure<uint64_t>(nQb + 1);
  uint32_t* ql = qlocs_.ensure<uint32_t>(nQb + 1);
  launch_seed_enumerate(q_res_.as<uint8_t>(), q_seq_of_.as<uint32_t>(), q_offs_.as<uint32_t>(), 0,
                        q_total_, idx_->seed_params.as<SeedParams>(), idx_->nshapes, idx_->key_bits, bc, bo,
                        b_scan_.ensure<uint64_t>(scan_tmp_elems(nbk)), qk, ql, st_);
  launch_copy_u64(bo + nbk - 1, dc + kCountSeeds, 1, st_);
  const uint64_t* nQd = dc + kCountSeeds;
  {
    uint64_t* k2 = qkeys2_.ensure<uint64_t>(nQb + 1);
    uint32_t* l2 = qlocs2_.ensure<uint32_t>(nQb + 1);
    radix_sort_pairs(qk, ql, k2, l2, nQb, total_bits, b_hist_.ensure<uint32_t>(radix_hist_elems(nQb)),
                     b_scan_.ensure<uint64_t>(radix_scan_tmp_elems(nQb)), st_, nQd);
  }
  stage_event(1);

  // ---- join: runs of equal query keys x DB chunks --------------------------
  // nq counts a key's entries over the whole batch (src/search.rs:109-131),
  // or over one record when opt.per_record (a record searched alone).
  uint32_t* fl = flags_.ensure<uint32_t>(nQb + 1);
  if (opt.per_record)
    launch_run_flags_grouped(qk, ql, q_seq_of_.as<uint32_t>(), q_group_.as<uint32_t>(), nQb, fl, st_, nQd);
  else
    launch_run_flags(qk, nQb, fl, st_, nQd);
  launch_fill_u32(fl + nQb, 1, 0, st_);
  uint64_t* ps = pos_.ensure<uint64_t>(nQb + 1);
  scan_exclusive_u32_to_u64(fl, ps, nQb + 1, b_scan_.ensure<uint64_t>(scan_tmp_elems(nQb + 1)), st_);
  launch_copy_u64(ps + nQb, dc + kCountRuns, 1, st_);
  const uint64_t* nrd = dc + kCountRuns;
  // Query runs stay in their own buffer: every slice reads them.
  uint32_t* rs = qruns_.ensure<uint32_t>(nQb + 1);
  launch_compact_flags(fl, ps, nQb, rs, st_);
  const uint32_t nch = (uint32_t)idx_->chunks.size();
  const uint64_t npairs = nQb * nch;  // bound of the (run, chunk) pairs
  uint32_t* cnt = jc_counts_.ensure<uint32_t>(npairs + 1);
  uint64_t* rlo = jc_rlo_.ensure<uint64_t>(npairs + 1);
  launch_join_count(qk, rs, nQb, nQb, idx_->chunk_desc.as<RefChunkDev>(), nch, opt.max_expand, cnt, rlo, st_,
                    nrd, nQd);
  launch_fill_u32(cnt + npairs, 1, 0, st_);
  // All hits at once when they fit the slice budget; else slices of targets
  // by their hit counts (bounded device memory; the cap decisions above are
  // the batch's, so slicing changes no result).
  uint64_t* hoff = jc_off_.ensure<uint64_t>(npairs + 1);
  scan_exclusive_u32_to_u64(cnt, hoff, npairs + 1, b_scan_.ensure<uint64_t>(scan_tmp_elems(npairs + 1)), st_);
  launch_copy_u64(hoff + npairs, dc + kCountHits, 1, st_);
  uint64_t* hc = counts_pin_.ensure<uint64_t>(kCountSlots);
  DMR_HIP(hipMemcpyAsync(hc, dc, 3 * sizeof(uint64_t), hipMemcpyDeviceToHost, st_));
  DMR_HIP(hipStreamSynchronize(st_));
  const uint64_t nQ = hc[kCountSeeds], nruns = hc[kCountRuns], all_hits = hc[kCountHits];
  S.query_seeds = nQ;
  S.runs = nruns;
  const uint64_t budget = opt.slice_hits ? opt.slice_hits : slice_hits_;
  std::vector<std::pair<uint32_t, uint32_t>> slices;
  std::vector<uint64_t> slice_total;
  if (all_hits <= budget) {
    slices.push_back({0, q_n_});
    slice_total.push_back(all_hits);
  } else {
    uint64_t* thits = thits_.ensure<uint64_t>(q_n_ + 1);
    launch_fill_u32(reinterpret_cast<uint32_t*>(thits), 2ull * q_n_, 0, st_);
    launch_target_hits(ql, rs, nQb, nQb, nch, cnt, q_seq_of_.as<uint32_t>(), thits, st_, nrd, nQd);
    uint64_t* h_thits = thits_pin_.ensure<uint64_t>(q_n_ + 1);
    DMR_HIP(hipMemcpyAsync(h_thits, thits, q_n_ * 8ull, hipMemcpyDeviceToHost, st_));
    DMR_HIP(hipStreamSynchronize(st_));
    uint32_t t0 = 0;
    uint64_t acc = 0;
    for (uint32_t t = 0; t < q_n_; ++t) {
      if (t > t0 && acc + h_thits[t] > budget) {
        slices.push_back({t0, t});
        slice_total.push_back(acc);
        t0 = t;
        acc = 0;
      }
      acc += h_thits[t];
    }
    slices.push_back({t0, q_n_});
    slice_total.push_back(acc);
  }
  stage_event(2);
  DMR_HIP(hipStreamSynchronize(st_));
  S.ms_seed = stage_ms(0, 1);
  S.ms_join = stage_ms(1, 2);
  slices_ = (uint32_t)slices.size();
  for (uint64_t h : slice_total) S.hits += h;
  HitGeom g;
  g.bt = bits_for(q_n_ > 0 ? q_n_ - 1 : 0);
  g.bs = bits_for(db_->n > 0 ? db_->n - 1 : 0);
  g.bp = bits_for(db_->max_len);
  g.bd = bits_for((uint64_t)q_max_len_ + db_->max_len);
  g.diag_bias = db_->max_len;
  const int hit_bits = (int)(g.bt + g.bs + g.bd + g.bp);
  if (hit_bits > 64) throw Error(-5, "hit key does not fit 64 bits for this batch");
  for (int c = 0; c < kClasses; ++c) last_ms_class_[c] = 0;
  if (profile_) {
    for (int c = 0; c < kClasses; ++c) prof_jobs_[c] = prof_cells_[c] = 0;
    prof_dp_cells_ = 0;
  }
  for (size_t sl = 0; sl < slices.size(); ++sl) {
    const uint32_t t0 = slices[sl].first, t1 = slices[sl].second;
    const uint64_t nhits = slice_total[sl];
    stage_event(2);
    uint64_t* hits = hits_.ensure<uint64_t>(nhits + 1);
    // The emission writes exactly the hits its offsets' scan totals; a
    // slice's total is checked against its per-target count before any write.
    const bool whole = slices.size() == 1;
    uint32_t* scnt = whole ? cnt : jc_scnt_.ensure<uint32_t>(npairs + 1);
    uint32_t* sst = whole ? nullptr : jc_sst_.ensure<uint32_t>(npairs + 1);
    uint32_t* snq = whole ? nullptr : jc_snq_.ensure<uint32_t>(npairs + 1);
    if (!whole) {
      launch_join_slice(ql, rs, nQb, nQb, nch, cnt, q_seq_of_.as<uint32_t>(), t0, t1, scnt, sst, snq, st_, nrd,
                        nQd);
      launch_fill_u32(scnt + npairs, 1, 0, st_);
      scan_exclusive_u32_to_u64(scnt, hoff, npairs + 1, b_scan_.ensure<uint64_t>(scan_tmp_elems(npairs + 1)), st_);
      if (read1(hoff + npairs) != nhits) throw Error(-2, "join offsets disagree with the per-target hit counts");
    }
    launch_join_emit(ql, rs, nQb, nQb, idx_->chunk_desc.as<RefChunkDev>(), nch, scnt, rlo, hoff, nhits,
                     q_seq_of_.as<uint32_t>(), q_offs_.as<uint32_t>(), db_->seq_of.as<uint32_t>(),
                     db_->offs.as<uint32_t>(), g, hits, st_, sst, snq, nrd, nQd);
    stage_event(11);
    radix_sort_keys(hits, hits2_.ensure<uint64_t>(nhits + 1), nhits, hit_bits,
                    b_hist_.ensure<uint32_t>(radix_hist_elems(nhits)),
                    b_scan_.ensure<uint64_t>(radix_scan_tmp_elems(nhits)), st_);
    stage_event(3);
    search_slice(opt, P, g, hits, nhits, qs, dbs, &S, &result, ops_stores);
    S.ms_join += stage_ms(2, 11);
    S.ms_sort_hits += stage_ms(11, 3);
  }
  last_slices_ = slices_;
  prof_valid_ = profile_;
  S.ms_total = now_ms() - t_start;
  if (stats) *stats = S;
  return result;
}

// One slice of targets, from its sorted hits to its HSPs in result[target]:
// ungapped extension, bands, the score pass, selection, traceback.
void Engine::search_slice(const SearchOptions& opt, const AlpParams& P, const HitGeom& g, const uint64_t* hits,
                          uint64_t nhits, SeqSet qs, SeqSet dbs, StageStats* Sp,
                          std::vector<std::vector<HspRec>>* result_out,
                          std::vector<std::shared_ptr<const OpsStore>>* ops_stores) {
  StageStats& S = *Sp;
  std::vector<std::vector<HspRec>>& result = *result_out;
  const int xdrop = raw_score(P, 12.3);  // --xdrop 12.3 bits (src/search.rs:51-53)
  const int cutoff = 19;                  // src/search.rs:164
  // The chain from the sorted hits to the score pass's host read, run
  // directly or (layout key "graph") captured and replayed as one graph.
  SwJob* jobs = nullptr;
  int32_t* scores = nullptr;
  uint64_t* hc = nullptr;
  StageStats SS;
  auto chain = [&]() -> uint64_t {
    // ---- ungapped: runs of (target, subject, diagonal) -----------------------
    // Every later count (hit runs, segments, groups, jobs) is at most the
    // hits': buffers are sized by nhits and the counts stay on the device until
    // the score pass's host read.
    uint64_t* dc = counts_dev_.ensure<uint64_t>(kCountSlots);
    uint32_t* fl = flags_.ensure<uint32_t>(nhits + 1);
    launch_key_flags(hits, nhits, (int)g.bp, fl, st_);
    launch_fill_u32(fl + nhits, 1, 0, st_);
    uint64_t* ps = pos_.ensure<uint64_t>(nhits + 1);
    scan_exclusive_u32_to_u64(fl, ps, nhits + 1, b_scan_.ensure<uint64_t>(scan_tmp_elems(nhits + 1)), st_);
    launch_copy_u64(ps + nhits, dc + kCountHitRuns, 1, st_);
    uint32_t* hrs = runs_.ensure<uint32_t>(nhits + 1);
    launch_compact_flags(fl, ps, nhits, hrs, st_);
    SegRec* sraw = segs_raw_.ensure<SegRec>(nhits + 1);
    uint32_t* keep = keep_.ensure<uint32_t>(nhits + 1);
    launch_xdrop_runs(hits, hrs, nhits, nhits, g, qs, dbs, matrix32_.as<int8_t>(), xdrop, cutoff, sraw, keep, st_,
                      dc + kCountHitRuns);
    launch_fill_u32(keep + nhits, 1, 0, st_);
    scan_exclusive_u32_to_u64(keep, ps, nhits + 1, b_scan_.ensure<uint64_t>(scan_tmp_elems(nhits + 1)), st_);
    launch_copy_u64(ps + nhits, dc + kCountSegs, 1, st_);
    SegRec* segs = segs_.ensure<SegRec>(nhits + 1);
    uint64_t* gk = gkeys_.ensure<uint64_t>(nhits + 1);
    launch_compact_segments(sraw, hits, keep, ps, nhits, (int)(g.bd + g.bp), segs, gk, st_);
    stage_event(4);

    // ---- bands per (target, subject) -----------------------------------------
    launch_key_flags(gk, nhits, 0, fl, st_, dc + kCountSegs);
    launch_fill_u32(fl + nhits, 1, 0, st_);
    scan_exclusive_u32_to_u64(fl, ps, nhits + 1, b_scan_.ensure<uint64_t>(scan_tmp_elems(nhits + 1)), st_);
    launch_copy_u64(ps + nhits, dc + kCountGroups, 1, st_);
    uint32_t* gs = gstart_.ensure<uint32_t>(nhits + 1);
    launch_compact_flags(fl, ps, nhits, gs, st_);
    uint32_t* nbd = nbands_.ensure<uint32_t>(nhits + 1);
    SwJob* tj = tmpjobs_.ensure<SwJob>(nhits + 1);
    launch_bands(segs, gk, gs, nhits, nhits, g.bs, qs, dbs, bd0_.ensure<int32_t>(nhits + 1),
                 bd1_.ensure<int32_t>(nhits + 1), bbest_.ensure<uint32_t>(nhits + 1), tj, nbd, st_,
                 dc + kCountGroups, dc + kCountSegs);
    launch_fill_u32(nbd + nhits, 1, 0, st_);
    uint64_t* jo = joboff_.ensure<uint64_t>(nhits + 1);
    scan_exclusive_u32_to_u64(nbd, jo, nhits + 1, b_scan_.ensure<uint64_t>(scan_tmp_elems(nhits + 1)), st_);
    launch_copy_u64(jo + nhits, dc + kCountJobs, 1, st_);
    jobs = jobs_.ensure<SwJob>(nhits + 1);
    launch_jobs_compact(tj, gs, nbd, jo, nhits, jobs, st_, dc + kCountGroups);
    hc = counts_pin_.ensure<uint64_t>(kCountSlots);
    DMR_HIP(hipMemcpyAsync(hc + kCountHitRuns, dc + kCountHitRuns, 4 * sizeof(uint64_t), hipMemcpyDeviceToHost,
                           st_));
    stage_event(5);

    // ---- banded DP, score pass -----------------------------------------------
    scores = scores_.ensure<int32_t>(nhits + 1);
    stage_event(6);
    SS = StageStats();
    return score_pass(jobs, nhits, (int)std::max(g.bt, g.bs), qs, dbs, opt, scores, &SS, dc + kCountJobs);
  };
  uint64_t njobs = 0;
  bool direct = true;
  if (nhits > 0 && nhits <= graph_cap_ && resolve_layout(opt.sensitivity).graph) {
    DMR_HIP(hipStreamBeginCapture(st_, hipStreamCaptureModeThreadLocal));
    capturing_ = true;
    g_capture_active = true;
    auto abandon = [&] {
      g_capture_active = false;
      if (!capturing_) return;
      capturing_ = false;
      hipGraph_t gr = nullptr;
      (void)hipStreamEndCapture(st_, &gr);
      if (gr) (void)hipGraphDestroy(gr);
      (void)hipGetLastError();
    };
    try {
      njobs = chain();  // the score pass ends the capture and launches the graph
      direct = false;
    } catch (const CaptureGrow&) {  // a buffer must grow: run the chain directly
      abandon();
      graph_cap_ = 0;
    } catch (...) {
      abandon();
      throw;
    }
  }
  if (direct) {
    njobs = chain();
    graph_cap_ = std::max(graph_cap_, nhits);
  }
  if (!nhits) DMR_HIP(hipStreamSynchronize(st_));
  stage_event(7);
  const uint64_t nsegs = hc[kCountSegs];
  S.hit_runs += hc[kCountHitRuns];
  S.segments += nsegs;
  S.groups += hc[kCountGroups];
  S.jobs += njobs;
  if (hc[kCountJobs] != njobs) throw Error(-2, "score pass job count disagrees with the bands'");
  S.sw16_lanes += SS.sw16_lanes;
  S.sw_generic += SS.sw_generic;
  S.sw_redo += SS.sw_redo;
  for (int c = 0; c < kClasses; ++c) {
    S.ms_class[c] += SS.ms_class[c];
    last_ms_class_[c] += SS.ms_class[c];
  }
  last_njobs_ = njobs;
  if (profile_) accumulate_profile(njobs);

  // ---- selection by E-value -------------------------------------------------
  uint32_t* fl = flags_.ensure<uint32_t>(njobs + 1);
  launch_evalue_filter(jobs, scores, njobs, qs, dbs, P, db_->stat_letters, opt.evalue, fl, st_);
  launch_fill_u32(fl + njobs, 1, 0, st_);
  uint64_t* ps = pos_.ensure<uint64_t>(njobs + 1);
  scan_exclusive_u32_to_u64(fl, ps, njobs + 1, b_scan_.ensure<uint64_t>(scan_tmp_elems(njobs + 1)), st_);
  const uint64_t nmaybe = read1(ps + njobs);
  S.maybe += nmaybe;
  uint32_t* midx = sel_idx_.ensure<uint32_t>(nmaybe + 1);
  launch_compact_flags(fl, ps, njobs, midx, st_);
  // The candidates (job index, job, score) come back through pinned memory
  // and are read in place (valid until the next search).
  const uint64_t pin_job = (nmaybe * 4 + 15) / 16 * 16;
  const uint64_t pin_sc = pin_job + (nmaybe * sizeof(SwJob) + 15) / 16 * 16;
  uint8_t* spin = sel_pin_.ensure<uint8_t>(pin_sc + nmaybe * 4 + 16);
  uint32_t* h_midx = reinterpret_cast<uint32_t*>(spin);
  SwJob* h_mjobs = reinterpret_cast<SwJob*>(spin + pin_job);
  int32_t* h_mscore = reinterpret_cast<int32_t*>(spin + pin_sc);
  if (nmaybe) {
    SwJob* gj = tmpjobs_.ensure<SwJob>(std::max<uint64_t>(nmaybe, nsegs) + 1);
    int32_t* gs = sel_scores_.ensure<int32_t>(nmaybe + 1);
    launch_gather(jobs, scores, midx, nmaybe, gj, gs, st_);
    DMR_HIP(hipMemcpyAsync(h_midx, midx, nmaybe * 4, hipMemcpyDeviceToHost, st_));
    DMR_HIP(hipMemcpyAsync(h_mjobs, gj, nmaybe * sizeof(SwJob), hipMemcpyDeviceToHost, st_));
    DMR_HIP(hipMemcpyAsync(h_mscore, gs, nmaybe * 4, hipMemcpyDeviceToHost, st_));
  }
  stage_event(8);
  DMR_HIP(hipStreamSynchronize(st_));

  // ---- traceback, launched before the selection -------------------------------
  // The reported DPs are among the candidates with a positive score (H6), so
  // all of those are traced while the host computes the exact E-values and
  // selects; the few that are not selected are dropped at assembly. The
  // score pass's best scores (by job) let each traceback DP stop at the
  // row of its first maximum.
  std::vector<uint32_t> tb_list, tb_slen;
  std::vector<SwJob> tb_jobs;
  std::vector<uint64_t> tb_pos(nmaybe, ~0ull);  // maybe index -> traced index
  for (uint64_t k = 0; k < nmaybe; ++k) {
    if (h_mscore[k] <= 0) continue;
    tb_pos[k] = tb_list.size();
    tb_list.push_back(h_midx[k]);
    tb_jobs.push_back(h_mjobs[k]);
    tb_slen.push_back(db_->h_offs[h_mjobs[k].subject + 1] - db_->h_offs[h_mjobs[k].subject]);
  }
  stage_event(9);
  TbRun tb_run;
  traceback_launch(&tb_run, jobs, tb_list, tb_jobs, tb_slen, qs, dbs, opt.gap_open, opt.gap_extend, scores);

  double th0 = now_ms();
  // Exact E-values of the candidates (glibc f64, as the reference), computed
  // in parallel; the selection below reads them in order.
  std::vector<double> h_ev(nmaybe, 0.0);
  host_parallel_for(nmaybe, 256, [&](uint64_t b, uint64_t e) {
    for (uint64_t k = b; k < e; ++k) {
      const SwJob& j = h_mjobs[k];
      const int32_t sc = h_mscore[k];
      if (sc <= 0) continue;
      const uint64_t qlen = tset_.offs[j.target + 1] - tset_.offs[j.target];
      const uint64_t slen = db_->h_offs[j.subject + 1] - db_->h_offs[j.subject];
      h_ev[k] = alp_evalue(P, sc, qlen, slen, db_->stat_letters);
    }
  });
  // Per group: bands in rank order; the first max_hsps whose exact E-value
  // passes are kept (src/search.rs:232-255).
  struct Cand {
    uint64_t job;  // index into the maybe list
    HspRec h;
  };
  std::vector<std::vector<std::vector<Cand>>> per_target(q_n_);
  std::vector<uint32_t> touched;
  {
    uint64_t k = 0;
    while (k < nmaybe) {
      uint32_t grp = h_mjobs[k].group;
      std::vector<Cand> hs;
      for (; k < nmaybe && h_mjobs[k].group == grp; ++k) {
        if (hs.size() >= opt.max_hsps) continue;
        const SwJob& j = h_mjobs[k];
        int32_t sc = h_mscore[k];
        if (sc <= 0) continue;
        const double e = h_ev[k];
        if (e <= opt.evalue) {
          Cand c;
          c.job = k;
          c.h.query = j.target;
          c.h.subject = j.subject;
          c.h.frame = 0;
          c.h.score = sc;
          c.h.evalue = e;
          c.h.bits = bit_score(P, sc);
          hs.push_back(std::move(c));
        }
      }
      if (!hs.empty()) {
        std::stable_sort(hs.begin(), hs.end(),
                         [](const Cand& a, const Cand& b) { return hsp_less(a.h, b.h); });
        const uint32_t tq = hs[0].h.query;
        if (per_target[tq].empty()) touched.push_back(tq);
        per_target[tq].push_back(std::move(hs));
      }
    }
  }
  uint64_t nsel = 0;
  for (uint32_t tq : touched) {
    auto& found = per_target[tq];
    std::stable_sort(found.begin(), found.end(),
                     [](const std::vector<Cand>& a, const std::vector<Cand>& b) {
                       return hsp_less(a[0].h, b[0].h);
                     });
    if (found.size() > opt.max_target_seqs) found.resize(opt.max_target_seqs);
    for (auto& hs : found) {
      if (hs.size() > opt.max_hsps) hs.resize(opt.max_hsps);
      nsel += hs.size();
    }
  }
  S.ms_host += now_ms() - th0;
  S.selected += nsel;

  std::vector<SwOut> h_touts;
  std::vector<OpsRef> h_tops;
  std::shared_ptr<OpsStore> tb_store;
  uint64_t tcells = 0;
  traceback_collect(&tb_run, &h_touts, &h_tops, &tb_store, &tcells);
  S.traceback_cells += tcells;
  ops_stores->push_back(tb_store);
  stage_event(10);
  DMR_HIP(hipStreamSynchronize(st_));

  // ---- assemble HSPs in align_target order ----------------------------------
  for (uint32_t tq : touched) {
    for (auto& hs : per_target[tq]) {
      for (auto& c : hs) {
        const SwJob& j = h_mjobs[c.job];
        const uint64_t k = tb_pos[c.job];  // selected DPs have a positive score
        const SwOut& o = h_touts[k];
        HspRec h = c.h;
        h.q_begin = j.q0 + o.q_begin;
        h.q_end = j.q0 + o.best_i;
        h.s_begin = o.s_begin;
        h.s_end = o.best_j;
        h.identities = o.identities;
        h.positives = o.positives;
        h.mismatches = o.mismatches;
        h.gap_openings = o.gap_openings;
        h.gaps = o.gaps;
        h.length = o.length;
        h.ops = h_tops[k];
        if (o.score != c.h.score)
          throw Error(-2, "traceback score differs from the score pass");
        result[h.query].push_back(std::move(h));
      }
    }
  }
  S.ms_ungapped += stage_ms(3, 4);
  S.ms_bands += stage_ms(4, 5);
  S.ms_sw += stage_ms(6, 7);
  S.ms_select += stage_ms(7, 8);
  S.ms_traceback += stage_ms(9, 10);
}

// Score pass over all DP jobs: jobs of the standard band widths go to the
// packed int16 kernel (two DPs of one query per lane), the rest (merged or
// clipped bands, other gap costs) and any int16-saturated DP to the exact
// int32 kernel.
uint64_t Engine::score_pass(const SwJob* jobs, uint64_t njobs, int seq_bits, SeqSet qs, SeqSet dbs,
                            const SearchOptions& opt, int32_t* scores, StageStats* S, const uint64_t* njobs_dev) {
  for (int c = 0; c < kClasses; ++c) S->ms_class[c] = 0;
  if (!njobs) return 0;
  layout_ = resolve_layout(opt.sensitivity);
  layout_preset_ = opt.sensitivity;
  const ScoreLayout L = layout_;
  const uint64_t mmap = sw16_merged_map(L.merged_mask);
  const bool fast = opt.gap_open == 11 && opt.gap_extend == 1 && !force_generic_;
  if (fast && (!qs.pad || !qs.pad_offs || !dbs.pad || !dbs.pad_offs))
    throw Error(-2, "score pass without padded letters");
  uint64_t* keys = sw_keys_.ensure<uint64_t>(njobs + 1);
  uint32_t* idx = sw_idx_.ensure<uint32_t>(njobs + 1);
  // With njobs_dev, njobs bounds the count until the summary read below.
  const uint32_t transpose = fast ? transpose_modes(L) : 0u;
  launch_classify(jobs, njobs, qs, dbs, seq_bits, mmap, transpose, keys, idx, st_, njobs_dev);
  static_assert(kClasses <= 32, "the class takes 5 key bits");
  const int key_bits = 5 + 1 + seq_bits + kRowKeyBits;
  radix_sort_pairs(keys, idx, sw_keys2_.ensure<uint64_t>(njobs + 1), sw_idx2_.ensure<uint32_t>(njobs + 1),
                   njobs, key_bits, b_hist_.ensure<uint32_t>(radix_hist_elems(njobs)),
                   b_scan_.ensure<uint64_t>(radix_scan_tmp_elems(njobs)), st_, njobs_dev);
  // Runs of equal (class, orientation, row sequence) and their starts; lane
  // heads (even positions of a run in a packed class), their scan, the lanes
  // and the per-class summary, all on the device; one small read brings the
  // class ranges and lane counts back.
  uint32_t* fl = flags_.ensure<uint32_t>(njobs + 1);
  launch_key_flags(keys, njobs, kRowKeyBits, fl, st_, njobs_dev);
  launch_fill_u32(fl + njobs, 1, 0, st_);
  uint64_t* ps = pos_.ensure<uint64_t>(njobs + 1);
  scan_exclusive_u32_to_u64(fl, ps, njobs + 1, b_scan_.ensure<uint64_t>(scan_tmp_elems(njobs + 1)), st_);
  uint32_t* rs = runs_.ensure<uint32_t>(njobs + 1);
  launch_compact_flags(fl, ps, njobs, rs, st_);
  const int cls_shift = seq_bits + 1 + kRowKeyBits;
  const int npacked = fast ? kPackedClasses : 0;
  // Summary: class ranges, lane ranges and the lane total, then per packed
  // class the query-row lane count (kSumRow) and the first subject-row job
  // (kSumTr).
  constexpr int kSumRow = 2 * (kClasses + kPackedClasses) + 1, kSumTr = kSumRow + kPackedClasses;
  constexpr int kSummary = kSumTr + kPackedClasses;
  uint32_t* summ = sw_summary_.ensure<uint32_t>(kSummary);
  launch_fill_u32(summ, kSummary, 0, st_);
  uint32_t* heads = sw_heads_.ensure<uint32_t>(njobs + 1);
  launch_lane_heads(keys, fl, ps, rs, njobs, cls_shift, npacked, kClasses, kSumTr, heads, summ, st_, njobs_dev);
  launch_fill_u32(heads + njobs, 1, 0, st_);
  uint64_t* hpos = sw_hpos_.ensure<uint64_t>(njobs + 1);
  scan_exclusive_u32_to_u64(heads, hpos, njobs + 1, b_scan_.ensure<uint64_t>(scan_tmp_elems(njobs + 1)), st_);
  uint2* paired = sw_lanes2_.ensure<uint2>(njobs + 1);
  launch_make_lanes(idx, fl, heads, hpos, njobs, paired, st_, njobs_dev);
  launch_lane_summary(hpos, njobs, npacked, kClasses, kSumTr, kSumRow, summ, st_);
  uint32_t* h_summ = sw_pin_summary_.ensure<uint32_t>(kSummary);
  DMR_HIP(hipMemcpyAsync(h_summ, summ, kSummary * 4, hipMemcpyDeviceToHost, st_));
  uint64_t* h_njobs = sw_pin_count_.ensure<uint64_t>(1);
  if (njobs_dev) DMR_HIP(hipMemcpyAsync(h_njobs, njobs_dev, 8, hipMemcpyDeviceToHost, st_));
  if (capturing_) end_capture_and_launch();
  DMR_HIP(hipStreamSynchronize(st_));
  if (njobs_dev) {
    if (*h_njobs > njobs) throw Error(-2, "score pass: job count past its bound");
    njobs = *h_njobs;
    if (!njobs) return 0;
  }
  // Class ranges in the sorted order; lanes of the packed classes.
  uint64_t cbeg[kClasses], cend[kClasses];
  for (int c = 0; c < kClasses; ++c) {
    cbeg[c] = h_summ[c];
    cend[c] = h_summ[kClasses + c];
  }
  uint32_t class_lanes[kPackedClasses] = {}, class_lane0[kPackedClasses] = {}, class_rows[kPackedClasses] = {};
  for (int c = 0; c < npacked; ++c) {
    class_lane0[c] = h_summ[2 * kClasses + c];
    class_lanes[c] = h_summ[2 * kClasses + npacked + c];
    class_rows[c] = h_summ[kSumRow + c];
  }
  const uint32_t nl = h_summ[2 * (kClasses + npacked)];
  // Counters of the score pass: the redo list's length, one task counter per
  // packed and wave class, the fallback list's length, the redo kernel's
  // task counter.
  constexpr int kCntRedo = 0, kCntPacked = 1, kCntWave = kCntPacked + kPackedClasses,
                kCntFallback = kCntWave + kWaveKinds, kCntRedoWave = kCntFallback + 1, kCntRedoWide = kCntRedoWave + 1,
                kCntWords = kCntRedoWide + 1;
  uint32_t* cnt = sw_counters_.ensure<uint32_t>(kCntWords);
  launch_fill_u32(cnt, kCntWords, 0, st_);
  uint32_t* redo = sw_redo_.ensure<uint32_t>(njobs + 1);
  uint32_t* fallback = sw_fallback_.ensure<uint32_t>(njobs + 1);
  const int cus = cu_count_;
  uint2* lanes = nullptr;
  if (nl) {
    // Lanes of each class in order of row span, longest first: a wave (or a
    // block of lane groups) then steps lanes of about the same length, and
    // the longest start first. Classes stay contiguous (class is the major
    // key), so class_lane0/class_lanes still hold.
    uint64_t* lk = sw_keys_.as<uint64_t>();
    uint32_t* li = sw_idx2_.as<uint32_t>();
    launch_lane_keys(paired, nl, jobs, qs, dbs, mmap, transpose, lk, li, st_);
    static_assert(kPackedClasses <= 16, "the lane class takes 4 key bits");
    radix_sort_pairs(lk, li, sw_keys2_.as<uint64_t>(), sw_lane_idx_.ensure<uint32_t>(nl + 1), nl,
                     4 + 1 + kLaneSpanBits, b_hist_.ensure<uint32_t>(radix_hist_elems(nl)),
                     b_scan_.ensure<uint64_t>(radix_scan_tmp_elems(nl)), st_);
    lanes = sw_lanes_.ensure<uint2>(nl + 1);
    launch_gather_lanes(paired, li, nl, lanes, st_);
    if (lane_report_ && qs.offs == q_offs_.as<uint32_t>())
      report_lanes(jobs, njobs, lanes, nl, class_lane0, class_lanes, class_rows, L);
  }
  // The class kernels are independent (distinct jobs and counters; the redo
  // and fallback lists are filled through atomics): they run side by side on
  // the engine stream and the three auxiliary streams. The classes of few,
  // long DPs (wide wave classes, merged bands) are latency-bound and start
  // first on streams of their own; the VALU-bound packed classes follow one
  // another on a third (DMR_SCORE_PLAN selects other arrangements).
  fork();
  // Fused launch (tuning key `fuse`): the classes of DMR_SW16_FUSED in one
  // persistent grid; they are not launched on their own. Its time counts for
  // each of them.
  bool fused[kPackedClasses] = {};
  const bool fuse = L.fuse && nl && !score_serial_;
  if (fuse)
    for (int k = 0; k < kFusedKinds; ++k) fused[sw16_fused_class(k)] = true;
  const int fused_slot = sw16_fused_class(0);  // its events and task counter
  auto launch_fused = [&](hipStream_t s) {
    FusedSections fs = {};
    bool any = false;
    for (int k = 0; k < kFusedKinds; ++k) {
      const int c = sw16_fused_class(k);
      fs.lanes[k] = lanes + class_lane0[c];
      fs.nlanes[k] = class_lanes[c];
      fs.nrow[k] = class_rows[c];
      any |= class_lanes[c] > 0;
    }
    if (!any) return;
    DMR_HIP(hipEventRecord(ev_[12 + 2 * fused_slot], s));
    if (!launch_sw16_fused(L.f16, fs, jobs, qs, dbs, (L.f16 ? pair_tab16_ : pair_tab_).as<int>(),
                           cnt + kCntPacked + fused_slot, scores, redo, cnt + kCntRedo, cus, s))
      throw Error(-1, "fused packed launch: lane counts out of range");
    DMR_HIP(hipEventRecord(ev_[13 + 2 * fused_slot], s));
  };
  auto launch_packed = [&](int c, hipStream_t s) {
    if (!nl || !class_lanes[c] || fused[c]) return;
    DMR_HIP(hipEventRecord(ev_[12 + 2 * c], s));
    if (!launch_sw16(c, L.lanes[c], L.threads[c], L.f16, lanes + class_lane0[c], class_lanes[c], class_rows[c],
                     jobs, qs, dbs, (L.f16 ? pair_tab16_ : pair_tab_).as<int>(), matrix32_.as<int8_t>(),
                     cnt + kCntPacked + c, scores, redo, cnt + kCntRedo, cus, s))
      throw Error(-1, "no compiled packed kernel for class " + packed_class_name(c));
    DMR_HIP(hipEventRecord(ev_[13 + 2 * c], s));
  };
  // Wave-per-DP int32 kernel: classes 7-12 (and every width class when the
  // gap costs have no 16-bit table). A class's jobs are one range of the
  // sorted order, used in place; only the fallback for other gap costs
  // concatenates ranges.
  auto launch_wave = [&](int c, hipStream_t s) {
    const int K = 1 << (c - kPackedClasses);
    std::vector<std::pair<uint64_t, uint64_t>> ranges;
    if (cend[c] > cbeg[c]) ranges.push_back({cbeg[c], cend[c]});
    if (!fast)
      for (int w = 0; w < kPackedClasses; ++w)
        if (cend[w] > cbeg[w] && swwave_k_for(sw16_width(w)) == K) ranges.push_back({cbeg[w], cend[w]});
    uint64_t total = 0;
    for (auto& rg : ranges) total += rg.second - rg.first;
    if (!total) return;
    const uint32_t* list = idx + ranges[0].first;
    if (ranges.size() > 1) {
      uint32_t* cat = sw_klist_[c - kPackedClasses].ensure<uint32_t>(total + 1);
      uint64_t o = 0;
      for (auto& rg : ranges) {
        DMR_HIP(hipMemcpyAsync(cat + o, idx + rg.first, (rg.second - rg.first) * 4, hipMemcpyDeviceToDevice, s));
        o += rg.second - rg.first;
      }
      list = cat;
    }
    DMR_HIP(hipEventRecord(ev_[12 + 2 * c], s));
    launch_sw_wave(K, false, jobs, list, (uint32_t)total, qs, dbs, matrix32_.as<int8_t>(),
                   opt.gap_open, opt.gap_extend, cnt + kCntWave + (c - kPackedClasses), nullptr, scores, nullptr,
                   nullptr, nullptr, nullptr, fallback, cnt + kCntFallback, cus, s);
    DMR_HIP(hipEventRecord(ev_[13 + 2 * c], s));
  };
  static_assert(kExactClasses == 5 && kWaveKinds == 6, "the launch plan names every class");
  const int W0 = kPackedClasses;  // first wave class (K = 1)
  auto launch_merged = [&](hipStream_t s) {  // widest first
    for (int c = kPackedClasses - 1; c >= kExactClasses; --c) launch_packed(c, s);
  };
  if (score_serial_) {
    for (int c = kPackedClasses - 1; c >= 0; --c) launch_packed(c, st_);
    for (int c = W0; c < W0 + kWaveKinds; ++c) launch_wave(c, st_);
  } else if (L.plan == 2) {
    // Every throughput class on a stream of its own beside the others.
    for (int c = W0 + kWaveKinds - 1; c >= W0 + 3; --c) launch_wave(c, st_);  // K = 32, 16, 8
    launch_merged(aux_[2]);
    if (fuse) launch_fused(aux_[1]);
    launch_packed(4, aux_[0]);
    launch_packed(3, aux_[1]);
    for (int c = 2; c >= 0; --c) launch_packed(c, aux_[2]);
    for (int c = W0 + 2; c >= W0; --c) launch_wave(c, aux_[1]);  // K = 4, 2, 1
  } else {
    // The throughput-bound band_for classes one after another on one stream
    // (plan 0: the smaller classes first, then the widest, which then has the
    // chip nearly to itself; plan 1: widest first); the merged bands and the
    // latency-bound classes (few, long DPs) beside them on the other streams.
    for (int c = W0 + kWaveKinds - 1; c >= W0 + 3; --c) launch_wave(c, st_);  // K = 32, 16, 8
    launch_merged(aux_[2]);
    if (fuse) launch_fused(aux_[0]);
    if (L.plan == 1)
      for (int c = 4; c >= 0; --c) launch_packed(c, aux_[0]);
    else
      for (int c : {3, 2, 1, 0, 4}) launch_packed(c, aux_[0]);
    for (int c = W0 + 2; c >= W0; --c) launch_wave(c, aux_[1]);  // K = 4, 2, 1
  }
  join();
  // Bands wider than a wave covers: one block of 16 waves per DP (widths up
  // to 16384; wider ones, and DPs whose best may lie in the clamped rows, go
  // on to the fallback list of the row-exact kernel).
  const uint64_t nwide = cend[kRowExact] > cbeg[kRowExact] ? cend[kRowExact] - cbeg[kRowExact] : 0;
  float wide_ms = 0;
  if (nwide) {
    DMR_HIP(hipEventRecord(ev_[12 + 2 * kRowExact], st_));
    launch_sw_block(false, jobs, idx + cbeg[kRowExact], (uint32_t)nwide, qs, dbs, matrix32_.as<int8_t>(),
                    opt.gap_open, opt.gap_extend, nullptr, scores, nullptr, nullptr, nullptr, nullptr, fallback,
                    cnt + kCntFallback,
                    nullptr, 0, st_);
    DMR_HIP(hipEventRecord(ev_[13 + 2 * kRowExact], st_));
  }
  // DPs at the 16-bit limit (f16 1900, int16 32000, or a best below 11 with
  // clamped rows that may score): re-scored by the int32 wave kernel, bands
  // up to 256 wide with four offsets per lane and, side by side on an
  // auxiliary stream, the tagged wider ones (the m385 and m513 classes) with
  // eight. They read the list's length on the device, so no host read
  // precedes them; one read afterwards brings the counts.
  // Tagged entries exist only when a wide class ran, and then its re-run
  // is always launched.
  const bool wide_redo = nl && (class_lanes[kPackedClasses - 1] || class_lanes[kPackedClasses - 2]);
  const bool side = wide_redo && !score_serial_;
  if (side) fork();
  if (wide_redo)
    launch_sw_wave(8, false, jobs, redo, (uint32_t)njobs, qs, dbs, matrix32_.as<int8_t>(), opt.gap_open,
                   opt.gap_extend, cnt + kCntRedoWide, nullptr, scores, nullptr, nullptr, nullptr, nullptr, fallback,
                   cnt + kCntFallback, cus, side ? aux_[0] : st_, nullptr, cnt + kCntRedo, 2);
  launch_sw_wave(4, false, jobs, redo, (uint32_t)njobs, qs, dbs, matrix32_.as<int8_t>(), opt.gap_open,
                 opt.gap_extend, cnt + kCntRedoWave, nullptr, scores, nullptr, nullptr, nullptr, nullptr, fallback,
                 cnt + kCntFallback, cus, st_, nullptr, cnt + kCntRedo, 1);
  if (side) join();
  uint32_t* h_cnt = sw_pin_summary_.ensure<uint32_t>(kCntWords);
  DMR_HIP(hipMemcpyAsync(h_cnt, cnt, kCntWords * 4, hipMemcpyDeviceToHost, st_));
  DMR_HIP(hipStreamSynchronize(st_));
  const uint32_t nr16 = h_cnt[kCntRedo], nfb = h_cnt[kCntFallback];
  for (int c = 0; c < kRowExact; ++c)
    if (c < kPackedClasses && fused[c] && class_lanes[c])
      S->ms_class[c] = stage_ms(12 + 2 * fused_slot, 13 + 2 * fused_slot);
    else if ((c < kPackedClasses && nl && class_lanes[c]) ||
             (c >= kPackedClasses && swk_launched(c, cbeg, cend, fast)))
      S->ms_class[c] = stage_ms(12 + 2 * c, 13 + 2 * c);
  if (nwide) wide_ms = stage_ms(12 + 2 * kRowExact, 13 + 2 * kRowExact);
  S->ms_class[kRowExact] = wide_ms;
  // Row-exact int32 kernel: the DPs the other kernels handed on (bands wider
  // than the block kernel's reach, and DPs whose band never reaches the last
  // in-band row but whose clamped rows may score).
  const uint64_t nlist = nfb;
  S->sw16_lanes = nl;
  S->sw_redo = nr16;
  S->sw_generic = nlist;
  last_cbeg_.assign(cbeg, cbeg + kClasses);
  last_cend_.assign(cend, cend + kClasses);
  if (nlist) {
    uint32_t* list = sw_list_.ensure<uint32_t>(nlist + 1);
    DMR_HIP(hipMemcpyAsync(list, fallback, nfb * 4, hipMemcpyDeviceToDevice, st_));
    uint32_t* stride = stride_.ensure<uint32_t>(nlist + 1);
    launch_job_geometry(jobs, list, nlist, dbs, nullptr, nullptr, stride, st_);
    uint64_t* soff = scr_off_.ensure<uint64_t>(nlist + 1);
    launch_u32_times(stride, nlist, 4, soff, st_);
    scan_exclusive_u64(soff, soff, nlist + 1, b_scan_.ensure<uint64_t>(scan_tmp_elems(nlist + 1)), st_);
    const uint64_t scratch_total = read1(soff + nlist);
    DMR_HIP(hipEventRecord(ev_[12 + 2 * kRowExact], st_));
    launch_sw_generic(jobs, list, nlist, qs, dbs, matrix32_.as<int8_t>(), opt.gap_open,
                      opt.gap_extend, scratch_.ensure<int32_t>(scratch_total + 1), soff, nullptr,
                      nullptr, nullptr, nullptr, nullptr, false, st_, /*scatter=*/true, scores);
    DMR_HIP(hipEventRecord(ev_[13 + 2 * kRowExact], st_));
    DMR_HIP(hipEventSynchronize(ev_[13 + 2 * kRowExact]));
    S->ms_class[kRowExact] = wide_ms + stage_ms(12 + 2 * kRowExact, 13 + 2 * kRowExact);
  }
  return njobs;
}

// Diagnostics (DMR_LANE_REPORT=1): per packed class, the band cells of its
// DPs against the cell slots its lanes and waves step through (a lane runs
// the union of its two DPs' rows; a wave runs as many rows as its longest
// lane or lane group; the first class_rows lanes of a class have the query
// on the rows, the others the subject, and no wave mixes the two).
void Engine::report_lanes(const SwJob* jobs, uint64_t njobs, const uint2* lanes, uint32_t nl,
                          const uint32_t* class_lane0, const uint32_t* class_lanes, const uint32_t* class_rows,
                          const ScoreLayout& L) {
  std::vector<SwJob> hj(njobs);
  std::vector<uint2> hl(nl);
  DMR_HIP(hipMemcpyAsync(hj.data(), jobs, njobs * sizeof(SwJob), hipMemcpyDeviceToHost, st_));
  DMR_HIP(hipMemcpyAsync(hl.data(), lanes, nl * sizeof(uint2), hipMemcpyDeviceToHost, st_));
  DMR_HIP(hipStreamSynchronize(st_));
  auto rows = [&](uint32_t k, bool tr, int64_t* lo, int64_t* hi) {
    const SwJob& j = hj[k];
    const int64_t n = (int64_t)j.q1 - j.q0;
    const int64_t m = (int64_t)db_->h_offs[j.subject + 1] - db_->h_offs[j.subject];
    *lo = tr ? std::max<int64_t>(1, 1 - (int64_t)j.d1) : std::max<int64_t>(1, (int64_t)j.d0 + 1);
    *hi = tr ? std::min<int64_t>(m, n - (int64_t)j.d0) : std::min<int64_t>(n, m + j.d1);
  };
  for (int c = 0; c < kPackedClasses; ++c) {
    if (!class_lanes[c]) continue;
    const int64_t W = sw16_width(c);
    const int P = L.lanes[c];
    const int per_wave = 64 / P;
    uint64_t band = 0, pair_slots = 0, wave_slots = 0;
    int64_t wave_max = 0, in_wave = 0;
    for (uint32_t t = 0; t < class_lanes[c]; ++t) {
      const uint2 ln = hl[class_lane0[c] + t];
      const bool tr = t >= class_rows[c];
      if (t == class_rows[c] && in_wave) {  // the subject-row lanes start a wave
        wave_slots += (uint64_t)wave_max * W * 2 * per_wave;
        wave_max = 0;
        in_wave = 0;
      }
      int64_t lo = INT64_MAX, hi = INT64_MIN, a, b;
      for (uint32_t k : {ln.x, ln.y}) {
        if (k == ~0u) continue;
        rows(k, tr, &a, &b);
        band += (uint64_t)std::max<int64_t>(0, b - a + 1) * W;
        lo = std::min(lo, a);
        hi = std::max(hi, b);
      }
      const int64_t span = std::max<int64_t>(0, hi - lo + 1);
      pair_slots += (uint64_t)span * W * 2;
      wave_max = std::max(wave_max, span);
      if (++in_wave == per_wave || t + 1 == class_lanes[c]) {
        wave_slots += (uint64_t)wave_max * W * 2 * per_wave;
        wave_max = 0;
        in_wave = 0;
      }
    }
    std::fprintf(stderr, "[lanes] class %d (W %lld, %d lanes per pair): %u pairs (%u subject rows), band cells %.3g, "
                 "pair slots %.3g (x%.2f), wave slots %.3g (x%.2f)\n", c, (long long)W, P, class_lanes[c],
                 class_lanes[c] - class_rows[c],
                 (double)band, (double)pair_slots, (double)pair_slots / std::max<double>(1, band),
                 (double)wave_slots, (double)wave_slots / std::max<double>(1, band));
  }
}

bool Engine::swk_launched(int c, const uint64_t* cbeg, const uint64_t* cend, bool fast) const {
  if (cend[c] > cbeg[c]) return true;
  if (fast) return false;
  for (int w = 0; w < kPackedClasses; ++w)
    if (cend[w] > cbeg[w] && swwave_k_for(sw16_width(w)) == (1 << (c - kPackedClasses))) return true;
  return false;
}

// Traceback DPs for jobs[list[k]] (hjobs[k] is the host copy, slen[k] the
// subject length). Wave-per-DP kernel by band width; the row-exact kernel for
// bands wider than 512 and for DPs whose best cell may sit in the clamped
// rows below the band. known (optional, device): the score pass's scores by
// job; with them each wave-kernel DP stops at the row of its first maximum.
void Engine::traceback_pass(const SwJob* djobs, const std::vector<uint32_t>& list,
                            const std::vector<SwJob>& hjobs, const std::vector<uint32_t>& slen,
                            SeqSet qs, SeqSet dbs, int go, int ge, std::vector<SwOut>* outs,
                            std::vector<std::vector<uint32_t>>* ops, uint64_t* cells_out,
                            const int32_t* known) {
  TbRun run;
  traceback_launch(&run, djobs, list, hjobs, slen, qs, dbs, go, ge, known);
  std::vector<OpsRef> refs;
  std::shared_ptr<OpsStore> store;
  traceback_collect(&run, outs, &refs, &store, cells_out);
  ops->resize(refs.size());
  for (size_t k = 0; k < refs.size(); ++k) (*ops)[k].assign(refs[k].begin(), refs[k].end());
}

// First half of traceback_pass: every wave-kernel DP launched, its ops packed
// and the first round of result copies queued. The host may work until
// traceback_collect; list, hjobs and slen must stay alive until then.
void Engine::traceback_launch(TbRun* run, const SwJob* djobs, const std::vector<uint32_t>& list,
                              const std::vector<SwJob>& hjobs, const std::vector<uint32_t>& slen,
                              SeqSet qs, SeqSet dbs, int go, int ge, const int32_t* known) {
  const uint64_t ns = list.size();
  *run = TbRun{};
  run->djobs = djobs;
  run->list = &list;
  run->hjobs = &hjobs;
  run->slen = &slen;
  run->qs = qs;
  run->dbs = dbs;
  run->go = go;
  run->ge = ge;
  run->known = known;
  run->ns = ns;
  if (!ns) return;
  // Group g < kWaveKinds: wave kernel with K = 2^g; the last group: row-exact.
  std::vector<std::vector<uint64_t>>& group = run->group;
  group.assign(kTbGroups, {});
  for (uint64_t k = 0; k < ns; ++k) {
    const SwJob& j = hjobs[k];
    const int K = swwave_k_for((int64_t)j.d1 - j.d0 + 1);
    int g = kWaveKinds;
    if (K) g = __builtin_ctz((unsigned)K);
    group[g].push_back(k);
  }
  uint32_t* cnt = tb_counters_.ensure<uint32_t>(2 * kWaveKinds);
  launch_fill_u32(cnt, 2 * kWaveKinds, 0, st_);
  const int cus = cu_count_;
  uint64_t cells = 0;
  // Tasks of all wave groups in group order (group g from out_base[g]).
  uint64_t* out_base = run->out_base;
  uint64_t nout = 0;
  for (int g = 0; g < kWaveKinds; ++g) {
    out_base[g] = nout;
    nout += group[g].size();
  }
  run->nout = nout;
  // One pinned block goes to the device in one copy: per group [list u32 x
  // ni, padded to 8 bytes][trace offsets u64 x ni+1], then the ops offsets of
  // all groups' tasks (u64 x nout+1, one array: ops and packed ops are laid
  // out over all groups).
  uint64_t* in_off = run->in_off;
  uint64_t in_words = 0;  // in u64 units
  for (int g = 0; g < kWaveKinds; ++g) {
    const uint64_t ni = group[g].size();
    if (!ni) continue;
    in_off[g] = in_words;
    in_words += (ni + 1) / 2 + (ni + 1);
  }
  run->ooff_at = in_words;
  in_words += nout + 1;
  uint64_t* pin_in = tb_pin_in_.ensure<uint64_t>(in_words + 1);
  uint64_t* ooff = pin_in + run->ooff_at;
  ooff[0] = 0;
  for (int g = 0; g < kWaveKinds; ++g) {
    auto& items = group[g];
    const uint64_t ni = items.size();
    if (!ni) continue;
    // Rows and cells of each job's band; the longest jobs start first so
    // they do not finish last (sorted on precomputed keys, ties by index).
    std::vector<std::pair<uint64_t, uint64_t>> order(ni);  // (cells, item)
    for (uint64_t t = 0; t < ni; ++t) {
      const SwJob& j = hjobs[items[t]];
      const int64_t n = (int64_t)j.q1 - j.q0, m = slen[items[t]];
      const int64_t rlo = std::max<int64_t>(1, (int64_t)j.d0 + 1), rhi = std::min<int64_t>(n, m + j.d1);
      const int64_t rows = rhi >= rlo ? rhi - rlo + 1 : 0;
      order[t] = {(uint64_t)(rows * ((int64_t)j.d1 - j.d0 + 1)), items[t]};
    }
    std::sort(order.begin(), order.end(), [](const std::pair<uint64_t, uint64_t>& a,
                                             const std::pair<uint64_t, uint64_t>& b) {
      return a.first != b.first ? a.first > b.first : a.second < b.second;
    });
    uint32_t* jl = reinterpret_cast<uint32_t*>(pin_in + in_off[g]);
    uint64_t* toff = pin_in + in_off[g] + (ni + 1) / 2;
    toff[0] = 0;
    const int64_t stride = 32ll << g;  // trace row: 32 * K bytes, 4 bits per cell
    for (uint64_t t = 0; t < ni; ++t) {
      items[t] = order[t].second;
      const SwJob& j = hjobs[items[t]];
      jl[t] = list[items[t]];
      const int64_t n = (int64_t)j.q1 - j.q0, m = slen[items[t]];
      const int64_t w = (int64_t)j.d1 - j.d0 + 1;
      const int64_t rows = w > 0 ? (int64_t)(order[t].first / (uint64_t)w) : 0;
      toff[t + 1] = toff[t] + (uint64_t)(rows * stride);
      const uint64_t T = out_base[g] + t;
      ooff[T + 1] = ooff[T] + (uint64_t)(n + m + 1);
      cells += order[t].first;
    }
  }
  uint64_t* din = tb_in_.ensure<uint64_t>(in_words + 1);
  DMR_HIP(hipMemcpyAsync(din, pin_in, in_words * 8, hipMemcpyHostToDevice, st_));
  const uint64_t* dooff = din + run->ooff_at;
  SwOut* dout_all = tb_out_all_.ensure<SwOut>(nout + 1);
  uint32_t* dops_all = tb_ops_all_.ensure<uint32_t>(ooff[nout] + 1);
  // The groups run side by side on the auxiliary streams, widest first (its
  // DPs run longest).
  fork();
  for (int g = kWaveKinds - 1; g >= 0; --g) {
    const uint64_t ni = group[g].size();
    if (!ni) continue;
    // The two largest groups (K = 8, K = 4) on streams of their own, the
    // widest two (few DPs) on the engine stream, K = 1, 2 on the third.
    static const int kStreamOf[kWaveKinds] = {2, 2, 1, 0, -1, -1};
    hipStream_t s = kStreamOf[g] < 0 ? st_ : aux_[kStreamOf[g]];
    const uint32_t* dl = reinterpret_cast<const uint32_t*>(din + in_off[g]);
    const uint64_t* dt = din + in_off[g] + (ni + 1) / 2;
    const uint64_t* htoff = pin_in + in_off[g] + (ni + 1) / 2;
    uint8_t* tr = tb_trace_[g].ensure<uint8_t>(htoff[ni] + 1);
    uint32_t* fb = tb_fb_[g].ensure<uint32_t>(ni + 1);
    launch_sw_wave(1 << g, true, djobs, dl, (uint32_t)ni, qs, dbs, matrix32_.as<int8_t>(), go, ge,
                   cnt + g, dout_all + out_base[g], nullptr, tr, dt, dops_all, dooff + out_base[g], fb,
                   cnt + kWaveKinds + g, cus, s, known);
  }
  join();
  // Results: the ops of all groups packed on the device (n_ops words per
  // DP, one array), then outputs, pack offsets, packed ops and fallback
  // lists through pinned memory: two rounds of copies, one wait each (the
  // first queued here, the rest in traceback_collect).
  SwOut* pin_out = tb_pin_out_.ensure<SwOut>(nout + 1);
  uint64_t* pin_poff = tb_pin_poff_.ensure<uint64_t>(nout + 1);
  uint32_t* pin_fb = tb_pin_fb_.ensure<uint32_t>(nout + kWaveKinds + 1);
  DMR_HIP(hipMemcpyAsync(pin_fb, cnt + kWaveKinds, kWaveKinds * 4, hipMemcpyDeviceToHost, st_));
  if (nout) {
    uint64_t* poff = tb_poff_all_.ensure<uint64_t>(nout + 1);
    launch_ops_pack(dout_all, dops_all, dooff, (uint32_t)nout, tb_sizes_.ensure<uint32_t>(nout + 1), poff,
                    b_scan_.ensure<uint64_t>(scan_tmp_elems(nout + 1)), tb_pack_all_.ensure<uint32_t>(ooff[nout] + 1),
                    st_);
    DMR_HIP(hipMemcpyAsync(pin_out, dout_all, nout * sizeof(SwOut), hipMemcpyDeviceToHost, st_));
    DMR_HIP(hipMemcpyAsync(pin_poff, poff, (nout + 1) * 8, hipMemcpyDeviceToHost, st_));
  }
  run->cells = cells;
}

// Second half of traceback_pass: waits for the wave-kernel DPs, brings their
// ops back and runs the row-exact fallback.
void Engine::traceback_collect(TbRun* run, std::vector<SwOut>* outs, std::vector<OpsRef>* ops,
                               std::shared_ptr<OpsStore>* store_out, uint64_t* cells_out) {
  const uint64_t ns = run->ns;
  outs->assign(ns, SwOut{});
  ops->assign(ns, OpsRef{});
  *store_out = std::make_shared<OpsStore>();
  if (cells_out) *cells_out = 0;
  if (!ns) return;
  const std::vector<uint32_t>& list = *run->list;
  const std::vector<SwJob>& hjobs = *run->hjobs;
  const std::vector<uint32_t>& slen = *run->slen;
  const std::vector<std::vector<uint64_t>>& group = run->group;
  const uint64_t* out_base = run->out_base;
  const uint64_t nout = run->nout;
  const SwJob* djobs = run->djobs;
  const SeqSet qs = run->qs, dbs = run->dbs;
  const int go = run->go, ge = run->ge;
  uint64_t cells = run->cells;
  std::vector<uint64_t> fallback_items;
  SwOut* pin_out = tb_pin_out_.as<SwOut>();
  uint64_t* pin_poff = tb_pin_poff_.as<uint64_t>();
  uint32_t* pin_fb = tb_pin_fb_.as<uint32_t>();
  DMR_HIP(hipStreamSynchronize(st_));
  uint32_t nfb[kWaveKinds];
  std::copy(pin_fb, pin_fb + kWaveKinds, nfb);
  uint64_t fb_base[kWaveKinds] = {}, nfbt = 0;
  for (int g = 0; g < kWaveKinds; ++g) {
    fb_base[g] = nfbt;
    nfbt += group[g].empty() ? 0 : nfb[g];
  }
  const uint64_t nops = nout ? pin_poff[nout] : 0;
  // The packed ops go straight into the pinned block the results keep.
  auto store = std::make_shared<OpsStore>(ops_pool_, nops + 1);
  *store_out = store;
  uint32_t* pin_ops = store->block();
  uint32_t* pin_fbl = tb_pin_fb_.ensure<uint32_t>(nfbt + 1);  // counts already read
  if (nops)
    DMR_HIP(hipMemcpyAsync(pin_ops, tb_pack_all_.as<uint32_t>(), nops * 4, hipMemcpyDeviceToHost, st_));
  for (int g = 0; g < kWaveKinds; ++g)
    if (!group[g].empty() && nfb[g])
      DMR_HIP(hipMemcpyAsync(pin_fbl + fb_base[g], tb_fb_[g].as<uint32_t>(), nfb[g] * 4,
                             hipMemcpyDeviceToHost, st_));
  DMR_HIP(hipStreamSynchronize(st_));
  for (int g = 0; g < kWaveKinds; ++g) {
    const auto& items = group[g];
    const uint64_t ni = items.size();
    if (!ni) continue;
    std::vector<char> is_fb(ni, 0);
    for (uint32_t f = 0; f < nfb[g]; ++f) {
      const uint32_t t = pin_fbl[fb_base[g] + f];
      is_fb[t] = 1;
      fallback_items.push_back(items[t]);
    }
    for (uint64_t t = 0; t < ni; ++t) {
      if (is_fb[t]) continue;
      const uint64_t T = out_base[g] + t;
      const SwOut& o = pin_out[T];
      (*outs)[items[t]] = o;
      (*ops)[items[t]] = OpsRef{pin_ops + pin_poff[T], o.n_ops};
    }
  }
  // Bands wider than a wave covers: the block kernel (with the score pass's
  // best scores); what it hands on joins the row-exact kernel's list.
  std::vector<uint64_t> block_items;
  for (uint64_t k : group[kWaveKinds]) {
    const SwJob& j = hjobs[k];
    if (run->known && swblock_k_for((int64_t)j.d1 - j.d0 + 1)) block_items.push_back(k);
    else fallback_items.push_back(k);
  }
  if (!block_items.empty()) {
    const uint64_t nb = block_items.size();
    std::vector<uint32_t> jl(nb);
    std::vector<uint64_t> toff(nb + 1, 0), ooff(nb + 1, 0);
    for (uint64_t t = 0; t < nb; ++t) {
      const uint64_t k = block_items[t];
      const SwJob& j = hjobs[k];
      jl[t] = list[k];
      const int64_t n = (int64_t)j.q1 - j.q0, m = slen[k];
      const int64_t rlo = std::max<int64_t>(1, (int64_t)j.d0 + 1), rhi = std::min<int64_t>(n, m + j.d1);
      const uint64_t rows = rhi >= rlo ? (uint64_t)(rhi - rlo + 1) : 0;
      toff[t + 1] = toff[t] + rows * swblock_row_bytes(swblock_k_for((int64_t)j.d1 - j.d0 + 1));
      ooff[t + 1] = ooff[t] + (uint64_t)(n + m + 1);
      cells += rows * (uint64_t)((int64_t)j.d1 - j.d0 + 1);
    }
    uint32_t* dl = bk_list_.ensure<uint32_t>(nb + 1);
    uint64_t* dto = bk_toff_.ensure<uint64_t>(nb + 1);
    uint64_t* doo = bk_ooff_.ensure<uint64_t>(nb + 1);
    DMR_HIP(hipMemcpyAsync(dl, jl.data(), nb * 4, hipMemcpyHostToDevice, st_));
    DMR_HIP(hipMemcpyAsync(dto, toff.data(), (nb + 1) * 8, hipMemcpyHostToDevice, st_));
    DMR_HIP(hipMemcpyAsync(doo, ooff.data(), (nb + 1) * 8, hipMemcpyHostToDevice, st_));
    uint32_t* dcnt = bk_cnt_.ensure<uint32_t>(1);
    launch_fill_u32(dcnt, 1, 0, st_);
    uint32_t* dfb = bk_fb_.ensure<uint32_t>(nb + 1);
    SwOut* dout = bk_out_.ensure<SwOut>(nb + 1);
    uint32_t* dops = bk_ops_.ensure<uint32_t>(ooff[nb] + 1);
    launch_sw_block(true, djobs, dl, (uint32_t)nb, qs, dbs, matrix32_.as<int8_t>(), go, ge, dout, nullptr,
                    bk_trace_.ensure<uint8_t>(toff[nb] + 4), dto, dops, doo, dfb, dcnt, run->known, 0, st_);
    std::vector<SwOut> ho(nb);
    std::vector<uint32_t> hfb(nb + 1), hops(ooff[nb]);
    uint32_t hn = 0;
    DMR_HIP(hipMemcpyAsync(&hn, dcnt, 4, hipMemcpyDeviceToHost, st_));
    DMR_HIP(hipMemcpyAsync(hfb.data(), dfb, nb * 4, hipMemcpyDeviceToHost, st_));
    DMR_HIP(hipMemcpyAsync(ho.data(), dout, nb * sizeof(SwOut), hipMemcpyDeviceToHost, st_));
    if (ooff[nb]) DMR_HIP(hipMemcpyAsync(hops.data(), dops, ooff[nb] * 4, hipMemcpyDeviceToHost, st_));
    DMR_HIP(hipStreamSynchronize(st_));
    std::vector<char> is_fb(nb, 0);
    if (hn > nb) throw Error(-2, "block-kernel fallback count exceeds its tasks");
    for (uint32_t f = 0; f < hn; ++f) {
      if (hfb[f] >= nb) throw Error(-2, "block-kernel fallback task out of range");
      is_fb[hfb[f]] = 1;
      fallback_items.push_back(block_items[hfb[f]]);
    }
    store->parts.push_back(std::move(hops));
    const uint32_t* bops = store->parts.back().data();
    for (uint64_t t = 0; t < nb; ++t) {
      if (is_fb[t]) continue;
      (*outs)[block_items[t]] = ho[t];
      (*ops)[block_items[t]] = OpsRef{bops + ooff[t], ho[t].n_ops};
    }
  }
  if (!fallback_items.empty()) {
    const uint64_t nf = fallback_items.size();
    std::vector<uint32_t> jl(nf);
    for (uint64_t t = 0; t < nf; ++t) jl[t] = list[fallback_items[t]];
    uint32_t* dl = tb_list_[kWaveKinds].ensure<uint32_t>(nf + 1);
    DMR_HIP(hipMemcpyAsync(dl, jl.data(), nf * 4, hipMemcpyHostToDevice, st_));
    uint64_t* dc = cells_.ensure<uint64_t>(nf + 1);
    uint64_t* dw = width0_.ensure<uint64_t>(nf + 1);
    uint32_t* dsr = stride_.ensure<uint32_t>(nf + 1);
    launch_job_geometry(djobs, dl, nf, dbs, dc, dw, dsr, st_);
    std::vector<uint64_t> hc(nf), hw(nf);
    std::vector<uint32_t> hs(nf);
    DMR_HIP(hipMemcpyAsync(hc.data(), dc, nf * 8, hipMemcpyDeviceToHost, st_));
    DMR_HIP(hipMemcpyAsync(hw.data(), dw, nf * 8, hipMemcpyDeviceToHost, st_));
    DMR_HIP(hipMemcpyAsync(hs.data(), dsr, nf * 4, hipMemcpyDeviceToHost, st_));
    DMR_HIP(hipStreamSynchronize(st_));
    std::vector<uint64_t> toff(nf + 1, 0), so(nf + 1, 0), ooff(nf + 1, 0);
    for (uint64_t t = 0; t < nf; ++t) {
      const SwJob& j = hjobs[fallback_items[t]];
      toff[t + 1] = toff[t] + hc[t] + hw[t];
      so[t + 1] = so[t] + 4ull * hs[t];
      ooff[t + 1] = ooff[t] + (uint64_t)(j.q1 - j.q0) + slen[fallback_items[t]] + 1;
      cells += hc[t];
    }
    uint64_t* dto = tb_toff_[kWaveKinds].ensure<uint64_t>(nf + 1);
    uint64_t* dso = scr_off_.ensure<uint64_t>(nf + 1);
    uint64_t* doo = tb_ooff_[kWaveKinds].ensure<uint64_t>(nf + 1);
    DMR_HIP(hipMemcpyAsync(dto, toff.data(), (nf + 1) * 8, hipMemcpyHostToDevice, st_));
    DMR_HIP(hipMemcpyAsync(dso, so.data(), (nf + 1) * 8, hipMemcpyHostToDevice, st_));
    DMR_HIP(hipMemcpyAsync(doo, ooff.data(), (nf + 1) * 8, hipMemcpyHostToDevice, st_));
    uint32_t* dops = tb_ops_[kWaveKinds].ensure<uint32_t>(ooff[nf] + 1);
    SwOut* dout = tb_out_[kWaveKinds].ensure<SwOut>(nf + 1);
    launch_sw_generic(djobs, dl, nf, qs, dbs, matrix32_.as<int8_t>(), go, ge,
                      scratch_.ensure<int32_t>(so[nf] + 1), dso, tb_trace_[kWaveKinds].ensure<uint8_t>(toff[nf] + 1),
                      dto, dops, doo, dout, true, st_);
    std::vector<SwOut> ho(nf);
    std::vector<uint32_t> hops(ooff[nf]);
    DMR_HIP(hipMemcpyAsync(ho.data(), dout, nf * sizeof(SwOut), hipMemcpyDeviceToHost, st_));
    DMR_HIP(hipMemcpyAsync(hops.data(), dops, ooff[nf] * 4, hipMemcpyDeviceToHost, st_));
    DMR_HIP(hipStreamSynchronize(st_));
    store->extra = std::move(hops);
    for (uint64_t t = 0; t < nf; ++t) {
      (*outs)[fallback_items[t]] = ho[t];
      (*ops)[fallback_items[t]] = OpsRef{store->extra.data() + ooff[t], ho[t].n_ops};
    }
  }
  if (cells_out) *cells_out = cells;
}

// Jobs and DP cells per score class of the slice just scored (its jobs in
// the class order of sw_idx_), added to the profile of the search.
void Engine::accumulate_profile(uint64_t njobs) {
  if (!njobs || last_cbeg_.size() != (size_t)kClasses) return;
  SeqSet dbs{db_->res.as<uint8_t>(), db_->offs.as<uint32_t>()};
  uint64_t* c = cells_.ensure<uint64_t>(njobs + 1);
  launch_job_geometry(jobs_.as<SwJob>(), sw_idx_.as<uint32_t>(), njobs, dbs, c, nullptr, nullptr, st_);
  // Per-class sums on the device: only kClasses words come back.
  uint64_t* rng = prof_ranges_.ensure<uint64_t>(3 * kClasses);
  uint64_t* sums = rng + 2 * kClasses;
  uint64_t* h = prof_pin_.ensure<uint64_t>(3 * kClasses);
  uint64_t longest = 0;
  for (int k = 0; k < kClasses; ++k) {
    h[2 * k] = last_cbeg_[k];
    h[2 * k + 1] = std::max(last_cbeg_[k], last_cend_[k]);
    h[2 * kClasses + k] = 0;
    longest = std::max(longest, h[2 * k + 1] - h[2 * k]);
  }
  DMR_HIP(hipMemcpyAsync(rng, h, 3 * kClasses * 8, hipMemcpyHostToDevice, st_));
  launch_range_sums(c, rng, kClasses, longest, sums, st_);
  DMR_HIP(hipMemcpyAsync(h + 2 * kClasses, sums, kClasses * 8, hipMemcpyDeviceToHost, st_));
  DMR_HIP(hipStreamSynchronize(st_));
  for (int k = 0; k < kClasses; ++k) {
    prof_jobs_[k] += h[2 * k + 1] - h[2 * k];
    prof_cells_[k] += h[2 * kClasses + k];
    prof_dp_cells_ += h[2 * kClasses + k];
  }
}

void Engine::last_class_profile(uint64_t jobs[kClasses], uint64_t cells[kClasses],
                                double ms[kClasses]) {
  for (int c = 0; c < kClasses; ++c) {
    jobs[c] = cells[c] = 0;
    ms[c] = last_ms_class_[c];
  }
  if (last_slices_ > 1) {
    if (!prof_valid_)
      throw Error(-5, "the last search ran in several slices: its class profile needs profiling on "
                      "(dmr_ctx_set_tuning \"profile\" 1) before the search");
    for (int c = 0; c < kClasses; ++c) {
      jobs[c] = prof_jobs_[c];
      cells[c] = prof_cells_[c];
    }
    return;
  }
  const uint64_t n = last_njobs_;
  if (!n || last_cbeg_.size() != (size_t)kClasses) return;
  SeqSet dbs{db_->res.as<uint8_t>(), db_->offs.as<uint32_t>()};
  uint64_t* c = cells_.ensure<uint64_t>(n + 1);
  launch_job_geometry(jobs_.as<SwJob>(), sw_idx_.as<uint32_t>(), n, dbs, c, nullptr, nullptr, st_);
  std::vector<uint64_t> hc(n);
  DMR_HIP(hipMemcpyAsync(hc.data(), c, n * 8, hipMemcpyDeviceToHost, st_));
  DMR_HIP(hipStreamSynchronize(st_));
  for (int k = 0; k < kClasses; ++k)
    for (uint64_t p = last_cbeg_[k]; p < last_cend_[k]; ++p) {
      jobs[k]++;
      cells[k] += hc[p];
    }
}

uint64_t Engine::count_last_dp_cells() {
  if (last_slices_ > 1) {
    if (!prof_valid_)
      throw Error(-5, "the last search ran in several slices: its DP cells need profiling on "
                      "(dmr_ctx_set_tuning \"profile\" 1) before the search");
    return prof_dp_cells_;
  }
  uint64_t n = last_njobs_;
  if (!n) return 0;
  SeqSet dbs{db_->res.as<uint8_t>(), db_->offs.as<uint32_t>()};
  uint64_t* cells = cells_.ensure<uint64_t>(n + 1);
  launch_job_geometry(jobs_.as<SwJob>(), nullptr, n, dbs, cells, nullptr, nullptr, st_);
  std::vector<uint64_t> h(n);
  DMR_HIP(hipMemcpyAsync(h.data(), cells, n * 8, hipMemcpyDeviceToHost, st_));
  DMR_HIP(hipStreamSynchronize(st_));
  return std::accumulate(h.begin(), h.end(), (uint64_t)0);
}

// ---------------------------------------------------------------------------
// Drop-in calls on host buffers.
// ---------------------------------------------------------------------------

// seed::build (src/seed.rs:76-104): entries sorted by (key, seq, pos).
void Engine::seed_build(const uint8_t* seqs, const uint64_t* offs, uint32_t nseq,
                        const ShapeDef& shape, const ReductionDef& red, uint32_t base, double cut,
                        std::vector<uint64_t>* keys, std::vector<uint32_t>* seq,
                        std::vector<uint32_t>* pos) {
  DMR_HIP(hipSetDevice(device_));
  keys->clear();
  seq->clear();
  pos->clear();
  uint64_t total = offs[nseq];
  if (nseq == 0 || total == 0) return;
  if (total >= (1ull << 32)) throw Error(-5, "input larger than 4G letters");
  std::vector<uint32_t> o32(nseq + 1), seq_of(total);
  for (uint32_t i = 0; i <= nseq; ++i) o32[i] = (uint32_t)offs[i];
  for (uint32_t i = 0; i < nseq; ++i)
    for (uint64_t k = offs[i]; k < offs[i + 1]; ++k) seq_of[k] = i;
  DevBuf res, dof, dso, prm, k0, l0, k1, l1;
  DMR_HIP(hipMemcpyAsync(res.ensure<uint8_t>(total + 64), seqs, total, hipMemcpyHostToDevice, st_));
  DMR_HIP(hipMemcpyAsync(dof.ensure<uint32_t>(nseq + 1), o32.data(), (nseq + 1) * 4, hipMemcpyHostToDevice, st_));
  DMR_HIP(hipMemcpyAsync(dso.ensure<uint32_t>(total), seq_of.data(), total * 4, hipMemcpyHostToDevice, st_));
  SeedParams P;
  std::memset(&P, 0, sizeof P);
  P.nshapes = 1;
  P.red_size = red.size;
  std::memcpy(P.red_map, red.map, 25);
  P.cut = cut;
  for (int k = 2; k < 33; ++k) P.lnfact[k] = P.lnfact[k - 1] + std::log((double)k);
  P.shapes[0].weight = shape.weight;
  P.shapes[0].span = shape.span;
  std::memcpy(P.shapes[0].pos, shape.pos, sizeof shape.pos);
  DMR_HIP(hipMemcpyAsync(prm.ensure<SeedParams>(1), &P, sizeof P, hipMemcpyHostToDevice, st_));
  long double mx = std::pow((long double)red.size, (long double)shape.weight) - 1;
  if (mx >= 18446744073709551615.0L) throw Error(-5, "seed key does not fit 64 bits");
  int kb = bits_for((uint64_t)mx);
  uint32_t nb = seed_enum_blocks(0, total);
  uint64_t nbk = nb + 1ull;
  uint32_t* bc = b_counts_.ensure<uint32_t>(nbk);
  launch_fill_u32(bc + nb, 1, 0, st_);
  uint64_t* bo = b_off_.ensure<uint64_t>(nbk);
  uint64_t* kk = k0.ensure<uint64_t>(total + 1);
  uint32_t* ll = l0.ensure<uint32_t>(total + 1);
  launch_seed_enumerate(res.as<uint8_t>(), dso.as<uint32_t>(), dof.as<uint32_t>(), 0, total,
                        prm.as<SeedParams>(), 1, kb, bc, bo, b_scan_.ensure<uint64_t>(scan_tmp_elems(nbk)),
                        kk, ll, st_);
  uint64_t n = read1(bo + nb);
  radix_sort_pairs(kk, ll, k1.ensure<uint64_t>(n + 1), l1.ensure<uint32_t>(n + 1), n, kb,
                   b_hist_.ensure<uint32_t>(radix_hist_elems(n)),
                   b_scan_.ensure<uint64_t>(radix_scan_tmp_elems(n)), st_);
  keys->resize(n);
  std::vector<uint32_t> locs(n);
  DMR_HIP(hipMemcpyAsync(keys->data(), kk, n * 8, hipMemcpyDeviceToHost, st_));
  DMR_HIP(hipMemcpyAsync(locs.data(), ll, n * 4, hipMemcpyDeviceToHost, st_));
  DMR_HIP(hipStreamSynchronize(st_));
  seq->resize(n);
  pos->resize(n);
  for (uint64_t i = 0; i < n; ++i) {
    uint32_t s = seq_of[locs[i]];
    (*seq)[i] = base + s;
    (*pos)[i] = locs[i] - o32[s];
  }
}

// seed::join_into (src/seed.rs:110-143) on key-sorted arrays; hits are
// returned sorted by (query, subject, diagonal, spos).
void Engine::seed_join(const uint64_t* qk, const uint32_t* qs, const uint32_t* qp, uint64_t nq,
                       const uint64_t* rk, const uint32_t* rs, const uint32_t* rp, uint64_t nr,
                       uint64_t max_expand, std::vector<uint32_t>* out4) {
  DMR_HIP(hipSetDevice(device_));
  out4->clear();
  if (!nq || !nr) return;
  for (uint64_t i = 1; i < nq; ++i)
    if (qk[i] < qk[i - 1]) throw Error(-1, "query seeds are not sorted by key");
  for (uint64_t i = 1; i < nr; ++i)
    if (rk[i] < rk[i - 1]) throw Error(-1, "reference seeds are not sorted by key");
  uint64_t maxkey = std::max(qk[nq - 1], rk[nr - 1]);
  uint32_t mq = 0, mqp = 0, ms = 0, msp = 0;
  for (uint64_t i = 0; i < nq; ++i) { mq = std::max(mq, qs[i]); mqp = std::max(mqp, qp[i]); }
  for (uint64_t i = 0; i < nr; ++i) { ms = std::max(ms, rs[i]); msp = std::max(msp, rp[i]); }
  int kb = bits_for(maxkey);
  int tb = std::min(kb, 22);
  DevBuf dqk, dqs, dqp, drk, drs, drp, tab, dfl, dps, druns, dcnt, drlo, doff, dh, dh2, desc;
  DMR_HIP(hipMemcpyAsync(dqk.ensure<uint64_t>(nq), qk, nq * 8, hipMemcpyHostToDevice, st_));
  DMR_HIP(hipMemcpyAsync(dqs.ensure<uint32_t>(nq), qs, nq * 4, hipMemcpyHostToDevice, st_));
  DMR_HIP(hipMemcpyAsync(dqp.ensure<uint32_t>(nq), qp, nq * 4, hipMemcpyHostToDevice, st_));
  DMR_HIP(hipMemcpyAsync(drk.ensure<uint64_t>(nr), rk, nr * 8, hipMemcpyHostToDevice, st_));
  DMR_HIP(hipMemcpyAsync(drs.ensure<uint32_t>(nr), rs, nr * 4, hipMemcpyHostToDevice, st_));
  DMR_HIP(hipMemcpyAsync(drp.ensure<uint32_t>(nr), rp, nr * 4, hipMemcpyHostToDevice, st_));
  RefChunkDev ch;
  ch.keys = drk.as<uint64_t>();
  ch.locs = nullptr;
  ch.n = nr;
  ch.shift = kb - tb;
  ch.nbuckets = (maxkey >> ch.shift) + 1;
  ch.table = tab.ensure<uint64_t>(ch.nbuckets + 1);
  ch.pad = 0;
  launch_bucket_table(ch.keys, nr, ch.shift, ch.nbuckets, tab.as<uint64_t>(), st_);
  DMR_HIP(hipMemcpyAsync(desc.ensure<RefChunkDev>(1), &ch, sizeof ch, hipMemcpyHostToDevice, st_));
  uint32_t* fl = dfl.ensure<uint32_t>(nq + 1);
  launch_run_flags(dqk.as<uint64_t>(), nq, fl, st_);
  launch_fill_u32(fl + nq, 1, 0, st_);
  uint64_t* ps = dps.ensure<uint64_t>(nq + 1);
  scan_exclusive_u32_to_u64(fl, ps, nq + 1, b_scan_.ensure<uint64_t>(scan_tmp_elems(nq + 1)), st_);
  uint64_t nruns = read1(ps + nq);
  uint32_t* runs = druns.ensure<uint32_t>(nruns + 1);
  launch_compact_flags(fl, ps, nq, runs, st_);
  uint32_t* cnt = dcnt.ensure<uint32_t>(nruns + 1);
  uint64_t* rlo = drlo.ensure<uint64_t>(nruns + 1);
  launch_join_count(dqk.as<uint64_t>(), runs, nruns, nq, desc.as<RefChunkDev>(), 1, max_expand, cnt,
                    rlo, st_);
  launch_fill_u32(cnt + nruns, 1, 0, st_);
  uint64_t* hoff = doff.ensure<uint64_t>(nruns + 1);
  scan_exclusive_u32_to_u64(cnt, hoff, nruns + 1, b_scan_.ensure<uint64_t>(scan_tmp_elems(nruns + 1)), st_);
  uint64_t nh = read1(hoff + nruns);
  HitGeom g;
  g.bt = bits_for(mq);
  g.bs = bits_for(ms);
  g.bp = bits_for(msp);
  g.bd = bits_for((uint64_t)mqp + msp + 1);
  g.diag_bias = (int64_t)msp;
  int hb = (int)(g.bt + g.bs + g.bd + g.bp);
  if (hb > 64) throw Error(-5, "hit key does not fit 64 bits");
  uint64_t* hits = dh.ensure<uint64_t>(nh + 1);
  launch_join_emit_entries(dqs.as<uint32_t>(), dqp.as<uint32_t>(), runs, nruns, nq, drs.as<uint32_t>(),
                           drp.as<uint32_t>(), cnt, rlo, hoff, g, hits, st_);
  radix_sort_keys(hits, dh2.ensure<uint64_t>(nh + 1), nh, hb, b_hist_.ensure<uint32_t>(radix_hist_elems(nh)),
                  b_scan_.ensure<uint64_t>(radix_scan_tmp_elems(nh)), st_);
  std::vector<uint64_t> h(nh);
  DMR_HIP(hipMemcpyAsync(h.data(), hits, nh * 8, hipMemcpyDeviceToHost, st_));
  DMR_HIP(hipStreamSynchronize(st_));
  out4->resize(nh * 4);
  const uint64_t pm = (1ull << g.bp) - 1, dm = (1ull << g.bd) - 1, sm = (1ull << g.bs) - 1;
  for (uint64_t i = 0; i < nh; ++i) {
    uint64_t k = h[i];
    uint32_t spos = (uint32_t)(k & pm);
    int64_t diag = (int64_t)((k >> g.bp) & dm) - g.diag_bias;
    uint32_t subj = (uint32_t)((k >> (g.bp + g.bd)) & sm);
    uint32_t qry = (uint32_t)(k >> (g.bp + g.bd + g.bs));
    (*out4)[4 * i] = qry;
    (*out4)[4 * i + 1] = (uint32_t)(diag + spos);
    (*out4)[4 * i + 2] = subj;
    (*out4)[4 * i + 3] = spos;
  }
}

// ungapped::xdrop (src/ungapped.rs:38-88) for many seeds of one sequence pair.
void Engine::xdrop_batch(const uint8_t* q, uint64_t qlen, const uint8_t* s, uint64_t slen,
                         const uint32_t* qa, const uint32_t* sa, uint64_t n, int xdrop, SegRec* out) {
  DMR_HIP(hipSetDevice(device_));
  if (!n) return;
  for (uint64_t i = 0; i < n; ++i)
    if (qa[i] > qlen || sa[i] > slen) throw Error(-1, "seed position outside the sequence");
  DevBuf dq, ds, dqa, dsa, dout;
  DMR_HIP(hipMemcpyAsync(dq.ensure<uint8_t>(qlen + 1), q, qlen, hipMemcpyHostToDevice, st_));
  DMR_HIP(hipMemcpyAsync(ds.ensure<uint8_t>(slen + 1), s, slen, hipMemcpyHostToDevice, st_));
  DMR_HIP(hipMemcpyAsync(dqa.ensure<uint32_t>(n), qa, n * 4, hipMemcpyHostToDevice, st_));
  DMR_HIP(hipMemcpyAsync(dsa.ensure<uint32_t>(n), sa, n * 4, hipMemcpyHostToDevice, st_));
  launch_xdrop_pairs(dq.as<uint8_t>(), qlen, ds.as<uint8_t>(), slen, dqa.as<uint32_t>(),
                     dsa.as<uint32_t>(), n, matrix32_.as<int8_t>(), xdrop, dout.ensure<SegRec>(n), st_);
  DMR_HIP(hipMemcpyAsync(out, dout.as<SegRec>(), n * sizeof(SegRec), hipMemcpyDeviceToHost, st_));
  DMR_HIP(hipStreamSynchronize(st_));
}

// gapped::extend_span (src/gapped.rs:67-280) for a batch of sequence pairs.
void Engine::extend_span_batch(const uint8_t* qbuf, uint64_t qbuf_len, const uint8_t* sbuf,
                               uint64_t sbuf_len, const SpanJob* jobs, uint64_t njobs, int gap_open,
                               int gap_extend, bool traceback, std::vector<SwOut>* outs,
                               std::vector<std::vector<uint32_t>>* ops) {
  DMR_HIP(hipSetDevice(device_));
  outs->assign(njobs, SwOut{});
  if (ops) ops->assign(njobs, {});
  if (!njobs) return;
  if (qbuf_len >= (1ull << 32) || sbuf_len >= (1ull << 32)) throw Error(-5, "buffers larger than 4G");
  // Each job becomes (target j, subject j) of a private offsets table.
  std::vector<uint32_t> qo(2 * njobs), so(2 * njobs);
  std::vector<SwJob> dj(njobs);
  for (uint64_t i = 0; i < njobs; ++i) {
    const SpanJob& J = jobs[i];
    if (J.q_off + J.qlen > qbuf_len || J.s_off + J.slen > sbuf_len)
      throw Error(-1, "job outside its buffer");
    int64_t d0 = std::min(J.d0, J.d1), d1 = std::max(J.d0, J.d1);
    if (d0 < INT32_MIN / 2 || d1 > INT32_MAX / 2) throw Error(-1, "diagonal out of range");
    qo[2 * i] = (uint32_t)J.q_off;
    qo[2 * i + 1] = (uint32_t)(J.q_off + J.qlen);
    so[2 * i] = (uint32_t)J.s_off;
    so[2 * i + 1] = (uint32_t)(J.s_off + J.slen);
    dj[i] = SwJob{(uint32_t)(2 * i), (uint32_t)(2 * i), 0, (uint32_t)J.qlen, (int32_t)d0, (int32_t)d1,
                  (uint32_t)i, 0};
  }
  DevBuf dq, ds, dqo, dso, djobs, dsc;
  DevBuf dqp, dsp, dqpo, dspo;
  DMR_HIP(hipMemcpyAsync(dq.ensure<uint8_t>(qbuf_len + 64), qbuf, qbuf_len, hipMemcpyHostToDevice, st_));
  DMR_HIP(hipMemcpyAsync(ds.ensure<uint8_t>(sbuf_len + 64), sbuf, sbuf_len, hipMemcpyHostToDevice, st_));
  DMR_HIP(hipMemcpyAsync(dqo.ensure<uint32_t>(2 * njobs), qo.data(), 8 * njobs, hipMemcpyHostToDevice, st_));
  DMR_HIP(hipMemcpyAsync(dso.ensure<uint32_t>(2 * njobs), so.data(), 8 * njobs, hipMemcpyHostToDevice, st_));
  DMR_HIP(hipMemcpyAsync(djobs.ensure<SwJob>(njobs), dj.data(), njobs * sizeof(SwJob), hipMemcpyHostToDevice, st_));
  SeqSet qs{dq.as<uint8_t>(), dqo.as<uint32_t>()}, ss{ds.as<uint8_t>(), dso.as<uint32_t>()};
  if (!traceback) {
    // Padded letters (times 4) of each job's query and subject for the packed
    // score kernels: entry 2i of the bases points at job i's copy, entry
    // 2i + 1 at the empty sequence at base 1.
    std::vector<uint8_t> qpad(2, 4 * 25), spad(2, 4 * 25);
    std::vector<uint32_t> qpo(2 * njobs, 1u), spo(2 * njobs, 1u);
    auto append = [](std::vector<uint8_t>* pad, const uint8_t* src, uint64_t len) {
      const uint64_t base = pad->size();
      pad->resize(base + len + 1);
      for (uint64_t k = 0; k < len; ++k) (*pad)[base + k] = (uint8_t)(4 * src[k]);
      (*pad)[base + len] = 4 * 25;
      return base;
    };
    for (uint64_t i = 0; i < njobs; ++i) {
      qpo[2 * i] = (uint32_t)append(&qpad, qbuf + jobs[i].q_off, jobs[i].qlen);
      spo[2 * i] = (uint32_t)append(&spad, sbuf + jobs[i].s_off, jobs[i].slen);
    }
    if (qpad.size() >= (1ull << 32) || spad.size() >= (1ull << 32)) throw Error(-5, "jobs larger than 4G letters");
    DMR_HIP(hipMemcpyAsync(dqp.ensure<uint8_t>(qpad.size()), qpad.data(), qpad.size(), hipMemcpyHostToDevice, st_));
    DMR_HIP(hipMemcpyAsync(dsp.ensure<uint8_t>(spad.size()), spad.data(), spad.size(), hipMemcpyHostToDevice, st_));
    DMR_HIP(hipMemcpyAsync(dqpo.ensure<uint32_t>(2 * njobs), qpo.data(), 8 * njobs, hipMemcpyHostToDevice, st_));
    DMR_HIP(hipMemcpyAsync(dspo.ensure<uint32_t>(2 * njobs), spo.data(), 8 * njobs, hipMemcpyHostToDevice, st_));
    qs.pad = dqp.as<uint8_t>();
    qs.pad_offs = dqpo.as<uint32_t>();
    ss.pad = dsp.as<uint8_t>();
    ss.pad_offs = dspo.as<uint32_t>();
  }
  if (traceback) {
    std::vector<uint32_t> list(njobs), slen(njobs);
    for (uint64_t i = 0; i < njobs; ++i) {
      list[i] = (uint32_t)i;
      slen[i] = (uint32_t)jobs[i].slen;
    }
    std::vector<SwOut> ho;
    std::vector<std::vector<uint32_t>> hops;
    traceback_pass(djobs.as<SwJob>(), list, dj, slen, qs, ss, gap_open, gap_extend, &ho, &hops, nullptr);
    *outs = ho;
    if (ops) *ops = hops;
    return;
  } else {
    // Score only: the same score pass as the batched search (packed int16
    // kernel for the standard band widths, exact int32 for the rest).
    SearchOptions so;
    so.gap_open = gap_open;
    so.gap_extend = gap_extend;
    StageStats S;
    int tb = 1;
    while ((1ull << tb) <= 2 * njobs) ++tb;
    int32_t* sc = dsc.ensure<int32_t>(njobs);
    score_pass(djobs.as<SwJob>(), njobs, tb, qs, ss, so, sc, &S);
    last_dropin_stats_ = S;
    // The class profile of a drop-in batch: kernel times only.
    for (int c = 0; c < kClasses; ++c) last_ms_class_[c] = S.ms_class[c];
    last_njobs_ = 0;
    last_slices_ = 1;
    std::vector<int32_t> hsc(njobs);
    DMR_HIP(hipMemcpyAsync(hsc.data(), sc, njobs * 4, hipMemcpyDeviceToHost, st_));
    DMR_HIP(hipStreamSynchronize(st_));
    for (uint64_t i = 0; i < njobs; ++i) (*outs)[i].score = hsc[i];
  }
}

}  // namespace dmr
