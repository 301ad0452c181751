// Host orchestration of the MI355X search core.
//
// Restates src/search.rs:95-396 around device kernels: presets, resident seed
// index per DB chunk, query seeding, join, ungapped extension, band merge,
// banded DP, selection by E-value, traceback, culling and frame merging.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "dmr_alp.hpp"
#include "dmr_device.hpp"
#include "dmr_kernels.hpp"
#include "dmr_tables.hpp"
#include "tuning.hpp"

namespace dmr {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define DMR_HIP(x)                                                                          \
  do {                                                                                      \
    hipError_t e_ = (x);                                                                    \
    if (e_ != hipSuccess)                                                                   \
      throw ::dmr::Error(e_ == hipErrorOutOfMemory ? -3 : -2,                               \
                         std::string(#x) + ": " + hipGetErrorString(e_));                   \
  } while (0)

// Growable device buffer (reallocated only when a larger size is requested).
// Set while a stream capture records the chain into a graph: a buffer that
// would have to grow then throws CaptureGrow (allocation is no stream work),
// and the chain runs again without the graph.
extern thread_local bool g_capture_active;
struct CaptureGrow {};

struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), cap(o.cap) { o.p = nullptr; o.cap = 0; }
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
  template <class T>
  T* as() const { return static_cast<T*>(p); }
  template <class T>
  T* ensure(size_t n) {
    size_t bytes = n * sizeof(T) + 256;
    if (bytes > cap) {
      if (g_capture_active) throw CaptureGrow{};
      release();
      size_t c = bytes + bytes / 4;
      DMR_HIP(hipMalloc(&p, c));
      cap = c;
    }
    return static_cast<T*>(p);
  }
};

// Pinned host memory (hipHostMalloc), grown like DevBuf: device-to-host and
// host-to-device copies through it run as DMA without staging.
struct PinBuf {
  void* p = nullptr;
  size_t cap = 0;
  PinBuf() = default;
  PinBuf(const PinBuf&) = delete;
  PinBuf& operator=(const PinBuf&) = delete;
  ~PinBuf() { release(); }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
  }
  template <class T>
  T* ensure(size_t n) {
    size_t bytes = n * sizeof(T) + 256;
    if (bytes > cap) {
      if (g_capture_active) throw CaptureGrow{};
      release();
      size_t c = bytes + bytes / 4;
      DMR_HIP(hipHostMalloc(&p, c, hipHostMallocDefault));
      cap = c;
    }
    return static_cast<T*>(p);
  }
  template <class T>
  T* as() const { return static_cast<T*>(p); }
};

struct SearchOptions {
  std::string sensitivity = "default";
  double evalue = 0.001;
  uint32_t max_target_seqs = 25, max_hsps = 1;
  int index_chunks = -1;  // <0: preset
  double seed_cut = -1;   // <0: preset
  int gap_open = 11, gap_extend = 1;
  uint64_t max_expand = 4096;
  // The cap counts a key's query entries per record instead of over the
  // batch: each record's hits are those of searching it alone.
  bool per_record = false;
  // Hits per slice of targets (0 = the engine's setting).
  uint64_t slice_hits = 0;
};

// Run-length alignment ops of one HSP (op | count << 2; 0 = M, 1 = I, 2 = D),
// inside an OpsStore that the holder of the HSP list keeps alive.
struct OpsRef {
  const uint32_t* p = nullptr;
  uint32_t n = 0;
  const uint32_t* begin() const { return p; }
  const uint32_t* end() const { return p + n; }
  size_t size() const { return n; }
};

// Pinned host blocks, reused: a search's traceback ops are copied from the
// device straight into one block, which its result keeps (OpsStore) and gives
// back when it goes, so results cost no host copy, allocation or page fault
// per HSP. Thread-safe (results may be destroyed on any thread).
class PinPool {
 public:
  ~PinPool();
  // A block of at least `bytes`; *cap receives its size.
  void* take(size_t bytes, size_t* cap);
  void give(void* p, size_t cap);

 private:
  std::mutex m_;
  std::vector<std::pair<void*, size_t>> free_;
};

// The ops of a list of HSPs: one pinned block from a PinPool (the packed
// traceback ops) and a host vector for ops produced on the host side.
class OpsStore {
 public:
  OpsStore() = default;
  OpsStore(std::shared_ptr<PinPool> pool, size_t words);
  OpsStore(const OpsStore&) = delete;
  OpsStore& operator=(const OpsStore&) = delete;
  ~OpsStore();
  uint32_t* block() { return static_cast<uint32_t*>(blk_); }
  std::vector<uint32_t> extra;  // not resized once HSPs point into it
  // Per-HSP vectors (host-side producers); moving the outer vector keeps
  // every inner buffer in place.
  std::vector<std::vector<uint32_t>> parts;

 private:
  std::shared_ptr<PinPool> pool_;
  void* blk_ = nullptr;
  size_t cap_ = 0;
};

struct HspRec {
  uint32_t query, subject, frame;
  uint32_t q_begin, q_end, s_begin, s_end;  // 0-based half-open, frame coordinates
  int32_t score;
  double bits, evalue;
  uint32_t identities, positives, mismatches, gap_openings, gaps, length;
  OpsRef ops;
};

struct StageStats {
  uint64_t query_seeds = 0, runs = 0, hits = 0, hit_runs = 0, segments = 0, groups = 0;
  uint64_t jobs = 0, dp_cells = 0, maybe = 0, selected = 0, traceback_cells = 0;
  uint64_t sw16_lanes = 0, sw_generic = 0, sw_redo = 0;
  // Score-kernel time per class: 0-4 packed 16-bit (widths 25/33/61/81/129),
  // 5-12 packed 16-bit merged bands (m97 .. m513), 13-18 wave-per-DP int32
  // (K = 1..32), 19 row-exact int32.
  double ms_class[kPackedClasses + kWaveKinds + 1] = {};
  double ms_seed = 0, ms_join = 0, ms_sort_hits = 0, ms_ungapped = 0, ms_bands = 0, ms_sw = 0,
         ms_select = 0, ms_traceback = 0, ms_total = 0, ms_host = 0;
  double ms_translate = 0;  // blastx: six frames, masking and targets on the device
};

// A batch of query targets (one per protein query, or <= 6 frames per read).
struct TargetSet {
  std::vector<uint8_t> letters;
  std::vector<uint32_t> offs;      // n+1
  std::vector<uint32_t> query_of;  // record index of each target
  std::vector<uint8_t> frame_of;
};

// One DB chunk of a resident seed index: tagged keys sorted, the global
// letter index of each seed, and a bucket table of lower bounds over the
// top key bits.
struct IndexChunk {
  DevBuf keys, locs, table;
  uint64_t n = 0, nbuckets = 0;
  int shift = 0;
};

// The seed index of one preset (shapes, reduction, cut) over the chunks of
// the database (src/search.rs:121-133), resident until the database goes.
struct IndexState {
  std::vector<IndexChunk> chunks;
  DevBuf chunk_desc, seed_params;
  int key_bits = 0, shape_bits = 0;
  uint32_t nshapes = 0;
  double build_ms = 0;
};

// The resident database of a device and its seed indexes, shared by the
// engines that search it (each with its own streams and scratch): letters,
// offsets, owner of each letter, padded letters, host copies, and one index
// per preset built so far (built once, under the lock).
struct DbState {
  DevBuf res, offs, seq_of, pad, pad_offs;
  uint32_t n = 0;
  uint64_t letters = 0;       // letters loaded
  uint64_t stat_letters = 0;  // letters E-values use (Database::total_letters)
  uint32_t max_len = 0;
  std::vector<uint32_t> h_offs;
  std::shared_ptr<const std::vector<uint32_t>> h_offs_sp;
  std::shared_ptr<const std::vector<uint8_t>> h_res_sp;
  std::mutex m;
  std::vector<std::pair<std::string, std::shared_ptr<IndexState>>> indexes;
};

class Engine {
 public:
  explicit Engine(int device);
  ~Engine();

  int device() const { return device_; }
  hipStream_t stream() const { return st_; }

  // Searches the resident database (and indexes) of another engine of the
  // same device: engines of one device share it in HBM.
  void share_db(const Engine& other);
  // Resident database (codes 0..24, offsets n+1). stat_letters: the letter
  // count E-values use (Database::total_letters, src/db.rs:236-238, read
  // from a file's header); NULL = the letters loaded.
  void load_db(const uint8_t* letters, const uint64_t* offsets, uint32_t n,
               const uint64_t* stat_letters = nullptr);
  uint32_t db_size() const { return db_->n; }
  uint64_t db_letters() const { return db_->stat_letters; }
  const std::vector<uint32_t>& db_offsets_host() const { return db_->h_offs; }
  const std::vector<uint8_t>& db_letters_host() const { return *db_->h_res_sp; }
  std::shared_ptr<const std::vector<uint32_t>> db_offsets_shared() const { return db_->h_offs_sp; }
  std::shared_ptr<const std::vector<uint8_t>> db_letters_shared() const { return db_->h_res_sp; }

  // Resident seed index for a preset (rebuilt only when the preset changes).
  void build_index(const std::string& sensitivity, int index_chunks, double seed_cut);
  double last_index_ms() const { return idx_ ? idx_->build_ms : 0.0; }
  uint64_t index_entries() const;

  // Upload a target batch (kept resident until the next call).
  void load_targets(const TargetSet& t);
  // blastx targets made on the device (src/workflow.rs:112-127): the six
  // frames of every read (ASCII bases reads[roffs[r]..roffs[r+1]]), masked
  // with min_orf[r], frames without a letter below X dropped; kept resident
  // as load_targets keeps them, and returned in *host (the result's copy).
  void load_reads(const char* reads, const uint64_t* roffs, uint32_t n, const uint32_t* min_orf,
                  TargetSet* host);
  // translate::six_frames + mask_short_orfs of every read: all 6n frames in
  // slot order (letters and 6n + 1 offsets).
  void six_frames(const char* reads, const uint64_t* roffs, uint32_t n, const uint32_t* min_orf,
                  std::vector<uint8_t>* letters, std::vector<uint64_t>* offsets);
  double last_translate_ms() const { return translate_ms_; }
  // Search the resident targets; HSPs per target, ordered as align_target
  // returns them (src/search.rs:262-270).
  // The HSPs' ops live in *ops_store, which the caller keeps with them.
  std::vector<std::vector<HspRec>> search_targets(const SearchOptions& opt, StageStats* stats,
                                                  std::vector<std::shared_ptr<const OpsStore>>* ops_stores);
  // Slices of targets the last search ran in (hits bounded per slice).
  uint32_t last_slices() const { return last_slices_; }
  uint64_t slice_hits() const { return slice_hits_; }
  void set_slice_hits(uint64_t v) { slice_hits_ = v ? v : 1; }
  // Profiling: per-class jobs and DP cells of every slice of a search
  // (dmr_last_class_profile / dmr_last_dp_cells of sliced searches).
  void set_profile(bool on) { profile_ = on; }
  // Score-pass kernel settings (tuning.hpp): the tuning file's settings per
  // preset ("" clears them), keys pinned over them (validated against the
  // compiled kernels; a bad key leaves the settings as they were), and the
  // layout of the last search (before the first: the default preset's).
  void load_tuning(const std::string& path);
  void set_layout_key(const std::string& key, double value);
  double layout_key(const std::string& key) const { return layout_value(layout_, key); }
  const ScoreLayout& layout() const { return layout_; }
  size_t tuned_presets() const { return tuned_.size(); }
  uint64_t graph_launches() const { return graph_launches_; }
  uint64_t graph_builds() const { return graph_builds_; }
  const std::string& tuning_file() const { return tuning_file_; }
  // The tuning file and pinned keys of another engine (serving workers).
  void copy_tuning(const Engine& o);
  // Cells of every DP of the last search (one extra pass, for GCUPS).
  uint64_t count_last_dp_cells();
  static constexpr int kClasses = kPackedClasses + kWaveKinds + 1;  // k_classify in k_sw16.hip
  static constexpr int kRowExact = kClasses - 1;
  // Per score-kernel class of the last search (see StageStats::ms_class):
  // jobs, DP cells and kernel milliseconds.
  void last_class_profile(uint64_t jobs[kClasses], uint64_t cells[kClasses], double ms[kClasses]);
  // Traceback DPs for djobs[list[k]] (host copies hjobs[k], subject lengths slen[k]).
  void traceback_pass(const SwJob* djobs, const std::vector<uint32_t>& list,
                      const std::vector<SwJob>& hjobs, const std::vector<uint32_t>& slen, SeqSet qs,
                      SeqSet dbs, int go, int ge, std::vector<SwOut>* outs,
                      std::vector<std::vector<uint32_t>>* ops, uint64_t* cells_out,
                      const int32_t* known = nullptr);
  // A traceback between its launch and the collection of its results.
  struct TbRun {
    const SwJob* djobs = nullptr;
    const std::vector<uint32_t>* list = nullptr;
    const std::vector<SwJob>* hjobs = nullptr;
    const std::vector<uint32_t>* slen = nullptr;
    SeqSet qs{}, dbs{};
    int go = 0, ge = 0;
    const int32_t* known = nullptr;  // score-pass scores by job (device)
    uint64_t ns = 0, cells = 0;
    std::vector<std::vector<uint64_t>> group;
    // Per wave-kernel group: its block in the pinned input and its first
    // task in the all-group order (outputs, ops offsets and packed ops are
    // laid out over all groups in that order); nout tasks in all, ooff_at:
    // the all-group ops offsets (nout + 1) in the pinned input.
    uint64_t in_off[kWaveKinds] = {}, out_base[kWaveKinds] = {};
    uint64_t nout = 0, ooff_at = 0;
  };
  void traceback_launch(TbRun* run, const SwJob* djobs, const std::vector<uint32_t>& list,
                        const std::vector<SwJob>& hjobs, const std::vector<uint32_t>& slen, SeqSet qs,
                        SeqSet dbs, int go, int ge, const int32_t* known);
  void traceback_collect(TbRun* run, std::vector<SwOut>* outs, std::vector<OpsRef>* ops,
                         std::shared_ptr<OpsStore>* store,
                         uint64_t* cells_out);

  // --- drop-in calls on host buffers -------------------------------------
  void seed_build(const uint8_t* seqs, const uint64_t* offs, uint32_t nseq, const ShapeDef& shape,
                  const ReductionDef& red, uint32_t base, double cut, std::vector<uint64_t>* keys,
                  std::vector<uint32_t>* seq, std::vector<uint32_t>* pos);
  void seed_join(const uint64_t* qk, const uint32_t* qs, const uint32_t* qp, uint64_t nq,
                 const uint64_t* rk, const uint32_t* rs, const uint32_t* rp, uint64_t nr,
                 uint64_t max_expand, std::vector<uint32_t>* out4);
  void xdrop_batch(const uint8_t* q, uint64_t qlen, const uint8_t* s, uint64_t slen,
                   const uint32_t* qa, const uint32_t* sa, uint64_t n, int xdrop, SegRec* out);
  struct SpanJob {
    uint64_t q_off, qlen, s_off, slen;
    int64_t d0, d1;
  };
  void extend_span_batch(const uint8_t* qbuf, uint64_t qbuf_len, const uint8_t* sbuf,
                         uint64_t sbuf_len, const SpanJob* jobs, uint64_t njobs, int gap_open,
                         int gap_extend, bool traceback, std::vector<SwOut>* outs,
                         std::vector<std::vector<uint32_t>>* ops);

  AlpParams alp(int gap_open, int gap_extend) const;

  // gapped::extend_span with traceback for one sequence pair, allocating
  // nothing: its device memory comes from a static per-device array of
  // kNoAllocScratchBytes (noalloc.hip). *ops_fit = false when the ops did not
  // fit ops_cap (out->n_ops holds the count needed; no ops are copied).
  static constexpr uint64_t kNoAllocScratchBytes = 32ull << 20;
  void extend_span_noalloc(const uint8_t* q, uint64_t qlen, const uint8_t* s, uint64_t slen, int64_t d0,
                           int64_t d1, int gap_open, int gap_extend, SwOut* out, uint32_t* ops,
                           uint32_t ops_cap, bool* ops_fit);

 private:
  // Device address of the no-allocation array on the current device.
  static void* noalloc_scratch_address();
  void* noalloc_ = nullptr;
  void upload_seed_params(const std::string& sens, double cut, DevBuf* params, uint32_t* nshapes, int* key_bits,
                          int* shape_bits);
  bool swk_launched(int c, const uint64_t* cbeg, const uint64_t* cend, bool fast) const;
  // Score pass over the jobs; njobs_dev: njobs is a bound and the count is
  // read on the device (the pass's one host read brings it back). Returns the
  // job count.
  uint64_t score_pass(const SwJob* jobs, uint64_t njobs, int seq_bits, SeqSet qs, SeqSet dbs,
                      const SearchOptions& opt, int32_t* scores, StageStats* S, const uint64_t* njobs_dev = nullptr);
  void search_slice(const SearchOptions& opt, const AlpParams& P, const HitGeom& g, const uint64_t* hits,
                    uint64_t nhits, SeqSet qs, SeqSet dbs, StageStats* S,
                    std::vector<std::vector<HspRec>>* result,
                    std::vector<std::shared_ptr<const OpsStore>>* ops_stores);
  void accumulate_profile(uint64_t njobs);
  uint32_t slices_ = 0, last_slices_ = 0;
  uint64_t slice_hits_ = 300ull << 20;
  bool profile_ = false, prof_valid_ = false;
  uint64_t prof_jobs_[kPackedClasses + kWaveKinds + 1] = {}, prof_cells_[kPackedClasses + kWaveKinds + 1] = {};
  uint64_t prof_dp_cells_ = 0;
  DevBuf thits_, jc_scnt_, jc_sst_, jc_snq_, q_group_, qruns_, prof_ranges_;
  PinBuf prof_pin_;
  PinBuf thits_pin_;
  void stage_event(int k);
  float stage_ms(int a, int b);
  template <class T>
  T read1(const T* dptr);

  int device_ = 0;
  hipStream_t st_ = nullptr;
  // 0-11 stage boundaries, 12 + 2c / 13 + 2c around score class c.
  hipEvent_t ev_[12 + 2 * kClasses] = {};
  // Auxiliary streams (default priority) for kernels that may run side by
  // side: fork() makes them wait for the work queued on st_, join() makes st_
  // wait for everything queued on them.
  // Three, so that with the engine stream each has a hardware queue of its
  // own (HIP gives a process four by default).
  static constexpr int kAux = 3;
  hipStream_t aux_[kAux] = {};
  hipEvent_t fork_ev_ = nullptr, join_ev_[kAux] = {};
  void fork();
  void join();

  // database and the index of the current preset (shared)
  std::shared_ptr<DbState> db_ = std::make_shared<DbState>();
  std::shared_ptr<IndexState> idx_;

  // matrix (32 x 32 int8, padded letters score -64)
  DevBuf matrix32_;

  // targets
  TargetSet tset_;
  DevBuf q_res_, q_offs_, q_seq_of_;
  DevBuf q_pad_, q_pad_offs_;
  uint32_t q_n_ = 0, q_max_len_ = 0;
  uint64_t q_total_ = 0;

  // blastx reads: bases, offsets, min_orf, frame offsets, translated and
  // masked frames, per-frame counts and the compaction's scans.
  DevBuf r_nt_, r_off_, r_minorf_, r_foff_, r_frames_, r_masked_, r_good_, r_keep_, r_klen_, r_kept_,
      r_tlet_, r_qof_, r_fof_;
  uint64_t frames_dev(const char* reads, const uint64_t* roffs, uint32_t n, const uint32_t* min_orf,
                      std::vector<uint64_t>* foff);
  double translate_ms_ = 0;

  // per-search scratch
  DevBuf b_counts_, b_off_, b_scan_, b_hist_;
  DevBuf qkeys_, qlocs_, qkeys2_, qlocs2_, flags_, pos_, runs_;
  DevBuf jc_counts_, jc_rlo_, jc_off_;
  DevBuf hits_, hits2_;
  DevBuf segs_raw_, keep_, segs_, gkeys_, gstart_;
  DevBuf bd0_, bd1_, bbest_, tmpjobs_, nbands_, joboff_, jobs_;
  DevBuf stride_, scr_off_, scratch_, scores_;
  DevBuf sel_idx_, cells_, width0_, trace_off_, trace_, ops_off_, ops_, sel_scores_;
  DevBuf sw_keys_, sw_idx_, sw_keys2_, sw_idx2_, sw_lanes_, sw_lanes2_, sw_lane_idx_, sw_counters_,
      sw_redo_, sw_list_, sw_summary_, sw_heads_, sw_hpos_;
  PinBuf sw_pin_summary_;
  DevBuf pair_tab_, pair_tab16_, sw_fallback_, sw_klist_[kWaveKinds];
  // Traceback groups: one per wave kernel K, then the row-exact kernel.
  static constexpr int kTbGroups = kWaveKinds + 1;
  DevBuf tb_counters_, tb_list_[kTbGroups], tb_toff_[kTbGroups], tb_ooff_[kTbGroups],
      tb_trace_[kTbGroups], tb_ops_[kTbGroups], tb_out_[kTbGroups], tb_fb_[kTbGroups], tb_in_,
      tb_sizes_, tb_poff_[kTbGroups], tb_pack_[kTbGroups];
  PinBuf tb_pin_in_, tb_pin_out_, tb_pin_poff_, tb_pin_fb_;
  std::shared_ptr<PinPool> ops_pool_ = std::make_shared<PinPool>();
  DevBuf tb_out_all_, tb_ops_all_, tb_poff_all_, tb_pack_all_;
  DevBuf bk_list_, bk_toff_, bk_ooff_, bk_trace_, bk_ops_, bk_out_, bk_fb_, bk_cnt_;
  PinBuf sel_pin_;  // selection candidates (index, job, score-pass output)
  // Stage counts of a search kept on the device (kCount* slots) and their
  // pinned copies, read at the chain's host syncs.
  static constexpr int kCountSeeds = 0, kCountRuns = 1, kCountHits = 2, kCountHitRuns = 3, kCountSegs = 4,
                       kCountGroups = 5, kCountJobs = 6, kCountSlots = 8;
  DevBuf counts_dev_;
  PinBuf counts_pin_, sw_pin_count_;
  // The ungapped .. score-pass preprocessing chain as a graph (layout key
  // "graph"): captured per slice, updated into one executable graph (built
  // again when its shape changes) and launched; graph_cap_ = the most hits a
  // run without the graph has sized every buffer of the chain for.
  hipGraphExec_t gexec_ = nullptr;
  bool capturing_ = false;
  uint64_t graph_cap_ = 0, graph_launches_ = 0, graph_builds_ = 0;
  void end_capture_and_launch();
  // Host copies of each group's job list and offsets; they stay alive until
  // the copies queued from them have run.
  std::vector<uint64_t> last_cbeg_, last_cend_;
  int cu_count_ = 256;
  bool force_generic_ = false;
  bool lane_report_ = false;
  bool score_serial_ = false;
  std::map<std::string, Settings> tuned_;
  Settings pinned_;
  std::string tuning_file_;
  ScoreLayout layout_ = default_layout();
  std::string layout_preset_ = "default";
  // Layout for a search preset: defaults, the file's "*" and preset entries,
  // then the pinned keys (pinned: in place of pinned_).
  ScoreLayout resolve_layout(const std::string& preset, const Settings* pinned = nullptr) const;
  void report_lanes(const SwJob* jobs, uint64_t njobs, const uint2* lanes, uint32_t nl,
                    const uint32_t* class_lane0, const uint32_t* class_lanes, const uint32_t* class_rows,
                    const ScoreLayout& L);
  uint64_t last_njobs_ = 0;
  double last_ms_class_[kClasses] = {};

 public:
  StageStats last_dropin_stats_;
};

// search::run restated on the host with one drop-in call (one kernel launch)
// per seed::build / seed::join_into / ungapped::xdrop / gapped::extend_span.
// The HSPs' ops live in *ops_store, which the caller keeps with them.
std::vector<std::vector<HspRec>> search_plain(Engine& eng, const TargetSet& t,
                                              const SearchOptions& opt, StageStats* stats,
                                              std::shared_ptr<const OpsStore>* ops_store);

}  // namespace dmr
