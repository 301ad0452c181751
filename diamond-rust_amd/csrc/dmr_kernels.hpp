// Kernel launchers of the seed / extension / gapped stages.
#pragma once
#include "dmr_alp.hpp"
#include "dmr_device.hpp"

#include <vector>

namespace dmr {

// One DB chunk of the resident seed index (all shapes of the preset).
struct RefChunkDev {
  const uint64_t* keys;   // tagged keys, sorted
  const uint32_t* locs;   // global residue index of each seed
  const uint64_t* table;  // bucket lower bounds, nbuckets + 1 entries
  uint64_t n;
  uint64_t nbuckets;
  int shift;              // bucket = key >> shift
  int pad;
};

// ---- seeds ---------------------------------------------------------------
void launch_seed_enumerate(const uint8_t* res, const uint32_t* seq_of, const uint32_t* offs,
                           uint64_t r0, uint64_t r1, const SeedParams* P, uint32_t nshapes,
                           int key_bits, uint32_t* block_counts, uint64_t* block_off,
                           uint64_t* scan_tmp, uint64_t* keys, uint32_t* locs, hipStream_t st);
// The two halves of launch_seed_enumerate: per-block counts and their scan
// (the total is block_off[nblocks * nshapes]), then the emission.
void launch_seed_count(const uint8_t* res, const uint32_t* seq_of, const uint32_t* offs, uint64_t r0,
                       uint64_t r1, const SeedParams* P, uint32_t nshapes, int key_bits, uint32_t* block_counts,
                       uint64_t* block_off, uint64_t* scan_tmp, hipStream_t st);
void launch_seed_emit(const uint8_t* res, const uint32_t* seq_of, const uint32_t* offs, uint64_t r0,
                      uint64_t r1, const SeedParams* P, uint32_t nshapes, int key_bits, const uint64_t* block_off,
                      uint64_t* keys, uint32_t* locs, hipStream_t st);
uint32_t seed_enum_blocks(uint64_t r0, uint64_t r1);
void launch_bucket_table(const uint64_t* keys, uint64_t n, int shift, uint64_t nbuckets,
                         uint64_t* table, hipStream_t st);
// Device counts: launchers that take n_dev (or nruns_dev / nq_dev) run over
// the host count as a bound (the grid) and take the device value as the
// count; their flags and counts past it are 0, so a scan over the bound
// totals the same. No host read of the count is needed.
void launch_run_flags(const uint64_t* keys, uint64_t n, uint32_t* flags, hipStream_t st,
                      const uint64_t* n_dev = nullptr);
void launch_compact_flags(const uint32_t* flags, const uint64_t* pos, uint64_t n, uint32_t* out,
                          hipStream_t st);
void launch_join_count(const uint64_t* qkeys, const uint32_t* run_start, uint64_t nruns,
                       uint64_t nq_total, const RefChunkDev* chunks, uint32_t nchunks,
                       uint64_t max_expand, uint32_t* counts, uint64_t* rlo, hipStream_t st,
                       const uint64_t* nruns_dev = nullptr, const uint64_t* nq_dev = nullptr);
// hit_off: npairs + 1 = nruns * nchunks + 1 exclusive offsets of counts, the
// last the total nhits (which the host knows).
void launch_join_emit(const uint32_t* qlocs, const uint32_t* run_start, uint64_t nruns,
                      uint64_t nq_total, const RefChunkDev* chunks, uint32_t nchunks,
                      const uint32_t* counts, const uint64_t* rlo, const uint64_t* hit_off, uint64_t nhits,
                      const uint32_t* q_seq_of, const uint32_t* q_offs, const uint32_t* db_seq_of,
                      const uint32_t* db_offs, HitGeom g, uint64_t* hits, hipStream_t st,
                      const uint32_t* slice_start = nullptr, const uint32_t* slice_nq = nullptr,
                      const uint64_t* nruns_dev = nullptr, const uint64_t* nq_dev = nullptr);
// Run flags that also split a key's entries by record (group[target]).
void launch_run_flags_grouped(const uint64_t* keys, const uint32_t* locs, const uint32_t* seq_of,
                              const uint32_t* group, uint64_t n, uint32_t* flags, hipStream_t st,
                              const uint64_t* n_dev = nullptr);
// thits[target] += hits of every kept (run, chunk) pair (thits zeroed first).
void launch_target_hits(const uint32_t* qlocs, const uint32_t* run_start, uint64_t nruns, uint64_t nq_total,
                        uint32_t nchunks, const uint32_t* counts, const uint32_t* q_seq_of, uint64_t* thits,
                        hipStream_t st, const uint64_t* nruns_dev = nullptr, const uint64_t* nq_dev = nullptr);
// Per (run, chunk): the hits, first entry and entry count of the part whose
// targets lie in [t0, t1).
void launch_join_slice(const uint32_t* qlocs, const uint32_t* run_start, uint64_t nruns, uint64_t nq_total,
                       uint32_t nchunks, const uint32_t* counts, const uint32_t* q_seq_of, uint32_t t0,
                       uint32_t t1, uint32_t* slice_counts, uint32_t* slice_start, uint32_t* slice_nq,
                       hipStream_t st, const uint64_t* nruns_dev = nullptr,
                       const uint64_t* nq_dev = nullptr);

// ---- ungapped + bands ----------------------------------------------------
struct SeqSet {
  const uint8_t* res;
  const uint32_t* offs;  // n+1
  // Padded letters for the packed score kernels (null where unused): letter
  // codes times 4, sequence k starting at pad[pad_offs[k]], with the padding
  // code (25 * 4) at pad[pad_offs[k] - 1] and pad[pad_offs[k] + len_k], and
  // at pad[0] and pad[1] (base 1 serves as an empty sequence).
  const uint8_t* pad = nullptr;
  const uint32_t* pad_offs = nullptr;
};

// Bytes of the padded letters of n sequences of `total` letters laid out one
// after another: two padding codes, then each sequence followed by one.
inline uint64_t pad_letters_bytes(uint64_t n, uint64_t total) { return total + n + 2; }
// Builds that layout from s.res/s.offs (offs non-decreasing from 0): pad_offs[k]
// = offs[k] + k + 2.
void launch_pad_letters(SeqSet s, uint32_t n, uint8_t* pad, uint32_t* pad_offs, hipStream_t st);

// Six-frame translation and --min-orf masking (k_translate.hip). Reads are
// ASCII bases at nt[roff[r]..roff[r+1]]; frame k of read r is slot 6r + k
// at foff[6r + k] (6n + 1 offsets); min_orf holds one length per read.
void launch_translate_frames(const uint8_t* nt, const uint64_t* roff, uint32_t n, const uint64_t* foff,
                             uint8_t* out, hipStream_t st);
void launch_mask_orfs(const uint64_t* foff, uint32_t nframes, const uint32_t* min_orf, const uint8_t* in,
                      uint8_t* out, uint32_t* good, hipStream_t st);
void launch_kept_lengths(const uint64_t* foff, const uint32_t* good, uint32_t nframes, uint32_t* keep,
                         uint32_t* klen, hipStream_t st);
void launch_compact_frames(const uint64_t* foff, const uint32_t* good, const uint64_t* kept_before,
                           const uint64_t* tletters, const uint8_t* in, uint8_t* letters, uint32_t* toffs,
                           uint32_t* query_of, uint8_t* frame_of, uint32_t nframes, hipStream_t st);
// seq_of[p] = k for every letter p of sequence k (offs: n + 1 entries).
void launch_seq_of(const uint32_t* offs, uint32_t n, uint32_t* seq_of, hipStream_t st);

// Key flags over packed hits: 1 where (hit >> shift) changes.
void launch_key_flags(const uint64_t* keys, uint64_t n, int shift, uint32_t* flags, hipStream_t st,
                      const uint64_t* n_dev = nullptr);
// dst[k] = src[k], k < n <= 64 (one small kernel, for counts).
void launch_copy_u64(const uint64_t* src, uint64_t* dst, uint32_t n, hipStream_t st);
void launch_xdrop_runs(const uint64_t* hits, const uint32_t* run_start, uint64_t nruns,
                       uint64_t nhits, HitGeom g, SeqSet q, SeqSet db, const int8_t* matrix32,
                       int xdrop, int cutoff, SegRec* segs, uint32_t* keep, hipStream_t st,
                       const uint64_t* nruns_dev = nullptr);
void launch_compact_segments(const SegRec* segs, const uint64_t* hits, const uint32_t* keep,
                             const uint64_t* pos, uint64_t nhits, int group_shift, SegRec* out,
                             uint64_t* out_gkey, hipStream_t st);
void launch_bands(const SegRec* segs, const uint64_t* gkey, const uint32_t* group_start,
                  uint64_t ngroups, uint64_t nsegs, uint32_t bs, SeqSet q, SeqSet db,
                  int32_t* tmp_d0, int32_t* tmp_d1, uint32_t* tmp_best, SwJob* tmp_jobs,
                  uint32_t* nbands, hipStream_t st, const uint64_t* ngroups_dev = nullptr,
                  const uint64_t* nsegs_dev = nullptr);
void launch_jobs_compact(const SwJob* tmp_jobs, const uint32_t* group_start, const uint32_t* nbands,
                         const uint64_t* job_off, uint64_t ngroups, SwJob* jobs, hipStream_t st,
                         const uint64_t* ngroups_dev = nullptr);

// ---- gapped --------------------------------------------------------------
// Per job: DP cells (sum of row widths, rows 1..n) and the int32 row stride.
void launch_job_geometry(const SwJob* jobs, const uint32_t* job_idx, uint64_t njobs, SeqSet db,
                         uint64_t* cells, uint64_t* width0, uint32_t* stride, hipStream_t st);
void launch_sw_generic(const SwJob* jobs, const uint32_t* job_idx, uint64_t njobs, SeqSet q,
                       SeqSet db, const int8_t* matrix32, int gap_open, int gap_extend,
                       int32_t* scratch, const uint64_t* scratch_off, uint8_t* trace,
                       const uint64_t* trace_off, uint32_t* ops, const uint64_t* ops_off,
                       SwOut* out, bool traceback, hipStream_t st, bool scatter = false,
                       int32_t* scores = nullptr);  // scores: score mode writes scores[job] only

// Packed int16 score pass (k_sw16.hip).
int pair_table_entries();
// Pair-score table of the packed kernels: entry (qa, sA, sB) holds
// (score(qa, sA) + o, score(qa, sB) + o) as two int16 or, f16, two halves.
void fill_pair_table(const int8_t (&m)[25][25], bool f16, std::vector<int>* out);
// Packed score classes: 0-4 the band_for widths (2 * band + 1 = 25, 33, 61,
// 81, 129); kExactClasses + i for merged bands (widths 62..512 other than 81
// and 129; per-DP walls): i < 6 up to width 97 + 32 i (m97 .. m225, then
// m257 for widths up to 256), i = 6, 7 up to 384 and 512 (m385, m513).
constexpr int kExactClasses = 5;
constexpr int kMergedClasses = 8;
// Merged kernel always in use (m257: every merged width up to 256 is
// packed); m385 and m513 are optional (without them, widths past 256 go to
// the int32 wave-per-DP kernels, which finish a few long DPs sooner).
constexpr int kMergedAlways = 5;
constexpr int kPackedClasses = kExactClasses + kMergedClasses;
// Width of class cls (0-4) or the width bound of a merged class.
int sw16_width(int cls);
// Class table of the merged widths for the merged kernels in use (bit i of
// mask = class kExactClasses + i; m257 is always in use): a merged width
// goes to the narrowest kernel in use that holds it (4 bits per width
// bucket: 62..97, then 32 widths each up to 256, then 64 each up to 512; 15
// where no kernel in use holds the bucket).
uint64_t sw16_merged_map(unsigned mask);
// A compiled packed kernel: class, its width (bound), lanes per DP pair (1 =
// k_sw16, the whole band in one lane; otherwise k_sw16p split over that many
// lanes) and threads per block.
struct Sw16Variant {
  int cls, width, lanes, threads;
};
const std::vector<Sw16Variant>& sw16_variants();
bool sw16_has_variant(int cls, int lanes, int threads);
// Score-pass sort keys (class | orientation | row sequence | rows, see
// k_classify). seq_bits: bits of the larger of a target and a subject index;
// transpose: per packed class c, bits 2c..2c+1 say which sequence runs along
// the rows of its DPs: 0 the query, 1 the one giving fewer rows, 2 the
// subject (every result is the same).
void launch_classify(const SwJob* jobs, uint64_t njobs, SeqSet q, SeqSet db, int seq_bits, uint64_t merged_map,
                     uint32_t transpose, uint64_t* keys, uint32_t* idx, hipStream_t st,
                     const uint64_t* njobs_dev = nullptr);
// Lane pairing of the sorted score-pass jobs (see k_lane_heads): head flags
// and class ranges, then lanes at the heads' scan positions, then the lanes
// per class. summary holds 2 * (nclasses + npacked) + 1 words, then at
// tfirst and trow npacked words each (first job with the subject on the rows,
// lanes with the query on the rows).
void launch_lane_heads(const uint64_t* keys, const uint32_t* flags, const uint64_t* pos, const uint32_t* run_start,
                       uint64_t n, int cls_shift, int npacked, int nclasses, int tfirst, uint32_t* heads,
                       uint32_t* summary, hipStream_t st, const uint64_t* n_dev = nullptr);
void launch_make_lanes(const uint32_t* sorted_idx, const uint32_t* flags, const uint32_t* heads,
                       const uint64_t* hpos, uint64_t n, uint2* lanes, hipStream_t st,
                       const uint64_t* n_dev = nullptr);
void launch_lane_summary(const uint64_t* hpos, uint64_t n, int npacked, int nclasses, int tfirst, int trow,
                         uint32_t* summary, hipStream_t st);
// Lane order of the packed classes: key = class << (kLaneSpanBits + 1) |
// orientation << kLaneSpanBits | (kLaneSpanMax - row span), so a sort keeps
// each class's query-row lanes first and puts the longest lanes of each
// orientation first.
constexpr int kLaneSpanBits = 13;
constexpr uint32_t kLaneSpanMax = (1u << kLaneSpanBits) - 1;
void launch_lane_keys(const uint2* lanes, uint32_t nlanes, const SwJob* jobs, SeqSet q, SeqSet db,
                      uint64_t merged_map, uint32_t transpose, uint64_t* keys, uint32_t* idx, hipStream_t st);
void launch_gather_lanes(const uint2* in, const uint32_t* idx, uint32_t n, uint2* out, hipStream_t st);
// Packed score pass of class cls over nlanes lanes (k_sw16.hip, merged
// classes k_sw16m.hip) with the compiled kernel (cls, lanes_per_pair,
// threads) of sw16_variants(); false when there is none or nrow > nlanes.
// Lanes [0, nrow) have the query on the rows, the rest the subject. f16 =
// half-precision arithmetic (pair_tab must then be the f16 table).
bool launch_sw16(int cls, int lanes_per_pair, int threads, bool f16, const uint2* lanes, uint32_t nlanes,
                 uint32_t nrow, const SwJob* jobs, SeqSet q, SeqSet db, const int* pair_tab, const int8_t* m32,
                 uint32_t* counter, int32_t* scores, uint32_t* redo, uint32_t* nredo, int cu_count, hipStream_t st);
bool launch_sw16_merged(int cls, int lanes_per_pair, int threads, bool f16, const uint2* lanes, uint32_t nlanes,
                        uint32_t nrow, const SwJob* jobs, SeqSet q, SeqSet db, const int* pair_tab, uint32_t* counter,
                        int32_t* scores, uint32_t* redo, uint32_t* nredo, int cu_count, hipStream_t st);
// Fused launch of the packed classes m513, m385 (sixteen lanes per DP pair),
// m257, m225, m193, m161, m129, m97 (eight lanes; all with per-DP walls),
// w81 and w61 (two lanes per pair): one persistent
// grid of 512-thread blocks, kinds in that order (k_sw16p_fused). Per kind:
// its lanes, lane count and query-row lane count; task0 are the first task
// of each kind (computed by launch_sw16_fused).
constexpr int kFusedKinds = 10;
struct FusedSections {
  const uint2* lanes[kFusedKinds];
  uint32_t nlanes[kFusedKinds], nrow[kFusedKinds], task0[kFusedKinds];
  uint32_t ntasks;
};
// Packed class of each fused kind (the merged bands widest first, then w81
// and w61).
int sw16_fused_class(int kind);
bool launch_sw16_fused(bool f16, FusedSections s, const SwJob* jobs, SeqSet q, SeqSet db, const int* pair_tab,
                       uint32_t* counter, int32_t* scores, uint32_t* redo, uint32_t* nredo, int cu_count,
                       hipStream_t st);
void launch_evalue_filter(const SwJob* jobs, const int32_t* scores, uint64_t njobs, SeqSet q,
                          SeqSet db, AlpParams P, uint64_t db_letters, double threshold,
                          uint32_t* flags, hipStream_t st);
void launch_u32_times(const uint32_t* in, uint64_t n, uint32_t mul, uint64_t* out, hipStream_t st);
// sums[j] += v[ranges[2j] .. ranges[2j + 1]) for j < k; longest = the
// largest range length (sums are not cleared).
void launch_range_sums(const uint64_t* v, const uint64_t* ranges, uint32_t k, uint64_t longest, uint64_t* sums,
                       hipStream_t st);
void launch_gather_u64(const uint64_t* in, const uint32_t* idx, uint64_t n, uint64_t* out, hipStream_t st);

// Persistent kernels (k_sw16, k_sw_wave) launch as many blocks as can be
// resident on cu_count CUs at once (from the occupancy of the instantiation),
// capped by the work available.
template <class Kernel>
unsigned resident_blocks(Kernel kernel, int threads, int cu_count) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  return (unsigned)per_cu * (unsigned)(cu_count > 0 ? cu_count : 1);
}

// Score-pass sort key of a DP job: class | orientation | row sequence | row
// field, the row field being (row span / 32, first row / 32), 9 bits each.
constexpr int kRowKeyBits = 18;

// Wave-per-DP int32 kernel (k_swwave.hip): K offsets per lane, band width
// <= 64*K, K a power of two up to kMaxWaveK (0 = wider than 64*kMaxWaveK).
constexpr int kMaxWaveK = 32;
// Block-per-DP int32 kernel (k_swblock.hip): 16 waves, K offsets per thread,
// widths up to 1024*K for K = 4, 8, 16 (0 = wider).
constexpr int kMaxBlockK = 16;
int swblock_k_for(int64_t width);
uint64_t swblock_row_bytes(int K);  // trace bytes of one row
// Every DP of list whose width lies in (2048, 16384] (only_k: one K's range,
// 0 = all); score mode writes scores[job], trace mode outs[task] and the
// task's ops, and needs the score pass's scores (known, by job). Wider DPs
// and DPs whose best may lie in the clamped rows go to the fallback list (job
// ids in score mode, tasks in trace mode).
void launch_sw_block(bool traceback, const SwJob* jobs, const uint32_t* list, uint32_t njobs, SeqSet q, SeqSet db,
                     const int8_t* m32, int gap_open, int gap_extend, SwOut* outs, int32_t* scores,
                     uint8_t* trace, const uint64_t* trace_off, uint32_t* ops, const uint64_t* ops_off,
                     uint32_t* fallback, uint32_t* nfallback, const int32_t* known, int only_k, hipStream_t st);
constexpr int kWaveKinds = 6;  // K = 1, 2, 4, 8, 16, 32
int swwave_k_for(int64_t width);
// Packs the run-length ops of n traceback DPs (capacity slots at ops_off)
// into packed, at pack_off (n + 1 offsets; pack_off[n] = total words).
// sizes: n + 1 words of scratch; scan_tmp: scan_tmp_elems(n + 1).
void launch_ops_pack(const SwOut* outs, const uint32_t* ops, const uint64_t* ops_off, uint32_t n,
                     uint32_t* sizes, uint64_t* pack_off, uint64_t* scan_tmp, uint32_t* packed,
                     hipStream_t st);
// Wave-per-DP int32 kernel over jobs[list[0..njobs)]: score mode writes
// scores[job]; trace mode writes outs[task] and the task's ops, stopping each
// DP at the first cell that reaches known[job] (the score pass's scores).
// list_sel 1 / 2: the list holds tagged entries (bit 31 set for bands wider
// than 256, the packed kernels' redo list) and the kernel takes only the
// untagged / tagged ones.
void launch_sw_wave(int K, bool traceback, const SwJob* jobs, const uint32_t* list, uint32_t njobs,
                    SeqSet q, SeqSet db, const int8_t* m32, int gap_open, int gap_extend,
                    uint32_t* counter, SwOut* outs, int32_t* scores, uint8_t* trace, const uint64_t* trace_off,
                    uint32_t* ops, const uint64_t* ops_off, uint32_t* fallback, uint32_t* nfallback,
                    int cu_count, hipStream_t st, const int32_t* known = nullptr,
                    const uint32_t* njobs_dev = nullptr,  // njobs_dev: a device-side count (njobs bounds it)
                    uint32_t list_sel = 0);

}  // namespace dmr
