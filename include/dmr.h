/*
 * dmr.h — C ABI of the MI355X (gfx950) search core.
 *
 * Each entry point replaces one call of diamond-rust's search core; the
 * comment above it names the reference function (file:line in the
 * diamond-rust repository) whose arguments and results it mirrors. All
 * buffers are host memory owned by the caller; sizes are element counts.
 * Functions return DMR_OK or a negative status; dmr_last_error() gives the
 * message of the last failure on the calling thread.
 *
 * Residues are NCBI-order letter codes 0..24
 * (A R N D C Q E G H I L K M F P S T W Y V B J Z X *), as produced by
 * alphabet::encode (src/alphabet.rs:54-63). Sequence sets are passed as one
 * concatenated buffer plus nseq+1 offsets.
 */
#ifndef DMR_H
#define DMR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DMR_OK 0
#define DMR_E_ARG (-1)         /* invalid argument (bad shape, gap costs, sizes) */
#define DMR_E_DEVICE (-2)      /* HIP runtime failure */
#define DMR_E_OOM (-3)         /* device allocation failed */
#define DMR_E_CAPACITY (-4)    /* output buffer too small; *n_out holds the size needed */
#define DMR_E_UNSUPPORTED (-5) /* input outside what the library handles */
#define DMR_E_TIMEOUT (-6)     /* dmr_server_wait: the request is not done yet */

typedef struct dmr_ctx dmr_ctx;
typedef struct dmr_result dmr_result;

/* seed::SeedEntry (src/seed.rs:14-19) */
typedef struct {
  uint64_t key;
  uint32_t seq;
  uint32_t pos;
} dmr_seed_entry;

/* seed::SeedHit (src/seed.rs:21-27) */
typedef struct {
  uint32_t query, qpos, subject, spos;
} dmr_seed_hit;

/* ungapped::Segment (src/ungapped.rs:10-16) */
typedef struct {
  uint64_t q_start, s_start, len;
  int32_t score;
  int32_t reserved;
} dmr_segment;

/* hsp::HspStats (src/hsp.rs:29-42) */
typedef struct {
  uint32_t identities, positives, mismatches, gap_openings, gaps, length;
} dmr_hsp_stats;

/* gapped::Aligned (src/gapped.rs:34-43); found = 0 stands for None.
 * Alignment ops are returned separately as uint32 (op | count << 2) with
 * op 0 = Match, 1 = Ins (gap in subject), 2 = Del (gap in query). */
typedef struct {
  int32_t found;
  int32_t score;
  uint64_t q_begin, q_end, s_begin, s_end;
  dmr_hsp_stats stats;
  uint32_t n_ops;
  uint32_t reserved;
} dmr_aligned;

/* One extend_span call for dmr_extend_span_batch: query and subject slices
 * of the caller's buffers plus the diagonal range [d0, d1]. */
typedef struct {
  uint64_t q_off, q_len, s_off, s_len;
  int64_t d0, d1;
} dmr_span_job;

/* search::SearchConfig (src/search.rs:19-33) and the blastp/blastx options
 * the search reads (src/cli.rs:125-269). NULL sensitivity = "default";
 * index_chunks < 0 and seed_cut < 0 keep the preset; gap costs -1 keep the
 * matrix default (BLOSUM62 11/1). */
typedef struct {
  const char* sensitivity;
  double evalue;
  uint32_t max_target_seqs;
  uint32_t max_hsps;
  int32_t index_chunks;
  int32_t gap_open;
  int32_t gap_extend;
  int32_t min_orf; /* blastx --min-orf; < 0 = length-dependent default */
  double seed_cut;
} dmr_search_config;

/* Fills the reference defaults: default sensitivity, evalue 1e-3, k 25,
 * max-hsps 1 (src/search.rs:36-48, src/cli.rs:125-140). */
void dmr_search_config_default(dmr_search_config* cfg);

const char* dmr_last_error(void);
const char* dmr_version(void);

/* Context: one HIP device, one stream, the resident database and seed index. */
int dmr_ctx_create(int device, dmr_ctx** out);
void dmr_ctx_destroy(dmr_ctx* ctx);
/* ctx takes the resident database (letters, ids, headers) and seed indexes
 * of `from` (same device) in place of its own, with no copy; no counterpart
 * in the reference, whose rayon workers share one &Database
 * (src/search.rs:135-145). Contexts sharing a database may search
 * concurrently from different host threads, each with its own query batch
 * and stream, so one batch's host work and latency-bound kernels overlap
 * another's score pass; an index missing for a preset is built once, under
 * the database's lock. The database lives while some context holds it; a
 * later dmr_db_load* on either context replaces only that context's. */
int dmr_ctx_share_db(dmr_ctx* ctx, const dmr_ctx* from);
int dmr_device_count(int* n);

/* Engine settings of a context; no counterpart in the reference (none of
 * them changes a result). Keys:
 *   "slice_hits"  seed hits per slice of query targets: a search runs its
 *                 targets in slices whose hits stay under this (device
 *                 memory grows with it, ~52 bytes per hit and ~150 per DP);
 *   "profile"     1 = keep per-class DP jobs and cells over every slice of a
 *                 search (for dmr_last_class_profile / dmr_last_dp_cells);
 * and the score-pass kernels, per packed class <c> (w25 w33 w61 w81 w129:
 * band_for widths; m97 m129 m161 m193 m225 m257 m385 m513: merged bands up
 * to that width, m257 up to 256):
 *   "lanes_<c>"   lanes per DP pair (1 = the whole band in one lane, w*
 *                 only; else a compiled split kernel, dmr_sw16_variants);
 *                 also sets the class's first compiled block size for it;
 *   "threads_<c>" threads per block of a compiled kernel for those lanes;
 *   "merged_mask" merged kernels in use (bit i = class m97 .. m513 in that
 *                 order; m257 always; without m385 and m513, bands wider
 *                 than 256 run the int32 wave-per-DP kernels):
 *                 a merged band goes to the narrowest one that holds it;
 *   "f16"         1 = packed pairs in f16, 0 = saturating int16;
 *   "score_plan"  how the score-class kernels share the streams (0, 1, 2);
 *   "graph"       1 = the chain from the sorted hits to the score pass's
 *                 class summary (ungapped extension, bands, job keys, sort,
 *                 lane pairing) is captured and launched as one hipGraph,
 *                 its executable graph updated in place from batch to batch;
 *   "transpose_<c>" which sequence runs along the rows of the class's DPs:
 *                 0 the query, 1 the one giving fewer rows (fewer band cells
 *                 past the matrix), 2 the subject; the best score of a band
 *                 is the same either way ("transpose" sets every class and
 *                 reads -1 when they differ);
 *   "fuse"        1 = the merged-band classes in use (eight lanes per
 *                 pair) and w81, w61 (two lanes) run in one persistent
 *                 launch (their lanes_/threads_ keys then do not apply), the
 *                 merged bands' tasks first, widest first.
 * Score-pass settings come from built-in defaults, then the tuning file (the
 * file DMR_TUNING names at context creation, empty for none; else
 * tuning_<arch>.json beside the library when it exists; settings per search
 * preset, written by the autotuner), then the keys set here, which stay
 * pinned over every preset. A key with no compiled kernel is refused
 * (DMR_E_ARG) and changes nothing. dmr_ctx_get_tuning reads the settings of
 * the last search (before the first: those of the default preset), plus
 * "last_slices" (slices of the last search), "tuned_presets" (presets in
 * the loaded tuning file), "graph_launches" and "graph_builds" (graph
 * launches and instantiations so far). */
int dmr_ctx_set_tuning(dmr_ctx* ctx, const char* key, double value);
int dmr_ctx_get_tuning(const dmr_ctx* ctx, const char* key, double* value);
/* Replaces the context's tuning file (NULL or "" = none); the pinned keys
 * stay. A malformed file, or one naming an unknown key or a kernel that is
 * not compiled, is refused (DMR_E_ARG) and changes nothing. */
int dmr_ctx_load_tuning(dmr_ctx* ctx, const char* path);
/* Compiled packed score kernels: *n entries of (class, width or width bound,
 * lanes per DP pair, threads per block) into out (4 * cap int32). out NULL:
 * *n only; cap too small: DMR_E_CAPACITY. */
int dmr_sw16_variants(int32_t* out, uint32_t cap, uint32_t* n);

/* gapped::band_for (src/gapped.rs:27-32) */
uint32_t dmr_band_for(uint64_t len);

/* ScoreMatrix::raw_score / bit_score / evalue for BLOSUM62 with the given
 * gap costs (src/matrix.rs:182-200, src/alp.rs:108-122). */
int dmr_evalue(int32_t gap_open, int32_t gap_extend, int32_t raw, uint64_t query_len,
               uint64_t subject_len, uint64_t db_letters, double* evalue, double* bits);

/* seed::build (src/seed.rs:76-104): every seed of seqs (shape code of 0/1,
 * reduction as DIAMOND group strings, e.g. "A KR EDNQ C G H ILVM FYW P ST"),
 * sequence numbers offset by base, sorted by (key, seq, pos). */
int dmr_seed_build(dmr_ctx* ctx, const uint8_t* seqs, const uint64_t* offsets, uint32_t nseq,
                   const char* shape_code, const char* reduction_def, uint32_t base,
                   double complexity_cut, dmr_seed_entry* out, uint64_t cap, uint64_t* n_out);

/* seed::join_into (src/seed.rs:110-143) for key-sorted arrays; hits are
 * returned sorted by (query, subject, qpos - spos, spos), the order
 * align_target consumes them (src/search.rs:167-178). */
int dmr_seed_join(dmr_ctx* ctx, const dmr_seed_entry* query, uint64_t nq,
                  const dmr_seed_entry* reference, uint64_t nr, uint64_t max_expand,
                  dmr_seed_hit* out, uint64_t cap, uint64_t* n_out);

/* ungapped::xdrop (src/ungapped.rs:38-88). */
int dmr_xdrop(dmr_ctx* ctx, const uint8_t* query, uint64_t qlen, const uint8_t* subject,
              uint64_t slen, uint64_t qa, uint64_t sa, int32_t xdrop, dmr_segment* out);
/* The same for n seeds of one sequence pair. */
int dmr_xdrop_batch(dmr_ctx* ctx, const uint8_t* query, uint64_t qlen, const uint8_t* subject,
                    uint64_t slen, const uint32_t* qa, const uint32_t* sa, uint64_t n,
                    int32_t xdrop, dmr_segment* out);

/* gapped::extend_span (src/gapped.rs:67-280). ops receives up to ops_cap
 * run-length ops; out->n_ops is the number produced (DMR_E_CAPACITY when
 * larger than ops_cap). */
int dmr_extend_span(dmr_ctx* ctx, const uint8_t* query, uint64_t qlen, const uint8_t* subject,
                    uint64_t slen, int64_t d0, int64_t d1, int32_t gap_open, int32_t gap_extend,
                    dmr_aligned* out, uint32_t* ops, uint32_t ops_cap);

/* No counterpart in the reference: dmr_extend_span for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). Same arguments and results; it calls no
 * hipMalloc, hipFree or other allocating routine. Its device memory (the
 * sequences, the DP rows, the trace and the ops) comes from one static
 * array of 32 MiB per GPU, held until the process ends; a DP needing more
 * (about qlen x band width trace bytes plus 16 bytes per band cell of a row
 * and 4 bytes per letter of both sequences) fails with DMR_E_UNSUPPORTED.
 * All calls on one GPU share that array: calls that overlap on one GPU (from
 * two threads, or two contexts of the same device) race. */
int dmr_extend_span_noalloc(dmr_ctx* ctx, const uint8_t* query, uint64_t qlen, const uint8_t* subject,
                            uint64_t slen, int64_t d0, int64_t d1, int32_t gap_open, int32_t gap_extend,
                            dmr_aligned* out, uint32_t* ops, uint32_t ops_cap);

/* gapped::extend (src/gapped.rs:50-60): band around the anchor diagonal. */
int dmr_extend(dmr_ctx* ctx, const uint8_t* query, uint64_t qlen, const uint8_t* subject,
               uint64_t slen, uint64_t anchor_q, uint64_t anchor_s, uint64_t band,
               int32_t gap_open, int32_t gap_extend, dmr_aligned* out, uint32_t* ops,
               uint32_t ops_cap);

/* Many extend_span calls in one launch. With traceback = 0 only score is
 * filled (found = score > 0). ops_offsets (njobs + 1) index ops_pool;
 * *ops_needed receives the pool size required. */
int dmr_extend_span_batch(dmr_ctx* ctx, const uint8_t* qbuf, uint64_t qbuf_len,
                          const uint8_t* sbuf, uint64_t sbuf_len, const dmr_span_job* jobs,
                          uint64_t njobs, int32_t gap_open, int32_t gap_extend, int traceback,
                          dmr_aligned* out, uint32_t* ops_pool, uint64_t ops_cap,
                          uint64_t* ops_offsets, uint64_t* ops_needed);

/* Database::open + build (src/db.rs:159-291) as a resident device copy:
 * encoded letters and nseq + 1 offsets. Replaces any previous database. */
int dmr_db_load(dmr_ctx* ctx, const uint8_t* letters, const uint64_t* offsets, uint32_t nseq);

/* Database::open (src/db.rs:159-224, with db::resolve, src/db.rs:39-60) of a
 * DRDMND01 file made by `diamond-rust makedb`, loaded as the resident
 * database as by dmr_db_load; *nseq (may be NULL) receives its sequence
 * count. Errors carry the reference's messages (file not found, file too
 * short, bad magic, unsupported version, truncated, corrupt offset table);
 * offsets and letter codes the reference reads only on use are checked at
 * load. On error the previous database stays resident. */
int dmr_db_load_file(dmr_ctx* ctx, const char* path, uint32_t* nseq);

/* Database::id of every sequence of the database that dmr_db_load_file
 * loaded (src/db.rs:262-265, alphabet::seq_id src/alphabet.rs:87-93), each
 * followed by '\n': the subject_ids of dmr_result_format. *len receives the
 * byte length; the text (not NUL-terminated) is written when cap >= *len. */
int dmr_db_ids(const dmr_ctx* ctx, char* buf, uint64_t cap, uint64_t* len);

/* Database::header of every sequence of the database that dmr_db_load_file
 * loaded (src/db.rs:262-266: the header bytes as text, empty when they are
 * not UTF-8), each followed by '\n': the subject_titles of dmr_result_format
 * (stitle). *len receives the byte length; the text is written when
 * cap >= *len. DMR_E_UNSUPPORTED when a header holds a newline. */
int dmr_db_headers(const dmr_ctx* ctx, char* buf, uint64_t cap, uint64_t* len);

/* Builds (or keeps) the resident seed index for a sensitivity preset:
 * seed::chunks(n, index_chunks) slices x all shapes (src/search.rs:121-133). */
int dmr_db_build_index(dmr_ctx* ctx, const char* sensitivity, int32_t index_chunks,
                       double seed_cut, double* build_ms);

/* search::run + merge_frames (src/search.rs:95-146, 357-396) for blastp:
 * queries are encoded letters. Results stay on the handle. */
int dmr_blastp(dmr_ctx* ctx, const uint8_t* queries, const uint64_t* q_offsets, uint32_t nq,
               const dmr_search_config* cfg, dmr_result** out);

/* blastx: queries are nucleotide ASCII; six frames with --min-orf masking
 * (src/translate.rs:54-137, src/workflow.rs:112-127), translated on the
 * device; frames without a letter below X are not searched. */
int dmr_blastx(dmr_ctx* ctx, const char* reads, const uint64_t* r_offsets, uint32_t nreads,
               const dmr_search_config* cfg, dmr_result** out);

/* translate::six_frames + mask_short_orfs (src/translate.rs:54-121) of every
 * read, on the device: frame k of read r (0-2 forward offsets 0-2, 3-5 the
 * reverse complement's) at letters[frame_offsets[6r + k] ..
 * frame_offsets[6r + k + 1]], as letter codes (X for ambiguous codons, * for
 * stops). min_orf < 0 = translate::min_orf_len's default per read length.
 * frame_offsets holds 6 * nreads + 1 entries (may be NULL); *n_letters
 * receives the letter count; the letters are written when cap >= it. */
int dmr_six_frames(dmr_ctx* ctx, const char* reads, const uint64_t* r_offsets, uint32_t nreads, int32_t min_orf,
                   uint8_t* letters, uint64_t cap, uint64_t* frame_offsets, uint64_t* n_letters);

/* Uploads a blastp query batch and keeps it resident on the device, so
 * dmr_search_loaded can run without any host-to-device copy of the inputs
 * (the same search as dmr_blastp). */
int dmr_queries_load(dmr_ctx* ctx, const uint8_t* queries, const uint64_t* q_offsets, uint32_t nq);
int dmr_search_loaded(dmr_ctx* ctx, const dmr_search_config* cfg, dmr_result** out);

/* The plain drop-in form: search::run restated on the host exactly as the
 * reference structures it (src/search.rs:95-271), with each seed::build,
 * seed::join_into, ungapped::xdrop and gapped::extend_span call replaced by
 * one call of the entry points above (one kernel launch per call). Same
 * results as dmr_blastp; kept as the per-call baseline of the batched form. */
int dmr_blastp_plain(dmr_ctx* ctx, const uint8_t* queries, const uint64_t* q_offsets, uint32_t nq,
                     const dmr_search_config* cfg, dmr_result** out);

/* Number of HSPs (all queries, output order). */
int64_t dmr_result_count(const dmr_result* r);
/* 16 int64 per HSP: query, subject, frame, q_begin, q_end, s_begin, s_end,
 * score, identities, positives, mismatches, gap_openings, gaps, length,
 * n_ops, 0; and 2 doubles: bitscore, evalue. Coordinates are 0-based
 * half-open in the frame's amino-acid space. */
int dmr_result_hsps(const dmr_result* r, int64_t* out16, double* out2);
/* Alignment ops of HSP i (op | count << 2). */
int dmr_result_ops(const dmr_result* r, int64_t i, uint32_t* ops, uint32_t cap, uint32_t* n);
/* outfmt 6 text (Tabular::write, src/output.rs:150-223). query_ids and
 * subject_ids are newline-separated (Record::id / Database::id);
 * query_titles and subject_titles are the full headers, newline-separated
 * (Record::header / Database::header: qtitle and stitle,
 * src/output.rs:201-202), and may be NULL when neither field is asked for.
 * fields is a space-separated list of outfmt-6 field names (NULL or "" = the
 * 12 defaults). *len receives the byte length; the text is written when buf
 * holds cap >= *len bytes. */
int dmr_result_format(const dmr_result* r, const char* query_ids, const char* subject_ids,
                      const char* query_titles, const char* subject_titles, const char* fields,
                      char* buf, uint64_t cap, uint64_t* len);
/* Stage counters and GPU times: query_seeds, runs, hits, hit_runs,
 * segments, groups, jobs, maybe, selected, traceback_cells, ms_seed,
 * ms_join, ms_sort_hits, ms_ungapped, ms_bands, ms_sw, ms_select,
 * ms_traceback, ms_total, ms_host, ms_translate (blastx: six frames, masking
 * and the target set on the device) (21 doubles). */
int dmr_result_stats(const dmr_result* r, double* out21);
/* Cells of every banded DP of the last search on ctx (for GCUPS). */
int dmr_last_dp_cells(dmr_ctx* ctx, uint64_t* cells);
/* Score-pass profile of the last search on ctx, per kernel class
 * (DMR_SCORE_CLASSES entries): 0-4 packed 16-bit kernel for band widths
 * 25/33/61/81/129, 5-12 packed kernel for merged bands of widths up to
 * 97/129/161/193/225/256/384/512 (those in use; see "merged_mask"), 13-18
 * wave-per-DP int32 kernel for widths up to 64/128/256/512/1024/2048 (bands
 * the packed kernels do not take), 19 block-per-DP and row-exact int32
 * kernels (wider bands, and bands that end before their clamped rows): DP
 * jobs, DP cells and kernel milliseconds. Diagnostics only; no counterpart in
 * the reference. */
#define DMR_SCORE_CLASSES 20
int dmr_last_class_profile(dmr_ctx* ctx, uint64_t* jobs, uint64_t* cells, double* ms);
/* Layout of the packed classes 0-4 in the last search on ctx: bit c of
 * *split_mask set = class c ran split over several lanes per DP pair
 * (k_sw16p), clear = one lane (k_sw16); *f16 = 1 when the packed kernels
 * computed in f16 pairs, 0 in saturating int16 pairs. Either may be NULL.
 * Diagnostics only; no counterpart in the reference. */
int dmr_score_layout(dmr_ctx* ctx, uint32_t* split_mask, uint32_t* f16);
void dmr_result_destroy(dmr_result* r);

/* ---- Serving loop ----------------------------------------------------------
 * No counterpart in the reference (its CLI searches one batch per process):
 * a request loop over the context's resident database and seed indexes.
 * Requests (one encoded protein query each, with its own config: preset,
 * evalue, limits) are submitted at any time; a dispatcher thread forms
 * batches of queued requests with equal configs, oldest first, as workers
 * come free and the device-memory budget admits them, and each worker
 * searches its batch on its own streams, so batches overlap on the GPU.
 * The max_expand cap (src/search.rs:109-131, src/seed.rs:123-125) counts a
 * key's query entries per request, so every result equals that of
 * dmr_blastp on the request alone. The context must stay alive and keep its
 * database while the server runs. */
typedef struct dmr_server dmr_server;
/* workers: 1..8 batches in flight; hbm_budget: device bytes the admitted
 * requests' estimates may total (at least 64 MiB); presets: comma-separated
 * sensitivities whose seed indexes are built now (may be NULL). */
int dmr_server_create(dmr_ctx* ctx, uint32_t workers, uint64_t hbm_budget, const char* presets,
                      dmr_server** out);
/* Queues one query; *ticket identifies it for dmr_server_wait. */
int dmr_server_submit(dmr_server* s, const uint8_t* query, uint64_t len, const dmr_search_config* cfg,
                      uint64_t* ticket);
/* Waits up to timeout_ms (< 0: without limit) for a request; on success
 * *out receives its result (one query record, index 0; free it with
 * dmr_result_destroy) and *latency_ms (may be NULL) the time from submit to
 * its result. DMR_E_TIMEOUT leaves the request pending; any other return
 * collects the ticket. */
int dmr_server_wait(dmr_server* s, uint64_t ticket, int64_t timeout_ms, dmr_result** out, double* latency_ms);
/* requests done, batches, mean and largest batch, peak admitted bytes,
 * longest queue, mean queue wait (ms), budget (8 doubles). */
int dmr_server_stats(dmr_server* s, double* out8);
/* Finishes the queued and running requests, then stops; results not
 * collected are freed. */
void dmr_server_destroy(dmr_server* s);

/* ---- FASTA ingress (host) -------------------------------------------------
 * io::parse_fasta (src/io.rs:40-102) of a FASTA text held in memory: one
 * record per '>' line, the header from that line (lossy UTF-8, trimmed),
 * Record::id (its first ASCII-whitespace token, "seq<len>" when empty), and
 * the residues of the lines up to the next header, upper-cased, with ASCII
 * whitespace, '-' and '.' dropped. Lines before the first header may be
 * empty or start with ';' or '#'. Fails with DMR_E_ARG and the reference's
 * message for sequence data before the first header and for non-empty text
 * without a header. Needs no context or device. */
typedef struct dmr_fasta dmr_fasta;
int dmr_fasta_parse(const char* data, uint64_t len, dmr_fasta** out);
/* Record count, residue count, byte lengths of the ids and headers texts,
 * and io::looks_like_dna (src/io.rs:104-121). Any may be NULL. */
int dmr_fasta_sizes(const dmr_fasta* f, uint64_t* nrec, uint64_t* nres, uint64_t* ids_len,
                    uint64_t* headers_len, int* looks_like_dna);
/* Residues as parsed (raw: the blastx reads of dmr_blastx), encoded with
 * alphabet::encode (letters: src/alphabet.rs:54-63, the blastp queries of
 * dmr_blastp and the database of dmr_db_load), and nrec + 1 offsets into
 * either. Buffers hold nres bytes and nrec + 1 offsets; any may be NULL. */
int dmr_fasta_residues(const dmr_fasta* f, char* raw, uint8_t* letters, uint64_t* offsets);
/* The ids and headers, each followed by '\n' (the query_ids and subject_ids
 * of dmr_result_format), not NUL-terminated; either may be NULL. */
int dmr_fasta_text(const dmr_fasta* f, char* ids, char* headers);
void dmr_fasta_destroy(dmr_fasta* f);

#ifdef __cplusplus
}
#endif

#endif /* DMR_H */
