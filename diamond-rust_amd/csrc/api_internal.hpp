// Types and helpers the C ABI's source files share (api.cpp, serve.cpp):
// the context and result handles and the status mapping.
#pragma once
#include "../../include/dmr.h"

#include <memory>
#include <string>
#include <vector>

#include "engine.hpp"

struct dmr_ctx {
  std::unique_ptr<dmr::Engine> eng;
  dmr::TargetSet loaded;  // resident query batch (dmr_queries_load)
  uint32_t loaded_nq = 0;
  bool has_loaded = false;
  // Ids and headers of the resident database, one per line
  // (dmr_db_load_file only); db_header_nl: some header holds a newline.
  std::string db_ids, db_headers;
  bool db_has_ids = false, db_header_nl = false;
};

struct dmr_result {
  std::vector<dmr::HspRec> hsps;     // output order
  std::vector<std::shared_ptr<const dmr::OpsStore>> ops_stores;  // where the HSPs' ops live
  std::vector<int> target_of;        // target index of each HSP
  dmr::TargetSet targets;
  std::vector<uint64_t> dna_len;     // per query record (blastx)
  bool translated = false;
  std::shared_ptr<const std::vector<uint32_t>> db_offs;  // host copy of the database
  std::shared_ptr<const std::vector<uint8_t>> db_res;
  dmr::StageStats stats;
  // Text of the last dmr_result_format call and the inputs it was made from,
  // so the usual size-then-fill call pair renders once.
  // Last rendered text, as one block of lines per host thread, and the
  // inputs it was rendered from (id lists and fields, compared on the next call).
  mutable std::string fmt_q, fmt_s, fmt_qt, fmt_st, fmt_f;
  mutable std::vector<std::string> fmt_parts;
  mutable uint64_t fmt_len = 0;
  mutable bool fmt_valid = false;
};

namespace dmr_api {

// Sets the calling thread's last error message and returns code.
int fail(int code, const std::string& msg);

// Runs f and maps exceptions to status codes (no exception crosses the ABI).
template <class F>
int guard(F&& f) {
  try {
    return f();
  } catch (const dmr::Error& e) {
    return fail(e.code, e.what());
  } catch (const std::bad_alloc&) {
    return fail(DMR_E_OOM, "host allocation failed");
  } catch (const std::exception& e) {
    return fail(DMR_E_DEVICE, e.what());
  }
}

// Throws DMR_E_ARG when a letter code is outside the 25-letter alphabet.
void check_letters(const uint8_t* p, uint64_t n, const char* what);
// search options of a config (NULL = the reference defaults).
dmr::SearchOptions options(const dmr_search_config* c);
// HSPs per target into r (merge_frames, targets, database copies).
void assemble(dmr_ctx* ctx, dmr_result* r, std::vector<std::vector<dmr::HspRec>>& per_target,
              const dmr::SearchOptions& o, size_t nqueries);

}  // namespace dmr_api
