// Serving loop over a resident database (C ABI dmr_server_*, include/dmr.h).
//
// Requests (one protein query each, with its own preset and options) are
// submitted at any time. A dispatcher thread forms batches of queued
// requests with the same options as workers come free and the device-memory
// budget allows, oldest request first; each worker searches its batch on an
// engine of its own (streams and scratch) against the database and seed
// indexes the context keeps in HBM, so the batches of different workers run
// on the device at the same time. The max_expand cap counts a key's query
// entries per request (one record) instead of over the batch, so a request's
// result is exactly the result of searching it alone (src/search.rs:109-131
// with a batch of one).
#include "api_internal.hpp"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>

namespace {

double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

struct ServeReq {
  uint64_t ticket = 0;
  std::vector<uint8_t> letters;
  dmr::SearchOptions opt;
  std::string key;  // requests with equal keys batch together
  uint64_t est = 0;  // admitted device bytes
  double t_submit = 0, t_start = 0, t_done = 0;
  std::unique_ptr<dmr_result> result;
  int status = 1;  // 1 pending, 0 done, < 0 failed
  std::string err;
};

struct Worker {
  std::unique_ptr<dmr::Engine> eng;
  std::thread th;
  std::vector<ServeReq*> batch;
  uint64_t batch_est = 0;
  bool busy = false;
  std::condition_variable cv;
};

// Options that change results or the index: requests batch only with equal
// ones.
std::string options_key(const dmr::SearchOptions& o) {
  char b[256];
  std::snprintf(b, sizeof b, "%s|%.17g|%u|%u|%d|%.17g|%d|%d", o.sensitivity.c_str(), o.evalue, o.max_target_seqs,
                o.max_hsps, o.index_chunks, o.seed_cut, o.gap_open, o.gap_extend);
  return b;
}

}  // namespace

struct dmr_server {
  dmr_ctx* ctx = nullptr;
  uint64_t budget = 0;
  uint64_t slice_hits = 0;
  uint32_t max_batch = 1024;
  std::vector<std::unique_ptr<Worker>> workers;
  std::mutex m;
  std::condition_variable cv_disp, cv_done;
  std::deque<ServeReq*> queue;
  std::unordered_map<uint64_t, std::unique_ptr<ServeReq>> reqs;
  uint64_t next_ticket = 1, in_flight = 0;
  bool stop = false;
  std::thread dispatcher;
  // Seed hits per query letter seen so far, per options key (admission
  // estimates start from the preset's shapes and grow with what batches
  // show).
  std::map<std::string, double> hits_per_letter;
  // statistics
  uint64_t n_batches = 0, n_done = 0, sum_batch = 0, max_batch_seen = 0, peak_in_flight = 0, max_queue = 0;
  double sum_wait = 0;

  uint64_t estimate(const ServeReq& r) {
    const double nshapes = (double)dmr::shape_codes(r.opt.sensitivity).size();
    double hpl = 0;
    auto it = hits_per_letter.find(r.key);
    if (it != hits_per_letter.end()) hpl = it->second;
    else hpl = 64.0 * nshapes;
    const double L = (double)r.letters.size();
    // query seeds (key + letter, twice for the sort), hits (~52 bytes each
    // through the ungapped stage, ~150 per DP job at about one job in three
    // hits), at most one slice of hits in flight per batch.
    const double seeds = L * nshapes * 24.0;
    const double hits = std::min(L * hpl, (double)slice_hits);
    return (uint64_t)(seeds + hits * 110.0) + 65536;
  }

  void run_dispatcher();
  void run_worker(Worker* w);
};

void dmr_server::run_dispatcher() {
  std::unique_lock<std::mutex> g(m);
  for (;;) {
    Worker* idle = nullptr;
    cv_disp.wait(g, [&] {
      if (stop) return true;
      if (queue.empty()) return false;
      idle = nullptr;
      for (auto& w : workers)
        if (!w->busy) {
          idle = w.get();
          break;
        }
      if (!idle) return false;
      // Admission: the oldest request must fit the memory left, unless
      // nothing is in flight (a request larger than the budget runs alone).
      return in_flight == 0 || in_flight + queue.front()->est <= budget;
    });
    if (stop) return;
    // The oldest request's options decide the batch; requests with the same
    // options join it in arrival order while the budget allows.
    const std::string key = queue.front()->key;
    std::vector<ServeReq*> batch;
    uint64_t est = 0;
    for (auto it = queue.begin(); it != queue.end() && batch.size() < max_batch;) {
      ServeReq* r = *it;
      if (r->key != key) {
        ++it;
        continue;
      }
      if (!batch.empty() && in_flight + est + r->est > budget) break;
      est += r->est;
      batch.push_back(r);
      it = queue.erase(it);
    }
    in_flight += est;
    peak_in_flight = std::max(peak_in_flight, in_flight);
    const double t = now_ms();
    for (ServeReq* r : batch) {
      r->t_start = t;
      sum_wait += t - r->t_submit;
    }
    ++n_batches;
    sum_batch += batch.size();
    max_batch_seen = std::max<uint64_t>(max_batch_seen, batch.size());
    idle->batch = std::move(batch);
    idle->batch_est = est;
    idle->busy = true;
    idle->cv.notify_one();
  }
}

void dmr_server::run_worker(Worker* w) {
  for (;;) {
    std::vector<ServeReq*> batch;
    {
      std::unique_lock<std::mutex> g(m);
      w->cv.wait(g, [&] { return stop || !w->batch.empty(); });
      if (w->batch.empty()) return;  // stopping with nothing assigned
      batch = w->batch;
    }
    std::vector<std::unique_ptr<dmr_result>> results(batch.size());
    int status = DMR_OK;
    std::string err;
    uint64_t hits = 0, letters = 0;
    try {
      dmr::TargetSet t;
      t.offs.push_back(0);
      for (uint32_t i = 0; i < batch.size(); ++i) {
        t.letters.insert(t.letters.end(), batch[i]->letters.begin(), batch[i]->letters.end());
        t.offs.push_back((uint32_t)t.letters.size());
        t.query_of.push_back(i);
        t.frame_of.push_back(0);
      }
      letters = t.letters.size();
      dmr::SearchOptions opt = batch[0]->opt;
      opt.per_record = true;
      opt.slice_hits = slice_hits;
      w->eng->load_targets(t);
      dmr::StageStats st;
      std::vector<std::shared_ptr<const dmr::OpsStore>> stores;
      auto per_target = w->eng->search_targets(opt, &st, &stores);
      hits = st.hits;
      for (uint32_t i = 0; i < batch.size(); ++i) {
        auto r = std::make_unique<dmr_result>();
        r->targets.letters = batch[i]->letters;
        r->targets.offs = {0, (uint32_t)batch[i]->letters.size()};
        r->targets.query_of = {0};
        r->targets.frame_of = {0};
        r->dna_len = {0};
        r->ops_stores = stores;
        r->stats = st;
        std::vector<std::vector<dmr::HspRec>> mine(1);
        mine[0] = std::move(per_target[i]);
        // A request's HSPs keep query 0 (its only record).
        for (auto& h : mine[0]) h.query = 0;
        dmr_api::assemble(ctx, r.get(), mine, opt, 1);
        results[i] = std::move(r);
      }
    } catch (const dmr::Error& e) {
      status = e.code;
      err = e.what();
    } catch (const std::exception& e) {
      status = DMR_E_DEVICE;
      err = e.what();
    }
    {
      std::lock_guard<std::mutex> g(m);
      const double t = now_ms();
      for (uint32_t i = 0; i < batch.size(); ++i) {
        ServeReq* r = batch[i];
        r->t_done = t;
        r->status = status;
        r->err = err;
        r->result = std::move(results[i]);
      }
      n_done += batch.size();
      if (status == DMR_OK && letters) {
        double& h = hits_per_letter[batch[0]->key];
        h = std::max((double)hits / (double)letters, 0.5 * h + 0.5 * (double)hits / (double)letters);
      }
      in_flight -= w->batch_est;
      w->batch.clear();
      w->batch_est = 0;
      w->busy = false;
    }
    cv_done.notify_all();
    cv_disp.notify_one();
  }
}

using namespace dmr_api;

extern "C" {

int dmr_server_create(dmr_ctx* ctx, uint32_t workers, uint64_t hbm_budget, const char* presets,
                      dmr_server** out) {
  if (!ctx || !out) return fail(DMR_E_ARG, "NULL argument");
  if (workers == 0 || workers > 8) return fail(DMR_E_ARG, "workers must be 1..8");
  if (hbm_budget < (64ull << 20)) return fail(DMR_E_ARG, "device-memory budget below 64 MiB");
  if (ctx->eng->db_size() == 0) return fail(DMR_E_ARG, "no resident database");
  return guard([&]() -> int {
    auto s = std::make_unique<dmr_server>();
    s->ctx = ctx;
    s->budget = hbm_budget;
    // Each worker's batch holds at most this many hits at a time (the
    // engine slices larger batches).
    s->slice_hits = std::max<uint64_t>(1 << 16, hbm_budget / workers / 110);
    // The indexes of the listed presets are built now, so no request waits
    // for one.
    if (presets && *presets) {
      std::string p = presets;
      size_t a = 0;
      while (a <= p.size()) {
        size_t b = p.find(',', a);
        if (b == std::string::npos) b = p.size();
        const std::string name = p.substr(a, b - a);
        if (!name.empty()) ctx->eng->build_index(name, -1, -1.0);
        a = b + 1;
      }
    }
    for (uint32_t k = 0; k < workers; ++k) {
      auto w = std::make_unique<Worker>();
      w->eng = std::make_unique<dmr::Engine>(ctx->eng->device());
      w->eng->share_db(*ctx->eng);
      w->eng->copy_tuning(*ctx->eng);
      s->workers.push_back(std::move(w));
    }
    dmr_server* sp = s.get();
    for (auto& w : s->workers) {
      Worker* wp = w.get();
      w->th = std::thread([sp, wp] { sp->run_worker(wp); });
    }
    s->dispatcher = std::thread([sp] { sp->run_dispatcher(); });
    *out = s.release();
    return DMR_OK;
  });
}

int dmr_server_submit(dmr_server* s, const uint8_t* query, uint64_t len, const dmr_search_config* cfg,
                      uint64_t* ticket) {
  if (!s || !ticket || (len && !query)) return fail(DMR_E_ARG, "NULL argument");
  if (len == 0) return fail(DMR_E_ARG, "empty query");
  dmr::SearchOptions o = options(cfg);
  dmr::GapRow gr;
  if (!dmr::blosum62_gap_row(o.gap_open, o.gap_extend, &gr))
    return fail(DMR_E_ARG, "Gap penalty settings are outside the supported range for BLOSUM62");
  if (len >= (1ull << 31)) return fail(DMR_E_UNSUPPORTED, "query longer than 2^31 letters");
  return guard([&]() -> int {
    check_letters(query, len, "query");
    auto r = std::make_unique<ServeReq>();
    r->letters.assign(query, query + len);
    r->opt = o;
    r->key = options_key(o);
    r->t_submit = now_ms();
    {
      std::lock_guard<std::mutex> g(s->m);
      r->ticket = s->next_ticket++;
      r->est = s->estimate(*r);
      *ticket = r->ticket;
      s->queue.push_back(r.get());
      s->max_queue = std::max<uint64_t>(s->max_queue, s->queue.size());
      s->reqs.emplace(r->ticket, std::move(r));
    }
    s->cv_disp.notify_one();
    return DMR_OK;
  });
}

int dmr_server_wait(dmr_server* s, uint64_t ticket, int64_t timeout_ms, dmr_result** out, double* latency_ms) {
  if (!s || !out) return fail(DMR_E_ARG, "NULL argument");
  std::unique_lock<std::mutex> g(s->m);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return fail(DMR_E_ARG, "unknown or already collected ticket");
  ServeReq* r = it->second.get();
  auto done = [&] { return r->status != 1; };
  if (timeout_ms < 0) s->cv_done.wait(g, done);
  else if (!s->cv_done.wait_for(g, std::chrono::milliseconds(timeout_ms), done))
    return fail(DMR_E_TIMEOUT, "request not finished within the timeout");
  if (latency_ms) *latency_ms = r->t_done - r->t_submit;
  std::unique_ptr<ServeReq> own = std::move(it->second);
  s->reqs.erase(it);
  if (own->status != DMR_OK) return fail(own->status, own->err);
  *out = own->result.release();
  return DMR_OK;
}

int dmr_server_stats(dmr_server* s, double* out8) {
  if (!s || !out8) return fail(DMR_E_ARG, "NULL argument");
  std::lock_guard<std::mutex> g(s->m);
  const double v[8] = {(double)s->n_done,
                       (double)s->n_batches,
                       s->n_batches ? (double)s->sum_batch / (double)s->n_batches : 0.0,
                       (double)s->max_batch_seen,
                       (double)s->peak_in_flight,
                       (double)s->max_queue,
                       s->sum_batch ? s->sum_wait / (double)s->sum_batch : 0.0,
                       (double)s->budget};
  std::memcpy(out8, v, sizeof v);
  return DMR_OK;
}

void dmr_server_destroy(dmr_server* s) {
  if (!s) return;
  {
    std::unique_lock<std::mutex> g(s->m);
    // Requests already queued or running finish first.
    s->cv_done.wait(g, [&] {
      if (!s->queue.empty()) return false;
      for (auto& w : s->workers)
        if (w->busy) return false;
      return true;
    });
    s->stop = true;
  }
  s->cv_disp.notify_all();
  for (auto& w : s->workers) w->cv.notify_all();
  if (s->dispatcher.joinable()) s->dispatcher.join();
  for (auto& w : s->workers)
    if (w->th.joinable()) w->th.join();
  delete s;
}

}  // extern "C"
