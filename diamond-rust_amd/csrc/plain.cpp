// The plain drop-in form: search::run (src/search.rs:95-271) kept on the
// host as the reference structures it, with every seed::build,
// seed::join_into, ungapped::xdrop and gapped::extend_span call going to the
// GPU through the drop-in entry points — one kernel launch per call.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <map>

#include "engine.hpp"

namespace dmr {

namespace {

struct Hit {
  uint32_t qpos, spos;
};

bool hsp_less(const HspRec& a, const HspRec& b) {
  if (a.score != b.score) return a.score > b.score;
  if (a.evalue < b.evalue) return true;
  if (b.evalue < a.evalue) return false;
  return a.subject < b.subject;
}

}  // namespace

std::vector<std::vector<HspRec>> search_plain(Engine& eng, const TargetSet& t,
                                              const SearchOptions& opt, StageStats* stats,
                                              std::shared_ptr<const OpsStore>* ops_store) {
  auto t0 = std::chrono::steady_clock::now();
  auto store = std::make_shared<OpsStore>();
  *ops_store = store;
  const uint32_t nt = (uint32_t)t.offs.size() - 1;
  std::vector<std::vector<HspRec>> out(nt);
  const uint32_t ndb = eng.db_size();
  if (nt == 0 || ndb == 0) return out;
  const auto& dbo = eng.db_offsets_host();
  const auto& dbl = eng.db_letters_host();
  const AlpParams P = eng.alp(opt.gap_open, opt.gap_extend);
  const auto& codes = shape_codes(opt.sensitivity);
  ReductionDef red = parse_reduction(reduction_def(opt.sensitivity));
  std::vector<ShapeDef> shapes(codes.size());
  for (size_t i = 0; i < codes.size(); ++i) parse_shape(codes[i], &shapes[i]);
  double sc = opt.seed_cut >= 0 ? opt.seed_cut : preset_seed_cut(opt.sensitivity);
  double cut = sc * 0.6931471805599453 * (double)shapes[0].weight;
  int nch = opt.index_chunks >= 0 ? std::max(opt.index_chunks, 1) : preset_index_chunks(opt.sensitivity);

  // Query seeds once per shape (src/search.rs:109-119).
  std::vector<uint64_t> qoff(t.offs.begin(), t.offs.end());
  std::vector<std::vector<uint64_t>> qk(shapes.size());
  std::vector<std::vector<uint32_t>> qs(shapes.size()), qp(shapes.size());
  for (size_t s = 0; s < shapes.size(); ++s)
    eng.seed_build(t.letters.data(), qoff.data(), nt, shapes[s], red, 0, cut, &qk[s], &qs[s], &qp[s]);

  // Chunk x shape: reference seeds and join (src/search.rs:121-133).
  std::vector<std::map<uint32_t, std::vector<Hit>>> hits(nt);
  std::vector<std::pair<uint32_t, uint32_t>> ranges;
  {
    uint32_t parts = std::min<uint32_t>((uint32_t)std::max(nch, 1), ndb), start = 0;
    for (uint32_t p = 0; p < parts; ++p) {
      uint32_t end = (p + 1 == parts) ? ndb : start + ndb / parts + (p < ndb % parts ? 1 : 0);
      ranges.push_back({start, end});
      start = end;
    }
  }
  uint64_t nhits = 0;
  for (auto [c0, c1] : ranges) {
    std::vector<uint64_t> ro(c1 - c0 + 1);
    for (uint32_t i = c0; i <= c1; ++i) ro[i - c0] = dbo[i] - dbo[c0];
    for (size_t s = 0; s < shapes.size(); ++s) {
      std::vector<uint64_t> rk;
      std::vector<uint32_t> rs, rp, h;
      eng.seed_build(dbl.data() + dbo[c0], ro.data(), c1 - c0, shapes[s], red, c0, cut, &rk, &rs, &rp);
      eng.seed_join(qk[s].data(), qs[s].data(), qp[s].data(), qk[s].size(), rk.data(), rs.data(),
                    rp.data(), rk.size(), opt.max_expand, &h);
      for (size_t k = 0; k < h.size() / 4; ++k) hits[h[4 * k]][h[4 * k + 2]].push_back({h[4 * k + 1], h[4 * k + 3]});
      nhits += h.size() / 4;
    }
  }

  const int xdrop = raw_score(P, 12.3);
  const int cutoff = 19;
  uint64_t njobs = 0, nseg = 0;
  // align_target per target (src/search.rs:148-271).
  for (uint32_t ti = 0; ti < nt; ++ti) {
    const uint8_t* q = t.letters.data() + t.offs[ti];
    const uint64_t qlen = t.offs[ti + 1] - t.offs[ti];
    if (qlen == 0 || hits[ti].empty()) continue;
    std::vector<std::vector<HspRec>> found;
    for (auto& [sid, hv] : hits[ti]) {  // subjects ascending
      const uint8_t* s = dbl.data() + dbo[sid];
      const uint64_t slen = dbo[sid + 1] - dbo[sid];
      if (slen == 0) continue;
      std::sort(hv.begin(), hv.end(), [](const Hit& a, const Hit& b) {
        int64_t da = (int64_t)a.qpos - a.spos, db = (int64_t)b.qpos - b.spos;
        return da != db ? da < db : a.spos < b.spos;
      });
      std::vector<SegRec> segs;
      for (const Hit& h : hv) {
        int64_t diag = (int64_t)h.qpos - h.spos;
        if (!segs.empty()) {
          const SegRec& l = segs.back();
          if ((int64_t)l.q_start - l.s_start == diag && (int64_t)l.s_start + l.len >= h.spos) continue;
        }
        SegRec g;
        eng.xdrop_batch(q, qlen, s, slen, &h.qpos, &h.spos, 1, xdrop, &g);
        if (g.score >= cutoff) segs.push_back(g);
      }
      if (segs.empty()) continue;
      nseg += segs.size();
      const int64_t band = band_for(qlen);
      std::vector<size_t> order(segs.size());
      for (size_t i = 0; i < order.size(); ++i) order[i] = i;
      std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
        int64_t da = (int64_t)segs[a].q_start - segs[a].s_start, db = (int64_t)segs[b].q_start - segs[b].s_start;
        return da < db;
      });
      struct Band { int64_t b0, b1; size_t best; };
      std::vector<Band> bands;
      for (size_t idx : order) {
        int64_t d = (int64_t)segs[idx].q_start - segs[idx].s_start;
        if (!bands.empty() && d - band <= bands.back().b1) {
          bands.back().b1 = std::max(bands.back().b1, d + band);
          if (segs[idx].score > segs[bands.back().best].score) bands.back().best = idx;
        } else {
          bands.push_back({d - band, d + band, idx});
        }
      }
      std::stable_sort(bands.begin(), bands.end(),
                       [&](const Band& a, const Band& b) { return segs[a.best].score > segs[b.best].score; });
      std::vector<HspRec> hsps;
      for (const Band& b : bands) {
        if (hsps.size() >= opt.max_hsps) break;
        // gapped_hsp window clipping (src/search.rs:298-312).
        uint64_t row_cells = (uint64_t)std::max<int64_t>(b.b1 - b.b0, 0) + 1;
        uint64_t max_rows = std::max<uint64_t>(16000000 / row_cells, 2 * band + 16);
        uint64_t q0 = 0, q1 = qlen;
        if (q1 - q0 > max_rows) {
          uint64_t a = segs[b.best].q_start;
          q0 = a > max_rows / 2 ? a - max_rows / 2 : 0;
          q1 = std::min(q0 + max_rows, qlen);
          q0 = q1 > max_rows ? q1 - max_rows : 0;
        }
        Engine::SpanJob j{0, q1 - q0, 0, slen, b.b0 - (int64_t)q0, b.b1 - (int64_t)q0};
        std::vector<SwOut> o;
        std::vector<std::vector<uint32_t>> ops;
        eng.extend_span_batch(q + q0, q1 - q0, s, slen, &j, 1, opt.gap_open, opt.gap_extend, true, &o, &ops);
        ++njobs;
        if (o[0].score <= 0) continue;
        double e = alp_evalue(P, o[0].score, qlen, slen, eng.db_letters());
        if (!(e <= opt.evalue)) continue;
        HspRec h;
        h.query = ti;
        h.subject = sid;
        h.frame = 0;
        h.q_begin = (uint32_t)(q0 + o[0].q_begin);
        h.q_end = (uint32_t)(q0 + o[0].best_i);
        h.s_begin = o[0].s_begin;
        h.s_end = o[0].best_j;
        h.score = o[0].score;
        h.bits = bit_score(P, o[0].score);
        h.evalue = e;
        h.identities = o[0].identities;
        h.positives = o[0].positives;
        h.mismatches = o[0].mismatches;
        h.gap_openings = o[0].gap_openings;
        h.gaps = o[0].gaps;
        h.length = o[0].length;
        store->parts.push_back(std::move(ops[0]));
        h.ops = OpsRef{store->parts.back().data(), (uint32_t)store->parts.back().size()};
        hsps.push_back(std::move(h));
      }
      if (!hsps.empty()) {
        std::stable_sort(hsps.begin(), hsps.end(), hsp_less);
        found.push_back(std::move(hsps));
      }
    }
    std::stable_sort(found.begin(), found.end(), [](const std::vector<HspRec>& a, const std::vector<HspRec>& b) {
      return hsp_less(a[0], b[0]);
    });
    if (found.size() > opt.max_target_seqs) found.resize(opt.max_target_seqs);
    for (auto& hs : found) {
      if (hs.size() > opt.max_hsps) hs.resize(opt.max_hsps);
      for (auto& h : hs) out[ti].push_back(std::move(h));
    }
  }
  if (stats) {
    *stats = StageStats{};
    stats->hits = nhits;
    stats->segments = nseg;
    stats->jobs = njobs;
    stats->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  return out;
}

}  // namespace dmr
