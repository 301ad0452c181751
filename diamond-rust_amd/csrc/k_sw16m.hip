// Packed score pass of merged bands (classes kExactClasses..kPackedClasses-1):
// split kernels with a wall at each DP's own width, one per width bound and
// lane-group size of DMR_SW16_MERGED (kernels in sw16_kernels.hpp). A
// separate translation unit from the band_for widths, so the two compile in
// parallel.
#include "sw16_kernels.hpp"

namespace dmr {

bool launch_sw16_merged(int cls, int lanes_per_pair, int threads, bool f16, const uint2* lanes, uint32_t nlanes,
                        uint32_t nrow, const SwJob* jobs, SeqSet q, SeqSet db, const int* pair_tab, uint32_t* counter, int32_t* scores,
                        uint32_t* redo, uint32_t* nredo, int cu_count, hipStream_t st) {
#define DMR_SW16_MERGED_CASE(c, w, p, t)                                                                      \
  if (cls == c && lanes_per_pair == p && threads == t) {                                                     \
    if (f16)                                                                                                 \
      launch_split<w, p, t, F16, true>(lanes, nlanes, nrow, jobs, q, db, pair_tab, counter, scores, redo, nredo,     \
                                       cu_count, st);                                                        \
    else                                                                                                     \
      launch_split<w, p, t, I16, true>(lanes, nlanes, nrow, jobs, q, db, pair_tab, counter, scores, redo, nredo,     \
                                       cu_count, st);                                                        \
    return true;                                                                                             \
  }
  DMR_SW16_MERGED(DMR_SW16_MERGED_CASE)
#undef DMR_SW16_MERGED_CASE
  return false;
}

}  // namespace dmr
