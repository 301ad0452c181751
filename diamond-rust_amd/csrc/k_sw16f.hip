// Fused packed score pass (k_sw16p_fused in sw16_kernels.hpp): the six
// merged-band classes and the band_for widths 81 and 61 in one persistent
// launch. A translation unit of its own, so it compiles beside the others.
#include "sw16_kernels.hpp"

namespace dmr {

int sw16_fused_class(int kind) {
  switch (kind) {
#define DMR_SW16_FUSED_CLASS(i, c, w, p, wall) \
  case i:                                     \
    return c;
    DMR_SW16_FUSED(DMR_SW16_FUSED_CLASS)
#undef DMR_SW16_FUSED_CLASS
    default:
      return -1;
  }
}

bool launch_sw16_fused(bool f16, FusedSections s, const SwJob* jobs, SeqSet q, SeqSet db, const int* pair_tab,
                       uint32_t* counter, int32_t* scores, uint32_t* redo, uint32_t* nredo, int cu_count,
                       hipStream_t st) {
  static const int kLanesPer[kFusedKinds] = {
#define DMR_SW16_FUSED_P(i, c, w, p, wall) p,
      DMR_SW16_FUSED(DMR_SW16_FUSED_P)
#undef DMR_SW16_FUSED_P
  };
  uint32_t t = 0;
  for (int k = 0; k < kFusedKinds; ++k) {
    if (s.nrow[k] > s.nlanes[k]) return false;
    const uint32_t per = 64 / kLanesPer[k];
    s.task0[k] = t;
    t += (s.nrow[k] + per - 1) / per + (s.nlanes[k] - s.nrow[k] + per - 1) / per;
  }
  s.ntasks = t;
  if (!t) return true;
  if (f16) {
    static const unsigned resident = resident_blocks(k_sw16p_fused<F16>, 512, cu_count);
    const unsigned blocks = std::min<unsigned>(resident, (t + 7) / 8);
    hipLaunchKernelGGL(k_sw16p_fused<F16>, dim3(blocks), dim3(512), 0, st, s, jobs, q, db, pair_tab, counter,
                       scores, redo, nredo);
  } else {
    static const unsigned resident = resident_blocks(k_sw16p_fused<I16>, 512, cu_count);
    const unsigned blocks = std::min<unsigned>(resident, (t + 7) / 8);
    hipLaunchKernelGGL(k_sw16p_fused<I16>, dim3(blocks), dim3(512), 0, st, s, jobs, q, db, pair_tab, counter,
                       scores, redo, nredo);
  }
  return true;
}

}  // namespace dmr
