// The queue launch of a batch's BabyBear LDE passes (ntt_impl.h),
// in a translation unit of its own: its one kernel carries 38 pass variants.
#include "ntt_impl.h"

namespace p3h {

int ntt_prepare_queue_bb() { return ntt_prepare_queue<BabyBear>(); }

template bool plan_lde_queue<BabyBear>(const std::vector<LdeJob<uint32_t>>&, LdeQueuePlan<uint32_t>&);
template int launch_lde_queue<BabyBear>(const LdeQueuePlan<uint32_t>&, const NttQueueItem<uint32_t>*, uint32_t*,
                                        hipStream_t);

}  // namespace p3h
