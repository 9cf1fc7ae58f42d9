// The BabyBear NTT / coset LDE kernels and their launchers (ntt_impl.h), and
// the once-per-device set-up of every pass kernel of both fields.
#include "ntt_impl.h"

namespace p3h {

int ntt_prepare_device() {
    int e = ntt_prepare_field<BabyBear>();
    if (!e) e = ntt_prepare_gl();
    return e ? e : ntt_prepare_queue_bb();
}

template int butterfly_probe_device<BabyBear>(uint32_t*, const uint32_t*, uint32_t, uint32_t, hipStream_t);
template std::vector<LdePassGroup<uint32_t>> plan_lde_batch<BabyBear>(const std::vector<LdeJob<uint32_t>>&);
template int launch_lde_group<BabyBear>(const LdePassGroup<uint32_t>&, const NttWork<uint32_t>*, hipStream_t);
template int lde_device<BabyBear>(const uint32_t*, uint32_t*, uint32_t, uint32_t, uint32_t, const uint32_t*,
                                  const uint32_t*, const NttTuning&, hipStream_t);
template int dft_device<BabyBear>(int, const uint32_t*, uint32_t*, uint32_t*, uint32_t, uint32_t,
                                  const uint32_t*, const uint32_t*, const NttTuning&, hipStream_t);
template int powers_device<BabyBear>(uint32_t*, uint64_t, uint32_t, uint32_t, uint32_t, hipStream_t);
template int lde_twist_device<BabyBear>(uint32_t*, uint32_t, uint32_t, const uint32_t*, uint32_t, uint32_t,
                                        hipStream_t);
template int coset_to_natural_device<uint32_t>(const uint32_t*, uint32_t*, uint32_t, uint32_t, uint32_t,
                                               hipStream_t);

}  // namespace p3h
