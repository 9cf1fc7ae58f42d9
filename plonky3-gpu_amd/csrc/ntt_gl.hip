// The Goldilocks NTT / coset LDE kernels and their launchers (ntt_impl.h).
#include "ntt_impl.h"

namespace p3h {

int ntt_prepare_gl() { return ntt_prepare_field<Goldilocks>(); }

template int butterfly_probe_device<Goldilocks>(uint64_t*, const uint64_t*, uint32_t, uint32_t, hipStream_t);
template std::vector<LdePassGroup<uint64_t>> plan_lde_batch<Goldilocks>(const std::vector<LdeJob<uint64_t>>&);
template int launch_lde_group<Goldilocks>(const LdePassGroup<uint64_t>&, const NttWork<uint64_t>*, hipStream_t);
template int lde_device<Goldilocks>(const uint64_t*, uint64_t*, uint32_t, uint32_t, uint32_t, const uint64_t*,
                                    const uint64_t*, const NttTuning&, hipStream_t);
template int dft_device<Goldilocks>(int, const uint64_t*, uint64_t*, uint64_t*, uint32_t, uint32_t,
                                    const uint64_t*, const uint64_t*, const NttTuning&, hipStream_t);
template int powers_device<Goldilocks>(uint64_t*, uint64_t, uint64_t, uint64_t, uint64_t, hipStream_t);
template int lde_twist_device<Goldilocks>(uint64_t*, uint32_t, uint32_t, const uint64_t*, uint64_t, uint64_t,
                                          hipStream_t);
template int coset_to_natural_device<uint64_t>(const uint64_t*, uint64_t*, uint32_t, uint32_t, uint32_t,
                                               hipStream_t);

}  // namespace p3h
