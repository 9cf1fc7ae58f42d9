// Host-side launcher for the FRI fold kernels (fri.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "field.h"

namespace p3h {

// Fold a BabyBear^4 vector of 2^log_h elements (canonical, 4 words each) by
// arity 2^log_arity into n_out = 2^(log_h - log_arity) elements.  beta is
// canonical (4 words), on the host, or -- when beta_dev is not null -- read
// from device memory by the kernel.  Arity 2 uses tinv_m (see fri.hip); other
// arities use omega_h_m = omega_{2^log_h}, w_m = omega_{arity}, arity_inv_m =
// 1/arity, all in Montgomery form.
int fold_device(const uint32_t* in, uint32_t* out, uint64_t n_out, uint32_t log_h, uint32_t log_arity,
                const uint32_t* beta, const uint32_t* beta_dev, const uint32_t* tinv_m, uint32_t omega_h_m,
                uint32_t w_m, uint32_t arity_inv_m, hipStream_t st);

// The allocation-free fold keeps its arity-2 table in a static device array
// (fri.hip) of fold_scratch_words() words per GPU.  fold_scratch_setup(device),
// called once per device with that device current, looks up the array's
// device address; fold_scratch_ptr(device) returns it (null before setup).
uint64_t fold_scratch_words();
int fold_scratch_setup(int device);
uint32_t* fold_scratch_ptr(int device);
// fold_device with the arity-2 table built in scratch_ptr (n_out words) first:
// tinv_base_m = omega_{2^log_h}^-1 and half_m = 1/2 (Montgomery); no device
// allocation.  Returns a HIP error code (hipErrorInvalidValue when n_out
// exceeds the scratch at arity 2).
int fold_noalloc_device(const uint32_t* in, uint32_t* out, uint64_t n_out, uint32_t log_h, uint32_t log_arity,
                        const uint32_t* beta, uint32_t* scratch_ptr, uint32_t tinv_base_m, uint32_t half_m,
                        uint32_t omega_h_m, uint32_t w_m, uint32_t arity_inv_m, hipStream_t st);

}  // namespace p3h
