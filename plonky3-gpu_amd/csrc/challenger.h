// Device-resident duplex-sponge transcript (BabyBear, Poseidon2 width 16,
// rate 8) and the proof-of-work search, for the C ABI's p3h_challenger_*.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "poseidon2_f64.h"

namespace p3h {

// Transcript state, canonical words.  The input buffer holds in_len pending
// observations; output[0..out_len) are the unconsumed samples (popped from
// the back).
struct DevChal {
    uint32_t sponge[16];
    uint32_t input[8];
    uint32_t output[8];
    uint32_t in_len, out_len;
    uint32_t pad[6];
};

enum ChalCode : uint32_t {
    CH_OBSERVE = 0,      // observe n words from src (device memory)
    CH_OBSERVE_IMM = 1,  // observe the word arg
    CH_WITNESS = 2,      // observe_witness(arg): its low and high 32-bit halves
    CH_SAMPLE = 3        // n samples to dst (device memory)
};

struct ChalOp {
    uint32_t code;
    uint32_t n;
    uint64_t arg;
    const uint32_t* src;
    uint32_t* dst;
};

constexpr uint32_t CHAL_MAX_OPS = 12;
struct ChalProgram {
    ChalOp op[CHAL_MAX_OPS];
    uint32_t n_ops;
};

// Run a program of transcript operations on *st (one wavefront).
int chal_run_device(DevChal* st, const ChalProgram& prog, const P2ParamsF64& p, hipStream_t s);

// For witnesses w in [base, base + count): the sample that follows
// observe_witness(w) on a copy of *st; the least w whose sample has `bits`
// zero low bits is folded into *found with an atomic minimum.  Workgroups take
// consecutive chunks in increasing order from the counter *next (zero before
// the launch) and stop taking them once *found lies below the next chunk, so
// the search ends soon after the least witness instead of covering the whole
// range; *found (all ones before the launch) is final when the grid drains.
int grind_search_device(const DevChal* st, uint64_t base, uint64_t count, uint32_t bits,
                        unsigned long long* found, unsigned long long* next, const P2ParamsF64& p,
                        hipStream_t s);

}  // namespace p3h
