// Duplex-sponge transcript on the device (reference: DuplexChallenger,
// duplex_challenger.hpp:42-268) and the grinding search.
//
// The transcript runs in one wavefront: the sponge state is spread over a
// 16-lane row (Poseidon2F64Lanes), the input and output buffers sit in LDS,
// and the buffer lengths are wave-uniform.  Semantics:
//   observe(v): output buffer cleared; v appended to the input buffer; a full
//               buffer (8) triggers a duplexing;
//   duplexing:  state[0..len) = inputs, inputs cleared, permute, outputs =
//               state[0..8);
//   sample():   duplexing if inputs are pending or no output is left; pop the
//               last output.
//   observe_witness(w): output cleared; absorb w mod 2^32 and w >> 32 (each
//               reduced mod p).

#include "challenger.h"

namespace p3h {

namespace {

constexpr uint32_t BB_PRIME = 0x78000001u;

static_assert(sizeof(ChalProgram) + sizeof(P2ParamsF64) + sizeof(void*) <= 4096,
              "transcript kernel arguments past 4 KiB");

__global__ __launch_bounds__(64) void chal_run_kernel(DevChal* st, ChalProgram prog, P2ParamsF64 p) {
    __shared__ uint32_t inb[8], outb[8];
    const uint32_t lane = threadIdx.x;
    const auto pc = Poseidon2F64Lanes::consts(p, lane & 15);
    double s = lane < 16 ? double(st->sponge[lane]) : 0.0;
    if (lane < 8) {
        inb[lane] = st->input[lane];
        outb[lane] = st->output[lane];
    }
    uint32_t in_len = st->in_len, out_len = st->out_len;
    __syncthreads();

    auto duplex = [&]() {
        if (lane < in_len) s = double(inb[lane]);
        s = Poseidon2F64Lanes::permute(s, pc, lane & 15);
        const uint32_t v = BBF64::canonical(s);
        s = double(v);
        if (lane < 8) outb[lane] = v;
        in_len = 0;
        out_len = 8;
        __syncthreads();
    };
    auto absorb = [&](uint32_t v) {
        if (lane == 0) inb[in_len] = v % BB_PRIME;
        __syncthreads();
        if (++in_len == 8) duplex();
    };

    for (uint32_t i = 0; i < prog.n_ops; ++i) {
        const ChalOp op = prog.op[i];
        if (op.code == CH_OBSERVE) {
            out_len = 0;
            for (uint32_t j = 0; j < op.n; ++j) absorb(op.src[j]);
        } else if (op.code == CH_OBSERVE_IMM) {
            out_len = 0;
            absorb(uint32_t(op.arg));
        } else if (op.code == CH_WITNESS) {
            out_len = 0;
            absorb(uint32_t(op.arg & 0xFFFFFFFFu));
            absorb(uint32_t(op.arg >> 32));
        } else if (op.code == CH_SAMPLE) {
            for (uint32_t j = 0; j < op.n; ++j) {
                if (in_len > 0 || out_len == 0) duplex();
                --out_len;
                if (lane == 0) op.dst[j] = outb[out_len];
            }
        }
    }
    // write the state back: lane i < 16 its sponge word, lanes < 8 the buffers
    if (lane < 16) st->sponge[lane] = BBF64::canonical(s);
    if (lane < 8) {
        st->input[lane] = inb[lane];
        st->output[lane] = outb[lane];
    }
    if (lane == 0) {
        st->in_len = in_len;
        st->out_len = out_len;
    }
}

// One thread per witness.  With l pending inputs, observe_witness + sample is
// one permutation of [inputs, lo, hi, state[l+2..16)] when l <= 6 (for l = 6
// the second word fills the buffer and the sample is taken from that
// duplexing's outputs), and for l = 7 two: [inputs, lo, state[8..]] and then
// [hi, state[1..]].  The sample is word 7 of the last permutation's output.
constexpr uint32_t GRIND_THREADS = 256;  // one witness per thread per chunk

__device__ __forceinline__ uint32_t grind_sample(const DevChal* st, uint64_t w, const P2ParamsF64& p) {
    const uint32_t lo = uint32_t(w & 0xFFFFFFFFu) % BB_PRIME, hi = uint32_t(w >> 32) % BB_PRIME;
    const uint32_t ell = st->in_len;
    double s[16];
#pragma unroll
    for (uint32_t i = 0; i < 16; ++i) {
        const uint32_t inp = i < 8 ? st->input[i] : 0u;
        const uint32_t v = i < ell ? inp : (i == ell ? lo : (i == ell + 1 && ell < 7 ? hi : st->sponge[i]));
        s[i] = double(v);
    }
    // the grid loops over witness chunks: constants re-read per permutation
    Poseidon2F64::permute<true, true>(s, p);
    if (ell == 7) {
        s[0] = double(hi);
        Poseidon2F64::permute<true, true>(s, p);
    }
    return BBF64::canonical(s[7]);
}

// Work queue over chunks of GRIND_THREADS witnesses.  Every workgroup leaves
// the loop once the counter passes the end of the range or the least passing
// witness found so far lies below its next chunk; chunks below that witness
// were all taken before it and finish, so the grid drains with the minimum.
__global__ __launch_bounds__(GRIND_THREADS) void grind_search_kernel(const DevChal* st, uint64_t base, uint64_t count,
                                                                     uint32_t bits, unsigned long long* found,
                                                                     unsigned long long* next, P2ParamsF64 p) {
    __shared__ unsigned long long chunk;
    const uint32_t mask = bits >= 32 ? 0xFFFFFFFFu : ((1u << bits) - 1u);
    const uint64_t end = base + count;
    while (true) {
        if (threadIdx.x == 0) {
            const uint64_t c = atomicAdd(next, 1ull);
            const uint64_t start = base + c * GRIND_THREADS;
            const uint64_t f = __hip_atomic_load(found, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            chunk = (c >= (count + GRIND_THREADS - 1) / GRIND_THREADS || f < start) ? ~0ull : start;
        }
        __syncthreads();
        const uint64_t start = chunk;
        __syncthreads();  // every thread has read it before thread 0 writes the next
        if (start == ~0ull) break;
        const uint64_t w = start + threadIdx.x;
        if (w < end && (grind_sample(st, w, p) & mask) == 0) atomicMin(found, (unsigned long long)w);
    }
}

}  // namespace

int chal_run_device(DevChal* st, const ChalProgram& prog, const P2ParamsF64& p, hipStream_t s) {
    if (prog.n_ops > CHAL_MAX_OPS) return int(hipErrorInvalidValue);
    hipLaunchKernelGGL(chal_run_kernel, dim3(1), dim3(64), 0, s, st, prog, p);
    return int(hipGetLastError());
}

int grind_search_device(const DevChal* st, uint64_t base, uint64_t count, uint32_t bits,
                        unsigned long long* found, unsigned long long* next, const P2ParamsF64& p,
                        hipStream_t s) {
    if (count == 0) return 0;
    // about as many workgroups as the device holds at once; extra ones exit
    // at their first chunk once the search is over
    const uint64_t chunks = (count + GRIND_THREADS - 1) / GRIND_THREADS;
    const uint32_t grid = uint32_t(chunks < 4096 ? chunks : 4096);
    hipLaunchKernelGGL(grind_search_kernel, dim3(grid), dim3(GRIND_THREADS), 0, s, st, base, count, bits, found, next,
                       p);
    return int(hipGetLastError());
}

}  // namespace p3h
