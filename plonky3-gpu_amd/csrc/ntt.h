// Host-side launchers for the NTT / LDE kernels (ntt_impl.h; built per field by
// ntt_bb.hip and ntt_gl.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "field.h"

namespace p3h {

// Pass-shape knobs picked by the autotuner (0 = built-in default).
struct NttTuning {
    uint32_t log_cols;    // log2 columns per tile (default 4)
    uint32_t max_stages;  // cap on radix-2 stages per LDS pass (default: tile-limited)
    uint32_t log_slab;    // LDE only: run all passes on slabs of 2^log_slab columns in turn (0 = whole width)
};

// Stage twiddles: omega_{2^(q+1)}^(+-j) at (2^q - 1) + j for q < 10, j < 2^q,
// multiply-constant form; the same for every transform size.  The twiddle
// buffer of a 2^L transform is [omega^-e][omega^e] (e < n/2), then the
// inverse and the forward stage table.
constexpr uint32_t TWC_LEN = 1023;
template <class T>
inline const T* twc_of(const T* tw_inv_fwd, uint32_t L, bool forward) {
    return tw_inv_fwd + (L ? (uint64_t(1) << L) : 0) + (forward ? TWC_LEN : 0);
}
inline uint64_t twiddle_buffer_len(uint32_t L) { return (L ? (uint64_t(1) << L) : 0) + 2 * TWC_LEN; }

// Arguments of one NTT pass over a row-major matrix (see ntt_impl.h).
template <class T>
struct NttArgs {
    const T* src;
    T* dst;
    const T* tw;     // twiddles omega^(+-e), e < n/2, in multiply-constant form
    const T* twc;    // stage twiddles of the same direction (TWC_LEN entries, see twc_of)
    const T* scale;  // mode-dependent table, or nullptr
    uint64_t src_region_rows;  // row offset between regions in src
    uint64_t dst_region_rows;  // row offset between regions in dst
    uint32_t L;                // log2 transform size n
    uint32_t a;                // the pass covers index bits [a, a+S)
    uint32_t w;                // columns transformed (from src/dst column 0)
    uint32_t ld;               // row stride of src and dst in elements (>= w)
    uint32_t logC;             // columns per tile = 2^logC
    uint32_t n_col_tiles;
    uint32_t n_regions;        // cosets (LDE) or 1
};

// One descriptor of a work-list launch; block_start = first block of it.
template <class T>
struct NttWork {
    NttArgs<T> args;
    uint32_t block_start;
    uint32_t pad;
};

// One coset LDE of a batch (same meaning as lde_device's arguments).
template <class T>
struct LdeJob {
    const T* in;
    T* out;
    uint32_t L, w, k;
    const T* tw;
    const T* twist;
    NttTuning tune;
};

// All passes of a batch that share mode and width run as one launch.
template <class T>
struct LdePassGroup {
    int mode;
    uint32_t s;
    uint32_t blocks;
    uint32_t lds_bytes;  // dynamic LDS of the launch: the largest of its passes
    std::vector<NttWork<T>> work;
};

// Plan the passes of many LDEs: groups in execution order.
template <class F>
std::vector<LdePassGroup<typename F::T>> plan_lde_batch(const std::vector<LdeJob<typename F::T>>& jobs);

// One pass of one LDE in the queue launch (launch_lde_queue): its blocks are
// taken `chunk` at a time, one ticket per chunk, tickets [first, first +
// ceil(blocks / chunk)); it reads what item `dep` (the same LDE's previous
// pass, -1 for none) writes.
template <class T>
struct NttQueueItem {
    NttArgs<T> args;
    uint32_t first;
    uint32_t blocks;
    uint32_t chunk;
    int32_t dep;
    uint16_t mode;  // kernel variant (after the split-form choice)
    uint16_t s;     // pass width in index bits
};

// Every pass of many LDEs as one work list, in ticket order: slot by slot
// (the inverse outer passes farthest first, the fused inner pass, then the
// forward outer passes), within a slot LDE by LDE, so every dependency has
// lower tickets than its dependents.
template <class T>
struct LdeQueuePlan {
    std::vector<NttQueueItem<T>> items;  // ascending `first`
    uint32_t total = 0;                  // tickets (chunks of every pass; the launch's workgroups)
    uint32_t lds_bytes = 0;              // dynamic LDS: the largest pass's
};

// Plan the queue launch of many LDEs.  Returns false when a pass has a shape
// the queue kernel does not carry (a fused inner pass of 0, 4 or 5 stages: a
// transform of 1, 16 or 32 rows, or a tuning capped at 4 or 5 stages), when
// a pass's LDS tile would leave one workgroup per CU (a full tile that is not
// split, as on regions of 4 GiB and more), or when the blocks exceed one
// dispatch (2^22 workgroups); the grouped launches then run the batch.
template <class F>
bool plan_lde_queue(const std::vector<LdeJob<typename F::T>>& jobs, LdeQueuePlan<typename F::T>& plan);

// Words of the queue launch's counter block: the ticket counter, the error
// word (set when a wait exceeded its spin limit) and one completion count per
// item.  launch_lde_queue zeroes them with a kernel of its own first, so the
// pair can be captured into a HIP graph and replayed.
inline size_t lde_queue_counter_words(size_t items) { return 2 + items; }

// The queue launch: plan.total workgroups each take a ticket in order from the
// counter block, wait for the pass their ticket's pass reads, run the chunk's
// blocks and count them done.  d_items holds plan.items in device memory.
template <class F>
int launch_lde_queue(const LdeQueuePlan<typename F::T>& plan, const NttQueueItem<typename F::T>* d_items,
                     uint32_t* d_counters, hipStream_t st);

// Once per device, before the first pass launch (the context does it at
// creation): lets every pass kernel take up to 160 KiB of dynamic LDS.
int ntt_prepare_device();
int ntt_prepare_gl();        // the Goldilocks passes (ntt_gl.hip)
int ntt_prepare_queue_bb();  // the BabyBear queue kernel (ntt_queue.hip)

// Launch one group; d_work holds g.work in device memory.
template <class F>
int launch_lde_group(const LdePassGroup<typename F::T>& g, const NttWork<typename F::T>* d_work, hipStream_t st);

// Coset LDE, coset-major output: out[(r * n + i) * w + c] = LDE row (i << k | r).
// tw_inv_fwd: [omega_n^-e for e < n/2][omega_n^e for e < n/2] (multiply-constant form);
// twist: per coset r and bit-reversed position p, (shift * omega_N^r)^bitrev(p) / n.
// `in` is read only; out has (n << k) rows.
template <class F>
int lde_device(const typename F::T* in, typename F::T* out, uint32_t L, uint32_t w, uint32_t k,
               const typename F::T* tw_inv_fwd, const typename F::T* twist, const NttTuning& tune,
               hipStream_t st);

// kind 0 dft, 1 idft, 2 coset_dft, 3 coset_idft; natural order in and out.
// row_scale: forward kinds -> multiplier of input row i (shift^i) or nullptr;
// inverse kinds -> multiplier of output row j (n^-1 * shift^-j), required.
// tmp: n x w scratch used by inverse kinds when more than one pass is needed.
template <class F>
int dft_device(int kind, const typename F::T* in, typename F::T* out, typename F::T* tmp, uint32_t L,
               uint32_t w, const typename F::T* tw_inv_fwd, const typename F::T* row_scale,
               const NttTuning& tune, hipStream_t st);

// out[i] = base^i * mult for i < count, all in multiply-constant form.
template <class F>
int powers_device(typename F::T* out, uint64_t count, typename F::T base_c, typename F::T mult_c,
                  typename F::T one_c, hipStream_t st);

// out[r * n + p] = coset_base[r]^bitrev(p) * inv_n (multiply-constant form).
template <class F>
int lde_twist_device(typename F::T* out, uint32_t L, uint32_t n_cosets, const typename F::T* coset_base_c,
                     typename F::T inv_n_c, typename F::T one_c, hipStream_t st);

template <class T>
int coset_to_natural_device(const T* in, T* out, uint32_t L, uint32_t k, uint32_t w, hipStream_t st);
// VALU ceiling probe of the NTT butterflies: blocks x 256 threads, each
// `iters` x 32 butterflies on registers (out: blocks x 256 words; tw: 16
// canonical twiddles in the field's multiply-constant form).
template <class F>
int butterfly_probe_device(typename F::T* out, const typename F::T* tw, uint32_t blocks, uint32_t iters,
                           hipStream_t st);

}  // namespace p3h
