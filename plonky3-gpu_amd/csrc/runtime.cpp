// libp3hip runtime: contexts, the caching device allocator with an HBM
// budget, device-resident matrices and trees, table caches, and every C-ABI
// entry point of include/p3hip.h.  All device work of a context is ordered on
// its one HIP stream; host copies synchronise that stream.

#include "p3hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cstdio>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <cstring>
#include <cstdlib>
#include <functional>
#include <map>
#include <numeric>
#include <random>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "fri.h"
#include "host_plan.h"
#include "challenger.h"
#include "pcs.h"
#include "merkle.h"
#include "ntt.h"

using namespace p3h;

namespace {

// ------------------------------------------------------------- host math ---
constexpr uint32_t BB_P = BabyBear::P;
constexpr uint64_t GL_P = Goldilocks::P;

uint32_t bb_mul(uint32_t a, uint32_t b) { return uint32_t(uint64_t(a) * b % BB_P); }
uint32_t bb_pow(uint32_t a, uint64_t e) {
    uint32_t r = 1;
    while (e) {
        if (e & 1) r = bb_mul(r, a);
        a = bb_mul(a, a);
        e >>= 1;
    }
    return r;
}
uint32_t bb_inv(uint32_t a) { return bb_pow(a, BB_P - 2); }
uint32_t bb_mont(uint32_t a) { return uint32_t((uint64_t(a) << 32) % BB_P); }  // a * 2^32 mod p
uint32_t bb_root(uint32_t bits) { return bb_pow(440564289u, uint64_t(1) << (27 - bits)); }

// BabyBear^4 = F_p[a]/(a^4 - 11) on the host (small per-call constants only).
using HE4 = std::array<uint32_t, 4>;
HE4 he4_add(const HE4& a, const HE4& b) {
    HE4 r;
    for (int i = 0; i < 4; ++i) r[i] = uint32_t((uint64_t(a[i]) + b[i]) % BB_P);
    return r;
}
HE4 he4_sub(const HE4& a, const HE4& b) {
    HE4 r;
    for (int i = 0; i < 4; ++i) r[i] = uint32_t((uint64_t(a[i]) + BB_P - b[i]) % BB_P);
    return r;
}
HE4 he4_mul(const HE4& a, const HE4& b) {
    uint64_t t[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) t[i + j] = (t[i + j] + uint64_t(a[i]) * b[j]) % BB_P;
    HE4 r;
    for (int i = 0; i < 4; ++i) r[i] = uint32_t((t[i] + 11 * (i + 4 < 7 ? t[i + 4] : 0)) % BB_P);
    return r;
}
HE4 he4_pow(HE4 a, uint64_t e) {
    HE4 r = {1, 0, 0, 0};
    while (e) {
        if (e & 1) r = he4_mul(r, a);
        a = he4_mul(a, a);
        e >>= 1;
    }
    return r;
}
HE4 he4_mont(const HE4& a) { return HE4{bb_mont(a[0]), bb_mont(a[1]), bb_mont(a[2]), bb_mont(a[3])}; }

uint64_t gl_mul(uint64_t a, uint64_t b) { return uint64_t((unsigned __int128)a * b % GL_P); }
uint64_t gl_pow(uint64_t a, uint64_t e) {
    uint64_t r = 1;
    while (e) {
        if (e & 1) r = gl_mul(r, a);
        a = gl_mul(a, a);
        e >>= 1;
    }
    return r;
}
uint64_t gl_inv(uint64_t a) { return gl_pow(a, GL_P - 2); }
// Goldilocks two-adic structure (not defined by the reference): generator 7,
// omega_{2^32} = 7^((p-1)/2^32).
uint64_t gl_root(uint32_t bits) { return gl_pow(gl_pow(7, (GL_P - 1) >> 32), uint64_t(1) << (32 - bits)); }

// Per-field constant helpers: "c" = multiply-constant form used by kernels.
struct HBB {
    using T = uint32_t;
    using F = BabyBear;
    using D = BBDom;
    static constexpr int FIELD = P3H_BABYBEAR;
    static constexpr uint32_t MAX_BITS = 27;
    static T c(T x) { return bb_mont(x); }
    static T mul(T a, T b) { return bb_mul(a, b); }
    static T pow(T a, uint64_t e) { return bb_pow(a, e); }
    static T inv(T a) { return bb_inv(a); }
    static T root(uint32_t bits) { return bb_root(bits); }
    static bool canonical(uint64_t x) { return x < BB_P; }
};
struct HGL {
    using T = uint64_t;
    using F = Goldilocks;
    using D = GLDom;
    static constexpr int FIELD = P3H_GOLDILOCKS;
    static constexpr uint32_t MAX_BITS = 32;
    static T c(T x) { return x; }
    static T mul(T a, T b) { return gl_mul(a, b); }
    static T pow(T a, uint64_t e) { return gl_pow(a, e); }
    static T inv(T a) { return gl_inv(a); }
    static T root(uint32_t bits) { return gl_root(bits); }
    static bool canonical(uint64_t x) { return x < GL_P; }
};

struct Err {
    int code;
    std::string msg;
};

size_t elem_size(int field) { return field == P3H_GOLDILOCKS ? 8 : 4; }
uint32_t log2_exact(uint64_t n) {
    if (n == 0 || (n & (n - 1))) throw Err{P3H_EINVAL, "height must be a non-zero power of two"};
    return uint32_t(__builtin_ctzll(n));
}
void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw Err{P3H_EDEVICE, std::string(what) + ": " + hipGetErrorString(e)};
}
void launch_check(int e, const char* what) { hip_check(hipError_t(e), what); }

}  // namespace

// ======================================================= context & memory ===

struct DevBlock {
    p3h_ctx* ctx = nullptr;
    void* ptr = nullptr;
    size_t cls = 0;
    DevBlock() = default;
    DevBlock(const DevBlock&) = delete;
    DevBlock& operator=(const DevBlock&) = delete;
    ~DevBlock();
};
using DevPtr = std::shared_ptr<DevBlock>;

struct PhaseStat {
    uint64_t launches = 0;
    double ms = 0;
};

struct p3h_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    uint64_t budget = 0, in_use = 0, cached = 0;
    uint64_t table_bytes = 0;  // part of in_use held by the table caches below
    // allocation-failure injection (P3H_FAIL_ALLOC_AFTER): -1 = off
    int64_t fail_after = -1;
    uint64_t n_allocs = 0;
    std::multimap<size_t, void*> free_blocks;
    std::string err;
    bool p2_set[2] = {false, false};
    P2Params<uint32_t, 16> bbp{};
    P2Params<uint64_t, 8> glp{};
    P2ParamsF64 bbp64{};     // BabyBear constants as doubles (binary64 engine)
    int hash_engine = 1;     // BabyBear Poseidon2: 0 = uint32 Montgomery, 1 = binary64 (default)
    std::map<std::tuple<int, uint32_t>, DevPtr> twiddles;                      // (field, L)
    std::map<std::tuple<int, uint32_t, uint32_t, uint64_t>, DevPtr> twists;    // (field, L, k, shift)
    std::map<std::tuple<int, uint32_t, int, uint64_t>, DevPtr> row_scales;     // (field, L, kind, shift)
    std::map<uint32_t, DevPtr> fri_tinv;                                        // LH
    NttTuning tune{};
    // autotuned pass shapes per (field, log_n, width, added bits)
    std::map<std::tuple<int, uint32_t, uint32_t, uint32_t>, NttTuning> tuned;
    bool profiling = false;
    bool capturing = false;
    // pinned staging ring for small copies in stream order (h2d_async, d2h_gather)
    uint8_t* stage = nullptr;
    size_t stage_head = 0;
    std::map<std::string, PhaseStat> stats;
    struct Pending {
        std::string phase;
        hipEvent_t a, b;
    };
    std::vector<Pending> pending;
};

DevBlock::~DevBlock() {
    if (ctx && ptr) {
        ctx->in_use -= cls;
        ctx->cached += cls;
        ctx->free_blocks.emplace(cls, ptr);
    }
}

namespace {

void trim(p3h_ctx* ctx) {
    (void)hipStreamSynchronize(ctx->stream);
    for (auto& kv : ctx->free_blocks) (void)hipFree(kv.second);
    ctx->free_blocks.clear();
    ctx->cached = 0;
}

DevPtr dalloc(p3h_ctx* ctx, size_t bytes) {
    size_t cls = size_class(bytes);
    if (ctx->fail_after >= 0 && ctx->n_allocs >= uint64_t(ctx->fail_after))
        throw Err{P3H_ENOMEM, "injected allocation failure (P3H_FAIL_ALLOC_AFTER=" +
                                  std::to_string(ctx->fail_after) + ")"};
    if (ctx->in_use + cls > ctx->budget)
        throw Err{P3H_EBUDGET, "request of " + std::to_string(cls) + " bytes exceeds the HBM budget (" +
                                   std::to_string(ctx->in_use) + " of " + std::to_string(ctx->budget) + " in use)"};
    auto blk = std::make_shared<DevBlock>();
    auto it = ctx->free_blocks.find(cls);
    if (it != ctx->free_blocks.end()) {
        blk->ptr = it->second;
        ctx->free_blocks.erase(it);
        ctx->cached -= cls;
    } else {
        void* p = nullptr;
        if (hipMalloc(&p, cls) != hipSuccess) {
            (void)hipGetLastError();
            trim(ctx);
            if (hipMalloc(&p, cls) != hipSuccess) {
                (void)hipGetLastError();
                throw Err{P3H_ENOMEM, "hipMalloc of " + std::to_string(cls) + " bytes failed"};
            }
        }
        blk->ptr = p;
    }
    blk->ctx = ctx;
    blk->cls = cls;
    ctx->in_use += cls;
    ++ctx->n_allocs;
    return blk;
}

template <class T>
T* dptr(const DevPtr& p) {
    return static_cast<T*>(p->ptr);
}

void h2d(p3h_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!bytes) return;
    hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream), "hipMemcpyAsync H2D");
    hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
}
void d2h(p3h_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!bytes) return;
    hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream), "hipMemcpyAsync D2H");
    hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
}

// Small copies through the context's pinned staging ring.  A region of the
// ring is written again only after a stream synchronisation, so no copy still
// queued on the stream reads bytes that have been overwritten.
constexpr size_t STAGE_BYTES = size_t(4) << 20;
constexpr size_t STAGE_MAX_ONE = size_t(64) << 10;

uint8_t* stage_take(p3h_ctx* ctx, size_t bytes) {
    if (!ctx->stage)
        hip_check(hipHostMalloc(reinterpret_cast<void**>(&ctx->stage), STAGE_BYTES, hipHostMallocDefault),
                  "hipHostMalloc");
    const size_t need = (bytes + 255) & ~size_t(255);
    if (ctx->stage_head + need > STAGE_BYTES) {
        hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        ctx->stage_head = 0;
    }
    uint8_t* p = ctx->stage + ctx->stage_head;
    ctx->stage_head += need;
    return p;
}

// Host-to-device copy that returns without waiting for the device.
void h2d_async(p3h_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!bytes) return;
    if (bytes > STAGE_MAX_ONE || ctx->capturing) {
        h2d(ctx, dst, src, bytes);
        return;
    }
    uint8_t* p = stage_take(ctx, bytes);
    std::memcpy(p, src, bytes);
    hip_check(hipMemcpyAsync(dst, p, bytes, hipMemcpyHostToDevice, ctx->stream), "hipMemcpyAsync H2D");
}

// Several device-to-host copies with one synchronisation.
struct Readback {
    void* dst;
    const void* src;
    size_t bytes;
};
void d2h_gather(p3h_ctx* ctx, const std::vector<Readback>& rb) {
    size_t total = 0;
    for (const auto& r : rb) total += (r.bytes + 255) & ~size_t(255);
    if (total > STAGE_BYTES / 2) {
        for (const auto& r : rb) d2h(ctx, r.dst, r.src, r.bytes);
        return;
    }
    std::vector<uint8_t*> at;
    for (const auto& r : rb) {
        uint8_t* p = stage_take(ctx, r.bytes);
        if (r.bytes)
            hip_check(hipMemcpyAsync(p, r.src, r.bytes, hipMemcpyDeviceToHost, ctx->stream), "hipMemcpyAsync D2H");
        at.push_back(p);
    }
    hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    for (size_t i = 0; i < rb.size(); ++i)
        if (rb[i].bytes) std::memcpy(rb[i].dst, at[i], rb[i].bytes);
}

// RAII phase timer: HIP events around the phase's launches on the context stream.
struct Phase {
    p3h_ctx* ctx;
    std::string name;
    hipEvent_t a = nullptr;
    Phase(p3h_ctx* c, const char* n) : ctx(c), name(n) {
        if (ctx->profiling && !ctx->capturing) {
            (void)hipEventCreate(&a);
            (void)hipEventRecord(a, ctx->stream);
        }
    }
    ~Phase() {
        if (a) {
            hipEvent_t b;
            (void)hipEventCreate(&b);
            (void)hipEventRecord(b, ctx->stream);
            ctx->pending.push_back({name, a, b});
        }
    }
};

void resolve_pending(p3h_ctx* ctx) {
    for (auto& p : ctx->pending) {
        (void)hipEventSynchronize(p.b);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, p.a, p.b);
        auto& s = ctx->stats[p.phase];
        s.launches++;
        s.ms += ms;
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    ctx->pending.clear();
}

}  // namespace

// ============================================================= handles ===

struct p3h_mat {
    p3h_ctx* ctx;
    int field;
    uint64_t height;
    uint32_t width;
    DevPtr buf;
    uint64_t offset_elems = 0;  // view offset into buf
    uint32_t k = 0;             // coset-major layout: 2^k cosets of height 2^L
    uint32_t L = 0;
    std::atomic<int> refs{1};
    template <class T>
    T* data() const {
        return static_cast<T*>(buf->ptr) + offset_elems;
    }
};

struct p3h_tree {
    p3h_ctx* ctx;
    int field;
    bool flat;
    uint32_t cap_height = 0;
    std::vector<DevPtr> layers;
    std::vector<uint64_t> counts;
    std::vector<p3h_mat*> mats;  // commit order, referenced
    std::vector<uint64_t> heights;
    std::vector<uint32_t> widths;
    std::vector<DevPtr> seg_tables;  // keep device segment tables alive
    // flat trees built over a virtual concatenation (PCS): the segments
    std::vector<char> flat_segs_host;  // RowSeg<T> array (host copy)
    uint32_t total_width = 0;
};

struct p3h_fri {
    p3h_ctx* ctx;
    std::vector<p3h_tree*> trees;
    std::vector<p3h_mat*> layers;
};

// Device-resident transcript (DevChal) of one context, plus a small device
// buffer for samples and the grinding result.
struct p3h_chal {
    p3h_ctx* ctx;
    DevPtr st;
    DevPtr io;  // 64 sample words, then the 64-bit grinding minimum and the search's work counter
};

namespace {

void mat_ref(p3h_mat* m) { m->refs.fetch_add(1); }
void mat_unref(p3h_mat* m) {
    if (m && m->refs.fetch_sub(1) == 1) delete m;
}

p3h_mat* new_mat(p3h_ctx* ctx, int field, uint64_t h, uint32_t w) {
    auto* m = new p3h_mat{};
    m->ctx = ctx;
    m->field = field;
    m->height = h;
    m->width = w;
    try {
        m->buf = dalloc(ctx, std::max<size_t>(1, h * w * elem_size(field)));
    } catch (...) {
        delete m;
        throw;
    }
    return m;
}

template <class H>
RowSeg<typename H::T> seg_of(const p3h_mat* m) {
    RowSeg<typename H::T> s{};
    s.ptr = m->data<typename H::T>();
    s.hmask = ~0ull;
    s.w = m->width;
    s.k = m->k;
    s.L = m->L;
    return s;
}

template <class H>
const P2Params<typename H::T, H::D::W>& p2(p3h_ctx* ctx);
template <>
const P2Params<uint32_t, 16>& p2<HBB>(p3h_ctx* ctx) {
    if (!ctx->p2_set[P3H_BABYBEAR]) throw Err{P3H_ESTATE, "Poseidon2 constants for BabyBear not set"};
    return ctx->bbp;
}
template <>
const P2Params<uint64_t, 8>& p2<HGL>(p3h_ctx* ctx) {
    if (!ctx->p2_set[P3H_GOLDILOCKS]) throw Err{P3H_ESTATE, "Poseidon2 constants for Goldilocks not set"};
    return ctx->glp;
}

// [omega^-e][omega^e], e < n/2, then the inverse and forward stage tables
// (omega_{2^(q+1)}^(-+j) at (2^q - 1) + j, q < 10), multiply-constant form
template <class H>
const typename H::T* twiddles(p3h_ctx* ctx, uint32_t L) {
    using T = typename H::T;
    auto key = std::make_tuple(H::FIELD, L);
    auto it = ctx->twiddles.find(key);
    if (it != ctx->twiddles.end()) return dptr<T>(it->second);
    uint64_t half = L ? (1ull << (L - 1)) : 0;
    DevPtr buf = dalloc(ctx, twiddle_buffer_len(L) * sizeof(T));
    std::vector<T> stages(2 * TWC_LEN);
    for (uint32_t q = 0; (1u << (q + 1)) - 1 <= TWC_LEN; ++q) {
        const T wf = H::root(q + 1), wi = H::inv(wf);
        T xf = 1, xi = 1;
        for (uint32_t j = 0; j < (1u << q); ++j) {
            stages[(1u << q) - 1 + j] = H::c(xi);
            stages[TWC_LEN + (1u << q) - 1 + j] = H::c(xf);
            xf = H::mul(xf, wf);
            xi = H::mul(xi, wi);
        }
    }
    h2d(ctx, const_cast<T*>(twc_of(dptr<T>(buf), L, false)), stages.data(), stages.size() * sizeof(T));
    if (half) {
        T w = H::root(L);
        launch_check(powers_device<typename H::F>(dptr<T>(buf), half, H::c(H::inv(w)), H::c(1), H::c(1), ctx->stream),
                     "twiddles");
        launch_check(powers_device<typename H::F>(dptr<T>(buf) + half, half, H::c(w), H::c(1), H::c(1), ctx->stream),
                     "twiddles");
    }
    ctx->twiddles[key] = buf;
    ctx->table_bytes += buf->cls;
    return dptr<T>(buf);
}

template <class H>
const typename H::T* lde_twist(p3h_ctx* ctx, uint32_t L, uint32_t k, typename H::T shift) {
    using T = typename H::T;
    auto key = std::make_tuple(H::FIELD, L, k, uint64_t(shift));
    auto it = ctx->twists.find(key);
    if (it != ctx->twists.end()) return dptr<T>(it->second);
    uint32_t nc = 1u << k;
    std::vector<T> bases(nc);
    T wN = H::root(L + k), cur = shift;
    for (uint32_t r = 0; r < nc; ++r) {
        bases[r] = H::c(cur);
        cur = H::mul(cur, wN);
    }
    DevPtr bb = dalloc(ctx, nc * sizeof(T));
    h2d(ctx, bb->ptr, bases.data(), nc * sizeof(T));
    DevPtr buf = dalloc(ctx, (uint64_t(nc) << L) * sizeof(T));
    T inv_n = H::inv(H::pow(2, L));
    launch_check(lde_twist_device<typename H::F>(dptr<T>(buf), L, nc, dptr<T>(bb), H::c(inv_n), H::c(1), ctx->stream),
                 "lde twist");
    hip_check(hipStreamSynchronize(ctx->stream), "sync");
    ctx->twists[key] = buf;
    ctx->table_bytes += buf->cls;
    return dptr<T>(buf);
}

// kind 2: shift^i; kind 1: 1/n; kind 3: shift^-i / n
template <class H>
const typename H::T* row_scale(p3h_ctx* ctx, uint32_t L, int kind, typename H::T shift) {
    using T = typename H::T;
    if (kind == 0) return nullptr;
    auto key = std::make_tuple(H::FIELD, L, kind, uint64_t(shift));
    auto it = ctx->row_scales.find(key);
    if (it != ctx->row_scales.end()) return dptr<T>(it->second);
    uint64_t n = 1ull << L;
    DevPtr buf = dalloc(ctx, n * sizeof(T));
    T inv_n = H::inv(H::pow(2, L));
    T base = 1, mult = 1;
    if (kind == 2) base = shift;
    if (kind == 1) mult = inv_n;
    if (kind == 3) {
        base = H::inv(shift);
        mult = inv_n;
    }
    launch_check(powers_device<typename H::F>(dptr<T>(buf), n, H::c(base), H::c(mult), H::c(1), ctx->stream),
                 "row scale");
    ctx->row_scales[key] = buf;
    ctx->table_bytes += buf->cls;
    return dptr<T>(buf);
}

// natural-order device copy of m (m itself when already natural)
template <class H>
const typename H::T* natural_data(p3h_ctx* ctx, const p3h_mat* m, DevPtr& tmp) {
    using T = typename H::T;
    if (m->k == 0) return m->data<T>();
    tmp = dalloc(ctx, m->height * m->width * sizeof(T));
    launch_check(coset_to_natural_device<T>(m->data<T>(), dptr<T>(tmp), m->L, m->k, m->width, ctx->stream),
                 "coset_to_natural");
    return dptr<T>(tmp);
}

// ------------------------------------------------------------------ LDE ---
NttTuning tuning_for(const p3h_ctx* ctx, int field, uint32_t L, uint32_t w, uint32_t k) {
    auto it = ctx->tuned.find(std::make_tuple(field, L, w, k));
    return it == ctx->tuned.end() ? ctx->tune : it->second;
}

template <class H>
p3h_mat* do_lde(p3h_ctx* ctx, const p3h_mat* in, uint32_t k, uint64_t shift64) {
    using T = typename H::T;
    if (!H::canonical(shift64)) throw Err{P3H_EINVAL, "shift is not a canonical field element"};
    uint32_t L = log2_exact(in->height);
    if (L + k > H::MAX_BITS) throw Err{P3H_EINVAL, "LDE height exceeds the field's two-adicity"};
    if (in->width == 0) throw Err{P3H_EINVAL, "matrix width is zero"};
    DevPtr tmp;
    const T* src = natural_data<H>(ctx, in, tmp);
    const T* tw = twiddles<H>(ctx, L);
    const T* twist = lde_twist<H>(ctx, L, k, T(shift64));
    p3h_mat* out = new_mat(ctx, H::FIELD, in->height << k, in->width);
    out->k = k;
    out->L = L;
    try {
        Phase ph(ctx, "lde");
        launch_check(lde_device<typename H::F>(src, out->data<T>(), L, in->width, k, tw, twist,
                                               tuning_for(ctx, H::FIELD, L, in->width, k), ctx->stream),
                     "lde");
    } catch (...) {
        mat_unref(out);
        throw;
    }
    if (k == 0) out->k = 0;
    return out;
}

template <class H>
p3h_mat* do_dft(p3h_ctx* ctx, int kind, const p3h_mat* in, uint64_t shift64) {
    using T = typename H::T;
    if (kind < 0 || kind > 3) throw Err{P3H_EINVAL, "dft kind must be 0..3"};
    if (!H::canonical(shift64)) throw Err{P3H_EINVAL, "shift is not a canonical field element"};
    if ((kind == 2 || kind == 3) && shift64 == 0) throw Err{P3H_EINVAL, "coset shift must be non-zero"};
    uint32_t L = log2_exact(in->height);
    if (L > H::MAX_BITS) throw Err{P3H_EINVAL, "height exceeds the field's two-adicity"};
    if (in->width == 0) throw Err{P3H_EINVAL, "matrix width is zero"};
    DevPtr tmpn;
    const T* src = natural_data<H>(ctx, in, tmpn);
    const T* tw = twiddles<H>(ctx, L);
    const T* sc = row_scale<H>(ctx, L, kind, T(shift64));
    p3h_mat* out = new_mat(ctx, H::FIELD, in->height, in->width);
    try {
        DevPtr tmp;
        if (kind == 1 || kind == 3) tmp = dalloc(ctx, in->height * in->width * sizeof(T));
        Phase ph(ctx, "dft");
        launch_check(dft_device<typename H::F>(kind, src, out->data<T>(), tmp ? dptr<T>(tmp) : nullptr, L, in->width,
                                               tw, sc, ctx->tune, ctx->stream),
                     "dft");
        if (tmp) hip_check(hipStreamSynchronize(ctx->stream), "sync");
    } catch (...) {
        mat_unref(out);
        throw;
    }
    return out;
}

// ----------------------------------------------------------------- trees ---
template <class H>
DevPtr upload_segs(p3h_ctx* ctx, const std::vector<RowSeg<typename H::T>>& segs) {
    DevPtr d = dalloc(ctx, segs.size() * sizeof(RowSeg<typename H::T>));
    h2d_async(ctx, d->ptr, segs.data(), segs.size() * sizeof(RowSeg<typename H::T>));
    return d;
}

// A recorded device launch on the context stream (throws on launch errors).
using LaunchFn = std::function<void(p3h_ctx*)>;
using LaunchList = std::vector<std::pair<const char*, LaunchFn>>;

// Multi-height tree: allocate every layer and upload the segment tables now,
// return the kernel launches (leaf hashing, then one compress-with-injection
// launch per level) for the caller to run or capture.
template <class H>
p3h_tree* prepare_mmcs(p3h_ctx* ctx, const p3h_mat* const* mats, uint32_t n, uint32_t cap_height, LaunchList& out) {
    using T = typename H::T;
    using D = typename H::D;
    const P2Params<T, D::W> pp = p2<H>(ctx);
    if (n == 0) throw Err{P3H_EINVAL, "no matrices to commit"};
    for (uint32_t i = 0; i < n; ++i) {
        if (!mats[i] || mats[i]->field != H::FIELD) throw Err{P3H_EINVAL, "matrix field mismatch"};
        log2_exact(mats[i]->height);
    }
    std::vector<uint32_t> order(n);
    for (uint32_t i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t a, uint32_t b) { return mats[a]->height > mats[b]->height; });
    const uint64_t maxh = mats[order[0]]->height;
    auto* t = new p3h_tree{};
    t->ctx = ctx;
    t->field = H::FIELD;
    t->flat = false;
    t->cap_height = cap_height;
    try {
        for (uint32_t i = 0; i < n; ++i) {
            t->mats.push_back(const_cast<p3h_mat*>(mats[i]));
            mat_ref(const_cast<p3h_mat*>(mats[i]));
            t->heights.push_back(mats[i]->height);
            t->widths.push_back(mats[i]->width);
            t->total_width += mats[i]->width;
        }
        size_t next = 0;
        auto take = [&](uint64_t h, uint32_t& w) {
            std::vector<RowSeg<T>> segs;
            w = 0;
            while (next < n && mats[order[next]]->height == h) {
                segs.push_back(seg_of<H>(mats[order[next]]));
                w += mats[order[next]]->width;
                ++next;
            }
            return segs;
        };
        uint32_t w0 = 0;
        auto top = take(maxh, w0);
        DevPtr dsegs = upload_segs<H>(ctx, top);
        t->seg_tables.push_back(dsegs);
        DevPtr leaves = dalloc(ctx, maxh * D::DIG * sizeof(T));
        {
            const RowSeg<T>* ds = dptr<RowSeg<T>>(dsegs);
            uint32_t ns = uint32_t(top.size());
            RowSeg<T> s0 = top[0];
            T* lv = dptr<T>(leaves);
            const bool f64 = H::FIELD == P3H_BABYBEAR && ctx->hash_engine == 1;
            const P2ParamsF64 p64 = ctx->bbp64;
            out.emplace_back("leaf_hash", [=](p3h_ctx* c) {
                if constexpr (H::FIELD == P3H_BABYBEAR) {
                    if (f64) {
                        launch_check(hash_rows_f64_device(ds, ns, w0, s0, maxh, lv, p64, c->stream), "leaf hash");
                        return;
                    }
                }
                launch_check(hash_rows_device<D>(ds, ns, w0, s0, maxh, lv, pp, c->stream), "leaf hash");
            });
        }
        t->layers.push_back(leaves);
        t->counts.push_back(maxh);
        uint64_t level = maxh, cap = 1ull << cap_height;
        const bool f64_tail = H::FIELD == P3H_BABYBEAR && ctx->hash_engine == 1;
        while (level > cap) {
            // the last levels, when no matrix is left to inject, in one launch
            const bool injections_left = next < n && mats[order[next]]->height >= cap;
            if (f64_tail && level <= 128 && !injections_left) {
                std::vector<uint64_t> counts;
                std::vector<uint32_t*> outs;
                for (uint64_t c = level >> 1; c >= cap; c >>= 1) {
                    DevPtr nl = dalloc(ctx, c * D::DIG * sizeof(T));
                    outs.push_back(reinterpret_cast<uint32_t*>(dptr<T>(nl)));
                    counts.push_back(c);
                    t->layers.push_back(nl);
                    t->counts.push_back(c);
                    if (c == cap) break;
                }
                const uint32_t* prev = reinterpret_cast<const uint32_t*>(dptr<T>(t->layers[t->layers.size() - 1 - outs.size()]));
                const uint64_t pc = level;
                const P2ParamsF64 p64t = ctx->bbp64;
                out.emplace_back("compress", [=](p3h_ctx* c) {
                    launch_check(compress_tail_f64_device(prev, pc, outs.data(), counts.data(), uint32_t(outs.size()),
                                                          p64t, c->stream),
                                 "compress");
                });
                break;
            }
            level >>= 1;
            DevPtr nl = dalloc(ctx, level * D::DIG * sizeof(T));
            uint32_t wi = 0;
            auto inj = take(level, wi);
            DevPtr di;
            if (!inj.empty()) {
                di = upload_segs<H>(ctx, inj);
                t->seg_tables.push_back(di);
            }
            const T* prev = dptr<T>(t->layers.back());
            T* nxt = dptr<T>(nl);
            const RowSeg<T>* dinj = di ? dptr<RowSeg<T>>(di) : nullptr;
            uint32_t ninj = uint32_t(inj.size());
            uint64_t lv = level;
            const bool f64c = H::FIELD == P3H_BABYBEAR && ctx->hash_engine == 1;
            const P2ParamsF64 p64c = ctx->bbp64;
            out.emplace_back("compress", [=](p3h_ctx* c) {
                if constexpr (H::FIELD == P3H_BABYBEAR) {
                    if (f64c) {
                        launch_check(compress_layer_f64_device(prev, lv * 2, nxt, lv, dinj, ninj, wi, p64c, c->stream),
                                     "compress");
                        return;
                    }
                }
                launch_check(compress_layer_device<D>(prev, lv * 2, nxt, lv, dinj, ninj, wi, pp, c->stream),
                             "compress");
            });
            t->layers.push_back(nl);
            t->counts.push_back(level);
        }
    } catch (...) {
        p3h_tree_free(t);
        throw;
    }
    return t;
}

// Consecutive launches of one phase share one pair of timing events.
void run_launches(p3h_ctx* ctx, const LaunchList& ls) {
    std::unique_ptr<Phase> ph;
    for (const auto& l : ls) {
        if (!ph || ph->name != l.first) {
            ph.reset();
            ph.reset(new Phase(ctx, l.first));
        }
        l.second(ctx);
    }
}

template <class H>
p3h_tree* do_mmcs_commit(p3h_ctx* ctx, const p3h_mat* const* mats, uint32_t n, uint32_t cap_height) {
    LaunchList ls;
    p3h_tree* t = prepare_mmcs<H>(ctx, mats, n, cap_height, ls);
    try {
        run_launches(ctx, ls);
    } catch (...) {
        p3h_tree_free(t);
        throw;
    }
    return t;
}

// flat tree over one (possibly virtual) matrix given by segments
template <class H>
p3h_tree* flat_tree(p3h_ctx* ctx, const std::vector<RowSeg<typename H::T>>& segs, uint64_t height, uint32_t width,
                    const std::vector<p3h_mat*>& refs) {
    using T = typename H::T;
    using D = typename H::D;
    const auto& pp = p2<H>(ctx);
    if (height == 0 || width == 0) throw Err{P3H_EINVAL, "commit_matrix: empty matrix"};
    auto* t = new p3h_tree{};
    t->ctx = ctx;
    t->field = H::FIELD;
    t->flat = true;
    t->total_width = width;
    try {
        for (auto* m : refs) {
            t->mats.push_back(m);
            mat_ref(m);
        }
        t->heights.push_back(height);
        t->widths.push_back(width);
        t->flat_segs_host.resize(segs.size() * sizeof(RowSeg<T>));
        std::memcpy(t->flat_segs_host.data(), segs.data(), t->flat_segs_host.size());
        bool f64 = false;
        if constexpr (H::FIELD == P3H_BABYBEAR) f64 = ctx->hash_engine == 1;
        // the device segment table is uploaded only for a kernel that reads it
        // (openings upload their own from flat_segs_host)
        DevPtr dsegs;
        bool need_table = true;
        if constexpr (H::FIELD == P3H_BABYBEAR)
            if (f64) need_table = hash_rows_f64_reads_table(uint32_t(segs.size()), segs[0], height);
        if (need_table) {
            dsegs = upload_segs<H>(ctx, segs);
            t->seg_tables.push_back(dsegs);
        }
        const RowSeg<T>* dsegs_ptr = need_table ? dptr<RowSeg<T>>(dsegs) : nullptr;
        DevPtr leaves = dalloc(ctx, height * D::DIG * sizeof(T));
        {
            Phase ph(ctx, "leaf_hash");
            bool done = false;
            if constexpr (H::FIELD == P3H_BABYBEAR) {
                if (f64) {
                    launch_check(hash_rows_f64_device(dsegs_ptr, uint32_t(segs.size()), width, segs[0], height,
                                                      dptr<T>(leaves), ctx->bbp64, ctx->stream),
                                 "leaf hash");
                    done = true;
                }
            }
            if (!done)
                launch_check(hash_rows_device<D>(dsegs_ptr, uint32_t(segs.size()), width, segs[0], height,
                                                 dptr<T>(leaves), pp, ctx->stream),
                             "leaf hash");
        }
        t->layers.push_back(leaves);
        t->counts.push_back(height);
        uint64_t cnt = height;
        std::unique_ptr<Phase> cph;  // one timing phase for all levels
        if (cnt > 1) cph.reset(new Phase(ctx, "compress"));
        while (cnt > 1) {
            if (H::FIELD == P3H_BABYBEAR && ctx->hash_engine == 1 && cnt <= 128) {
                // the last levels in one launch
                std::vector<uint64_t> counts;
                std::vector<uint32_t*> outs;
                const uint32_t* prev = reinterpret_cast<const uint32_t*>(dptr<T>(t->layers.back()));
                for (uint64_t c = (cnt + 1) / 2;; c = (c + 1) / 2) {
                    DevPtr nl = dalloc(ctx, c * D::DIG * sizeof(T));
                    outs.push_back(reinterpret_cast<uint32_t*>(dptr<T>(nl)));
                    counts.push_back(c);
                    t->layers.push_back(nl);
                    t->counts.push_back(c);
                    if (c == 1) break;
                }
                launch_check(compress_tail_f64_device(prev, cnt, outs.data(), counts.data(), uint32_t(outs.size()),
                                                      ctx->bbp64, ctx->stream),
                             "compress");
                break;
            }
            uint64_t nxt = (cnt + 1) / 2;
            DevPtr nl = dalloc(ctx, nxt * D::DIG * sizeof(T));
            bool done = false;
            if constexpr (H::FIELD == P3H_BABYBEAR) {
                if (ctx->hash_engine == 1) {
                    launch_check(compress_layer_f64_device(dptr<T>(t->layers.back()), cnt, dptr<T>(nl), nxt, nullptr,
                                                           0, 0, ctx->bbp64, ctx->stream),
                                 "compress");
                    done = true;
                }
            }
            if (!done)
                launch_check(compress_layer_device<D>(dptr<T>(t->layers.back()), cnt, dptr<T>(nl), nxt, nullptr, 0, 0,
                                                      pp, ctx->stream),
                             "compress");
            t->layers.push_back(nl);
            t->counts.push_back(nxt);
            cnt = nxt;
        }
    } catch (...) {
        p3h_tree_free(t);
        throw;
    }
    return t;
}

// Flat trees over several natural-order BabyBear matrices (widths multiples
// of 8), built together with the binary64 engine: one leaf-hash launch for
// all matrices, then one launch per level for every tree that still has that
// level.  Each tree is the one flat_tree builds on its matrix alone.
std::vector<p3h_tree*> flat_trees_multi(p3h_ctx* ctx, const std::vector<p3h_mat*>& mats) {
    std::vector<p3h_tree*> out;
    try {
        for (p3h_mat* m : mats) {
            if (m->field != P3H_BABYBEAR || m->k != 0 || m->height == 0 || m->width == 0 || m->width % 8)
                throw Err{P3H_EINVAL, "batched commit needs natural-order BabyBear matrices of width 8k"};
            auto* t = new p3h_tree{};
            out.push_back(t);
            t->ctx = ctx;
            t->field = P3H_BABYBEAR;
            t->flat = true;
            t->total_width = m->width;
            t->mats.push_back(m);
            mat_ref(m);
            t->heights.push_back(m->height);
            t->widths.push_back(m->width);
            const std::vector<RowSeg<uint32_t>> segs{seg_of<HBB>(m)};
            t->flat_segs_host.resize(sizeof(RowSeg<uint32_t>));  // openings upload their table from this
            std::memcpy(t->flat_segs_host.data(), segs.data(), sizeof(RowSeg<uint32_t>));
            for (uint64_t c = m->height;; c = (c + 1) / 2) {
                t->layers.push_back(dalloc(ctx, c * 8 * sizeof(uint32_t)));
                t->counts.push_back(c);
                if (c == 1) break;
            }
        }
        std::vector<MultiJob> jobs;
        for (size_t i = 0; i < mats.size(); ++i)
            jobs.push_back(MultiJob{mats[i]->data<uint32_t>(), dptr<uint32_t>(out[i]->layers[0]), 0, mats[i]->height,
                                    mats[i]->width});
        {
            Phase ph(ctx, "leaf_hash");
            launch_check(hash_rows_multi_f64_device(jobs.data(), uint32_t(jobs.size()), ctx->bbp64, ctx->stream),
                         "leaf hash");
        }
        Phase ph(ctx, "compress");
        for (size_t lev = 1;; ++lev) {
            jobs.clear();
            for (p3h_tree* t : out)
                if (lev < t->layers.size())
                    jobs.push_back(MultiJob{dptr<uint32_t>(t->layers[lev - 1]), dptr<uint32_t>(t->layers[lev]),
                                            t->counts[lev - 1], t->counts[lev], 0});
            if (jobs.empty()) break;
            launch_check(compress_multi_f64_device(jobs.data(), uint32_t(jobs.size()), ctx->bbp64, ctx->stream),
                         "compress");
        }
    } catch (...) {
        for (p3h_tree* t : out) p3h_tree_free(t);
        throw;
    }
    return out;
}

// MerkleTreeMmcs::open_batch (merkle_tree_mmcs.hpp:120-160) and the flat
// open_row (fri_merkle_tree_mmcs.hpp:129-141) for a batch of indices: the rows
// of every committed matrix at the index (scaled to each matrix's height) and
// the sibling digest of every layer below the top, zero where the sibling falls
// past an odd layer's end.  One table upload, one launch, one copy back.
template <class H>
void tree_open(p3h_ctx* ctx, const p3h_tree* t, const uint64_t* idx, uint32_t nq, typename H::T* values,
               typename H::T* proof) {
    using T = typename H::T;
    constexpr uint32_t DIG = H::D::DIG;
    const uint64_t maxh = t->counts[0];
    for (uint32_t q = 0; q < nq; ++q)
        if (idx[q] >= maxh) throw Err{P3H_EINVAL, "opening index out of range"};
    if (nq == 0) return;
    // rows: one segment per committed matrix (or the flat tree's segments) and
    // the row each query reads in it
    std::vector<RowSeg<T>> segs;
    std::vector<uint64_t> rows;
    uint32_t total_w = 0;
    if (values) {
        if (t->flat) {
            segs.resize(t->flat_segs_host.size() / sizeof(RowSeg<T>));
            std::memcpy(segs.data(), t->flat_segs_host.data(), t->flat_segs_host.size());
            total_w = t->total_width;
        } else {
            for (size_t m = 0; m < t->mats.size(); ++m) {
                segs.push_back(seg_of<H>(t->mats[m]));
                total_w += t->widths[m];
            }
        }
        rows.resize(uint64_t(nq) * segs.size());
        for (uint32_t q = 0; q < nq; ++q)
            for (size_t m = 0; m < segs.size(); ++m)
                rows[q * segs.size() + m] = t->flat ? idx[q] : idx[q] / (maxh / t->heights[m]);
    }
    const uint32_t steps = proof ? uint32_t(t->layers.size() - 1) : 0;
    std::vector<const T*> lptr(steps);
    std::vector<uint64_t> sib(uint64_t(nq) * steps);
    for (uint32_t l = 0; l < steps; ++l) {
        lptr[l] = dptr<T>(t->layers[l]);
        for (uint32_t q = 0; q < nq; ++q) {
            const uint64_t s = (idx[q] >> l) ^ 1;
            sib[uint64_t(q) * steps + l] = s < t->counts[l] ? s : ~0ull;
        }
    }
    if (total_w == 0 && steps == 0) return;
    // one table: segments | rows | layer pointers | siblings (16-byte aligned parts)
    auto up16 = [](size_t b) { return (b + 15) & ~size_t(15); };
    const size_t o_rows = up16(segs.size() * sizeof(RowSeg<T>));
    const size_t o_lay = o_rows + up16(rows.size() * sizeof(uint64_t));
    const size_t o_sib = o_lay + up16(lptr.size() * sizeof(const T*));
    const size_t tbytes = o_sib + sib.size() * sizeof(uint64_t);
    std::vector<uint8_t> tab(tbytes, 0);
    std::memcpy(tab.data(), segs.data(), segs.size() * sizeof(RowSeg<T>));
    std::memcpy(tab.data() + o_rows, rows.data(), rows.size() * sizeof(uint64_t));
    std::memcpy(tab.data() + o_lay, lptr.data(), lptr.size() * sizeof(const T*));
    std::memcpy(tab.data() + o_sib, sib.data(), sib.size() * sizeof(uint64_t));
    DevPtr dtab = dalloc(ctx, tbytes);
    h2d(ctx, dtab->ptr, tab.data(), tbytes);
    const uint64_t nvals = uint64_t(nq) * total_w, nproof = uint64_t(nq) * steps * DIG;
    DevPtr out = dalloc(ctx, (nvals + nproof) * sizeof(T));
    uint8_t* base = static_cast<uint8_t*>(dtab->ptr);
    launch_check(open_rows_device<T>(reinterpret_cast<const RowSeg<T>*>(base), uint32_t(segs.size()), total_w,
                                     reinterpret_cast<const uint64_t*>(base + o_rows),
                                     reinterpret_cast<const T* const*>(base + o_lay), steps, DIG,
                                     reinterpret_cast<const uint64_t*>(base + o_sib), nq, dptr<T>(out),
                                     dptr<T>(out) + nvals, ctx->stream),
                 "open rows");
    if (nvals)
        hip_check(hipMemcpyAsync(values, dptr<T>(out), nvals * sizeof(T), hipMemcpyDeviceToHost, ctx->stream),
                  "hipMemcpyAsync D2H");
    if (nproof)
        hip_check(hipMemcpyAsync(proof, dptr<T>(out) + nvals, nproof * sizeof(T), hipMemcpyDeviceToHost, ctx->stream),
                  "hipMemcpyAsync D2H");
    hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
}

// ------------------------------------------------------------------- FRI ---
const uint32_t* fri_tinv(p3h_ctx* ctx, uint32_t LH) {
    auto it = ctx->fri_tinv.find(LH);
    if (it != ctx->fri_tinv.end()) return dptr<uint32_t>(it->second);
    uint32_t bits = LH - 1;
    DevPtr buf = dalloc(ctx, (uint64_t(1) << bits) * sizeof(uint32_t));
    uint32_t base = bb_mont(bb_inv(bb_root(LH)));
    DevPtr bb = dalloc(ctx, sizeof(uint32_t));
    h2d(ctx, bb->ptr, &base, sizeof(uint32_t));
    // out[p] = (omega_{2^LH}^-1)^bitrev(p, LH-1) * (1/2)
    launch_check(lde_twist_device<BabyBear>(dptr<uint32_t>(buf), bits, 1, dptr<uint32_t>(bb), bb_mont(bb_inv(2)),
                                            bb_mont(1), ctx->stream),
                 "fri tinv");
    hip_check(hipStreamSynchronize(ctx->stream), "sync");
    ctx->fri_tinv[LH] = buf;
    ctx->table_bytes += buf->cls;
    return dptr<uint32_t>(buf);
}

// beta: canonical, host; or beta_dev (not null): canonical, device memory.
p3h_mat* do_fold(p3h_ctx* ctx, const p3h_mat* cur, uint32_t log_arity, const uint32_t* beta,
                 const uint32_t* beta_dev = nullptr) {
    if (cur->field != P3H_BABYBEAR || cur->width != 4 || cur->k != 0)
        throw Err{P3H_EINVAL, "fold input must be a natural-order (n x 4) BabyBear matrix of BabyBear^4 values"};
    uint32_t lh = log2_exact(cur->height);
    if (log_arity < 1 || log_arity > 3 || log_arity > lh) throw Err{P3H_EINVAL, "log_arity must be 1, 2 or 3"};
    if (lh > 27) throw Err{P3H_EINVAL, "height exceeds two-adicity"};
    if (!beta_dev)
        for (int q = 0; q < 4; ++q)
            if (beta[q] >= BB_P) throw Err{P3H_EINVAL, "beta is not canonical"};
    uint64_t n_out = cur->height >> log_arity;
    p3h_mat* out = new_mat(ctx, P3H_BABYBEAR, n_out, 4);
    try {
        const uint32_t* tinv = log_arity == 1 && lh >= 1 ? fri_tinv(ctx, lh) : nullptr;
        Phase ph(ctx, "fold");
        launch_check(fold_device(cur->data<uint32_t>(), out->data<uint32_t>(), n_out, lh, log_arity, beta, beta_dev,
                                 tinv,
                                 bb_mont(bb_root(lh)), bb_mont(bb_root(log_arity)),
                                 bb_mont(bb_pow(bb_inv(2), log_arity)), ctx->stream),
                     "fold");
    } catch (...) {
        mat_unref(out);
        throw;
    }
    return out;
}

// ------------------------------------------------------------ challenger ---
DevChal* chal_state(p3h_chal* ch) { return dptr<DevChal>(ch->st); }
uint32_t* chal_io(p3h_chal* ch) { return dptr<uint32_t>(ch->io); }
unsigned long long* chal_found(p3h_chal* ch) { return reinterpret_cast<unsigned long long*>(chal_io(ch) + 64); }

void chal_run(p3h_chal* ch, const ChalProgram& prog) {
    p3h_ctx* ctx = ch->ctx;
    if (!ctx->p2_set[P3H_BABYBEAR]) throw Err{P3H_ESTATE, "Poseidon2 constants for BabyBear not set"};
    launch_check(chal_run_device(chal_state(ch), prog, ctx->bbp64, ctx->stream), "challenger");
}

ChalOp chal_op(uint32_t code, uint32_t n = 0, uint64_t arg = 0, const uint32_t* src = nullptr,
               uint32_t* dst = nullptr) {
    return ChalOp{code, n, arg, src, dst};
}

// The least witness whose following sample has `bits` zero low bits
// (DuplexChallenger::grind, duplex_challenger.hpp:236-254), searched on the
// device over increasing witnesses; the search launch stops taking work once
// it has found the least one, so one launch covers a range of 2^36 witnesses
// (the chance that none of them passes at bits <= 30 is below e^-64); the
// transcript is not changed.
uint64_t chal_search(p3h_chal* ch, uint32_t bits) {
    p3h_ctx* ctx = ch->ctx;
    if (bits >= 31) throw Err{P3H_EINVAL, "proof-of-work bits must be below 31"};
    constexpr uint64_t RANGE = 1ull << 36;
    const unsigned long long none = ~0ull;
    unsigned long long* found = chal_found(ch);
    unsigned long long* next = found + 1;  // the work counter after the minimum
    for (uint64_t base = 0;; base += RANGE) {
        const unsigned long long init[2] = {none, 0ull};
        h2d(ctx, found, init, sizeof(init));
        launch_check(grind_search_device(chal_state(ch), base, RANGE, bits, found, next, ctx->bbp64, ctx->stream),
                     "grind");
        unsigned long long w = none;
        d2h(ctx, &w, found, sizeof(w));
        if (w != none) return w;
    }
}

// grind(bits): search, then observe the witness and consume one sample.
uint64_t chal_grind(p3h_chal* ch, uint32_t bits) {
    const uint64_t w = bits == 0 ? 0 : chal_search(ch, bits);  // with 0 bits every witness passes
    ChalProgram prog{};
    prog.op[0] = chal_op(CH_WITNESS, 0, w);
    prog.op[1] = chal_op(CH_SAMPLE, 1, 0, nullptr, chal_io(ch));
    prog.n_ops = 2;
    chal_run(ch, prog);
    return w;
}

p3h_mat* view_of(p3h_mat* base, uint64_t h, uint32_t w) {
    auto* v = new p3h_mat{};
    v->ctx = base->ctx;
    v->field = base->field;
    v->height = h;
    v->width = w;
    v->buf = base->buf;
    v->offset_elems = base->offset_elems;
    return v;
}

// Batched commit of many BabyBear traces: every buffer allocated and every
// table uploaded at planning time; the LDE passes of all traces run as
// work-list launches grouped by (mode, pass width) or, with
// P3H_BATCH_QUEUE_LDE, as one work-queue launch, and the tree as leaf hashing
// + one launch per level.  With P3H_BATCH_GRAPH the whole sequence is
// captured once as a HIP graph.
struct BatchPlan {
    p3h_ctx* ctx = nullptr;
    std::vector<p3h_mat*> traces;  // referenced
    std::vector<p3h_mat*> ldes;    // referenced (the tree holds its own)
    std::vector<LdePassGroup<uint32_t>> groups;
    std::vector<DevPtr> work_tables;
    bool queue = false;
    LdeQueuePlan<uint32_t> qplan;
    DevPtr q_items, q_ctr;
    p3h_tree* tree = nullptr;
    LaunchList tree_launches;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;

    void launch_all() {
        if (queue) {
            Phase ph(ctx, "lde");
            launch_check(launch_lde_queue<BabyBear>(qplan, dptr<NttQueueItem<uint32_t>>(q_items),
                                                    dptr<uint32_t>(q_ctr), ctx->stream),
                         "batched lde queue");
        } else {
            for (size_t g = 0; g < groups.size(); ++g) {
                Phase ph(ctx, "lde");
                launch_check(
                    launch_lde_group<BabyBear>(groups[g], dptr<NttWork<uint32_t>>(work_tables[g]), ctx->stream),
                    "batched lde");
            }
        }
        run_launches(ctx, tree_launches);
    }
    // after a run: the queue's error word (a wait that exceeded its spin limit)
    void check_queue() {
        if (!queue) return;
        uint32_t err = 0;
        d2h(ctx, &err, dptr<uint32_t>(q_ctr) + 1, sizeof(err));
        if (err) throw Err{P3H_EDEVICE, "batched LDE queue: a pass waited beyond its spin limit"};
    }
    ~BatchPlan() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (tree) p3h_tree_free(tree);
        for (auto* m : ldes) mat_unref(m);
        for (auto* m : traces) mat_unref(m);
    }
};

BatchPlan* make_batch_plan(p3h_ctx* ctx, const p3h_mat* const* traces, const uint32_t* added_bits, uint32_t n,
                           uint64_t shift, uint32_t cap_height, int flags) {
    const bool use_graph = (flags & P3H_BATCH_GRAPH) != 0;
    if (!traces || !added_bits || n == 0) throw Err{P3H_EINVAL, "null argument"};
    if (shift >= BB_P) throw Err{P3H_EINVAL, "shift is not a canonical field element"};
    auto* bp = new BatchPlan{};
    bp->ctx = ctx;
    try {
        std::vector<LdeJob<uint32_t>> jobs;
        for (uint32_t i = 0; i < n; ++i) {
            const p3h_mat* tr = traces[i];
            if (!tr || tr->field != P3H_BABYBEAR || tr->k != 0)
                throw Err{P3H_EINVAL, "batch traces must be natural-order BabyBear matrices"};
            if (tr->width == 0) throw Err{P3H_EINVAL, "matrix width is zero"};
            uint32_t L = log2_exact(tr->height), k = added_bits[i];
            if (L + k > 27) throw Err{P3H_EINVAL, "LDE height exceeds the field's two-adicity"};
            bp->traces.push_back(const_cast<p3h_mat*>(tr));
            mat_ref(const_cast<p3h_mat*>(tr));
            p3h_mat* out = new_mat(ctx, P3H_BABYBEAR, tr->height << k, tr->width);
            out->k = k;
            out->L = L;
            bp->ldes.push_back(out);
            jobs.push_back(LdeJob<uint32_t>{tr->data<uint32_t>(), out->data<uint32_t>(), L, tr->width, k,
                                            twiddles<HBB>(ctx, L), lde_twist<HBB>(ctx, L, k, uint32_t(shift)),
                                            tuning_for(ctx, P3H_BABYBEAR, L, tr->width, k)});
        }
        if ((flags & P3H_BATCH_QUEUE_LDE) && plan_lde_queue<BabyBear>(jobs, bp->qplan)) {
            bp->queue = true;
            const size_t ib = bp->qplan.items.size() * sizeof(NttQueueItem<uint32_t>);
            bp->q_items = dalloc(ctx, ib ? ib : 1);
            if (ib) h2d(ctx, bp->q_items->ptr, bp->qplan.items.data(), ib);
            bp->q_ctr = dalloc(ctx, lde_queue_counter_words(bp->qplan.items.size()) * sizeof(uint32_t));
        } else {
            bp->groups = plan_lde_batch<BabyBear>(jobs);
        }
        for (const auto& g : bp->groups) {
            DevPtr d = dalloc(ctx, g.work.size() * sizeof(NttWork<uint32_t>));
            h2d(ctx, d->ptr, g.work.data(), g.work.size() * sizeof(NttWork<uint32_t>));
            bp->work_tables.push_back(d);
        }
        std::vector<const p3h_mat*> lp(bp->ldes.begin(), bp->ldes.end());
        bp->tree = prepare_mmcs<HBB>(ctx, lp.data(), n, cap_height, bp->tree_launches);
        hip_check(hipStreamSynchronize(ctx->stream), "sync");
        if (use_graph) {
            ctx->capturing = true;
            hipError_t e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
            if (e == hipSuccess) {
                try {
                    bp->launch_all();
                } catch (...) {
                    hipGraph_t g = nullptr;
                    (void)hipStreamEndCapture(ctx->stream, &g);
                    if (g) (void)hipGraphDestroy(g);
                    ctx->capturing = false;
                    throw;
                }
                e = hipStreamEndCapture(ctx->stream, &bp->graph);
            }
            ctx->capturing = false;
            hip_check(e, "graph capture");
            hip_check(hipGraphInstantiate(&bp->exec, bp->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
        }
    } catch (...) {
        delete bp;
        throw;
    }
    return bp;
}

void run_batch_plan(BatchPlan* bp) {
    p3h_ctx* ctx = bp->ctx;
    if (bp->exec) {
        Phase ph(ctx, "batch_graph");
        hip_check(hipGraphLaunch(bp->exec, ctx->stream), "hipGraphLaunch");
    } else {
        bp->launch_all();
    }
}

template <class Fn>
int api(p3h_ctx* ctx, Fn&& fn) {
    try {
        fn();
        return P3H_OK;
    } catch (const Err& e) {
        if (ctx) ctx->err = e.msg;
        return e.code;
    } catch (const std::bad_alloc&) {
        if (ctx) ctx->err = "host allocation failed";
        return P3H_ENOMEM;
    } catch (const std::exception& e) {
        if (ctx) ctx->err = e.what();
        return P3H_EDEVICE;
    }
}

}  // namespace

// ================================================================== C ABI ===
namespace {

// prove_fri's commit loop (fri_prover.hpp:83-190) for one input codeword.
// Each round commits the current vector as a (len/arity) x (4*arity) matrix
// (ExtensionMmcs over the flat MMCS) and folds it with that round's beta:
// betas[4*round..] from the caller, or -- with a challenger -- the transcript
// observes the root, grinds pow_bits and samples beta on the device, and the
// fold reads beta from device memory, so no round waits for the host unless
// pow_bits > 0.  The transcript then observes the final polynomial's
// coefficients and every round's log arity (fri_prover.hpp:180-189).
// `extra`: further inputs of strictly decreasing heights below the
// codeword's; each is added, times beta^arity, to the folded vector of its
// height (fri_prover.hpp:123-135), and each round's arity stops at the next
// input's height (compute_log_arity_for_round, fri_params.hpp:36-57).
void do_fri_commit(p3h_ctx* ctx, const p3h_mat* codeword, uint32_t log_blowup, uint32_t log_final_poly_len,
                   uint32_t max_log_arity, const uint32_t* betas, p3h_chal* ch, uint32_t pow_bits,
                   uint32_t* roots_out, uint32_t* betas_out, uint64_t* witnesses_out, uint32_t* final_poly_out,
                   uint32_t* n_rounds, p3h_fri** keep, const std::vector<p3h_mat*>& extra = {}) {
    if (!codeword || !roots_out || !final_poly_out) throw Err{P3H_EINVAL, "null argument"};
    if (ch && (!betas_out || !witnesses_out)) throw Err{P3H_EINVAL, "null argument"};
    if (codeword->field != P3H_BABYBEAR || codeword->width != 4 || codeword->k != 0)
        throw Err{P3H_EINVAL, "codeword must be a natural-order (n x 4) BabyBear matrix"};
    if (max_log_arity < 1 || max_log_arity > 3) throw Err{P3H_EINVAL, "max_log_arity must be 1..3"};
    if (ch && pow_bits >= 31) throw Err{P3H_EINVAL, "proof-of-work bits must be below 31"};
    uint32_t lh0 = log2_exact(codeword->height);
    uint32_t lfh = log_blowup + log_final_poly_len;
    if (lh0 <= lfh) throw Err{P3H_EINVAL, "codeword must be longer than the final domain"};
    uint64_t prev_h = codeword->height;
    for (const p3h_mat* e : extra) {
        if (e->field != P3H_BABYBEAR || e->width != 4 || e->k != 0 || e->height >= prev_h ||
            log2_exact(e->height) <= lfh)
            throw Err{P3H_EINVAL, "FRI inputs must be natural (n x 4) matrices of decreasing heights above the final domain"};
        prev_h = e->height;
    }
    auto* f = new p3h_fri{};
    f->ctx = ctx;
    try {
        DevPtr betas_dev = ch ? dalloc(ctx, 4 * 32 * sizeof(uint32_t)) : DevPtr();
        std::vector<uint32_t> arities;
        p3h_mat* cur = const_cast<p3h_mat*>(codeword);
        mat_ref(cur);
        f->layers.push_back(cur);
        uint32_t round = 0;
        size_t next_in = 0;
        // With the challenges given, no layer depends on a root: every layer is
        // folded first and all layer trees are then built together, level by
        // level (binary64 engine).  With a transcript each root is observed
        // before the next challenge, so each tree is built in its round.
        const bool together = !ch && ctx->hash_engine == 1;
        std::vector<p3h_mat*> views;
        struct ViewRefs {
            std::vector<p3h_mat*>& v;
            ~ViewRefs() {
                for (p3h_mat* m : v) mat_unref(m);
            }
        } view_refs{views};
        while (cur->height > (1ull << lfh)) {
            uint32_t lh = log2_exact(cur->height);
            const bool has_next = next_in < extra.size();
            const uint32_t target = has_next ? log2_exact(extra[next_in]->height) : lfh;
            if (lh <= target) break;
            uint32_t la = std::min(lh - target, max_log_arity);
            uint32_t arity = 1u << la;
            p3h_mat* view = view_of(cur, cur->height >> la, 4 * arity);
            if (together) {
                views.push_back(view);
            } else {
                p3h_tree* t = nullptr;
                try {
                    t = flat_tree<HBB>(ctx, {seg_of<HBB>(view)}, view->height, view->width, {view});
                } catch (...) {
                    mat_unref(view);
                    throw;
                }
                mat_unref(view);  // the tree holds its own reference
                f->trees.push_back(t);
            }
            const uint32_t* beta_dev = nullptr;
            if (ch) {
                uint32_t* bd = dptr<uint32_t>(betas_dev) + 4 * round;
                const uint32_t* root = dptr<uint32_t>(f->trees.back()->layers.back());
                ChalProgram pr{};
                pr.op[pr.n_ops++] = chal_op(CH_OBSERVE, 8, 0, root);  // observe_commitment
                uint64_t w = 0;
                if (pow_bits > 0) {
                    chal_run(ch, pr);
                    w = chal_search(ch, pow_bits);
                    pr.n_ops = 0;
                }
                pr.op[pr.n_ops++] = chal_op(CH_WITNESS, 0, w);                       // grind: witness ...
                pr.op[pr.n_ops++] = chal_op(CH_SAMPLE, 1, 0, nullptr, chal_io(ch));  // ... and its sample
                pr.op[pr.n_ops++] = chal_op(CH_SAMPLE, 4, 0, nullptr, bd);           // sample_challenge
                chal_run(ch, pr);
                witnesses_out[round] = w;
                beta_dev = bd;
            }
            p3h_mat* nxt = do_fold(ctx, cur, la, ch ? nullptr : betas + 4 * round, beta_dev);
            f->layers.push_back(nxt);
            if (has_next && extra[next_in]->height == nxt->height) {
                launch_check(fri_mix_device(nxt->data<uint32_t>(), extra[next_in]->data<uint32_t>(), nxt->height,
                                            ch ? nullptr : betas + 4 * round, beta_dev, la, ctx->stream),
                             "fri mix");
                ++next_in;
            }
            cur = nxt;
            arities.push_back(la);
            ++round;
        }
        if (together) f->trees = flat_trees_multi(ctx, views);
        std::vector<Readback> rb;
        for (uint32_t r = 0; r < round; ++r) rb.push_back({roots_out + 8 * r, f->trees[r]->layers.back()->ptr, 32});
        if (ch) rb.push_back({betas_out, dptr<uint32_t>(betas_dev), 4 * round * sizeof(uint32_t)});
        // final polynomial: truncate to 2^log_final_poly_len, bit-reverse, inverse DFT (host, tiny)
        uint32_t n = 1u << log_final_poly_len;
        std::vector<uint32_t> last(uint64_t(cur->height) * 4);
        rb.push_back({last.data(), cur->data<uint32_t>(), last.size() * 4});
        d2h_gather(ctx, rb);
        std::vector<uint32_t> br(4 * n);
        for (uint32_t i = 0; i < n; ++i) {
            uint32_t j = 0;
            for (uint32_t b = 0; b < log_final_poly_len; ++b) j |= ((i >> b) & 1) << (log_final_poly_len - 1 - b);
            for (int q = 0; q < 4; ++q) br[4 * j + q] = last[4 * i + q];
        }
        uint32_t winv = bb_inv(bb_root(log_final_poly_len)), ninv = bb_inv(n);
        for (uint32_t k = 0; k < n; ++k)
            for (int q = 0; q < 4; ++q) {
                uint64_t acc = 0;
                for (uint32_t j = 0; j < n; ++j)
                    acc = (acc + uint64_t(bb_mul(br[4 * j + q], bb_pow(winv, uint64_t(k) * j)))) % BB_P;
                final_poly_out[4 * k + q] = bb_mul(uint32_t(acc), ninv);
            }
        if (ch) {
            // observe_challenge for each coefficient, then observe_arity for each round
            std::vector<uint32_t> obs(final_poly_out, final_poly_out + 4 * n);
            obs.insert(obs.end(), arities.begin(), arities.end());
            DevPtr d = dalloc(ctx, obs.size() * sizeof(uint32_t));
            h2d_async(ctx, d->ptr, obs.data(), obs.size() * sizeof(uint32_t));
            ChalProgram pr{};
            pr.op[0] = chal_op(CH_OBSERVE, uint32_t(obs.size()), 0, dptr<uint32_t>(d));
            pr.n_ops = 1;
            chal_run(ch, pr);
            hip_check(hipStreamSynchronize(ctx->stream), "sync");
        }
        if (n_rounds) *n_rounds = round;
    } catch (...) {
        p3h_fri_free(f);
        throw;
    }
    if (keep)
        *keep = f;
    else
        p3h_fri_free(f);
}

}  // namespace

namespace {

// n host states permuted on the device with the context's engine (canonical in
// and out).
void permute_host(p3h_ctx* ctx, int field, void* states, uint64_t n) {
    if (field == P3H_BABYBEAR) {
        const auto& pp = p2<HBB>(ctx);
        DevPtr d = dalloc(ctx, std::max<uint64_t>(1, n * 16 * 4));
        h2d(ctx, d->ptr, states, n * 16 * 4);
        if (ctx->hash_engine == 1)
            launch_check(permute_f64_device(dptr<uint32_t>(d), n, ctx->bbp64, ctx->stream), "permute");
        else
            launch_check(permute_device<BBDom>(dptr<uint32_t>(d), n, pp, ctx->stream), "permute");
        d2h(ctx, states, d->ptr, n * 16 * 4);
    } else if (field == P3H_GOLDILOCKS) {
        const auto& pp = p2<HGL>(ctx);
        DevPtr d = dalloc(ctx, std::max<uint64_t>(1, n * 8 * 8));
        h2d(ctx, d->ptr, states, n * 8 * 8);
        launch_check(permute_device<GLDom>(dptr<uint64_t>(d), n, pp, ctx->stream), "permute");
        d2h(ctx, states, d->ptr, n * 8 * 8);
    } else {
        throw Err{P3H_EINVAL, "unknown field"};
    }
}

// The sponge (PaddingFreeSponge::hash_iter_slices, padding_free_sponge.hpp:63-113)
// and the 2-to-1 compression (TruncatedPermutation, truncated_permutation.hpp:61-76)
// over host digests, each permutation run on the device.  The verifiers below
// walk one authentication path, so they are a short chain of dependent
// permutations.
template <class Raw, int W, int RATE, int DIG>
struct PathHasher {
    using Digest = std::array<Raw, DIG>;
    p3h_ctx* ctx;
    int field;
    Raw modulus;

    void perm(std::array<Raw, W>& s) const { permute_host(ctx, field, s.data(), 1); }
    // state = 0; each chunk of up to RATE elements overwrites the front of the
    // state and is followed by a permutation; no input -> zero digest.
    Digest hash(const std::vector<std::pair<const Raw*, uint64_t>>& parts) const {
        std::array<Raw, W> s{};
        uint32_t fill = 0;
        for (const auto& [ptr, len] : parts)
            for (uint64_t i = 0; i < len; ++i) {
                if (ptr[i] >= modulus) throw Err{P3H_EINVAL, "opened value is not a canonical field element"};
                s[fill++] = ptr[i];
                if (fill == RATE) {
                    perm(s);
                    fill = 0;
                }
            }
        if (fill) perm(s);
        Digest d;
        std::copy(s.begin(), s.begin() + DIG, d.begin());
        return d;
    }
    Digest compress(const Digest& l, const Digest& r) const {
        std::array<Raw, W> s{};
        std::copy(l.begin(), l.end(), s.begin());
        std::copy(r.begin(), r.end(), s.begin() + DIG);
        perm(s);
        Digest d;
        std::copy(s.begin(), s.begin() + DIG, d.begin());
        return d;
    }
    Digest digest_at(const void* p, uint64_t i) const {
        if (!p) throw Err{P3H_EINVAL, "null digest"};
        Digest d;
        std::memcpy(d.data(), static_cast<const Raw*>(p) + i * DIG, sizeof(d));
        for (Raw x : d)
            if (x >= modulus) throw Err{P3H_EINVAL, "digest element is not a canonical field element"};
        return d;
    }

    // MerkleTreeMmcs::verify_batch (merkle_tree_mmcs.hpp:168-254).
    bool verify_batch(const void* cap, uint32_t cap_len, const uint64_t* heights, const uint32_t* widths,
                      uint32_t n, uint64_t index, const Raw* values, const void* proof, uint32_t proof_len) const {
        if (n == 0) return false;
        std::vector<uint64_t> offset(n + 1, 0);
        for (uint32_t i = 0; i < n; ++i) offset[i + 1] = offset[i] + widths[i];
        std::vector<uint32_t> order(n);
        std::iota(order.begin(), order.end(), 0u);
        std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return heights[a] > heights[b]; });
        const uint64_t max_h = heights[order[0]];
        if (max_h == 0) return false;
        uint32_t levels = 0;
        for (uint64_t h = max_h; h > cap_len; h /= 2) ++levels;
        if (proof_len != levels) return false;
        size_t pos = 0;
        auto group = [&](uint64_t h) {
            std::vector<std::pair<const Raw*, uint64_t>> parts;
            while (pos < order.size() && heights[order[pos]] == h) {
                const uint32_t m = order[pos++];
                parts.push_back({values + offset[m], widths[m]});
            }
            return parts;
        };
        Digest cur = hash(group(max_h));
        uint64_t idx = index;
        for (uint32_t step = 0; step < levels; ++step) {
            const Digest sib = digest_at(proof, step);
            cur = (idx & 1) == 0 ? compress(cur, sib) : compress(sib, cur);
            idx >>= 1;
            const auto inject = group(max_h >> (step + 1));
            if (!inject.empty()) cur = compress(cur, hash(inject));
        }
        if (idx >= cap_len) return false;
        return cur == digest_at(cap, idx);
    }
    // p3_merkle::MerkleTreeMmcs::verify_row (fri_merkle_tree_mmcs.hpp:146-168).
    bool verify_row(const void* root, uint32_t n_root, uint64_t index, const Raw* row, uint32_t width,
                    const void* proof, uint32_t proof_len) const {
        Digest cur = hash({{row, width}});
        uint64_t idx = index;
        for (uint32_t step = 0; step < proof_len; ++step) {
            const Digest sib = digest_at(proof, step);
            cur = idx % 2 == 0 ? compress(cur, sib) : compress(sib, cur);
            idx >>= 1;
        }
        if (n_root == 0) return false;
        return cur == digest_at(root, 0);
    }
};

using BBPath = PathHasher<uint32_t, 16, 8, 8>;
using GLPath = PathHasher<uint64_t, 8, 4, 4>;

}  // namespace

extern "C" {

int p3h_ctx_create(int device, uint64_t budget, p3h_ctx** out) {
    if (!out) return P3H_EINVAL;
    *out = nullptr;
    auto* ctx = new p3h_ctx{};
    ctx->device = device;
    int rc = api(ctx, [&] {
        int n = 0;
        hip_check(hipGetDeviceCount(&n), "hipGetDeviceCount");
        if (device < 0 || device >= n) throw Err{P3H_EINVAL, "no such device"};
        hip_check(hipSetDevice(device), "hipSetDevice");
        hip_check(hipError_t(ntt_prepare_device()), "NTT kernel LDS limits");
        hip_check(hipError_t(fold_scratch_setup(device)), "hipGetSymbolAddress(scratch)");
        hip_check(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking), "hipStreamCreate");
        size_t fr = 0, tot = 0;
        hip_check(hipMemGetInfo(&fr, &tot), "hipMemGetInfo");
        ctx->budget = budget ? budget : tot;
        if (const char* f = std::getenv("P3H_FAIL_ALLOC_AFTER")) {
            char* end = nullptr;
            const long long v = std::strtoll(f, &end, 10);
            if (end != f && *end == '\0' && v >= 0) ctx->fail_after = v;
        }
    });
    if (rc != P3H_OK) {
        delete ctx;
        return rc;
    }
    *out = ctx;
    return P3H_OK;
}

int p3h_ctx_destroy(p3h_ctx* ctx) {
    if (!ctx) return P3H_EINVAL;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    resolve_pending(ctx);
    ctx->twiddles.clear();
    ctx->twists.clear();
    ctx->row_scales.clear();
    ctx->fri_tinv.clear();
    trim(ctx);
    if (ctx->stage) (void)hipHostFree(ctx->stage);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return P3H_OK;
}

const char* p3h_last_error(const p3h_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int p3h_ctx_sync(p3h_ctx* ctx) {
    return api(ctx, [&] { hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize"); });
}

int p3h_ctx_memory(const p3h_ctx* ctx, uint64_t* in_use, uint64_t* budget, uint64_t* cached) {
    if (!ctx) return P3H_EINVAL;
    if (in_use) *in_use = ctx->in_use;
    if (budget) *budget = ctx->budget;
    if (cached) *cached = ctx->cached;
    return P3H_OK;
}

int p3h_host_alloc(uint64_t bytes, void** out) {
    if (!out) return P3H_EINVAL;
    *out = nullptr;
    if (bytes == 0) return P3H_EINVAL;
    void* p = nullptr;
    if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess || !p) {
        (void)hipGetLastError();  // clear the sticky error of the failed allocation
        return P3H_ENOMEM;
    }
    *out = p;
    return P3H_OK;
}

int p3h_host_free(void* p) {
    if (!p) return P3H_OK;
    return hipHostFree(p) == hipSuccess ? P3H_OK : P3H_EDEVICE;
}

int p3h_seeded_fill(int field, uint64_t seed, void* out, uint64_t n) {
    if (!out && n) return P3H_EINVAL;
    std::mt19937_64 rng(seed);
    if (field == P3H_BABYBEAR) {
        uint32_t* o = static_cast<uint32_t*>(out);
        for (uint64_t i = 0; i < n; ++i) o[i] = uint32_t(rng() % 0x78000001ull);
    } else if (field == P3H_GOLDILOCKS) {
        uint64_t* o = static_cast<uint64_t*>(out);
        for (uint64_t i = 0; i < n; ++i) o[i] = rng() % 0xFFFFFFFF00000001ull;
    } else {
        return P3H_EINVAL;
    }
    return P3H_OK;
}

int p3h_probe_butterflies(p3h_ctx* ctx, int field, uint32_t blocks, uint32_t iters, double* ms,
                          double* butterflies) {
    return api(ctx, [&] {
        if (!ms || !butterflies || blocks == 0 || iters == 0) throw Err{P3H_EINVAL, "bad probe arguments"};
        if (field != P3H_BABYBEAR && field != P3H_GOLDILOCKS) throw Err{P3H_EINVAL, "unknown field"};
        const size_t eb = field == P3H_BABYBEAR ? 4 : 8;
        DevPtr out = dalloc(ctx, size_t(blocks) * 256 * eb), tw = dalloc(ctx, 32 * eb);
        uint64_t host[32];
        for (int i = 0; i < 32; ++i) host[i] = 1000003ull * uint64_t(i + 1) + 17;
        if (field == P3H_BABYBEAR) {
            uint32_t h32[32];
            for (int i = 0; i < 32; ++i) h32[i] = bb_mont(uint32_t(host[i] % BB_P));
            h2d(ctx, tw->ptr, h32, sizeof(h32));
        } else {
            h2d(ctx, tw->ptr, host, sizeof(host));
        }
        auto launch = [&](uint32_t it) {
            launch_check(field == P3H_BABYBEAR
                             ? butterfly_probe_device<BabyBear>(dptr<uint32_t>(out), dptr<uint32_t>(tw), blocks, it,
                                                                ctx->stream)
                             : butterfly_probe_device<Goldilocks>(dptr<uint64_t>(out), dptr<uint64_t>(tw), blocks,
                                                                  it, ctx->stream),
                         "butterfly probe");
        };
        launch(2);  // warm the clocks and the code object
        hipEvent_t a, b;
        hip_check(hipEventCreate(&a), "hipEventCreate");
        hip_check(hipEventCreate(&b), "hipEventCreate");
        hip_check(hipEventRecord(a, ctx->stream), "hipEventRecord");
        launch(iters);
        hip_check(hipEventRecord(b, ctx->stream), "hipEventRecord");
        hip_check(hipEventSynchronize(b), "hipEventSynchronize");
        float t = 0;
        hip_check(hipEventElapsedTime(&t, a, b), "hipEventElapsedTime");
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        *ms = t;
        *butterflies = double(blocks) * 256.0 * double(iters) * 32.0;
    });
}

int p3h_ctx_table_bytes(const p3h_ctx* ctx, uint64_t* bytes) {
    if (!ctx || !bytes) return P3H_EINVAL;
    *bytes = ctx->table_bytes;
    return P3H_OK;
}

int p3h_ctx_trim(p3h_ctx* ctx) {
    return api(ctx, [&] { trim(ctx); });
}

int p3h_ctx_set_profiling(p3h_ctx* ctx, int on) {
    return api(ctx, [&] {
        hip_check(hipStreamSynchronize(ctx->stream), "sync");
        resolve_pending(ctx);
        ctx->stats.clear();
        ctx->profiling = on != 0;
    });
}

int p3h_phase_stats(const p3h_ctx* cctx, const char* phase, uint64_t* launches, double* total_ms) {
    auto* ctx = const_cast<p3h_ctx*>(cctx);
    return api(ctx, [&] {
        resolve_pending(ctx);
        auto it = ctx->stats.find(phase ? phase : "");
        if (launches) *launches = it == ctx->stats.end() ? 0 : it->second.launches;
        if (total_ms) *total_ms = it == ctx->stats.end() ? 0 : it->second.ms;
    });
}

int p3h_ctx_set_ntt_tuning(p3h_ctx* ctx, uint32_t log_cols, uint32_t max_stages, uint32_t log_slab) {
    return api(ctx, [&] {
        if (log_cols > 10 || max_stages > 13 || log_slab > 16) throw Err{P3H_EINVAL, "tuning out of range"};
        ctx->tune.log_cols = log_cols;
        ctx->tune.max_stages = max_stages;
        ctx->tune.log_slab = log_slab;
    });
}

int p3h_ctx_set_hash_engine(p3h_ctx* ctx, int engine) {
    return api(ctx, [&] {
        if (engine != 0 && engine != 1) throw Err{P3H_EINVAL, "hash engine must be 0 (uint32) or 1 (binary64)"};
        ctx->hash_engine = engine;
    });
}

// MDS-light of the width-16 BabyBear Poseidon2 (4x4 circulant blocks, then
// each lane plus the sum of its column over the blocks) applied to integers.
static void bb_mds_light_int(const int64_t* x, int64_t* y) {
    for (int b = 0; b < 16; b += 4) {
        const int64_t* v = x + b;
        const int64_t t01 = v[0] + v[1], t23 = v[2] + v[3], t0123 = t01 + t23;
        y[b] = t0123 + v[1] + t01;
        y[b + 1] = t0123 + v[1] + 2 * v[2];
        y[b + 2] = t0123 + v[3] + t23;
        y[b + 3] = t0123 + v[3] + 2 * v[0];
    }
    int64_t sums[4];
    for (int k = 0; k < 4; ++k) sums[k] = y[k] + y[k + 4] + y[k + 8] + y[k + 12];
    for (int i = 0; i < 16; ++i) y[i] += sums[i & 3];
}

// x with MDS-light(x) = v (mod p), canonical, by Gauss-Jordan elimination over
// F_p on the matrix whose column j is MDS-light of the j-th unit vector
// (determinant 5^4 35^4, invertible mod p).
static void bb_mds_light_solve(const uint32_t* v, uint32_t* x) {
    uint32_t a[16][17];
    for (int j = 0; j < 16; ++j) {
        int64_t e[16] = {}, col[16];
        e[j] = 1;
        bb_mds_light_int(e, col);
        for (int i = 0; i < 16; ++i) a[i][j] = uint32_t(col[i] % BB_P);
    }
    for (int i = 0; i < 16; ++i) a[i][16] = v[i] % BB_P;
    for (int c = 0; c < 16; ++c) {
        int piv = c;
        while (a[piv][c] == 0) ++piv;
        if (piv != c)
            for (int k = 0; k < 17; ++k) std::swap(a[piv][k], a[c][k]);
        const uint32_t inv = bb_inv(a[c][c]);
        for (int k = 0; k < 17; ++k) a[c][k] = bb_mul(a[c][k], inv);
        for (int i = 0; i < 16; ++i) {
            if (i == c || a[i][c] == 0) continue;
            const uint32_t f = a[i][c];
            for (int k = 0; k < 17; ++k) a[i][k] = uint32_t((a[i][k] + uint64_t(BB_P - f) * a[c][k]) % BB_P);
        }
    }
    for (int i = 0; i < 16; ++i) x[i] = a[i][16];
}

// The binary64 engine's constants from the reference layout
// initial[hf][16], terminal[hf][16], internal[rp] (any representatives).
// Besides the plain tables, mc_init[r] / mc_term[r] hold MAGIC + c with c the
// balanced residue of MDS-light^-1 of the constants of the round after round r
// (init r + 1, then partial round 0 on lane 0; term r + 1; none after the
// last), which the engine adds in round r's S-boxes: MDS-light(y + c) = MDS-light(y)
// + next constants (mod p).
static P2ParamsF64 bb_f64_params(const uint32_t* c, uint32_t hf, uint32_t rp) {
    P2ParamsF64 q{};
    uint32_t init[P2_MAX_HALF_FULL][16], term[P2_MAX_HALF_FULL][16];
    size_t k = 0;
    for (uint32_t r = 0; r < hf; ++r)
        for (int i = 0; i < 16; ++i) init[r][i] = c[k++] % BB_P;
    for (uint32_t r = 0; r < hf; ++r)
        for (int i = 0; i < 16; ++i) term[r][i] = c[k++] % BB_P;
    for (uint32_t r = 0; r < hf; ++r)
        for (int i = 0; i < 16; ++i) {
            q.init[r][i] = double(init[r][i]);
            q.term[r][i] = double(term[r][i]);
        }
    for (uint32_t r = 0; r < rp; ++r) q.internal[r] = double(c[k++] % BB_P);
    auto fold = [](const uint32_t* next, double* mc) {
        uint32_t x[16] = {};
        if (next) bb_mds_light_solve(next, x);
        for (int i = 0; i < 16; ++i) {
            const int64_t bal = x[i] > BB_P / 2 ? int64_t(x[i]) - int64_t(BB_P) : int64_t(x[i]);
            mc[i] = BBF64::MAGIC + double(bal);
        }
    };
    uint32_t p0[16] = {};
    if (rp) p0[0] = uint32_t(q.internal[0]);
    for (uint32_t r = 0; r < hf; ++r) {
        fold(r + 1 < hf ? init[r + 1] : (rp ? p0 : nullptr), q.mc_init[r]);
        fold(r + 1 < hf ? term[r + 1] : nullptr, q.mc_term[r]);
    }
    for (int i = 0; i < 16; ++i) q.diag[i] = double(1u << i);
    q.half_full = hf;
    q.partial = rp;
    return q;
}

int p3h_poseidon2_set(p3h_ctx* ctx, int field, uint32_t width, uint32_t rf, uint32_t rp, const void* consts) {
    return api(ctx, [&] {
        if (!consts) throw Err{P3H_EINVAL, "null constants"};
        if (rf % 2 || rf / 2 > uint32_t(P2_MAX_HALF_FULL) || rp > uint32_t(P2_MAX_PARTIAL))
            throw Err{P3H_EINVAL, "round counts out of range (rounds_f even <= 8, rounds_p <= 32)"};
        if (field == P3H_BABYBEAR) {
            if (width != 16) throw Err{P3H_EINVAL, "BabyBear Poseidon2 width must be 16"};
            const uint32_t* c = static_cast<const uint32_t*>(consts);
            P2Params<uint32_t, 16> p{};
            size_t k = 0;
            for (uint32_t r = 0; r < rf / 2; ++r)
                for (int i = 0; i < 16; ++i) p.init[r][i] = bb_mont(c[k++] % BB_P);
            for (uint32_t r = 0; r < rf / 2; ++r)
                for (int i = 0; i < 16; ++i) p.term[r][i] = bb_mont(c[k++] % BB_P);
            for (uint32_t r = 0; r < rp; ++r) p.internal[r] = bb_mont(c[k++] % BB_P);
            for (int i = 0; i < 16; ++i) p.diag[i] = bb_mont(bb_pow(2, i));
            p.half_full = rf / 2;
            p.partial = rp;
            ctx->bbp = p;
            ctx->bbp64 = bb_f64_params(c, rf / 2, rp);
            ctx->p2_set[P3H_BABYBEAR] = true;
        } else if (field == P3H_GOLDILOCKS) {
            if (width != 8) throw Err{P3H_EINVAL, "Goldilocks Poseidon2 width must be 8"};
            const uint64_t* c = static_cast<const uint64_t*>(consts);
            P2Params<uint64_t, 8> p{};
            size_t k = 0;
            for (uint32_t r = 0; r < rf / 2; ++r)
                for (int i = 0; i < 8; ++i) p.init[r][i] = c[k++] % GL_P;
            for (uint32_t r = 0; r < rf / 2; ++r)
                for (int i = 0; i < 8; ++i) p.term[r][i] = c[k++] % GL_P;
            for (uint32_t r = 0; r < rp; ++r) p.internal[r] = c[k++] % GL_P;
            for (int i = 0; i < 8; ++i) p.diag[i] = gl_pow(2, i);
            p.half_full = rf / 2;
            p.partial = rp;
            ctx->glp = p;
            ctx->p2_set[P3H_GOLDILOCKS] = true;
        } else {
            throw Err{P3H_EINVAL, "unknown field"};
        }
    });
}

int p3h_poseidon2_permute(p3h_ctx* ctx, int field, void* states, uint64_t n) {
    return api(ctx, [&] {
        if (!states && n) throw Err{P3H_EINVAL, "null states"};
        permute_host(ctx, field, states, n);
    });
}

int p3h_mmcs_verify(p3h_ctx* ctx, int field, const void* cap, uint32_t cap_len, const uint64_t* heights,
                    const uint32_t* widths, uint32_t n_mats, uint64_t index, const void* values, const void* proof,
                    uint32_t proof_len, int* ok) {
    return api(ctx, [&] {
        if (!ok || (n_mats && (!heights || !widths || !values))) throw Err{P3H_EINVAL, "null argument"};
        *ok = 0;
        bool r;
        if (field == P3H_BABYBEAR)
            r = BBPath{ctx, field, uint32_t(BB_P)}.verify_batch(cap, cap_len, heights, widths, n_mats, index,
                                                                 static_cast<const uint32_t*>(values), proof,
                                                                 proof_len);
        else if (field == P3H_GOLDILOCKS)
            r = GLPath{ctx, field, GL_P}.verify_batch(cap, cap_len, heights, widths, n_mats, index,
                                                      static_cast<const uint64_t*>(values), proof, proof_len);
        else
            throw Err{P3H_EINVAL, "unknown field"};
        *ok = r ? 1 : 0;
    });
}

int p3h_flat_verify(p3h_ctx* ctx, int field, const void* root, uint32_t n_root, uint64_t index, const void* row,
                    uint32_t width, const void* proof, uint32_t proof_len, int* ok) {
    return api(ctx, [&] {
        if (!ok || (width && !row)) throw Err{P3H_EINVAL, "null argument"};
        *ok = 0;
        bool r;
        if (field == P3H_BABYBEAR)
            r = BBPath{ctx, field, uint32_t(BB_P)}.verify_row(root, n_root, index, static_cast<const uint32_t*>(row),
                                                               width, proof, proof_len);
        else if (field == P3H_GOLDILOCKS)
            r = GLPath{ctx, field, GL_P}.verify_row(root, n_root, index, static_cast<const uint64_t*>(row), width,
                                                    proof, proof_len);
        else
            throw Err{P3H_EINVAL, "unknown field"};
        *ok = r ? 1 : 0;
    });
}

int p3h_mat_upload(p3h_ctx* ctx, int field, const void* host, uint64_t h, uint32_t w, p3h_mat** out) {
    return api(ctx, [&] {
        if (!out || (!host && h * w)) throw Err{P3H_EINVAL, "null argument"};
        if (field != P3H_BABYBEAR && field != P3H_GOLDILOCKS) throw Err{P3H_EINVAL, "unknown field"};
        p3h_mat* m = new_mat(ctx, field, h, w);
        try {
            h2d(ctx, m->data<char>(), host, h * w * elem_size(field));
        } catch (...) {
            mat_unref(m);
            throw;
        }
        *out = m;
    });
}

int p3h_mat_alloc(p3h_ctx* ctx, int field, uint64_t h, uint32_t w, p3h_mat** out) {
    return api(ctx, [&] {
        if (!out) throw Err{P3H_EINVAL, "null argument"};
        if (field != P3H_BABYBEAR && field != P3H_GOLDILOCKS) throw Err{P3H_EINVAL, "unknown field"};
        *out = new_mat(ctx, field, h, w);
    });
}

int p3h_mat_download_rows(p3h_ctx* ctx, const p3h_mat* m, uint64_t row0, uint64_t nrows, void* host) {
    return api(ctx, [&] {
        if (!m || (!host && nrows)) throw Err{P3H_EINVAL, "null argument"};
        if (row0 + nrows > m->height) throw Err{P3H_EINVAL, "row range out of bounds"};
        size_t es = elem_size(m->field);
        size_t row_bytes = size_t(m->width) * es;
        if (m->k == 0) {
            d2h(ctx, host, m->data<char>() + row0 * row_bytes, nrows * row_bytes);
            return;
        }
        DevPtr tmp = dalloc(ctx, m->height * row_bytes);
        if (m->field == P3H_BABYBEAR)
            launch_check(coset_to_natural_device<uint32_t>(m->data<uint32_t>(), dptr<uint32_t>(tmp), m->L, m->k,
                                                           m->width, ctx->stream),
                         "coset_to_natural");
        else
            launch_check(coset_to_natural_device<uint64_t>(m->data<uint64_t>(), dptr<uint64_t>(tmp), m->L, m->k,
                                                           m->width, ctx->stream),
                         "coset_to_natural");
        d2h(ctx, host, static_cast<char*>(tmp->ptr) + row0 * row_bytes, nrows * row_bytes);
    });
}

int p3h_mat_download(p3h_ctx* ctx, const p3h_mat* m, void* host) {
    if (!m) return P3H_EINVAL;
    return p3h_mat_download_rows(ctx, m, 0, m->height, host);
}

int p3h_mat_info(const p3h_mat* m, int* field, uint64_t* h, uint32_t* w) {
    if (!m) return P3H_EINVAL;
    if (field) *field = m->field;
    if (h) *h = m->height;
    if (w) *w = m->width;
    return P3H_OK;
}

int p3h_mat_device_ptr(const p3h_mat* m, void** out) {
    if (!m || !out) return P3H_EINVAL;
    *out = m->field == P3H_BABYBEAR ? static_cast<void*>(m->data<uint32_t>()) : static_cast<void*>(m->data<uint64_t>());
    return P3H_OK;
}

int p3h_mat_free(p3h_mat* m) {
    if (!m) return P3H_EINVAL;
    mat_unref(m);
    return P3H_OK;
}

int p3h_dft(p3h_ctx* ctx, int kind, const p3h_mat* in, uint64_t shift, p3h_mat** out) {
    return api(ctx, [&] {
        if (!in || !out) throw Err{P3H_EINVAL, "null argument"};
        *out = in->field == P3H_BABYBEAR ? do_dft<HBB>(ctx, kind, in, shift) : do_dft<HGL>(ctx, kind, in, shift);
    });
}

int p3h_coset_lde(p3h_ctx* ctx, const p3h_mat* in, uint32_t added_bits, uint64_t shift, p3h_mat** out) {
    return api(ctx, [&] {
        if (!in || !out) throw Err{P3H_EINVAL, "null argument"};
        *out = in->field == P3H_BABYBEAR ? do_lde<HBB>(ctx, in, added_bits, shift)
                                         : do_lde<HGL>(ctx, in, added_bits, shift);
    });
}

int p3h_mmcs_commit(p3h_ctx* ctx, const p3h_mat* const* mats, uint32_t n, uint32_t cap_height, p3h_tree** out) {
    return api(ctx, [&] {
        if (!mats || !out || n == 0 || !mats[0]) throw Err{P3H_EINVAL, "null argument"};
        *out = mats[0]->field == P3H_BABYBEAR ? do_mmcs_commit<HBB>(ctx, mats, n, cap_height)
                                              : do_mmcs_commit<HGL>(ctx, mats, n, cap_height);
    });
}

int p3h_flat_commit(p3h_ctx* ctx, const p3h_mat* m, p3h_tree** out) {
    return api(ctx, [&] {
        if (!m || !out) throw Err{P3H_EINVAL, "null argument"};
        std::vector<p3h_mat*> refs{const_cast<p3h_mat*>(m)};
        if (m->field == P3H_BABYBEAR)
            *out = flat_tree<HBB>(ctx, {seg_of<HBB>(m)}, m->height, m->width, refs);
        else
            *out = flat_tree<HGL>(ctx, {seg_of<HGL>(m)}, m->height, m->width, refs);
    });
}

int p3h_tree_cap(p3h_ctx* ctx, const p3h_tree* t, void* cap_out, uint32_t* cap_len) {
    return api(ctx, [&] {
        if (!t) throw Err{P3H_EINVAL, "null tree"};
        uint64_t cnt = t->counts.back();
        size_t dig_bytes = t->field == P3H_BABYBEAR ? 32 : 32;
        if (cap_len) *cap_len = uint32_t(cnt);
        if (cap_out) d2h(ctx, cap_out, t->layers.back()->ptr, cnt * dig_bytes);
    });
}

int p3h_tree_info(const p3h_tree* t, uint32_t* n_layers, uint64_t* n_leaves, uint32_t* total_width) {
    if (!t) return P3H_EINVAL;
    if (n_layers) *n_layers = uint32_t(t->layers.size());
    if (n_leaves) *n_leaves = t->counts[0];
    if (total_width) *total_width = t->total_width;
    return P3H_OK;
}

int p3h_tree_matrices(const p3h_tree* t, uint32_t* n_mats, uint64_t* heights, uint32_t* widths) {
    if (!t || !n_mats) return P3H_EINVAL;
    *n_mats = uint32_t(t->widths.size());
    for (size_t i = 0; i < t->widths.size(); ++i) {
        if (heights) heights[i] = t->heights[i];
        if (widths) widths[i] = t->widths[i];
    }
    return P3H_OK;
}

int p3h_tree_layer(p3h_ctx* ctx, const p3h_tree* t, uint32_t layer, void* out) {
    return api(ctx, [&] {
        if (!t || !out || layer >= t->layers.size()) throw Err{P3H_EINVAL, "bad layer"};
        d2h(ctx, out, t->layers[layer]->ptr, t->counts[layer] * 32);
    });
}

int p3h_tree_open(p3h_ctx* ctx, const p3h_tree* t, const uint64_t* idx, uint32_t n, void* values, void* proof) {
    return api(ctx, [&] {
        if (!t || (!idx && n)) throw Err{P3H_EINVAL, "null argument"};
        if (t->field == P3H_BABYBEAR)
            tree_open<HBB>(ctx, t, idx, n, static_cast<uint32_t*>(values), static_cast<uint32_t*>(proof));
        else
            tree_open<HGL>(ctx, t, idx, n, static_cast<uint64_t*>(values), static_cast<uint64_t*>(proof));
    });
}

int p3h_tree_free(p3h_tree* t) {
    if (!t) return P3H_EINVAL;
    for (auto* m : t->mats) mat_unref(m);
    delete t;
    return P3H_OK;
}

int p3h_fri_fold(p3h_ctx* ctx, const p3h_mat* cur, uint32_t log_arity, const uint32_t* beta, p3h_mat** out) {
    return api(ctx, [&] {
        if (!cur || !beta || !out) throw Err{P3H_EINVAL, "null argument"};
        *out = do_fold(ctx, cur, log_arity, beta);
    });
}

int p3h_fri_fold_device(p3h_ctx* ctx, const uint32_t* d_in, uint32_t* d_out, const uint32_t* beta, uint32_t log_arity,
                        uint32_t log_height, uint64_t n) {
    return api(ctx, [&] {
        if (!beta || (n && (!d_in || !d_out))) throw Err{P3H_EINVAL, "null argument"};
        if (log_arity < 1 || log_arity > 3) throw Err{P3H_EINVAL, "log_arity must be 1, 2, or 3"};
        if (log_height < log_arity || log_height > 27)
            throw Err{P3H_EINVAL, "log_height must lie in [log_arity, 27]"};
        if (n > (uint64_t(1) << (log_height - log_arity)))
            throw Err{P3H_EINVAL, "n exceeds the folded height 2^(log_height - log_arity)"};
        for (int q = 0; q < 4; ++q)
            if (beta[q] >= BB_P) throw Err{P3H_EINVAL, "beta is not canonical"};
        if (n == 0) return;
        const uint32_t* tinv = log_arity == 1 ? fri_tinv(ctx, log_height) : nullptr;
        Phase ph(ctx, "fold");
        launch_check(fold_device(d_in, d_out, n, log_height, log_arity, beta, nullptr, tinv,
                                 bb_mont(bb_root(log_height)), bb_mont(bb_root(log_arity)),
                                 bb_mont(bb_pow(bb_inv(2), log_arity)), ctx->stream),
                     "fold");
        hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    });
}

int p3h_fri_fold_device_noalloc(p3h_ctx* ctx, const uint32_t* d_in, uint32_t* d_out, const uint32_t* beta,
                                uint32_t log_arity, uint32_t log_height, uint64_t n) {
    return api(ctx, [&] {
        if (!ctx || !beta || (n && (!d_in || !d_out))) throw Err{P3H_EINVAL, "null argument"};
        if (log_arity < 1 || log_arity > 3) throw Err{P3H_EINVAL, "log_arity must be 1, 2, or 3"};
        if (log_height < log_arity || log_height > 27)
            throw Err{P3H_EINVAL, "log_height must lie in [log_arity, 27]"};
        if (n > (uint64_t(1) << (log_height - log_arity)))
            throw Err{P3H_EINVAL, "n exceeds the folded height 2^(log_height - log_arity)"};
        if (log_arity == 1 && n > fold_scratch_words())
            throw Err{P3H_EINVAL, "arity-2 fold of more than 2^22 rows: beyond the allocation-free scratch"};
        for (int q = 0; q < 4; ++q)
            if (beta[q] >= BB_P) throw Err{P3H_EINVAL, "beta is not canonical"};
        if (n == 0) return;
        uint32_t* sp = fold_scratch_ptr(ctx->device);
        if (!sp) throw Err{P3H_ESTATE, "scratch address not set up for this device"};
        launch_check(fold_noalloc_device(d_in, d_out, n, log_height, log_arity, beta, sp,
                                         bb_mont(bb_inv(bb_root(log_height))), bb_mont(bb_inv(2)),
                                         bb_mont(bb_root(log_height)), bb_mont(bb_root(log_arity)),
                                         bb_mont(bb_pow(bb_inv(2), log_arity)), ctx->stream),
                     "fold (no allocation)");
    });
}

int p3h_fri_commit_phase(p3h_ctx* ctx, const p3h_mat* codeword, uint32_t log_blowup, uint32_t log_final_poly_len,
                         uint32_t max_log_arity, const uint32_t* betas, uint32_t* roots_out,
                         uint32_t* final_poly_out, uint32_t* n_rounds, p3h_fri** keep) {
    return api(ctx, [&] {
        if (!betas) throw Err{P3H_EINVAL, "null argument"};
        do_fri_commit(ctx, codeword, log_blowup, log_final_poly_len, max_log_arity, betas, nullptr, 0, roots_out,
                      nullptr, nullptr, final_poly_out, n_rounds, keep);
    });
}

int p3h_fri_commit_phase_chal(p3h_ctx* ctx, const p3h_mat* codeword, uint32_t log_blowup,
                              uint32_t log_final_poly_len, uint32_t max_log_arity, p3h_chal* ch, uint32_t pow_bits,
                              uint32_t* roots_out, uint32_t* betas_out, uint64_t* witnesses_out,
                              uint32_t* final_poly_out, uint32_t* n_rounds, p3h_fri** keep) {
    return api(ctx, [&] {
        if (!ch || ch->ctx != ctx) throw Err{P3H_EINVAL, "challenger missing or of another context"};
        do_fri_commit(ctx, codeword, log_blowup, log_final_poly_len, max_log_arity, nullptr, ch, pow_bits,
                      roots_out, betas_out, witnesses_out, final_poly_out, n_rounds, keep);
    });
}

// ------------------------------------------------------------ challenger ---
int p3h_challenger_create(p3h_ctx* ctx, p3h_chal** out) {
    return api(ctx, [&] {
        if (!out) throw Err{P3H_EINVAL, "null argument"};
        auto* ch = new p3h_chal{ctx, dalloc(ctx, sizeof(DevChal)), dalloc(ctx, 64 * sizeof(uint32_t) + 16)};
        const DevChal zero{};
        try {
            h2d(ctx, ch->st->ptr, &zero, sizeof(zero));
        } catch (...) {
            delete ch;
            throw;
        }
        *out = ch;
    });
}

int p3h_challenger_clone(const p3h_chal* ch, p3h_chal** out) {
    if (!ch) return P3H_EINVAL;
    p3h_ctx* ctx = ch->ctx;
    return api(ctx, [&] {
        if (!out) throw Err{P3H_EINVAL, "null argument"};
        auto* c = new p3h_chal{ctx, dalloc(ctx, sizeof(DevChal)), dalloc(ctx, 64 * sizeof(uint32_t) + 16)};
        if (hipMemcpyAsync(c->st->ptr, ch->st->ptr, sizeof(DevChal), hipMemcpyDeviceToDevice, ctx->stream) !=
            hipSuccess) {
            delete c;
            throw Err{P3H_EDEVICE, "challenger copy failed"};
        }
        *out = c;
    });
}

int p3h_challenger_free(p3h_chal* ch) {
    if (!ch) return P3H_EINVAL;
    delete ch;
    return P3H_OK;
}

int p3h_challenger_observe(p3h_chal* ch, const uint32_t* values, uint32_t n) {
    if (!ch) return P3H_EINVAL;
    return api(ch->ctx, [&] {
        if (n == 0) return;  // an empty slice changes nothing (duplex_challenger.hpp:105-108)
        if (!values) throw Err{P3H_EINVAL, "null argument"};
        DevPtr d = dalloc(ch->ctx, uint64_t(n) * sizeof(uint32_t));
        h2d(ch->ctx, d->ptr, values, uint64_t(n) * sizeof(uint32_t));
        ChalProgram pr{};
        pr.op[pr.n_ops++] = chal_op(CH_OBSERVE, n, 0, dptr<uint32_t>(d));
        chal_run(ch, pr);
    });
}

int p3h_challenger_observe_cap(p3h_chal* ch, const p3h_tree* t) {
    if (!ch) return P3H_EINVAL;
    return api(ch->ctx, [&] {
        if (!t || t->field != P3H_BABYBEAR || t->ctx != ch->ctx)
            throw Err{P3H_EINVAL, "tree missing, not BabyBear, or of another context"};
        ChalProgram pr{};
        pr.op[pr.n_ops++] = chal_op(CH_OBSERVE, uint32_t(t->counts.back() * 8), 0, dptr<uint32_t>(t->layers.back()));
        chal_run(ch, pr);
    });
}

int p3h_challenger_sample(p3h_chal* ch, uint32_t n, uint32_t* out) {
    if (!ch) return P3H_EINVAL;
    return api(ch->ctx, [&] {
        if (n && !out) throw Err{P3H_EINVAL, "null argument"};
        for (uint32_t off = 0; off < n; off += 64) {
            const uint32_t m = std::min(64u, n - off);
            ChalProgram pr{};
            pr.op[pr.n_ops++] = chal_op(CH_SAMPLE, m, 0, nullptr, chal_io(ch));
            chal_run(ch, pr);
            d2h(ch->ctx, out + off, chal_io(ch), m * sizeof(uint32_t));
        }
    });
}

int p3h_challenger_sample_bits(p3h_chal* ch, uint32_t bits, uint64_t* out) {
    if (!ch) return P3H_EINVAL;
    return api(ch->ctx, [&] {
        if (!out || bits > 64) throw Err{P3H_EINVAL, "bits must be 0..64"};
        const uint32_t n = (bits + 30) / 31;
        uint32_t v[3] = {0, 0, 0};
        if (n) {
            ChalProgram pr{};
            pr.op[pr.n_ops++] = chal_op(CH_SAMPLE, n, 0, nullptr, chal_io(ch));
            chal_run(ch, pr);
            d2h(ch->ctx, v, chal_io(ch), n * sizeof(uint32_t));
        }
        uint64_t r = 0;
        uint32_t rem = bits;
        for (uint32_t i = 0; i < n; ++i) {
            const uint32_t chunk = std::min(rem, 31u);
            r |= uint64_t(v[i] & ((1u << chunk) - 1u)) << (bits - rem);
            rem -= chunk;
        }
        *out = r;
    });
}

int p3h_challenger_grind(p3h_chal* ch, uint32_t bits, uint64_t* witness) {
    if (!ch) return P3H_EINVAL;
    return api(ch->ctx, [&] {
        if (!witness) throw Err{P3H_EINVAL, "null argument"};
        *witness = chal_grind(ch, bits);
        hip_check(hipStreamSynchronize(ch->ctx->stream), "sync");
    });
}

int p3h_challenger_check_witness(p3h_chal* ch, uint32_t bits, uint64_t witness, int* ok) {
    if (!ch) return P3H_EINVAL;
    return api(ch->ctx, [&] {
        if (!ok || bits > 63) throw Err{P3H_EINVAL, "bits must be 0..63"};
        ChalProgram pr{};
        pr.op[pr.n_ops++] = chal_op(CH_WITNESS, 0, witness);
        pr.op[pr.n_ops++] = chal_op(CH_SAMPLE, 1, 0, nullptr, chal_io(ch));
        chal_run(ch, pr);
        uint32_t v = 0;
        d2h(ch->ctx, &v, chal_io(ch), sizeof(v));
        *ok = (uint64_t(v) & ((1ull << bits) - 1)) == 0 ? 1 : 0;
    });
}

int p3h_challenger_state(p3h_chal* ch, uint32_t* out34) {
    if (!ch) return P3H_EINVAL;
    return api(ch->ctx, [&] {
        if (!out34) throw Err{P3H_EINVAL, "null argument"};
        DevChal h{};
        d2h(ch->ctx, &h, chal_state(ch), sizeof(h));
        std::memcpy(out34, h.sponge, 16 * 4);
        std::memcpy(out34 + 16, h.input, 8 * 4);
        std::memcpy(out34 + 24, h.output, 8 * 4);
        out34[32] = h.in_len;
        out34[33] = h.out_len;
    });
}

// ------------------------------------------------------------- PCS open ---
int p3h_pcs_open(p3h_ctx* ctx, const p3h_tree* tree, const uint64_t* shifts, const uint32_t* n_points,
                 const uint32_t* points, p3h_chal* ch, uint32_t log_final_poly_len, uint32_t max_log_arity,
                 uint32_t num_queries, uint32_t commit_pow_bits, uint32_t query_pow_bits, uint32_t* opened_out,
                 uint32_t* roots_out, uint64_t* commit_witnesses_out, uint32_t* final_poly_out,
                 uint64_t* query_witness_out, uint64_t* query_indices_out, uint32_t* query_rows_out,
                 uint32_t* n_rounds, p3h_fri** keep) {
    return api(ctx, [&] {
        if (!tree || !shifts || !n_points || !ch || !roots_out || !commit_witnesses_out || !final_poly_out ||
            !query_witness_out || (num_queries && (!query_indices_out || !query_rows_out)))
            throw Err{P3H_EINVAL, "null argument"};
        if (ch->ctx != ctx || tree->ctx != ctx) throw Err{P3H_EINVAL, "challenger or tree of another context"};
        if (!tree->flat || tree->field != P3H_BABYBEAR || tree->mats.empty())
            throw Err{P3H_EINVAL, "expects the tree of p3h_pcs_commit"};
        const auto& mats = tree->mats;
        const uint32_t nm = uint32_t(mats.size());
        const uint32_t kb = mats[0]->k;
        uint64_t total_open = 0, total_w = 0;
        for (uint32_t m = 0; m < nm; ++m) {
            if (mats[m]->field != P3H_BABYBEAR || mats[m]->k != kb)
                throw Err{P3H_EINVAL, "PCS matrices must be BabyBear LDEs of one blowup"};
            if (shifts[m] % BB_P == 0) throw Err{P3H_EINVAL, "coset shift must be non-zero"};
            total_open += uint64_t(n_points[m]) * mats[m]->width;
            total_w += mats[m]->width;
        }
        if (total_open && (!points || !opened_out)) throw Err{P3H_EINVAL, "null argument"};
        auto lde_view = [&](const p3h_mat* m) {
            const uint32_t lN = log2_exact(m->height);
            return LdeView{m->data<uint32_t>(), lN - m->k, m->k, m->width};
        };
        // alpha (two_adic_fri_pcs.hpp:277)
        ChalProgram pr{};
        pr.op[pr.n_ops++] = chal_op(CH_SAMPLE, 4, 0, nullptr, chal_io(ch));
        chal_run(ch, pr);
        HE4 alpha;
        d2h(ctx, alpha.data(), chal_io(ch), 16);
        // opened values: barycentric interpolation on each LDE coset
        DevPtr opened_dev = dalloc(ctx, std::max<uint64_t>(1, total_open) * 16);
        std::vector<uint32_t> opened_host(total_open * 4);
        {
            Phase ph(ctx, "pcs_open_values");
            uint64_t off = 0, pt = 0;
            for (uint32_t m = 0; m < nm; ++m) {
                const p3h_mat* L = mats[m];
                const uint64_t N = L->height;
                const uint32_t lN = log2_exact(N), w = L->width, g = uint32_t(shifts[m] % BB_P);
                const uint32_t omega_m = bb_mont(bb_root(lN));
                const uint32_t common_m = bb_mont(bb_inv(bb_mul(uint32_t(N % BB_P), bb_pow(g, N - 1))));
                const uint32_t rows_per_block = uint32_t(std::max<uint64_t>(64, N / 2048));
                const uint32_t n_blocks = uint32_t((N + rows_per_block - 1) / rows_per_block);
                for (uint32_t p0 = 0; p0 < n_points[m]; p0 += PCS_MAX_POINTS) {
                    const uint32_t P = std::min(PCS_MAX_POINTS, n_points[m] - p0);
                    std::vector<uint32_t> z(4 * P), van(4 * P);
                    for (uint32_t p = 0; p < P; ++p) {
                        HE4 zz;
                        for (int q = 0; q < 4; ++q) zz[q] = points[4 * (pt + p0 + p) + q] % BB_P;
                        const HE4 v = he4_sub(he4_pow(zz, N), HE4{bb_pow(g, N), 0, 0, 0});
                        for (int q = 0; q < 4; ++q) {
                            z[4 * p + q] = zz[q];
                            van[4 * p + q] = v[q];
                        }
                    }
                    DevPtr zd = dalloc(ctx, z.size() * 4), vd = dalloc(ctx, van.size() * 4);
                    h2d_async(ctx, zd->ptr, z.data(), z.size() * 4);
                    h2d_async(ctx, vd->ptr, van.data(), van.size() * 4);
                    DevPtr u = dalloc(ctx, N * P * 16);
                    DevPtr part = dalloc(ctx, uint64_t(n_blocks) * P * w * 16);
                    const LdeView v = lde_view(L);
                    launch_check(pcs_bary_weights_device(v, g, omega_m, common_m, dptr<uint32_t>(zd), P,
                                                         dptr<uint32_t>(u), ctx->stream),
                                 "pcs weights");
                    launch_check(pcs_bary_dot_device(v, dptr<uint32_t>(u), P, rows_per_block, dptr<uint32_t>(part),
                                                     ctx->stream),
                                 "pcs dot");
                    launch_check(pcs_bary_finish_device(dptr<uint32_t>(part), n_blocks, P, w, dptr<uint32_t>(vd),
                                                        dptr<uint32_t>(opened_dev) + 4 * off, ctx->stream),
                                 "pcs finish");
                    off += uint64_t(P) * w;
                }
                pt += n_points[m];
            }
            if (total_open) d2h(ctx, opened_host.data(), dptr<uint32_t>(opened_dev), total_open * 16);
        }
        if (total_open) std::memcpy(opened_out, opened_host.data(), total_open * 16);
        // the transcript observes every opened value (observe_challenge)
        for (uint64_t o = 0; o < total_open * 4; o += 0xFFFF0000u) {
            ChalProgram po{};
            po.op[po.n_ops++] = chal_op(CH_OBSERVE, uint32_t(std::min<uint64_t>(0xFFFF0000u, total_open * 4 - o)), 0,
                                        dptr<uint32_t>(opened_dev) + o);
            chal_run(ch, po);
        }
        // quotients, one vector per LDE height (descending), bit-reversed
        std::map<uint32_t, std::vector<uint32_t>, std::greater<uint32_t>> groups;
        for (uint32_t m = 0; m < nm; ++m) groups[log2_exact(mats[m]->height)].push_back(m);
        std::vector<uint64_t> aoff(nm), ooff(nm);
        {
            uint64_t a = 0;
            for (uint32_t m = 0; m < nm; ++m) {
                aoff[m] = a;
                ooff[m] = a;  // opened values are laid out in the same [matrix][point][col] order
                a += uint64_t(n_points[m]) * mats[m]->width;
            }
        }
        std::vector<p3h_mat*> inputs;
        try {
            std::unique_ptr<Phase> ph(new Phase(ctx, "pcs_quotients"));
            for (auto& [lN, ms] : groups) {
                p3h_mat* q = new_mat(ctx, P3H_BABYBEAR, 1ull << lN, 4);
                inputs.push_back(q);
                hip_check(hipMemsetAsync(q->data<uint32_t>(), 0, (16ull << lN), ctx->stream), "memset");
                for (uint32_t m : ms) {
                    const p3h_mat* L = mats[m];
                    const uint32_t w = L->width, P = n_points[m], g = uint32_t(shifts[m] % BB_P);
                    if (P == 0) continue;
                    // alpha^c (Montgomery), S_p = sum_c alpha^c f_p[c], alpha^(a + p w) (Montgomery)
                    std::vector<uint32_t> apow(4 * w), z(4 * P), sv(4 * P), ao(4 * P);
                    std::vector<HE4> ap(w);
                    HE4 cur = {1, 0, 0, 0};
                    for (uint32_t c = 0; c < w; ++c) {
                        ap[c] = cur;
                        const HE4 cm = he4_mont(cur);
                        for (int qq = 0; qq < 4; ++qq) apow[4 * c + qq] = cm[qq];
                        cur = he4_mul(cur, alpha);
                    }
                    uint64_t pbase = 0;
                    for (uint32_t mm = 0; mm < m; ++mm) pbase += n_points[mm];
                    for (uint32_t p = 0; p < P; ++p) {
                        HE4 acc = {0, 0, 0, 0};
                        for (uint32_t c = 0; c < w; ++c) {
                            const uint32_t* f = &opened_host[4 * (ooff[m] + uint64_t(p) * w + c)];
                            acc = he4_add(acc, he4_mul(ap[c], HE4{f[0], f[1], f[2], f[3]}));
                        }
                        const HE4 a = he4_mont(he4_pow(alpha, aoff[m] + uint64_t(p) * w));
                        for (int qq = 0; qq < 4; ++qq) {
                            z[4 * p + qq] = points[4 * (pbase + p) + qq] % BB_P;
                            sv[4 * p + qq] = acc[qq];
                            ao[4 * p + qq] = a[qq];
                        }
                    }
                    DevPtr ad = dalloc(ctx, apow.size() * 4), zd = dalloc(ctx, z.size() * 4),
                           sd = dalloc(ctx, sv.size() * 4), od = dalloc(ctx, ao.size() * 4),
                           abd = dalloc(ctx, L->height * 32);
                    h2d_async(ctx, ad->ptr, apow.data(), apow.size() * 4);
                    h2d_async(ctx, zd->ptr, z.data(), z.size() * 4);
                    h2d_async(ctx, sd->ptr, sv.data(), sv.size() * 4);
                    h2d_async(ctx, od->ptr, ao.data(), ao.size() * 4);
                    launch_check(pcs_quotient_device(lde_view(L), g, bb_mont(bb_root(lN)), dptr<uint32_t>(ad),
                                                     dptr<uint32_t>(zd), dptr<uint32_t>(sd), dptr<uint32_t>(od), P,
                                                     dptr<uint32_t>(abd), q->data<uint32_t>(), ctx->stream),
                                 "pcs quotient");
                }
            }
            ph.reset();
            // prove_fri over the quotient vectors, then its query phase
            p3h_fri* f = nullptr;
            std::vector<uint32_t> betas(4 * 32);
            std::vector<p3h_mat*> extra(inputs.begin() + 1, inputs.end());
            do_fri_commit(ctx, inputs[0], kb, log_final_poly_len, max_log_arity, nullptr, ch, commit_pow_bits,
                          roots_out, betas.data(), commit_witnesses_out, final_poly_out, n_rounds, &f, extra);
            std::unique_ptr<p3h_fri, int (*)(p3h_fri*)> fguard(f, p3h_fri_free);
            Phase pq(ctx, "fri_queries");
            *query_witness_out = chal_grind(ch, query_pow_bits);
            const uint32_t log_max = log2_exact(inputs[0]->height);
            const uint32_t per = (log_max + 30) / 31;
            std::vector<uint32_t> smp(uint64_t(num_queries) * per);
            for (uint64_t off = 0; off < smp.size(); off += 64) {
                const uint32_t cnt = uint32_t(std::min<uint64_t>(64, smp.size() - off));
                ChalProgram ps{};
                ps.op[ps.n_ops++] = chal_op(CH_SAMPLE, cnt, 0, nullptr, chal_io(ch));
                chal_run(ch, ps);
                d2h(ctx, smp.data() + off, chal_io(ch), cnt * 4);
            }
            for (uint32_t q = 0; q < num_queries; ++q) {  // sample_bits (duplex_challenger.hpp:137-148)
                uint64_t r = 0;
                uint32_t rem = log_max;
                for (uint32_t i = 0; i < per; ++i) {
                    const uint32_t chunk = std::min(rem, 31u);
                    r |= uint64_t(smp[q * per + i] & ((1u << chunk) - 1u)) << (log_max - rem);
                    rem -= chunk;
                }
                query_indices_out[q] = r;
            }
            // PcsQueryInputMmcs::open: per height group the concatenated LDE rows at
            // natural index bitrev(index >> (log_max - log_h), log_h)
            if (num_queries) {
                uint64_t col0 = 0;
                DevPtr rows_d = dalloc(ctx, uint64_t(num_queries) * 8);
                for (auto& [lN, ms] : groups) {
                    std::vector<RowSeg<uint32_t>> segs;
                    uint32_t gw = 0;
                    for (uint32_t m : ms) {
                        segs.push_back(seg_of<HBB>(mats[m]));
                        gw += mats[m]->width;
                    }
                    std::vector<uint64_t> rows(num_queries);
                    for (uint32_t q = 0; q < num_queries; ++q) {
                        const uint64_t red = query_indices_out[q] >> (log_max - lN);
                        uint64_t k = 0;
                        for (uint32_t b = 0; b < lN; ++b) k |= ((red >> b) & 1ull) << (lN - 1 - b);
                        rows[q] = k;
                    }
                    DevPtr sd = upload_segs<HBB>(ctx, segs);
                    h2d(ctx, rows_d->ptr, rows.data(), rows.size() * 8);
                    DevPtr out = dalloc(ctx, uint64_t(num_queries) * gw * 4);
                    launch_check(gather_rows_device<uint32_t>(dptr<RowSeg<uint32_t>>(sd), uint32_t(segs.size()), gw,
                                                              dptr<uint64_t>(rows_d), num_queries,
                                                              dptr<uint32_t>(out), ctx->stream),
                                 "query rows");
                    std::vector<uint32_t> host(uint64_t(num_queries) * gw);
                    d2h(ctx, host.data(), dptr<uint32_t>(out), host.size() * 4);
                    for (uint32_t q = 0; q < num_queries; ++q)
                        std::memcpy(query_rows_out + uint64_t(q) * total_w + col0, &host[uint64_t(q) * gw], gw * 4);
                    col0 += gw;
                }
            }
            if (keep) *keep = fguard.release();
        } catch (...) {
            for (auto* q : inputs) mat_unref(q);
            throw;
        }
        for (auto* q : inputs) mat_unref(q);
    });
}

int p3h_fri_layer_tree(const p3h_fri* f, uint32_t round, const p3h_tree** out) {
    if (!f || !out || round >= f->trees.size()) return P3H_EINVAL;
    *out = f->trees[round];
    return P3H_OK;
}

int p3h_fri_free(p3h_fri* f) {
    if (!f) return P3H_EINVAL;
    for (auto* t : f->trees) p3h_tree_free(t);
    for (auto* m : f->layers) mat_unref(m);
    delete f;
    return P3H_OK;
}

int p3h_pcs_commit(p3h_ctx* ctx, const p3h_mat* const* mats, const uint64_t* shifts, uint32_t n, uint32_t log_blowup,
                   p3h_tree** out) {
    return api(ctx, [&] {
        if (!mats || !shifts || !out || n == 0) throw Err{P3H_EINVAL, "null argument"};
        std::vector<p3h_mat*> ldes;
        try {
            uint64_t maxh = 0;
            uint32_t tw = 0;
            for (uint32_t i = 0; i < n; ++i) {
                if (mats[i]->field != P3H_BABYBEAR) throw Err{P3H_EINVAL, "PCS commit is BabyBear only"};
                ldes.push_back(do_lde<HBB>(ctx, mats[i], log_blowup, shifts[i]));
                maxh = std::max(maxh, ldes.back()->height);
                tw += mats[i]->width;
            }
            if (tw == 0) throw Err{P3H_EINVAL, "TwoAdicFriPcs::commit: empty matrices (total_width=0)"};
            uint32_t pw = 1;
            while (pw < tw) pw <<= 1;
            std::vector<RowSeg<uint32_t>> segs;
            for (auto* l : ldes) {
                RowSeg<uint32_t> s = seg_of<HBB>(l);
                s.hmask = l->height - 1;
                segs.push_back(s);
            }
            if (pw > tw) segs.push_back(RowSeg<uint32_t>{nullptr, ~0ull, pw - tw, 0, 0, 0});
            *out = flat_tree<HBB>(ctx, segs, maxh, pw, ldes);
        } catch (...) {
            for (auto* l : ldes) mat_unref(l);
            throw;
        }
        for (auto* l : ldes) mat_unref(l);
    });
}

int p3h_batch_commit(p3h_ctx* ctx, const p3h_mat* const* traces, const uint32_t* added_bits, uint32_t n,
                     uint64_t shift, uint32_t cap_height, p3h_tree** out) {
    return api(ctx, [&] {
        if (!out) throw Err{P3H_EINVAL, "null argument"};
        BatchPlan* bp = make_batch_plan(ctx, traces, added_bits, n, shift, cap_height, 0);
        try {
            run_batch_plan(bp);
            bp->check_queue();
        } catch (...) {
            delete bp;
            throw;
        }
        *out = bp->tree;  // hand the tree over; it keeps the LDEs alive
        bp->tree = nullptr;
        delete bp;
    });
}

struct p3h_batch {
    BatchPlan* plan;
};

int p3h_batch_plan(p3h_ctx* ctx, const p3h_mat* const* traces, const uint32_t* added_bits, uint32_t n, uint64_t shift,
                   uint32_t cap_height, int flags, p3h_batch** out) {
    return api(ctx, [&] {
        if (!out) throw Err{P3H_EINVAL, "null argument"};
        if (flags & ~(P3H_BATCH_GRAPH | P3H_BATCH_QUEUE_LDE)) throw Err{P3H_EINVAL, "unknown batch flags"};
        BatchPlan* bp = make_batch_plan(ctx, traces, added_bits, n, shift, cap_height, flags);
        *out = new p3h_batch{bp};
    });
}

int p3h_batch_lde_launches(const p3h_batch* b, uint32_t* launches) {
    if (!b || !launches) return P3H_EINVAL;
    *launches = b->plan->queue ? 1u : uint32_t(b->plan->groups.size());
    return P3H_OK;
}

int p3h_batch_run(p3h_batch* b, void* cap_out) {
    if (!b) return P3H_EINVAL;
    p3h_ctx* ctx = b->plan->ctx;
    return api(ctx, [&] {
        run_batch_plan(b->plan);
        b->plan->check_queue();
        if (cap_out) d2h(ctx, cap_out, b->plan->tree->layers.back()->ptr, b->plan->tree->counts.back() * 32);
    });
}

int p3h_batch_tree(const p3h_batch* b, const p3h_tree** out) {
    if (!b || !out) return P3H_EINVAL;
    *out = b->plan->tree;
    return P3H_OK;
}

int p3h_batch_free(p3h_batch* b) {
    if (!b) return P3H_EINVAL;
    (void)hipStreamSynchronize(b->plan->ctx->stream);
    delete b->plan;
    delete b;
    return P3H_OK;
}

int p3h_mat_write(p3h_ctx* ctx, p3h_mat* m, const void* host) {
    return api(ctx, [&] {
        if (!m || !host) throw Err{P3H_EINVAL, "null argument"};
        if (m->k != 0) throw Err{P3H_EINVAL, "only natural-order matrices can be overwritten"};
        h2d(ctx, m->data<char>(), host, m->height * m->width * elem_size(m->field));
    });
}

// Time one LDE shape under candidate pass shapes and keep the fastest.
int p3h_autotune_lde(p3h_ctx* ctx, int field, uint32_t log_n, uint32_t width, uint32_t added_bits, uint32_t iters,
                     uint32_t* best_log_cols, uint32_t* best_max_stages, uint32_t* best_log_slab, double* best_ms) {
    return api(ctx, [&] {
        if (field != P3H_BABYBEAR && field != P3H_GOLDILOCKS) throw Err{P3H_EINVAL, "unknown field"};
        if (width == 0 || iters == 0) throw Err{P3H_EINVAL, "width and iters must be positive"};
        uint64_t h = 1ull << log_n;
        p3h_mat* in = new_mat(ctx, field, h, width);
        std::unique_ptr<p3h_mat, void (*)(p3h_mat*)> guard(in, mat_unref);
        hip_check(hipMemsetAsync(in->buf->ptr, 0, h * width * elem_size(field), ctx->stream), "memset");
        // passes of at most 10 stages (the tile holds 2^15 BabyBear or 2^14
        // Goldilocks words, at least 16 columns at 10 stages)
        const uint32_t smax = 10;
        std::vector<NttTuning> cands;
        for (uint32_t sl : {0u, 4u, 5u, 6u}) {
            if (sl && (1u << sl) >= width) continue;
            for (uint32_t lc = 2; lc <= 7; ++lc) {
                if (sl && lc > sl) continue;
                for (uint32_t ms : {0u, smax - 1, smax - 2, smax - 3}) cands.push_back(NttTuning{lc, ms, sl});
            }
        }
        auto key = std::make_tuple(field, log_n, width, added_bits);
        NttTuning saved = ctx->tune;
        bool had = ctx->tuned.count(key) != 0;
        NttTuning prev = had ? ctx->tuned[key] : NttTuning{};
        double best = 1e30;
        NttTuning bt{};
        hipEvent_t e0, e1;
        hip_check(hipEventCreate(&e0), "event");
        hip_check(hipEventCreate(&e1), "event");
        for (const auto& c : cands) {
            ctx->tuned[key] = c;
            double tmin = 1e30;
            for (uint32_t it = 0; it <= iters; ++it) {
                (void)hipEventRecord(e0, ctx->stream);
                p3h_mat* o = field == P3H_BABYBEAR ? do_lde<HBB>(ctx, in, added_bits, 31)
                                                   : do_lde<HGL>(ctx, in, added_bits, 7);
                (void)hipEventRecord(e1, ctx->stream);
                (void)hipEventSynchronize(e1);
                mat_unref(o);
                float ms = 0;
                (void)hipEventElapsedTime(&ms, e0, e1);
                if (it > 0 && ms < tmin) tmin = ms;  // first run warms the table caches
            }
            if (tmin < best) {
                best = tmin;
                bt = c;
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        ctx->tune = saved;
        ctx->tuned[key] = bt;
        (void)prev;
        if (best_log_cols) *best_log_cols = bt.log_cols;
        if (best_max_stages) *best_max_stages = bt.max_stages;
        if (best_log_slab) *best_log_slab = bt.log_slab;
        if (best_ms) *best_ms = best;
    });
}

// Tuning table as text lines "field log_n width added log_cols max_stages log_slab".
int p3h_tuning_save(const p3h_ctx* ctx, const char* path) {
    auto* c = const_cast<p3h_ctx*>(ctx);
    return api(c, [&] {
        FILE* f = std::fopen(path, "w");
        if (!f) throw Err{P3H_EINVAL, std::string("cannot write ") + path};
        for (const auto& kv : c->tuned)
            std::fprintf(f, "%d %u %u %u %u %u %u\n", std::get<0>(kv.first), std::get<1>(kv.first),
                         std::get<2>(kv.first), std::get<3>(kv.first), kv.second.log_cols, kv.second.max_stages,
                         kv.second.log_slab);
        std::fclose(f);
    });
}

int p3h_tuning_load(p3h_ctx* ctx, const char* path) {
    return api(ctx, [&] {
        FILE* f = std::fopen(path, "r");
        if (!f) throw Err{P3H_EINVAL, std::string("cannot read ") + path};
        int field;
        unsigned L, w, k, lc, ms, sl;
        while (std::fscanf(f, "%d %u %u %u %u %u %u", &field, &L, &w, &k, &lc, &ms, &sl) == 7)
            ctx->tuned[std::make_tuple(field, L, w, k)] = NttTuning{lc, ms, sl};
        std::fclose(f);
    });
}

}  // extern "C"

// ============================================================ serving ===
// A commit server: n_slots worker threads, each owning a context (its own HIP
// stream and allocator), pull requests from one queue.  A worker takes every
// queued request that fits the remaining HBM budget (up to max_batch), runs
// their LDEs as one work-list launch group sequence, then each request's tree,
// and publishes the roots.  Results (LDE + all Merkle layers) stay resident
// until released, so openings are served from HBM.  Requests of different
// slots run concurrently on the device.
struct ServeReq {
    uint64_t id;
    const uint32_t* trace;        // caller's host rows, read once at upload
    uint64_t height;
    uint32_t width, k;
    uint32_t shift;
    uint64_t reserve;             // bytes held against the budget until release
    std::chrono::steady_clock::time_point t_submit, t_done;
    bool done = false;
    int status = P3H_OK;
    std::string err;
    uint32_t root[8];
    p3h_tree* tree = nullptr;
    int slot = -1;
};

struct p3h_serve {
    int device;
    uint64_t budget;
    uint64_t reserved = 0;
    uint32_t max_batch;
    std::vector<p3h_ctx*> slots;
    std::vector<std::unique_ptr<std::mutex>> slot_mu;
    std::vector<std::thread> workers;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<ServeReq*> queue;
    std::unordered_map<uint64_t, std::unique_ptr<ServeReq>> reqs;
    uint64_t next_id = 1;
    bool stop = false;
    int callers = 0;  // wait/poll/open/release calls in progress (destroy waits for them)
    uint64_t completed = 0, batches = 0, committed_elems = 0;
    std::string err;
};

namespace {

void serve_process(p3h_serve* sv, int slot, std::vector<ServeReq*>& batch) {
    p3h_ctx* ctx = sv->slots[slot];
    std::lock_guard<std::mutex> lk(*sv->slot_mu[slot]);
    std::vector<p3h_mat*> traces(batch.size(), nullptr), ldes(batch.size(), nullptr);
    std::string err;
    int code = P3H_OK;
    try {
        std::vector<LdeJob<uint32_t>> jobs;
        for (size_t i = 0; i < batch.size(); ++i) {
            ServeReq* r = batch[i];
            traces[i] = new_mat(ctx, P3H_BABYBEAR, r->height, r->width);
            h2d(ctx, traces[i]->data<char>(), r->trace, r->height * r->width * 4);
            r->trace = nullptr;
            uint32_t L = log2_exact(r->height);
            ldes[i] = new_mat(ctx, P3H_BABYBEAR, r->height << r->k, r->width);
            ldes[i]->k = r->k;
            ldes[i]->L = L;
            jobs.push_back(LdeJob<uint32_t>{traces[i]->data<uint32_t>(), ldes[i]->data<uint32_t>(), L, r->width, r->k,
                                            twiddles<HBB>(ctx, L), lde_twist<HBB>(ctx, L, r->k, r->shift),
                                            tuning_for(ctx, P3H_BABYBEAR, L, r->width, r->k)});
        }
        auto groups = plan_lde_batch<BabyBear>(jobs);
        std::vector<DevPtr> tables;
        for (const auto& g : groups) {
            DevPtr d = dalloc(ctx, g.work.size() * sizeof(NttWork<uint32_t>));
            h2d(ctx, d->ptr, g.work.data(), g.work.size() * sizeof(NttWork<uint32_t>));
            tables.push_back(d);
        }
        for (size_t g = 0; g < groups.size(); ++g)
            launch_check(launch_lde_group<BabyBear>(groups[g], dptr<NttWork<uint32_t>>(tables[g]), ctx->stream),
                         "serve lde");
        for (size_t i = 0; i < batch.size(); ++i) {
            const p3h_mat* one = ldes[i];
            batch[i]->tree = do_mmcs_commit<HBB>(ctx, &one, 1, 0);
        }
        for (size_t i = 0; i < batch.size(); ++i)
            hip_check(hipMemcpyAsync(batch[i]->root, batch[i]->tree->layers.back()->ptr, 32, hipMemcpyDeviceToHost,
                                     ctx->stream),
                      "root copy");
        hip_check(hipStreamSynchronize(ctx->stream), "sync");
    } catch (const Err& e) {
        code = e.code;
        err = e.msg;
    } catch (const std::exception& e) {
        code = P3H_EDEVICE;
        err = e.what();
    }
    for (auto* m : traces) mat_unref(m);
    for (auto* m : ldes) mat_unref(m);  // trees hold their own references
    auto now = std::chrono::steady_clock::now();
    std::lock_guard<std::mutex> g(sv->mu);
    for (auto* r : batch) {
        r->t_done = now;
        r->done = true;
        r->status = code;
        r->err = err;
        r->slot = slot;
        if (code != P3H_OK) {
            if (r->tree) p3h_tree_free(r->tree);
            r->tree = nullptr;
            sv->reserved -= r->reserve;
            r->reserve = 0;
        } else {
            sv->committed_elems += (r->height << r->k) * r->width;
        }
        ++sv->completed;
    }
    ++sv->batches;
    sv->cv_done.notify_all();
    sv->cv_work.notify_all();
}

// A wait/poll/open/release call in progress: p3h_serve_destroy waits until
// none is left before it frees the server.
struct ServeCall {
    p3h_serve* sv;
    explicit ServeCall(p3h_serve* s) : sv(s) {
        std::lock_guard<std::mutex> lk(sv->mu);
        ++sv->callers;
    }
    ~ServeCall() {
        std::lock_guard<std::mutex> lk(sv->mu);
        --sv->callers;
        sv->cv_done.notify_all();
    }
};

void serve_worker(p3h_serve* sv, int slot) {
    (void)hipSetDevice(sv->device);
    std::unique_lock<std::mutex> lk(sv->mu);
    std::vector<uint64_t> fp;
    while (true) {
        sv->cv_work.wait(lk, [&] {
            return sv->stop || (!sv->queue.empty() && sv->queue.front()->reserve <= sv->budget - sv->reserved);
        });
        if (sv->stop) return;
        fp.clear();
        for (size_t i = 0; i < sv->queue.size() && i < sv->max_batch; ++i) fp.push_back(sv->queue[i]->reserve);
        const size_t take = serve_admit(sv->reserved, sv->budget, sv->max_batch, fp.data(), fp.size());
        std::vector<ServeReq*> batch;
        for (size_t i = 0; i < take; ++i) {
            ServeReq* r = sv->queue.front();
            sv->queue.pop_front();
            sv->reserved += r->reserve;
            batch.push_back(r);
        }
        lk.unlock();
        serve_process(sv, slot, batch);
        lk.lock();
    }
}

}  // namespace

extern "C" {

int p3h_serve_create(int device, uint64_t hbm_budget, uint32_t n_slots, uint32_t max_batch, const uint32_t* consts,
                     uint32_t rounds_f, uint32_t rounds_p, p3h_serve** out) {
    if (!out || !consts || n_slots == 0 || n_slots > 16 || max_batch == 0) return P3H_EINVAL;
    auto* sv = new p3h_serve{};
    sv->device = device;
    sv->max_batch = max_batch;
    for (uint32_t i = 0; i < n_slots; ++i) {
        p3h_ctx* c = nullptr;
        int rc = p3h_ctx_create(device, 0, &c);
        if (rc == P3H_OK) rc = p3h_poseidon2_set(c, P3H_BABYBEAR, 16, rounds_f, rounds_p, consts);
        if (rc != P3H_OK) {
            for (auto* s : sv->slots) p3h_ctx_destroy(s);
            if (c) p3h_ctx_destroy(c);
            delete sv;
            return rc;
        }
        sv->slots.push_back(c);
        sv->slot_mu.emplace_back(new std::mutex);
    }
    sv->budget = hbm_budget ? hbm_budget : sv->slots[0]->budget;
    for (uint32_t i = 0; i < n_slots; ++i) sv->workers.emplace_back(serve_worker, sv, int(i));
    *out = sv;
    return P3H_OK;
}

int p3h_serve_submit(p3h_serve* sv, const uint32_t* trace, uint64_t height, uint32_t width, uint32_t added_bits,
                     uint64_t shift, uint64_t* id_out) {
    if (!sv || !trace || !id_out || width == 0 || height == 0 || (height & (height - 1)) || shift >= BB_P ||
        __builtin_ctzll(height) + added_bits > 27)
        return P3H_EINVAL;
    auto t0 = std::chrono::steady_clock::now();
    auto r = std::make_unique<ServeReq>();
    r->trace = trace;
    r->height = height;
    r->width = width;
    r->k = added_bits;
    r->shift = uint32_t(shift);
    r->reserve = serve_footprint(height, width, added_bits);
    std::lock_guard<std::mutex> lk(sv->mu);
    if (sv->stop) {
        sv->err = "the server is being destroyed";
        return P3H_ESTATE;
    }
    if (r->reserve > sv->budget) {
        sv->err = "request footprint exceeds the server's HBM budget";
        return P3H_EBUDGET;
    }
    r->id = sv->next_id++;
    r->t_submit = t0;
    *id_out = r->id;
    sv->queue.push_back(r.get());
    sv->reqs[r->id] = std::move(r);
    sv->cv_work.notify_one();
    return P3H_OK;
}

int p3h_serve_wait(p3h_serve* sv, uint64_t id, uint32_t* root_out, double* latency_ms) {
    if (!sv) return P3H_EINVAL;
    const ServeCall call(sv);
    std::unique_lock<std::mutex> lk(sv->mu);
    auto it = sv->reqs.find(id);
    if (it == sv->reqs.end()) return P3H_EINVAL;
    ServeReq* r = it->second.get();
    sv->cv_done.wait(lk, [&] { return r->done; });
    if (r->status != P3H_OK) {
        sv->err = r->err;
        return r->status;
    }
    if (root_out) std::memcpy(root_out, r->root, 32);
    if (latency_ms) latency_ms[0] = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
    return P3H_OK;
}

int p3h_serve_poll(p3h_serve* sv, uint64_t id, int* done) {
    if (!sv || !done) return P3H_EINVAL;
    const ServeCall call(sv);
    std::lock_guard<std::mutex> lk(sv->mu);
    auto it = sv->reqs.find(id);
    if (it == sv->reqs.end()) return P3H_EINVAL;
    *done = it->second->done ? 1 : 0;
    return P3H_OK;
}

int p3h_serve_open(p3h_serve* sv, uint64_t id, const uint64_t* indices, uint32_t n, uint32_t* values,
                   uint32_t* proof) {
    if (!sv) return P3H_EINVAL;
    const ServeCall call(sv);
    ServeReq* r = nullptr;
    {
        std::unique_lock<std::mutex> lk(sv->mu);
        auto it = sv->reqs.find(id);
        if (it == sv->reqs.end()) return P3H_EINVAL;
        r = it->second.get();
        sv->cv_done.wait(lk, [&] { return r->done; });
        if (r->status != P3H_OK || !r->tree) return P3H_ESTATE;
    }
    std::lock_guard<std::mutex> g(*sv->slot_mu[r->slot]);
    int rc = p3h_tree_open(sv->slots[r->slot], r->tree, indices, n, values, proof);
    if (rc != P3H_OK) {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->err = p3h_last_error(sv->slots[r->slot]);
    }
    return rc;
}

int p3h_serve_release(p3h_serve* sv, uint64_t id) {
    if (!sv) return P3H_EINVAL;
    const ServeCall call(sv);
    std::unique_ptr<ServeReq> r;
    {
        std::unique_lock<std::mutex> lk(sv->mu);
        auto it = sv->reqs.find(id);
        if (it == sv->reqs.end()) return P3H_EINVAL;
        sv->cv_done.wait(lk, [&] { return it->second->done; });
        r = std::move(it->second);
        sv->reqs.erase(it);
    }
    if (r->tree) {
        std::lock_guard<std::mutex> g(*sv->slot_mu[r->slot]);
        p3h_tree_free(r->tree);
        // keep the slot's block cache within its share of the budget
        p3h_ctx* c = sv->slots[r->slot];
        if (c->cached > sv->budget / sv->slots.size()) trim(c);
    }
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->reserved -= r->reserve;
    sv->cv_work.notify_all();
    return P3H_OK;
}

int p3h_serve_stats(p3h_serve* sv, uint64_t* completed, uint64_t* batches, uint64_t* reserved_bytes,
                    uint64_t* committed_elems) {
    if (!sv) return P3H_EINVAL;
    std::lock_guard<std::mutex> lk(sv->mu);
    if (completed) *completed = sv->completed;
    if (batches) *batches = sv->batches;
    if (reserved_bytes) *reserved_bytes = sv->reserved;
    if (committed_elems) *committed_elems = sv->committed_elems;
    return P3H_OK;
}

const char* p3h_serve_last_error(p3h_serve* sv) { return sv ? sv->err.c_str() : "null server"; }

int p3h_serve_destroy(p3h_serve* sv) {
    if (!sv) return P3H_EINVAL;
    {
        // requests still queued fail; their waiters wake with P3H_ESTATE
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->stop = true;
        for (ServeReq* r : sv->queue) {
            r->done = true;
            r->status = P3H_ESTATE;
            r->err = "the server was destroyed before the request ran";
            sv->reserved -= r->reserve;
            r->reserve = 0;
        }
        sv->queue.clear();
    }
    sv->cv_work.notify_all();
    sv->cv_done.notify_all();
    for (auto& t : sv->workers) t.join();  // batches in flight complete and publish
    {
        std::unique_lock<std::mutex> lk(sv->mu);
        sv->cv_done.wait(lk, [&] { return sv->callers == 0; });
    }
    for (auto& kv : sv->reqs)
        if (kv.second->tree) p3h_tree_free(kv.second->tree);
    sv->reqs.clear();
    for (auto* c : sv->slots) p3h_ctx_destroy(c);
    delete sv;
    return P3H_OK;
}

}  // extern "C"
