// VALU issue rates on gfx950: cycles per wave64 instruction of single VALU
// instructions (and a few dependent pairs), each in 8 independent chains per
// lane, 8192 workgroups of 256 lanes, timed with HIP events.  The NTT and
// Poseidon2 arithmetic is chosen by these figures.  Build: make -C
// plonky3-gpu_amd/csrc valu_rates; run: plonky3-gpu_amd/valu_rates.  Cycles are
// counted at 2.4 GHz over the 1024 SIMDs (256 CUs x 4).
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdint>

#define CHECK(x) do { hipError_t e = (x); if (e != hipSuccess) { printf("HIP error %s at %d\n", hipGetErrorString(e), __LINE__); exit(1);} } while (0)

template <int OP, int CH>
__global__ __launch_bounds__(256) void probe(uint32_t* out, uint32_t iters, uint32_t seed) {
    uint32_t tid = blockIdx.x * 256 + threadIdx.x;
    uint32_t a[CH], b[CH];
    uint64_t s[CH];
    double d[CH];
    uint32_t c = seed * 7u + 3u;
    uint64_t msk = 0x5555555555555555ull ^ seed;
#pragma unroll
    for (int k = 0; k < CH; ++k) { a[k] = tid * 13u + k; b[k] = tid ^ (k * 77u); d[k] = double(tid + k); s[k] = 0; }
    for (uint32_t it = 0; it < iters; ++it) {
#pragma unroll
        for (int r = 0; r < 8; ++r) {
#pragma unroll
            for (int k = 0; k < CH; ++k) {
                if constexpr (OP == 0) asm volatile("v_add_u32 %0, %0, %1" : "+v"(a[k]) : "v"(c));
                if constexpr (OP == 1) asm volatile("v_sub_u32 %0, %0, %1" : "+v"(a[k]) : "v"(c));
                if constexpr (OP == 2) asm volatile("v_and_b32 %0, %0, %1" : "+v"(a[k]) : "v"(c));
                if constexpr (OP == 3) asm volatile("v_xor_b32 %0, %0, %1" : "+v"(a[k]) : "v"(c));
                if constexpr (OP == 4) asm volatile("v_lshlrev_b32 %0, 3, %0" : "+v"(a[k]));
                if constexpr (OP == 5) asm volatile("v_min_u32 %0, %0, %1" : "+v"(a[k]) : "v"(c));
                if constexpr (OP == 6) asm volatile("v_max_u32 %0, %0, %1" : "+v"(a[k]) : "v"(c));
                if constexpr (OP == 7) asm volatile("v_cndmask_b32_e64 %0, %0, %1, %2" : "+v"(a[k]) : "v"(c), "s"(msk));
                if constexpr (OP == 8) asm volatile("v_add_co_u32 %0, %1, %0, %2" : "+v"(a[k]), "=s"(s[k]) : "v"(c));
                if constexpr (OP == 9) asm volatile("v_add_co_u32 %0, vcc, %0, %2\n\tv_addc_co_u32 %1, vcc, %1, %2, vcc" : "+v"(a[k]), "+v"(b[k]) : "v"(c) : "vcc");
                if constexpr (OP == 10) asm volatile("v_add_co_u32 %0, %2, %0, %3\n\tv_addc_co_u32 %1, %2, %1, %3, %2" : "+v"(a[k]), "+v"(b[k]), "=&s"(s[k]) : "v"(c));
                if constexpr (OP == 11) asm volatile("v_mul_lo_u32 %0, %0, %1" : "+v"(a[k]) : "v"(c));
                if constexpr (OP == 12) {
                    uint64_t t = (uint64_t(b[k]) << 32) | a[k];
                    asm volatile("v_mad_u64_u32 %0, %1, %2, %3, %0" : "+v"(t), "=s"(s[k]) : "v"(a[k]), "v"(c));
                    a[k] = uint32_t(t); b[k] = uint32_t(t >> 32);
                }
                if constexpr (OP == 13) asm volatile("v_lshl_add_u32 %0, %0, 3, %1" : "+v"(a[k]) : "v"(c));
                if constexpr (OP == 14) asm volatile("v_add3_u32 %0, %0, %1, %2" : "+v"(a[k]) : "v"(c), "v"(b[k]));
                if constexpr (OP == 15) asm volatile("v_mad_u32_u24 %0, %0, %1, %2" : "+v"(a[k]) : "v"(c), "v"(b[k]));
                if constexpr (OP == 16) asm volatile("v_sub_co_u32 %0, vcc, %0, %2\n\tv_cndmask_b32 %1, %1, %0, vcc" : "+v"(a[k]), "+v"(b[k]) : "v"(c) : "vcc");
                if constexpr (OP == 17) asm volatile("v_fma_f64 %0, %0, %1, %1" : "+v"(d[k]) : "v"(double(c)));
                if constexpr (OP == 18) {
                    uint64_t t = (uint64_t(b[k]) << 32) | a[k];
                    asm volatile("v_lshlrev_b64 %0, 3, %0" : "+v"(t));
                    a[k] = uint32_t(t); b[k] = uint32_t(t >> 32);
                }
                if constexpr (OP == 19) asm volatile("v_alignbit_b32 %0, %0, %1, 7" : "+v"(a[k]) : "v"(c));
                if constexpr (OP == 20) asm volatile("v_cvt_f64_u32 %0, %1" : "=v"(d[k]) : "v"(a[k]));
                if constexpr (OP == 21) asm volatile("v_bfi_b32 %0, %0, %1, %2" : "+v"(a[k]) : "v"(c), "v"(b[k]));
                if constexpr (OP == 22) asm volatile("v_sub_u32 %0, %0, %2\n\tv_min_u32 %1, %1, %0" : "+v"(a[k]), "+v"(b[k]) : "v"(c));
                if constexpr (OP == 23) asm volatile("v_sub_co_u32 %0, %2, %0, %3\n\tv_cndmask_b32_e64 %1, %1, %0, %2" : "+v"(a[k]), "+v"(b[k]), "=&s"(s[k]) : "v"(c));
                if constexpr (OP == 24) asm volatile("v_mul_hi_u32 %0, %0, %1" : "+v"(a[k]) : "v"(c));
                if constexpr (OP == 25) asm volatile("v_add_f64 %0, %0, %1" : "+v"(d[k]) : "v"(double(c)));
                if constexpr (OP == 26) asm volatile("v_sub_co_u32 %0, vcc, %0, %1" : "+v"(a[k]) : "v"(c) : "vcc");
                if constexpr (OP == 27) asm volatile("v_addc_co_u32 %0, vcc, %0, %1, vcc" : "+v"(a[k]) : "v"(c) : "vcc");
                if constexpr (OP == 28) asm volatile("v_max_i32 %0, %0, %1" : "+v"(a[k]) : "v"(c));
                if constexpr (OP == 29) asm volatile("v_sub_u32 %0, %0, %1 clamp" : "+v"(a[k]) : "v"(c));
                if constexpr (OP == 30) asm volatile("v_add_u32_e64 %0, %0, %1" : "+v"(a[k]) : "v"(c));
                if constexpr (OP == 31) asm volatile("v_cmp_gt_u32 vcc, %1, %2\n\tv_cndmask_b32 %0, %1, %0, vcc" : "+v"(a[k]) : "v"(c), "v"(b[k]) : "vcc");
                if constexpr (OP == 32) asm volatile("v_mov_b32 %0, %1" : "=v"(a[k]) : "v"(b[k]));
                if constexpr (OP == 33) asm volatile("v_med3_u32 %0, %0, %1, %2" : "+v"(a[k]) : "v"(c), "v"(b[k]));
            }
        }
    }
    uint32_t acc = 0;
#pragma unroll
    for (int k = 0; k < CH; ++k) acc ^= a[k] ^ b[k] ^ uint32_t(s[k]) ^ uint32_t(__double_as_longlong(d[k]));
    out[tid] = acc;
}

static const char* NAMES[] = {"v_add_u32", "v_sub_u32", "v_and_b32", "v_xor_b32", "v_lshlrev_b32", "v_min_u32",
    "v_max_u32", "v_cndmask_b32_e64 (sgpr mask)", "v_add_co_u32 (sgpr carry)", "v_add_co+v_addc vcc [2]",
    "v_add_co+v_addc sgpr [2]", "v_mul_lo_u32", "v_mad_u64_u32", "v_lshl_add_u32", "v_add3_u32", "v_mad_u32_u24",
    "v_sub_co+v_cndmask vcc [2]", "v_fma_f64", "v_lshlrev_b64", "v_alignbit_b32", "v_cvt_f64_u32", "v_bfi_b32",
    "v_sub_u32+v_min_u32 [2]", "v_sub_co+v_cndmask sgpr [2]", "v_mul_hi_u32", "v_add_f64", "v_sub_co_u32 vcc",
    "v_addc_co_u32 vcc", "v_max_i32", "v_sub_u32 clamp", "v_add_u32_e64", "v_cmp_gt_u32+v_cndmask vcc [2]", "v_mov_b32", "v_med3_u32"};
static const int NINSTR[] = {1,1,1,1,1,1,1,1,1,2,2,1,1,1,1,1,2,1,1,1,1,1,2,2,1,1,1,1,1,1,1,2,1,1};

template <int OP>
void run(uint32_t* d_out, uint32_t blocks, uint32_t iters, bool print = true) {
    constexpr int CH = 8;
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0)); CHECK(hipEventCreate(&e1));
    hipLaunchKernelGGL((probe<OP, CH>), dim3(blocks), dim3(256), 0, 0, d_out, 4u, 1u);
    CHECK(hipDeviceSynchronize());
    CHECK(hipEventRecord(e0));
    hipLaunchKernelGGL((probe<OP, CH>), dim3(blocks), dim3(256), 0, 0, d_out, iters, 1u);
    CHECK(hipEventRecord(e1));
    CHECK(hipEventSynchronize(e1));
    float ms; CHECK(hipEventElapsedTime(&ms, e0, e1));
    const double waves = double(blocks) * 4.0;
    const double winstr = waves * iters * 8.0 * CH * NINSTR[OP];
    const double cyc = ms * 1e-3 * 2.4e9 * 1024.0 / winstr;
    if (print) printf("| `%s` | %.3f | %.2f |\n", NAMES[OP], ms, cyc);
    CHECK(hipEventDestroy(e0)); CHECK(hipEventDestroy(e1));
}

template <int... I>
void all(uint32_t* d, uint32_t blocks, uint32_t iters, std::integer_sequence<int, I...>) { (run<I>(d, blocks, iters), ...); }

int main() {
    const uint32_t blocks = 8192, iters = 256;
    uint32_t* d;
    CHECK(hipMalloc(&d, size_t(blocks) * 256 * 4));
    run<0>(d, blocks, iters, false);  // brings the clocks up
    printf("| instruction (pairs: per instruction) | ms | cycles per wave64 instruction |\n|---|---:|---:|\n");
    all(d, blocks, iters, std::make_integer_sequence<int, 34>{});
    CHECK(hipFree(d));
    printf("done\n");
    return 0;
}
