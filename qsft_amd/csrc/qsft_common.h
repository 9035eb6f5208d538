// Device-side helpers shared by the q-SFT kernels (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/qsft_amd.h"

namespace qsft {

// Pointers read from a descriptor are generic; accesses through them compile to
// flat_* instructions, which count on both the vector-memory and the LDS wait
// counters (so an LDS wait also waits for them). Kernels mark global data with
// gptr() so the accesses compile to global_* instructions.
#define QSFT_GLOBAL __attribute__((address_space(1)))
template <typename T> __device__ __forceinline__ QSFT_GLOBAL T* gptr(T* p) { return (QSFT_GLOBAL T*)p; }
// Data that no thread writes while a kernel runs may be read through the
// constant address space: wave-uniform reads then compile to scalar loads
// (the LDS/scalar wait counter, not the vector-memory one).
#define QSFT_CONST __attribute__((address_space(4)))
template <typename T> __device__ __forceinline__ const QSFT_CONST T* cptr(const T* p) { return (const QSFT_CONST T*)p; }
// Built-in vector types for such accesses (HIP's vector classes only copy in
// the generic address space).
typedef double gd2 __attribute__((ext_vector_type(2)));
typedef int gi4 __attribute__((ext_vector_type(4)));
typedef unsigned int gu4 __attribute__((ext_vector_type(4)));

// Largest i with pre[i] <= blk: the instance that owns work block blk. The
// prefix array is written by the host before the launch, so the search reads
// it with scalar loads (scalar cache).
__device__ __forceinline__ int find_inst(const int32_t* __restrict__ pre, int n_inst, int blk) {
    const __attribute__((address_space(4))) int32_t* cp = (const __attribute__((address_space(4))) int32_t*)pre;
    int lo = 0, hi = n_inst - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (cp[mid] <= blk) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__host__ __device__ __forceinline__ int ipow(int q, int e) {
    int r = 1;
    for (int i = 0; i < e; ++i) r *= q;
    return r;
}

// The matrix-core evaluation runs eight-wave workgroups (512 base queries)
// for q in {2, 4} with one digit block (n_pad = 64) and q^b >= 512, four-wave
// ones (256) otherwise; bit 11 of its launch word says which.
__host__ __device__ __forceinline__ bool mfma_wide(int q, int b, int n_pad) {
    return (q == 2 || q == 4) && n_pad == 64 && ipow(q, b) >= 512;
}

// x mod q for 0 <= x < 2^32 / q, by multiply-high with m = floor(2^32 / q) + 1:
// floor(x m / 2^32) = floor(x / q + x (m - 2^32/q) / 2^32) and the error term is
// < x / 2^32 < 1 / q, which cannot cross the next integer.
struct ModQ {
    uint32_t q, m;
    __host__ __device__ ModQ() : q(1), m(0) {}
    __host__ __device__ explicit ModQ(uint32_t q_) : q(q_), m((uint32_t)(0xFFFFFFFFull / q_) + 1u) {}
    __device__ __forceinline__ uint32_t operator()(uint32_t x) const {
        const uint32_t d = __umulhi(x, m);
        return x - d * q;
    }
};

template <typename T> struct cplx;
template <> struct cplx<float> { using type = float2; };
template <> struct cplx<double> { using type = double2; };
template <typename T> using cplx_t = typename cplx<T>::type;

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC 2011): multipliers 0xD2511F53 / 0xCD9E8D57, Weyl key
// increments 0x9E3779B9 / 0xBB67AE85, ten rounds.
struct u4 { uint32_t x, y, z, w; };
__host__ __device__ __forceinline__ u4 philox4x32_10(u4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c = u4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

// Two independent N(0, 1) draws from one Philox block (Box-Muller on two
// 53-bit uniforms; u1 in (0, 1] so the log is finite).
__device__ __forceinline__ double2 normal2(uint64_t counter, uint64_t seed) {
    const u4 r = philox4x32_10(u4{(uint32_t)counter, (uint32_t)(counter >> 32), 0u, 0u},
                               (uint32_t)seed, (uint32_t)(seed >> 32));
    const uint64_t a = ((uint64_t)r.x << 21) ^ (uint64_t)r.y;   // 53 significant bits
    const uint64_t b = ((uint64_t)r.z << 21) ^ (uint64_t)r.w;
    const double u1 = (double)((a & ((1ull << 53) - 1)) + 1) * 0x1.0p-53;
    const double u2 = (double)(b & ((1ull << 53) - 1)) * 0x1.0p-53;
    const double rad = sqrt(-2.0 * log(u1));
    double s, c;
    sincospi(2.0 * u2, &s, &c);
    return make_double2(rad * c, rad * s);
}

// Four independent N(0, 1) draws from one Philox block: Box-Muller in FP32 on
// two pairs of 24-bit uniforms, u1 in (0, 1) so the log is finite (largest
// radius sqrt(50 ln 2) = 5.9). v_log_f32 is log2 and v_sin_f32 / v_cos_f32
// take revolutions, so the angle needs no 2 pi scaling; v_sqrt_f32 (about
// 1 ulp) replaces the correctly rounded square root, whose expansion costs
// a dozen more instructions for no visible difference in a noise draw.
__device__ __forceinline__ float4 normal4(uint64_t counter, uint64_t seed) {
    const u4 r = philox4x32_10(u4{(uint32_t)counter, (uint32_t)(counter >> 32), 0u, 0u},
                               (uint32_t)seed, (uint32_t)(seed >> 32));
    const float s24 = 0x1.0p-24f;
    const float u1a = ((float)(r.x >> 8) + 0.5f) * s24, u2a = (float)(r.y >> 8) * s24;
    const float u1b = ((float)(r.z >> 8) + 0.5f) * s24, u2b = (float)(r.w >> 8) * s24;
    const float m2ln2 = -1.3862943611198906f;                      // -2 ln 2
    const float ra = __builtin_amdgcn_sqrtf(m2ln2 * __builtin_amdgcn_logf(u1a));
    const float rb = __builtin_amdgcn_sqrtf(m2ln2 * __builtin_amdgcn_logf(u1b));
    return make_float4(ra * __builtin_amdgcn_cosf(u2a), ra * __builtin_amdgcn_sinf(u2a),
                       rb * __builtin_amdgcn_cosf(u2b), rb * __builtin_amdgcn_sinf(u2b));
}

// Transform-domain noise of the working element (c, p, j) (row p of group c,
// bin j; `rows` rows per group, Q bins): one Philox block per bin pair
// (2h, 2h + 1) of a row, components (x, y) for the even bin and (z, w) for the
// odd one, each scaled by the per-component sd and added with one rounding.
// The assembler and the peeler's rebuild of an unwritten column use exactly
// this, so both give the same bits.
__device__ __forceinline__ uint64_t noise_counter(int c, int p, int rows, int64_t Q, int64_t j) {
    return ((uint64_t)c * rows + p) * (uint64_t)((Q + 1) >> 1) + (uint64_t)(j >> 1);
}
__device__ __forceinline__ double2 noise_add(double2 v, float zx, float zy, double comp_sd) {
    return make_double2(fma(comp_sd, (double)zx, v.x), fma(comp_sd, (double)zy, v.y));
}
__device__ __forceinline__ double2 add_philox(double2 v, int c, int p, int rows, int64_t Q, int64_t j,
                                              double comp_sd, uint64_t seed) {
    const float4 z = normal4(noise_counter(c, p, rows, Q, j), seed);
    return (j & 1) ? noise_add(v, z.z, z.w, comp_sd) : noise_add(v, z.x, z.y, comp_sd);
}

// Cross-lane moves for all-lane reductions, on the VALU (no LDS traffic).
// Each step pairs every lane with one partner and both compute the same
// a + b (commutative), so a reduction ends with the same bits in every lane.
// DPP controls: quad_perm [1,0,3,2] (lane ^ 1), quad_perm [2,3,0,1]
// (lane ^ 2), row_half_mirror (7 - i within 8 lanes), row_mirror (15 - i
// within 16 lanes); then v_permlane16_swap pairs rows 0/1 and 2/3, and
// v_permlane32_swap pairs the two halves of the wave. Every lane must be
// active.
template <int CTRL> __device__ __forceinline__ uint32_t dpp_u32(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xF, 0xF, false);
}
template <int CTRL> __device__ __forceinline__ double dpp_f64(double v) {
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const uint64_t r = (uint64_t)dpp_u32<CTRL>((uint32_t)b) | ((uint64_t)dpp_u32<CTRL>((uint32_t)(b >> 32)) << 32);
    return __builtin_bit_cast(double, r);
}
// (own row pair member a, partner b) such that both lanes of the pair see the
// values in the same order: .x from the lower row / half, .y from the upper
__device__ __forceinline__ void swap16_u64(uint64_t v, uint64_t& lo_part, uint64_t& hi_part) {
    const auto l = __builtin_amdgcn_permlane16_swap((uint32_t)v, (uint32_t)v, false, false);
    const auto h = __builtin_amdgcn_permlane16_swap((uint32_t)(v >> 32), (uint32_t)(v >> 32), false, false);
    lo_part = (uint64_t)l[0] | ((uint64_t)h[0] << 32);
    hi_part = (uint64_t)l[1] | ((uint64_t)h[1] << 32);
}
__device__ __forceinline__ void swap32_u64(uint64_t v, uint64_t& lo_part, uint64_t& hi_part) {
    const auto l = __builtin_amdgcn_permlane32_swap((uint32_t)v, (uint32_t)v, false, false);
    const auto h = __builtin_amdgcn_permlane32_swap((uint32_t)(v >> 32), (uint32_t)(v >> 32), false, false);
    lo_part = (uint64_t)l[0] | ((uint64_t)h[0] << 32);
    hi_part = (uint64_t)l[1] | ((uint64_t)h[1] << 32);
}

// Sum over the 64 lanes of a wave; every lane gets the same bits.
__device__ __forceinline__ double wave_sum(double v) {
    v += dpp_f64<0xB1>(v);
    v += dpp_f64<0x4E>(v);
    v += dpp_f64<0x141>(v);
    v += dpp_f64<0x140>(v);
    uint64_t a, b;
    swap16_u64(__builtin_bit_cast(uint64_t, v), a, b);
    v = __builtin_bit_cast(double, a) + __builtin_bit_cast(double, b);
    swap32_u64(__builtin_bit_cast(uint64_t, v), a, b);
    return __builtin_bit_cast(double, a) + __builtin_bit_cast(double, b);
}

// XOR over the 64 lanes of a wave.
__device__ __forceinline__ uint64_t wave_xor64(uint64_t v) {
    auto step = [](uint64_t x, uint64_t y) { return x ^ y; };
    v = step(v, (uint64_t)dpp_u32<0xB1>((uint32_t)v) | ((uint64_t)dpp_u32<0xB1>((uint32_t)(v >> 32)) << 32));
    v = step(v, (uint64_t)dpp_u32<0x4E>((uint32_t)v) | ((uint64_t)dpp_u32<0x4E>((uint32_t)(v >> 32)) << 32));
    v = step(v, (uint64_t)dpp_u32<0x141>((uint32_t)v) | ((uint64_t)dpp_u32<0x141>((uint32_t)(v >> 32)) << 32));
    v = step(v, (uint64_t)dpp_u32<0x140>((uint32_t)v) | ((uint64_t)dpp_u32<0x140>((uint32_t)(v >> 32)) << 32));
    uint64_t a, b;
    swap16_u64(v, a, b);
    swap32_u64(a ^ b, a, b);
    return a ^ b;
}

// Sum over the 16 lanes of each row (lanes 16 r .. 16 r + 15).
__device__ __forceinline__ int row_sum_i(int v) {
    v += (int)dpp_u32<0xB1>((uint32_t)v);
    v += (int)dpp_u32<0x4E>((uint32_t)v);
    v += (int)dpp_u32<0x141>((uint32_t)v);
    v += (int)dpp_u32<0x140>((uint32_t)v);
    return v;
}

// Rows per lane of a bin column held by one wave: row p = lane + 64 r (P <= 256).
constexpr int kColRowsPerLane = 4;

// |col|^2 as every zeroton test computes it: per-lane sums of re^2 + im^2 over
// the lane's rows in order with explicit FMAs, then one xor butterfly (the
// reference's np.linalg.norm(col) ** 2 up to summation order and the rounding
// of its square root and square, a few ulp). The assembler's energies, the
// peeler's screen and its per-pass detection all call this, so they classify
// a column alike.
template <int N> __device__ __forceinline__ double bin_energy(const double2 (&u)[N]) {
    double s = 0.0;
#pragma unroll
    for (int r = 0; r < N; ++r) {
        s = __fma_rn(u[r].x, u[r].x, s);
        s = __fma_rn(u[r].y, u[r].y, s);
    }
    return wave_sum(s);
}

// LDS byte address of a __shared__ object
__device__ __forceinline__ uint32_t lds_addr(const void* p) {
    return (uint32_t)(uintptr_t)(const __attribute__((address_space(3))) char*)p;
}

// 16 bytes per lane from global memory straight into LDS: lane L writes
// lds_base + 16 L (lds_base wave-uniform). Issued from asm so the compiler
// does not track the LDS write: its own tracking cannot tell separate LDS
// buffers apart and would wait for these loads before every read of another
// buffer.
// The instruction writes no VGPR; m0 is read when it issues and is restored
// right after. Every read of a destination slot follows an explicit wait_vm.
// The global address is a wave-uniform base (SGPR pair) plus a 32-bit
// per-lane byte offset.
__device__ __forceinline__ uint64_t uniform_base(const void* sbase) {
    // the base is wave-uniform by contract; readfirstlane puts it in SGPRs. It
    // returns int: widen through uint32_t, or a low word >= 2^31 sign-extends
    // into the high word and the address is garbage.
    const uint64_t bv = (uint64_t)(uintptr_t)sbase;
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(bv >> 32));
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)bv);
    return ((uint64_t)hi << 32) | (uint64_t)lo;
}

// dma16 with the global base already wave-uniform (uniform_base)
__device__ __forceinline__ void dma16_u(uint64_t bs, uint32_t voff, uint32_t lds_base) {
    uint32_t saved;
    asm volatile(
        "s_mov_b32 %0, m0\n\t"
        "s_mov_b32 m0, %3\n\t"
        "s_nop 0\n\t"
        "global_load_lds_dwordx4 %1, %2\n\t"
        "s_mov_b32 m0, %0"
        : "=&s"(saved)
        : "v"(voff), "s"(bs), "s"(lds_base)
        : "memory");
}

__device__ __forceinline__ void dma16(const void* sbase, uint32_t voff, uint32_t lds_base) {
    const uint64_t bs = uniform_base(sbase);
    uint32_t saved;
    asm volatile(
        "s_mov_b32 %0, m0\n\t"
        "s_mov_b32 m0, %3\n\t"
        "s_nop 0\n\t"
        "global_load_lds_dwordx4 %1, %2\n\t"
        "s_mov_b32 m0, %0"
        : "=&s"(saved)
        : "v"(voff), "s"(bs), "s"(lds_base)
        : "memory");
}

// The same for 4 bytes, issued by lane 0 alone, which writes lds_base.
__device__ __forceinline__ void dma4(const void* sbase, uint32_t voff, uint32_t lds_base) {
    const uint64_t bs = uniform_base(sbase);
    uint32_t saved;
    asm volatile(
        "s_mov_b32 %0, m0\n\t"
        "s_mov_b32 m0, %3\n\t"
        "s_nop 0\n\t"
        "global_load_lds_dword %1, %2\n\t"
        "s_mov_b32 m0, %0"
        : "=&s"(saved)
        : "v"(voff), "s"(bs), "s"(lds_base)
        : "memory");
}

// s_waitcnt immediate: vector-memory count <= n, the other counters unconstrained (gfx9 encoding)
template <int N>
__device__ __forceinline__ void wait_vm() {
    static_assert(N >= 0 && N < 64, "vmcnt range");
    __builtin_amdgcn_s_waitcnt((N & 0xF) | ((N >> 4) << 14) | (0x7 << 4) | (0xF << 8));
}

// 64-bit mix of a digit vector (FNV-1a over the bytes, then the splitmix64
// finaliser); equal vectors give equal hashes.
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

}  // namespace qsft
