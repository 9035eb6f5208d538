// Stage (a)+(b): query generation and sparse-signal evaluation on gfx950.
//
// For every query m of a coset {M_c l + d_p mod q : l in Z_q^b} the signal
// x(m) = sum_s a_s w^{<m, k_s>} (w = e^{2 pi i / q}) is evaluated the way the
// reference's SyntheticSubsampledSignal.subsample does
// (synt_exp/synt_src/synthetic_signal.py:97-118), with the phase exponent
// <m, k_s> computed exactly:
//
//   * one wave owns 16 queries; each lane builds 16 (or 32/48/64) query digits
//     of one query in registers (stage (a): exact integer work, never stored);
//   * per block of 16 supports, v_mfma_i32_16x16x64_i8 forms E = m . k^T over
//     the digit axis (int8 x int8 -> int32, exact: E <= n (q-1)^2);
//   * each lane reduces its 4 entries of E mod q, reads w^e from an LDS table
//     and accumulates a_s w^e in FP32 or FP64 (the VALU part that bounds the
//     kernel);
//   * the 16 lanes that share a query row are reduced with xor-shuffles.
//
// MFMA fragment layout (16x16x64 int8): lane L supplies row L%16 of A and
// column L%16 of B, each as 16 consecutive k-bytes of k-group L/16; the
// int32 result D[i][j] sits in lane j + 16*(i/4), register i%4. A and B use
// the same lane/k mapping, so any fixed digit order inside a 64-digit block
// gives the same dot product.
#include "qsft_common.h"

namespace qsft {

typedef int v4i __attribute__((ext_vector_type(4)));

template <typename T>
__device__ __forceinline__ void fill_wtable(cplx_t<T>* tab, int q) {
    for (int r = threadIdx.x; r < q; r += blockDim.x) {
        double s, c;
        sincospi(2.0 * (double)r / (double)q, &s, &c);
        tab[r].x = (T)c;
        tab[r].y = (T)s;
    }
}

// Pack 16 digits (values < 128) into a 16-byte MFMA fragment.
__device__ __forceinline__ v4i pack16(const int (&d)[16]) {
    v4i r;
#pragma unroll
    for (int w = 0; w < 4; ++w)
        r[w] = (d[4 * w] & 0xFF) | ((d[4 * w + 1] & 0xFF) << 8) | ((d[4 * w + 2] & 0xFF) << 16) |
               ((d[4 * w + 3] & 0xFF) << 24);
    return r;
}

__device__ __forceinline__ int byte_of(const v4i& v, int j) {
    return (v[j >> 2] >> (8 * (j & 3))) & 0xFF;
}

// Accumulate sum_s a_s w^{E mod q} for the wave's 16 queries over all S
// supports; afrag holds this lane's query digits. Returns this lane's 4
// partial sums (rows 4*(lane/16)+v, the column of this lane).
template <typename T, bool POW2, int NKB>
__device__ __forceinline__ void accumulate(const v4i (&afrag)[NKB], const int8_t* __restrict__ K,
                                           const cplx_t<T>* __restrict__ a, int S, int n_pad,
                                           const cplx_t<T>* tab, ModQ mq, uint32_t qmask,
                                           cplx_t<T> (&acc)[4]) {
    const int lane = threadIdx.x & 63;
    const int g = lane >> 4, r16 = lane & 15;
#pragma unroll
    for (int v = 0; v < 4; ++v) acc[v].x = acc[v].y = (T)0;
    for (int s0 = 0; s0 < S; s0 += 16) {
        const int s = s0 + r16;
        const bool ok = s < S;
        v4i e = {0, 0, 0, 0};
#pragma unroll
        for (int kb = 0; kb < NKB; ++kb) {
            v4i bf = {0, 0, 0, 0};
            if (ok) bf = *reinterpret_cast<const v4i*>(K + (size_t)s * n_pad + 64 * kb + 16 * g);
            e = __builtin_amdgcn_mfma_i32_16x16x64_i8(afrag[kb], bf, e, 0, 0, 0);
        }
        cplx_t<T> as;
        as.x = as.y = (T)0;
        if (ok) as = a[s];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const uint32_t r = POW2 ? ((uint32_t)e[v] & qmask) : mq((uint32_t)e[v]);
            const cplx_t<T> w = tab[r];
            acc[v].x += as.x * w.x - as.y * w.y;
            acc[v].y += as.x * w.y + as.y * w.x;
        }
    }
    // reduce over the 16 lanes (columns) that share the same 4 rows
#pragma unroll
    for (int off = 1; off < 16; off <<= 1) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            acc[v].x += __shfl_xor(acc[v].x, off);
            acc[v].y += __shfl_xor(acc[v].y, off);
        }
    }
}

template <typename T>
__device__ __forceinline__ cplx_t<T> pick4(const cplx_t<T> (&acc)[4], int v) {
    cplx_t<T> r = acc[0];
    if (v == 1) r = acc[1];
    if (v == 2) r = acc[2];
    if (v == 3) r = acc[3];
    return r;
}

template <typename T, bool POW2, int NKB>
__global__ __launch_bounds__(256) void k_eval_cosets(const qsft_eval_desc* __restrict__ descs,
                                                     const int32_t* __restrict__ pre, int n_inst) {
    using C2 = cplx_t<T>;
    __shared__ C2 tab[QSFT_MAX_Q];
    const int inst = find_inst(pre, n_inst, blockIdx.x);
    const qsft_eval_desc& d = descs[inst];
    const int q = d.q, b = d.b, n_pad = d.n_pad;
    const int Q = ipow(q, b);
    const int tiles = (Q + 63) >> 6;
    const int local = blockIdx.x - pre[inst];
    const int row = local / tiles;               // c * P + p
    const int tile = local - row * tiles;
    const int c = row / d.P;
    fill_wtable<T>(tab, q);
    __syncthreads();

    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int g = lane >> 4, r16 = lane & 15;
    const int lq0 = tile * 64 + wave * 16;
    const int l = lq0 + r16;                     // this lane's query (A row)
    const ModQ mq((uint32_t)q);
    const int8_t* __restrict__ Mt = d.Mt + (size_t)c * b * n_pad;
    const int8_t* __restrict__ drow = d.D + (size_t)row * n_pad;

    v4i afrag[NKB];
#pragma unroll
    for (int kb = 0; kb < NKB; ++kb) {
        const int i0 = 64 * kb + 16 * g;         // first digit of this lane's slice
        const v4i dv = *reinterpret_cast<const v4i*>(drow + i0);
        int acc[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) acc[j] = byte_of(dv, j);
        int rem = l, pw = Q;
        for (int t = 0; t < b; ++t) {            // digit l_t of l, MSB first
            pw /= q;
            const int lt = rem / pw;
            rem -= lt * pw;
            const v4i mv = *reinterpret_cast<const v4i*>(Mt + (size_t)t * n_pad + i0);
#pragma unroll
            for (int j = 0; j < 16; ++j) acc[j] += lt * byte_of(mv, j);
        }
#pragma unroll
        for (int j = 0; j < 16; ++j) acc[j] = (int)mq((uint32_t)acc[j]);
        afrag[kb] = pack16(acc);
    }

    C2 sum[4];
    accumulate<T, POW2, NKB>(afrag, d.K, reinterpret_cast<const C2*>(d.a), d.S, n_pad, tab, mq,
                             (uint32_t)(q - 1), sum);
    if (r16 < 4) {
        const int ql = lq0 + 4 * g + r16;
        if (ql < Q) reinterpret_cast<C2*>(d.x)[(size_t)row * Q + ql] = pick4<T>(sum, r16);
    }
}

template <typename T, bool POW2, int NKB>
__global__ __launch_bounds__(256) void k_eval_points(const qsft_eval_points_desc* __restrict__ descs,
                                                     const int32_t* __restrict__ pre, int n_inst) {
    using C2 = cplx_t<T>;
    __shared__ C2 tab[QSFT_MAX_Q];
    const int inst = find_inst(pre, n_inst, blockIdx.x);
    const qsft_eval_points_desc& d = descs[inst];
    const int q = d.q, n_pad = d.n_pad;
    const int64_t local = blockIdx.x - pre[inst];
    fill_wtable<T>(tab, q);
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int g = lane >> 4, r16 = lane & 15;
    const int64_t lq0 = local * 64 + wave * 16;
    const int64_t l = lq0 + r16;
    const ModQ mq((uint32_t)q);
    v4i afrag[NKB];
#pragma unroll
    for (int kb = 0; kb < NKB; ++kb) {
        v4i z = {0, 0, 0, 0};
        if (l < d.N) z = *reinterpret_cast<const v4i*>(d.m + (size_t)l * n_pad + 64 * kb + 16 * g);
        afrag[kb] = z;
    }
    C2 sum[4];
    accumulate<T, POW2, NKB>(afrag, d.K, reinterpret_cast<const C2*>(d.a), d.S, n_pad, tab, mq,
                             (uint32_t)(q - 1), sum);
    if (r16 < 4) {
        const int64_t ql = lq0 + 4 * g + r16;
        if (ql < d.N) reinterpret_cast<C2*>(d.x)[ql] = pick4<T>(sum, r16);
    }
}


// Stage (a) alone: write the digits of every query (M_c l + d_p) mod q as
// int8 [C][P][Q][n_pad] (for user callbacks that need the indices).
__global__ __launch_bounds__(256) void k_query_digits(const qsft_eval_desc* __restrict__ descs,
                                                      const int32_t* __restrict__ pre, int n_inst) {
    const int inst = find_inst(pre, n_inst, blockIdx.x);
    const qsft_eval_desc& d = descs[inst];
    const int q = d.q, b = d.b, n = d.n, n_pad = d.n_pad;
    const int Q = ipow(q, b);
    const int tiles = (Q + 255) >> 8;
    const int local = blockIdx.x - pre[inst];
    const int row = local / tiles;
    const int l = (local - row * tiles) * 256 + threadIdx.x;
    if (l >= Q) return;
    const int c = row / d.P;
    const ModQ mq((uint32_t)q);
    const int8_t* __restrict__ Mt = d.Mt + (size_t)c * b * n_pad;
    const int8_t* __restrict__ drow = d.D + (size_t)row * n_pad;
    int8_t* __restrict__ out = reinterpret_cast<int8_t*>(d.x) + ((size_t)row * Q + l) * n_pad;
    for (int i = 0; i < n_pad; ++i) {
        uint32_t acc = 0;
        if (i < n) {
            acc = (uint32_t)drow[i];
            int rem = l, pw = Q;
            for (int t = 0; t < b; ++t) {
                pw /= q;
                const int lt = rem / pw;
                rem -= lt * pw;
                acc += (uint32_t)(lt * Mt[(size_t)t * n_pad + i]);
            }
            acc = mq(acc);
        }
        out[i] = (int8_t)acc;
    }
}

}  // namespace qsft

using namespace qsft;

namespace {

template <template <typename, bool, int> class Dummy>
struct unused {};

template <typename T, bool POW2>
hipError_t launch_cosets(int nkb, const qsft_eval_desc* d, const int32_t* pre, int n_inst, int blocks,
                         hipStream_t st) {
    switch (nkb) {
        case 1: hipLaunchKernelGGL((k_eval_cosets<T, POW2, 1>), dim3(blocks), dim3(256), 0, st, d, pre, n_inst); break;
        case 2: hipLaunchKernelGGL((k_eval_cosets<T, POW2, 2>), dim3(blocks), dim3(256), 0, st, d, pre, n_inst); break;
        case 3: hipLaunchKernelGGL((k_eval_cosets<T, POW2, 3>), dim3(blocks), dim3(256), 0, st, d, pre, n_inst); break;
        default: hipLaunchKernelGGL((k_eval_cosets<T, POW2, 4>), dim3(blocks), dim3(256), 0, st, d, pre, n_inst); break;
    }
    return hipGetLastError();
}

template <typename T, bool POW2>
hipError_t launch_points(int nkb, const qsft_eval_points_desc* d, const int32_t* pre, int n_inst,
                         int blocks, hipStream_t st) {
    switch (nkb) {
        case 1: hipLaunchKernelGGL((k_eval_points<T, POW2, 1>), dim3(blocks), dim3(256), 0, st, d, pre, n_inst); break;
        case 2: hipLaunchKernelGGL((k_eval_points<T, POW2, 2>), dim3(blocks), dim3(256), 0, st, d, pre, n_inst); break;
        case 3: hipLaunchKernelGGL((k_eval_points<T, POW2, 3>), dim3(blocks), dim3(256), 0, st, d, pre, n_inst); break;
        default: hipLaunchKernelGGL((k_eval_points<T, POW2, 4>), dim3(blocks), dim3(256), 0, st, d, pre, n_inst); break;
    }
    return hipGetLastError();
}

}  // namespace

// implemented in abi.cpp
extern "C" void qsft_set_error(const char* msg);
extern "C" int qsft_hip_status(hipError_t e, const char* where);

// A launch covers instances that share q's power-of-two class and the digit
// block count; the binding groups instances accordingly (one launch per key).
extern "C" int qsft_eval_blocks(const qsft_eval_desc* h) {
    if (!h || h->q < 2 || h->q > QSFT_MAX_Q || h->b < 1 || h->b > QSFT_MAX_B || h->n < 1 ||
        h->n > h->n_pad || h->n_pad % 64 || h->n_pad > QSFT_MAX_N || h->C < 1 || h->P < 1 || h->S < 0)
        return -1;
    const long long Q = (long long)ipow(h->q, h->b);
    const long long blocks = (long long)h->C * h->P * ((Q + 63) / 64);
    return blocks > 0x7FFFFFFF ? -1 : (int)blocks;
}

extern "C" int qsft_eval_points_blocks(const qsft_eval_points_desc* h) {
    if (!h || h->q < 2 || h->q > QSFT_MAX_Q || h->n < 1 || h->n > h->n_pad || h->n_pad % 64 ||
        h->n_pad > QSFT_MAX_N || h->N < 0 || h->S < 0)
        return -1;
    const long long blocks = (h->N + 63) / 64;
    return blocks > 0x7FFFFFFF ? -1 : (int)blocks;
}

// `precision` selects the accumulation type; bit 8 marks power-of-two q for
// every instance in the launch; bits 4..7 give the 64-digit block count (1..4).
extern "C" int qsft_eval_cosets(const qsft_eval_desc* d, const int32_t* pre, int n_inst, int blocks,
                                int precision, void* stream) {
    if (!d || !pre || n_inst < 1 || blocks < 0) { qsft_set_error("qsft_eval_cosets: bad arguments"); return QSFT_ERR_ARG; }
    if (blocks == 0) return QSFT_OK;
    const int prec = precision & 0xF;
    const int nkb = (precision >> 4) & 0xF;
    const bool pow2 = (precision >> 8) & 1;
    if (nkb < 1 || nkb > 4 || (prec != QSFT_F32 && prec != QSFT_F64)) {
        qsft_set_error("qsft_eval_cosets: bad precision word");
        return QSFT_ERR_ARG;
    }
    hipStream_t st = (hipStream_t)stream;
    hipError_t e;
    if (prec == QSFT_F32) e = pow2 ? launch_cosets<float, true>(nkb, d, pre, n_inst, blocks, st)
                                   : launch_cosets<float, false>(nkb, d, pre, n_inst, blocks, st);
    else e = pow2 ? launch_cosets<double, true>(nkb, d, pre, n_inst, blocks, st)
                  : launch_cosets<double, false>(nkb, d, pre, n_inst, blocks, st);
    return qsft_hip_status(e, "qsft_eval_cosets");
}

extern "C" int qsft_eval_points(const qsft_eval_points_desc* d, const int32_t* pre, int n_inst, int blocks,
                                int precision, void* stream) {
    if (!d || !pre || n_inst < 1 || blocks < 0) { qsft_set_error("qsft_eval_points: bad arguments"); return QSFT_ERR_ARG; }
    if (blocks == 0) return QSFT_OK;
    const int prec = precision & 0xF;
    const int nkb = (precision >> 4) & 0xF;
    const bool pow2 = (precision >> 8) & 1;
    if (nkb < 1 || nkb > 4 || (prec != QSFT_F32 && prec != QSFT_F64)) {
        qsft_set_error("qsft_eval_points: bad precision word");
        return QSFT_ERR_ARG;
    }
    hipStream_t st = (hipStream_t)stream;
    hipError_t e;
    if (prec == QSFT_F32) e = pow2 ? launch_points<float, true>(nkb, d, pre, n_inst, blocks, st)
                                   : launch_points<float, false>(nkb, d, pre, n_inst, blocks, st);
    else e = pow2 ? launch_points<double, true>(nkb, d, pre, n_inst, blocks, st)
                  : launch_points<double, false>(nkb, d, pre, n_inst, blocks, st);
    return qsft_hip_status(e, "qsft_eval_points");
}

extern "C" int qsft_query_digits_blocks(const qsft_eval_desc* h) {
    if (qsft_eval_blocks(h) < 0) return -1;
    const long long Q = (long long)ipow(h->q, h->b);
    const long long blocks = (long long)h->C * h->P * ((Q + 255) / 256);
    return blocks > 0x7FFFFFFF ? -1 : (int)blocks;
}

extern "C" int qsft_query_digits(const qsft_eval_desc* d, const int32_t* pre, int n_inst, int blocks, void* stream) {
    if (!d || !pre || n_inst < 1 || blocks < 0) { qsft_set_error("qsft_query_digits: bad arguments"); return QSFT_ERR_ARG; }
    if (blocks == 0) return QSFT_OK;
    hipLaunchKernelGGL(k_query_digits, dim3(blocks), dim3(256), 0, (hipStream_t)stream, d, pre, n_inst);
    return qsft_hip_status(hipGetLastError(), "qsft_query_digits");
}
