// Stage (c): forward-scaled multidimensional q-point DFTs on gfx950.
//
// A row of Q = q^b samples is a [q]*b tensor (first axis most significant,
// as numpy's reshape([q]*b)); U = fftn(x) / q^b applies a q-point DFT
// X[k] = sum_r x[r] e^{-2 pi i r k / q} along every axis -- there are no
// inter-axis twiddles. A pass transforms the axes [d0, d1) of every row: the
// row is viewed as [A, M, I] (A = q^d0, M = q^(d1-d0), I = q^(b-d1)), and
// one workgroup stages a tile of M x TI elements (TI consecutive inner
// indices, so global loads/stores are contiguous runs of TI elements) in LDS,
// runs the mb = d1-d0 axis transforms there with a barrier between axes, and
// writes the tile back. One pass suffices when the whole row fits the LDS
// budget (64 KiB: 8192 FP32 or 4096 FP64 complex); larger rows take two or
// more passes. The last pass applies the 1/q^b scale (a division in FP64,
// like numpy's `/ q**n`; a multiply by the rounded reciprocal in FP32).
//
// HBM traffic per pass is one read and one write of the rows: 16 B per
// complex FP32 sample (8 in + 8 out), 32 B in FP64.
#include "qsft_common.h"

namespace qsft {

constexpr int kDftThreads = 256;
constexpr int kDftLdsBytes = 65536;                 // dynamic LDS per workgroup
constexpr int kDftTwBytes = QSFT_MAX_Q * 16;        // twiddle table at the end of it
constexpr int kDftTileBytes = kDftLdsBytes - kDftTwBytes;

template <typename T>
__device__ __forceinline__ cplx_t<T> cmul(cplx_t<T> a, cplx_t<T> b) {
    cplx_t<T> r;
    r.x = a.x * b.x - a.y * b.y;
    r.y = a.x * b.y + a.y * b.x;
    return r;
}

// In-place q-point forward DFT of the pencil v[0], v[st], ..., v[(q-1) st].
template <typename T, int QR>
__device__ __forceinline__ void pencil(cplx_t<T>* v, int st, const cplx_t<T>* tw) {
    using C2 = cplx_t<T>;
    if constexpr (QR == 2) {
        const C2 a = v[0], b = v[st];
        v[0] = C2{a.x + b.x, a.y + b.y};
        v[st] = C2{a.x - b.x, a.y - b.y};
    } else if constexpr (QR == 4) {
        const C2 x0 = v[0], x1 = v[st], x2 = v[2 * st], x3 = v[3 * st];
        const C2 s02{x0.x + x2.x, x0.y + x2.y}, d02{x0.x - x2.x, x0.y - x2.y};
        const C2 s13{x1.x + x3.x, x1.y + x3.y}, d13{x1.x - x3.x, x1.y - x3.y};
        v[0] = C2{s02.x + s13.x, s02.y + s13.y};
        v[2 * st] = C2{s02.x - s13.x, s02.y - s13.y};
        // (-i) * d13 = (d13.y, -d13.x)
        v[st] = C2{d02.x + d13.y, d02.y - d13.x};
        v[3 * st] = C2{d02.x - d13.y, d02.y + d13.x};
    } else {
        C2 x[QR];
#pragma unroll
        for (int r = 0; r < QR; ++r) x[r] = v[r * st];
#pragma unroll
        for (int k = 0; k < QR; ++k) {
            C2 acc = x[0];
#pragma unroll
            for (int r = 1; r < QR; ++r) {
                const C2 w = tw[(r * k) % QR];
                acc.x += x[r].x * w.x - x[r].y * w.y;
                acc.y += x[r].x * w.y + x[r].y * w.x;
            }
            v[k * st] = acc;
        }
    }
}

// Generic q (not compiled for a fixed radix): out-of-place through a second
// LDS buffer, one output per thread iteration.
template <typename T>
__device__ __forceinline__ void pencil_generic_out(const cplx_t<T>* src, cplx_t<T>* dst, int st, int q,
                                                   const cplx_t<T>* tw) {
    for (int k = 0; k < q; ++k) {
        cplx_t<T> acc = src[0];
        for (int r = 1; r < q; ++r) {
            const cplx_t<T> w = tw[(r * k) % q];
            const cplx_t<T> x = src[r * st];
            acc.x += x.x * w.x - x.y * w.y;
            acc.y += x.x * w.y + x.y * w.x;
        }
        dst[k * st] = acc;
    }
}

template <typename T, int QR>
__global__ __launch_bounds__(kDftThreads) void k_dft_pass(const qsft_dft_desc* __restrict__ descs,
                                                          const int32_t* __restrict__ pre, int n_inst,
                                                          int pass, int lds_bytes) {
    using C2 = cplx_t<T>;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C2* lds = reinterpret_cast<C2*>(smem_raw);

    const int inst = find_inst(pre, n_inst, blockIdx.x);
    const qsft_dft_desc& d = descs[inst];
    if (pass >= d.n_pass) return;
    const int q = (QR > 0) ? QR : d.q;
    const int b = d.b;
    const int d0 = d.pass_d0[pass], d1 = d.pass_d1[pass], ti = d.pass_ti[pass];
    const int mb = d1 - d0;
    const int Q = ipow(q, b);
    const int A = ipow(q, d0), M = ipow(q, mb), I = ipow(q, b - d1);
    const int ichunks = I / ti;
    const int local = blockIdx.x - pre[inst];
    const int tiles_per_row = A * ichunks;
    const int row = local / tiles_per_row;
    const int rt = local - row * tiles_per_row;
    const int a_idx = rt / ichunks;
    const int i0 = (rt - a_idx * ichunks) * ti;
    const int64_t base = (int64_t)a_idx * M * I + i0;       // element offset inside the row
    const int E = M * ti;
    // twiddles follow the tile (two tiles for the generic radix's ping-pong)
    C2* tw = lds + (QR > 0 ? E : 2 * E);
    if ((int)(((QR > 0 ? E : 2 * E) + q) * sizeof(C2)) > lds_bytes) return;    // launch sized too small

    for (int r = threadIdx.x; r < q; r += blockDim.x) {
        double s, c;
        sincospi(-2.0 * (double)r / (double)q, &s, &c);
        tw[r].x = (T)c;
        tw[r].y = (T)s;
    }
    // load the tile
    if (pass == 0) {
        const C2* __restrict__ in = reinterpret_cast<const C2*>(d.in) + (int64_t)row * d.in_row_stride;
        for (int e = threadIdx.x; e < E; e += blockDim.x) {
            const int m = e / ti, ii = e - m * ti;
            lds[e] = in[(base + (int64_t)m * I + ii) * d.in_stride];
        }
    } else {
        const C2* __restrict__ in = reinterpret_cast<const C2*>(d.out) + (int64_t)row * Q;
        for (int e = threadIdx.x; e < E; e += blockDim.x) {
            const int m = e / ti, ii = e - m * ti;
            lds[e] = in[base + (int64_t)m * I + ii];
        }
    }
    __syncthreads();

    C2* cur = lds;
    C2* alt = lds + E;            // only used by the generic radix
    const int pencils = (M / q) * ti;
    for (int t = 0; t < mb; ++t) {
        const int st = ipow(q, mb - 1 - t);          // stride of local axis t in M units
        for (int pc = threadIdx.x; pc < pencils; pc += blockDim.x) {
            const int u = pc / ti, ii = pc - u * ti;
            const int hi = u / st, lo = u - hi * st;
            const int m0 = hi * st * q + lo;
            if constexpr (QR > 0) {
                pencil<T, QR>(cur + m0 * ti + ii, st * ti, tw);
            } else {
                pencil_generic_out<T>(cur + m0 * ti + ii, alt + m0 * ti + ii, st * ti, q, tw);
            }
        }
        __syncthreads();
        if constexpr (QR == 0) {
            C2* tmp = cur;
            cur = alt;
            alt = tmp;
        }
    }

    const bool last = (pass == d.n_pass - 1);
    C2* __restrict__ out = reinterpret_cast<C2*>(d.out) + (int64_t)row * Q;
    if (last) {
        if constexpr (sizeof(T) == 8) {
            const double qd = (double)Q;
            for (int e = threadIdx.x; e < E; e += blockDim.x) {
                const int m = e / ti, ii = e - m * ti;
                C2 v = cur[e];
                v.x = v.x / qd;
                v.y = v.y / qd;
                out[base + (int64_t)m * I + ii] = v;
            }
        } else {
            const T sc = (T)(1.0 / (double)Q);
            for (int e = threadIdx.x; e < E; e += blockDim.x) {
                const int m = e / ti, ii = e - m * ti;
                C2 v = cur[e];
                v.x *= sc;
                v.y *= sc;
                out[base + (int64_t)m * I + ii] = v;
            }
        }
    } else {
        for (int e = threadIdx.x; e < E; e += blockDim.x) {
            const int m = e / ti, ii = e - m * ti;
            out[base + (int64_t)m * I + ii] = cur[e];
        }
    }
}

// q = 4, whole rows of Q = 4^B samples (B even), one workgroup per row,
// 4^B / 16 threads. Base-4 digit t of the element index sits at bits 2t, 2t+1.
// A thread holds 16 samples that differ in two digits, so two of the axis
// transforms run in registers; two digits per register pass, B / 2 passes,
// the samples re-dealt through LDS between passes (one pad element per 16
// keeps the strided LDS accesses spread over the banks). The first pass reads
// 16 consecutive samples per thread, the last stores 64 consecutive samples
// per wave instruction. Same 4-point butterflies and scaling as k_dft_pass.
template <typename T>
__device__ __forceinline__ void dft4(cplx_t<T>& x0, cplx_t<T>& x1, cplx_t<T>& x2, cplx_t<T>& x3) {
    using C2 = cplx_t<T>;
    const C2 s02{x0.x + x2.x, x0.y + x2.y}, d02{x0.x - x2.x, x0.y - x2.y};
    const C2 s13{x1.x + x3.x, x1.y + x3.y}, d13{x1.x - x3.x, x1.y - x3.y};
    x0 = C2{s02.x + s13.x, s02.y + s13.y};
    x2 = C2{s02.x - s13.x, s02.y - s13.y};
    x1 = C2{d02.x + d13.y, d02.y - d13.x};                        // d02 + (-i) d13
    x3 = C2{d02.x - d13.y, d02.y + d13.x};
}

// element index of register j of thread tid when the thread holds digits 2p, 2p+1
__device__ __forceinline__ int dft4_index(int p, int tid, int j) {
    const int lo = 4 * p;
    return ((tid >> lo) << (lo + 4)) | (j << lo) | (tid & ((1 << lo) - 1));
}
__device__ __forceinline__ int dft4_pad(int i) { return i + (i >> 4); }

template <typename T, int B>
__global__ __launch_bounds__(1 << (2 * B - 4)) void k_dft4_rows(const qsft_dft_desc* __restrict__ descs,
                                                                const int32_t* __restrict__ pre, int n_inst) {
    using C2 = cplx_t<T>;
    constexpr int Q = 1 << (2 * B);
    static_assert(B % 2 == 0 && B >= 4, "even digit count");
    __shared__ C2 lds[Q + Q / 16];
    const int inst = find_inst(pre, n_inst, blockIdx.x);
    const qsft_dft_desc& d = descs[inst];
    const int row = blockIdx.x - pre[inst];
    if (d.q != 4 || d.b != B || d.in_stride != 1 || row >= d.rows) return;   // launched for other shapes
    const int tid = threadIdx.x;
    const C2* __restrict__ in = reinterpret_cast<const C2*>(d.in) + (int64_t)row * d.in_row_stride;
    C2* __restrict__ out = reinterpret_cast<C2*>(d.out) + (int64_t)row * Q;

    C2 v[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) v[j] = in[tid * 16 + j];
#pragma unroll
    for (int p = 0; p < B / 2; ++p) {
        if (p > 0) {
            __syncthreads();                                       // the previous reads of lds are done
#pragma unroll
            for (int j = 0; j < 16; ++j) lds[dft4_pad(dft4_index(p - 1, tid, j))] = v[j];
            __syncthreads();
#pragma unroll
            for (int j = 0; j < 16; ++j) v[j] = lds[dft4_pad(dft4_index(p, tid, j))];
        }
#pragma unroll
        for (int a = 0; a < 4; ++a) dft4<T>(v[4 * a], v[4 * a + 1], v[4 * a + 2], v[4 * a + 3]);   // digit 2p
#pragma unroll
        for (int a = 0; a < 4; ++a) dft4<T>(v[a], v[a + 4], v[a + 8], v[a + 12]);                 // digit 2p+1
    }
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        C2 x = v[j];
        if constexpr (sizeof(T) == 8) {
            x.x = x.x / (double)Q;
            x.y = x.y / (double)Q;
        } else {
            const T sc = (T)(1.0 / (double)Q);
            x.x *= sc;
            x.y *= sc;
        }
        out[dft4_index(B / 2 - 1, tid, j)] = x;
    }
}

}  // namespace qsft

using namespace qsft;

extern "C" void qsft_set_error(const char* msg);
extern "C" int qsft_hip_status(hipError_t e, const char* where);

static bool radix_compiled(int q) { return q == 2 || q == 3 || q == 4 || q == 5 || q == 20; }

// Split the b axes into passes, innermost first: each pass takes as many
// axes as fit the LDS budget with TI contiguous inner elements per tile.
extern "C" int qsft_dft_plan(qsft_dft_desc* h, int precision) {
    if (!h || h->q < 2 || h->q > QSFT_MAX_Q || h->b < 1 || h->b > QSFT_MAX_B || h->rows < 0 ||
        h->in_stride < 1 || (precision != QSFT_F32 && precision != QSFT_F64)) {
        qsft_set_error("qsft_dft_plan: bad descriptor");
        return QSFT_ERR_ARG;
    }
    const int q = h->q, b = h->b;
    const long long elem = (precision == QSFT_F32) ? 8 : 16;
    long long cap = kDftTileBytes / elem;
    if (!radix_compiled(q)) cap /= 2;                         // generic radix ping-pongs
    int np = 0;
    int d1 = b;
    while (d1 > 0) {
        if (np == QSFT_DFT_MAX_PASSES) { qsft_set_error("qsft_dft_plan: too many passes"); return QSFT_ERR_ARG; }
        const long long I = ipow(q, b - d1);
        // widest divisor of I up to one 128-byte line that leaves room for one axis
        const long long ti_max = 128 / elem;
        long long ti = 1;
        for (long long c = (I < ti_max ? I : ti_max); c >= 1; --c)
            if (I % c == 0 && (long long)q * c <= cap) { ti = c; break; }
        if ((long long)q * ti > cap) { qsft_set_error("qsft_dft_plan: q too large for LDS"); return QSFT_ERR_ARG; }
        int d0 = d1 - 1;
        while (d0 > 0 && ipow(q, d1 - d0 + 1) * ti <= cap) --d0;
        const long long A = ipow(q, d0);
        const long long blocks = (long long)h->rows * A * (I / ti);
        if (blocks > 0x7FFFFFFF) { qsft_set_error("qsft_dft_plan: too many blocks"); return QSFT_ERR_ARG; }
        h->pass_d0[np] = d0;
        h->pass_d1[np] = d1;
        h->pass_ti[np] = (int)ti;
        h->pass_blocks[np] = (int)blocks;
        ++np;
        d1 = d0;
    }
    h->n_pass = np;
    for (int i = np; i < QSFT_DFT_MAX_PASSES; ++i) h->pass_d0[i] = h->pass_d1[i] = h->pass_ti[i] = h->pass_blocks[i] = 0;
    return QSFT_OK;
}

template <typename T>
static hipError_t launch_dft(int radix, const qsft_dft_desc* d, const int32_t* pre, int n_inst, int blocks,
                             int pass, int lds, hipStream_t st) {
    const size_t sh = (size_t)lds;
    switch (radix) {
        case 2: hipLaunchKernelGGL((k_dft_pass<T, 2>), dim3(blocks), dim3(kDftThreads), sh, st, d, pre, n_inst, pass, lds); break;
        case 3: hipLaunchKernelGGL((k_dft_pass<T, 3>), dim3(blocks), dim3(kDftThreads), sh, st, d, pre, n_inst, pass, lds); break;
        case 4: hipLaunchKernelGGL((k_dft_pass<T, 4>), dim3(blocks), dim3(kDftThreads), sh, st, d, pre, n_inst, pass, lds); break;
        case 5: hipLaunchKernelGGL((k_dft_pass<T, 5>), dim3(blocks), dim3(kDftThreads), sh, st, d, pre, n_inst, pass, lds); break;
        case 20: hipLaunchKernelGGL((k_dft_pass<T, 20>), dim3(blocks), dim3(kDftThreads), sh, st, d, pre, n_inst, pass, lds); break;
        default: hipLaunchKernelGGL((k_dft_pass<T, 0>), dim3(blocks), dim3(kDftThreads), sh, st, d, pre, n_inst, pass, lds); break;
    }
    return hipGetLastError();
}

extern "C" int qsft_dft_lds_bytes(const qsft_dft_desc* h, int pass, int precision) {
    if (!h || pass < 0 || pass >= h->n_pass || (precision != QSFT_F32 && precision != QSFT_F64)) return -1;
    const long long elem = (precision == QSFT_F32) ? 8 : 16;
    const long long E = (long long)ipow(h->q, h->pass_d1[pass] - h->pass_d0[pass]) * h->pass_ti[pass];
    const long long bytes = (radix_compiled(h->q) ? E : 2 * E) * elem + (long long)h->q * elem;
    return bytes > kDftLdsBytes ? -1 : (int)bytes;
}

// All instances of one launch share q (radix kernels are compiled per q);
// `precision` low nibble = QSFT_F32/F64, bits 8..15 = q. lds_bytes >= the
// qsft_dft_lds_bytes() of every instance's pass (a block whose tile does not
// fit does nothing).
extern "C" int qsft_dft_pass(const qsft_dft_desc* d, const int32_t* pre, int n_inst, int blocks, int pass,
                             int precision, int lds_bytes, void* stream) {
    if (!d || !pre || n_inst < 1 || blocks < 0 || pass < 0 || pass >= QSFT_DFT_MAX_PASSES || lds_bytes < 16 ||
        lds_bytes > kDftLdsBytes) {
        qsft_set_error("qsft_dft_pass: bad arguments");
        return QSFT_ERR_ARG;
    }
    if (blocks == 0) return QSFT_OK;
    const int prec = precision & 0xF;
    const int q = (precision >> 8) & 0xFF;
    if ((prec != QSFT_F32 && prec != QSFT_F64) || q < 2 || q > QSFT_MAX_Q) {
        qsft_set_error("qsft_dft_pass: bad precision word");
        return QSFT_ERR_ARG;
    }
    const int radix = radix_compiled(q) ? q : 0;
    hipStream_t st = (hipStream_t)stream;
    const hipError_t e = (prec == QSFT_F32) ? launch_dft<float>(radix, d, pre, n_inst, blocks, pass, lds_bytes, st)
                                            : launch_dft<double>(radix, d, pre, n_inst, blocks, pass, lds_bytes, st);
    return qsft_hip_status(e, "qsft_dft_pass");
}

// q = 4 rows of 4^b samples, b in {4, 6}, contiguous input (in_stride 1):
// one workgroup per row. word: low nibble QSFT_F32/F64, bits 8..15 = b.
extern "C" int qsft_dft4_rows(const qsft_dft_desc* d, const int32_t* pre, int n_inst, int blocks, int word,
                              void* stream) {
    const int prec = word & 0xF, b = (word >> 8) & 0xFF;
    if (!d || !pre || n_inst < 1 || blocks < 0 || (prec != QSFT_F32 && prec != QSFT_F64) || (b != 4 && b != 6)) {
        qsft_set_error("qsft_dft4_rows: bad arguments");
        return QSFT_ERR_ARG;
    }
    if (blocks == 0) return QSFT_OK;
    hipStream_t st = (hipStream_t)stream;
    const dim3 g(blocks);
    if (prec == QSFT_F32) {
        if (b == 6) hipLaunchKernelGGL((k_dft4_rows<float, 6>), g, dim3(256), 0, st, d, pre, n_inst);
        else hipLaunchKernelGGL((k_dft4_rows<float, 4>), g, dim3(16), 0, st, d, pre, n_inst);
    } else {
        if (b == 6) hipLaunchKernelGGL((k_dft4_rows<double, 6>), g, dim3(256), 0, st, d, pre, n_inst);
        else hipLaunchKernelGGL((k_dft4_rows<double, 4>), g, dim3(16), 0, st, d, pre, n_inst);
    }
    return qsft_hip_status(hipGetLastError(), "qsft_dft4_rows");
}
