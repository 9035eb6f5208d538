// Stage (c): forward-scaled multidimensional q-point DFTs on gfx950.
//
// A row of Q = q^b samples is a [q]*b tensor (first axis most significant,
// as numpy's reshape([q]*b)); U = fftn(x) / q^b applies a q-point DFT
// X[k] = sum_r x[r] e^{-2 pi i r k / q} along every axis -- there are no
// inter-axis twiddles. A pass transforms the axes [d0, d1) of every row: the
// row is viewed as [A, M, I] (A = q^d0, M = q^(d1-d0), I = q^(b-d1)), and
// one workgroup stages a tile of TA x M x TI elements in LDS (TI consecutive
// inner indices, a divisor of I; when TI = I the tile is TA whole
// consecutive [M, I] blocks, one contiguous run), runs the axis transforms
// there innermost axis first with a barrier between axes, and writes the
// tile back. Every schedule therefore applies the b axis transforms in the
// same order (innermost first) with the same arithmetic, so the results do
// not depend on how the axes are grouped into passes. The last pass applies
// the 1/q^b scale (a division in FP64, like numpy's `/ q**n`; a multiply by
// the rounded reciprocal in FP32).
//
// Radix kernels: 2, 4 (butterflies), 3 and 5 (the symmetric-pair forms with
// cos/sin constants), 20 (prime-factor 4 x 5: Ruritanian input map
// n = 5 n1 + 4 n2 mod 20, four-point transforms over n1, five-point over n2,
// CRT output map k = 5 k1 + 16 k2 mod 20: no twiddles), any other q by a
// direct sum over a twiddle table (out of place through a second tile).
//
// Schedules (passes, axis groups, TI, TA, LDS budget) are enumerated by the
// host planner and ranked by a bandwidth/occupancy model; qsft_dft_autotune
// times the best-ranked candidates on the device for one (q, b, precision)
// and records the fastest, which qsft_dft_plan then returns.
//
// HBM traffic per pass is one read and one write of the rows: 16 B per
// complex FP32 sample (8 in + 8 out), 32 B in FP64.
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <vector>

#include "qsft_common.h"

namespace qsft {

constexpr int kDftThreads = 256;
constexpr int kDftLdsBytes = 65536;                 // dynamic LDS per workgroup (at most)
constexpr int kDftTwBytes = QSFT_MAX_Q * 16;        // twiddle table of the generic radix, after the tiles

template <typename T>
__device__ __forceinline__ cplx_t<T> cmul(cplx_t<T> a, cplx_t<T> b) {
    cplx_t<T> r;
    r.x = a.x * b.x - a.y * b.y;
    r.y = a.x * b.y + a.y * b.x;
    return r;
}

// x / d for x d < 2^32 by one multiply-high (m = floor(2^32 / d) + 1; the
// error term x / 2^32 < 1 / d cannot cross the next integer).
struct FastDiv {
    uint32_t d, m;
    __device__ explicit FastDiv(uint32_t d_) : d(d_), m(d_ > 1 ? (uint32_t)(0xFFFFFFFFull / d_) + 1u : 0u) {}
    __device__ __forceinline__ uint32_t div(uint32_t x) const { return d == 1 ? x : __umulhi(x, m); }
};

template <typename T>
__device__ __forceinline__ void dft4_reg(cplx_t<T>& x0, cplx_t<T>& x1, cplx_t<T>& x2, cplx_t<T>& x3) {
    using C2 = cplx_t<T>;
    const C2 s02{x0.x + x2.x, x0.y + x2.y}, d02{x0.x - x2.x, x0.y - x2.y};
    const C2 s13{x1.x + x3.x, x1.y + x3.y}, d13{x1.x - x3.x, x1.y - x3.y};
    x0 = C2{s02.x + s13.x, s02.y + s13.y};
    x2 = C2{s02.x - s13.x, s02.y - s13.y};
    x1 = C2{d02.x + d13.y, d02.y - d13.x};                        // d02 + (-i) d13
    x3 = C2{d02.x - d13.y, d02.y + d13.x};
}

// Five-point forward DFT in registers: with a1 = x1 + x4, a2 = x2 + x3,
// b1 = x1 - x4, b2 = x2 - x3, c_k = cos(2 pi k / 5), s_k = sin(2 pi k / 5):
// X1,4 = x0 + c1 a1 + c2 a2 -/+ i (s1 b1 + s2 b2), X2,3 = x0 + c2 a1 + c1 a2 -/+ i (s2 b1 - s1 b2).
template <typename T>
__device__ __forceinline__ void dft5_reg(cplx_t<T>& x0, cplx_t<T>& x1, cplx_t<T>& x2, cplx_t<T>& x3,
                                         cplx_t<T>& x4) {
    using C2 = cplx_t<T>;
    const T c1 = (T)0.30901699437494742410, c2 = (T)-0.80901699437494742410;
    const T s1 = (T)0.95105651629515357212, s2 = (T)0.58778525229247312917;
    const C2 a1{x1.x + x4.x, x1.y + x4.y}, a2{x2.x + x3.x, x2.y + x3.y};
    const C2 b1{x1.x - x4.x, x1.y - x4.y}, b2{x2.x - x3.x, x2.y - x3.y};
    const C2 t1{x0.x + c1 * a1.x + c2 * a2.x, x0.y + c1 * a1.y + c2 * a2.y};
    const C2 t2{x0.x + c2 * a1.x + c1 * a2.x, x0.y + c2 * a1.y + c1 * a2.y};
    const C2 u1{s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y};
    const C2 u2{s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y};
    x0 = C2{x0.x + a1.x + a2.x, x0.y + a1.y + a2.y};
    x1 = C2{t1.x + u1.y, t1.y - u1.x};                            // t1 - i u1
    x4 = C2{t1.x - u1.y, t1.y + u1.x};
    x2 = C2{t2.x + u2.y, t2.y - u2.x};
    x3 = C2{t2.x - u2.y, t2.y + u2.x};
}

// In-place q-point forward DFT of the pencil v[0], v[st], ..., v[(q-1) st].
template <typename T, int QR>
__device__ __forceinline__ void pencil(cplx_t<T>* v, int st, const cplx_t<T>* tw) {
    using C2 = cplx_t<T>;
    if constexpr (QR == 2) {
        const C2 a = v[0], b = v[st];
        v[0] = C2{a.x + b.x, a.y + b.y};
        v[st] = C2{a.x - b.x, a.y - b.y};
    } else if constexpr (QR == 3) {
        const T h = (T)0.86602540378443864676;                    // sin(2 pi / 3)
        const C2 x0 = v[0], x1 = v[st], x2 = v[2 * st];
        const C2 a{x1.x + x2.x, x1.y + x2.y}, d{x1.x - x2.x, x1.y - x2.y};
        const C2 t{x0.x - (T)0.5 * a.x, x0.y - (T)0.5 * a.y};
        v[0] = C2{x0.x + a.x, x0.y + a.y};
        v[st] = C2{t.x + h * d.y, t.y - h * d.x};                 // t - i h d
        v[2 * st] = C2{t.x - h * d.y, t.y + h * d.x};
    } else if constexpr (QR == 4) {
        C2 x0 = v[0], x1 = v[st], x2 = v[2 * st], x3 = v[3 * st];
        dft4_reg<T>(x0, x1, x2, x3);
        v[0] = x0;
        v[st] = x1;
        v[2 * st] = x2;
        v[3 * st] = x3;
    } else if constexpr (QR == 5) {
        C2 x0 = v[0], x1 = v[st], x2 = v[2 * st], x3 = v[3 * st], x4 = v[4 * st];
        dft5_reg<T>(x0, x1, x2, x3, x4);
        v[0] = x0;
        v[st] = x1;
        v[2 * st] = x2;
        v[3 * st] = x3;
        v[4 * st] = x4;
    } else if constexpr (QR == 20) {
        // prime-factor 4 x 5: y[n1][n2] = x[(5 n1 + 4 n2) mod 20]
        C2 y[4][5];
#pragma unroll
        for (int n1 = 0; n1 < 4; ++n1)
#pragma unroll
            for (int n2 = 0; n2 < 5; ++n2) y[n1][n2] = v[((5 * n1 + 4 * n2) % 20) * st];
#pragma unroll
        for (int n2 = 0; n2 < 5; ++n2) dft4_reg<T>(y[0][n2], y[1][n2], y[2][n2], y[3][n2]);
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) dft5_reg<T>(y[k1][0], y[k1][1], y[k1][2], y[k1][3], y[k1][4]);
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1)
#pragma unroll
            for (int k2 = 0; k2 < 5; ++k2) v[((5 * k1 + 16 * k2) % 20) * st] = y[k1][k2];
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
// LDS buffer.
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

__host__ __device__ __forceinline__ bool radix_compiled(int q) {
    return q == 2 || q == 3 || q == 4 || q == 5 || q == 20;
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
    const int d0 = d.pass_d0[pass], d1 = d.pass_d1[pass], ti = d.pass_ti[pass], ta = d.pass_ta[pass];
    const int mb = d1 - d0;
    const int Q = ipow(q, b);
    const int A = ipow(q, d0), M = ipow(q, mb), I = ipow(q, b - d1);
    const bool contiguous = ti == I;                              // the tile is ta whole [M, I] blocks
    const int ichunks = I / ti, achunks = A / ta;
    const int local = blockIdx.x - pre[inst];
    const int tiles_per_row = achunks * ichunks;
    const int row = local / tiles_per_row;
    const int rt = local - row * tiles_per_row;
    const int ac = rt / ichunks;
    const int i0 = (rt - ac * ichunks) * ti;
    const int64_t base = (int64_t)ac * ta * M * I + i0;           // element offset inside the row
    const int E = ta * M * ti;
    C2* tw = lds + (QR > 0 ? E : 2 * E);
    if ((int)(((QR > 0 ? E : 2 * E) + (QR > 0 ? 0 : q)) * sizeof(C2)) > lds_bytes) return;   // launch sized too small
    const FastDiv fti((uint32_t)ti);

    if constexpr (QR == 0) {
        for (int r = threadIdx.x; r < q; r += blockDim.x) {
            double s, c;
            sincospi(-2.0 * (double)r / (double)q, &s, &c);
            tw[r].x = (T)c;
            tw[r].y = (T)s;
        }
    }
    auto goff = [&](int e) -> int64_t {                          // tile element -> row offset
        if (contiguous) return base + e;
        const int m = (int)fti.div((uint32_t)e);
        return base + (int64_t)m * I + (e - m * ti);
    };
    // load the tile
    {
        const C2* __restrict__ in = pass == 0
            ? reinterpret_cast<const C2*>(d.in) + (int64_t)row * d.in_row_stride
            : reinterpret_cast<const C2*>(d.out) + (int64_t)row * Q;
        const int64_t stride = pass == 0 ? d.in_stride : 1;
        for (int e = threadIdx.x; e < E; e += blockDim.x) lds[e] = in[goff(e) * stride];
    }
    __syncthreads();

    C2* cur = lds;
    C2* alt = lds + E;            // only used by the generic radix
    const int pencils = E / q;
    // local axes innermost first: axis t has stride G = q^(mb-1-t) ti in the tile
    int G = ti;
    for (int t = mb - 1; t >= 0; --t) {
        const FastDiv fd((uint32_t)G);
        for (int pc = threadIdx.x; pc < pencils; pc += blockDim.x) {
            const int X = (int)fd.div((uint32_t)pc);
            const int p0 = X * q * G + (pc - X * G);
            if constexpr (QR > 0) {
                pencil<T, QR>(cur + p0, G, tw);
            } else {
                pencil_generic_out<T>(cur + p0, alt + p0, G, q, tw);
            }
        }
        __syncthreads();
        if constexpr (QR == 0) {
            C2* tmp = cur;
            cur = alt;
            alt = tmp;
        }
        G *= q;
    }

    const bool last = (pass == d.n_pass - 1);
    C2* __restrict__ out = reinterpret_cast<C2*>(d.out) + (int64_t)row * Q;
    if (last) {
        if constexpr (sizeof(T) == 8) {
            const double qd = (double)Q;
            for (int e = threadIdx.x; e < E; e += blockDim.x) {
                C2 v = cur[e];
                v.x = v.x / qd;
                v.y = v.y / qd;
                out[goff(e)] = v;
            }
        } else {
            const T sc = (T)(1.0 / (double)Q);
            for (int e = threadIdx.x; e < E; e += blockDim.x) {
                C2 v = cur[e];
                v.x *= sc;
                v.y *= sc;
                out[goff(e)] = v;
            }
        }
    } else {
        for (int e = threadIdx.x; e < E; e += blockDim.x) out[goff(e)] = cur[e];
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

namespace {

struct Sched {
    int np = 0;
    int d0[QSFT_DFT_MAX_PASSES] = {}, d1[QSFT_DFT_MAX_PASSES] = {}, ti[QSFT_DFT_MAX_PASSES] = {},
        ta[QSFT_DFT_MAX_PASSES] = {};
    double cost = 0.0;
};

std::mutex g_sched_lock;
std::map<int, Sched> g_tuned;      // measured winners (qsft_dft_autotune)
std::map<int, Sched> g_model;      // model-ranked best, memoised

int sched_key(int q, int b, int prec) { return (q << 16) | (b << 8) | prec; }

long long ipow_ll(int q, int e) {
    long long r = 1;
    for (int i = 0; i < e; ++i) r *= q;
    return r;
}

// TI and TA of one pass over axes [d0, d1) with at most `cap` tile elements:
// TI = the largest divisor of I up to min(ti_max, cap / M); when TI = I the
// tile also takes TA = q^j whole [M, I] blocks (one contiguous run).
bool pass_shape(int q, int b, int d0, int d1, long long cap, long long ti_max, int* ti_out, int* ta_out) {
    const long long M = ipow_ll(q, d1 - d0), I = ipow_ll(q, b - d1), A = ipow_ll(q, d0);
    if (M > cap || M < q) return false;
    long long lim = cap / M < ti_max ? cap / M : ti_max;
    if (lim < 1) return false;
    long long ti = 1;
    if (I <= lim) {
        ti = I;
    } else {
        for (long long c = lim; c >= 1; --c)
            if (I % c == 0) { ti = c; break; }
    }
    long long ta = 1;
    if (ti == I)
        while (A % (ta * q) == 0 && ta * q <= A && ta * q * M * I <= cap) ta *= q;
    *ti_out = (int)ti;
    *ta_out = (int)ta;
    return true;
}

// Relative cost of a pass (lower is better): HBM efficiency of its access
// runs, the residency its LDS tile allows, and how evenly its pencils fill
// the workgroup.
double pass_cost(int q, int b, int d0, int d1, int ti, int ta, int elem) {
    const long long M = ipow_ll(q, d1 - d0), I = ipow_ll(q, b - d1);
    const long long E = (long long)ta * M * ti;
    const long long tile = (radix_compiled(q) ? E : 2 * E) * elem + (radix_compiled(q) ? 0 : q * elem);
    const double run = (double)((ti == I) ? E : ti) * elem;
    const double coal = run >= 256.0 ? 1.0 : run / 256.0 + 0.05;
    long long blocks_cu = (160 * 1024) / (tile > 0 ? tile : 1);
    if (blocks_cu > 8) blocks_cu = 8;
    if (blocks_cu < 1) blocks_cu = 1;
    const double occ = blocks_cu * 4 >= 16 ? 1.0 : (double)(blocks_cu * 4) / 16.0;
    const double pencils = (double)E / q;
    const double rounds = pencils / kDftThreads;
    const double util = pencils / (kDftThreads * (double)(long long)(rounds + 0.999999));
    return 1.0 / (coal * occ) + 0.25 / util;
}

// Every split of the b axes into at most QSFT_DFT_MAX_PASSES groups
// (innermost group first), under several LDS budgets and TI limits: the
// feasible schedules, best model cost first.
std::vector<Sched> candidates(int q, int b, int prec) {
    const int elem = prec == QSFT_F32 ? 8 : 16;
    const long long budgets[] = {8192, 16384, 32768, kDftLdsBytes - kDftTwBytes};
    const long long ti_maxes[] = {16, 64, 256, 1 << 13};
    std::vector<Sched> out;
    std::vector<int> parts;
    // enumerate compositions of b (parts[0] = innermost group size)
    std::vector<std::vector<int>> comps;
    std::vector<int> cur;
    auto rec = [&](auto&& self, int left) -> void {
        if (left == 0) { comps.push_back(cur); return; }
        if ((int)cur.size() == QSFT_DFT_MAX_PASSES) return;
        for (int k = 1; k <= left; ++k) {
            cur.push_back(k);
            self(self, left - k);
            cur.pop_back();
        }
    };
    rec(rec, b);
    for (const auto& comp : comps)
        for (long long bud : budgets)
            for (long long tmax : ti_maxes) {
                long long cap = bud / elem;
                if (!radix_compiled(q)) cap = (bud - (long long)q * elem) / (2 * elem);
                Sched sc;
                bool ok = true;
                int d1 = b;
                for (int k : comp) {
                    const int d0 = d1 - k;
                    int ti = 0, ta = 0;
                    if (!pass_shape(q, b, d0, d1, cap, tmax, &ti, &ta)) { ok = false; break; }
                    sc.d0[sc.np] = d0; sc.d1[sc.np] = d1; sc.ti[sc.np] = ti; sc.ta[sc.np] = ta;
                    sc.cost += pass_cost(q, b, d0, d1, ti, ta, elem);
                    ++sc.np;
                    d1 = d0;
                }
                if (!ok) continue;
                bool dup = false;
                for (const Sched& o : out)
                    if (o.np == sc.np && !memcmp(o.d0, sc.d0, sizeof sc.d0) && !memcmp(o.ti, sc.ti, sizeof sc.ti) &&
                        !memcmp(o.ta, sc.ta, sizeof sc.ta))
                        dup = true;
                if (!dup) out.push_back(sc);
            }
    std::sort(out.begin(), out.end(), [](const Sched& a, const Sched& c) { return a.cost < c.cost; });
    return out;
}

bool schedule_for(int q, int b, int prec, Sched* sc, bool* tuned) {
    const int key = sched_key(q, b, prec);
    std::lock_guard<std::mutex> g(g_sched_lock);
    auto it = g_tuned.find(key);
    if (it != g_tuned.end()) { *sc = it->second; *tuned = true; return true; }
    *tuned = false;
    auto im = g_model.find(key);
    if (im != g_model.end()) { *sc = im->second; return true; }
    const std::vector<Sched> c = candidates(q, b, prec);
    if (c.empty()) return false;
    g_model[key] = c[0];
    *sc = c[0];
    return true;
}

int apply_schedule(qsft_dft_desc* h, const Sched& sc) {
    const int q = h->q, b = h->b;
    for (int p = 0; p < sc.np; ++p) {
        const long long A = ipow_ll(q, sc.d0[p]), I = ipow_ll(q, b - sc.d1[p]);
        const long long blocks = (long long)h->rows * (A / sc.ta[p]) * (I / sc.ti[p]);
        if (blocks > 0x7FFFFFFF) { qsft_set_error("qsft_dft_plan: too many blocks"); return QSFT_ERR_ARG; }
        h->pass_d0[p] = sc.d0[p];
        h->pass_d1[p] = sc.d1[p];
        h->pass_ti[p] = sc.ti[p];
        h->pass_ta[p] = sc.ta[p];
        h->pass_blocks[p] = (int)blocks;
    }
    h->n_pass = sc.np;
    for (int i = sc.np; i < QSFT_DFT_MAX_PASSES; ++i)
        h->pass_d0[i] = h->pass_d1[i] = h->pass_ti[i] = h->pass_ta[i] = h->pass_blocks[i] = 0;
    return QSFT_OK;
}

}  // namespace

// The schedule of a (q, b, precision): the autotuned one when
// qsft_dft_autotune has measured it in this process, else the model's best.
extern "C" int qsft_dft_plan(qsft_dft_desc* h, int precision) {
    if (!h || h->q < 2 || h->q > QSFT_MAX_Q || h->b < 1 || h->b > QSFT_MAX_B || h->rows < 0 ||
        h->in_stride < 1 || (precision != QSFT_F32 && precision != QSFT_F64)) {
        qsft_set_error("qsft_dft_plan: bad descriptor");
        return QSFT_ERR_ARG;
    }
    Sched sc;
    bool tuned = false;
    if (!schedule_for(h->q, h->b, precision, &sc, &tuned)) {
        qsft_set_error("qsft_dft_plan: no schedule fits the LDS (q too large)");
        return QSFT_ERR_ARG;
    }
    return apply_schedule(h, sc);
}

// The planner's candidate number `index` (0 = the model's best; the order
// qsft_dft_autotune tries them in); QSFT_ERR_ARG past the last one.
extern "C" int qsft_dft_plan_candidate(qsft_dft_desc* h, int precision, int index) {
    if (!h || h->q < 2 || h->q > QSFT_MAX_Q || h->b < 1 || h->b > QSFT_MAX_B || h->rows < 0 ||
        h->in_stride < 1 || (precision != QSFT_F32 && precision != QSFT_F64) || index < 0) {
        qsft_set_error("qsft_dft_plan_candidate: bad arguments");
        return QSFT_ERR_ARG;
    }
    const std::vector<Sched> c = candidates(h->q, h->b, precision);
    if (index >= (int)c.size()) {
        qsft_set_error("qsft_dft_plan_candidate: no such candidate");
        return QSFT_ERR_ARG;
    }
    return apply_schedule(h, c[index]);
}

extern "C" int qsft_dft_lds_bytes(const qsft_dft_desc* h, int pass, int precision) {
    if (!h || pass < 0 || pass >= h->n_pass || (precision != QSFT_F32 && precision != QSFT_F64)) return -1;
    const long long elem = (precision == QSFT_F32) ? 8 : 16;
    const long long E = ipow_ll(h->q, h->pass_d1[pass] - h->pass_d0[pass]) * h->pass_ti[pass] * h->pass_ta[pass];
    const long long bytes = radix_compiled(h->q) ? E * elem : 2 * E * elem + (long long)h->q * elem;
    return bytes > kDftLdsBytes ? -1 : (int)(bytes < 16 ? 16 : bytes);
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

// Times the model's best-ranked schedules (at most `max_candidates`) on
// `rows` rows of q^b samples held in `buf` (transformed in place, contents
// irrelevant; the last 4 KiB hold the descriptor and block prefixes), one
// warm-up and three timed runs each, and records the fastest for qsft_dft_plan.
extern "C" int qsft_dft_autotune(int q, int b, int precision, int rows, void* buf, size_t buf_bytes,
                                 int max_candidates, void* stream, float* best_ms) {
    if (q < 2 || q > QSFT_MAX_Q || b < 1 || b > QSFT_MAX_B || rows < 1 || !buf || max_candidates < 1 ||
        (precision != QSFT_F32 && precision != QSFT_F64)) {
        qsft_set_error("qsft_dft_autotune: bad arguments");
        return QSFT_ERR_ARG;
    }
    const long long elem = precision == QSFT_F32 ? 8 : 16;
    const long long Q = ipow_ll(q, b);
    const size_t data = (size_t)(rows * Q * elem + 255) / 256 * 256;
    if (buf_bytes < data + 4096) {
        qsft_set_error("qsft_dft_autotune: buffer smaller than rows * q^b elements + 4096 bytes");
        return QSFT_ERR_ARG;
    }
    const std::vector<Sched> cand = candidates(q, b, precision);
    if (cand.empty()) { qsft_set_error("qsft_dft_autotune: no schedule fits"); return QSFT_ERR_ARG; }
    hipStream_t st = (hipStream_t)stream;
    char* dev_desc = static_cast<char*>(buf) + data;
    int32_t* dev_pre = reinterpret_cast<int32_t*>(dev_desc + 1024);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        if (e0) (void)hipEventDestroy(e0);
        qsft_set_error("qsft_dft_autotune: hipEventCreate failed");
        return QSFT_ERR_HIP;
    }
    float best = 1e30f;
    int best_i = -1, rc = QSFT_OK;
    const int n_try = (int)cand.size() < max_candidates ? (int)cand.size() : max_candidates;
    for (int c = 0; c < n_try && rc == QSFT_OK; ++c) {
        qsft_dft_desc h{};
        h.q = q; h.b = b; h.rows = rows; h.in_stride = 1; h.in_row_stride = Q; h.in = buf; h.out = buf;
        if (apply_schedule(&h, cand[c]) != QSFT_OK) continue;
        int32_t pre[2 * QSFT_DFT_MAX_PASSES];
        int lds[QSFT_DFT_MAX_PASSES];
        bool fits = true;
        for (int p = 0; p < h.n_pass; ++p) {
            pre[2 * p] = 0;
            pre[2 * p + 1] = h.pass_blocks[p];
            lds[p] = qsft_dft_lds_bytes(&h, p, precision);
            fits = fits && lds[p] > 0;
        }
        if (!fits) continue;
        hipError_t e = hipMemcpyAsync(dev_desc, &h, sizeof(h), hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(dev_pre, pre, sizeof(pre), hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) { rc = qsft_hip_status(e, "qsft_dft_autotune(upload)"); break; }
        float ms = 1e30f;
        for (int rep = 0; rep < 4 && rc == QSFT_OK; ++rep) {
            (void)hipEventRecord(e0, st);
            for (int p = 0; p < h.n_pass && rc == QSFT_OK; ++p)
                rc = qsft_dft_pass((const qsft_dft_desc*)dev_desc, dev_pre + 2 * p, 1, h.pass_blocks[p], p,
                                   precision | (q << 8), lds[p], st);
            (void)hipEventRecord(e1, st);
            if (hipEventSynchronize(e1) != hipSuccess) { rc = qsft_hip_status(hipGetLastError(), "qsft_dft_autotune"); break; }
            float t = 0.0f;
            (void)hipEventElapsedTime(&t, e0, e1);
            if (rep > 0 && t < ms) ms = t;                          // the first run warms up
        }
        if (rc == QSFT_OK && ms < best) { best = ms; best_i = c; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != QSFT_OK) return rc;
    if (best_i < 0) { qsft_set_error("qsft_dft_autotune: no candidate ran"); return QSFT_ERR_ARG; }
    {
        std::lock_guard<std::mutex> g(g_sched_lock);
        g_tuned[sched_key(q, b, precision)] = cand[best_i];
    }
    if (best_ms) *best_ms = best;
    return QSFT_OK;
}

// The schedule qsft_dft_plan uses for (q, b, precision): passes, axis groups,
// TI and TA per pass; *tuned = 1 when it was measured by qsft_dft_autotune.
extern "C" int qsft_dft_schedule(int q, int b, int precision, int* n_pass, int* d0, int* d1, int* ti, int* ta,
                                 int* tuned) {
    if (q < 2 || q > QSFT_MAX_Q || b < 1 || b > QSFT_MAX_B || (precision != QSFT_F32 && precision != QSFT_F64) ||
        !n_pass || !d0 || !d1 || !ti || !ta || !tuned) {
        qsft_set_error("qsft_dft_schedule: bad arguments");
        return QSFT_ERR_ARG;
    }
    Sched sc;
    bool t = false;
    if (!schedule_for(q, b, precision, &sc, &t)) { qsft_set_error("qsft_dft_schedule: no schedule fits"); return QSFT_ERR_ARG; }
    *n_pass = sc.np;
    for (int p = 0; p < QSFT_DFT_MAX_PASSES; ++p) {
        d0[p] = sc.d0[p]; d1[p] = sc.d1[p]; ti[p] = sc.ti[p]; ta[p] = sc.ta[p];
    }
    *tuned = t ? 1 : 0;
    return QSFT_OK;
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
