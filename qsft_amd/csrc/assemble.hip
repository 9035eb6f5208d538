// Working transform of one QSFT.transform call (see qsft_amd.h): gather the
// selected groups/repeats of the stored sub-transforms, promote to FP64 and
// add transform-domain noise. Logical element (c, p = r*p1 + i, j) is
//   src[gsel[c]][rsel[r]*p1 + i][j] + noise(c, p, j)
// and is stored row-major [C][R*p1][Q] or bin-major [C][Q][R*p1] (the
// peeler's layout). The bin-major store goes through a 32-row x 64-bin LDS
// tile so both the reads (rows of the source) and the writes (columns of U)
// are contiguous. HBM-bound: 8 or 16 B read + 16 B write (+16 B noise read)
// per element. Philox noise is indexed by the logical element, so it does
// not depend on the layout.
#include "qsft_common.h"

namespace qsft {

constexpr int kAsmThreads = 256;
constexpr int kAsmPerThread = 4;
constexpr int kTileP = 32;
constexpr int kTileJ = 64;

__device__ __forceinline__ double2 asm_value(const qsft_assemble_desc& d, int c, int p, int64_t j,
                                             uint64_t seed_offset) {
    const int Q = d.Q, rows = d.R * d.p1;
    const int r = p / d.p1, i = p - r * d.p1;
    const int64_t src_idx = ((int64_t)d.gsel[c] * d.src_rows + (int64_t)d.rsel[r] * d.p1 + i) * Q + j;
    double vx, vy;
    if (d.src_prec == QSFT_F32) {
        const float2 v = reinterpret_cast<const float2*>(d.src)[src_idx];
        vx = v.x;
        vy = v.y;
    } else {
        const double2 v = reinterpret_cast<const double2*>(d.src)[src_idx];
        vx = v.x;
        vy = v.y;
    }
    const int64_t e = ((int64_t)c * rows + p) * Q + j;             // logical element index
    if (d.noise_mode == QSFT_NOISE_ARRAY) {
        const double2 z = reinterpret_cast<const double2*>(d.noise)[e];
        vx = vx + z.x;
        vy = vy + z.y;
    } else if (d.noise_mode == QSFT_NOISE_PHILOX) {
        const double comp_sd = d.sigma * 0.70710678118654752440;  // per component: sqrt(sigma^2 / 2)
        const double2 z = normal2((uint64_t)e, d.seed + seed_offset);
        vx = vx + comp_sd * z.x;
        vy = vy + comp_sd * z.y;
    }
    return make_double2(vx, vy);
}

__global__ __launch_bounds__(kAsmThreads) void k_assemble(const qsft_assemble_desc* __restrict__ descs,
                                                          const int32_t* __restrict__ pre, int n_inst,
                                                          uint64_t seed_offset) {
    __shared__ double2 tile[kTileP][kTileJ + 1];
    const int inst = find_inst(pre, n_inst, blockIdx.x);
    const qsft_assemble_desc& d = descs[inst];
    const int Q = d.Q;
    const int rows = d.R * d.p1;
    const int local = blockIdx.x - pre[inst];
    if (!d.bin_major) {
        const int64_t total = (int64_t)d.C * rows * Q;
        const int64_t e0 = (int64_t)local * (kAsmThreads * kAsmPerThread) + threadIdx.x;
        for (int k = 0; k < kAsmPerThread; ++k) {
            const int64_t e = e0 + (int64_t)k * kAsmThreads;
            if (e >= total) break;
            const int64_t j = e % Q;
            const int64_t rowg = e / Q;
            const int c = (int)(rowg / rows);
            const int p = (int)(rowg - (int64_t)c * rows);
            reinterpret_cast<double2*>(d.U)[e] = asm_value(d, c, p, j, seed_offset);
        }
        return;
    }
    const int nj = (Q + kTileJ - 1) / kTileJ, np = (rows + kTileP - 1) / kTileP;
    const int c = local / (nj * np);
    const int rem = local - c * nj * np;
    const int jt = rem / np, pt = rem - jt * np;
    for (int e = threadIdx.x; e < kTileP * kTileJ; e += kAsmThreads) {
        const int pr = e / kTileJ, jj = e - pr * kTileJ;
        const int p = pt * kTileP + pr, j = jt * kTileJ + jj;
        if (p < rows && j < Q) tile[pr][jj] = asm_value(d, c, p, j, seed_offset);
    }
    __syncthreads();
    double2* __restrict__ U = reinterpret_cast<double2*>(d.U);
    for (int e = threadIdx.x; e < kTileP * kTileJ; e += kAsmThreads) {
        const int jj = e / kTileP, pr = e - jj * kTileP;
        const int p = pt * kTileP + pr, j = jt * kTileJ + jj;
        if (p < rows && j < Q) U[((int64_t)c * Q + j) * rows + p] = tile[pr][jj];
    }
}

}  // namespace qsft

using namespace qsft;

extern "C" void qsft_set_error(const char* msg);
extern "C" int qsft_hip_status(hipError_t e, const char* where);

extern "C" int qsft_assemble_blocks(const qsft_assemble_desc* h) {
    if (!h || h->Q < 1 || h->C < 1 || h->C > QSFT_MAX_C || h->R < 1 || h->R > QSFT_MAX_R || h->p1 < 1 ||
        h->src_rows < h->p1 || (h->bin_major != 0 && h->bin_major != 1))
        return -1;
    const long long rows = (long long)h->R * h->p1;
    long long blocks;
    if (h->bin_major) {
        blocks = (long long)h->C * ((h->Q + kTileJ - 1) / kTileJ) * ((rows + kTileP - 1) / kTileP);
    } else {
        const long long total = (long long)h->C * rows * h->Q;
        const long long per = kAsmThreads * kAsmPerThread;
        blocks = (total + per - 1) / per;
    }
    return blocks > 0x7FFFFFFF ? -1 : (int)blocks;
}

extern "C" int qsft_assemble(const qsft_assemble_desc* d, const int32_t* pre, int n_inst, int blocks,
                             uint64_t seed_offset, void* stream) {
    if (!d || !pre || n_inst < 1 || blocks < 0) { qsft_set_error("qsft_assemble: bad arguments"); return QSFT_ERR_ARG; }
    if (blocks == 0) return QSFT_OK;
    hipLaunchKernelGGL(k_assemble, dim3(blocks), dim3(kAsmThreads), 0, (hipStream_t)stream, d, pre, n_inst,
                       seed_offset);
    return qsft_hip_status(hipGetLastError(), "qsft_assemble");
}
