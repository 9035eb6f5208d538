// Working transform of one QSFT.transform call (see qsft_amd.h): gather the
// selected groups/repeats of the stored sub-transforms, promote to FP64 and
// add transform-domain noise. Element (c, r, i, j) of the output is
//   U[c][r*p1 + i][j] = src[gsel[c]][rsel[r]*p1 + i][j] + noise
// HBM-bound: 8 or 16 B read + 16 B write (+16 B noise read) per element.
#include "qsft_common.h"

namespace qsft {

constexpr int kAsmThreads = 256;
constexpr int kAsmPerThread = 4;

__global__ __launch_bounds__(kAsmThreads) void k_assemble(const qsft_assemble_desc* __restrict__ descs,
                                                          const int32_t* __restrict__ pre, int n_inst,
                                                          uint64_t seed_offset) {
    const int inst = find_inst(pre, n_inst, blockIdx.x);
    const qsft_assemble_desc& d = descs[inst];
    const int64_t Q = d.Q;
    const int rows = d.R * d.p1;
    const int64_t total = (int64_t)d.C * rows * Q;
    const int64_t e0 = (int64_t)(blockIdx.x - pre[inst]) * (kAsmThreads * kAsmPerThread) + threadIdx.x;
    const double comp_sd = d.sigma * 0.70710678118654752440;     // per component: sqrt(sigma^2 / 2)
    for (int k = 0; k < kAsmPerThread; ++k) {
        const int64_t e = e0 + (int64_t)k * kAsmThreads;
        if (e >= total) break;
        const int64_t j = e % Q;
        const int64_t rowg = e / Q;                 // c * rows + r * p1 + i
        const int c = (int)(rowg / rows);
        const int ri = (int)(rowg - (int64_t)c * rows);
        const int r = ri / d.p1, i = ri - r * d.p1;
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
        if (d.noise_mode == QSFT_NOISE_ARRAY) {
            const double2 z = reinterpret_cast<const double2*>(d.noise)[e];
            vx = vx + z.x;
            vy = vy + z.y;
        } else if (d.noise_mode == QSFT_NOISE_PHILOX) {
            const double2 z = normal2((uint64_t)e, d.seed + seed_offset);
            vx = vx + comp_sd * z.x;
            vy = vy + comp_sd * z.y;
        }
        reinterpret_cast<double2*>(d.U)[e] = make_double2(vx, vy);
    }
}

}  // namespace qsft

using namespace qsft;

extern "C" void qsft_set_error(const char* msg);
extern "C" int qsft_hip_status(hipError_t e, const char* where);

extern "C" int qsft_assemble_blocks(const qsft_assemble_desc* h) {
    if (!h || h->Q < 1 || h->C < 1 || h->C > QSFT_MAX_C || h->R < 1 || h->R > QSFT_MAX_R || h->p1 < 1 ||
        h->src_rows < h->p1)
        return -1;
    const long long total = (long long)h->C * h->R * h->p1 * h->Q;
    const long long per = kAsmThreads * kAsmPerThread;
    const long long blocks = (total + per - 1) / per;
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
