// Working transform of one QSFT.transform call (see qsft_amd.h): gather the
// selected groups/repeats of the stored sub-transforms, promote to FP64 and
// add transform-domain noise. Logical element (c, p = r*p1 + i, j) is
//   src[gsel[c]][rsel[r]*p1 + i][j] + noise(c, p, j)
// and is stored row-major [C][R*p1][Q] or bin-major [C][Q][R*p1] (the
// peeler's layout). HBM-bound: 8 or 16 B read + 16 B write (+16 B noise read
// in the array mode) per element.
//
// Bin-major: a workgroup owns 16 bins x up to 128 rows; it reads the rows'
// 16-bin segments into an LDS tile, then writes the 16 bins' columns, which
// are one contiguous run of U, a pair of bins per wave with lanes along the
// rows. Philox noise is drawn in the write phase, one Philox block for the two
// elements (c, p, 2h) and (c, p, 2h + 1). The noise is a function of the
// logical element only, so both layouts draw the same values. With `energy`,
// each wave reduces its two columns from registers exactly as the peeler does
// (all rows of a bin are in one tile when R*p1 <= 128).
#include "qsft_common.h"

namespace qsft {

constexpr int kAsmThreads = 256;
constexpr int kAsmPerThread = 4;
constexpr int kBmBins = 16;        // bins per bin-major tile
constexpr int kBmRows = 128;       // rows per bin-major tile (row chunks beyond)

__device__ __forceinline__ double2 src_value(const qsft_assemble_desc& d, int c, int p, int64_t j) {
    const int r = p / d.p1, i = p - r * d.p1;
    const int64_t src_idx = ((int64_t)d.gsel[c] * d.src_rows + (int64_t)d.rsel[r] * d.p1 + i) * d.Q + j;
    if (d.src_prec == QSFT_F32) {
        const float2 v = reinterpret_cast<const float2*>(d.src)[src_idx];
        return make_double2(v.x, v.y);
    }
    return reinterpret_cast<const double2*>(d.src)[src_idx];
}

__device__ __forceinline__ double2 add_array_noise(const qsft_assemble_desc& d, double2 v, int64_t e) {
    if (d.noise_mode == QSFT_NOISE_ARRAY) {
        const double2 z = reinterpret_cast<const double2*>(d.noise)[e];
        v.x = v.x + z.x;
        v.y = v.y + z.y;
    }
    return v;
}

__device__ __forceinline__ double2 to_d2(double2 v) { return v; }
__device__ __forceinline__ double2 to_d2(float2 v) { return make_double2(v.x, v.y); }

// TT: element type of the LDS tile. float2 (half the LDS, twice the
// workgroups per CU) when every instance of the launch reads an FP32 source
// without a noise array (the host passes that fact); double2 otherwise.
template <typename TT>
__global__ __launch_bounds__(kAsmThreads) void k_assemble(const qsft_assemble_desc* __restrict__ descs,
                                                          const int32_t* __restrict__ pre, int n_inst,
                                                          uint64_t seed_offset, const uint64_t* seed_ptr) {
    __shared__ TT tile[kBmRows][kBmBins + 1];
    __shared__ int64_t srow[kBmRows];
    const int inst = find_inst(pre, n_inst, blockIdx.x);
    const qsft_assemble_desc& d = descs[inst];
    const int64_t Q = d.Q;
    const int rows = d.R * d.p1;
    const int local = blockIdx.x - pre[inst];
    const double comp_sd = d.sigma * 0.70710678118654752440;      // per component: sqrt(sigma^2 / 2)
    const uint64_t seed = d.seed + (seed_ptr ? *seed_ptr : seed_offset);
    if (d.seed_out && local == 0 && threadIdx.x == 0) *d.seed_out = seed;
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
            double2 v = add_array_noise(d, src_value(d, c, p, j), e);
            if (d.noise_mode == QSFT_NOISE_PHILOX) v = add_philox(v, c, p, rows, Q, j, comp_sd, seed);
            reinterpret_cast<double2*>(d.U)[e] = v;
        }
        return;
    }
    const int nj = (int)((Q + kBmBins - 1) / kBmBins), npc = (rows + kBmRows - 1) / kBmRows;
    const int c = local / (nj * npc);
    const int rem = local - c * nj * npc;
    const int jt = rem / npc, pc = rem - jt * npc;
    const int p0 = pc * kBmRows, Pc = min(kBmRows, rows - p0);
    const int64_t j0 = (int64_t)jt * kBmBins;
    const int Jc = (int)min((int64_t)kBmBins, Q - j0);
    // source row of each working row of the chunk (one division per row)
    if (threadIdx.x < Pc) {
        const int p = p0 + threadIdx.x;
        const int r = p / d.p1, i = p - r * d.p1;
        srow[threadIdx.x] = ((int64_t)d.gsel[c] * d.src_rows + (int64_t)d.rsel[r] * d.p1 + i) * Q;
    }
    __syncthreads();
    // read: 16 lanes cover one row's 16-bin segment; all loads of a thread are
    // issued before any is stored
    constexpr int kReads = kBmRows * kBmBins / kAsmThreads;       // 8
    if constexpr (sizeof(TT) == 8) {                               // FP32 source, no noise array
        float2 v[kReads];
#pragma unroll
        for (int k = 0; k < kReads; ++k) {
            const int e = threadIdx.x + k * kAsmThreads;
            const int pr = e >> 4, jj = e & 15;
            v[k] = make_float2(0.0f, 0.0f);
            if (pr < Pc && jj < Jc) v[k] = reinterpret_cast<const float2*>(d.src)[srow[pr] + j0 + jj];
        }
#pragma unroll
        for (int k = 0; k < kReads; ++k) {
            const int e = threadIdx.x + k * kAsmThreads;
            const int pr = e >> 4, jj = e & 15;
            if (pr < Pc && jj < Jc) tile[pr][jj] = v[k];
        }
    } else {
        double2 v[kReads];
#pragma unroll
        for (int k = 0; k < kReads; ++k) {
            const int e = threadIdx.x + k * kAsmThreads;
            const int pr = e >> 4, jj = e & 15;
            v[k] = make_double2(0.0, 0.0);
            if (pr < Pc && jj < Jc) {
                const int64_t si = srow[pr] + j0 + jj;
                if (d.src_prec == QSFT_F32) {
                    const float2 f = reinterpret_cast<const float2*>(d.src)[si];
                    v[k] = make_double2(f.x, f.y);
                } else {
                    v[k] = reinterpret_cast<const double2*>(d.src)[si];
                }
            }
        }
#pragma unroll
        for (int k = 0; k < kReads; ++k) {
            const int e = threadIdx.x + k * kAsmThreads;
            const int pr = e >> 4, jj = e & 15;
            if (pr < Pc && jj < Jc)
                tile[pr][jj] = add_array_noise(d, v[k], ((int64_t)c * rows + p0 + pr) * Q + j0 + jj);
        }
    }
    __syncthreads();
    // write: wave w takes bin pairs (2h, 2h + 1), h = w, w + 4, ...; lane L holds
    // rows L and L + 64 of the chunk (the peeler's column layout), so each store
    // is 1 KiB contiguous and the energies reduce straight from registers. With
    // `written` (energies, so all rows are in this chunk) a column is stored
    // only if its energy is above the peeler's zeroton cutoff.
    double2* __restrict__ U = reinterpret_cast<double2*>(d.U);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int npairs = (Jc + 1) >> 1;
    const bool philox = d.noise_mode == QSFT_NOISE_PHILOX;
    const double thr = d.cutoff * (double)rows;                    // as the peeler's screen forms it
    for (int h = wave; h < npairs; h += kAsmThreads / 64) {
        const int jj = 2 * h;
        const int64_t j = j0 + jj;
        const bool two = jj + 1 < Jc;
        double2 u0[kColRowsPerLane], u1[kColRowsPerLane];
#pragma unroll
        for (int r = 0; r < kColRowsPerLane; ++r) {
            u0[r] = make_double2(0.0, 0.0);
            u1[r] = make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int r = 0; r < kBmRows / 64; ++r) {
            const int pr = lane + 64 * r;
            if (pr < Pc) {
                const int p = p0 + pr;
                double2 v0 = to_d2(tile[pr][jj]);
                double2 v1 = two ? to_d2(tile[pr][jj + 1]) : make_double2(0.0, 0.0);
                if (philox) {
                    const float4 z = normal4(noise_counter(c, p, rows, Q, j), seed);
                    v0 = noise_add(v0, z.x, z.y, comp_sd);
                    v1 = noise_add(v1, z.z, z.w, comp_sd);
                }
                u0[r] = v0;
                u1[r] = v1;
            }
        }
        bool w0 = true, w1 = two;
        if (d.energy) {                          // rows <= kBmRows here (checked on the host)
            const double e0 = bin_energy(u0), e1 = bin_energy(u1);
            if (d.written) {
                w0 = e0 > thr;
                w1 = two && e1 > thr;
            }
            if (lane == 0) {
                d.energy[(int64_t)c * Q + j] = e0;
                if (two) d.energy[(int64_t)c * Q + j + 1] = e1;
                if (d.written) {
                    d.written[(int64_t)c * Q + j] = w0 ? 1 : 0;
                    if (two) d.written[(int64_t)c * Q + j + 1] = w1 ? 1 : 0;
                }
            }
        }
#pragma unroll
        for (int r = 0; r < kBmRows / 64; ++r) {
            const int pr = lane + 64 * r;
            if (pr < Pc) {
                const int p = p0 + pr;
                if (w0) U[((int64_t)c * Q + j) * rows + p] = u0[r];
                if (w1) U[((int64_t)c * Q + j + 1) * rows + p] = u1[r];
            }
        }
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
    if (h->energy && (!h->bin_major || rows > kBmRows)) return -1;   // energies need whole columns per tile
    if (h->written && !h->energy) return -1;                          // the write test is the energy
    long long blocks;
    if (h->bin_major) {
        blocks = (long long)h->C * ((h->Q + kBmBins - 1) / kBmBins) * ((rows + kBmRows - 1) / kBmRows);
    } else {
        const long long total = (long long)h->C * rows * h->Q;
        const long long per = kAsmThreads * kAsmPerThread;
        blocks = (total + per - 1) / per;
    }
    return blocks > 0x7FFFFFFF ? -1 : (int)blocks;
}

extern "C" int qsft_assemble_word(const qsft_assemble_desc* h) {
    if (!h) return 0;
    return (h->src_prec == QSFT_F32 && h->noise_mode != QSFT_NOISE_ARRAY) ? QSFT_ASM_F32_TILE : 0;
}

extern "C" int qsft_assemble(const qsft_assemble_desc* d, const int32_t* pre, int n_inst, int blocks,
                             uint64_t seed_offset, int word, void* stream) {
    if (!d || !pre || n_inst < 1 || blocks < 0 || (word & ~(QSFT_ASM_F32_TILE | QSFT_ASM_SEED_PTR)) ||
        ((word & QSFT_ASM_SEED_PTR) && !seed_offset)) {
        qsft_set_error("qsft_assemble: bad arguments");
        return QSFT_ERR_ARG;
    }
    if (blocks == 0) return QSFT_OK;
    // QSFT_ASM_SEED_PTR: seed_offset is the device address of the offset (read by the kernel)
    const uint64_t* sp = (word & QSFT_ASM_SEED_PTR) ? reinterpret_cast<const uint64_t*>(seed_offset) : nullptr;
    const uint64_t so = sp ? 0 : seed_offset;
    if (word & QSFT_ASM_F32_TILE)
        hipLaunchKernelGGL(k_assemble<float2>, dim3(blocks), dim3(kAsmThreads), 0, (hipStream_t)stream, d, pre,
                           n_inst, so, sp);
    else
        hipLaunchKernelGGL(k_assemble<double2>, dim3(blocks), dim3(kAsmThreads), 0, (hipStream_t)stream, d, pre,
                           n_inst, so, sp);
    return qsft_hip_status(hipGetLastError(), "qsft_assemble");
}
