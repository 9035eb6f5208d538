// Stage (d): singleton detection and peeling, device resident (gfx950).
//
// Semantics follow QSFT.transform (qsft/qsft.py:142-255) and
// singleton_detection (qsft/reconstruct.py:12-31, 100-113):
//   * every pass classifies the bins (c, j) of a frozen U (Jacobi): zeroton if
//     |col|^2 <= cutoff * P; otherwise k from nso1 or the noiseless rule,
//     signature w^{D k}, rho = sig^H col / P, residual col - rho sig; a
//     singleton iff M_c^T k mod q equals the digits of j and the residual
//     energy is <= cutoff * P, a multiton otherwise;
//   * the pass stops the loop if it found no singleton or no multiton, but
//     still peels; at most max_iter passes and q^n peeled balls;
//   * every distinct k is peeled once per pass, with the rho of the last
//     (c, j) (in c-major, j-ascending order) that detected it, from its bin
//     in every group;
//   * the output value of k is the running mean, in detection order, of
//     every rho ever detected for it.
//
// Layout: U is bin-major [C][Q][P] complex FP64, so one bin's column is P
// contiguous 16-byte rows. One wave owns one bin at a time: lanes hold rows
// p = lane + 64 r (P <= 256) for the energy, rho and residual sums (FP64
// xor-butterfly reductions, bit-identical on every lane) and coordinates
// i = lane + 1 (+ 64) for nso1 (one atan2 per lane).
//
// Scan stamps: a screen stamps the non-zeroton bins for pass 0; a later pass
// is due on the bins that were non-zeroton on the previous pass or were
// peeled into by it (an untouched zeroton keeps its column, so it stays a
// zeroton). The per-pass kernels mark such a bin with a plain store of
// pass + 1 (every writer of one bin in one kernel stores the same value),
// and the next pass's waves test the stamps of their 64 bins with one load
// and a ballot: no list, no returning atomic per bin. A multiton carries an
// extra bit until a ball is queued on it; a due bin with that bit has the
// column it was a multiton with, so it is counted as one without being read.
//
// Kernel A (detect) first subtracts the balls queued on the bin by the
// previous pass and writes the column back, then classifies it and registers
// each singleton in an open-addressing hash table keyed by a 64-bit hash
// ("tag") of k: CAS claims a slot, atomicMin / atomicMax on
// (pass+1) << 32 | bin record the first and the last detection, and a record
// (time, rho) is pushed on the slot's lock-free list; a wave's singletons are
// registered together, one lane each. Kernel B (apply) checks that the
// detection's k equals its slot's key (a 64-bit tag collision sets an error
// bit instead of merging two keys), writes the key of new slots, and lets the
// last detection of each key queue the ball on its bin in every group. No
// kernel waits on another thread's progress.
#include "qsft_common.h"

namespace qsft {

constexpr int kWaves = 4;             // waves per block of the per-bin kernels
constexpr int kBinThreads = 64 * kWaves;
constexpr int kBinsPerWave = 64;      // consecutive bins owned by one wave (one per lane) in the per-bin kernels
constexpr int kRowsPerLane = kColRowsPerLane;   // P <= 256
constexpr int kSlotThreads = 256;
constexpr uint32_t kNil = 0xFFFFFFFFu;
constexpr int kHeldMulti = 1 << 30;   // mark bit: multiton when last scanned, no ball queued on it since

// numpy's remainder(x, 2 pi) for finite x: fmod, then shift negatives up.
__device__ __forceinline__ double rem_2pi(double x) {
    const double twopi = 6.283185307179586;      // 2 * np.pi
    double m = fmod(x, twopi);
    if (m != 0.0) {
        if (m < 0.0) m += twopi;
    } else {
        m = 0.0;
    }
    return m;
}

// <row, k> over nw 4-digit words (nw a multiple of 16, rows 16-byte aligned);
// k is the wave's LDS copy (broadcast reads), both read 16 bytes at a time.
__device__ __forceinline__ uint32_t dot_row(const QSFT_GLOBAL int8_t* __restrict__ row, const uint32_t* k, int nw) {
    const QSFT_GLOBAL gi4* __restrict__ r = reinterpret_cast<const QSFT_GLOBAL gi4*>(row);
    const uint4* k4 = reinterpret_cast<const uint4*>(k);
    int acc = 0;
    for (int w = 0; w < (nw >> 2); ++w) {
        const gi4 rv = r[w];
        const uint4 kv = k4[w];
        acc = __builtin_amdgcn_sdot4(rv.x, (int)kv.x, acc, false);
        acc = __builtin_amdgcn_sdot4(rv.y, (int)kv.y, acc, false);
        acc = __builtin_amdgcn_sdot4(rv.z, (int)kv.z, acc, false);
        acc = __builtin_amdgcn_sdot4(rv.w, (int)kv.w, acc, false);
    }
    return (uint32_t)acc;
}

// The same product with the row in LDS.
__device__ __forceinline__ uint32_t dot_row_lds(const int8_t* row, const uint32_t* k, int nw) {
    const uint4* r = reinterpret_cast<const uint4*>(row);
    const uint4* k4 = reinterpret_cast<const uint4*>(k);
    int acc = 0;
    for (int w = 0; w < (nw >> 2); ++w) {
        const uint4 rv = r[w];
        const uint4 kv = k4[w];
        acc = __builtin_amdgcn_sdot4((int)rv.x, (int)kv.x, acc, false);
        acc = __builtin_amdgcn_sdot4((int)rv.y, (int)kv.y, acc, false);
        acc = __builtin_amdgcn_sdot4((int)rv.z, (int)kv.z, acc, false);
        acc = __builtin_amdgcn_sdot4((int)rv.w, (int)kv.w, acc, false);
    }
    return (uint32_t)acc;
}

// Block-level LDS copies of small per-instance tables (filled once per block):
// the offset rows of structured delays (group 0 when shared) and every
// group's M^T rows. Larger tables stay in global memory.
// MAX: the buffer's size (the first-pass kernel has smaller ones).
constexpr int kStageOffBytes = 2048;
constexpr int kStageMtBytes = 4096;

template <int MAX = kStageOffBytes>
__device__ __forceinline__ bool stage_offsets(const qsft_peel_desc& d, int8_t* buf) {
    const int R = d.P / d.p1, n_pad = d.n_pad;
    if (!(d.flags & QSFT_PEEL_UNIT_D) || !(d.flags & QSFT_PEEL_SHARED_D) || R * n_pad > MAX)
        return false;
    const int words = R * n_pad / 16;
    for (int e = threadIdx.x; e < words; e += blockDim.x) {
        const int r = e / (n_pad / 16), w = e - r * (n_pad / 16);
        reinterpret_cast<gu4*>(buf)[e] = reinterpret_cast<const QSFT_GLOBAL gu4*>(gptr(d.D) + (size_t)r * d.p1 * n_pad)[w];
    }
    return true;
}

template <int MAX = kStageMtBytes>
__device__ __forceinline__ bool stage_mt(const qsft_peel_desc& d, int8_t* buf) {
    const int bytes = d.C * d.b * d.n_pad;
    if (bytes > MAX) return false;
    for (int e = threadIdx.x; e < bytes / 16; e += blockDim.x)
        reinterpret_cast<gu4*>(buf)[e] = reinterpret_cast<const QSFT_GLOBAL gu4*>(gptr(d.Mt))[e];
    return true;
}

// 64-bit hash of the wave's k (words kk[0 .. nw)): lane w mixes word w with
// its position, the lanes' values are XOR-combined and mixed again. Every
// lane returns the same value; equal k give equal hashes.
__device__ __forceinline__ uint64_t hash_k(const uint32_t* kk, int nw, int lane) {
    const uint64_t h = (lane < nw) ? mix64((uint64_t)kk[lane] ^ ((uint64_t)(lane + 1) * 0x9E3779B97F4A7C15ull)) : 0ull;
    return mix64(wave_xor64(h) ^ 0xCBF29CE484222325ull);
}

__device__ __forceinline__ void fill_wtable_d(double2* tab, int q) {
    for (int r = threadIdx.x; r < q; r += blockDim.x) {
        double s, c;
        sincospi(2.0 * (double)r / (double)q, &s, &c);
        tab[r] = make_double2(c, s);
    }
}

// LDS writes by one lane, read by other lanes of the same wave.
// A wave's LDS operations execute in order, so this only waits for the wave's
// outstanding LDS (and scalar-memory) operations and keeps the compiler from
// moving memory accesses across it; vector-memory loads in flight (the column
// prefetch) are not waited for.
__device__ __forceinline__ void wave_lds_sync() {
    __asm__ volatile("" ::: "memory");
    __builtin_amdgcn_s_waitcnt(0xF | (0x3 << 14) | (0x7 << 4));   // lgkmcnt(0)
    __builtin_amdgcn_wave_barrier();
    __asm__ volatile("" ::: "memory");
}

__global__ __launch_bounds__(kSlotThreads) void k_peel_init(const qsft_peel_desc* __restrict__ descs,
                                                            const int32_t* __restrict__ pre, int n_inst) {
    const int inst = find_inst(pre, n_inst, blockIdx.x);
    const qsft_peel_desc& d = descs[inst];
    const int local = blockIdx.x - pre[inst];
    const int H = 1 << d.hcap_log2;
    const int h = local * kSlotThreads + threadIdx.x;
    if (local == 0 && threadIdx.x < QSFT_CTR_WORDS) d.ctr[threadIdx.x] = (threadIdx.x == 3) ? 1u : 0u;
    if (h < H) {
        d.tag[h] = 0;
        d.owner[h] = ~0ull;
        d.last[h] = 0;
        d.head[h] = kNil;
        d.count[h] = 0;
    }
    if (h < d.C * d.Q) {
        d.sstamp[h] = 0;
        d.pend_head[h] = -1;
    }
}

__device__ __forceinline__ void load_col(double2 (&u)[kRowsPerLane], const double2* __restrict__ col, int P,
                                         int lane) {
#pragma unroll
    for (int r = 0; r < kRowsPerLane; ++r) {
        const int p = lane + 64 * r;
        u[r] = (p < P) ? col[p] : make_double2(0.0, 0.0);
    }
}

// First-pass screen. Wave w owns bins [64 w, 64 w + 64); lane L stamps bin
// 64 w + L for pass 0 iff it is not a zeroton, and -1 (never) otherwise. With
// energies from the assembler it compares one number per bin; otherwise four
// columns are loaded before their energies are reduced, so several loads per
// lane are in flight and U streams once at HBM speed. Kernel A then visits
// only the stamped bins.
__global__ __launch_bounds__(kBinThreads) void k_peel_screen(const qsft_peel_desc* __restrict__ descs,
                                                             const int32_t* __restrict__ pre, int n_inst) {
    const int inst = find_inst(pre, n_inst, blockIdx.x);
    const qsft_peel_desc& d = descs[inst];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int CQ = d.C * d.Q, P = d.P;
    const int base = ((blockIdx.x - pre[inst]) * kWaves + wave) * kBinsPerWave;
    if (base >= CQ) return;
    const int nb = min(kBinsPerWave, CQ - base);
    const double thr = d.cutoff * (double)P;
    if (d.energy) {
        if (base + lane < CQ) d.mark[base + lane] = (d.energy[base + lane] > thr) ? 0 : -1;
        return;
    }
    const double2* __restrict__ Ub = reinterpret_cast<const double2*>(d.U) + (size_t)base * P;
    bool nz = false;
    for (int g = 0; g < nb; g += 4) {
        double2 u[4][kRowsPerLane];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (g + k < nb) {
                load_col(u[k], Ub + (size_t)(g + k) * P, P, lane);
            } else {
#pragma unroll
                for (int r = 0; r < kRowsPerLane; ++r) u[k][r] = make_double2(0.0, 0.0);
            }
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const double e = bin_energy(u[k]);
            if (lane == g + k) nz = e > thr;
        }
    }
    if (base + lane < CQ) d.mark[base + lane] = nz ? 0 : -1;
}

// w^{<D_c row p, k>} for this lane's rows p = lane + 64 r (zero past P); k is
// the wave's LDS copy (words kk, digits as bytes). With QSFT_PEEL_UNIT_D,
// row p = r p1 + i is offset_r + s e_(i-1), so <row, k> = <offset_r, k> + s k_(i-1)
// (i > 0): lanes r < R form the R offset products, every row takes one of them
// and one digit. Otherwise each row is a full dot product.
template <int NR>
__device__ __forceinline__ void row_phases(const qsft_peel_desc& d, const QSFT_GLOBAL int8_t* __restrict__ Dc,
                                           const uint32_t* kk, const double2* tab, const ModQ& mq, int lane,
                                           double2 (&w)[NR], const int8_t* offs, bool offs_staged) {
    const int P = d.P, n_pad = d.n_pad, nw = n_pad >> 2;
    if (d.flags & QSFT_PEEL_UNIT_D) {
        const int p1 = d.p1, R = P / p1, q = d.q;
        const bool neg = (d.flags & QSFT_PEEL_UNIT_D_NEG) != 0;
        const int8_t* kb = reinterpret_cast<const int8_t*>(kk);
        const uint32_t m1 = (uint32_t)(0xFFFFFFFFull / (uint32_t)p1) + 1u;   // p / p1 exact for p < 256
        int doff = 0;
        if (lane < R)                                              // offset rows: LDS copy when staged
            doff = (int)mq(offs_staged ? dot_row_lds(offs + (size_t)lane * n_pad, kk, nw)
                                : dot_row(Dc + (size_t)lane * p1 * n_pad, kk, nw));
        // up to 4 offsets as wave-uniform values (selected per row), more by a shuffle per row
        const bool few = R <= 4;
        int o0 = 0, o1 = 0, o2 = 0, o3 = 0;
        if (few) {
            o0 = __builtin_amdgcn_readlane(doff, 0);
            o1 = __builtin_amdgcn_readlane(doff, 1);
            o2 = __builtin_amdgcn_readlane(doff, 2);
            o3 = __builtin_amdgcn_readlane(doff, 3);
        }
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const int p = lane + 64 * r;
            const uint32_t pc = (uint32_t)min(p, P - 1);
            const int blk = (int)__umulhi(pc, m1);
            const int i = (int)pc - blk * p1;
            const int base = few ? (blk == 0 ? o0 : blk == 1 ? o1 : blk == 2 ? o2 : o3)
                                 : __shfl(doff, blk);              // every lane takes part
            w[r] = make_double2(0.0, 0.0);
            if (p < P) {
                int e = base;
                if (i > 0) e += neg ? q - (int)kb[i - 1] : (int)kb[i - 1];
                if (e >= q) e -= q;
                w[r] = tab[e];
            }
        }
    } else {
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const int p = lane + 64 * r;
            w[r] = (p < P) ? tab[mq(dot_row(Dc + (size_t)p * n_pad, kk, nw))] : make_double2(0.0, 0.0);
        }
    }
}

// A queued ball: its entry, the next entry of its queue and its rho (see
// ld_prev for how the previous pass's arrays are read). Its k row
// goes into an LDS buffer of the wave, either by an LDS-direct load issued a
// bin ahead (prefetch_krow, waited for with the column) or by plain loads
// (load_krow_now).
struct Pending {
    int ent, next;
    double rx, ry;
};

// A load of the previous pass's arrays. Written by an earlier launch: a scalar
// load. Written earlier in the same launch (PERSIST, the fused tail kernel): a
// vector load, since vector stores do not update the scalar cache.
template <bool PERSIST>
__device__ __forceinline__ int32_t ld_prev(const int32_t* p) {
    if constexpr (PERSIST) return __hip_atomic_load(const_cast<int32_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    else return *cptr(p);
}
template <bool PERSIST>
__device__ __forceinline__ double ld_prev(const double* p) {
    if constexpr (PERSIST)
        return __builtin_bit_cast(double, __hip_atomic_load(reinterpret_cast<uint64_t*>(const_cast<double*>(p)),
                                                            __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP));
    else return *cptr(p);
}

template <bool PERSIST>
__device__ __forceinline__ Pending pending_head(const qsft_peel_desc& d, int ent, int from) {
    ent = __builtin_amdgcn_readfirstlane(ent);                     // wave-uniform
    Pending pd{ent, -1, 0.0, 0.0};
    if (ent < 0) return pd;
    const size_t CQ = (size_t)d.C * d.Q;
    const int sbin = ent / d.C;                                    // the ball's detection bin
    const double* rprev = d.rho_bin + (size_t)(from & 1) * CQ * 2;
    pd.rx = ld_prev<PERSIST>(rprev + 2 * (size_t)sbin);
    pd.ry = ld_prev<PERSIST>(rprev + 2 * (size_t)sbin + 1);
    pd.next = ld_prev<PERSIST>(d.pend_next + ent);
    return pd;
}

__device__ __forceinline__ const int8_t* krow_src(const qsft_peel_desc& d, int ent, int from) {
    const size_t CQ = (size_t)d.C * d.Q;
    return d.kbin + (size_t)(from & 1) * CQ * d.n_pad + (size_t)(ent / d.C) * d.n_pad;
}

// LDS copy of a queued ball, loaded a bin ahead by LDS-direct loads (the
// caller waits for the vector-memory counter): its k row (16 bytes per lane),
// its rho and the next entry of its queue.
struct PendingSlot {
    uint32_t krow[QSFT_MAX_N / 4];
    double2 rho;
    int32_t next;
    int32_t pad[3];
};

__device__ __forceinline__ void prefetch_pending(const qsft_peel_desc& d, int ent, int from, int lane, PendingSlot* ps) {
    if (lane < (d.n_pad >> 4))
        dma16(krow_src(d, ent, from), (uint32_t)lane * 16u, __builtin_amdgcn_readfirstlane(lds_addr(ps->krow)));
    if (lane == 0) {
        const size_t CQ = (size_t)d.C * d.Q;
        dma16(d.rho_bin + (size_t)(from & 1) * CQ * 2 + 2 * (size_t)(ent / d.C), 0u,
              __builtin_amdgcn_readfirstlane(lds_addr(&ps->rho)));
        dma4(d.pend_next + ent, 0u, __builtin_amdgcn_readfirstlane(lds_addr(&ps->next)));
    }
}

__device__ __forceinline__ void load_krow_now(const qsft_peel_desc& d, int ent, int from, int lane, uint32_t* krow) {
    if (lane < (d.n_pad >> 2))
        krow[lane] = reinterpret_cast<const QSFT_GLOBAL uint32_t*>(gptr(krow_src(d, ent, from)))[lane];
    wave_lds_sync();
}

// Subtract from bin t's column u (rows p = lane + 64 r, held by the wave) every
// ball queued on it by pass `from`: rho w^{<D_c row, k>} with the ball's k
// and rho from that pass's singleton arrays (the caller writes the column
// back with store_col). The first ball of the queue (pd) has its k row in kfirst; the
// others are loaded into kk. The queue order comes from an atomic race, so the
// first kSortedPending entries are applied in ascending entry order (the FP64
// result does not depend on the race); more balls than that on one bin in one
// pass follow in queue order.
constexpr int kSortedPending = 8;

template <int NR>
__device__ __forceinline__ void subtract_ball(const qsft_peel_desc& d, const QSFT_GLOBAL int8_t* __restrict__ Dc,
                                              const Pending& pd, double2 (&u)[NR], const uint32_t* krow,
                                              const double2* tab, const ModQ& mq, int lane, const int8_t* offs,
                                              bool offs_staged) {
    const int P = d.P;
    double2 w[NR];
    row_phases(d, Dc, krow, tab, mq, lane, w, offs, offs_staged);
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        const int p = lane + 64 * r;
        if (p < P) {
            const double sx = __dsub_rn(__dmul_rn(pd.rx, w[r].x), __dmul_rn(pd.ry, w[r].y));
            const double sy = __dadd_rn(__dmul_rn(pd.rx, w[r].y), __dmul_rn(pd.ry, w[r].x));
            u[r].x = __dsub_rn(u[r].x, sx);
            u[r].y = __dsub_rn(u[r].y, sy);
        }
    }
    wave_lds_sync();                                               // the k row buffer is rewritten next
}

template <bool PERSIST, int NR>
__device__ __forceinline__ void apply_pending(const qsft_peel_desc& d, int t, int c, int from, Pending pd,
                                              double2 (&u)[NR], const uint32_t* kfirst, uint32_t* kk,
                                              const double2* tab, const ModQ& mq, int lane, const int8_t* offs,
                                              bool offs_staged) {
    if (pd.ent < 0) return;
    const int P = d.P, n_pad = d.n_pad;
    const QSFT_GLOBAL int8_t* __restrict__ Dc =
        gptr(d.D) + (size_t)((d.flags & QSFT_PEEL_SHARED_D) ? 0 : c) * P * n_pad;
    // lane a holds the a-th queued entry (the first kSortedPending of them)
    int mine = (lane == 0) ? pd.ent : -1;
    int m = 1, rest = __builtin_amdgcn_readfirstlane(pd.next);
    while (rest >= 0 && m < kSortedPending) {
        if (lane == m) mine = rest;
        ++m;
        rest = __builtin_amdgcn_readfirstlane(ld_prev<PERSIST>(d.pend_next + rest));
    }
    int rank = 0;                                                  // position in ascending order
    for (int a = 0; a < m; ++a) {
        const int o = __shfl(mine, a);
        if (lane < m && o < mine) ++rank;                          // entries are distinct
    }
    for (int a = 0;; ++a) {
        int e;
        if (a < m) {
            e = __shfl(mine, __builtin_ctzll(__ballot(lane < m && rank == a)));
        } else if (rest >= 0) {                                    // beyond kSortedPending: queue order
            e = rest;
        } else {
            break;
        }
        e = __builtin_amdgcn_readfirstlane(e);
        if (e == pd.ent) {
            subtract_ball(d, Dc, pd, u, kfirst, tab, mq, lane, offs, offs_staged);
        } else {
            const Pending pa = pending_head<PERSIST>(d, e, from);
            if (a >= m) rest = pa.next;
            load_krow_now(d, e, from, lane, kk);
            subtract_ball(d, Dc, pa, u, kk, tab, mq, lane, offs, offs_staged);
        }
    }
}

// Sparse working transform in use: the flag plus the buffers it needs (a
// descriptor with the flag but without them is treated as a dense U).
__device__ __forceinline__ bool sparse_u(const qsft_peel_desc& d) {
    return (d.flags & QSFT_PEEL_SPARSE_U) && d.xsrc && d.written && d.noise_key;
}

// Write bin t's column back after its queue was applied, and clear the queue
// (with a sparse U the column now exists there).
template <int NR>
__device__ __forceinline__ void store_col(const qsft_peel_desc& d, int t, const double2 (&u)[NR], int lane) {
    const int P = d.P;
    QSFT_GLOBAL gd2* __restrict__ col = reinterpret_cast<QSFT_GLOBAL gd2*>(gptr(d.U)) + (size_t)t * P;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        const int p = lane + 64 * r;
        if (p < P) col[p] = gd2{u[r].x, u[r].y};
    }
    if (lane == 0) {
        gptr(d.pend_head)[t] = -1;                                 // the queue is consumed
        if (sparse_u(d)) gptr(d.written)[t] = 1;
    }
}

// Sparse U: the column of bin t = (c, j) that the assembler did not write,
// formed as the assembler forms it: the FP32 sample promoted to FP64 plus the
// Philox noise of the element (same counter, key and rounding).
template <int NR>
__device__ __forceinline__ void rebuild_col(const qsft_peel_desc& d, int c, int j, double2 (&u)[NR], int lane) {
    const int P = d.P;
    const int64_t Q = d.Q;
    const float2* __restrict__ xs = reinterpret_cast<const float2*>(d.xsrc) + (size_t)c * P * Q + j;
    const double comp_sd = d.noise_sigma * 0.70710678118654752440;
    const uint64_t key = *d.noise_key;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        const int p = lane + 64 * r;
        u[r] = make_double2(0.0, 0.0);
        if (p < P) {
            const float2 v = xs[(size_t)p * Q];
            u[r] = make_double2((double)v.x, (double)v.y);
            if (d.noise_sigma > 0.0) u[r] = add_philox(u[r], c, p, P, Q, j, comp_sd, key);
        }
    }
}

// After the last pass: apply what its kernel B queued, so U ends as the
// reference's loop leaves it. Same 64-bin wave partition over the slot grid
// (which covers at least C*Q threads); waves without a queued bin exit at once.
__global__ __launch_bounds__(kSlotThreads) void k_peel_flush(const qsft_peel_desc* __restrict__ descs,
                                                             const int32_t* __restrict__ pre, int n_inst) {
    __shared__ double2 tab[QSFT_MAX_Q];
    __shared__ __align__(16) uint32_t kws[kSlotThreads / 64][QSFT_MAX_N / 4];
    __shared__ __align__(16) uint32_t kfirst[kSlotThreads / 64][QSFT_MAX_N / 4];
    const int inst = find_inst(pre, n_inst, blockIdx.x);
    const qsft_peel_desc& d = descs[inst];
    __shared__ __align__(16) int8_t offbuf[kStageOffBytes];
    const uint32_t iters = d.ctr[QSFT_CTR_ITERS];
    if (iters == 0) return;
    fill_wtable_d(tab, d.q);
    const bool staged = stage_offsets(d, offbuf);
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int CQ = d.C * d.Q, P = d.P;
    const int base = ((blockIdx.x - pre[inst]) * (kSlotThreads / 64) + wave) * kBinsPerWave;
    if (base >= CQ) return;
    uint64_t todo = __ballot(base + lane < CQ && d.pend_head[base + lane] >= 0);
    const ModQ mq((uint32_t)d.q);
    while (todo) {
        const int t = base + __builtin_ctzll(todo);
        todo &= todo - 1;
        double2 u[kRowsPerLane];
        if (sparse_u(d) && !d.written[t]) rebuild_col(d, t / d.Q, t % d.Q, u, lane);
        else load_col(u, reinterpret_cast<const double2*>(d.U) + (size_t)t * P, P, lane);
        const Pending pd = pending_head<false>(d, d.pend_head[t], (int)iters - 1);
        load_krow_now(d, pd.ent, (int)iters - 1, lane, kfirst[wave]);
        apply_pending<false>(d, t, t / d.Q, (int)iters - 1, pd, u, kfirst[wave], kws[wave], tab, mq, lane, offbuf, staged);
        store_col(d, t, u, lane);
    }
}

// Digit of a phase: the k of the nearest multiple of 2 pi / q, as the exact
// FP64 rules below compute it, taken from a cheaper test wherever the phase is
// not within a small margin of a decision boundary (half-way between two
// multiples). For q = 4 and q = 2 the sector follows from comparisons of the
// components; otherwise an FP32 atan2 of the components, scaled by a power of
// two into FP32 range, is off by < 1e-6 rad, far inside the margin (1e-4 of a
// sector, >= 1e-5 rad for q <= 64).
// Returns -1 where the exact rule must decide (near a boundary, zero, inf, NaN).
__device__ __forceinline__ int fast_sector(double zr, double zi, int q) {
    const double ax = fabs(zr), ay = fabs(zi);
    if (!(ax < 1.0e300 && ay < 1.0e300)) return -1;               // inf or NaN
    const double m = fmax(ax, ay);
    if (q == 4) {
        if (!(fabs(ax - ay) > 1e-9 * m)) return -1;
        return ax > ay ? (zr > 0.0 ? 0 : 2) : (zi > 0.0 ? 1 : 3);
    }
    if (q == 2) {
        if (!(ax > 1e-9 * m)) return -1;
        return zr > 0.0 ? 0 : 1;
    }
    if (!(m > 1.0e-300)) return -1;
    const int ex = ilogb(m);
    const float fx = (float)ldexp(zr, -ex), fy = (float)ldexp(zi, -ex);
    float v = atan2f(fy, fx) * ((float)q * 0.15915494309189535f);   // q / (2 pi)
    if (v < 0.0f) v += (float)q;
    const float r = rintf(v);
    if (!(fabsf(v - r) < 0.5f - 1.0e-4f)) return -1;
    int k = (int)r;
    if (k >= q) k -= q;
    return k;
}

// nso1 digit of coordinate i (reference: the nso1 branch of
// singleton_detection, qsft/reconstruct.py:12-31): the phase of the mean over
// the R repeats of u_{r,0} conj(u_{r,i}), rounded to the nearest 2 pi t / q.
// The exact FP64 rule, for phases the FP32 sector test leaves open (within
// about 1e-4 of a sector boundary): kept out of line, so that its registers
// are not reserved across the detection loop (its spills run only when it
// is called).
__device__ __noinline__ uint32_t nso1_slow(double zr, double zi, int q, int R) {
    zr = __ddiv_rn(zr, (double)R);
    zi = __ddiv_rn(zi, (double)R);
    const double twopi = 6.283185307179586;
    const double th = rem_2pi(atan2(zi, zr));
    const double step = twopi / (double)q;                         // 2 * np.pi / q
    int best = 0;
    double bestv = fabs(__dsub_rn(0.0, th));
    for (int tt = 1; tt <= q; ++tt) {
        const double v = fabs(__dsub_rn(__dmul_rn(step, (double)tt), th));
        if (v < bestv) { bestv = v; best = tt; }
    }
    return (uint32_t)(best % q);
}

__device__ __forceinline__ uint32_t nso1_digit(double zr, double zi, int q, int R) {
    const int f = fast_sector(zr, zi, q);
    return f >= 0 ? (uint32_t)f : nso1_slow(zr, zi, q, R);
}

// Noiseless digit of coordinate i (reference: qsft/reconstruct.py:100-113):
// round(q (angle(u_i) - angle(u_0)) / (2 pi)) mod q, from the two FP64 angles
// (out of line, as nso1_slow) unless the FP32 sector test of the phase of
// u_i conj(u_0), exact to ~1e-16 rad when both magnitudes are far from the
// FP64 range ends, decides it.
__device__ __noinline__ uint32_t noiseless_slow(double2 v0, double2 vi, int q) {
    const double twopi = 6.283185307179586;
    const double a0 = atan2(v0.y, v0.x), ai = atan2(vi.y, vi.x);
    const double v = __ddiv_rn(__dmul_rn((double)q, __dsub_rn(ai, a0)), twopi);
    long long kq = (long long)rint(v) % q;
    if (kq < 0) kq += q;
    return (uint32_t)kq;
}

__device__ __forceinline__ uint32_t noiseless_digit(double2 v0, double2 vi, int q) {
    const double m0 = fmax(fabs(v0.x), fabs(v0.y)), mi = fmax(fabs(vi.x), fabs(vi.y));
    if (m0 > 1.0e-150 && m0 < 1.0e150 && mi > 1.0e-150 && mi < 1.0e150) {
        const double wr = vi.x * v0.x + vi.y * v0.y, wi = vi.y * v0.x - vi.x * v0.y;
        const int f = fast_sector(wr, wi, q);
        if (f >= 0) return (uint32_t)f;
    }
    return noiseless_slow(v0, vi, q);
}

// The detection's FP64 accumulations, with every rounding written out (no
// contraction left to the compiler), so that the one-bin-per-wave and the
// four-bins-per-wave detections give the same bits:
// rr += Re(conj(w) u), ri += Im(conj(w) u); s += |u - rho w|^2.
__device__ __forceinline__ void conj_dot_acc(const double2 w, const double2 u, double& rr, double& ri) {
    rr = __dadd_rn(rr, __fma_rn(w.x, u.x, __dmul_rn(w.y, u.y)));
    ri = __dadd_rn(ri, __fma_rn(w.x, u.y, -__dmul_rn(w.y, u.x)));
}
__device__ __forceinline__ void residual_acc(const double2 u, const double2 w, double rho_x, double rho_y,
                                             double& s) {
    const double dx = __dsub_rn(u.x, __fma_rn(rho_x, w.x, -__dmul_rn(rho_y, w.y)));
    const double dy = __dsub_rn(u.y, __fma_rn(rho_x, w.y, __dmul_rn(rho_y, w.x)));
    s = __dadd_rn(s, __fma_rn(dx, dx, __dmul_rn(dy, dy)));
}

// Digit of coordinate i (1..n) of the k of a bin from its column cw (rows in
// LDS): nso1 (the mean over the R repeats of u_{r,0} conj(u_{r,i})) or the
// noiseless rule.
__device__ __forceinline__ uint32_t coord_digit(const qsft_peel_desc& d, const double2* cw, int i, int q, int P) {
    if (d.method & QSFT_DETECT_NSO1) {
        const int p1 = d.p1, R = P / p1;
        double zr = 0.0, zi = 0.0;
        for (int r = 0; r < R; ++r) {
            const double2 u0 = cw[r * p1];
            const double2 ui = cw[r * p1 + i];
            // u0 * conj(ui), as numpy forms the complex product
            const double pr = __dsub_rn(__dmul_rn(u0.x, ui.x), __dmul_rn(u0.y, -ui.y));
            const double pi_ = __dadd_rn(__dmul_rn(u0.x, -ui.y), __dmul_rn(u0.y, ui.x));
            zr = __dadd_rn(zr, pr);
            zi = __dadd_rn(zi, pi_);
        }
        return nso1_digit(zr, zi, q, R);
    }
    return noiseless_digit(cw[0], cw[i], q);
}

// Residual energy |col - rho sig|^2 = |col|^2 - P |rho|^2 (|sig_p| = 1 and
// sig^H col = P rho), roundings written out.
__device__ __forceinline__ double residual_energy(double energy, int P, double rho_x, double rho_y) {
    return __fma_rn(-(double)P, __fma_rn(rho_x, rho_x, __dmul_rn(rho_y, rho_y)), energy);
}

// Per-wave LDS of the detection: the k digits, two column slots with the
// first ball queued on each, and the singletons buffered for registration.
// ---- coded source: Reed-Solomon syndrome decoding in one wave
constexpr int kRsMaxSyn = 128;                                     // syndrome digits 2 t s
constexpr int kRsMaxT2 = 32;                                       // 2 t
constexpr int kRsScratch = 168;                                    // S[32] lam[33] B[33] omega[32] T[33] L status

// GF(q^s): an element is the integer whose base-q digits are its polynomial
// coefficients; exp/log tables of the primitive element alpha = x.
struct GFDev {
    const int16_t* ex;
    const int16_t* lg;
    int q, s, nc;
    __device__ int add(int a, int b) const {
        int r = 0, p = 1;
        for (int i = 0; i < s; ++i) {
            int x = a % q + b % q;
            if (x >= q) x -= q;
            r += x * p;
            a /= q;
            b /= q;
            p *= q;
        }
        return r;
    }
    __device__ int neg(int a) const {
        int r = 0, p = 1;
        for (int i = 0; i < s; ++i) {
            const int x = a % q;
            r += (x ? q - x : 0) * p;
            a /= q;
            p *= q;
        }
        return r;
    }
    __device__ int mul(int a, int b) const { return (a == 0 || b == 0) ? 0 : (int)ex[lg[a] + lg[b]]; }
    __device__ int inv(int a) const { return (int)ex[nc - lg[a]]; }
    __device__ int alpha(int e) const {
        e %= nc;
        if (e < 0) e += nc;
        return (int)ex[e];
    }
};

// k from the 2t syndrome symbols (s digits each, highest degree first) in
// `syn`: Berlekamp-Massey on lane 0 (LDS state in `rs`), then a lane-parallel
// Chien search over the q^s - 1 code positions with Forney's error values
// e = -Omega(X^-1) / Lambda'(X^-1). Positions past the shortened part hold
// k. A failed decode (locator degree > t, roots != degree, a zero
// derivative, an error value outside GF(q)) leaves k = 0, as the received
// all-zero word. kb8 (n_pad bytes, zero on entry) receives k.
__device__ __noinline__ void rs_decode_wave(const qsft_peel_desc& d, const int8_t* syn, int8_t* kb8, int32_t* rs,
                                            int lane) {
    const int t2 = 2 * d.rs_t, s = d.rs_s, nc = d.rs_nc, q = d.q, n = d.n;
    if (!d.rs_tab || t2 < 2 || t2 > kRsMaxT2) return;              // no code given: k = 0
    const GFDev F{d.rs_tab, d.rs_tab + 2 * nc, q, s, nc};
    int32_t* S = rs;
    int32_t* lam = rs + 32;
    int32_t* bp = rs + 65;
    int32_t* om = rs + 98;
    int32_t* tmp = rs + 130;
    if (lane == 0) {
        bool any = false;
        for (int j = 0; j < t2; ++j) {
            int v = 0;
            for (int r = 0; r < s; ++r) v = v * q + (int)syn[s * j + r];
            S[j] = v;
            any = any || v != 0;
        }
        int status = 0, L = 0;
        if (any) {
            for (int i = 0; i <= t2; ++i) lam[i] = bp[i] = (i == 0);
            int m = 1, bd = 1;
            for (int r = 0; r < t2; ++r) {
                int disc = S[r];
                for (int i = 1; i <= L; ++i) disc = F.add(disc, F.mul(lam[i], S[r - i]));
                if (disc == 0) {
                    ++m;
                    continue;
                }
                const int coef = F.mul(disc, F.inv(bd));
                if (2 * L <= r) {
                    for (int i = 0; i <= t2; ++i) tmp[i] = lam[i];
                    for (int i = 0; i + m <= t2; ++i) lam[i + m] = F.add(lam[i + m], F.neg(F.mul(coef, bp[i])));
                    for (int i = 0; i <= t2; ++i) bp[i] = tmp[i];
                    L = r + 1 - L;
                    bd = disc;
                    m = 1;
                } else {
                    for (int i = 0; i + m <= t2; ++i) lam[i + m] = F.add(lam[i + m], F.neg(F.mul(coef, bp[i])));
                    ++m;
                }
            }
            bool ok = L <= t2 / 2 && lam[L] != 0;
            for (int i = L + 1; i <= t2; ++i) ok = ok && lam[i] == 0;
            if (ok) {
                for (int a = 0; a < t2; ++a) {                    // Omega = S(x) Lambda(x) mod x^2t
                    int v = 0;
                    for (int b2 = 0; b2 <= L && b2 <= a; ++b2) v = F.add(v, F.mul(S[a - b2], lam[b2]));
                    om[a] = v;
                }
                status = 2;
            }
        }
        rs[164] = L;
        rs[165] = status;
    }
    wave_lds_sync();
    if (rs[165] != 2) return;                                      // no syndrome, or failure: k = 0
    const int L = rs[164];
    int roots = 0;
    bool fail = false;
    for (int p0 = 0; p0 < nc; p0 += 64) {
        const int pos = p0 + lane;
        bool root = false, bad = false;
        int val = 0;
        if (pos < nc) {
            const int xinv = F.alpha(-(nc - 1 - pos));
            int v = 0;
            for (int i = L; i >= 0; --i) v = F.add(F.mul(v, xinv), lam[i]);
            if (v == 0) {
                root = true;
                int dv = 0;
                for (int i = L; i >= 1; --i) dv = F.add(F.mul(dv, xinv), F.mul(i % q, lam[i]));
                int ov = 0;
                for (int i = t2 - 1; i >= 0; --i) ov = F.add(F.mul(ov, xinv), om[i]);
                if (dv == 0) {
                    bad = true;
                } else {
                    val = F.neg(F.mul(ov, F.inv(dv)));
                    bad = val >= q;
                }
            }
        }
        roots += __popcll(__ballot(root));
        fail = fail || __ballot(bad) != 0ull;
        if (root && !bad && pos >= nc - n) kb8[pos - (nc - n)] = (int8_t)val;
    }
    wave_lds_sync();
    if (roots != L || fail) {
        uint32_t* kk = reinterpret_cast<uint32_t*>(kb8);
        if (lane < (d.n_pad >> 2)) kk[lane] = 0;
        wave_lds_sync();
    }
}

// Singletons a wave buffers in LDS before registering them (CAP of them; a
// full buffer is registered mid-loop), so that the buffers stay small enough
// for four detection blocks per CU. The fused tail, one workgroup per CU
// anyway, buffers all 64.
constexpr int kSbCap = 16;

template <int NR, int CAP = kSbCap>
struct DetectWaveLds {
    __align__(16) uint32_t kws[QSFT_MAX_N / 4];
    int8_t syn[kRsMaxSyn];                                         // coded source: syndrome digits
    int32_t rs[kRsScratch];                                        // coded source: decoder state
    __align__(16) PendingSlot pslot[2];                            // first queued ball, per column slot
    __align__(16) double2 colw[2][64 * NR];                        // current / next column
    uint64_t sb_h[CAP];
    double2 sb_rho[CAP];
    int32_t sb_t[CAP];
};

// The wave's bins due this pass among the 64 from `base` (bit i: bin base + i).
// A held multiton (a multiton last pass that no ball has been queued on since)
// keeps its column, so it is a multiton again: it is counted (nheld) and
// re-stamped without being read.
__device__ __forceinline__ uint64_t scan_marks(const qsft_peel_desc& d, int it, int base, uint32_t& nheld) {
    const int lane = threadIdx.x & 63;
    QSFT_GLOBAL int32_t* __restrict__ mark = gptr(d.mark);
    const int mk = (base + lane < d.C * d.Q) ? mark[base + lane] : -1;
    const bool due = (mk & ~kHeldMulti) == it;
    const bool held = due && (mk & kHeldMulti) != 0;
    nheld = (uint32_t)__popcll(__ballot(held));
    if (held) mark[base + lane] = (it + 1) | kHeldMulti;
    return __ballot(due && !held);
}

// One wave classifies the due bins `todo` among the 64 from `base` and
// registers its singletons (kernel A's body; the fused tail kernel runs it
// with PERSIST). tab, offbuf and mtbuf are the block's staged tables.
// End of a wave's detection: the pass counters (bins scanned, singletons,
// multitons: one atomic each per wave), then the hash registration of the
// wave's buffered singletons, one lane each, so the atomics of up to 64
// registrations are in flight together.
// Hash registration of the wave's ns buffered singletons (sb_*), one lane
// each. Every lane of the wave takes part (ballots).
__device__ __forceinline__ void register_singletons(const qsft_peel_desc& d, int it, int ns, const int32_t* sb_t,
                                                    const uint64_t* sb_h, const double2* sb_rho, int lane) {
    uint32_t* __restrict__ ctr = d.ctr;
    const int CQ = d.C * d.Q;
    int slot = -1;
    bool claimed = false;                                      // this lane's k took a new slot
    bool newball = false;
    int t = 0;
    uint64_t tm = 0;
    double2 rho = make_double2(0.0, 0.0);
    if (lane < ns) {
        t = sb_t[lane];
        rho = sb_rho[lane];
        const uint64_t hv = sb_h[lane];
        const uint64_t tg = hv | 1ull;
        const uint32_t mask = (1u << d.hcap_log2) - 1u;
        uint32_t pos = (uint32_t)(hv >> 32) & mask;
        for (uint32_t probe = 0; probe <= mask; ++probe) {
            const unsigned long long old = atomicCAS((unsigned long long*)&d.tag[pos], 0ull, (unsigned long long)tg);
            if (old == 0ull || old == tg) {
                slot = (int)pos;
                claimed = old == 0ull;
                break;
            }
            pos = (pos + 1) & mask;
        }
        if (slot < 0) {
            atomicOr(&ctr[QSFT_CTR_ERROR], (uint32_t)QSFT_DEVERR_TABLE_FULL);
        } else {
            d.slot_bin[t] = slot;
            d.sstamp[t] = it + 1;                                 // kernel B peels only registered bins
            tm = ((uint64_t)(it + 1) << 32) | (uint32_t)t;
            atomicMin((unsigned long long*)&d.owner[slot], (unsigned long long)tm);
            const unsigned long long prev = atomicMax((unsigned long long*)&d.last[slot], (unsigned long long)tm);
            newball = (prev >> 32) != (unsigned long long)(it + 1);
        }
    }
    const uint64_t nb_mask = __ballot(newball);
    const uint64_t ok_mask = __ballot(slot >= 0);
    const uint64_t cl_mask = __ballot(claimed);
    uint32_t rbase = 0, lbase = 0;
    if (lane == 0) {
        if (nb_mask) atomicAdd(&ctr[it * QSFT_CTR_ITER_STRIDE + 2], (uint32_t)__popcll(nb_mask));
        if (ok_mask) rbase = atomicAdd(&ctr[QSFT_CTR_LOG], (uint32_t)__popcll(ok_mask));
        if (cl_mask) lbase = atomicAdd(&ctr[QSFT_CTR_LIVE], (uint32_t)__popcll(cl_mask));
    }
    rbase = __shfl(rbase, 0);
    lbase = __shfl(lbase, 0);
    if (claimed) d.live[lbase + (uint32_t)__popcll(cl_mask & ((1ull << lane) - 1ull))] = (uint32_t)slot;
    if (slot >= 0) {
        const uint32_t rix = rbase + (uint32_t)__popcll(ok_mask & ((1ull << lane) - 1ull));
        if ((uint64_t)rix >= (uint64_t)d.max_iter * (uint64_t)CQ) {   // cannot happen: one record per bin per pass
            atomicOr(&ctr[QSFT_CTR_ERROR], (uint32_t)QSFT_DEVERR_BAD_SHAPE);
        } else {
            d.rec_time[rix] = tm;
            d.rec_rho[2 * (size_t)rix] = rho.x;
            d.rec_rho[2 * (size_t)rix + 1] = rho.y;
            d.rec_next[rix] = atomicExch(&d.head[slot], rix);
            atomicAdd(&d.count[slot], 1u);
        }
    }
}

// End of a wave's detection: the pass counters over the whole wave (scanned
// bins, ns singletons, nz - ns multitons) and the registration of the nbuf
// singletons still buffered.
__device__ __forceinline__ void finish_wave(const qsft_peel_desc& d, int it, uint32_t scanned, int nz, int ns,
                                            const int32_t* sb_t, const uint64_t* sb_h, const double2* sb_rho,
                                            int lane, int nbuf) {
    if (lane == 0) {
        uint32_t* pc = d.ctr + it * QSFT_CTR_ITER_STRIDE;
        atomicAdd(pc + 4, scanned);
        if (ns) atomicAdd(pc + 0, (uint32_t)ns);
        if (nz > ns) atomicAdd(pc + 1, (uint32_t)(nz - ns));
    }
    if (nbuf > 0) register_singletons(d, it, nbuf, sb_t, sb_h, sb_rho, lane);
}

template <int NR, bool PERSIST, int CAP>
__device__ __forceinline__ void detect_bins(const qsft_peel_desc& d, int it, int base, uint64_t todo, uint32_t nheld,
                                            DetectWaveLds<NR, CAP>& W, const double2* tab, const int8_t* offbuf,
                                            bool offs_staged, const int8_t* mtbuf, bool mt_staged) {
    uint32_t* __restrict__ ctr = d.ctr;
    const int lane = threadIdx.x & 63;
    if (todo == 0) {
        if (lane == 0 && nheld) atomicAdd(ctr + it * QSFT_CTR_ITER_STRIDE + 1, nheld);
        return;
    }
    const int q = d.q, n = d.n, b = d.b, C = d.C, P = d.P, Q = d.Q, n_pad = d.n_pad;
    const int nw = n_pad >> 2;
    const int CQ = C * Q;
    QSFT_GLOBAL int32_t* __restrict__ mark = gptr(d.mark);
    uint32_t* kk = W.kws;
    int8_t* kb8 = reinterpret_cast<int8_t*>(kk);
    const ModQ mq((uint32_t)q);
    const double thr = d.cutoff * (double)P;
    const double inv_p = 1.0 / (double)P;
    const double2* __restrict__ Ub = reinterpret_cast<const double2*>(d.U);
    QSFT_GLOBAL int8_t* __restrict__ kcur = gptr(d.kbin) + (size_t)(it & 1) * CQ * n_pad;   // this pass's k and rho
    QSFT_GLOBAL double* __restrict__ rcur = gptr(d.rho_bin) + (size_t)(it & 1) * CQ * 2;
    int wdig = 0;                                                  // q^(b - 1 - lane): weight of bin digit `lane`
    if (lane < b) {
        wdig = 1;
        for (int e = lane + 1; e < b; ++e) wdig *= q;
    }
    const uint32_t scanned = (uint32_t)__popcll(todo);
    int nz = (int)nheld, ns = 0;                                   // wave-uniform
    int nb = 0;                                                    // singletons buffered in W.sb_*
    // sparse U: due bins whose column the assembler skipped (only this wave
    // writes its bins' flags, in store_col, after their step)
    const uint64_t unwritten =
        sparse_u(d) ? __ballot(((todo >> lane) & 1) && !gptr(d.written)[base + lane]) : 0ull;
    int c = base / Q;                                              // group of the current bin

    // queue heads of the wave's bins (balls the previous pass peeled into them)
    const int qhead = (it > 0 && ((todo >> lane) & 1)) ? gptr(d.pend_head)[base + lane] : -1;
    // Bins stream through two LDS slots per wave: each holds a column and the
    // first queued ball (k row, rho, next entry), loaded by LDS-direct loads
    // (no registers held) a bin ahead. A step issues the next bin's loads,
    // classifies the current bin (no global loads or stores on the common
    // path), then waits once for every vector-memory operation in flight (the
    // next bin's data, which has had the whole step to land, and the previous
    // step's stores) and only then issues its own stores.
    const int nrow = (P + 63) >> 6;                                // loads per column
    const uint32_t rowoff = (uint32_t)min(lane, P - 1) * 16u;     // rows past P re-read row P - 1
    auto issue = [&](int tb, int ent, int slot) {
        if (ent >= 0) prefetch_pending(d, ent, it - 1, lane, &W.pslot[slot]);
        const double2* cb = Ub + (size_t)tb * P;
        for (int r = 0; r < nrow; ++r)
            dma16(cb, min(64 * r * 16u + rowoff, (uint32_t)(P - 1) * 16u),
                  __builtin_amdgcn_readfirstlane(lds_addr(&W.colw[slot][64 * r])));
    };
    int cur = __builtin_ctzll(todo);
    int cslot = 0;
    int curent = __builtin_amdgcn_readfirstlane(__shfl(qhead, cur));
    issue(base + cur, curent, 0);
    wait_vm<0>();
    __asm__ volatile("" ::: "memory");
    for (;;) {
        const int t = base + cur;
        double2* cw = W.colw[cslot];
        Pending pd{curent, -1, 0.0, 0.0};
        if (curent >= 0) {
            const double2 rv = W.pslot[cslot].rho;
            pd.rx = rv.x;
            pd.ry = rv.y;
            pd.next = W.pslot[cslot].next;
        }
        todo &= todo - 1;
        const bool more = todo != 0;
        const int nxt = more ? __builtin_ctzll(todo) : 0;
        const int nent = more ? __builtin_amdgcn_readfirstlane(__shfl(qhead, nxt)) : -1;
        if (more) issue(base + nxt, nent, cslot ^ 1);
        double2 u[NR];
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const int p = lane + 64 * r;
            u[r] = (p < P) ? cw[p] : make_double2(0.0, 0.0);
        }

        while (t >= (c + 1) * Q) ++c;
        const int j = t - c * Q;
        // A sparse U needs the assembler's energies, so this instance has
        // P <= 128, but NR follows the batch's largest P: NR = 4 kernels
        // rebuild skipped columns too (rows past P stay zero).
        const bool rebuilt = (unwritten >> cur) & 1;               // wave-uniform
        if (rebuilt) rebuild_col(d, c, j, u, lane);                // (its prefetched copy is not U)
        // balls queued on this bin by the previous pass
        apply_pending<PERSIST>(d, t, c, it - 1, pd, u, W.pslot[cslot].krow, kk, tab, mq, lane, offbuf, offs_staged);
        int mval = 0;                                              // mark to store (0: none)
        bool single = false;
        double rho_x = 0.0, rho_y = 0.0;
        const double energy = bin_energy(u);
        if (energy > thr) {                                        // not a zeroton (wave-uniform)
            ++nz;
            if (pd.ent >= 0 || rebuilt) {                          // the column changed: refresh the LDS copy
#pragma unroll
                for (int r = 0; r < NR; ++r) {
                    const int p = lane + 64 * r;
                    if (p < P) cw[p] = u[r];
                }
            }
            if (lane < nw) kk[lane] = 0;
            wave_lds_sync();

            // --- k: one coordinate per lane (coded source: one syndrome digit
            // per lane, decoded into k below)
            const bool coded = (d.method & QSFT_DETECT_CODED) != 0;
            const int nch = coded ? d.p1 - 1 : n;
            int8_t* dig = coded ? W.syn : kb8;
            for (int i0 = 1; i0 <= nch; i0 += 64) {
                const int i = i0 + lane;
                if (i <= nch) dig[i - 1] = (int8_t)coord_digit(d, cw, i, q, P);
            }
            wave_lds_sync();
            if (coded) rs_decode_wave(d, W.syn, kb8, W.rs, lane);

            // --- signature, rho, residual (rows across lanes); the column is
            // re-read from LDS so its registers are free during the k digits
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                const int p = lane + 64 * r;
                u[r] = (p < P) ? cw[p] : make_double2(0.0, 0.0);
            }
            const QSFT_GLOBAL int8_t* __restrict__ Dc =
                gptr(d.D) + (size_t)((d.flags & QSFT_PEEL_SHARED_D) ? 0 : c) * P * n_pad;
            double2 w[NR];
            row_phases(d, Dc, kk, tab, mq, lane, w, offbuf, offs_staged);
            double rr = 0.0, ri = 0.0;
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                const int p = lane + 64 * r;
                if (p < P) conj_dot_acc(w[r], u[r], rr, ri);       // conj(w) * u
            }
            rr = wave_sum(rr);
            ri = wave_sum(ri);
            rho_x = rr * inv_p;                                    // within 1 ulp of rr / P
            rho_y = ri * inv_p;
            // residual energy |col - rho sig|^2 = |col|^2 - P |rho|^2 (|sig_p| = 1 and
            // sig^H col = P rho); its rounding error is ~1e-15 |col|^2. Where that
            // could exceed 1e-3 of the threshold (|col|^2 > 1e12 thr: large
            // amplitudes against the noiseless cutoff 1e-9 P) the residual is
            // summed explicitly over the rows instead, as the reference forms it.
            double res = residual_energy(energy, P, rho_x, rho_y);
            if (energy > 1e12 * thr) {                             // wave-uniform
                double s = 0.0;
#pragma unroll
                for (int r = 0; r < NR; ++r) {
                    const int p = lane + 64 * r;
                    if (p < P) residual_acc(u[r], w[r], rho_x, rho_y, s);
                }
                res = wave_sum(s);
            }

            // --- bin match: the bin of k in group c, sum over digits dg (lane dg)
            // of (<M_c^T row dg, k> mod q) q^(b - 1 - dg), equals j
            int hd = 0;
            if (lane < b) {
                const size_t mo = ((size_t)c * b + lane) * n_pad;
                hd = (int)mq(mt_staged ? dot_row_lds(mtbuf + mo, kk, nw) : dot_row(gptr(d.Mt) + mo, kk, nw)) * wdig;
            }
            const bool match = __builtin_amdgcn_readfirstlane(row_sum_i(hd)) == j;   // b <= 16: lanes of row 0
            single = match && res <= thr;
            mval = single ? it + 1 : (it + 1) | kHeldMulti;        // non-zeroton: due next pass
            if (single) {
                const uint64_t hk = hash_k(kk, nw, lane);
                if (lane == 0) {
                    W.sb_t[nb] = t;
                    W.sb_h[nb] = hk;
                    W.sb_rho[nb] = make_double2(rho_x, rho_y);
                }
                ++ns;
                ++nb;
            }
        }
        wait_vm<0>();                                              // the next bin's data has landed
        __asm__ volatile("" ::: "memory");
        // this bin's stores
        if (pd.ent >= 0 || rebuilt) store_col(d, t, u, lane);
        if (mval && lane == 0) mark[t] = mval;
        if (single) {
            if (lane < nw) reinterpret_cast<QSFT_GLOBAL uint32_t*>(kcur + (size_t)t * n_pad)[lane] = kk[lane];
            if (lane == 0) {
                rcur[2 * (size_t)t] = rho_x;
                rcur[2 * (size_t)t + 1] = rho_y;
            }
        }
        if (!more) break;
        cur = nxt;
        curent = nent;
        cslot ^= 1;
        wave_lds_sync();                                           // kk / this slot are rewritten next
        if constexpr (CAP < kBinsPerWave) {
            if (nb == CAP) {                                       // the buffer is full: register it
                register_singletons(d, it, nb, W.sb_t, W.sb_h, W.sb_rho, lane);
                nb = 0;
                wave_lds_sync();
            }
        }
    }

    finish_wave(d, it, scanned, nz, ns, W.sb_t, W.sb_h, W.sb_rho, lane, nb);
}

// ---- The first pass, four bins per wave ------------------------------------
// The first pass has no queued balls and every bin it visits was written by
// the assembler, so a wave can classify four bins at once: 16 lanes per bin,
// rows p = l + 16 r (r < 8, P <= 128), coordinates i = 1 + l + 16 m. One
// instruction then serves four bins where the one-bin layout spends it on one
// (its sums keep only 2 rows per lane busy and its dot products, hash and
// bin match a few lanes). Every value is formed exactly as detect_bins forms
// it: lane l of a bin holds the rows of that layout's lanes l + 16 r' (r' < 4:
// rows l + 16 r' and l + 16 r' + 64), so its four per-lane partial sums are
// that layout's per-lane sums, the in-row butterfly is the same DPP sequence,
// and the four row totals combine as (R0 + R1) + (R2 + R3), as the cross-row
// swaps of wave_sum do: energies, rho, residuals, digits, phases and hashes
// come out bit for bit as in the one-bin layout.
constexpr int kQuad = 4;                 // bins a wave classifies at once
constexpr int kQRows = 8;                // rows per lane (P <= 128)

struct QuadWaveLds {
    __align__(16) double2 col[kQuad][128];         // the current quad's columns
    __align__(16) uint32_t kws[kQuad][16];         // each bin's k (n_pad = 64)
    __align__(16) int32_t off[kQuad][4];           // each bin's <offset_r, k> mod q (unit delays)
    uint64_t sb_h[kSbCap];
    double2 sb_rho[kSbCap];
    int32_t sb_t[kSbCap];
};

// In-row (16 lanes) part of wave_sum: the same DPP butterfly.
__device__ __forceinline__ double row16_sum(double v) {
    v += dpp_f64<0xB1>(v);
    v += dpp_f64<0x4E>(v);
    v += dpp_f64<0x141>(v);
    v += dpp_f64<0x140>(v);
    return v;
}
// The four virtual rows' partial sums (one per r') to the wave_sum total.
__device__ __forceinline__ double quad_total(const double (&s)[4]) {
    const double r0 = row16_sum(s[0]), r1 = row16_sum(s[1]), r2 = row16_sum(s[2]), r3 = row16_sum(s[3]);
    return (r0 + r1) + (r2 + r3);
}
// hash_k over the 16 words of a bin's k held by the 16 lanes of its row.
__device__ __forceinline__ uint64_t hash_k16(uint32_t word, int l16) {
    uint64_t v = mix64((uint64_t)word ^ ((uint64_t)(l16 + 1) * 0x9E3779B97F4A7C15ull));
    v ^= (uint64_t)dpp_u32<0xB1>((uint32_t)v) | ((uint64_t)dpp_u32<0xB1>((uint32_t)(v >> 32)) << 32);
    v ^= (uint64_t)dpp_u32<0x4E>((uint32_t)v) | ((uint64_t)dpp_u32<0x4E>((uint32_t)(v >> 32)) << 32);
    v ^= (uint64_t)dpp_u32<0x141>((uint32_t)v) | ((uint64_t)dpp_u32<0x141>((uint32_t)(v >> 32)) << 32);
    v ^= (uint64_t)dpp_u32<0x140>((uint32_t)v) | ((uint64_t)dpp_u32<0x140>((uint32_t)(v >> 32)) << 32);
    return mix64(v ^ 0xCBF29CE484222325ull);
}

// The wave's due bins (first pass), four at a time. Preconditions (checked by
// the caller): no coded source, n_pad = 64, P <= 128, unit delays with R <= 4
// staged offset rows (or general delays), every due bin written.
__device__ __forceinline__ void detect_quads(const qsft_peel_desc& d, int base, uint64_t todo, uint32_t nheld, QuadWaveLds& W,
                             const double2* tab, const int8_t* offbuf, const int8_t* mtbuf, bool mt_staged) {
    constexpr int it = 0;
    const int lane = threadIdx.x & 63, g = lane >> 4, l16 = lane & 15;
    const int q = d.q, n = d.n, b = d.b, P = d.P, Q = d.Q, n_pad = d.n_pad;
    const int nw = n_pad >> 2;                                     // 16
    const ModQ mq((uint32_t)q);
    const double thr = d.cutoff * (double)P;
    const double inv_p = 1.0 / (double)P;
    const double2* __restrict__ Ub = reinterpret_cast<const double2*>(d.U);
    const int CQ = d.C * Q;
    QSFT_GLOBAL int32_t* __restrict__ mark = gptr(d.mark);
    QSFT_GLOBAL int8_t* __restrict__ kcur = gptr(d.kbin);         // pass 0: slot (it & 1) = 0
    QSFT_GLOBAL double* __restrict__ rcur = gptr(d.rho_bin);
    const bool unit = (d.flags & QSFT_PEEL_UNIT_D) != 0;
    const bool neg = (d.flags & QSFT_PEEL_UNIT_D_NEG) != 0;
    const int p1 = d.p1, R = P / p1;
    const uint32_t m1 = (uint32_t)(0xFFFFFFFFull / (uint32_t)p1) + 1u;   // p / p1 exact for p < 256
    int wdig = 0;                                                  // q^(b - 1 - l16): weight of bin digit l16
    if (l16 < b) {
        wdig = 1;
        for (int e = l16 + 1; e < b; ++e) wdig *= q;
    }
    const uint32_t scanned = (uint32_t)__popcll(todo);
    int nz = (int)nheld, ns = 0;                                   // wave-uniform
    int nbuf = 0;                                                  // singletons buffered in W.sb_*
    const int nrow = (P + 63) >> 6;
    const uint32_t rowoff = (uint32_t)min(lane, P - 1) * 16u;
    // the next up to four due bins (wave-uniform), -1 past the end
    auto take = [&](uint64_t& m, int (&tb)[kQuad]) {
        int cnt = 0;
#pragma unroll
        for (int k = 0; k < kQuad; ++k) {
            tb[k] = -1;
            if (m) {
                tb[k] = __builtin_ctzll(m);
                m &= m - 1;
                ++cnt;
            }
        }
        return cnt;
    };
    auto issue = [&](const int (&tb)[kQuad]) {
#pragma unroll
        for (int k = 0; k < kQuad; ++k) {
            if (tb[k] < 0) continue;
            const double2* cb = Ub + (size_t)(base + tb[k]) * P;
            for (int r = 0; r < nrow; ++r)
                dma16(cb, min(64 * r * 16u + rowoff, (uint32_t)(P - 1) * 16u),
                      __builtin_amdgcn_readfirstlane(lds_addr(&W.col[k][64 * r])));
        }
    };
    uint64_t rest = todo;
    int cb[kQuad];
    take(rest, cb);
    const int cbase = base / Q;
    for (;;) {
        // the quad's columns into LDS (one buffer: the other waves of the
        // SIMD hide the load)
        issue(cb);
        wait_vm<0>();
        __asm__ volatile("" ::: "memory");
        int nb[kQuad];
        const int ncnt = take(rest, nb);
        const int my = g == 0 ? cb[0] : g == 1 ? cb[1] : g == 2 ? cb[2] : cb[3];
        const bool active = my >= 0;
        const int t = base + (active ? my : 0);
        int c = cbase;
        while (t >= (c + 1) * Q) ++c;
        const int j = t - c * Q;
        const double2* cw = W.col[g];
        // row r of the lane (p = l16 + 16 r) from the LDS column, read where it
        // is used so that no column is held in registers across the k digits
        auto row = [&](int r) {
            const int p = l16 + 16 * r;
            return (active && p < P) ? cw[p] : make_double2(0.0, 0.0);
        };
        // |col|^2: bin_energy's per-lane sums of the one-bin layout (rows L, L + 64)
        double es[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const double2 a = row(v), e = row(v + 4);
            double sv = 0.0;
            sv = __fma_rn(a.x, a.x, sv);
            sv = __fma_rn(a.y, a.y, sv);
            sv = __fma_rn(e.x, e.x, sv);
            sv = __fma_rn(e.y, e.y, sv);
            es[v] = sv;
        }
        const double energy = quad_total(es);
        const bool nonzero = active && energy > thr;               // uniform over the bin's 16 lanes
        nz += (int)__popcll(__ballot(l16 == 0 && nonzero));
        W.kws[g][l16] = 0u;
        wave_lds_sync();
        // --- k: coordinates i = 1 + l16 + 16 m of the bin
        int8_t* kb8 = reinterpret_cast<int8_t*>(W.kws[g]);
        if (nonzero) {
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const int i = 1 + l16 + 16 * m;
                if (i <= n) kb8[i - 1] = (int8_t)coord_digit(d, cw, i, q, P);
            }
        }
        wave_lds_sync();
        // --- signature phases, rho, residual
        const uint32_t* kk = W.kws[g];
        int o[4] = {0, 0, 0, 0};
        if (unit) {
            if (nonzero && l16 < R) W.off[g][l16] = (int)mq(dot_row_lds(offbuf + (size_t)l16 * n_pad, kk, nw));
            wave_lds_sync();
            const int4 ov = *reinterpret_cast<const int4*>(W.off[g]);
            o[0] = ov.x;
            o[1] = ov.y;
            o[2] = ov.z;
            o[3] = ov.w;
        }
        const QSFT_GLOBAL int8_t* __restrict__ Dc =
            gptr(d.D) + (size_t)((d.flags & QSFT_PEEL_SHARED_D) ? 0 : c) * P * n_pad;
        // signature phase of row r of the lane, formed where it is used
        auto phase = [&](int r) {
            const int p = l16 + 16 * r;
            double2 w = make_double2(0.0, 0.0);
            if (nonzero && p < P) {
                if (unit) {
                    const int blk = (int)__umulhi((uint32_t)p, m1);
                    const int i = p - blk * p1;
                    int e = blk == 0 ? o[0] : blk == 1 ? o[1] : blk == 2 ? o[2] : o[3];
                    if (i > 0) e += neg ? q - (int)kb8[i - 1] : (int)kb8[i - 1];
                    if (e >= q) e -= q;
                    w = tab[e];
                } else {
                    w = tab[mq(dot_row(Dc + (size_t)p * n_pad, kk, nw))];
                }
            }
            return w;
        };
        double rs[4], is[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            double rr = 0.0, ri = 0.0;
            if (l16 + 16 * v < P) conj_dot_acc(phase(v), row(v), rr, ri);
            if (l16 + 16 * v + 64 < P) conj_dot_acc(phase(v + 4), row(v + 4), rr, ri);
            rs[v] = rr;
            is[v] = ri;
        }
        const double rho_x = quad_total(rs) * inv_p, rho_y = quad_total(is) * inv_p;
        double res = residual_energy(energy, P, rho_x, rho_y);
        if (__ballot(nonzero && energy > 1e12 * thr)) {          // explicit residual (as detect_bins)
            double ss[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                double sv = 0.0;
                if (l16 + 16 * v < P) residual_acc(row(v), phase(v), rho_x, rho_y, sv);
                if (l16 + 16 * v + 64 < P) residual_acc(row(v + 4), phase(v + 4), rho_x, rho_y, sv);
                ss[v] = sv;
            }
            const double sres = quad_total(ss);
            if (energy > 1e12 * thr) res = sres;
        }
        // --- bin match: sum over digits dg (lane dg of the bin) of (<M_c^T row dg, k> mod q) q^(b - 1 - dg)
        int hd = 0;
        if (nonzero && l16 < b) {
            const size_t mo = ((size_t)c * b + l16) * n_pad;
            hd = (int)mq(mt_staged ? dot_row_lds(mtbuf + mo, kk, nw) : dot_row(gptr(d.Mt) + mo, kk, nw)) * wdig;
        }
        const bool match = row_sum_i(hd) == j;
        const bool single = nonzero && match && res <= thr;
        const int mval = nonzero ? (single ? it + 1 : (it + 1) | kHeldMulti) : 0;
        const uint64_t hk = hash_k16(kk[l16], l16);
        const uint64_t smask = __ballot(l16 == 0 && single);
        const int nnew = (int)__popcll(smask);
        if (nbuf + nnew > kSbCap) {                                // no room: register the buffer first
            register_singletons(d, it, nbuf, W.sb_t, W.sb_h, W.sb_rho, lane);
            nbuf = 0;
            wave_lds_sync();
        }
        if (l16 == 0 && single) {
            const int at = nbuf + (int)__popcll(smask & ((1ull << lane) - 1ull));
            W.sb_t[at] = t;
            W.sb_h[at] = hk;
            W.sb_rho[at] = make_double2(rho_x, rho_y);
        }
        ns += nnew;
        nbuf += nnew;
        if (mval && l16 == 0) mark[t] = mval;
        if (single) {
            reinterpret_cast<QSFT_GLOBAL uint32_t*>(kcur + (size_t)t * n_pad)[l16] = kk[l16];
            if (l16 == 0) {
                rcur[2 * (size_t)t] = rho_x;
                rcur[2 * (size_t)t + 1] = rho_y;
            }
        }
        if (!ncnt) break;
#pragma unroll
        for (int k = 0; k < kQuad; ++k) cb[k] = nb[k];
        wave_lds_sync();                                           // kws / the column buffer are rewritten next
    }
    (void)CQ;
    finish_wave(d, it, scanned, nz, ns, W.sb_t, W.sb_h, W.sb_rho, lane, nbuf);
}

// Kernel A of the first pass: detect_quads where its preconditions hold (and
// QSFT_PEEL_ONE_BIN_WAVES is not set), else detect_bins. Same grid and bin
// partition as k_peel_detect. The block's LDS (one column buffer of four bins
// per wave, 16-entry singleton buffers, half-size staging) is under 40 KiB, so
// four blocks fit a CU: four waves per SIMD (128 VGPRs) with four bins each.
// Measured on the K1 64-batch's first pass against the earlier forms: two
// waves with two column buffers 0.24 ms, three waves with one 0.21 ms, four
// waves 0.17 ms (the pass waits on memory; more bins in flight beat prefetch).
constexpr int kQuadOffBytes = 1024;
constexpr int kQuadMtBytes = 2048;

__global__ __launch_bounds__(kBinThreads) __attribute__((amdgpu_waves_per_eu(4))) void k_peel_detect0(
    const qsft_peel_desc* __restrict__ descs, const int32_t* __restrict__ pre, int n_inst) {
    __shared__ double2 tab[QSFT_MAX_Q];
    __shared__ union WaveLds {
        QuadWaveLds quad;
        DetectWaveLds<2> one;
    } wl[kWaves];
    __shared__ __align__(16) int8_t offbuf[kQuadOffBytes];
    __shared__ __align__(16) int8_t mtbuf[kQuadMtBytes];
    const int inst = find_inst(pre, n_inst, blockIdx.x);
    const qsft_peel_desc& d = descs[inst];
    uint32_t* __restrict__ ctr = d.ctr;
    constexpr int it = 0;
    if (it >= d.max_iter) return;
    if (cptr(ctr)[it * QSFT_CTR_ITER_STRIDE + 3] == 0) return;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int base = ((blockIdx.x - pre[inst]) * kWaves + wave) * kBinsPerWave;
    uint32_t nheld;
    const uint64_t todo = scan_marks(d, it, base, nheld);
    if (!__syncthreads_or(todo != 0)) {
        if (lane == 0 && nheld) atomicAdd(ctr + it * QSFT_CTR_ITER_STRIDE + 1, nheld);
        return;
    }
    fill_wtable_d(tab, d.q);
    const bool offs_staged = stage_offsets<kQuadOffBytes>(d, offbuf);
    const bool mt_staged = stage_mt<kQuadMtBytes>(d, mtbuf);
    __syncthreads();
    const bool unit = (d.flags & QSFT_PEEL_UNIT_D) != 0;
    // every due bin of the wave written by the assembler (sparse U)
    const bool written = !sparse_u(d) || __ballot(((todo >> lane) & 1) && !gptr(d.written)[base + lane]) == 0;
    const bool quads = !(d.flags & QSFT_PEEL_ONE_BIN_WAVES) && !(d.method & QSFT_DETECT_CODED) && d.n_pad == 64 &&
                       d.P <= 128 && (!unit || (offs_staged && d.P / d.p1 <= 4)) && written;
    if (quads) detect_quads(d, base, todo, nheld, wl[wave].quad, tab, offbuf, mtbuf, mt_staged);
    else detect_bins<2, false>(d, it, base, todo, nheld, wl[wave].one, tab, offbuf, offs_staged, mtbuf, mt_staged);
}

// Kernel A. Wave w of an instance owns bins [64 w, 64 w + 64): it classifies
// the ones this pass scans (mark == pass), prefetching
// the next bin's column while it works on the current one, and buffers its
// singletons in LDS. Then one lane per buffered singleton registers it in the
// hash table, so the atomics of up to 16 registrations are in flight
// together, and the pass counters take one atomic per wave. Four waves per
// SIMD fit without scratch for two rows per lane (at most 128 VGPRs; three for
// four rows): the FP64 atan2 rules, whose registers would otherwise be
// reserved across the whole loop, are out of line.
template <int NR>
__global__ __launch_bounds__(kBinThreads) __attribute__((amdgpu_waves_per_eu(NR == 2 ? 4 : 3))) void k_peel_detect(const qsft_peel_desc* __restrict__ descs,
                                                             const int32_t* __restrict__ pre, int n_inst, int it) {
    __shared__ double2 tab[QSFT_MAX_Q];
    __shared__ DetectWaveLds<NR> wl[kWaves];
    __shared__ __align__(16) int8_t offbuf[kStageOffBytes];
    __shared__ __align__(16) int8_t mtbuf[kStageMtBytes];
    const int inst = find_inst(pre, n_inst, blockIdx.x);
    const qsft_peel_desc& d = descs[inst];
    uint32_t* __restrict__ ctr = d.ctr;
    if (it >= d.max_iter) return;                                  // beyond the bound workspace
    if (cptr(ctr)[it * QSFT_CTR_ITER_STRIDE + 3] == 0) return;    // loop already stopped (set by the
                                                                   // previous pass; scalar load)
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int base = ((blockIdx.x - pre[inst]) * kWaves + wave) * kBinsPerWave;
    uint32_t nheld;
    const uint64_t todo = scan_marks(d, it, base, nheld);
    if (!__syncthreads_or(todo != 0)) {                            // nothing to read in this block
        if (lane == 0 && nheld) atomicAdd(ctr + it * QSFT_CTR_ITER_STRIDE + 1, nheld);
        return;
    }
    fill_wtable_d(tab, d.q);
    const bool offs_staged = stage_offsets(d, offbuf);
    const bool mt_staged = stage_mt(d, mtbuf);
    __syncthreads();
    detect_bins<NR, false>(d, it, base, todo, nheld, wl[wave], tab, offbuf, offs_staged, mtbuf, mt_staged);
}

constexpr int kKeyChunks = QSFT_MAX_N / 16;                        // 16-byte chunks of a k row

// The checks and queue pushes of kernel B for the singleton found in bin t.
// mt: the block's LDS copy of every group's M^T rows (stage_mt), or null to
// read them from global memory.
__device__ __forceinline__ void apply_singleton(const qsft_peel_desc& d, int t, int it, int q, int b, int C, int Q,
                                                int n_pad, int nq, uint32_t* ctr, const int8_t* mt) {
    const int CQ = C * Q;
    const ModQ mq((uint32_t)q);
    const int8_t* __restrict__ kcur = d.kbin + (size_t)(it & 1) * CQ * n_pad;
    const uint4* __restrict__ krow = reinterpret_cast<const uint4*>(kcur + (size_t)t * n_pad);
    uint4 k[kKeyChunks];
#pragma unroll
    for (int w = 0; w < kKeyChunks; ++w) k[w] = (w < nq) ? krow[w] : make_uint4(0u, 0u, 0u, 0u);
    const int slot = d.slot_bin[t];
    const uint64_t tm = ((uint64_t)(it + 1) << 32) | (uint32_t)t;
    const uint64_t own = d.owner[slot];
    const uint64_t lst = d.last[slot];
    // the detection's k must equal its slot's key
    uint4* kslot = reinterpret_cast<uint4*>(d.key + (size_t)slot * n_pad);
    bool same = true;
    if (own == tm) {                                               // first detection of the key: it writes the key
#pragma unroll
        for (int w = 0; w < kKeyChunks; ++w)
            if (w < nq) kslot[w] = k[w];
    } else {
        // a key first seen in this pass is compared with that detection's k
        // (its key row may not be written yet), an older key with its key row
        const uint4* ref = ((own >> 32) == (uint64_t)(it + 1))
                               ? reinterpret_cast<const uint4*>(kcur + (size_t)(uint32_t)own * n_pad)
                               : kslot;
#pragma unroll
        for (int w = 0; w < kKeyChunks; ++w) {
            if (w < nq) {
                const uint4 r = ref[w];
                same = same && r.x == k[w].x && r.y == k[w].y && r.z == k[w].z && r.w == k[w].w;
            }
        }
    }
    if (!same) atomicOr(&ctr[QSFT_CTR_ERROR], (uint32_t)QSFT_DEVERR_TAG_COLLISION);
    if (lst != tm) return;
    // the last detection of this key in this pass queues it on every group:
    // bin of group l = sum_dg (<M_l^T row dg, k> mod q) q^(b - 1 - dg), MSB first
    for (int l = 0; l < C; ++l) {
        int jl = 0;
        for (int dg = 0; dg < b; ++dg) {
            const size_t mo = ((size_t)l * b + dg) * n_pad;
            int acc = 0;
            if (mt) {                                              // (block-uniform)
                const uint4* mrow = reinterpret_cast<const uint4*>(mt + mo);
#pragma unroll
                for (int w = 0; w < kKeyChunks; ++w) {
                    if (w < nq) {
                        const uint4 mv = mrow[w];
                        acc = __builtin_amdgcn_sdot4((int)mv.x, (int)k[w].x, acc, false);
                        acc = __builtin_amdgcn_sdot4((int)mv.y, (int)k[w].y, acc, false);
                        acc = __builtin_amdgcn_sdot4((int)mv.z, (int)k[w].z, acc, false);
                        acc = __builtin_amdgcn_sdot4((int)mv.w, (int)k[w].w, acc, false);
                    }
                }
            } else {
                const uint4* __restrict__ mrow = reinterpret_cast<const uint4*>(d.Mt + mo);
#pragma unroll
                for (int w = 0; w < kKeyChunks; ++w) {
                    if (w < nq) {
                        const uint4 mv = mrow[w];
                        acc = __builtin_amdgcn_sdot4((int)mv.x, (int)k[w].x, acc, false);
                        acc = __builtin_amdgcn_sdot4((int)mv.y, (int)k[w].y, acc, false);
                        acc = __builtin_amdgcn_sdot4((int)mv.z, (int)k[w].z, acc, false);
                        acc = __builtin_amdgcn_sdot4((int)mv.w, (int)k[w].w, acc, false);
                    }
                }
            }
            jl = jl * q + (int)mq((uint32_t)acc);
        }
        const int target = l * Q + jl;
        const int e = t * C + l;
        d.pend_next[e] = atomicExch(&d.pend_head[target], e);
        d.mark[target] = it + 1;                                   // the peeled bin is rescanned
    }
}

// Kernel B. Same bin partition as kernel A; lane L of a wave handles bin
// 64 w + L if this pass found a singleton there (sstamp == pass + 1), all of
// the wave's singletons at once: the detection's k (held in registers) must
// equal its slot's key; the last detection of the key queues the ball on its
// bin in every group (the C x b bin digits are dot products of the k with the
// M^T rows; one atomicExch per target) and stamps the target for the next
// pass, whose kernel A subtracts it from the column it owns.

// Loop control after pass `it` (one thread; its detection counters are final):
// the iteration and peel counts, and whether pass it + 1 runs: only if this
// pass ran and found a singleton and a multiton, below the peel cap.
template <bool PERSIST>
__device__ __forceinline__ void loop_control(const qsft_peel_desc& d, int it, uint32_t active) {
    uint32_t* __restrict__ ctr = d.ctr;
    auto rd = [&](int w) -> uint32_t {
        if constexpr (PERSIST) return __hip_atomic_load(ctr + w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        else return ctr[w];
    };
    const uint32_t ns = rd(it * QSFT_CTR_ITER_STRIDE + 0);
    const uint32_t nm = rd(it * QSFT_CTR_ITER_STRIDE + 1);
    const uint32_t nb = rd(it * QSFT_CTR_ITER_STRIDE + 2);
    uint32_t peeled = rd(QSFT_CTR_PEELED);
    if (active) {
        ctr[QSFT_CTR_ITERS] = (uint32_t)(it + 1);
        peeled = (peeled + nb < peeled) ? 0xFFFFFFFFu : peeled + nb;
        ctr[QSFT_CTR_PEELED] = peeled;
    }
    const uint32_t next = (active && ns > 0 && nm > 0 && (double)peeled < d.peel_cap) ? 1u : 0u;
    if (it + 1 < 16) ctr[(it + 1) * QSFT_CTR_ITER_STRIDE + 3] = next;
}

__global__ __launch_bounds__(kBinThreads) void k_peel_apply(const qsft_peel_desc* __restrict__ descs,
                                                            const int32_t* __restrict__ pre, int n_inst, int it) {
    __shared__ __align__(16) int8_t mtbuf[kStageMtBytes];
    const int inst = find_inst(pre, n_inst, blockIdx.x);
    const qsft_peel_desc& d = descs[inst];
    uint32_t* __restrict__ ctr = d.ctr;
    const int local = blockIdx.x - pre[inst];
    if (it >= d.max_iter) return;
    const uint32_t active = cptr(ctr)[it * QSFT_CTR_ITER_STRIDE + 3];   // written before this kernel
    if (local == 0 && threadIdx.x == 0) loop_control<false>(d, it, active);   // the pass's counters are final
    if (!active) return;
    const int q = d.q, b = d.b, C = d.C, Q = d.Q, n_pad = d.n_pad;
    const int nq = n_pad >> 4;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int CQ = C * Q;
    const int t = (local * kWaves + wave) * kBinsPerWave + lane;
    const bool mine = t < CQ && d.sstamp[t] == it + 1;
    if (!__syncthreads_or(mine)) return;                           // (active is the same for the whole block)
    const bool mt_staged = stage_mt(d, mtbuf);                     // the targets' bin digits read M^T from LDS
    __syncthreads();
    if (mine) apply_singleton(d, t, it, q, b, C, Q, n_pad, nq, ctr, mt_staged ? mtbuf : nullptr);   // no cross-lane work
}

// Passes it0 .. it_end - 1 in one launch, one workgroup per instance (the
// fused tail of the loop: after the first passes most loops have stopped and
// the rest touch a few bins per pass, so per-pass launches would cost more
// than the work). Each pass runs kernel A's body over the instance's 64-bin
// groups (wave w takes groups w, w + kW, ...; the marks of kTailBatch groups
// are loaded together), a barrier, the loop control, kernel B's body over the
// instance's bins (kTailBatch stamps loaded together), and a barrier. The stop
// flag is read by one thread and broadcast through LDS, so every thread
// leaves the loop at the same pass. Only this workgroup touches its
// instance's arrays, so the barriers' workgroup-scope ordering is all the
// passes need; arrays written earlier in the launch are read by vector loads
// (PERSIST), never through the scalar cache.
template <int NR>
constexpr int tail_waves() { return NR == 2 ? 12 : 8; }           // no scratch: 170 / 256 VGPRs per lane
constexpr int kTailBatch = 8;

template <int NR>
__global__ __launch_bounds__(64 * tail_waves<NR>()) void k_peel_tail(const qsft_peel_desc* __restrict__ descs,
                                                                  int it0, int it_end) {
    constexpr int kW = tail_waves<NR>();
    __shared__ double2 tab[QSFT_MAX_Q];
    __shared__ DetectWaveLds<NR, kBinsPerWave> wl[kW];
    __shared__ __align__(16) int8_t offbuf[kStageOffBytes];
    __shared__ __align__(16) int8_t mtbuf[kStageMtBytes];
    __shared__ uint64_t tdue[kW][kTailBatch];                      // due masks of the wave's scanned groups
    __shared__ uint32_t theld[kW][kTailBatch];
    __shared__ uint32_t go;
    const qsft_peel_desc& d = descs[blockIdx.x];
    uint32_t* __restrict__ ctr = d.ctr;
    const int end = min(it_end, min(d.max_iter, 15));
    if (it0 >= end) return;                                        // (block-uniform)
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int CQ = d.C * d.Q, n_pad = d.n_pad;
    constexpr int kGroupStride = kW * kBinsPerWave;
    fill_wtable_d(tab, d.q);
    const bool offs_staged = stage_offsets(d, offbuf);
    const bool mt_staged = stage_mt(d, mtbuf);
    for (int it = it0; it < end; ++it) {
        if (threadIdx.x == 0)
            go = __hip_atomic_load(ctr + it * QSFT_CTR_ITER_STRIDE + 3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();                                           // go and (first pass) the tables
        if (go == 0) break;                                        // the loop has stopped (every thread)
        for (int g0 = wave * kBinsPerWave; g0 < CQ; g0 += kGroupStride * kTailBatch) {
            uint64_t due[kTailBatch];
            uint32_t nh[kTailBatch];
#pragma unroll
            for (int i = 0; i < kTailBatch; ++i) due[i] = scan_marks(d, it, g0 + i * kGroupStride, nh[i]);
            if (lane == 0) {
#pragma unroll
                for (int i = 0; i < kTailBatch; ++i) {
                    tdue[wave][i] = due[i];
                    theld[wave][i] = nh[i];
                }
            }
            wave_lds_sync();
            for (int i = 0; i < kTailBatch; ++i) {
                const uint64_t todo = tdue[wave][i];
                const uint32_t nheld = theld[wave][i];
                if (todo != 0 || nheld != 0)
                    detect_bins<NR, true>(d, it, g0 + i * kGroupStride, todo, nheld, wl[wave], tab, offbuf,
                                          offs_staged, mtbuf, mt_staged);
            }
            wave_lds_sync();                                       // tdue is rewritten next
        }
        __syncthreads();                                           // every detection registered, counters final
        if (threadIdx.x == 0) loop_control<true>(d, it, 1u);
        for (int t0 = threadIdx.x; t0 < CQ; t0 += blockDim.x * kTailBatch) {
            uint32_t hit = 0;
#pragma unroll
            for (int i = 0; i < kTailBatch; ++i) {
                const int t = t0 + i * (int)blockDim.x;
                if (t < CQ && gptr(d.sstamp)[t] == it + 1) hit |= 1u << i;
            }
            while (hit) {
                const int t = t0 + __builtin_ctz(hit) * (int)blockDim.x;
                hit &= hit - 1;
                apply_singleton(d, t, it, d.q, d.b, d.C, d.Q, n_pad, n_pad >> 4, ctr, mt_staged ? mtbuf : nullptr);
            }
        }
        __syncthreads();                                           // queues and stamps for the next pass
    }
}

// Output of one key: the running mean of its records in time order, gathered
// kFinRecs at a time by walks of the key's list.
constexpr int kFinRecs = 8;

__global__ __launch_bounds__(kSlotThreads) void k_peel_finalize(const qsft_peel_desc* __restrict__ descs,
                                                                const int32_t* __restrict__ pre, int n_inst) {
    // one thread per claimed slot (the claim order of the registration), so
    // the output row of a key is its position in that list
    const int inst = find_inst(pre, n_inst, blockIdx.x);
    const qsft_peel_desc& d = descs[inst];
    const uint32_t nlive = d.ctr[QSFT_CTR_LIVE];
    const uint32_t o = (uint32_t)(blockIdx.x - pre[inst]) * kSlotThreads + threadIdx.x;
    if (o == 0) {
        d.ctr[QSFT_CTR_OUT] = nlive;
        if (nlive > (uint32_t)d.out_cap) atomicOr(&d.ctr[QSFT_CTR_ERROR], (uint32_t)QSFT_DEVERR_OUT_FULL);
    }
    if (o >= nlive || o >= (uint32_t)d.out_cap) return;
    const uint32_t h = d.live[o];
    const uint32_t m = d.count[h];
    const uint32_t L = d.ctr[QSFT_CTR_LOG];
    // running mean over the records in time order: g <- (g * c + v) / (c + 1)
    double gx = 0.0, gy = 0.0;
    auto accumulate = [&](uint32_t cnt, double vx, double vy) {
        if (cnt == 0) {
            gx = vx;
            gy = vy;
        } else {
            const double cd = (double)cnt;
            gx = __ddiv_rn(__dadd_rn(__dmul_rn(gx, cd), vx), cd + 1.0);
            gy = __ddiv_rn(__dadd_rn(__dmul_rn(gy, cd), vy), cd + 1.0);
        }
    };
    // Each walk of the list keeps, in registers and sorted by time, the
    // kFinRecs earliest records later than the last one accumulated, then
    // accumulates them in order: ceil(m / kFinRecs) walks (one for most keys).
    uint64_t prev = 0;                                             // record times are > 0
    uint32_t done = 0;
    while (done < m) {
        uint64_t tr[kFinRecs];
        double rx[kFinRecs], ry[kFinRecs];
#pragma unroll
        for (int a = 0; a < kFinRecs; ++a) {
            tr[a] = ~0ull;
            rx[a] = ry[a] = 0.0;
        }
        uint32_t steps = 0;
        for (uint32_t r = d.head[h]; r != kNil && steps <= L; r = d.rec_next[r], ++steps) {
            uint64_t t = d.rec_time[r];
            if (t <= prev || t >= tr[kFinRecs - 1]) continue;
            double vx = d.rec_rho[2 * (size_t)r], vy = d.rec_rho[2 * (size_t)r + 1];
#pragma unroll
            for (int a = 0; a < kFinRecs; ++a) {                   // insertion: the later entries shift up
                if (t < tr[a]) {
                    const uint64_t tt = tr[a];
                    const double xx = rx[a], yy = ry[a];
                    tr[a] = t;
                    rx[a] = vx;
                    ry[a] = vy;
                    t = tt;
                    vx = xx;
                    vy = yy;
                }
            }
        }
        uint32_t got = 0;
#pragma unroll
        for (int a = 0; a < kFinRecs; ++a) {
            if (tr[a] != ~0ull && done < m) {
                accumulate(done, rx[a], ry[a]);
                prev = tr[a];
                ++done;
                ++got;
            }
        }
        if (got == 0) break;                                       // (a list shorter than its count)
    }
    const int nw = d.n_pad >> 2;
    const uint32_t* ksrc = reinterpret_cast<const uint32_t*>(d.key + (size_t)h * d.n_pad);
    uint32_t* kdst = reinterpret_cast<uint32_t*>(d.out_k + (size_t)o * d.n_pad);
    for (int w = 0; w < nw; ++w) kdst[w] = ksrc[w];
    d.out_val[2 * (size_t)o] = gx;
    d.out_val[2 * (size_t)o + 1] = gy;
    d.out_order[o] = d.owner[h];
}

// Exclusive scan of the per-instance output counts into the pack header:
// hdr[i] = count_i, hdr[n + i] = offset_i, hdr[2n] = total, hdr[2n + 1] = OR of
// error bits, hdr[2n + 2] = most passes run, hdr[2n + 3] = instances whose loop
// asked for a pass beyond those launched.
__global__ __launch_bounds__(64) void k_pack_scan(const qsft_peel_desc* __restrict__ descs, int n_inst,
                                                  int64_t* __restrict__ hdr, int64_t cap_rows) {
    const int lane = threadIdx.x;
    long long carry = 0;
    uint32_t err = 0;
    int iters = 0, unfinished = 0;
    for (int i0 = 0; i0 < n_inst; i0 += 64) {                     // 64 instances per step, lane-parallel
        const int i = i0 + lane;
        long long m = 0;
        if (i < n_inst) {
            const qsft_peel_desc& d = descs[i];
            uint32_t c = d.ctr[QSFT_CTR_OUT];
            if ((int)c > d.out_cap) c = (uint32_t)d.out_cap;
            err |= d.ctr[QSFT_CTR_ERROR];
            m = c;
            // passes that ran, and whether the loop asked for one more than were launched
            const int it = (int)d.ctr[QSFT_CTR_ITERS];
            iters = max(iters, it);
            if (it < d.max_iter && d.ctr[it * QSFT_CTR_ITER_STRIDE + 3] != 0) ++unfinished;
        }
        long long incl = m;                                        // inclusive scan over the lanes
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const long long v = __shfl_up(incl, off);
            if (lane >= off) incl += v;
        }
        if (i < n_inst) {
            hdr[i] = m;
            hdr[n_inst + i] = carry + incl - m;
        }
        carry += __shfl(incl, 63);
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        err |= (uint32_t)__shfl_xor((int)err, off);
        iters = max(iters, __shfl_xor(iters, off));
        unfinished += __shfl_xor(unfinished, off);
    }
    if (lane == 0) {
        if (carry > cap_rows) err |= QSFT_DEVERR_OUT_FULL;
        hdr[2 * n_inst] = carry;
        hdr[2 * n_inst + 1] = err;
        hdr[2 * n_inst + 2] = iters;
        hdr[2 * n_inst + 3] = unfinished;
    }
}

// One block per instance copies its rows: k digits (n_pad bytes), value (2 doubles), order (u64).
__global__ __launch_bounds__(256) void k_pack_rows(const qsft_peel_desc* __restrict__ descs, int n_inst,
                                                   const int64_t* __restrict__ hdr, unsigned char* __restrict__ rows,
                                                   int64_t cap_rows, int row_stride) {
    const int i = blockIdx.x;
    const qsft_peel_desc& d = descs[i];
    const int64_t m = hdr[i], off = hdr[n_inst + i];
    const int nw = d.n_pad >> 2;
    for (int64_t r = threadIdx.x; r < m; r += blockDim.x) {
        if (off + r >= cap_rows) break;
        unsigned char* dst = rows + (off + r) * row_stride;
        const uint32_t* ks = reinterpret_cast<const uint32_t*>(d.out_k + r * d.n_pad);
        uint32_t* kd = reinterpret_cast<uint32_t*>(dst);
        for (int w = 0; w < nw; ++w) kd[w] = ks[w];
        double* vd = reinterpret_cast<double*>(dst + d.n_pad);
        vd[0] = d.out_val[2 * r];
        vd[1] = d.out_val[2 * r + 1];
        *reinterpret_cast<uint64_t*>(dst + d.n_pad + 16) = d.out_order[r];
    }
}

// The same rows, each written at its instance offset plus its rank in first
// detection order (the reference's dict insertion order), so a caller reads
// every instance's rows already ordered. The rank of a row is the number of
// the instance's rows with a smaller order word (order words are distinct:
// one first detection per pass and bin), counted against LDS tiles of the
// order words. kRankGroups blocks share an instance (block g ranks rows
// g*256 + tid, stepping by kRankGroups*256), so an instance with thousands
// of rows spreads its m^2 comparisons over many CUs.
constexpr int kRankTile = 2048;
constexpr int kRankGroups = 32;

__global__ __launch_bounds__(256) void k_pack_rows_ordered(const qsft_peel_desc* __restrict__ descs, int n_inst,
                                                           const int64_t* __restrict__ hdr,
                                                           unsigned char* __restrict__ rows, int64_t cap_rows,
                                                           int row_stride) {
    __shared__ uint64_t tile[kRankTile];
    const int i = blockIdx.x / kRankGroups, g = blockIdx.x % kRankGroups;
    const qsft_peel_desc& d = descs[i];
    const int64_t m = hdr[i], off = hdr[n_inst + i];
    const int nw = d.n_pad >> 2;
    for (int64_t r0 = (int64_t)g * blockDim.x; r0 < m; r0 += (int64_t)kRankGroups * blockDim.x) {   // block-uniform
        const int64_t r = r0 + threadIdx.x;
        const uint64_t mine = r < m ? d.out_order[r] : ~0ull;
        int64_t rank = 0;
        for (int64_t t0 = 0; t0 < m; t0 += kRankTile) {
            const int tn = (int)min((int64_t)kRankTile, m - t0);
            __syncthreads();                                       // the previous tile is consumed
            for (int j = threadIdx.x; j < tn; j += blockDim.x) tile[j] = d.out_order[t0 + j];
            __syncthreads();
            for (int j = 0; j < tn; ++j) rank += tile[j] < mine ? 1 : 0;
        }
        if (r < m && off + rank < cap_rows) {
            unsigned char* dst = rows + (off + rank) * row_stride;
            const uint32_t* ks = reinterpret_cast<const uint32_t*>(d.out_k + r * d.n_pad);
            uint32_t* kd = reinterpret_cast<uint32_t*>(dst);
            for (int w = 0; w < nw; ++w) kd[w] = ks[w];
            double* vd = reinterpret_cast<double*>(dst + d.n_pad);
            vd[0] = d.out_val[2 * r];
            vd[1] = d.out_val[2 * r + 1];
            *reinterpret_cast<uint64_t*>(dst + d.n_pad + 16) = mine;
        }
    }
}

}  // namespace qsft

using namespace qsft;

extern "C" void qsft_set_error(const char* msg);
extern "C" int qsft_hip_status(hipError_t e, const char* where);

namespace {
size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }
int hcap_log2_for(long long CQ) {
    long long want = 2 * CQ;
    if (want < 1024) want = 1024;
    int l = 10;
    while ((1ll << l) < want) ++l;
    return l;
}
struct PeelLayout {
    size_t off[23];
    size_t total;
};
PeelLayout peel_layout(long long CQ, int C, int n_pad, int hl2, int max_iter, int out_cap) {
    const size_t H = (size_t)1 << hl2;
    const size_t L = (size_t)max_iter * CQ;
    const size_t sizes[] = {
        (size_t)2 * CQ * n_pad,  // 0 kbin (two pass parities)
        (size_t)2 * CQ * 16,     // 1 rho_bin
        (size_t)CQ * 4,          // 2 slot_bin
        H * 8, H * 8, H * 8,     // 3-5 tag owner last
        H * 4, H * 4,            // 6-7 head count
        H * (size_t)n_pad,       // 8 key
        L * 8, L * 16, L * 4,    // 9-11 rec_time rec_rho rec_next
        QSFT_CTR_WORDS * 4,      // 12 ctr
        (size_t)out_cap * n_pad, (size_t)out_cap * 16, (size_t)out_cap * 8,  // 13-15 outputs
        (size_t)CQ * 4, (size_t)CQ * 4,                                      // 16-17 mark sstamp
        (size_t)CQ * 4, (size_t)CQ * 4 * C,                                  // 18-19 pend_head pend_next
        H * 4,                                                               // 20 live
        (size_t)CQ, 8,                                                       // 21-22 written noise_key
    };
    PeelLayout lay{};
    size_t o = 0;
    for (size_t i = 0; i < sizeof(sizes) / sizeof(sizes[0]); ++i) {
        lay.off[i] = o;
        o += align256(sizes[i]);
    }
    lay.total = o;
    return lay;
}
}  // namespace

extern "C" size_t qsft_peel_workspace_bytes(int q, int n, int b, int C, int P, int max_iter, int out_cap) {
    if (q < 2 || q > QSFT_MAX_Q || n < 1 || n > QSFT_MAX_N || b < 1 || b > QSFT_MAX_B || C < 1 || P < 1 ||
        max_iter < 1 || max_iter > 15 || out_cap < 0)
        return 0;
    const long long CQ = (long long)C * ipow(q, b);
    const int n_pad = (n + 63) / 64 * 64;
    const int hl2 = hcap_log2_for(CQ);
    return peel_layout(CQ, C, n_pad, hl2, max_iter, out_cap > 0 ? out_cap : (1 << hl2)).total;
}

extern "C" int qsft_peel_bind(qsft_peel_desc* h, int q, int n, int b, int C, int P, int p1, int method,
                              int max_iter, double cutoff, int out_cap, const int8_t* Mt, const int8_t* D,
                              double* U, void* ws, size_t ws_bytes, int n_pad) {
    if (!h || !ws || q < 2 || q > QSFT_MAX_Q || n < 1 || n > QSFT_MAX_N || b < 1 || b > QSFT_MAX_B || C < 1 ||
        C > QSFT_MAX_C || P < 1 || P > 64 * kRowsPerLane || p1 < 2 || P % p1 || max_iter < 1 || max_iter > 15 ||
        n_pad % 64 || n_pad < n || n_pad > QSFT_MAX_N || (method & ~(QSFT_DETECT_NSO1 | QSFT_DETECT_CODED)) ||
        ((method & QSFT_DETECT_NSO1) == QSFT_DETECT_NOISELESS && P != p1) ||
        (!(method & QSFT_DETECT_CODED) && p1 != n + 1) ||
        ((method & QSFT_DETECT_CODED) && (p1 - 1 > kRsMaxSyn || (p1 - 1) % 2))) {
        qsft_set_error("qsft_peel_bind: bad shape (P <= 256, p1 = n + 1 or, coded, 2 t s + 1 <= 129; "
                       "noiseless needs P = p1)");
        return QSFT_ERR_ARG;
    }
    const long long Qll = ipow(q, b);
    const long long CQ = (long long)C * Qll;
    if (CQ > 0x3FFFFFFF) { qsft_set_error("qsft_peel_bind: too many bins"); return QSFT_ERR_ARG; }
    const int hl2 = hcap_log2_for(CQ);
    if (out_cap <= 0) out_cap = 1 << hl2;
    const PeelLayout lay = peel_layout(CQ, C, n_pad, hl2, max_iter, out_cap);
    if (ws_bytes < lay.total) { qsft_set_error("qsft_peel_bind: workspace too small"); return QSFT_ERR_ARG; }
    char* base = static_cast<char*>(ws);
    qsft_peel_desc d{};
    d.q = q; d.n = n; d.b = b; d.C = C; d.P = P; d.Q = (int)Qll;
    d.n_pad = n_pad; d.p1 = p1; d.method = method; d.max_iter = max_iter;
    d.hcap_log2 = hl2; d.out_cap = out_cap;
    d.cutoff = cutoff;
    double cap = 1.0;
    for (int i = 0; i < n && cap < 1e300; ++i) cap *= q;
    d.peel_cap = cap;
    d.Mt = Mt; d.D = D; d.U = U;
    d.kbin = reinterpret_cast<int8_t*>(base + lay.off[0]);
    d.rho_bin = reinterpret_cast<double*>(base + lay.off[1]);
    d.slot_bin = reinterpret_cast<int32_t*>(base + lay.off[2]);
    d.tag = reinterpret_cast<uint64_t*>(base + lay.off[3]);
    d.owner = reinterpret_cast<uint64_t*>(base + lay.off[4]);
    d.last = reinterpret_cast<uint64_t*>(base + lay.off[5]);
    d.head = reinterpret_cast<uint32_t*>(base + lay.off[6]);
    d.count = reinterpret_cast<uint32_t*>(base + lay.off[7]);
    d.key = reinterpret_cast<int8_t*>(base + lay.off[8]);
    d.rec_time = reinterpret_cast<uint64_t*>(base + lay.off[9]);
    d.rec_rho = reinterpret_cast<double*>(base + lay.off[10]);
    d.rec_next = reinterpret_cast<uint32_t*>(base + lay.off[11]);
    d.ctr = reinterpret_cast<uint32_t*>(base + lay.off[12]);
    d.out_k = reinterpret_cast<int8_t*>(base + lay.off[13]);
    d.out_val = reinterpret_cast<double*>(base + lay.off[14]);
    d.out_order = reinterpret_cast<uint64_t*>(base + lay.off[15]);
    d.mark = reinterpret_cast<int32_t*>(base + lay.off[16]);
    d.sstamp = reinterpret_cast<int32_t*>(base + lay.off[17]);
    d.pend_head = reinterpret_cast<int32_t*>(base + lay.off[18]);
    d.pend_next = reinterpret_cast<int32_t*>(base + lay.off[19]);
    d.live = reinterpret_cast<uint32_t*>(base + lay.off[20]);
    d.written = reinterpret_cast<uint8_t*>(base + lay.off[21]);
    d.noise_key = reinterpret_cast<uint64_t*>(base + lay.off[22]);
    d.xsrc = nullptr;
    d.noise_sigma = 0.0;
    d.flags = 0;
    d.energy = nullptr;
    d.rs_tab = nullptr;
    d.rs_t = d.rs_s = d.rs_nc = 0;
    *h = d;
    return QSFT_OK;
}

extern "C" int qsft_peel_set_coded(qsft_peel_desc* h, int t, int s, const int16_t* tables) {
    if (!h || !tables || !(h->method & QSFT_DETECT_CODED) || t < 1 || s < 1 || h->p1 - 1 != 2 * t * s ||
        2 * t > kRsMaxT2) {
        qsft_set_error("qsft_peel_set_coded: bad arguments (coded method, p1 - 1 = 2 t s, 2 t <= 32)");
        return QSFT_ERR_ARG;
    }
    long long order = 1;
    for (int i = 0; i < s; ++i) order *= h->q;
    if (order > 32768 || order - 1 < (long long)h->n) {
        qsft_set_error("qsft_peel_set_coded: q^s out of range (n <= q^s - 1 <= 32767)");
        return QSFT_ERR_ARG;
    }
    h->rs_tab = tables;
    h->rs_t = t;
    h->rs_s = s;
    h->rs_nc = (int)(order - 1);
    return QSFT_OK;
}

extern "C" int qsft_peel_bin_blocks(const qsft_peel_desc* h) {
    if (!h) return -1;
    const long long CQ = (long long)h->C * h->Q;
    const long long per = (long long)kWaves * kBinsPerWave;
    return (int)((CQ + per - 1) / per);
}

extern "C" int qsft_peel_slot_blocks(const qsft_peel_desc* h) {
    if (!h) return -1;
    const long long H = 1ll << h->hcap_log2;
    const long long CQ = (long long)h->C * h->Q;
    return (int)(((H > CQ ? H : CQ) + kSlotThreads - 1) / kSlotThreads);
}

extern "C" int qsft_peel_init(const qsft_peel_desc* d, const int32_t* pre_bins, const int32_t* pre_slots,
                              int n_inst, int blocks_bins, int blocks_slots, void* stream) {
    if (!d || !pre_bins || !pre_slots || n_inst < 1 || blocks_bins < 1 || blocks_slots < 1) {
        qsft_set_error("qsft_peel_init: bad arguments");
        return QSFT_ERR_ARG;
    }
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_peel_init, dim3(blocks_slots), dim3(kSlotThreads), 0, st, d, pre_slots, n_inst);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return qsft_hip_status(e, "qsft_peel_init");
    hipLaunchKernelGGL(k_peel_screen, dim3(blocks_bins), dim3(kBinThreads), 0, st, d, pre_bins, n_inst);
    return qsft_hip_status(hipGetLastError(), "qsft_peel_init(screen)");
}

extern "C" int qsft_peel_iteration(const qsft_peel_desc* d, const int32_t* pre_bins, int n_inst, int blocks_bins,
                                   int max_p, int it, void* stream) {
    if (!d || !pre_bins || n_inst < 1 || blocks_bins < 1 || it < 0 || it >= 15 || max_p < 1 ||
        max_p > 64 * kRowsPerLane) {
        qsft_set_error("qsft_peel_iteration: bad arguments");
        return QSFT_ERR_ARG;
    }
    hipStream_t st = (hipStream_t)stream;
    // rows per lane: 2 when every instance has P <= 128, else 4; the first
    // pass of the former four bins per wave where it can
    if (max_p <= 128 && it == 0)
        hipLaunchKernelGGL(k_peel_detect0, dim3(blocks_bins), dim3(kBinThreads), 0, st, d, pre_bins, n_inst);
    else if (max_p <= 128)
        hipLaunchKernelGGL(k_peel_detect<2>, dim3(blocks_bins), dim3(kBinThreads), 0, st, d, pre_bins, n_inst, it);
    else
        hipLaunchKernelGGL(k_peel_detect<4>, dim3(blocks_bins), dim3(kBinThreads), 0, st, d, pre_bins, n_inst, it);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return qsft_hip_status(e, "qsft_peel_iteration(detect)");
    hipLaunchKernelGGL(k_peel_apply, dim3(blocks_bins), dim3(kBinThreads), 0, st, d, pre_bins, n_inst, it);
    return qsft_hip_status(hipGetLastError(), "qsft_peel_iteration(apply)");
}

extern "C" int qsft_peel_finalize(const qsft_peel_desc* d, const int32_t* pre_slots, int n_inst, int blocks_slots,
                                  int word, void* stream) {
    if (!d || !pre_slots || n_inst < 1 || blocks_slots < 1 || (word & ~QSFT_FIN_FLUSH)) {
        qsft_set_error("qsft_peel_finalize: bad arguments");
        return QSFT_ERR_ARG;
    }
    if (word & QSFT_FIN_FLUSH) {
        hipLaunchKernelGGL(k_peel_flush, dim3(blocks_slots), dim3(kSlotThreads), 0, (hipStream_t)stream, d, pre_slots,
                           n_inst);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return qsft_hip_status(e, "qsft_peel_finalize(flush)");
    }
    hipLaunchKernelGGL(k_peel_finalize, dim3(blocks_slots), dim3(kSlotThreads), 0, (hipStream_t)stream, d, pre_slots, n_inst);
    return qsft_hip_status(hipGetLastError(), "qsft_peel_finalize");
}

extern "C" int qsft_peel_passes(const qsft_peel_desc* d, int n_inst, int max_p, int it0, int it_end, void* stream) {
    if (!d || n_inst < 1 || it0 < 0 || it_end > 15 || it0 > it_end || max_p < 1 || max_p > 64 * kRowsPerLane) {
        qsft_set_error("qsft_peel_passes: bad arguments");
        return QSFT_ERR_ARG;
    }
    if (it0 == it_end) return QSFT_OK;
    hipStream_t st = (hipStream_t)stream;
    if (max_p <= 128)
        hipLaunchKernelGGL(k_peel_tail<2>, dim3(n_inst), dim3(64 * tail_waves<2>()), 0, st, d, it0, it_end);
    else
        hipLaunchKernelGGL(k_peel_tail<4>, dim3(n_inst), dim3(64 * tail_waves<4>()), 0, st, d, it0, it_end);
    return qsft_hip_status(hipGetLastError(), "qsft_peel_passes");
}
extern "C" int qsft_peel_run(const qsft_peel_desc* d, const int32_t* pre_bins, const int32_t* pre_slots, int n_inst,
                             int blocks_bins, int blocks_slots, int max_p, int max_iter, void* stream) {
    if (max_iter < 1 || max_iter > 15) { qsft_set_error("qsft_peel_run: max_iter must be in [1, 15]"); return QSFT_ERR_ARG; }
    int rc = qsft_peel_init(d, pre_bins, pre_slots, n_inst, blocks_bins, blocks_slots, stream);
    for (int it = 0; rc == QSFT_OK && it < max_iter; ++it)
        rc = qsft_peel_iteration(d, pre_bins, n_inst, blocks_bins, max_p, it, stream);
    if (rc == QSFT_OK) rc = qsft_peel_finalize(d, pre_slots, n_inst, blocks_slots, QSFT_FIN_FLUSH, stream);
    return rc;
}

extern "C" size_t qsft_peel_pack_header_bytes(int n_inst) {
    return n_inst < 1 ? 0 : (((size_t)(2 * n_inst + 4) * 8 + 255) & ~(size_t)255);
}

extern "C" int qsft_peel_pack(const qsft_peel_desc* d, int n_inst, void* dst, int64_t cap_rows, int row_stride,
                              void* stream) {
    if (!d || !dst || n_inst < 1 || cap_rows < 0 || row_stride < 24 + 64 || row_stride % 8) {
        qsft_set_error("qsft_peel_pack: bad arguments");
        return QSFT_ERR_ARG;
    }
    hipStream_t st = (hipStream_t)stream;
    int64_t* hdr = static_cast<int64_t*>(dst);
    unsigned char* rows = static_cast<unsigned char*>(dst) + qsft_peel_pack_header_bytes(n_inst);
    hipLaunchKernelGGL(k_pack_scan, dim3(1), dim3(64), 0, st, d, n_inst, hdr, cap_rows);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return qsft_hip_status(e, "qsft_peel_pack(scan)");
    hipLaunchKernelGGL(k_pack_rows, dim3(n_inst), dim3(256), 0, st, d, n_inst, hdr, rows, cap_rows, row_stride);
    return qsft_hip_status(hipGetLastError(), "qsft_peel_pack(rows)");
}

extern "C" int qsft_peel_pack_ordered(const qsft_peel_desc* d, int n_inst, void* dst, int64_t cap_rows,
                                      int row_stride, void* stream) {
    if (!d || !dst || n_inst < 1 || cap_rows < 0 || row_stride < 24 + 64 || row_stride % 8) {
        qsft_set_error("qsft_peel_pack_ordered: bad arguments");
        return QSFT_ERR_ARG;
    }
    hipStream_t st = (hipStream_t)stream;
    int64_t* hdr = static_cast<int64_t*>(dst);
    unsigned char* rows = static_cast<unsigned char*>(dst) + qsft_peel_pack_header_bytes(n_inst);
    hipLaunchKernelGGL(k_pack_scan, dim3(1), dim3(64), 0, st, d, n_inst, hdr, cap_rows);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return qsft_hip_status(e, "qsft_peel_pack_ordered(scan)");
    hipLaunchKernelGGL(k_pack_rows_ordered, dim3(n_inst * kRankGroups), dim3(256), 0, st, d, n_inst, hdr, rows,
                       cap_rows, row_stride);
    return qsft_hip_status(hipGetLastError(), "qsft_peel_pack_ordered(rows)");
}
