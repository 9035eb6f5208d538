// Stage (d): singleton detection and peeling, device resident (gfx950).
//
// Semantics follow QSFT.transform (qsft/qsft.py:142-255) and
// singleton_detection (qsft/reconstruct.py:12-31, 100-113):
//   * every pass scans all bins (c, j) of a frozen U (Jacobi): zeroton if
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
// Layout: U is [C][P][Q] complex FP64 (row-major, the DFT's output); one
// thread owns one bin and walks its P-column with stride Q, so a wave reads
// 64 consecutive bins of a row per access (1 KiB, coalesced).
//
// Kernel A (detect) registers each singleton in an open-addressing hash
// table keyed by a 64-bit hash ("tag") of k: CAS claims a slot, atomicMin /
// atomicMax on (pass+1) << 32 | bin record the first and the last detection,
// and a record (time, rho) is pushed on the slot's lock-free list. Kernel B
// (peel) checks that every detection's k equals its slot's key (a 64-bit tag
// collision sets an error bit instead of merging two keys), writes the key
// of new slots, and lets the last detection of each key peel it with FP64
// atomics. No kernel waits on another thread's progress.
#include "qsft_common.h"

namespace qsft {

constexpr int kPeelThreads = 128;
constexpr int kSlotThreads = 256;
constexpr uint32_t kNil = 0xFFFFFFFFu;

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

// <row, k> over n_pad digits; k lives in LDS as words ks[w * kPeelThreads + tid].
__device__ __forceinline__ uint32_t dot_k(const int8_t* __restrict__ row, const uint32_t* ks, int nw) {
    const int tid = threadIdx.x;
    int acc = 0;
    const uint32_t* r32 = reinterpret_cast<const uint32_t*>(row);
    for (int w = 0; w < nw; ++w) acc = __builtin_amdgcn_sdot4((int)r32[w], (int)ks[w * kPeelThreads + tid], acc, false);
    return (uint32_t)acc;
}

__device__ __forceinline__ void set_digit(uint32_t* ks, int i, uint32_t v) {
    const int tid = threadIdx.x;
    uint32_t* w = ks + (i >> 2) * kPeelThreads + tid;
    const int sh = 8 * (i & 3);
    *w = (*w & ~(0xFFu << sh)) | ((v & 0xFFu) << sh);
}

__device__ __forceinline__ uint64_t hash_k(const uint32_t* ks, int nw) {
    const int tid = threadIdx.x;
    uint64_t h = 0xCBF29CE484222325ull;
    for (int w = 0; w < nw; ++w) {
        h ^= ks[w * kPeelThreads + tid];
        h *= 0x100000001B3ull;
    }
    return mix64(h);
}

__device__ __forceinline__ void fill_wtable_d(double2* tab, int q) {
    for (int r = threadIdx.x; r < q; r += blockDim.x) {
        double s, c;
        sincospi(2.0 * (double)r / (double)q, &s, &c);
        tab[r] = make_double2(c, s);
    }
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
}

__global__ __launch_bounds__(kPeelThreads) void k_peel_detect(const qsft_peel_desc* __restrict__ descs,
                                                              const int32_t* __restrict__ pre, int n_inst,
                                                              int it) {
    __shared__ double2 tab[QSFT_MAX_Q];
    __shared__ uint32_t ks[(QSFT_MAX_N / 4) * kPeelThreads];
    const int inst = find_inst(pre, n_inst, blockIdx.x);
    const qsft_peel_desc& d = descs[inst];
    uint32_t* __restrict__ ctr = d.ctr;
    if (it >= d.max_iter) return;                                  // beyond the bound workspace
    if (ctr[it * QSFT_CTR_ITER_STRIDE + 3] == 0) return;          // loop already stopped
    const int q = d.q, n = d.n, b = d.b, C = d.C, P = d.P, Q = d.Q, n_pad = d.n_pad;
    const int nw = n_pad >> 2;
    const int tid = threadIdx.x;
    const int t = (blockIdx.x - pre[inst]) * kPeelThreads + tid;
    fill_wtable_d(tab, q);
    for (int w = 0; w < nw; ++w) ks[w * kPeelThreads + tid] = 0;
    __syncthreads();
    if (t >= C * Q) return;

    const ModQ mq((uint32_t)q);
    const int c = t / Q, j = t - c * Q;
    const double2* __restrict__ col = reinterpret_cast<const double2*>(d.U) + (size_t)c * P * Q + j;
    const double thr = d.cutoff * (double)P;

    double sre = 0.0, sim = 0.0;
    for (int p = 0; p < P; ++p) {
        const double2 u = col[(size_t)p * Q];
        sre += u.x * u.x;
        sim += u.y * u.y;
    }
    double energy = sqrt(sre + sim);
    energy = energy * energy;
    int status = 0;                                  // 0 zeroton, 1 singleton, 2 multiton
    double rho_x = 0.0, rho_y = 0.0;
    if (energy > thr) {
        const double twopi = 6.283185307179586;
        if (d.method == QSFT_DETECT_NSO1) {
            const int p1 = d.p1;
            const int R = P / p1;
            const double step = twopi / (double)q;   // 2 * np.pi / q
            for (int i = 1; i < p1; ++i) {
                double zr = 0.0, zi = 0.0;
                for (int r = 0; r < R; ++r) {
                    const double2 u0 = col[(size_t)(r * p1) * Q];
                    const double2 ui = col[(size_t)(r * p1 + i) * Q];
                    // u0 * conj(ui), as numpy forms the complex product
                    const double pr = __dsub_rn(__dmul_rn(u0.x, ui.x), __dmul_rn(u0.y, -ui.y));
                    const double pi_ = __dadd_rn(__dmul_rn(u0.x, -ui.y), __dmul_rn(u0.y, ui.x));
                    zr = __dadd_rn(zr, pr);
                    zi = __dadd_rn(zi, pi_);
                }
                zr = __ddiv_rn(zr, (double)R);
                zi = __ddiv_rn(zi, (double)R);
                const double th = rem_2pi(atan2(zi, zr));
                int best = 0;
                double bestv = fabs(__dsub_rn(0.0, th));
                for (int tt = 1; tt <= q; ++tt) {
                    const double v = fabs(__dsub_rn(__dmul_rn(step, (double)tt), th));
                    if (v < bestv) { bestv = v; best = tt; }
                }
                set_digit(ks, i - 1, (uint32_t)(best % q));
            }
        } else {
            const double2 v0 = col[0];
            const double a0 = atan2(v0.y, v0.x);
            for (int i = 1; i <= n; ++i) {
                const double2 vi = col[(size_t)i * Q];
                const double ai = atan2(vi.y, vi.x);
                const double v = __ddiv_rn(__dmul_rn((double)q, __dsub_rn(ai, a0)), twopi);
                long long kk = (long long)rint(v);
                kk %= q;
                if (kk < 0) kk += q;
                set_digit(ks, i - 1, (uint32_t)kk);
            }
        }
        // signature, rho
        const int8_t* __restrict__ Dc = d.D + (size_t)c * P * n_pad;
        double rr = 0.0, ri = 0.0;
        for (int p = 0; p < P; ++p) {
            const double2 w = tab[mq(dot_k(Dc + (size_t)p * n_pad, ks, nw))];
            const double2 u = col[(size_t)p * Q];
            rr += w.x * u.x + w.y * u.y;              // conj(w) * u
            ri += w.x * u.y - w.y * u.x;
        }
        rho_x = __ddiv_rn(rr, (double)P);
        rho_y = __ddiv_rn(ri, (double)P);
        double qre = 0.0, qim = 0.0;
        for (int p = 0; p < P; ++p) {
            const double2 w = tab[mq(dot_k(Dc + (size_t)p * n_pad, ks, nw))];
            const double2 u = col[(size_t)p * Q];
            const double sx = __dsub_rn(__dmul_rn(rho_x, w.x), __dmul_rn(rho_y, w.y));
            const double sy = __dadd_rn(__dmul_rn(rho_x, w.y), __dmul_rn(rho_y, w.x));
            const double ex = __dsub_rn(u.x, sx), ey = __dsub_rn(u.y, sy);
            qre += ex * ex;
            qim += ey * ey;
        }
        double res = sqrt(qre + qim);
        res = res * res;
        // bin match: M_c^T k mod q == digits of j (MSB first)
        const int8_t* __restrict__ Mc = d.Mt + (size_t)c * b * n_pad;
        bool match = true;
        int rem = j, pw = Q;
        for (int tt = 0; tt < b; ++tt) {
            pw /= q;
            const int dig = rem / pw;
            rem -= dig * pw;
            match = match && ((int)mq(dot_k(Mc + (size_t)tt * n_pad, ks, nw)) == dig);
        }
        status = (match && res <= thr) ? 1 : 2;
    }

    int32_t* slot_bin = d.slot_bin;
    if (status == 2) {
        atomicAdd(&ctr[it * QSFT_CTR_ITER_STRIDE + 1], 1u);
        slot_bin[t] = -1;
        return;
    }
    if (status == 0) {
        slot_bin[t] = -1;
        return;
    }
    atomicAdd(&ctr[it * QSFT_CTR_ITER_STRIDE + 0], 1u);
    // record k and rho of this singleton bin
    uint32_t* kdst = reinterpret_cast<uint32_t*>(d.kbin + (size_t)t * n_pad);
    for (int w = 0; w < nw; ++w) kdst[w] = ks[w * kPeelThreads + tid];
    d.rho_bin[2 * (size_t)t] = rho_x;
    d.rho_bin[2 * (size_t)t + 1] = rho_y;
    // claim or find the slot of k
    const uint64_t hv = hash_k(ks, nw);
    const uint64_t tg = hv | 1ull;
    const uint32_t mask = (1u << d.hcap_log2) - 1u;
    uint32_t pos = (uint32_t)(hv >> 32) & mask;
    int slot = -1;
    for (uint32_t probe = 0; probe <= mask; ++probe) {
        const unsigned long long old = atomicCAS((unsigned long long*)&d.tag[pos], 0ull, (unsigned long long)tg);
        if (old == 0ull || old == tg) { slot = (int)pos; break; }
        pos = (pos + 1) & mask;
    }
    if (slot < 0) {
        atomicOr(&ctr[QSFT_CTR_ERROR], (uint32_t)QSFT_DEVERR_TABLE_FULL);
        slot_bin[t] = -1;
        return;
    }
    slot_bin[t] = slot;
    const uint64_t tm = ((uint64_t)(it + 1) << 32) | (uint32_t)t;
    atomicMin((unsigned long long*)&d.owner[slot], (unsigned long long)tm);
    const unsigned long long prev = atomicMax((unsigned long long*)&d.last[slot], (unsigned long long)tm);
    if ((prev >> 32) != (unsigned long long)(it + 1)) atomicAdd(&ctr[it * QSFT_CTR_ITER_STRIDE + 2], 1u);
    const uint32_t r = atomicAdd(&ctr[QSFT_CTR_LOG], 1u);
    if ((uint64_t)r >= (uint64_t)d.max_iter * (uint64_t)(C * Q)) {   // cannot happen: one record per bin per pass
        atomicOr(&ctr[QSFT_CTR_ERROR], (uint32_t)QSFT_DEVERR_BAD_SHAPE);
        return;
    }
    d.rec_time[r] = tm;
    d.rec_rho[2 * (size_t)r] = rho_x;
    d.rec_rho[2 * (size_t)r + 1] = rho_y;
    d.rec_next[r] = atomicExch(&d.head[slot], r);
    atomicAdd(&d.count[slot], 1u);
}

__global__ __launch_bounds__(kPeelThreads) void k_peel_apply(const qsft_peel_desc* __restrict__ descs,
                                                             const int32_t* __restrict__ pre, int n_inst,
                                                             int it) {
    __shared__ double2 tab[QSFT_MAX_Q];
    __shared__ uint32_t ks[(QSFT_MAX_N / 4) * kPeelThreads];
    const int inst = find_inst(pre, n_inst, blockIdx.x);
    const qsft_peel_desc& d = descs[inst];
    uint32_t* __restrict__ ctr = d.ctr;
    const int local = blockIdx.x - pre[inst];
    if (it >= d.max_iter) return;
    const uint32_t active = ctr[it * QSFT_CTR_ITER_STRIDE + 3];
    if (local == 0 && threadIdx.x == 0) {
        // loop control for the next pass (the counters of this pass are final)
        const uint32_t ns = ctr[it * QSFT_CTR_ITER_STRIDE + 0];
        const uint32_t nm = ctr[it * QSFT_CTR_ITER_STRIDE + 1];
        const uint32_t nb = ctr[it * QSFT_CTR_ITER_STRIDE + 2];
        uint32_t peeled = ctr[QSFT_CTR_PEELED];
        if (active) {
            ctr[QSFT_CTR_ITERS] = (uint32_t)(it + 1);
            peeled = (peeled + nb < peeled) ? 0xFFFFFFFFu : peeled + nb;
            ctr[QSFT_CTR_PEELED] = peeled;
        }
        const uint32_t next = (active && ns > 0 && nm > 0 && (double)peeled < d.peel_cap) ? 1u : 0u;
        if (it + 1 < 16) ctr[(it + 1) * QSFT_CTR_ITER_STRIDE + 3] = next;
    }
    if (!active) return;
    const int q = d.q, b = d.b, C = d.C, P = d.P, Q = d.Q, n_pad = d.n_pad;
    const int nw = n_pad >> 2;
    const int tid = threadIdx.x;
    const int t = local * kPeelThreads + tid;
    fill_wtable_d(tab, q);
    __syncthreads();
    if (t >= C * Q) return;
    const int slot = d.slot_bin[t];
    if (slot < 0) return;
    const uint64_t tm = ((uint64_t)(it + 1) << 32) | (uint32_t)t;
    // bring k into LDS and check it against the slot's key
    const uint32_t* ksrc = reinterpret_cast<const uint32_t*>(d.kbin + (size_t)t * n_pad);
    for (int w = 0; w < nw; ++w) ks[w * kPeelThreads + tid] = ksrc[w];
    const uint64_t own = d.owner[slot];
    uint32_t* kslot = reinterpret_cast<uint32_t*>(d.key + (size_t)slot * n_pad);
    if ((own >> 32) == (uint64_t)(it + 1)) {
        const uint32_t ot = (uint32_t)own;
        if (own == tm) {
            for (int w = 0; w < nw; ++w) kslot[w] = ks[w * kPeelThreads + tid];
        } else {
            const uint32_t* kown = reinterpret_cast<const uint32_t*>(d.kbin + (size_t)ot * n_pad);
            for (int w = 0; w < nw; ++w)
                if (kown[w] != ks[w * kPeelThreads + tid]) atomicOr(&ctr[QSFT_CTR_ERROR], (uint32_t)QSFT_DEVERR_TAG_COLLISION);
        }
    } else {
        for (int w = 0; w < nw; ++w)
            if (kslot[w] != ks[w * kPeelThreads + tid]) atomicOr(&ctr[QSFT_CTR_ERROR], (uint32_t)QSFT_DEVERR_TAG_COLLISION);
    }
    if (d.last[slot] != tm) return;
    // the last detection of this key in this pass peels it from every group
    const ModQ mq((uint32_t)q);
    const double rx = d.rho_bin[2 * (size_t)t], ry = d.rho_bin[2 * (size_t)t + 1];
    for (int l = 0; l < C; ++l) {
        const int8_t* __restrict__ Ml = d.Mt + (size_t)l * b * n_pad;
        int jl = 0;
        for (int tt = 0; tt < b; ++tt) jl = jl * q + (int)mq(dot_k(Ml + (size_t)tt * n_pad, ks, nw));
        const int8_t* __restrict__ Dl = d.D + (size_t)l * P * n_pad;
        double* __restrict__ Ul = d.U + 2 * ((size_t)l * P * Q + jl);
        for (int p = 0; p < P; ++p) {
            const double2 w = tab[mq(dot_k(Dl + (size_t)p * n_pad, ks, nw))];
            const double sx = __dsub_rn(__dmul_rn(rx, w.x), __dmul_rn(ry, w.y));
            const double sy = __dadd_rn(__dmul_rn(rx, w.y), __dmul_rn(ry, w.x));
            unsafeAtomicAdd(Ul + 2 * (size_t)p * Q, -sx);
            unsafeAtomicAdd(Ul + 2 * (size_t)p * Q + 1, -sy);
        }
    }
}

__global__ __launch_bounds__(kSlotThreads) void k_peel_finalize(const qsft_peel_desc* __restrict__ descs,
                                                                const int32_t* __restrict__ pre, int n_inst) {
    const int inst = find_inst(pre, n_inst, blockIdx.x);
    const qsft_peel_desc& d = descs[inst];
    const int H = 1 << d.hcap_log2;
    const int h = (blockIdx.x - pre[inst]) * kSlotThreads + threadIdx.x;
    if (h >= H || d.tag[h] == 0) return;
    const uint32_t m = d.count[h];
    if (m == 0) return;
    const uint32_t L = d.ctr[QSFT_CTR_LOG];
    // running mean over the records in time order: g <- (g * c + v) / (c + 1)
    uint64_t prev = 0;
    double gx = 0.0, gy = 0.0;
    for (uint32_t cnt = 0; cnt < m; ++cnt) {
        uint64_t best = ~0ull;
        uint32_t bi = kNil;
        uint32_t steps = 0;
        for (uint32_t r = d.head[h]; r != kNil && steps <= L; r = d.rec_next[r], ++steps) {
            const uint64_t tmr = d.rec_time[r];
            if (tmr > prev && tmr < best) { best = tmr; bi = r; }
        }
        if (bi == kNil) break;
        const double vx = d.rec_rho[2 * (size_t)bi], vy = d.rec_rho[2 * (size_t)bi + 1];
        if (cnt == 0) {
            gx = vx;
            gy = vy;
        } else {
            const double cd = (double)cnt;
            gx = __ddiv_rn(__dadd_rn(__dmul_rn(gx, cd), vx), cd + 1.0);
            gy = __ddiv_rn(__dadd_rn(__dmul_rn(gy, cd), vy), cd + 1.0);
        }
        prev = best;
    }
    const uint32_t o = atomicAdd(&d.ctr[QSFT_CTR_OUT], 1u);
    if ((int)o >= d.out_cap) {
        atomicOr(&d.ctr[QSFT_CTR_ERROR], (uint32_t)QSFT_DEVERR_OUT_FULL);
        return;
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
// hdr[i] = count_i, hdr[n + i] = offset_i, hdr[2n] = total, hdr[2n + 1] = OR of error bits.
__global__ __launch_bounds__(64) void k_pack_scan(const qsft_peel_desc* __restrict__ descs, int n_inst,
                                                  int64_t* __restrict__ hdr, int64_t cap_rows) {
    if (threadIdx.x != 0) return;
    int64_t off = 0;
    uint32_t err = 0;
    for (int i = 0; i < n_inst; ++i) {
        const qsft_peel_desc& d = descs[i];
        uint32_t m = d.ctr[QSFT_CTR_OUT];
        if ((int)m > d.out_cap) m = (uint32_t)d.out_cap;
        err |= d.ctr[QSFT_CTR_ERROR];
        hdr[i] = m;
        hdr[n_inst + i] = off;
        off += m;
    }
    if (off > cap_rows) err |= QSFT_DEVERR_OUT_FULL;
    hdr[2 * n_inst] = off;
    hdr[2 * n_inst + 1] = err;
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
    size_t off[20];
    size_t total;
};
PeelLayout peel_layout(long long CQ, int n_pad, int hl2, int max_iter, int out_cap) {
    const size_t H = (size_t)1 << hl2;
    const size_t L = (size_t)max_iter * CQ;
    const size_t sizes[] = {
        (size_t)CQ * n_pad,      // kbin
        (size_t)CQ * 16,         // rho_bin
        (size_t)CQ * 4,          // slot_bin
        H * 8, H * 8, H * 8,     // tag owner last
        H * 4, H * 4,            // head count
        H * (size_t)n_pad,       // key
        L * 8, L * 16, L * 4,    // rec_time rec_rho rec_next
        QSFT_CTR_WORDS * 4,      // ctr
        (size_t)out_cap * n_pad, (size_t)out_cap * 16, (size_t)out_cap * 8,  // outputs
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
    return peel_layout(CQ, n_pad, hl2, max_iter, out_cap > 0 ? out_cap : (1 << hl2)).total;
}

extern "C" int qsft_peel_bind(qsft_peel_desc* h, int q, int n, int b, int C, int P, int p1, int method,
                              int max_iter, double cutoff, int out_cap, const int8_t* Mt, const int8_t* D,
                              double* U, void* ws, size_t ws_bytes, int n_pad) {
    if (!h || !ws || q < 2 || q > QSFT_MAX_Q || n < 1 || n > QSFT_MAX_N || b < 1 || b > QSFT_MAX_B || C < 1 ||
        C > QSFT_MAX_C || P < 1 || p1 < 2 || P % p1 || max_iter < 1 || max_iter > 15 || n_pad % 64 ||
        n_pad < n || n_pad > QSFT_MAX_N || (method != QSFT_DETECT_NSO1 && method != QSFT_DETECT_NOISELESS) ||
        (method == QSFT_DETECT_NOISELESS && P != n + 1) || (method == QSFT_DETECT_NSO1 && p1 != n + 1)) {
        qsft_set_error("qsft_peel_bind: bad shape");
        return QSFT_ERR_ARG;
    }
    const long long Qll = ipow(q, b);
    const long long CQ = (long long)C * Qll;
    if (CQ > 0x3FFFFFFF) { qsft_set_error("qsft_peel_bind: too many bins"); return QSFT_ERR_ARG; }
    const int hl2 = hcap_log2_for(CQ);
    if (out_cap <= 0) out_cap = 1 << hl2;
    const PeelLayout lay = peel_layout(CQ, n_pad, hl2, max_iter, out_cap);
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
    *h = d;
    return QSFT_OK;
}

extern "C" int qsft_peel_bin_blocks(const qsft_peel_desc* h) {
    if (!h) return -1;
    const long long CQ = (long long)h->C * h->Q;
    return (int)((CQ + kPeelThreads - 1) / kPeelThreads);
}

extern "C" int qsft_peel_slot_blocks(const qsft_peel_desc* h) {
    if (!h) return -1;
    return (int)(((1ll << h->hcap_log2) + kSlotThreads - 1) / kSlotThreads);
}

extern "C" int qsft_peel_init(const qsft_peel_desc* d, const int32_t* pre_slots, int n_inst, int blocks_slots,
                              void* stream) {
    if (!d || !pre_slots || n_inst < 1 || blocks_slots < 1) { qsft_set_error("qsft_peel_init: bad arguments"); return QSFT_ERR_ARG; }
    hipLaunchKernelGGL(k_peel_init, dim3(blocks_slots), dim3(kSlotThreads), 0, (hipStream_t)stream, d, pre_slots, n_inst);
    return qsft_hip_status(hipGetLastError(), "qsft_peel_init");
}

extern "C" int qsft_peel_iteration(const qsft_peel_desc* d, const int32_t* pre_bins, int n_inst, int blocks_bins,
                                   int it, void* stream) {
    if (!d || !pre_bins || n_inst < 1 || blocks_bins < 1 || it < 0 || it >= 15) {
        qsft_set_error("qsft_peel_iteration: bad arguments");
        return QSFT_ERR_ARG;
    }
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_peel_detect, dim3(blocks_bins), dim3(kPeelThreads), 0, st, d, pre_bins, n_inst, it);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return qsft_hip_status(e, "qsft_peel_iteration(detect)");
    hipLaunchKernelGGL(k_peel_apply, dim3(blocks_bins), dim3(kPeelThreads), 0, st, d, pre_bins, n_inst, it);
    return qsft_hip_status(hipGetLastError(), "qsft_peel_iteration(apply)");
}

extern "C" int qsft_peel_finalize(const qsft_peel_desc* d, const int32_t* pre_slots, int n_inst, int blocks_slots,
                                  void* stream) {
    if (!d || !pre_slots || n_inst < 1 || blocks_slots < 1) { qsft_set_error("qsft_peel_finalize: bad arguments"); return QSFT_ERR_ARG; }
    hipLaunchKernelGGL(k_peel_finalize, dim3(blocks_slots), dim3(kSlotThreads), 0, (hipStream_t)stream, d, pre_slots, n_inst);
    return qsft_hip_status(hipGetLastError(), "qsft_peel_finalize");
}

extern "C" int qsft_peel_run(const qsft_peel_desc* d, const int32_t* pre_bins, const int32_t* pre_slots, int n_inst,
                             int blocks_bins, int blocks_slots, int max_iter, void* stream) {
    if (max_iter < 1 || max_iter > 15) { qsft_set_error("qsft_peel_run: max_iter must be in [1, 15]"); return QSFT_ERR_ARG; }
    int rc = qsft_peel_init(d, pre_slots, n_inst, blocks_slots, stream);
    for (int it = 0; rc == QSFT_OK && it < max_iter; ++it) rc = qsft_peel_iteration(d, pre_bins, n_inst, blocks_bins, it, stream);
    if (rc == QSFT_OK) rc = qsft_peel_finalize(d, pre_slots, n_inst, blocks_slots, stream);
    return rc;
}

extern "C" size_t qsft_peel_pack_header_bytes(int n_inst) {
    return n_inst < 1 ? 0 : (((size_t)(2 * n_inst + 2) * 8 + 255) & ~(size_t)255);
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
