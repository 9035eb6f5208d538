// Stage (a)+(b) on the matrix cores: coset evaluation as a complex GEMM.
//
// Every query of group c is m = (M_c l + d_p) mod q, so its phase against a
// support vector splits exactly (mod q is a ring homomorphism):
//   <m, k_s> = <(M_c l) mod q, k_s> + <d_p, k_s>   (mod q)
// and the sample is
//   x[c][p][l] = sum_s W_c[l][s] B_c[s][p],
//   W_c[l][s] = w^{E0},  E0 = <(M_c l) mod q, k_s> mod q,
//   B_c[s][p] = a_s w^{<d_p, k_s> mod q}.
// Each sample is still the full S-term sum of a_s w^{<m, k_s>}; the B*S
// complex multiply-adds run as a (Q x S) x (S x P) complex GEMM on
// v_mfma_f32_16x16x32_f16, and the w-table epilogue runs C*Q*S times
// instead of C*P*Q*S times.
//
//   * E0 comes from v_mfma_i32_16x16x64_i8 (support digits x base-query digits,
//     exact int32); its D fragment (lane = query column l, 4 rows = supports
//     4g..4g+3) is exactly the f16 A fragment needed next: slots 0..3 hold
//     Re W for those 4 supports and slots 4..7 Im W.
//   * B is precomputed (prep kernel) in FP64, scaled by 2^bexp so that its
//     largest entry lies in [2^13, 2^14), split into two fp16 limbs (hi + lo:
//     the lo limb carries the rounding error of hi, ~22 significant bits
//     relative to max|B| and never overflowing), and stored already in the
//     B-fragment order of the f16 MFMA, with the complex product laid out as
//     a real GEMM: column 2p' holds Re x, column 2p'+1 holds Im x, and the
//     "Im W" k-slots carry -Im B / Re B. When all groups share their delay
//     rows (the reference's get_Ms_and_Ds), B is computed once.
//   * w^e is exact in fp16 for q in {2, 4} (0, +-1); other q split W into
//     two limbs too and keep the three limb products of order <= 1.
//   * accumulation is FP32 in the MFMA; x is scaled back by 2^-bexp (exact)
//     and written as complex FP32.
//
// A workgroup owns 256 base queries (4 waves x 4 subtiles of 16) x up to 8
// p-tiles (64 delay rows) of one group (5..8 for q in {2, 4} with n <= 64, split
// evenly over the group's tiles). Per 16-support block it stages the p-tiles' x 2
// B fragments (up to 16 KiB) in LDS and each wave reuses every fragment for its 4
// subtiles; the fragments of the next three blocks are loaded straight into
// LDS (a four-slot ring) while the current block's MFMAs run. Logical blocks are
// dealt out XCD by XCD so the workgroups that share B fragments share an L2.
#include <type_traits>

#include "qsft_common.h"

namespace qsft {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef unsigned int v4u __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef float v2f __attribute__((ext_vector_type(2)));
typedef _Float16 v8h __attribute__((ext_vector_type(8)));
typedef _Float16 v16h __attribute__((ext_vector_type(16)));
typedef unsigned int v8u __attribute__((ext_vector_type(8)));

constexpr int kMfThreads = 256;     // workgroup of the generic kernels (4 waves, 256 base queries)
constexpr int kPtPerBlock = 8;       // most p-tiles (8 complex delay rows each) per workgroup
constexpr int kLimbs = 2;            // fp16 limbs of B (and of W for q not in {2, 4})
constexpr int kStage = kPtPerBlock * kLimbs * 64;      // v4u per staged support block
constexpr int kXcds = 8;

// fp16 bits of f (round to nearest even)
__device__ __forceinline__ uint32_t f16_bits(float f) {
    return (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)f);
}
__device__ __forceinline__ double f16_to_double(uint32_t b) {
    return (double)__builtin_bit_cast(_Float16, (uint16_t)b);
}
// v = hi + lo + err, |err| <= 2^-11 |v - hi| <= 2^-22 |v| (for |v| < 2^15 and
// |v - hi| above the fp16 subnormal range; below it err is < 2^-25 absolute)
__device__ __forceinline__ void split2(double v, uint32_t (&l)[2]) {
    l[0] = f16_bits((float)v);
    l[1] = f16_bits((float)(v - f16_to_double(l[0])));
}

__device__ __forceinline__ v4i pack16b(const int (&d)[16]) {
    v4i r;
#pragma unroll
    for (int w = 0; w < 4; ++w)
        r[w] = (d[4 * w] & 0xFF) | ((d[4 * w + 1] & 0xFF) << 8) | ((d[4 * w + 2] & 0xFF) << 16) |
               ((d[4 * w + 3] & 0xFF) << 24);
    return r;
}

__device__ __forceinline__ int byte_at(const v4i& v, int j) { return (v[j >> 2] >> (8 * (j & 3))) & 0xFF; }

// Physical workgroup -> logical workgroup so that each XCD (workgroups are
// dealt round-robin to the 8 XCDs) runs a contiguous range of logical ids.
__device__ __forceinline__ int xcd_logical_block() {
    const int nb = gridDim.x, bid = blockIdx.x;
    const int xcd = bid % kXcds, slot = bid / kXcds;
    const int per = nb / kXcds, rem = nb % kXcds;
    return (xcd < rem) ? xcd * (per + 1) + slot : rem * (per + 1) + (xcd - rem) * per + slot;
}

// Prep: B_c[s][p] = 2^bexp a_s w^{<d_cp, k_s>} in fp16-limb B-fragment order.
// Dense (SP = false), for v_mfma_f32_16x16x32_f16: fragment (c, sb, pt, limb)
// is 64 lanes x 8 fp16; lane = 16 g + n with n = 2 (p % 8) + {0: Re column,
// 1: Im column}; slot j < 4 <-> "Re W" row of support s = 16 sb + 4 g + j,
// slot 4 + j <-> "Im W" row of the same support. One thread fills lanes
// (16 g + 2 pl, 16 g + 2 pl + 1) of every limb for one (c, sb, pt, g, pl): two
// adjacent 16-byte lanes per limb.
// Sparse (SP = true), for v_smfmac_f32_16x16x64_f16 (gfx950 lane layout: B
// element J of lane 16 g + n is k-row 32 (J / 8) + 8 g + J % 8 of column n):
// fragment (c, SB, pt, limb) is 64 lanes x 16 fp16 over 32 supports, k-row
// 2 s' + {0: Re W, 1: Im W} for support s' of the block. Half h = J / 8 of
// lane 16 g + n holds supports 32 SB + 16 (g % 2) + 8 h + 4 (g / 2) + j,
// j = 0..3, as (Re W row, Im W row) pairs; that order matches the compressed
// A fragments built from the E0 tiles in the main kernel. Each half is
// stored lane-major (half h of lane L at 16-byte unit 64 h + L of the
// fragment), so a wave reads a half as one contiguous, conflict-free 1 KiB.
// One thread fills half h of lanes (16 g + 2 pl, 16 g + 2 pl + 1) of every limb.
template <bool SP>
__global__ __launch_bounds__(256) void k_eval_mfma_prep(const qsft_eval_mfma_desc* __restrict__ descs,
                                                        const int32_t* __restrict__ pre, int n_inst) {
    __shared__ double2 tab[QSFT_MAX_Q];
    const int inst = find_inst(pre, n_inst, blockIdx.x);
    const qsft_eval_mfma_desc& d = descs[inst];
    const int q = d.q, n_pad = d.n_pad, nsb = d.nsb, npt = d.npt;
    for (int r = threadIdx.x; r < q; r += blockDim.x) {
        double s, c;
        sincospi(2.0 * (double)r / (double)q, &s, &c);
        tab[r] = make_double2(c, s);
    }
    __syncthreads();
    const int Cb = d.shared_d ? 1 : d.C;
    const int nsbf = SP ? (nsb + 1) / 2 : nsb;              // fragment blocks (32 or 16 supports)
    const int64_t idx = (int64_t)(blockIdx.x - pre[inst]) * blockDim.x + threadIdx.x;
    const int64_t per_c = (int64_t)nsbf * npt * (SP ? 64 : 32);   // (sb, pt, [h,] g, pl)
    if (idx >= per_c * Cb) return;
    const int c = (int)(idx / per_c);
    int64_t rem = idx - c * per_c;
    const int pl = (int)(rem & 7);
    rem >>= 3;
    const int g = (int)(rem & 3);
    rem >>= 2;
    int h = 0;
    if constexpr (SP) {
        h = (int)(rem & 1);
        rem >>= 1;
    }
    const int pt = (int)(rem % npt);
    const int sb = (int)(rem / npt);
    const int p = 8 * pt + pl;
    // first of the thread's four supports
    const int s0 = SP ? 32 * sb + 16 * (g & 1) + 8 * h + 4 * (g >> 1) : 16 * sb + 4 * g;
    const ModQ mq((uint32_t)q);
    const int bexp = d.bexp;
    // <d_p, k_s> for the thread's four supports: 16-byte loads of the delay row
    // and of the four digit rows (rows past P or S re-read the last one and
    // are zeroed below)
    const int4* __restrict__ dr = reinterpret_cast<const int4*>(d.D + ((size_t)c * d.P + min(p, d.P - 1)) * n_pad);
    const int4* __restrict__ kr[4];
#pragma unroll
    for (int j = 0; j < 4; ++j)
        kr[j] = reinterpret_cast<const int4*>(d.K + (size_t)min(s0 + j, d.S - 1) * n_pad);
    int acc[4] = {0, 0, 0, 0};
    for (int w = 0; w < (n_pad >> 4); ++w) {
        const int4 dv = dr[w];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int4 kv = kr[j][w];
            acc[j] = __builtin_amdgcn_sdot4(dv.x, kv.x, acc[j], false);
            acc[j] = __builtin_amdgcn_sdot4(dv.y, kv.y, acc[j], false);
            acc[j] = __builtin_amdgcn_sdot4(dv.z, kv.z, acc[j], false);
            acc[j] = __builtin_amdgcn_sdot4(dv.w, kv.w, acc[j], false);
        }
    }
    uint32_t lr[4][kLimbs], li[4][kLimbs];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int s = s0 + j;
        double br = 0.0, bi = 0.0;
        if (s < d.S && p < d.P) {
            const double2 wv = tab[mq((uint32_t)acc[j])];
            const double2 av = reinterpret_cast<const double2*>(d.a)[s];
            br = ldexp(av.x * wv.x - av.y * wv.y, bexp);
            bi = ldexp(av.x * wv.y + av.y * wv.x, bexp);
        }
        split2(br, lr[j]);
        split2(bi, li[j]);
    }
    v4u* __restrict__ out = reinterpret_cast<v4u*>(d.Bf);
    if constexpr (SP) {
#pragma unroll
        for (int limb = 0; limb < kLimbs; ++limb) {
            const size_t frag = ((((size_t)c * nsbf + sb) * npt + pt) * kLimbs + limb) * 128;
            // Re x column: (Br, -Bi) per support; Im x column: (Bi, Br)
            const v4u re = v4u{lr[0][limb] | ((li[0][limb] ^ 0x8000u) << 16), lr[1][limb] | ((li[1][limb] ^ 0x8000u) << 16),
                               lr[2][limb] | ((li[2][limb] ^ 0x8000u) << 16), lr[3][limb] | ((li[3][limb] ^ 0x8000u) << 16)};
            const v4u im = v4u{li[0][limb] | (lr[0][limb] << 16), li[1][limb] | (lr[1][limb] << 16),
                               li[2][limb] | (lr[2][limb] << 16), li[3][limb] | (lr[3][limb] << 16)};
            out[frag + 64 * h + 16 * g + 2 * pl] = re;
            out[frag + 64 * h + 16 * g + 2 * pl + 1] = im;
        }
        return;
    }
#pragma unroll
    for (int limb = 0; limb < kLimbs; ++limb) {
        const size_t frag = ((((size_t)c * nsb + sb) * npt + pt) * kLimbs + limb) * 64;
        // Re x column: Re W rows carry Br, Im W rows carry -Bi
        const v4u re = v4u{lr[0][limb] | (lr[1][limb] << 16), lr[2][limb] | (lr[3][limb] << 16),
                           (li[0][limb] ^ 0x8000u) | ((li[1][limb] ^ 0x8000u) << 16),
                           (li[2][limb] ^ 0x8000u) | ((li[3][limb] ^ 0x8000u) << 16)};
        // Im x column: Re W rows carry Bi, Im W rows carry Br
        const v4u im = v4u{li[0][limb] | (li[1][limb] << 16), li[2][limb] | (li[3][limb] << 16),
                           lr[0][limb] | (lr[1][limb] << 16), lr[2][limb] | (lr[3][limb] << 16)};
        out[frag + 16 * g + 2 * pl] = re;
        out[frag + 16 * g + 2 * pl + 1] = im;
    }
}


// LDS-count wait (all LDS ops of this wave done), the others unconstrained
__device__ __forceinline__ void wait_lgkm0() {
    __builtin_amdgcn_s_waitcnt(0xF | (0x3 << 14) | (0x7 << 4));
}

// Workgroup barrier for the ring: every wave has already waited for its own
// loads into the slot that becomes readable, so no memory wait is needed here
// (a fenced __syncthreads would drain the loads still in flight for the next
// slot). The empty asm statements keep the compiler from moving LDS accesses
// across the barrier.
__device__ __forceinline__ void ring_barrier() {
    __asm__ volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    __asm__ volatile("" ::: "memory");
}

// MODE 0: q in {2, 4} (w^e exact in fp16, q a power of two)
// MODE 1: other powers of two (2-limb W)   MODE 2: any other q (2-limb W)
//
// SP (MODE 0 only): the sparse form. For q in {2, 4} w^E0 is one of 1, i,
// -1, -i, so of the two k-rows (Re W, Im W) of a support exactly one is
// nonzero (+-1): the 2:4 structured sparsity of v_smfmac_f32_16x16x64_f16,
// which takes a compressed A (8 fp16 per lane) plus 2-bit row indices and
// covers 32 supports per instruction at the issue cost of the dense 16x16x32
// one. A pair of ring slots then holds one 32-support block (16 of its fp16
// B fragments per slot); the two E0 tiles of the block come from the two
// digit slots, and every product is the same exact fp16 product as in the
// dense form. gfx950 lane layout of the sparse instruction: A lane 16 g + m
// holds row m, k-rows 16 g + 4 (c / 2) + pos_c for its elements c = 0..7,
// pos_c = bits 2c..2c+1 of the index word (bits 0..15 with abid = 0); B and D
// as in the prep kernel and the dense instruction.
//
// Support blocks stream through four LDS slots, two pairs, filled by
// global->LDS loads (16 B per lane, lane L at slot base + 16 L): while the
// MFMAs of one pair run, the other pair's loads are in flight, and no
// registers hold staged data. The slots are separate arrays and the loop is
// unrolled by four, so every LDS access names its slot. One wait and one
// barrier per pair of support blocks.
template <int MODE, int NKB, int NPT, bool SP, int NW>
__device__ __forceinline__ void eval_mfma_body(const qsft_eval_mfma_desc* __restrict__ descs,
                                               const int32_t* __restrict__ pre, int n_inst) {
    static_assert(!SP || MODE == 0, "the sparse form needs w^e in {1, i, -1, -i}");
    constexpr int TT = 4;                            // subtiles of 16 base queries per wave
    constexpr int kThreads = 64 * NW;                // NW waves of TT subtiles
    constexpr int kRowsPerBlock = 16 * TT * NW;      // base queries per workgroup
    constexpr int kStagePerThread = kStage / kThreads;   // B loads per lane per slot
    constexpr int WL = (MODE == 0) ? 1 : kLimbs;
    __shared__ __align__(16) v4u bs0[kStage], bs1[kStage], bs2[kStage], bs3[kStage];   // 4 x 16 KiB of B fragments
    __shared__ __align__(16) v4i ks0[NKB * 64], ks1[NKB * 64], ks2[NKB * 64], ks3[NKB * 64];   // support digits
    __shared__ uint32_t wtab[kLimbs][QSFT_MAX_Q];                 // (Re | Im << 16) fp16 bits per limb

    const int lblock = xcd_logical_block();
    const int inst = find_inst(pre, n_inst, lblock);
    const qsft_eval_mfma_desc& d = descs[inst];
    if (d.ptb != NPT) return;                                      // launched for another split (host error)
    if (mfma_wide(d.q, d.b, d.n_pad) != (kRowsPerBlock == 512)) return;   // ... or another workgroup size
    const int q = d.q, b = d.b, n_pad = d.n_pad, npt = d.npt, P = d.P, S = d.S;
    const int nsb = SP ? 2 * ((d.nsb + 1) / 2) : d.nsb;          // 16-support blocks (even when sparse)
    const int Q = ipow(q, b);
    const int nlb = (Q + kRowsPerBlock - 1) / kRowsPerBlock;
    const int ncg = (npt + NPT - 1) / NPT;
    int local = lblock - pre[inst];
    const int cg = local % ncg;
    local /= ncg;
    const int lb = local % nlb;
    const int c = local / nlb;
    const int cb = d.shared_d ? 0 : c;                            // group whose B fragments apply
    const int pt0 = cg * NPT;
    const int npt_here = min(NPT, npt - pt0);
    const int nstage = npt_here * kLimbs * (SP ? 128 : 64);        // 16-byte units of one (pair of) slot(s)

    for (int r = threadIdx.x; r < q; r += blockDim.x) {
        double s, cc;
        sincospi(2.0 * (double)r / (double)q, &s, &cc);
        uint32_t lc[kLimbs], ls[kLimbs];
        split2(cc, lc);
        split2(s, ls);
        for (int i = 0; i < kLimbs; ++i) wtab[i][r] = lc[i] | (ls[i] << 16);
    }

    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int g = lane >> 4, r16 = lane & 15;
    const ModQ mq((uint32_t)q);
    // MODE 0 tables (see compute): q = 4: Re w^f = 1, 0, -1, 0 and Im w^f = 0, 1, 0, -1
    // for f = 0..3; q = 2: Re w^f = 1, -1 for f = 0, 1
    const uint32_t wmask = (q == 2) ? 0x01000100u : 0x03000300u;
    const uint32_t wre = (q == 2) ? 0x0000BC3Cu : 0x00BC003Cu;
    const uint32_t wim = (q == 2) ? 0u : 0xBC003C00u;
    const int lbase = lb * kRowsPerBlock + wave * 16 * TT;

    // base-query digit fragments (B operand of the int8 MFMA): column l = lbase + 16 t + r16
    v4i bq[TT][NKB];
    const int8_t* __restrict__ Mt = d.Mt + (size_t)c * b * n_pad;
#pragma unroll
    for (int t = 0; t < TT; ++t) {
        const int l = lbase + 16 * t + r16;
#pragma unroll
        for (int kb = 0; kb < NKB; ++kb) {
            const int i0 = 64 * kb + 16 * g;
            int acc[16];
#pragma unroll
            for (int jj = 0; jj < 16; ++jj) acc[jj] = 0;
            int rem = l < Q ? l : 0, pw = Q;
            for (int tt = 0; tt < b; ++tt) {
                pw /= q;
                const int lt = rem / pw;
                rem -= lt * pw;
                const v4i mv = *reinterpret_cast<const v4i*>(Mt + (size_t)tt * n_pad + i0);
#pragma unroll
                for (int jj = 0; jj < 16; ++jj) acc[jj] += lt * byte_at(mv, jj);
            }
#pragma unroll
            for (int jj = 0; jj < 16; ++jj) acc[jj] = (int)mq((uint32_t)acc[jj]);
            if constexpr (SP) {
                // digits times 64 (q = 4) or 128 (q = 2), mod 256: the low byte of
                // E0 then holds E0 mod 4 in bits 6..7 (or E0 mod 2 in bit 7), bits
                // 0..5 zero -- the sign / Im flags the sparse W fragments need
#pragma unroll
                for (int jj = 0; jj < 16; ++jj) acc[jj] <<= (q == 2) ? 7 : 6;
            }
            bq[t][kb] = pack16b(acc);
        }
    }

    v4f acc[TT][NPT];
#pragma unroll
    for (int t = 0; t < TT; ++t)
#pragma unroll
        for (int pt = 0; pt < NPT; ++pt) acc[t][pt] = v4f{0.f, 0.f, 0.f, 0.f};

    const v4u* __restrict__ Bg = reinterpret_cast<const v4u*>(d.Bf);
    const int8_t* __restrict__ Kd = d.K;
    // Loads of one support block: kStagePerThread B fragments per lane (p-tiles
    // beyond npt_here re-read the last valid fragment: they only feed
    // accumulators that are never stored), and wave kb < NKB loads digit
    // block kb of the 16 supports (rows past S re-read row S - 1, whose B rows
    // are zero). Group size in loads per lane: kStagePerThread (+1 for those waves).
    auto issue = [&](int sb, v4u* bdst, v4i* kdst) {
        const v4u* bbase = Bg + (((size_t)cb * nsb + sb) * npt + pt0) * kLimbs * 64;
#pragma unroll
        for (int i = 0; i < kStagePerThread; ++i) {
            const uint32_t e = (uint32_t)min(tid + i * kThreads, nstage - 1);
            dma16(bbase, e * 16u, __builtin_amdgcn_readfirstlane(lds_addr(bdst + i * kThreads + wave * 64)));
        }
        if (wave < NKB) {
            const int s = min(16 * sb + r16, S - 1);
            dma16(Kd, (uint32_t)(s * n_pad + 64 * wave + 16 * g),
                  __builtin_amdgcn_readfirstlane(lds_addr(kdst + wave * 64)));
        }
    };
    // all but the newest `keep` groups (0..2) complete
    auto wait_groups = [&](int keep) {
        if (wave < NKB) {
            if (keep >= 2) wait_vm<2 * (kStagePerThread + 1)>();
            else if (keep == 1) wait_vm<kStagePerThread + 1>();
            else wait_vm<0>();
        } else {
            if (keep >= 2) wait_vm<2 * kStagePerThread>();
            else if (keep == 1) wait_vm<kStagePerThread>();
            else wait_vm<0>();
        }
    };

    // E0 tile of subtile t: support digits (A) x base-query digits (B), exact int32
    auto e0_tile = [&](const v4i (&kf)[NKB], int t) {
        v4i e = {0, 0, 0, 0};
#pragma unroll
        for (int kb = 0; kb < NKB; ++kb) e = __builtin_amdgcn_mfma_i32_16x16x64_i8(kf[kb], bq[t][kb], e, 0, 0, 0);
        return e;
    };
    // W fragment(s) of one subtile from its E0 tile
    auto w_tile = [&](const v4i e, v4u (&wt)[WL]) {
        if constexpr (MODE == 0) {
            // q in {2, 4}: w^E0 is 0 or +-1 in each of Re, Im, two supports per 32-bit word
            // by byte permutes. The selector holds f = E0 mod q in bytes 1 and 3 (bytes 0 and
            // 2 select a zero byte); the tables hold the high byte of the fp16 value of
            // Re / Im w^f (1.0 = 0x3C00, -1.0 = 0xBC00) at byte f.
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const uint32_t sel =
                    (__builtin_amdgcn_perm((uint32_t)e[2 * h + 1], (uint32_t)e[2 * h], 0x040C000Cu) & wmask) |
                    0x000C000Cu;
                wt[0][h] = __builtin_amdgcn_perm(0u, wre, sel);              // Re words 0, 1
                wt[0][2 + h] = __builtin_amdgcn_perm(0u, wim, sel);          // Im words 2, 3
            }
        } else {
            uint32_t r[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) r[v] = (MODE == 1) ? ((uint32_t)e[v] & (uint32_t)(q - 1)) : mq((uint32_t)e[v]);
#pragma unroll
            for (int li = 0; li < WL; ++li) {
                uint32_t w[4];
#pragma unroll
                for (int v = 0; v < 4; ++v) w[v] = wtab[li][r[v]];
                wt[li] = v4u{(w[0] & 0xFFFFu) | (w[1] << 16), (w[2] & 0xFFFFu) | (w[3] << 16),
                             (w[0] >> 16) | (w[1] & 0xFFFF0000u), (w[2] >> 16) | (w[3] & 0xFFFF0000u)};
            }
        }
    };
    // W fragments of one support block
    auto wbuild = [&](const v4i* kcur, v4u (&wa)[4][WL]) {
        v4i kf[NKB];
#pragma unroll
        for (int kb = 0; kb < NKB; ++kb) kf[kb] = kcur[kb * 64 + lane];
        v4i e4[4];                                   // all four int8 products first, so their latency overlaps
#pragma unroll
        for (int t = 0; t < 4; ++t) e4[t] = e0_tile(kf, t);
#pragma unroll
        for (int t = 0; t < 4; ++t) w_tile(e4[t], wa[t]);
    };
    // The f16 GEMM of one support block. With PIPE (MODE 0, needs NPT >= 5) the
    // next block's W fragments (support digits kn) are built inside the first
    // limb's MFMA sequence: its int8 products after p-tile 0, one subtile's
    // byte permutes after each of p-tiles 1..4, so that VALU work runs while
    // this wave's MFMAs execute; scheduling barriers keep that order.
    auto gemm = [&](const v4u* bcur, const v4u (&wa)[4][WL], const v4i* kn, v4u (&wn)[4][WL], auto pipe) {
        constexpr bool PIPE = decltype(pipe)::value;
        v4i kfn[NKB];
        v4i en[4];
        if constexpr (PIPE) {
#pragma unroll
            for (int kb = 0; kb < NKB; ++kb) kfn[kb] = kn[kb * 64 + lane];
        }
        // limb-major order: an accumulator is revisited only every 8 x 4 MFMAs
#pragma unroll
        for (int wl = 0; wl < WL; ++wl) {
#pragma unroll
            for (int bl = 0; bl < kLimbs; ++bl) {
                if (wl + bl > 1) continue;                  // drop the lo x lo product (~2^-22 relative)
                // two fragments in registers: the read of p-tile pt + 1 is in
                // flight while the MFMAs of p-tile pt run
                v4u bnext = bcur[bl * 64 + lane];
#pragma unroll
                for (int pt = 0; pt < NPT; ++pt) {
                    const v8h bf = __builtin_bit_cast(v8h, bnext);
                    if (pt + 1 < NPT) bnext = bcur[((pt + 1) * kLimbs + bl) * 64 + lane];
#pragma unroll
                    for (int t = 0; t < 4; ++t)
                        acc[t][pt] = __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(v8h, wa[t][wl]), bf,
                                                                          acc[t][pt], 0, 0, 0);
                    if constexpr (PIPE) {
                        if (bl == 0) {
                            if (pt == 0) {
#pragma unroll
                                for (int t = 0; t < 4; ++t) en[t] = e0_tile(kfn, t);
                            } else if (pt <= 4) {
                                w_tile(en[pt - 1], wn[pt - 1]);
#pragma unroll
                                for (int li = 0; li < WL; ++li)       // materialise here: keeps the
                                    __asm__ volatile("" : "+v"(wn[pt - 1][li]));   // permutes in this slot
                            }
                        }
                        __builtin_amdgcn_sched_barrier(0);
                    }
                }
            }
        }
    };
    // Sparse W fragments of subtile t from its two E0 tiles (elements c = 4 u + v
    // from tile u, lane v): support c of the lane's group is +-1 in its Re W
    // (pos 2 (c % 2)) or Im W (pos 2 (c % 2) + 1) k-row. With the scaled query
    // digits the low byte of each E0 is f << 6, f = E0 mod 4 (q = 4; q = 2:
    // f = 2 (E0 mod 2)): bit 7 is the sign of w^f, bit 6 its Im flag. VALU work
    // competes with the MFMAs for issue slots, so this is 16 operations per
    // subtile: two byte permutes + one and-or per pair of fp16 values, and the
    // index from two gathers, two masks and two byte dot products.
    const uint32_t sp_one = 0x3C003C00u;                         // two fp16 1.0
    auto sp_w = [&](const v4i ea, const v4i eb, v4u& wa, uint32_t& wi) {
        uint32_t pk[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const v4i& e = (k < 2) ? ea : eb;
            const int h = k & 1;
            // bytes [0, E0 byte of element 2k, 0, E0 byte of element 2k + 1]
            pk[k] = __builtin_amdgcn_perm((uint32_t)e[2 * h + 1], (uint32_t)e[2 * h], 0x040C000Cu);
            // +-1.0 in fp16: (pk & 0x80008000) | 0x3C003C00 as one VOP3 (one
            // constant in an SGPR, the other in a VGPR: one constant-bus read;
            // the compiler would split it into two VOP2s with literals)
            __asm__("v_and_or_b32 %0, %1, %2, %3" : "=v"(wa[k]) : "v"(pk[k]), "s"(0x80008000u), "v"(sp_one));
        }
        // E0 bytes of elements 0..3 and 4..7; the Im flags (bit 6 of each byte)
        // weighted by 4^c give im_c at bit 2c (times 64); index bits 2c..2c+1
        // = 2 (c % 2) + im_c, i.e. 0x8888 plus im_c at bit 2c
        const uint32_t q0 = __builtin_amdgcn_perm(pk[1], pk[0], 0x07050301u);
        const uint32_t q1 = __builtin_amdgcn_perm(pk[3], pk[2], 0x07050301u);
        const uint32_t d0 = __builtin_amdgcn_udot4(q0 & 0x40404040u, 0x40100401u, 0x8888u << 6, false);
        const uint32_t d1 = __builtin_amdgcn_udot4(q1 & 0x40404040u, 0x40100401u, 0u, false);
        wi = (d0 >> 6) | (d1 << 2);
    };
    // Sparse ring: a pair of B slots holds one 32-support block m (its 16
    // fragments, f = 2 pt + limb in slot f / 8); the digits travel one block
    // ahead: the loads of block m's B fragments carry block m + 1's digits,
    // so that block m + 1's W fragments are built while block m's MFMAs run.
    const int nsp = nsb / 2;                                       // 32-support blocks
    // wave-uniform parts of the load addresses, computed once (each VALU
    // operation in the loop costs MFMA issue slots)
    const uint32_t wave_lds = (uint32_t)__builtin_amdgcn_readfirstlane(wave) * 1024u;   // 64 lanes x 16 B
    const uint64_t kd_base = uniform_base(Kd);
    uint32_t boff[2][kStagePerThread];                             // per-lane B offsets (bytes)
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
        for (int i = 0; i < kStagePerThread; ++i) boff[u][i] = 16u * (uint32_t)min(u * kStage + tid + i * kThreads, nstage - 1);
    auto issue_sp = [&](int m, v4u* d0, v4u* d1) {                 // B fragments of block m
        const uint64_t gb = uniform_base(Bg + (((size_t)cb * nsp + m) * npt + pt0) * kLimbs * 128);
        const uint32_t l0 = lds_addr(d0) + wave_lds, l1 = lds_addr(d1) + wave_lds;
#pragma unroll
        for (int u = 0; u < 2; ++u)
#pragma unroll
            for (int i = 0; i < kStagePerThread; ++i) dma16_u(gb, boff[u][i], (u ? l1 : l0) + 16u * i * kThreads);
    };
    auto issue_kd = [&](int m, v4i* d0, v4i* d1) {                 // digits of block m (waves < NKB)
        if (wave < NKB) {
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int s = min(32 * m + 16 * u + r16, S - 1);
                dma16_u(kd_base, (uint32_t)(s * n_pad + 64 * wave + 16 * g), lds_addr(u ? d1 : d0) + wave_lds);
            }
        }
    };
    // all W fragments of a block at once (the first block)
    auto sp_wall = [&](const v4i* k0, const v4i* k1, v4u (&wa)[TT], uint32_t (&wi)[TT]) {
        v4i kf0[NKB], kf1[NKB];
#pragma unroll
        for (int kb = 0; kb < NKB; ++kb) {
            kf0[kb] = k0[kb * 64 + lane];
            kf1[kb] = k1[kb * 64 + lane];
        }
        v4i ea[TT], eb[TT];
#pragma unroll
        for (int t = 0; t < TT; ++t) {
            ea[t] = e0_tile(kf0, t);
            eb[t] = e0_tile(kf1, t);
        }
#pragma unroll
        for (int t = 0; t < TT; ++t) sp_w(ea[t], eb[t], wa[t], wi[t]);
    };
    // The GEMM of one block with W fragments (wa, wi). With NEXT, the next
    // block's W fragments (digits in kn0, kn1) are built in the first TT + 2 of
    // the 2 NPT fragment steps: digit reads after step 0, subtile t's two int8
    // products after step t + 1 and its byte work after step t + 2, spread
    // between this wave's MFMAs; scheduling barriers keep that order.
    auto sp_gemm = [&](const v4u* b0, const v4u* b1, const v4u (&wa)[TT], const uint32_t (&wi)[TT], const v4i* kn0,
                       const v4i* kn1, v4u (&wn)[TT], uint32_t (&win)[TT], auto next) {
        constexpr bool NEXT = decltype(next)::value;
        auto frag = [&](int f, v4u& lo, v4u& hi) {
            const v4u* base = (f < 8) ? b0 : b1;
            const int off = (f & 7) * 128 + lane;
            lo = base[off];
            hi = base[off + 64];
        };
        v4i kf0[NKB], kf1[NKB];
        v4i ea, eb;
        v4u nlo, nhi;
        frag(0, nlo, nhi);
#pragma unroll
        for (int st = 0; st < kLimbs * NPT; ++st) {
            const int bl = st / NPT, pt = st % NPT;
            const v16h bf = __builtin_bit_cast(v16h, v8u{nlo[0], nlo[1], nlo[2], nlo[3], nhi[0], nhi[1], nhi[2], nhi[3]});
            if (st + 1 < kLimbs * NPT) frag(2 * ((st + 1) % NPT) + (st + 1) / NPT, nlo, nhi);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int t = 0; t < TT; ++t)
                acc[t][pt] = __builtin_amdgcn_smfmac_f32_16x16x64_f16(__builtin_bit_cast(v8h, wa[t]), bf, acc[t][pt],
                                                                     (int)wi[t], 0, 0);
            if constexpr (NEXT) {
                if (st == 0) {
#pragma unroll
                    for (int kb = 0; kb < NKB; ++kb) {
                        kf0[kb] = kn0[kb * 64 + lane];
                        kf1[kb] = kn1[kb * 64 + lane];
                    }
                }
                if (st >= 2 && st <= TT + 1) {
                    const int t = st - 2;
                    sp_w(ea, eb, wn[t], win[t]);
                    __asm__ volatile("" : "+v"(wn[t]), "+v"(win[t]));   // materialise here
                }
                if (st >= 1 && st <= TT) {
                    ea = e0_tile(kf0, st - 1);
                    eb = e0_tile(kf1, st - 1);
                }
            }
            __builtin_amdgcn_sched_barrier(0);
        }
    };

    if constexpr (SP) {
        // prologue: digits of blocks 0 and 1, B of block 0; W of block 0; then
        // B of block 1 with the digits of block 2 (into block 0's digit slots)
        issue_kd(0, ks0, ks1);
        issue_sp(0, bs0, bs1);
        if (nsp > 1) issue_kd(1, ks2, ks3);
        wait_vm<0>();
        ring_barrier();
        v4u wa0[TT], wa1[TT];
        uint32_t wi0[TT], wi1[TT];
        sp_wall(ks0, ks1, wa0, wi0);
        wait_lgkm0();
        ring_barrier();                              // every wave is done with block 0's digit slots
        if (nsp > 1) {
            issue_sp(1, bs2, bs3);
            if (nsp > 2) issue_kd(2, ks0, ks1);
        }
        // block m in B pair m % 2 with W (wa, wi) of parity m % 2; block m + 1's
        // digits in digit pair (m + 1) % 2
        for (int m = 0; m < nsp; m += 2) {
            // the last block also builds "next" W fragments from stale digit
            // slots (no load is in flight into them); they are never used, and
            // one copy of the GEMM code keeps the register allocation spill-free
            sp_gemm(bs0, bs1, wa0, wi0, ks2, ks3, wa1, wi1, std::true_type{});
            if (m + 1 >= nsp) break;
            wait_vm<0>();                            // B of block m + 1, digits of block m + 2
            ring_barrier();                          // ... for every wave; B pair 0, digit pair 1 free
            if (m + 2 < nsp) {
                issue_sp(m + 2, bs0, bs1);
                if (m + 3 < nsp) issue_kd(m + 3, ks2, ks3);
            }
            sp_gemm(bs2, bs3, wa1, wi1, ks0, ks1, wa0, wi0, std::true_type{});
            if (m + 2 >= nsp) break;
            wait_vm<0>();                            // B of block m + 2, digits of block m + 3
            ring_barrier();                          // ... for every wave; B pair 1, digit pair 0 free
            if (m + 3 < nsp) {
                issue_sp(m + 3, bs2, bs3);
                if (m + 4 < nsp) issue_kd(m + 4, ks0, ks1);
            }
        }
    } else {
    // The four slots form two pairs; a barrier closes each pair of support
    // blocks. While pair A is computed, pair B's loads are in flight; after
    // the barrier (every wave is done with pair A) pair A is refilled with the
    // blocks four ahead.
    issue(0, bs0, ks0);
    if (nsb > 1) issue(1, bs1, ks1);
    if (nsb > 2) issue(2, bs2, ks2);
    if (nsb > 3) issue(3, bs3, ks3);
    wait_groups(max(0, min(nsb - 2, 2)));            // blocks 0 and 1 have landed
    wait_lgkm0();                                    // this wave's wtab writes
    ring_barrier();                                  // ... for every wave; wtab visible

    // a pair of blocks: the second block's W fragments are built while the
    // first block's MFMAs run (independent instructions the scheduler interleaves)
    v4u w0[4][WL], w1[4][WL];
    constexpr bool kPipe = MODE == 0 && NPT >= 5;
    auto compute_pair = [&](const v4u* b0, const v4i* k0, const v4u* b1, const v4i* k1, bool two) {
        wbuild(k0, w0);
        if constexpr (kPipe) {
            // the second slot's digits are read even when it holds no block (its W is then unused)
            gemm(b0, w0, k1, w1, std::true_type{});
        } else {
            if (two) wbuild(k1, w1);
            gemm(b0, w0, k1, w1, std::false_type{});
        }
        if (two) gemm(b1, w1, k1, w1, std::false_type{});
    };
    for (int sb = 0; sb < nsb; sb += 4) {
        compute_pair(bs0, ks0, bs1, ks1, sb + 1 < nsb);
        if (sb + 2 >= nsb) break;
        wait_groups(0);                              // pair B (blocks sb + 2, sb + 3) has landed
        ring_barrier();                              // ... for every wave; pair A is free
        if (sb + 4 < nsb) issue(sb + 4, bs0, ks0);
        if (sb + 5 < nsb) issue(sb + 5, bs1, ks1);
        compute_pair(bs2, ks2, bs3, ks3, sb + 3 < nsb);
        if (sb + 4 >= nsb) break;
        wait_groups(0);                              // pair A (blocks sb + 4, sb + 5) has landed
        ring_barrier();                              // ... for every wave; pair B is free
        if (sb + 6 < nsb) issue(sb + 6, bs2, ks2);
        if (sb + 7 < nsb) issue(sb + 7, bs3, ks3);
    }
    }

    // store: lane holds column n = r16 -> delay row p = 8 (pt0 + pt) + n / 2, Re (n even) / Im (n odd),
    // for base queries l = lbase + 16 t + 4 g + v. Pair Re/Im across lanes (xor 1); the even lane
    // writes queries l0, l0 + 1 and the odd lane l0 + 2, l0 + 3, each pair as one 16-byte store
    // (l0 is a multiple of 4; Q is a power of q >= 2, so an odd Q is only the q^b < 4 case that
    // the per-query branch handles).
    float2* __restrict__ x = reinterpret_cast<float2*>(d.x) + (size_t)c * P * Q;
    const bool odd = r16 & 1;
    const bool even_q = (Q & 1) == 0;
    const float unscale = ldexpf(1.0f, -d.bexp);                  // exact: 2^-bexp is a normal float
#pragma unroll
    for (int t = 0; t < TT; ++t) {
        const int l0 = lbase + 16 * t + 4 * g + (odd ? 2 : 0);
#pragma unroll
        for (int pt = 0; pt < NPT; ++pt) {
            if (pt < npt_here) {
                const v4f a = acc[t][pt];
                v4f o;
#pragma unroll
                for (int v = 0; v < 4; ++v) o[v] = __shfl_xor(a[v], 1);
                const int p = 8 * (pt0 + pt) + (r16 >> 1);
                if (p < P && l0 < Q) {
                    QSFT_GLOBAL float* row = gptr(reinterpret_cast<float*>(x + (size_t)p * Q + l0));
                    const v4f v = odd ? v4f{o[2] * unscale, a[2] * unscale, o[3] * unscale, a[3] * unscale}
                                      : v4f{a[0] * unscale, o[0] * unscale, a[1] * unscale, o[1] * unscale};
                    if (even_q) {                                  // l0 even, so l0 + 1 < Q
                        __builtin_nontemporal_store(v, reinterpret_cast<QSFT_GLOBAL v4f*>(row));   // streamed out
                    } else {
                        *reinterpret_cast<QSFT_GLOBAL v2f*>(row) = v2f{v[0], v[1]};
                        if (l0 + 1 < Q) *reinterpret_cast<QSFT_GLOBAL v2f*>(row + 2) = v2f{v[2], v[3]};
                    }
                }
            }
        }
    }
}

// Default entry: one 256-thread block per CU slot, the register budget the body
// needs without spilling for every (MODE, NKB).
template <int MODE, int NKB>
__global__ __launch_bounds__(kMfThreads) void k_eval_mfma(const qsft_eval_mfma_desc* __restrict__ descs,
                                                          const int32_t* __restrict__ pre, int n_inst) {
    eval_mfma_body<MODE, NKB, kPtPerBlock, false, 4>(descs, pre, n_inst);
}

// Two waves per SIMD: caps the body at 256 VGPR+AGPR. Only the one-limb phase
// (q in {2, 4}) with one k-block (n <= 64) fits that without scratch; each
// SIMD then hides the LDS and global latency of one wave behind the MFMAs of
// the other. NW = 8 waves (512 base queries, for q^b >= 512) per workgroup,
// one workgroup per CU: every staged B fragment feeds twice the queries of a
// four-wave workgroup, which halves the L2 -> LDS traffic; NW = 4 for smaller
// q^b. NPT (5..8) p-tiles per workgroup, chosen so that the tiles of the delay
// rows split evenly over the workgroups (13 tiles: 2 x 7, not 2 x 8).
template <int MODE, int NKB, int NPT, int NW>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_eval_mfma_occ2(
    const qsft_eval_mfma_desc* __restrict__ descs, const int32_t* __restrict__ pre, int n_inst) {
    eval_mfma_body<MODE, NKB, NPT, false, NW>(descs, pre, n_inst);
}

// The sparse form of MODE 0, same two entries.
template <int NKB>
__global__ __launch_bounds__(kMfThreads) void k_eval_smfmac(const qsft_eval_mfma_desc* __restrict__ descs,
                                                            const int32_t* __restrict__ pre, int n_inst) {
    eval_mfma_body<0, NKB, kPtPerBlock, true, 4>(descs, pre, n_inst);
}
template <int NKB, int NPT, int NW>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_eval_smfmac_occ2(
    const qsft_eval_mfma_desc* __restrict__ descs, const int32_t* __restrict__ pre, int n_inst) {
    eval_mfma_body<0, NKB, NPT, true, NW>(descs, pre, n_inst);
}

}  // namespace qsft

using namespace qsft;

extern "C" void qsft_set_error(const char* msg);
extern "C" int qsft_hip_status(hipError_t e, const char* where);

static bool mfma_desc_ok(const qsft_eval_mfma_desc* h) {
    return h && h->q >= 2 && h->q <= QSFT_MAX_Q && h->b >= 1 && h->b <= QSFT_MAX_B && h->n >= 1 && h->n <= h->n_pad &&
           h->n_pad % 64 == 0 && h->n_pad <= QSFT_MAX_N && h->C >= 1 && h->P >= 1 && h->S >= 1 &&
           h->nsb == (h->S + 15) / 16 && h->npt == (h->P + 7) / 8 && (h->shared_d == 0 || h->shared_d == 1) &&
           h->bexp >= -100 && h->bexp <= 100 && h->ptb >= 1 && h->ptb <= kPtPerBlock &&
           (((uintptr_t)h->D | (uintptr_t)h->K | (uintptr_t)h->a) & 15) == 0;
}

// Sized for the sparse layout (support blocks rounded up to a multiple of 32),
// which also holds the dense one.
extern "C" size_t qsft_eval_mfma_workspace_bytes(int C, int P, int S) {
    if (C < 1 || P < 1 || S < 0) return 0;
    return (size_t)C * (2 * ((S + 31) / 32)) * ((P + 7) / 8) * kLimbs * 1024;
}

// base queries per workgroup of the main kernel
static int mfma_rows_per_block(const qsft_eval_mfma_desc* h) {
    return mfma_wide(h->q, h->b, h->n_pad) ? 2 * kMfThreads : kMfThreads;
}

extern "C" int qsft_eval_mfma_blocks(const qsft_eval_mfma_desc* h, int* prep_blocks, int* main_blocks) {
    if (!mfma_desc_ok(h) || !prep_blocks || !main_blocks) return -1;
    const long long Q = ipow(h->q, h->b);
    const long long Cb = h->shared_d ? 1 : h->C;
    // threads for the sparse layout (16-support blocks rounded up to even);
    // the dense prep returns early on the extra ones
    const long long prep = (Cb * (2 * ((h->nsb + 1) / 2)) * h->npt * 32 + 255) / 256;
    const long long rows = mfma_rows_per_block(h);
    const long long mainb = (long long)h->C * ((Q + rows - 1) / rows) *
                            ((h->npt + h->ptb - 1) / h->ptb);
    if (prep > 0x7FFFFFFF || mainb > 0x7FFFFFFF) return -1;
    *prep_blocks = (int)prep;
    *main_blocks = (int)mainb;
    return 0;
}

template <int NW>
static void launch_smfmac_occ2(int ptb, const qsft_eval_mfma_desc* d, const int32_t* pre, int n, int blocks,
                               hipStream_t st) {
    switch (ptb) {
        case 5: hipLaunchKernelGGL((k_eval_smfmac_occ2<1, 5, NW>), dim3(blocks), dim3(64 * NW), 0, st, d, pre, n); break;
        case 6: hipLaunchKernelGGL((k_eval_smfmac_occ2<1, 6, NW>), dim3(blocks), dim3(64 * NW), 0, st, d, pre, n); break;
        case 7: hipLaunchKernelGGL((k_eval_smfmac_occ2<1, 7, NW>), dim3(blocks), dim3(64 * NW), 0, st, d, pre, n); break;
        default: hipLaunchKernelGGL((k_eval_smfmac_occ2<1, 8, NW>), dim3(blocks), dim3(64 * NW), 0, st, d, pre, n); break;
    }
}
template <int NW>
static void launch_mfma_occ2(int ptb, const qsft_eval_mfma_desc* d, const int32_t* pre, int n, int blocks,
                             hipStream_t st) {
    switch (ptb) {
        case 5: hipLaunchKernelGGL((k_eval_mfma_occ2<0, 1, 5, NW>), dim3(blocks), dim3(64 * NW), 0, st, d, pre, n); break;
        case 6: hipLaunchKernelGGL((k_eval_mfma_occ2<0, 1, 6, NW>), dim3(blocks), dim3(64 * NW), 0, st, d, pre, n); break;
        case 7: hipLaunchKernelGGL((k_eval_mfma_occ2<0, 1, 7, NW>), dim3(blocks), dim3(64 * NW), 0, st, d, pre, n); break;
        default: hipLaunchKernelGGL((k_eval_mfma_occ2<0, 1, 8, NW>), dim3(blocks), dim3(64 * NW), 0, st, d, pre, n); break;
    }
}

// rows: base queries per workgroup (mfma_rows_per_block of the launch's descriptors)
static hipError_t launch_smfmac(int nkb, int ptb, int rows, const qsft_eval_mfma_desc* d, const int32_t* pre, int n,
                                int blocks, hipStream_t st) {
    switch (nkb) {
        case 1:
            if (rows == 2 * kMfThreads) launch_smfmac_occ2<8>(ptb, d, pre, n, blocks, st);
            else launch_smfmac_occ2<4>(ptb, d, pre, n, blocks, st);
            break;
        case 2: hipLaunchKernelGGL((k_eval_smfmac<2>), dim3(blocks), dim3(kMfThreads), 0, st, d, pre, n); break;
        case 3: hipLaunchKernelGGL((k_eval_smfmac<3>), dim3(blocks), dim3(kMfThreads), 0, st, d, pre, n); break;
        default: hipLaunchKernelGGL((k_eval_smfmac<4>), dim3(blocks), dim3(kMfThreads), 0, st, d, pre, n); break;
    }
    return hipGetLastError();
}

template <int MODE>
static hipError_t launch_mfma(int nkb, int ptb, int rows, const qsft_eval_mfma_desc* d, const int32_t* pre, int n,
                              int blocks, hipStream_t st) {
    switch (nkb) {
        case 1:
            if constexpr (MODE == 0) {
                if (rows == 2 * kMfThreads) launch_mfma_occ2<8>(ptb, d, pre, n, blocks, st);
                else launch_mfma_occ2<4>(ptb, d, pre, n, blocks, st);
            } else {
                hipLaunchKernelGGL((k_eval_mfma<MODE, 1>), dim3(blocks), dim3(kMfThreads), 0, st, d, pre, n);
            }
            break;
        case 2: hipLaunchKernelGGL((k_eval_mfma<MODE, 2>), dim3(blocks), dim3(kMfThreads), 0, st, d, pre, n); break;
        case 3: hipLaunchKernelGGL((k_eval_mfma<MODE, 3>), dim3(blocks), dim3(kMfThreads), 0, st, d, pre, n); break;
        default: hipLaunchKernelGGL((k_eval_mfma<MODE, 4>), dim3(blocks), dim3(kMfThreads), 0, st, d, pre, n); break;
    }
    return hipGetLastError();
}

// word: bits 4..7 digit-block count (1..4), bits 8..9 mode (0: q in {2,4}; 1: other powers of two; 2: other q),
// bit 10: mode 0 in the dense form (default: sparse), bit 11: eight-wave workgroups (mfma_wide of the
// descriptors), bits 12..15 p-tiles per workgroup (the descriptors' ptb; 5..8 for mode 0 with one digit
// block, else 8). All instances of one launch share the word.
extern "C" int qsft_eval_mfma(const qsft_eval_mfma_desc* d, const int32_t* pre_prep, const int32_t* pre_main,
                              int n_inst, int prep_blocks, int main_blocks, int word, void* stream) {
    if (!d || !pre_prep || !pre_main || n_inst < 1 || prep_blocks < 0 || main_blocks < 0) {
        qsft_set_error("qsft_eval_mfma: bad arguments");
        return QSFT_ERR_ARG;
    }
    const int nkb = (word >> 4) & 0xF, mode = (word >> 8) & 0x3, ptb = (word >> 12) & 0xF;
    const bool dense0 = (word >> 10) & 1, wide = (word >> 11) & 1;
    const int rows = wide ? 2 * kMfThreads : kMfThreads;
    const bool split_ok = (mode == 0 && nkb == 1) ? (ptb >= 5 && ptb <= kPtPerBlock) : (ptb == kPtPerBlock);
    if (nkb < 1 || nkb > 4 || mode > 2 || !split_ok || (wide && (mode != 0 || nkb != 1))) {
        qsft_set_error("qsft_eval_mfma: bad word");
        return QSFT_ERR_ARG;
    }
    hipStream_t st = (hipStream_t)stream;
    if (prep_blocks > 0) {
        if (mode == 0 && !dense0)
            hipLaunchKernelGGL(k_eval_mfma_prep<true>, dim3(prep_blocks), dim3(256), 0, st, d, pre_prep, n_inst);
        else
            hipLaunchKernelGGL(k_eval_mfma_prep<false>, dim3(prep_blocks), dim3(256), 0, st, d, pre_prep, n_inst);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return qsft_hip_status(e, "qsft_eval_mfma(prep)");
    }
    if (main_blocks == 0) return QSFT_OK;
    hipError_t e;
    if (mode == 0 && !dense0) e = launch_smfmac(nkb, ptb, rows, d, pre_main, n_inst, main_blocks, st);
    else if (mode == 0) e = launch_mfma<0>(nkb, ptb, rows, d, pre_main, n_inst, main_blocks, st);
    else if (mode == 1) e = launch_mfma<1>(nkb, ptb, rows, d, pre_main, n_inst, main_blocks, st);
    else e = launch_mfma<2>(nkb, ptb, rows, d, pre_main, n_inst, main_blocks, st);
    return qsft_hip_status(e, "qsft_eval_mfma");
}
