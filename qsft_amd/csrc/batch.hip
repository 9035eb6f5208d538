// Batch runtime: every stage of many synthetic transform requests planned,
// bound and enqueued by native code, so a serving loop pays a few host
// microseconds per request instead of building descriptors per stage.
//
// A batch is created from request shapes (sizes only), its inputs are packed
// into a host staging arena, it is bound to caller-owned device arenas
// (inputs, work buffers, plan) and an output buffer, and each launch enqueues
//   upload (optional) -> eval prep + GEMM -> sub-DFTs -> assemble + noise ->
//   peel passes 0 .. fuse_from-1 -> fused passes fuse_from .. max_iter-1 ->
//   (tail stream) finalize -> pack in first-detection order
// through the library's own entry points. Nothing here allocates device
// memory; the two HIP events per batch are created once, at the first bind.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <map>
#include <new>
#include <thread>
#include <tuple>
#include <vector>

#include "qsft_common.h"

extern "C" void qsft_set_error(const char* msg);
extern "C" int qsft_hip_status(hipError_t e, const char* where);

namespace {

size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

long long ipow_ll(int q, int e) {
    long long r = 1;
    for (int i = 0; i < e; ++i) r *= q;
    return r;
}

// Fake arena bases for the sizing pass (only offsets and alignment matter there).
constexpr uintptr_t kFakeBase = (uintptr_t)1 << 40;

struct ReqLayout {
    int n_pad, P, Q;
    bool energy, sparse;
    // inside the request's input block
    size_t mt, d, kraw, a, in_size;
    // inside the request's work block
    size_t x, u, en, bf, ws, k, work_size, ws_bytes, bf_bytes;
    // block offsets in the arenas
    size_t in_off, work_off;
};

bool layout_of(const qsft_request& r, int max_iter, ReqLayout* L) {
    if (r.q < 2 || r.q > QSFT_MAX_Q || r.n < 1 || r.n > QSFT_MAX_N || r.b < 1 || r.b > QSFT_MAX_B || r.C < 1 ||
        r.C > QSFT_MAX_C || r.R < 1 || r.R > QSFT_MAX_R || r.S < 1 ||
        (r.method != QSFT_DETECT_NSO1 && r.method != QSFT_DETECT_NOISELESS) ||
        (r.method == QSFT_DETECT_NOISELESS && r.R != 1) || !(r.sigma >= 0.0) || !(r.amax >= 0.0))
        return false;
    ReqLayout l{};
    l.n_pad = (r.n + 63) / 64 * 64;
    l.P = r.R * (r.n + 1);
    const long long Q = ipow_ll(r.q, r.b);
    if (l.P > 256 || Q * r.C > 0x3FFFFFFF) return false;
    l.Q = (int)Q;
    l.energy = l.P <= 128;
    l.sparse = l.energy;
    size_t o = 0;
    l.mt = o; o += al256((size_t)r.C * r.b * l.n_pad);
    l.d = o;  o += al256((size_t)l.P * l.n_pad);                  // one copy: every group shares the rows
    l.kraw = o; o += al256((size_t)r.n * r.S);                   // locq as given, [n][S]
    l.a = o;  o += al256((size_t)r.S * 16);
    l.in_size = o;
    const size_t cells = (size_t)r.C * l.P * Q;
    l.bf_bytes = qsft_eval_mfma_workspace_bytes(1, l.P, r.S);
    l.ws_bytes = qsft_peel_workspace_bytes(r.q, r.n, r.b, r.C, l.P, max_iter, 0);
    if (l.bf_bytes == 0 || l.ws_bytes == 0) return false;
    o = 0;
    l.x = o;  o += al256(cells * 8);
    l.u = o;  o += al256(cells * 16);
    l.en = o; o += l.energy ? al256((size_t)r.C * Q * 8) : 0;
    l.bf = o; o += al256(l.bf_bytes);
    l.ws = o; o += al256(l.ws_bytes);
    l.k = o;  o += al256((size_t)r.S * l.n_pad);                 // K [S][n_pad], staged on the device
    l.work_size = o;
    *L = l;
    return true;
}

int mfma_mode_of(int q) {
    if (q == 2 || q == 4) return 0;
    return (q & (q - 1)) == 0 ? 1 : 2;
}

int mfma_ptb_of(int mode, int nkb, int npt) {
    if (mode != 0 || nkb != 1) return 8;
    const int groups = (npt + 7) / 8;
    const int per = (npt + groups - 1) / groups;
    return per < 5 ? 5 : per;
}

int mfma_bexp_of(double amax) {
    if (!(amax > 0.0)) return 0;
    int e = 0;
    frexp(amax, &e);
    int v = 14 - e;
    return v < -100 ? -100 : (v > 100 ? 100 : v);
}

// Support digits arrive as the reference holds them, locq [n][S]; the
// evaluator reads K [S][n_pad]. One launch transposes every request's block
// through LDS (64 supports per workgroup, coalesced 32-bit stores, zero
// padding past n), so the host only copies bytes.
struct StageDesc {
    const int8_t* src;   // [n][S]
    int8_t* dst;         // [S][n_pad]
    int32_t n, S, n_pad, _pad;
};

__global__ __launch_bounds__(256) void k_stage_support(const StageDesc* __restrict__ descs,
                                                       const int32_t* __restrict__ pre, int n_req) {
    __shared__ __align__(4) int8_t tile[64][QSFT_MAX_N + 4];
    const int r = qsft::find_inst(pre, n_req, blockIdx.x);
    const StageDesc& d = descs[r];
    const int s0 = (blockIdx.x - pre[r]) * 64;
    const int n = d.n, S = d.S, np = d.n_pad;
    for (int idx = threadIdx.x; idx < 64 * np; idx += 256) {
        const int j = idx >> 6, sl = idx & 63;
        tile[sl][j] = (j < n && s0 + sl < S) ? d.src[(size_t)j * S + s0 + sl] : (int8_t)0;
    }
    __syncthreads();
    const int wpr = np >> 2;                                       // 32-bit words per row
    for (int w = threadIdx.x; w < 64 * wpr; w += 256) {
        const int sl = w / wpr, jw = w - sl * wpr;
        if (s0 + sl < S)
            *reinterpret_cast<uint32_t*>(d.dst + (size_t)(s0 + sl) * np + 4 * jw) =
                *reinterpret_cast<const uint32_t*>(&tile[sl][4 * jw]);
    }
}

struct MfmaLaunch { size_t desc, pre_prep, pre_main; int n, prep_blocks, main_blocks, word; };
struct Rows4Launch { size_t desc, pre; int n, blocks, word; };
struct DftLaunch { size_t desc, pre[QSFT_DFT_MAX_PASSES]; int n, npass, blocks[QSFT_DFT_MAX_PASSES],
                   lds[QSFT_DFT_MAX_PASSES], word; };

struct Batch {
    std::vector<qsft_request> req;
    std::vector<ReqLayout> lay;
    std::vector<int> sign;                 // detected delay structure per request (+1, -1, 0)
    int max_iter = 15, fuse_from = 3, max_p = 1, max_n_pad = 64;
    qsft_batch_sizes sz{};
    // bound state
    char* in_dev = nullptr;
    char* work_dev = nullptr;
    char* plan_host = nullptr;
    char* plan_dev = nullptr;
    void* out = nullptr;
    bool bound = false, launched = false;
    std::vector<MfmaLaunch> mfma;
    std::vector<Rows4Launch> rows4;
    std::vector<DftLaunch> dft;
    size_t asm_desc = 0, asm_pre = 0;
    int asm_blocks = 0, asm_word = 0;
    size_t peel_desc = 0, peel_pre_bins = 0, peel_pre_slots = 0;
    int peel_bins = 0, peel_slots = 0;
    size_t stage_desc = 0, stage_pre = 0;
    int stage_blocks = 0;
    size_t seed_ctr = 0;                    // plan offset of the replay seed counter
    hipEvent_t fork = nullptr, done = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
};

// One replay of a captured batch: the next seed offset.
__global__ void k_seed_bump(uint64_t* ctr) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *ctr = *ctr + 1;
}

// Appends descriptor arrays and block prefixes to the plan; with `write`,
// into plan_host (bases are the real arenas), else only sizes (fake bases).
struct PlanWriter {
    char* host;
    bool write;
    size_t used = 0;
    size_t put(const void* src, size_t n) {
        const size_t at = used;
        if (write && n) memcpy(host + at, src, n);
        used += al256(n);
        return at;
    }
    size_t prefix(const std::vector<int>& blocks, int* total) {
        std::vector<int32_t> pre(blocks.size() + 1, 0);
        long long acc = 0;
        for (size_t i = 0; i < blocks.size(); ++i) {
            pre[i] = (int32_t)acc;
            acc += blocks[i];
        }
        pre[blocks.size()] = (int32_t)acc;
        *total = acc > 0x7FFFFFFF ? -1 : (int)acc;
        return put(pre.data(), pre.size() * 4);
    }
};

int build_plan(Batch& B, bool write) {
    const int nreq = (int)B.req.size();
    const uintptr_t in_base = write ? (uintptr_t)B.in_dev : kFakeBase;
    const uintptr_t work_base = write ? (uintptr_t)B.work_dev : kFakeBase;
    PlanWriter W{B.plan_host, write};
    B.mfma.clear();
    B.rows4.clear();
    B.dft.clear();
    auto in_ptr = [&](int i, size_t off) { return (void*)(in_base + B.lay[i].in_off + off); };
    auto work_ptr = [&](int i, size_t off) { return (void*)(work_base + B.lay[i].work_off + off); };

    {
        const uint64_t zero = 0;
        B.seed_ctr = W.put(&zero, sizeof(zero));                   // replay seed counter, starts at 0
    }
    // ---- support digits: locq [n][S] -> K [S][n_pad] on the device
    {
        std::vector<StageDesc> sd(nreq);
        std::vector<int> sb;
        for (int i = 0; i < nreq; ++i) {
            const qsft_request& r = B.req[i];
            sd[i].src = (const int8_t*)in_ptr(i, B.lay[i].kraw);
            sd[i].dst = (int8_t*)work_ptr(i, B.lay[i].k);
            sd[i].n = r.n;
            sd[i].S = r.S;
            sd[i].n_pad = B.lay[i].n_pad;
            sd[i]._pad = 0;
            sb.push_back((r.S + 63) / 64);
        }
        B.stage_desc = W.put(sd.data(), sd.size() * sizeof(StageDesc));
        B.stage_pre = W.prefix(sb, &B.stage_blocks);
    }

    // ---- stage (a)+(b): matrix-core evaluation, grouped by kernel variant
    std::map<std::tuple<int, int, int, int>, std::vector<int>> eg;
    for (int i = 0; i < nreq; ++i) {
        const qsft_request& r = B.req[i];
        const int nkb = B.lay[i].n_pad / 64, mode = mfma_mode_of(r.q);
        eg[std::make_tuple(nkb, mode, mfma_ptb_of(mode, nkb, (B.lay[i].P + 7) / 8),
                           (int)qsft::mfma_wide(r.q, r.b, B.lay[i].n_pad))].push_back(i);
    }
    for (auto& kv : eg) {
        const int nkb = std::get<0>(kv.first), mode = std::get<1>(kv.first), ptb = std::get<2>(kv.first);
        const int wide = std::get<3>(kv.first);
        std::vector<qsft_eval_mfma_desc> descs;
        std::vector<int> prep, mainb;
        for (int i : kv.second) {
            const qsft_request& r = B.req[i];
            const ReqLayout& l = B.lay[i];
            qsft_eval_mfma_desc d{};
            d.q = r.q; d.n = r.n; d.b = r.b; d.C = r.C; d.P = l.P; d.S = r.S; d.n_pad = l.n_pad;
            d.nsb = (r.S + 15) / 16; d.npt = (l.P + 7) / 8; d.shared_d = 1; d.bexp = mfma_bexp_of(r.amax);
            d.ptb = ptb;
            d.Mt = (const int8_t*)in_ptr(i, l.mt);
            d.D = (const int8_t*)in_ptr(i, l.d);
            d.K = (const int8_t*)work_ptr(i, l.k);
            d.a = (const double*)in_ptr(i, l.a);
            d.Bf = (uint16_t*)work_ptr(i, l.bf);
            d.x = work_ptr(i, l.x);
            int pb = 0, mb = 0;
            if (qsft_eval_mfma_blocks(&d, &pb, &mb) != 0) {
                qsft_set_error("qsft_batch: bad evaluation shape");
                return QSFT_ERR_ARG;
            }
            descs.push_back(d);
            prep.push_back(pb);
            mainb.push_back(mb);
        }
        MfmaLaunch L{};
        L.desc = W.put(descs.data(), descs.size() * sizeof(qsft_eval_mfma_desc));
        L.pre_prep = W.prefix(prep, &L.prep_blocks);
        L.pre_main = W.prefix(mainb, &L.main_blocks);
        if (L.prep_blocks < 0 || L.main_blocks < 0) { qsft_set_error("qsft_batch: launch too large"); return QSFT_ERR_ARG; }
        L.n = (int)descs.size();
        L.word = (nkb << 4) | (mode << 8) | (wide << 11) | (ptb << 12);
        B.mfma.push_back(L);
    }

    // ---- stage (c): sub-DFTs in place on x (whole-row kernel for q = 4, b in {4, 6})
    std::map<int, std::vector<int>> fast, slow;
    for (int i = 0; i < nreq; ++i) {
        const qsft_request& r = B.req[i];
        if (r.q == 4 && (r.b == 4 || r.b == 6)) fast[r.b].push_back(i);
        else slow[r.q].push_back(i);
    }
    auto dft_desc = [&](int i) {
        const qsft_request& r = B.req[i];
        const ReqLayout& l = B.lay[i];
        qsft_dft_desc d{};
        d.q = r.q; d.b = r.b; d.rows = r.C * l.P; d.in_stride = 1; d.in_row_stride = l.Q;
        d.in = work_ptr(i, l.x);
        d.out = work_ptr(i, l.x);
        return d;
    };
    for (auto& kv : fast) {
        std::vector<qsft_dft_desc> descs;
        std::vector<int> rows;
        for (int i : kv.second) {
            descs.push_back(dft_desc(i));
            rows.push_back(descs.back().rows);
        }
        Rows4Launch L{};
        L.desc = W.put(descs.data(), descs.size() * sizeof(qsft_dft_desc));
        L.pre = W.prefix(rows, &L.blocks);
        L.n = (int)descs.size();
        L.word = QSFT_F32 | (kv.first << 8);
        B.rows4.push_back(L);
    }
    for (auto& kv : slow) {
        std::vector<qsft_dft_desc> descs;
        for (int i : kv.second) {
            qsft_dft_desc d = dft_desc(i);
            const int rc = qsft_dft_plan(&d, QSFT_F32);
            if (rc != QSFT_OK) return rc;
            descs.push_back(d);
        }
        DftLaunch L{};
        L.npass = 0;
        for (auto& d : descs) L.npass = d.n_pass > L.npass ? d.n_pass : L.npass;
        L.desc = W.put(descs.data(), descs.size() * sizeof(qsft_dft_desc));
        for (int p = 0; p < L.npass; ++p) {
            std::vector<int> bl;
            int lds = 16;
            for (auto& d : descs) {
                bl.push_back(p < d.n_pass ? d.pass_blocks[p] : 0);
                if (p < d.n_pass) {
                    const int x = qsft_dft_lds_bytes(&d, p, QSFT_F32);
                    if (x < 0) { qsft_set_error("qsft_batch: DFT tile does not fit the LDS"); return QSFT_ERR_ARG; }
                    lds = x > lds ? x : lds;
                }
            }
            L.pre[p] = W.prefix(bl, &L.blocks[p]);
            L.lds[p] = lds;
        }
        L.n = (int)descs.size();
        L.word = QSFT_F32 | (kv.first << 8);
        B.dft.push_back(L);
    }

    // ---- peel descriptors first (the assembler points at their sparse-U words)
    std::vector<qsft_peel_desc> pdescs(nreq);
    std::vector<int> pbins, pslots;
    for (int i = 0; i < nreq; ++i) {
        const qsft_request& r = B.req[i];
        const ReqLayout& l = B.lay[i];
        qsft_peel_desc& d = pdescs[i];
        int rc = qsft_peel_bind(&d, r.q, r.n, r.b, r.C, l.P, r.n + 1, r.method, B.max_iter, r.cutoff, 0,
                                (const int8_t*)in_ptr(i, l.mt), (const int8_t*)in_ptr(i, l.d),
                                (double*)work_ptr(i, l.u), work_ptr(i, l.ws), l.ws_bytes, l.n_pad);
        if (rc != QSFT_OK) return rc;
        d.flags = QSFT_PEEL_SHARED_D;
        if (B.sign[i] != 0) d.flags |= QSFT_PEEL_UNIT_D | (B.sign[i] < 0 ? QSFT_PEEL_UNIT_D_NEG : 0);
        if (l.energy) d.energy = (const double*)work_ptr(i, l.en);
        if (l.sparse) {
            d.flags |= QSFT_PEEL_SPARSE_U;
            d.xsrc = work_ptr(i, l.x);
            d.noise_sigma = r.sigma;
        }
        pbins.push_back(qsft_peel_bin_blocks(&d));
        pslots.push_back(qsft_peel_slot_blocks(&d));
    }

    // ---- get_MDU step: FP64 working transform, bin-major, Philox noise
    std::vector<qsft_assemble_desc> adescs(nreq);
    std::vector<int> ablocks;
    int aword = -1;
    for (int i = 0; i < nreq; ++i) {
        const qsft_request& r = B.req[i];
        const ReqLayout& l = B.lay[i];
        qsft_assemble_desc& d = adescs[i];
        memset(&d, 0, sizeof(d));
        d.Q = l.Q; d.C = r.C; d.R = r.R; d.p1 = r.n + 1; d.src_rows = l.P; d.src_prec = QSFT_F32;
        d.bin_major = 1;
        for (int c = 0; c < r.C; ++c) d.gsel[c] = c;
        for (int j = 0; j < r.R; ++j) d.rsel[j] = j;
        d.src = work_ptr(i, l.x);
        if (r.sigma > 0.0) {
            d.noise_mode = QSFT_NOISE_PHILOX;
            d.sigma = r.sigma;
            d.seed = r.seed;
        } else {
            d.noise_mode = QSFT_NOISE_NONE;
        }
        d.U = (double*)work_ptr(i, l.u);
        if (l.energy) d.energy = (double*)work_ptr(i, l.en);
        if (l.sparse) {
            d.written = pdescs[i].written;
            d.cutoff = r.cutoff;
            d.seed_out = pdescs[i].noise_key;
        }
        const int nb = qsft_assemble_blocks(&d);
        if (nb < 0) { qsft_set_error("qsft_batch: bad assemble shape"); return QSFT_ERR_ARG; }
        ablocks.push_back(nb);
        aword &= qsft_assemble_word(&d);
    }
    B.asm_desc = W.put(adescs.data(), adescs.size() * sizeof(qsft_assemble_desc));
    B.asm_pre = W.prefix(ablocks, &B.asm_blocks);
    B.asm_word = aword;

    B.peel_desc = W.put(pdescs.data(), pdescs.size() * sizeof(qsft_peel_desc));
    B.peel_pre_bins = W.prefix(pbins, &B.peel_bins);
    B.peel_pre_slots = W.prefix(pslots, &B.peel_slots);
    if (B.asm_blocks < 0 || B.peel_bins < 0 || B.peel_slots < 0) {
        qsft_set_error("qsft_batch: launch too large");
        return QSFT_ERR_ARG;
    }
    B.sz.plan_bytes = W.used;
    return QSFT_OK;
}

// delay rows structure of one request: +1 / -1 if every repeat block is its row
// 0 plus s e_(i-1) (mod q) in row i (identity / NSO delays), else 0
// (digits are already checked to lie in [0, q), so no division is needed)
int delay_sign(const int8_t* D, int q, int n, int R) {
    const int p1 = n + 1;
    for (int s = -1; s <= 1; s += 2) {
        bool ok = true;
        for (int r = 0; r < R && ok; ++r) {
            const int8_t* row0 = D + (size_t)r * p1 * n;
            for (int i = 1; i <= n && ok; ++i) {
                const int8_t* row = row0 + (size_t)i * n;
                for (int j = 0; j < n; ++j) {
                    if (j == i - 1) continue;
                    if (row[j] != row0[j]) { ok = false; break; }
                }
                if (!ok) break;
                int want = row0[i - 1] + s;
                if (want < 0) want += q;
                if (want >= q) want -= q;
                ok = row[i - 1] == want;
            }
        }
        if (ok) return s;
    }
    return 0;
}

Batch* as_batch(void* h) { return static_cast<Batch*>(h); }

}  // namespace

extern "C" size_t qsft_request_footprint(const qsft_request* r, int max_iter) {
    ReqLayout l;
    if (!r || max_iter < 1 || max_iter > 15 || !layout_of(*r, max_iter, &l)) return 0;
    return l.in_size + l.work_size;
}

extern "C" void* qsft_batch_create(const qsft_request* reqs, int n_req, int max_iter, int fuse_from,
                                   qsft_batch_sizes* sizes) {
    if (!reqs || n_req < 1 || max_iter < 1 || max_iter > 15 || fuse_from < 0 || !sizes) {
        qsft_set_error("qsft_batch_create: bad arguments");
        return nullptr;
    }
    Batch* B = new (std::nothrow) Batch();
    if (!B) { qsft_set_error("qsft_batch_create: out of host memory"); return nullptr; }
    B->req.assign(reqs, reqs + n_req);
    B->lay.resize(n_req);
    B->sign.assign(n_req, 0);
    B->max_iter = max_iter;
    B->fuse_from = fuse_from < max_iter ? fuse_from : max_iter;
    size_t in_off = 0, work_off = 0;
    long long cap_rows = 0;
    for (int i = 0; i < n_req; ++i) {
        if (!layout_of(reqs[i], max_iter, &B->lay[i])) {
            qsft_set_error("qsft_batch_create: bad request shape (q <= 64, n <= 256, b <= 16, C, R <= 16, "
                           "R (n + 1) <= 256, S >= 1, max|a| > 0; noiseless detection needs R = 1)");
            delete B;
            return nullptr;
        }
        B->lay[i].in_off = in_off;
        B->lay[i].work_off = work_off;
        in_off += B->lay[i].in_size;
        work_off += B->lay[i].work_size;
        B->max_p = B->lay[i].P > B->max_p ? B->lay[i].P : B->max_p;
        B->max_n_pad = B->lay[i].n_pad > B->max_n_pad ? B->lay[i].n_pad : B->max_n_pad;
        cap_rows += 4ll * reqs[i].S + 1024;                       // planted keys plus noise detections
    }
    B->sz.input_bytes = in_off;
    B->sz.work_bytes = work_off;
    B->sz.cap_rows = cap_rows;
    B->sz.row_stride = B->max_n_pad + 24;
    B->sz.out_bytes = qsft_peel_pack_header_bytes(n_req) + (size_t)cap_rows * B->sz.row_stride;
    B->sz.n_req = n_req;
    if (build_plan(*B, false) != QSFT_OK) {
        delete B;
        return nullptr;
    }
    *sizes = B->sz;
    return B;
}

extern "C" int qsft_batch_input_offset(void* h, int i, size_t* offsets) {
    Batch* B = as_batch(h);
    if (!B || i < 0 || i >= (int)B->req.size() || !offsets) {
        qsft_set_error("qsft_batch_input_offset: bad arguments");
        return QSFT_ERR_ARG;
    }
    const ReqLayout& l = B->lay[i];
    offsets[0] = l.in_off + l.mt;
    offsets[1] = l.in_off + l.d;
    offsets[2] = l.in_off + l.kraw;
    offsets[3] = l.in_off + l.a;
    return QSFT_OK;
}

extern "C" int qsft_batch_pack_input(void* h, int i, const int8_t* M, const int8_t* D, const int8_t* locq,
                                     const double* a, void* input_host) {
    Batch* B = as_batch(h);
    if (!B || i < 0 || i >= (int)B->req.size() || !M || !D || !locq || !a || !input_host) {
        qsft_set_error("qsft_batch_pack_input: bad arguments");
        return QSFT_ERR_ARG;
    }
    qsft_request& r = B->req[i];
    const ReqLayout& l = B->lay[i];
    // locals only in the loops: int8 stores may alias anything, so fields read
    // through references would be reloaded on every store
    const int n = r.n, b = r.b, C = r.C, P = l.P, S = r.S, np = l.n_pad;
    if (!(r.amax > 0.0)) {
        // amax 0: max |a_s| from the strengths, as sqrt(max(re^2 + im^2)) with
        // every product and sum rounded (no fused multiply-add)
        double m2 = 0.0;
        bool nan = false;
        for (int s = 0; s < S; ++s) {
            const double re = a[2 * (size_t)s], im = a[2 * (size_t)s + 1];
            const double t = __builtin_fma(re, re, 0.0) + __builtin_fma(im, im, 0.0);   // each product rounded
            nan |= t != t;
            m2 = t > m2 ? t : m2;
        }
        if (nan || !(m2 > 0.0) || m2 == __builtin_inf()) {
            qsft_set_error("qsft_batch_pack_input: the strengths are all zero or not finite");
            return QSFT_ERR_ARG;
        }
        r.amax = __builtin_sqrt(m2);
    }
    const int q = r.q;
    char* base = static_cast<char*>(input_host) + l.in_off;
    int8_t* __restrict__ Mt = reinterpret_cast<int8_t*>(base + l.mt);
    int8_t* __restrict__ Dd = reinterpret_cast<int8_t*>(base + l.d);
    // every digit in [0, q) (branch-free passes over the int8 inputs)
    int bad = 0;
    for (size_t t = 0, e = (size_t)C * n * b; t < e; ++t) bad |= (uint8_t)M[t] >= q;
    for (size_t t = 0, e = (size_t)P * n; t < e; ++t) bad |= (uint8_t)D[t] >= q;
    for (size_t t = 0, e = (size_t)n * S; t < e; ++t) bad |= (uint8_t)locq[t] >= q;
    if (bad) {
        qsft_set_error("qsft_batch_pack_input: a digit of M, D or locq is outside [0, q)");
        return QSFT_ERR_ARG;
    }
    memset(base + l.mt, 0, l.kraw - l.mt);                         // Mt and D carry zero padding past n
    for (int c = 0; c < C; ++c)
        for (int i2 = 0; i2 < n; ++i2)
            for (int t = 0; t < b; ++t) Mt[((size_t)c * b + t) * np + i2] = M[((size_t)c * n + i2) * b + t];
    for (int p = 0; p < P; ++p) memcpy(Dd + (size_t)p * np, D + (size_t)p * n, (size_t)n);
    memcpy(base + l.kraw, locq, (size_t)n * S);                    // transposed on the device
    memcpy(base + l.a, a, (size_t)S * 16);
    B->sign[i] = delay_sign(D, r.q, n, r.R);
    return QSFT_OK;
}

// Every request's inputs, packed by up to `threads` host threads (requests
// are independent: each writes only its own block of the staging arena).
extern "C" int qsft_batch_pack_inputs(void* h, const int8_t* const* M, const int8_t* const* D,
                                      const int8_t* const* locq, const double* const* a, void* input_host,
                                      int threads) {
    Batch* B = as_batch(h);
    if (!B || !M || !D || !locq || !a || !input_host || threads < 1) {
        qsft_set_error("qsft_batch_pack_inputs: bad arguments");
        return QSFT_ERR_ARG;
    }
    const int n = (int)B->req.size();
    const int nt = threads < n ? threads : n;
    std::vector<int> rc(nt, QSFT_OK);
    auto work = [&](int w) {
        for (int i = w; i < n; i += nt) {
            const int r = qsft_batch_pack_input(h, i, M[i], D[i], locq[i], a[i], input_host);
            if (r != QSFT_OK) { rc[w] = r; return; }
        }
    };
    if (nt == 1) {
        work(0);
    } else {
        std::vector<std::thread> pool;
        pool.reserve(nt - 1);
        for (int w = 1; w < nt; ++w) pool.emplace_back(work, w);
        work(0);
        for (auto& t : pool) t.join();
    }
    for (int r : rc)
        if (r != QSFT_OK) return r;
    return QSFT_OK;
}

extern "C" int qsft_batch_bind(void* h, void* input_dev, void* work_dev, void* plan_host, void* plan_dev, void* out) {
    Batch* B = as_batch(h);
    if (!B || !input_dev || !work_dev || !plan_host || !plan_dev || !out ||
        (((uintptr_t)input_dev | (uintptr_t)work_dev | (uintptr_t)plan_dev | (uintptr_t)plan_host) & 255)) {
        qsft_set_error("qsft_batch_bind: bad arguments (arenas and plan buffers 256-byte aligned)");
        return QSFT_ERR_ARG;
    }
    if (!B->fork || !B->done) {                                    // created once, on the first bind
        if ((!B->fork && hipEventCreateWithFlags(&B->fork, hipEventDisableTiming) != hipSuccess) ||
            (!B->done && hipEventCreateWithFlags(&B->done, hipEventDisableTiming) != hipSuccess)) {
            qsft_set_error("qsft_batch_bind: hipEventCreate failed");
            return QSFT_ERR_HIP;
        }
    }
    B->in_dev = static_cast<char*>(input_dev);
    B->work_dev = static_cast<char*>(work_dev);
    B->plan_host = static_cast<char*>(plan_host);
    B->plan_dev = static_cast<char*>(plan_dev);
    B->out = out;
    const int rc = build_plan(*B, true);
    B->bound = rc == QSFT_OK;
    return rc;
}

namespace {

// Enqueue the pipeline. seed_ptr (device) replaces seed_offset when set;
// `capturing`: inside a stream capture (no wait on an event recorded outside
// it; the tail stream is joined back to `stream` before the capture ends).
int enqueue(Batch* B, int upload, const void* input_host, uint64_t seed_offset, const uint64_t* seed_ptr,
            hipStream_t st, hipStream_t tail, bool capturing) {
    hipError_t e;
    if (B->launched && !capturing) {                               // the last tail still reads the tables
        e = hipStreamWaitEvent(st, B->done, 0);
        if (e != hipSuccess) return qsft_hip_status(e, "qsft_batch_launch(wait)");
    }
    if (upload & 1) {
        e = hipMemcpyAsync(B->in_dev, input_host, B->sz.input_bytes, hipMemcpyHostToDevice, st);
        if (e != hipSuccess) return qsft_hip_status(e, "qsft_batch_launch(inputs)");
    }
    if (upload & 2) {
        e = hipMemcpyAsync(B->plan_dev, B->plan_host, B->sz.plan_bytes, hipMemcpyHostToDevice, st);
        if (e != hipSuccess) return qsft_hip_status(e, "qsft_batch_launch(plan)");
    }
    char* P = B->plan_dev;
    int rc = QSFT_OK;
    if (upload & 1) {
        hipLaunchKernelGGL(k_stage_support, dim3(B->stage_blocks), dim3(256), 0, st,
                           (const StageDesc*)(P + B->stage_desc), (const int32_t*)(P + B->stage_pre),
                           (int)B->req.size());
        e = hipGetLastError();
        if (e != hipSuccess) return qsft_hip_status(e, "qsft_batch_launch(stage)");
    }
    for (const MfmaLaunch& L : B->mfma) {
        rc = qsft_eval_mfma((const qsft_eval_mfma_desc*)(P + L.desc), (const int32_t*)(P + L.pre_prep),
                            (const int32_t*)(P + L.pre_main), L.n, L.prep_blocks, L.main_blocks, L.word, st);
        if (rc != QSFT_OK) return rc;
    }
    for (const Rows4Launch& L : B->rows4) {
        rc = qsft_dft4_rows((const qsft_dft_desc*)(P + L.desc), (const int32_t*)(P + L.pre), L.n, L.blocks, L.word, st);
        if (rc != QSFT_OK) return rc;
    }
    for (const DftLaunch& L : B->dft)
        for (int p = 0; p < L.npass; ++p) {
            rc = qsft_dft_pass((const qsft_dft_desc*)(P + L.desc), (const int32_t*)(P + L.pre[p]), L.n, L.blocks[p],
                               p, L.word, L.lds[p], st);
            if (rc != QSFT_OK) return rc;
        }
    const int n = (int)B->req.size();
    if (seed_ptr)
        rc = qsft_assemble((const qsft_assemble_desc*)(P + B->asm_desc), (const int32_t*)(P + B->asm_pre), n,
                           B->asm_blocks, (uint64_t)(uintptr_t)seed_ptr, B->asm_word | QSFT_ASM_SEED_PTR, st);
    else
        rc = qsft_assemble((const qsft_assemble_desc*)(P + B->asm_desc), (const int32_t*)(P + B->asm_pre), n,
                           B->asm_blocks, seed_offset, B->asm_word, st);
    if (rc != QSFT_OK) return rc;
    const qsft_peel_desc* pd = (const qsft_peel_desc*)(P + B->peel_desc);
    const int32_t* pb = (const int32_t*)(P + B->peel_pre_bins);
    const int32_t* ps = (const int32_t*)(P + B->peel_pre_slots);
    rc = qsft_peel_init(pd, pb, ps, n, B->peel_bins, B->peel_slots, st);
    for (int it = 0; rc == QSFT_OK && it < B->fuse_from; ++it)
        rc = qsft_peel_iteration(pd, pb, n, B->peel_bins, B->max_p, it, st);
    if (rc == QSFT_OK) rc = qsft_peel_passes(pd, n, B->max_p, B->fuse_from, B->max_iter, st);
    if (rc != QSFT_OK) return rc;
    hipStream_t ts = st;
    if (tail && tail != st) {
        ts = tail;
        e = hipEventRecord(B->fork, st);
        if (e == hipSuccess) e = hipStreamWaitEvent(ts, B->fork, 0);
        if (e != hipSuccess) return qsft_hip_status(e, "qsft_batch_launch(fork)");
    }
    rc = qsft_peel_finalize(pd, ps, n, B->peel_slots, 0, ts);
    if (rc == QSFT_OK) rc = qsft_peel_pack_ordered(pd, n, B->out, B->sz.cap_rows, B->sz.row_stride, ts);
    if (rc != QSFT_OK) return rc;
    e = hipEventRecord(B->done, ts);
    if (e == hipSuccess && capturing && ts != st) e = hipStreamWaitEvent(st, B->done, 0);   // join the tail
    if (e != hipSuccess) return qsft_hip_status(e, "qsft_batch_launch(done)");
    return QSFT_OK;
}

}  // namespace

extern "C" int qsft_batch_launch(void* h, int upload, const void* input_host, uint64_t seed_offset, void* stream,
                                 void* tail_stream) {
    Batch* B = as_batch(h);
    if (!B || !B->bound || (upload & ~3) || ((upload & 1) && !input_host)) {
        qsft_set_error("qsft_batch_launch: batch not bound or bad upload flags");
        return QSFT_ERR_ARG;
    }
    const int rc = enqueue(B, upload, input_host, seed_offset, nullptr, (hipStream_t)stream, (hipStream_t)tail_stream,
                           false);
    if (rc == QSFT_OK) B->launched = true;
    return rc;
}

extern "C" int qsft_batch_capture(void* h, void* stream, void* tail_stream) {
    Batch* B = as_batch(h);
    if (!B || !B->bound || !B->launched || !stream) {
        qsft_set_error("qsft_batch_capture: bind and launch the batch (with its upload) on a stream first");
        return QSFT_ERR_ARG;
    }
    hipError_t e = hipEventSynchronize(B->done);                   // inputs staged, last launch finished
    if (e != hipSuccess) return qsft_hip_status(e, "qsft_batch_capture(sync)");
    if (B->graph_exec) { (void)hipGraphExecDestroy(B->graph_exec); B->graph_exec = nullptr; }
    if (B->graph) { (void)hipGraphDestroy(B->graph); B->graph = nullptr; }
    hipStream_t st = (hipStream_t)stream;
    e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return qsft_hip_status(e, "qsft_batch_capture(begin)");
    uint64_t* ctr = reinterpret_cast<uint64_t*>(B->plan_dev + B->seed_ctr);
    hipLaunchKernelGGL(k_seed_bump, dim3(1), dim3(64), 0, st, ctr);
    int rc = qsft_hip_status(hipGetLastError(), "qsft_batch_capture(bump)");
    if (rc == QSFT_OK) rc = enqueue(B, 0, nullptr, 0, ctr, st, (hipStream_t)tail_stream, true);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(st, &g);
    if (rc != QSFT_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return qsft_hip_status(e, "qsft_batch_capture(end)");
    B->graph = g;
    e = hipGraphInstantiate(&B->graph_exec, g, nullptr, nullptr, 0);
    return qsft_hip_status(e, "qsft_batch_capture(instantiate)");
}

extern "C" int qsft_batch_replay(void* h, void* stream) {
    Batch* B = as_batch(h);
    if (!B || !B->graph_exec || !stream) {
        qsft_set_error("qsft_batch_replay: no captured graph");
        return QSFT_ERR_ARG;
    }
    hipStream_t st = (hipStream_t)stream;
    hipError_t e = hipStreamWaitEvent(st, B->done, 0);
    if (e == hipSuccess) e = hipGraphLaunch(B->graph_exec, st);
    if (e == hipSuccess) e = hipEventRecord(B->done, st);
    return qsft_hip_status(e, "qsft_batch_replay");
}

extern "C" int qsft_batch_done(void* h) {
    Batch* B = as_batch(h);
    if (!B) { qsft_set_error("qsft_batch_done: no batch"); return -QSFT_ERR_ARG; }
    if (!B->launched) return 1;
    const hipError_t e = hipEventQuery(B->done);
    if (e == hipSuccess) return 1;
    if (e == hipErrorNotReady) return 0;
    return -qsft_hip_status(e, "qsft_batch_done");
}

extern "C" int qsft_batch_wait(void* h) {
    Batch* B = as_batch(h);
    if (!B) { qsft_set_error("qsft_batch_wait: no batch"); return QSFT_ERR_ARG; }
    if (!B->launched) return QSFT_OK;
    return qsft_hip_status(hipEventSynchronize(B->done), "qsft_batch_wait");
}

extern "C" int qsft_batch_delay_sign(void* h, int i) {
    Batch* B = as_batch(h);
    if (!B || i < 0 || i >= (int)B->req.size()) return 0;
    return B->sign[i];
}

extern "C" void qsft_batch_free(void* h) {
    Batch* B = as_batch(h);
    if (!B) return;
    if (B->launched) (void)hipEventSynchronize(B->done);
    if (B->graph_exec) (void)hipGraphExecDestroy(B->graph_exec);
    if (B->graph) (void)hipGraphDestroy(B->graph);
    if (B->fork) (void)hipEventDestroy(B->fork);
    if (B->done) (void)hipEventDestroy(B->done);
    delete B;
}
