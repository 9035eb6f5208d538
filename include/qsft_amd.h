/*
 * qsft_amd.h -- C ABI of the MI355X (gfx950) q-SFT engine.
 *
 * Every entry point takes plain device pointers, sizes and a HIP stream
 * (passed as void*, NULL = the null stream); nothing here allocates device
 * memory. Host code (the qsft_amd Python package, or any FFI) owns the
 * buffers -- PyTorch's caching allocator in the shipped binding.
 *
 * Batched launches: each launch processes n_inst independent instances
 * ("uneven" shapes allowed). The caller uploads an array of descriptors and a
 * block-prefix array pre[0..n_inst] (pre[i] = first work block of instance i,
 * pre[n_inst] = total blocks, computed with the *_blocks() helpers). A work
 * block finds its instance by binary search over pre[].
 *
 * The reference (basics-lab/qsft) is pure Python/numpy and has no native
 * layer; each entry point names the reference function whose work it
 * replaces (paths relative to the reference repository root).
 *
 * Return value: QSFT_OK or an error code; qsft_last_error() describes the
 * last failure on the calling thread.
 */
#ifndef QSFT_AMD_H
#define QSFT_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QSFT_ABI_VERSION 2
#define QSFT_MAX_Q 64          /* alphabet size limit (digits are int8, w-tables live in LDS) */
#define QSFT_MAX_N 256         /* digit count limit (n_pad <= 256)                              */
#define QSFT_MAX_B 16          /* sub-sampling dimension limit                                  */
#define QSFT_MAX_C 16          /* groups per instance                                           */
#define QSFT_MAX_R 16          /* repeats per instance                                          */
#define QSFT_DFT_MAX_PASSES 4

enum { QSFT_OK = 0, QSFT_ERR_ARG = 1, QSFT_ERR_HIP = 2, QSFT_ERR_DEVICE = 3 };
enum { QSFT_F32 = 0, QSFT_F64 = 1 };
enum { QSFT_DETECT_NOISELESS = 0, QSFT_DETECT_NSO1 = 1 };
/* method bit: coded (Reed-Solomon) source delays; the channel detection
 * (NOISELESS or NSO1) yields p1 - 1 = 2 t s syndrome digits, decoded into k
 * by the device syndrome decoder (qsft_peel_set_coded supplies the code). */
#define QSFT_DETECT_CODED 2
enum { QSFT_NOISE_NONE = 0, QSFT_NOISE_ARRAY = 1, QSFT_NOISE_PHILOX = 2 };

/* device-side error bits reported in qsft_peel_desc::ctr[QSFT_CTR_ERROR] */
enum { QSFT_DEVERR_TABLE_FULL = 1, QSFT_DEVERR_TAG_COLLISION = 2, QSFT_DEVERR_OUT_FULL = 4,
       QSFT_DEVERR_BAD_SHAPE = 8 };

const char* qsft_last_error(void);
int qsft_abi_version(void);
/* sizeof() of the descriptor structs, for FFI layout checks:
 * which = 0 eval, 1 eval_points, 2 dft, 3 assemble, 4 peel, 5 eval_mfma, 6 request, 7 batch sizes. */
size_t qsft_sizeof(int which);

/* ------------------------------------------------------------------------
 * Stage (a)+(b): query digits M_c l + d_p mod q generated in registers and
 * the sparse signal x(m) = sum_s a_s w^{<m, k_s>} evaluated at every query,
 * with the phase <m, k_s> from an int8 MFMA contraction (exact int32), then
 * mod q and a q-entry w-table, accumulated in FP32 or FP64.
 *
 * Replaces SubsampledSignal._get_qsft_query_indices
 * (qsft/input_signal_subsampled.py:183-206) followed by
 * SyntheticSubsampledSignal.subsample (synt_exp/synt_src/synthetic_signal.py:97-118)
 * for every (group, repeat) pair of _subsample_qsft
 * (qsft/input_signal_subsampled.py:120-147).
 * ------------------------------------------------------------------------ */
typedef struct qsft_eval_desc {
    int32_t q, n, b;        /* alphabet, digits, sub-sampling dimension                     */
    int32_t C, P;           /* groups, delay rows per group (all repeats)                  */
    int32_t S;              /* rows of K / entries of a                                     */
    int32_t n_pad;          /* row stride (bytes) of Mt, D, K; multiple of 64, >= n         */
    int32_t _pad0;
    const int8_t* Mt;       /* [C][b][n_pad]   Mt[c][t][i] = M_c[i][t]                       */
    const int8_t* D;        /* [C][P][n_pad]   delay rows                                    */
    const int8_t* K;        /* [S][n_pad]      support digits k_s (entries >= n are zero)    */
    const void* a;          /* [S]             strengths, float2 (F32) or double2 (F64)      */
    void* x;                /* [C][P][q^b]     samples, float2 or double2, l MSB-first       */
} qsft_eval_desc;

int qsft_eval_blocks(const qsft_eval_desc* host_desc);
int qsft_eval_cosets(const qsft_eval_desc* dev_descs, const int32_t* dev_pre, int n_inst,
                     int total_blocks, int precision, void* stream);

/* The same samples on the matrix cores (FP32 result). The phase splits as
 * <(M_c l + d_p) mod q, k_s> = <(M_c l) mod q, k_s> + <d_p, k_s> (mod q), so
 * x[c][p][l] = sum_s W_c[l][s] B_c[s][p] with W_c = w^{E0} (E0 from an int8
 * MFMA over the q^b base queries) and B_c[s][p] = a_s w^{<d_p, k_s>}: a complex
 * GEMM on fp16 MFMA with 2^bexp B split into two fp16 limbs and FP32
 * accumulation (x is scaled back by 2^-bexp).
 * A prep kernel writes B_c into `Bf` (qsft_eval_mfma_workspace_bytes(C, P, S)
 * bytes) before the main kernel. Replaces the same reference code as
 * qsft_eval_cosets. */
typedef struct qsft_eval_mfma_desc {
    int32_t q, n, b, C, P, S;
    int32_t n_pad;          /* multiple of 64, >= n                                         */
    int32_t nsb;            /* (S + 15) / 16  support blocks                                */
    int32_t npt;            /* (P + 7) / 8    delay-row tiles                               */
    int32_t shared_d;       /* 1: every group has the same delay rows (B computed once)     */
    int32_t bexp;           /* 14 - e where max|a_s| = m 2^e, m in [0.5, 1): the largest    *
                             * scaled |B| lies in [2^13, 2^14); in [-100, 100]               */
    int32_t ptb;            /* p-tiles per workgroup: 5..8 for q in {2, 4} with n_pad = 64,  *
                             * 8 otherwise (same for every instance of a launch)           */
    const int8_t* Mt;       /* [C][b][n_pad]                                                */
    const int8_t* D;        /* [C][P][n_pad] ([P][n_pad] when shared_d), 16-byte aligned    */
    const int8_t* K;        /* [S][n_pad], 16-byte aligned                                  */
    const double* a;        /* [S][2] strengths, complex FP64, 16-byte aligned              */
    uint16_t* Bf;           /* workspace: B fragments, qsft_eval_mfma_workspace_bytes(shared_d ? 1 : C, P, S) */
    void* x;                /* [C][P][q^b] complex FP32 samples out                         */
} qsft_eval_mfma_desc;

size_t qsft_eval_mfma_workspace_bytes(int C, int P, int S);
int qsft_eval_mfma_blocks(const qsft_eval_mfma_desc* host_desc, int* prep_blocks, int* main_blocks);
/* word: bits 4..7 = n_pad / 64, bits 8..9 = 0 for q in {2, 4}, 1 for other powers of two, 2 otherwise;
 * one word per launch. */
int qsft_eval_mfma(const qsft_eval_mfma_desc* dev_descs, const int32_t* dev_pre_prep, const int32_t* dev_pre_main,
                   int n_inst, int prep_blocks, int main_blocks, int word, void* stream);

/* Stage (a) alone: the digits (M_c l + d_p) mod q of every query, written as
 * int8 [C][P][q^b][n_pad] to desc->x (a, K and S are ignored). Replaces the
 * index construction of _get_qsft_query_indices
 * (qsft/input_signal_subsampled.py:198-202) for user subsample() callbacks. */
int qsft_query_digits_blocks(const qsft_eval_desc* host_desc);
int qsft_query_digits(const qsft_eval_desc* dev_descs, const int32_t* dev_pre, int n_inst, int total_blocks,
                      void* stream);

/* Same evaluation at explicit query digits (the generic query path, e.g. the
 * held-out evaluation queries of qsft/test_helper.py:235-260). */
typedef struct qsft_eval_points_desc {
    int32_t q, n, S, n_pad;
    int64_t N;              /* number of queries                                          */
    const int8_t* m;        /* [N][n_pad] query digits                                     */
    const int8_t* K;        /* [S][n_pad]                                                  */
    const void* a;          /* [S] float2 or double2                                       */
    void* x;                /* [N] float2 or double2                                       */
} qsft_eval_points_desc;

int qsft_eval_points_blocks(const qsft_eval_points_desc* host_desc);
int qsft_eval_points(const qsft_eval_points_desc* dev_descs, const int32_t* dev_pre, int n_inst,
                     int total_blocks, int precision, void* stream);

/* ------------------------------------------------------------------------
 * Stage (c): forward-scaled b-dimensional q-point DFT of every row,
 * U[j] = q^-b sum_l x[l] w^{-<j,l>}, LDS-staged, in up to 4 passes over
 * groups of axes. Replaces SubsampledSignal._compute_subtransform
 * (qsft/input_signal_subsampled.py:150-153, 264-266) -> gwht
 * (qsft/utils.py:31-36, scipy.fft.fftn / q^b). `in_stride` > 1 reads every
 * in_stride-th sample, the row[::q^(b_max - b)] of the all_bs loop.
 * ------------------------------------------------------------------------ */
typedef struct qsft_dft_desc {
    int32_t q, b;           /* DFT along b axes of size q                                  */
    int32_t rows;           /* independent rows                                             */
    int32_t in_stride;      /* element stride inside an input row                          */
    int64_t in_row_stride;  /* elements between input rows                                 */
    const void* in;         /* input samples                                                */
    void* out;              /* [rows][q^b] output (passes after the first run in place)    */
    /* filled by qsft_dft_plan(): */
    int32_t n_pass;
    int32_t pass_d0[QSFT_DFT_MAX_PASSES];   /* first axis of the pass                       */
    int32_t pass_d1[QSFT_DFT_MAX_PASSES];   /* one past the last axis                       */
    int32_t pass_ti[QSFT_DFT_MAX_PASSES];   /* contiguous inner elements per tile (divides I) */
    int32_t pass_blocks[QSFT_DFT_MAX_PASSES];
    int32_t pass_ta[QSFT_DFT_MAX_PASSES];   /* whole [M, I] blocks per tile (pass_ti = I only) */
} qsft_dft_desc;

/* Fills the pass fields: the schedule qsft_dft_autotune measured fastest for
 * (q, b, precision) in this process, else the planner's model-ranked best
 * (every split of the b axes into at most 4 passes, under several LDS budgets
 * and inner-run lengths). Every schedule applies the same axis transforms in
 * the same order (innermost axis first), so results do not depend on it. */
int qsft_dft_plan(qsft_dft_desc* host_desc, int precision);
/* The planner's candidate `index` (0 = the model's best, in the order the
 * autotuner tries them); QSFT_ERR_ARG past the last. */
int qsft_dft_plan_candidate(qsft_dft_desc* host_desc, int precision, int index);
/* Autotuner: times the model's best-ranked schedules (at most max_candidates)
 * on `rows` rows held in `buf` (rows * q^b elements + 4096 bytes, device
 * memory, overwritten) on `stream`, synchronously, and records the fastest
 * for qsft_dft_plan; *best_ms = its time for the rows. Call once per shape
 * before the launches that should use it (not under graph capture). */
int qsft_dft_autotune(int q, int b, int precision, int rows, void* buf, size_t buf_bytes, int max_candidates,
                      void* stream, float* best_ms);
/* The schedule qsft_dft_plan returns for (q, b, precision): pass count, axis
 * groups [d0, d1), TI and TA per pass (arrays of 4); *tuned = 1 if measured. */
int qsft_dft_schedule(int q, int b, int precision, int* n_pass, int* d0, int* d1, int* ti, int* ta, int* tuned);
/* Dynamic LDS bytes one instance's pass needs (tile + twiddles, <= 64 KiB). */
int qsft_dft_lds_bytes(const qsft_dft_desc* host_desc, int pass, int precision);
/* lds_bytes: dynamic LDS of the launch, at least qsft_dft_lds_bytes() of every
 * instance's pass; smaller tiles leave more workgroups resident per CU. */
int qsft_dft_pass(const qsft_dft_desc* dev_descs, const int32_t* dev_pre, int n_inst,
                  int total_blocks, int pass, int precision, int lds_bytes, void* stream);
/* q = 4 and b in {4, 6} with in_stride 1 (the served sub-transforms): one
 * workgroup per row (dev_pre counts rows), the axis transforms two digits at a
 * time in registers, one launch instead of qsft_dft_pass's passes. word: low
 * nibble QSFT_F32 / QSFT_F64, bits 8..15 = b (all instances of the launch).
 * Same reference code as qsft_dft_pass. */
int qsft_dft4_rows(const qsft_dft_desc* dev_descs, const int32_t* dev_pre, int n_inst, int total_blocks,
                   int word, void* stream);

/* ------------------------------------------------------------------------
 * Working transform for one QSFT.transform call: gather the selected groups
 * and repeats of the stored sub-transforms, promote to FP64 and add the
 * transform-domain noise CN(0, sigma^2) (sigma^2 = noise_sd^2 / q^b).
 * Replaces SubsampledSignal.get_MDU (qsft/input_signal_subsampled.py:225-262),
 * the noise of SyntheticSubsampledSignal.get_MDU
 * (synt_exp/synt_src/synthetic_signal.py:120-130) and the vstack/np.array of
 * qsft/qsft.py:114-120.
 * ------------------------------------------------------------------------ */
typedef struct qsft_assemble_desc {
    int32_t Q;              /* points per row                                               */
    int32_t C, R, p1;       /* working groups, working repeats, rows per repeat             */
    int32_t src_rows;       /* rows per group in the source (= stored repeats * p1)        */
    int32_t src_prec;       /* QSFT_F32 / QSFT_F64                                          */
    int32_t noise_mode;     /* QSFT_NOISE_*                                                 */
    int32_t bin_major;      /* 1: write U as [C][Q][R*p1] (the peeler's layout)             */
    int32_t gsel[QSFT_MAX_C];   /* source group of working group c                         */
    int32_t rsel[QSFT_MAX_R];   /* source repeat of working repeat r                       */
    const void* src;        /* [C_src][src_rows][Q]                                         */
    const double* noise;    /* QSFT_NOISE_ARRAY: [C][R*p1][Q] complex (re, im) doubles     */
    double sigma;           /* QSFT_NOISE_PHILOX: per-component sd = sqrt(sigma^2 / 2)     */
    uint64_t seed;          /* QSFT_NOISE_PHILOX key                                        */
    double* U;              /* [C][R*p1][Q] (or [C][Q][R*p1]) complex FP64 out              */
    double* energy;         /* optional [C][Q]: |U column|^2 of every bin as the peeler     *
                             * computes it (bin-major and R*p1 <= 128 only), for its      *
                             * first-pass zeroton screen; NULL: not written                */
    uint8_t* written;       /* optional [C][Q] (needs energy): write a bin's column only    *
                             * if its energy exceeds cutoff * R*p1 (the peeler's zeroton   *
                             * test) and store 1 / 0 here; NULL: every column is written   */
    double cutoff;          /* with `written`: the peeler's cutoff                          */
    uint64_t* seed_out;     /* optional: the Philox key of the launch (seed + seed_offset)  */
} qsft_assemble_desc;

int qsft_assemble_blocks(const qsft_assemble_desc* host_desc);
/* Launch word bits of one instance; the launch takes the AND over its
 * instances. QSFT_ASM_F32_TILE: FP32 source and no noise array, so the
 * bin-major tile may be held in FP32 (half the LDS per workgroup). */
#define QSFT_ASM_F32_TILE 1
/* Launch word bit: seed_offset holds the device address of a uint64 offset,
 * read when the kernel runs (a captured graph replayed with fresh noise). */
#define QSFT_ASM_SEED_PTR 2
int qsft_assemble_word(const qsft_assemble_desc* host_desc);
/* seed_offset is added to every descriptor's Philox seed, so one uploaded
 * batch can draw fresh noise on every launch. */
int qsft_assemble(const qsft_assemble_desc* dev_descs, const int32_t* dev_pre, int n_inst,
                  int total_blocks, uint64_t seed_offset, int word, void* stream);

/* ------------------------------------------------------------------------
 * Stage (d): singleton detection + peeling, device resident.
 * Replaces the while-loop of QSFT.transform (qsft/qsft.py:142-255) with
 * singleton_detection (qsft/reconstruct.py:12-31, 100-113, 132-167).
 * U is bin-major: [C][Q][P] complex FP64, one bin's P rows contiguous.
 * A screening kernel stamps the non-zeroton bins for the first pass. Per
 * iteration: kernel A (one wave per 64 consecutive bins) scans the frozen U --
 * the stamped bins: on later passes those that were non-zeroton on the
 * previous pass or were peeled into (an unchanged zeroton stays a zeroton) --
 * classifies zeroton / singleton / multiton (k, rho, bin match,
 * residual) and registers the wave's singletons in a device hash table keyed
 * by k, one lane per singleton; kernel B (same bin partition) queues each
 * distinct k, with the rho of its last (group, bin), on its bin in every
 * group, and the next pass's kernel A subtracts the queued balls from the
 * columns it owns before classifying them (no FP64 atomics). The stop rule (no singletons or
 * no multitons, the pass still peels), the q^n ball cap and the iteration
 * cap run on the device: no host round trip. The final kernel forms the
 * append-order running mean of every detection of each k
 * (qsft/qsft.py:247-255).
 * ------------------------------------------------------------------------ */
typedef struct qsft_peel_desc {
    int32_t q, n, b, C, P, Q;
    int32_t n_pad;          /* stride of Mt/D rows and of per-bin k digits (multiple of 64) */
    int32_t p1;             /* source parity: rows per repeat block (n + 1)                 */
    int32_t method;         /* QSFT_DETECT_*                                                */
    int32_t max_iter;       /* 15 in the reference (at most 15)                             */
    int32_t hcap_log2;      /* log2 of hash slots H                                         */
    int32_t out_cap;        /* capacity of the output arrays                                */
    int32_t flags;          /* QSFT_PEEL_* bits, set by the caller after qsft_peel_bind()   */
    double cutoff;          /* threshold per row: |col|^2 > cutoff * P                      */
    double peel_cap;        /* q^n (as double, saturating): cap on peeled balls            */
    const int8_t* Mt;       /* [C][b][n_pad]                                                */
    const int8_t* D;        /* [C][P][n_pad] ([P][n_pad] with QSFT_PEEL_SHARED_D)           */
    double* U;              /* [C][Q][P] complex FP64 (bin-major), modified in place       */
    const double* energy;   /* optional [C][Q] first-pass |col|^2 from qsft_assemble (set   *
                             * after qsft_peel_bind); NULL: the screen reads U             */
    /* workspace, carved by qsft_peel_bind() */
    int8_t* kbin;           /* [2][C*Q][n_pad] k of singleton bins, by pass parity          */
    double* rho_bin;        /* [2][C*Q][2]     rho of singleton bins, by pass parity        */
    int32_t* slot_bin;      /* [C*Q]         hash slot of a singleton bin, -1 otherwise     */
    int32_t* mark;          /* [C*Q]         pass that scans the bin next (-1: none)         */
    int32_t* sstamp;        /* [C*Q]         pass + 1 in which the bin was a singleton      */
    int32_t* pend_head;     /* [C*Q]         balls to peel from the bin (list head, -1: none) */
    int32_t* pend_next;     /* [C*Q][C]      list links; entry = source bin * C + group     */
    uint64_t* tag;          /* [H] 0 = empty                                                */
    uint64_t* owner;        /* [H] first (iteration, bin) of the key                        */
    uint64_t* last;         /* [H] last (iteration, bin) of the key                         */
    uint32_t* head;         /* [H] head of the key's record list                            */
    uint32_t* count;        /* [H] records of the key                                       */
    uint32_t* live;         /* [H] claimed slots, in claim order (QSFT_CTR_LIVE of them)    */
    int8_t* key;            /* [H][n_pad] digits of the key, written by its first record    */
    uint64_t* rec_time;     /* [L] (iteration + 1) << 32 | bin                              */
    double* rec_rho;        /* [L][2]                                                       */
    uint32_t* rec_next;     /* [L]                                                          */
    uint32_t* ctr;          /* counters, QSFT_CTR_* layout                                  */
    /* outputs */
    int8_t* out_k;          /* [out_cap][n_pad] recovered k                                 */
    double* out_val;        /* [out_cap][2]     running-mean coefficient                    */
    uint64_t* out_order;    /* [out_cap]        first (iteration, bin) of k: dict order     */
    /* sparse working transform (QSFT_PEEL_SPARSE_U; the caller sets these after
     * qsft_peel_bind and points the assembler's written / seed_out at `written`
     * / `noise_key`): U holds only the columns the assembler wrote; a column it
     * skipped (a first-pass zeroton) is rebuilt from the samples and the same
     * Philox noise when a ball first lands on it, and written back then. */
    const void* xsrc;       /* [C][P][Q] complex FP32 samples, the assembler's source       */
    double noise_sigma;     /* the assembler's sigma (0: no noise)                          */
    uint8_t* written;       /* [C*Q] 1: the bin's column is in U (workspace)                */
    uint64_t* noise_key;    /* the assembler's Philox key of this launch (workspace)        */
    /* coded source (method & QSFT_DETECT_CODED), set by qsft_peel_set_coded */
    const int16_t* rs_tab;  /* GF(q^s) tables: exp[2 (q^s - 1)] then log[q^s] (log 0 = -1)   */
    int32_t rs_t, rs_s;     /* correctable errors t, symbol digits s (p1 - 1 = 2 t s)        */
    int32_t rs_nc, _pad1;   /* code length q^s - 1                                          */
} qsft_peel_desc;

/* qsft_peel_desc::flags */
#define QSFT_PEEL_SHARED_D 1    /* every group uses the delay rows of group 0            */
#define QSFT_PEEL_UNIT_D 2      /* delay row i of every repeat block r is row 0 of the   *
                                 * block + s e_(i-1) (mod q): the NSO delays (s = -1)     *
                                 * and the identity delays (s = +1); signatures then take *
                                 * R dot products per key instead of P                   */
#define QSFT_PEEL_UNIT_D_NEG 4  /* with QSFT_PEEL_UNIT_D: s = -1                          */
#define QSFT_PEEL_ONE_BIN_WAVES 16 /* first pass with one bin per wave (default: four   *
                                  * bins per wave where possible; same results)          */
#define QSFT_PEEL_SPARSE_U 8    /* U holds only the columns the assembler wrote (above;  *
                                 * ignored unless xsrc, written and noise_key are set)   */

/* counter layout (uint32 words); per iteration i at 8 i + :
 *   0 singletons, 1 multitons, 2 distinct balls, 3 active flag,
 *   4 bins scanned in pass i */
#define QSFT_CTR_ITER_STRIDE 8
#define QSFT_CTR_LOG 128        /* records appended                                    */
#define QSFT_CTR_OUT 129        /* outputs written                                     */
#define QSFT_CTR_ERROR 130      /* QSFT_DEVERR_* bits                                  */
#define QSFT_CTR_ITERS 131      /* iterations that ran                                 */
#define QSFT_CTR_PEELED 132     /* balls peeled (saturating)                           */
#define QSFT_CTR_LIVE 133       /* hash slots claimed (keys)                           */
#define QSFT_CTR_WORDS 136

size_t qsft_peel_workspace_bytes(int q, int n, int b, int C, int P, int max_iter, int out_cap);
/* Fill the shape fields and carve `workspace` into the pointer fields. */
int qsft_peel_bind(qsft_peel_desc* host_desc, int q, int n, int b, int C, int P, int p1,
                   int method, int max_iter, double cutoff, int out_cap,
                   const int8_t* Mt, const int8_t* D, double* U,
                   void* workspace, size_t workspace_bytes, int n_pad);
/* Coded source delays (qsft/query.py:88-92, qsft/reconstruct.py:34-51,
 * qsft/ReedSolomon.py:27-50): the shortened Reed-Solomon code over GF(q^s)
 * whose parity-check rows H[j][i] = alpha^{(1+j)(n-1-i)} give the delay rows.
 * The detection kernels decode each bin's 2t syndrome symbols by
 * Berlekamp-Massey, a Chien search over all q^s - 1 positions and Forney's
 * formula; a failed decode (locator roots != its degree, or an error value
 * outside GF(q)) gives k = 0. `tables` (device) must stay valid while the
 * descriptor is used. Requires method & QSFT_DETECT_CODED and p1 - 1 = 2 t s. */
int qsft_peel_set_coded(qsft_peel_desc* host_desc, int t, int s, const int16_t* tables);
int qsft_peel_bin_blocks(const qsft_peel_desc* host_desc);   /* blocks of the per-bin kernels  */
int qsft_peel_slot_blocks(const qsft_peel_desc* host_desc);  /* blocks of the per-slot kernels */

/* Whole loop: init, max_iter x (detect, peel), finalize -- no host sync.
 * max_p: the largest P over the instances (<= 256); P <= 128 selects the
 * detection kernel that holds two column rows per lane instead of four. */
int qsft_peel_run(const qsft_peel_desc* dev_descs, const int32_t* dev_pre_bins,
                  const int32_t* dev_pre_slots, int n_inst, int blocks_bins, int blocks_slots,
                  int max_p, int max_iter, void* stream);
/* The same steps one at a time, for a host-driven loop (the plain path).
 * qsft_peel_init clears the tables and screens every bin of the first pass
 * (|col|^2 > cutoff * P): only non-zeroton bins are stamped for pass 0. */
int qsft_peel_init(const qsft_peel_desc* dev_descs, const int32_t* dev_pre_bins, const int32_t* dev_pre_slots,
                   int n_inst, int blocks_bins, int blocks_slots, void* stream);
int qsft_peel_iteration(const qsft_peel_desc* dev_descs, const int32_t* dev_pre_bins, int n_inst,
                        int blocks_bins, int max_p, int iteration, void* stream);
/* Passes it0 .. it_end - 1 of the same loop (qsft/qsft.py:142-255) in one
 * launch, one workgroup per instance, for the tail of the loop where most
 * instances have stopped and the rest touch few bins per pass: the same
 * results as qsft_peel_iteration for each of those passes, without a launch
 * per pass. A loop that stops leaves at once; 0 <= it0 <= it_end <= 15. */
int qsft_peel_passes(const qsft_peel_desc* dev_descs, int n_inst, int max_p, int it0, int it_end, void* stream);
/* word: QSFT_FIN_FLUSH first subtracts the balls the last pass queued, so U
 * ends as the reference's loop leaves it; without it U keeps them queued
 * (the recovered rows are the same either way). With QSFT_PEEL_SPARSE_U the
 * columns that were never written stay unwritten (samples plus noise). */
#define QSFT_FIN_FLUSH 1
int qsft_peel_finalize(const qsft_peel_desc* dev_descs, const int32_t* dev_pre_slots, int n_inst,
                       int blocks_slots, int word, void* stream);

/* Gather every instance's recovered rows into one buffer `dst` (device or
 * host-mapped) so a batch comes back in one copy: an int64 header of
 * qsft_peel_pack_header_bytes(n_inst) bytes holding counts[n_inst],
 * offsets[n_inst], total, error bits, the most passes any instance ran, and
 * the number of instances whose loop asked for a pass beyond those launched
 * (a caller that launched fewer than max_iter passes runs the rest and
 * finalizes and packs again); then rows of row_stride bytes:
 * k digits (the instance's n_pad bytes), value (2 doubles), first detection
 * (uint64). Rows beyond cap_rows are dropped and flagged QSFT_DEVERR_OUT_FULL.
 * Replaces the result dict assembly of qsft/qsft.py:247-255. */
size_t qsft_peel_pack_header_bytes(int n_inst);
int qsft_peel_pack(const qsft_peel_desc* dev_descs, int n_inst, void* dst, int64_t cap_rows, int row_stride,
                   void* stream);

/* The same rows, each instance's in first-detection order (the reference's
 * dict insertion order, qsft/qsft.py:247-255): a row's position is its
 * instance offset plus its rank by the order word, so readers need no sort. */
int qsft_peel_pack_ordered(const qsft_peel_desc* d, int n_inst, void* dst, int64_t cap_rows, int row_stride,
                           void* stream);

/* ------------------------------------------------------------------------
 * Batch runtime: whole transforms of many synthetic requests, planned and
 * enqueued natively. Replaces, per request, what one
 * QSFT(**qsft_args).transform(get_random_subsampled_signal(...)) call does:
 * SubsampledSignal._subsample_qsft (qsft/input_signal_subsampled.py:107-155,
 * with SyntheticSubsampledSignal.subsample, synt_exp/synt_src/synthetic_signal.py:97-118),
 * get_MDU + noise (qsft/input_signal_subsampled.py:225-262,
 * synt_exp/synt_src/synthetic_signal.py:120-130) and the peeling loop with its
 * result dict (qsft/qsft.py:98-281) -- for all groups, repeats and b of one
 * request as the reference's loops do one by one, and for all requests of a
 * batch in the same launches (the reference serves one request per call).
 * Samples and sub-DFTs in FP32 on the matrix cores, working transform, peeling
 * and coefficients in FP64; every group shares the request's delay rows.
 *
 * Usage: qsft_batch_create (sizes) -> per request qsft_batch_pack_input into a
 * host staging arena -> qsft_batch_bind to device arenas of the returned sizes,
 * a host plan buffer and a device plan buffer (all 256-byte aligned) and the
 * packed output (pinned host memory or device memory; the qsft_peel_pack
 * format, rows ordered as qsft_peel_pack_ordered writes them) ->
 * qsft_batch_launch (upload bit 0: copy the staged inputs, bit 1: the plan;
 * later launches of the same batch may skip both and only draw fresh noise
 * with a new seed_offset) -> qsft_batch_done / qsft_batch_wait -> read `out`.
 * A launch first waits for the batch's previous launch to finish. With a
 * tail stream, finalize and pack run there after the peeling passes, so the
 * next batch's work on `stream` overlaps them.
 * ------------------------------------------------------------------------ */
typedef struct qsft_request {
    int32_t q, n, b, C, R, S;   /* P = R (n + 1) delay rows, shared by the C groups            */
    int32_t method;             /* QSFT_DETECT_NSO1, or QSFT_DETECT_NOISELESS with R = 1        */
    int32_t _pad0;
    double cutoff;              /* zeroton threshold per row (reference: 1e-9 + 1.5 noise_sd^2 / q^b) */
    double sigma;               /* transform-domain noise sd per bin (noise_sd / sqrt(q^b)); 0: none */
    double amax;                /* max |a_s| (> 0), or 0: computed from the strengths when the   *
                                 * request is packed (qsft_batch_pack_input)                    */
    uint64_t seed;              /* Philox key of the request's noise                            */
} qsft_request;

typedef struct qsft_batch_sizes {
    size_t input_bytes;         /* inputs arena: Mt, D, K, a of every request                   */
    size_t work_bytes;          /* work arena: samples, working transform, energies, B fragments, peel workspace */
    size_t plan_bytes;          /* descriptors and block prefixes (host plan and device plan)  */
    size_t out_bytes;           /* packed results: header + cap_rows rows of row_stride bytes   */
    int64_t cap_rows;
    int32_t row_stride;
    int32_t n_req;
} qsft_batch_sizes;

/* Device bytes one request holds while it runs (inputs + work): the admission unit. 0: bad shape. */
size_t qsft_request_footprint(const qsft_request* r, int max_iter);
/* NULL on a bad request (qsft_last_error says which). fuse_from: first peeling pass of the fused tail. */
void* qsft_batch_create(const qsft_request* reqs, int n_req, int max_iter, int fuse_from, qsft_batch_sizes* sizes);
/* Byte offsets of request i's Mt, D, K and a inside the inputs arena. */
int qsft_batch_input_offset(void* batch, int i, size_t* offsets4);
/* Request i's inputs into the host staging arena: M [C][n][b], D [R (n + 1)][n], locq [n][S] (int8 digits
 * in [0, q), else QSFT_ERR_ARG) and a [S][2] (complex128). Also detects the identity / NSO delay structure
 * (qsft_batch_delay_sign). */
int qsft_batch_pack_input(void* batch, int i, const int8_t* M, const int8_t* D, const int8_t* locq,
                          const double* a, void* input_host);
/* Every request at once (arrays of n_req pointers), on up to `threads` host threads. */
int qsft_batch_pack_inputs(void* batch, const int8_t* const* M, const int8_t* const* D, const int8_t* const* locq,
                           const double* const* a, void* input_host, int threads);
int qsft_batch_delay_sign(void* batch, int i);
int qsft_batch_bind(void* batch, void* input_dev, void* work_dev, void* plan_host, void* plan_dev, void* out);
int qsft_batch_launch(void* batch, int upload, const void* input_host, uint64_t seed_offset, void* stream,
                      void* tail_stream);
/* hipGraph of the batch's relaunch (no upload): the whole pipeline, tail
 * stream included, captured once on `stream` (after a launch that uploaded
 * the inputs; waits for it). Each replay first bumps a device counter that
 * the assembler adds to the Philox seeds, so replay k draws the noise of an
 * eager qsft_batch_launch with seed_offset = k (k = 1, 2, ... since capture). */
int qsft_batch_capture(void* batch, void* stream, void* tail_stream);
int qsft_batch_replay(void* batch, void* stream);
/* 1: the last launch's rows are in `out`; 0: not yet; < 0: error. */
int qsft_batch_done(void* batch);
int qsft_batch_wait(void* batch);
void qsft_batch_free(void* batch);

/* ------------------------------------------------------------------------
 * No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops and harnesses that
 * forbid hipMalloc): one instance's whole peeling loop, the same work and
 * results as qsft_peel_bind + qsft_peel_run + qsft_peel_pack(n_inst = 1),
 * with the descriptor, the block prefixes and the peel workspace held in a
 * static device array of the library instead of caller memory. The caller
 * passes only the inputs (Mt, D as for qsft_peel_bind, U bin-major FP64,
 * modified in place) and the packed output `dst` (header + rows, as
 * qsft_peel_pack writes them). flags: QSFT_PEEL_SHARED_D / UNIT_D /
 * UNIT_D_NEG (no sparse U, no assembler energies).
 * Limit: qsft_peel_workspace_bytes(q, n, b, C, P, max_iter, 0) must be at most
 * QSFT_NOALLOC_SCRATCH_BYTES - 768 (64 MiB scratch: covers C = 3 groups of
 * q^b = 4096 bins at n = 50 or 8000 bins at n = 16; larger shapes get
 * QSFT_ERR_ARG). Each GPU holds its own scratch copy from the first call on
 * it until the process ends. Every call on one GPU reuses that GPU's copy
 * (the loop's init kernel clears what must start empty), so calls that
 * overlap on one GPU -- on different streams or threads -- race: serialize
 * them, e.g. on one stream.
 * ------------------------------------------------------------------------ */
#define QSFT_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
size_t qsft_noalloc_scratch_bytes(void);
int qsft_peel_transform_noalloc(int q, int n, int b, int C, int P, int p1, int method, int max_iter,
                                double cutoff, int flags, const int8_t* Mt, const int8_t* D, double* U,
                                int n_pad, void* dst, int64_t cap_rows, int row_stride, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* QSFT_AMD_H */
