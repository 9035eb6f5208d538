"""ctypes binding of libqsft_amd.so (the C ABI declared in include/qsft_amd.h).

torch is imported first on purpose: the library links libamdhip64.so.7, and
loading it after torch makes the dynamic loader reuse the HIP runtime torch
already loaded (same soname), so both share one HIP context and device
pointers from torch's allocator are valid in the library.

There is no fallback: if the library is missing the import fails loudly.
"""
import ctypes
import os

import torch  # noqa: F401  (must precede the CDLL load, see module docstring)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("QSFT_AMD_LIB", os.path.join(HERE, "lib", "libqsft_amd.so"))

QSFT_OK, QSFT_ERR_ARG, QSFT_ERR_HIP, QSFT_ERR_DEVICE = 0, 1, 2, 3
QSFT_F32, QSFT_F64 = 0, 1
QSFT_DETECT_NOISELESS, QSFT_DETECT_NSO1 = 0, 1
QSFT_DETECT_CODED = 2
QSFT_NOISE_NONE, QSFT_NOISE_ARRAY, QSFT_NOISE_PHILOX = 0, 1, 2
QSFT_DEVERR_TABLE_FULL, QSFT_DEVERR_TAG_COLLISION, QSFT_DEVERR_OUT_FULL = 1, 2, 4
QSFT_MAX_Q, QSFT_MAX_N, QSFT_MAX_B, QSFT_MAX_C, QSFT_MAX_R = 64, 256, 16, 16, 16
QSFT_DFT_MAX_PASSES = 4
QSFT_PEEL_SHARED_D, QSFT_PEEL_UNIT_D, QSFT_PEEL_UNIT_D_NEG, QSFT_PEEL_SPARSE_U = 1, 2, 4, 8
QSFT_PEEL_ONE_BIN_WAVES = 16
QSFT_CTR_ITER_STRIDE = 8
QSFT_CTR_LOG, QSFT_CTR_OUT, QSFT_CTR_ERROR, QSFT_CTR_ITERS, QSFT_CTR_PEELED, QSFT_CTR_LIVE = 128, 129, 130, 131, 132, 133
QSFT_FIN_FLUSH = 1
QSFT_CTR_WORDS = 136

i32, i64, u64, f64, vp = ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64, ctypes.c_double, ctypes.c_void_p


class EvalDesc(ctypes.Structure):
    _fields_ = [("q", i32), ("n", i32), ("b", i32), ("C", i32), ("P", i32), ("S", i32),
                ("n_pad", i32), ("_pad0", i32),
                ("Mt", vp), ("D", vp), ("K", vp), ("a", vp), ("x", vp)]


class EvalPointsDesc(ctypes.Structure):
    _fields_ = [("q", i32), ("n", i32), ("S", i32), ("n_pad", i32), ("N", i64),
                ("m", vp), ("K", vp), ("a", vp), ("x", vp)]


class DftDesc(ctypes.Structure):
    _fields_ = [("q", i32), ("b", i32), ("rows", i32), ("in_stride", i32), ("in_row_stride", i64),
                ("in_", vp), ("out", vp), ("n_pass", i32),
                ("pass_d0", i32 * QSFT_DFT_MAX_PASSES), ("pass_d1", i32 * QSFT_DFT_MAX_PASSES),
                ("pass_ti", i32 * QSFT_DFT_MAX_PASSES), ("pass_blocks", i32 * QSFT_DFT_MAX_PASSES),
                ("pass_ta", i32 * QSFT_DFT_MAX_PASSES)]


class AssembleDesc(ctypes.Structure):
    _fields_ = [("Q", i32), ("C", i32), ("R", i32), ("p1", i32), ("src_rows", i32), ("src_prec", i32),
                ("noise_mode", i32), ("bin_major", i32),
                ("gsel", i32 * QSFT_MAX_C), ("rsel", i32 * QSFT_MAX_R),
                ("src", vp), ("noise", vp), ("sigma", f64), ("seed", u64), ("U", vp), ("energy", vp),
                ("written", vp), ("cutoff", f64), ("seed_out", vp)]


class PeelDesc(ctypes.Structure):
    _fields_ = [("q", i32), ("n", i32), ("b", i32), ("C", i32), ("P", i32), ("Q", i32),
                ("n_pad", i32), ("p1", i32), ("method", i32), ("max_iter", i32),
                ("hcap_log2", i32), ("out_cap", i32), ("flags", i32),
                ("cutoff", f64), ("peel_cap", f64),
                ("Mt", vp), ("D", vp), ("U", vp), ("energy", vp),
                ("kbin", vp), ("rho_bin", vp), ("slot_bin", vp), ("mark", vp), ("sstamp", vp),
                ("pend_head", vp), ("pend_next", vp),
                ("tag", vp), ("owner", vp),
                ("last", vp), ("head", vp), ("count", vp), ("live", vp), ("key", vp), ("rec_time", vp),
                ("rec_rho", vp), ("rec_next", vp), ("ctr", vp),
                ("out_k", vp), ("out_val", vp), ("out_order", vp),
                ("xsrc", vp), ("noise_sigma", f64), ("written", vp), ("noise_key", vp),
                ("rs_tab", vp), ("rs_t", i32), ("rs_s", i32), ("rs_nc", i32), ("_pad1", i32)]


class EvalMfmaDesc(ctypes.Structure):
    _fields_ = [("q", i32), ("n", i32), ("b", i32), ("C", i32), ("P", i32), ("S", i32),
                ("n_pad", i32), ("nsb", i32), ("npt", i32), ("shared_d", i32), ("bexp", i32), ("ptb", i32),
                ("Mt", vp), ("D", vp), ("K", vp), ("a", vp), ("Bf", vp), ("x", vp)]


class Request(ctypes.Structure):
    _fields_ = [("q", i32), ("n", i32), ("b", i32), ("C", i32), ("R", i32), ("S", i32), ("method", i32),
                ("_pad0", i32), ("cutoff", ctypes.c_double), ("sigma", ctypes.c_double), ("amax", ctypes.c_double),
                ("seed", ctypes.c_uint64)]


class BatchSizes(ctypes.Structure):
    _fields_ = [("input_bytes", ctypes.c_size_t), ("work_bytes", ctypes.c_size_t), ("plan_bytes", ctypes.c_size_t),
                ("out_bytes", ctypes.c_size_t), ("cap_rows", i64), ("row_stride", i32), ("n_req", i32)]


STRUCTS = {0: EvalDesc, 1: EvalPointsDesc, 2: DftDesc, 3: AssembleDesc, 4: PeelDesc, 5: EvalMfmaDesc, 6: Request,
           7: BatchSizes}

# exported symbols (every function declared in include/qsft_amd.h)
SIGNATURES = {
    "qsft_last_error": (ctypes.c_char_p, []),
    "qsft_abi_version": (ctypes.c_int, []),
    "qsft_sizeof": (ctypes.c_size_t, [ctypes.c_int]),
    "qsft_eval_blocks": (ctypes.c_int, [ctypes.POINTER(EvalDesc)]),
    "qsft_eval_cosets": (ctypes.c_int, [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp]),
    "qsft_eval_mfma_workspace_bytes": (ctypes.c_size_t, [ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "qsft_eval_mfma_blocks": (ctypes.c_int, [ctypes.POINTER(EvalMfmaDesc), ctypes.POINTER(ctypes.c_int),
                                             ctypes.POINTER(ctypes.c_int)]),
    "qsft_eval_mfma": (ctypes.c_int, [vp, vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp]),
    "qsft_query_digits_blocks": (ctypes.c_int, [ctypes.POINTER(EvalDesc)]),
    "qsft_query_digits": (ctypes.c_int, [vp, vp, ctypes.c_int, ctypes.c_int, vp]),
    "qsft_eval_points_blocks": (ctypes.c_int, [ctypes.POINTER(EvalPointsDesc)]),
    "qsft_eval_points": (ctypes.c_int, [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp]),
    "qsft_dft_plan": (ctypes.c_int, [ctypes.POINTER(DftDesc), ctypes.c_int]),
    "qsft_dft_lds_bytes": (ctypes.c_int, [ctypes.POINTER(DftDesc), ctypes.c_int, ctypes.c_int]),
    "qsft_dft_plan_candidate": (ctypes.c_int, [ctypes.POINTER(DftDesc), ctypes.c_int, ctypes.c_int]),
    "qsft_dft_autotune": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, ctypes.c_size_t,
                                         ctypes.c_int, vp, ctypes.POINTER(ctypes.c_float)]),
    "qsft_dft_schedule": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int] + [ctypes.POINTER(ctypes.c_int)] * 6),
    "qsft_dft_pass": (ctypes.c_int, [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     vp]),
    "qsft_dft4_rows": (ctypes.c_int, [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp]),
    "qsft_assemble_blocks": (ctypes.c_int, [ctypes.POINTER(AssembleDesc)]),
    "qsft_assemble_word": (ctypes.c_int, [ctypes.POINTER(AssembleDesc)]),
    "qsft_assemble": (ctypes.c_int, [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_int, vp]),
    "qsft_peel_workspace_bytes": (ctypes.c_size_t, [ctypes.c_int] * 7),
    "qsft_peel_bind": (ctypes.c_int, [ctypes.POINTER(PeelDesc), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_double, ctypes.c_int, vp, vp, vp, vp, ctypes.c_size_t,
                                      ctypes.c_int]),
    "qsft_peel_set_coded": (ctypes.c_int, [ctypes.POINTER(PeelDesc), ctypes.c_int, ctypes.c_int, vp]),
    "qsft_peel_bin_blocks": (ctypes.c_int, [ctypes.POINTER(PeelDesc)]),
    "qsft_peel_slot_blocks": (ctypes.c_int, [ctypes.POINTER(PeelDesc)]),
    "qsft_peel_run": (ctypes.c_int, [vp, vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     vp]),
    "qsft_peel_init": (ctypes.c_int, [vp, vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp]),
    "qsft_peel_iteration": (ctypes.c_int, [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp]),
    "qsft_peel_passes": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp]),
    "qsft_peel_finalize": (ctypes.c_int, [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp]),
    "qsft_peel_pack_header_bytes": (ctypes.c_size_t, [ctypes.c_int]),
    "qsft_peel_pack": (ctypes.c_int, [vp, ctypes.c_int, vp, ctypes.c_int64, ctypes.c_int, vp]),
    "qsft_peel_pack_ordered": (ctypes.c_int, [vp, ctypes.c_int, vp, ctypes.c_int64, ctypes.c_int, vp]),
    "qsft_request_footprint": (ctypes.c_size_t, [ctypes.POINTER(Request), ctypes.c_int]),
    "qsft_batch_create": (vp, [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(BatchSizes)]),
    "qsft_batch_input_offset": (ctypes.c_int, [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_size_t)]),
    "qsft_batch_pack_input": (ctypes.c_int, [vp, ctypes.c_int, vp, vp, vp, vp, vp]),
    "qsft_batch_pack_inputs": (ctypes.c_int, [vp, vp, vp, vp, vp, vp, ctypes.c_int]),
    "qsft_batch_delay_sign": (ctypes.c_int, [vp, ctypes.c_int]),
    "qsft_batch_bind": (ctypes.c_int, [vp, vp, vp, vp, vp, vp]),
    "qsft_batch_launch": (ctypes.c_int, [vp, ctypes.c_int, vp, ctypes.c_uint64, vp, vp]),
    "qsft_batch_capture": (ctypes.c_int, [vp, vp, vp]),
    "qsft_batch_replay": (ctypes.c_int, [vp, vp]),
    "qsft_batch_done": (ctypes.c_int, [vp]),
    "qsft_batch_wait": (ctypes.c_int, [vp]),
    "qsft_batch_free": (None, [vp]),
    "qsft_noalloc_scratch_bytes": (ctypes.c_size_t, []),
    "qsft_peel_transform_noalloc": (ctypes.c_int, [ctypes.c_int] * 8 + [ctypes.c_double, ctypes.c_int, vp, vp, vp,
                                                                        ctypes.c_int, vp, ctypes.c_int64,
                                                                        ctypes.c_int, vp]),
}


class QsftLibraryError(RuntimeError):
    pass


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"libqsft_amd.so not found at {LIB_PATH}; build it with "
                          "`python -m qsft_amd.build` (hipcc, gfx950)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    for which, cls in STRUCTS.items():
        got = lib.qsft_sizeof(which)
        if got != ctypes.sizeof(cls):
            raise ImportError(f"descriptor layout mismatch for {cls.__name__}: C {got} vs ctypes "
                              f"{ctypes.sizeof(cls)}")
    return lib


lib = _load()


def check(rc, what=""):
    if rc != QSFT_OK:
        msg = lib.qsft_last_error().decode(errors="replace")
        raise QsftLibraryError(f"{what or 'qsft call'} failed (code {rc}): {msg}")
