// C-ABI glue: thread-local error text, version and struct-size queries.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>

#include "../../include/qsft_amd.h"

namespace {
thread_local char g_err[512] = "";
}

extern "C" void qsft_set_error(const char* msg) {
    strncpy(g_err, msg ? msg : "", sizeof(g_err) - 1);
    g_err[sizeof(g_err) - 1] = 0;
}

extern "C" int qsft_hip_status(hipError_t e, const char* where) {
    if (e == hipSuccess) return QSFT_OK;
    snprintf(g_err, sizeof(g_err), "%s: %s", where, hipGetErrorString(e));
    return QSFT_ERR_HIP;
}

extern "C" const char* qsft_last_error(void) { return g_err; }

extern "C" int qsft_abi_version(void) { return QSFT_ABI_VERSION; }

extern "C" size_t qsft_sizeof(int which) {
    switch (which) {
        case 0: return sizeof(qsft_eval_desc);
        case 1: return sizeof(qsft_eval_points_desc);
        case 2: return sizeof(qsft_dft_desc);
        case 3: return sizeof(qsft_assemble_desc);
        case 4: return sizeof(qsft_peel_desc);
        case 5: return sizeof(qsft_eval_mfma_desc);
        case 6: return sizeof(qsft_request);
        case 7: return sizeof(qsft_batch_sizes);
        default: return 0;
    }
}
