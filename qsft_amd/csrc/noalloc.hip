// One peeling transform with no device allocation by the caller or the
// library: the descriptor, the block prefixes and the peel workspace live in
// a fixed static device array of this library (one copy per GPU, held until
// the process ends).
#include <hip/hip_runtime.h>

#include <mutex>

#include "qsft_common.h"

extern "C" void qsft_set_error(const char* msg);
extern "C" int qsft_hip_status(hipError_t e, const char* where);

// The scratch array. Host code never uses its name as an address: the
// device address of each GPU's copy comes from hipGetSymbolAddress (below).
static __device__ __align__(256) unsigned char scratch[QSFT_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr size_t kDescBytes = 512;                      // descriptor slot (sizeof(qsft_peel_desc) <= 512)
constexpr size_t kPreBytes = 256;                       // pre_bins[2], pre_slots[2]
constexpr size_t kHeadBytes = kDescBytes + kPreBytes;  // the workspace starts here
constexpr int kMaxDevices = 64;
static_assert(sizeof(qsft_peel_desc) <= kDescBytes, "descriptor slot too small");

std::mutex g_lock;
void* g_ptr[kMaxDevices] = {};

// The scratch address on the current device, looked up once per device with
// that device current.
int scratch_ptr(void** out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return qsft_hip_status(e, "qsft_peel_transform_noalloc(hipGetDevice)");
    if (dev < 0 || dev >= kMaxDevices) {
        qsft_set_error("qsft_peel_transform_noalloc: device index out of range");
        return QSFT_ERR_ARG;
    }
    std::lock_guard<std::mutex> guard(g_lock);
    if (!g_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return qsft_hip_status(e, "qsft_peel_transform_noalloc(hipGetSymbolAddress)");
        if (!p) {
            qsft_set_error("qsft_peel_transform_noalloc: no scratch address");
            return QSFT_ERR_HIP;
        }
        g_ptr[dev] = p;
    }
    *out = g_ptr[dev];
    return QSFT_OK;
}

// Writes the bound descriptor and the one-instance block prefixes into the
// scratch head (a kernel argument, so nothing on the host must outlive the
// call, and no copy or memset node is recorded).
__global__ __launch_bounds__(64) void k_noalloc_setup(qsft_peel_desc d, int blocks_bins, int blocks_slots,
                                                      qsft_peel_desc* __restrict__ dst, int32_t* __restrict__ pre) {
    const int t = threadIdx.x;
    constexpr int kWords = (int)(sizeof(qsft_peel_desc) / 4);
    const uint32_t* src = reinterpret_cast<const uint32_t*>(&d);
    uint32_t* out = reinterpret_cast<uint32_t*>(dst);
    for (int i = t; i < kWords; i += 64) out[i] = src[i];
    if (t == 0) {
        pre[0] = 0;
        pre[1] = blocks_bins;
        pre[2] = 0;
        pre[3] = blocks_slots;
    }
}

}  // namespace

extern "C" size_t qsft_noalloc_scratch_bytes(void) { return QSFT_NOALLOC_SCRATCH_BYTES; }

extern "C" int qsft_peel_transform_noalloc(int q, int n, int b, int C, int P, int p1, int method, int max_iter,
                                           double cutoff, int flags, const int8_t* Mt, const int8_t* D, double* U,
                                           int n_pad, void* dst, int64_t cap_rows, int row_stride, void* stream) {
    if (!Mt || !D || !U || !dst || (flags & ~(QSFT_PEEL_SHARED_D | QSFT_PEEL_UNIT_D | QSFT_PEEL_UNIT_D_NEG))) {
        qsft_set_error("qsft_peel_transform_noalloc: bad arguments (flags: SHARED_D, UNIT_D, UNIT_D_NEG only)");
        return QSFT_ERR_ARG;
    }
    const size_t ws = qsft_peel_workspace_bytes(q, n, b, C, P, max_iter, 0);
    if (ws == 0) {
        qsft_set_error("qsft_peel_transform_noalloc: bad shape");
        return QSFT_ERR_ARG;
    }
    if (ws > QSFT_NOALLOC_SCRATCH_BYTES - kHeadBytes) {
        qsft_set_error("qsft_peel_transform_noalloc: the peel workspace of this shape exceeds the static scratch "
                       "(qsft_noalloc_scratch_bytes() minus 768 bytes); use qsft_peel_run with a caller workspace");
        return QSFT_ERR_ARG;
    }
    void* ptr = nullptr;
    int rc = scratch_ptr(&ptr);
    if (rc != QSFT_OK) return rc;
    unsigned char* base = static_cast<unsigned char*>(ptr);
    qsft_peel_desc h;
    rc = qsft_peel_bind(&h, q, n, b, C, P, p1, method, max_iter, cutoff, 0, Mt, D, U, base + kHeadBytes, ws, n_pad);
    if (rc != QSFT_OK) return rc;
    h.flags = flags;
    const int blocks_bins = qsft_peel_bin_blocks(&h);
    const int blocks_slots = qsft_peel_slot_blocks(&h);
    qsft_peel_desc* d_dev = reinterpret_cast<qsft_peel_desc*>(base);
    int32_t* pre = reinterpret_cast<int32_t*>(base + kDescBytes);
    hipLaunchKernelGGL(k_noalloc_setup, dim3(1), dim3(64), 0, (hipStream_t)stream, h, blocks_bins, blocks_slots,
                       d_dev, pre);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return qsft_hip_status(e, "qsft_peel_transform_noalloc(setup)");
    rc = qsft_peel_run(d_dev, pre, pre + 2, 1, blocks_bins, blocks_slots, P, max_iter, stream);
    if (rc != QSFT_OK) return rc;
    return qsft_peel_pack(d_dev, 1, dst, cap_rows, row_stride, stream);
}
