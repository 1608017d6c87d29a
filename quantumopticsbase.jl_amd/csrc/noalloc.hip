// normalize!(ket) without any device allocation, for callers that cannot
// call hipMalloc (code under HIP graph capture, tight serving loops).
//
// qob_normalize_ket (api.cpp) keeps its two-stage FP64 norm partials and the
// norm itself in the per-stream workspace, which grows on first use.  This
// variant runs the same kernels (k_blas1.hip: launch_nrm2, then
// launch_scale_by_inv_norm) on a fixed file-scope device array instead:
//   scratch[0 .. NOALLOC_PARTS)   stage-1 partial sums (double2 each)
//   scratch[NOALLOC_PARTS]        the norm (a double in the first half)
// The stage-1 launch uses reduce_blocks(n) <= NOALLOC_PARTS workgroups for
// every n, and stage 2 reads exactly those partials and writes the norm
// before the scale kernel reads it, so nothing in the array has to start at
// zero and no clearing launch is needed.
//
// Each device holds its own copy of the array for the life of the process.
// Host code never passes the array's name to a kernel: the name is a host
// stand-in, not a device address.  qob_ctx_create calls noalloc_init with
// the context's device current, which looks up the device address once per
// device with hipGetSymbolAddress; the entry point uses only that address.
//
// Calls that overlap on one device (two host threads, or two streams with
// work in flight at once) share the array and race; callers serialise them
// (one stream per device, or an event between streams).
#include "qob_internal.h"

#include <mutex>

namespace qob {

constexpr int NOALLOC_PARTS = 1024;  // reduce_blocks() never exceeds this
constexpr int NOALLOC_SLOTS = NOALLOC_PARTS + 1;
constexpr int NOALLOC_MAX_DEVICES = 64;

static __device__ __align__(256) double2 scratch[NOALLOC_SLOTS];

static std::mutex g_noalloc_mu;
static void *g_noalloc_ptr[NOALLOC_MAX_DEVICES];  // device address per device, null until looked up

static_assert(sizeof(double2) * NOALLOC_SLOTS == 16 * 1025, "scratch layout");

// Look up this device's copy of the array (the device must be current).
qob_status noalloc_init(int device) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES)
        return set_error(QOB_EARG, "no-allocation scratch: device index out of range");
    std::lock_guard<std::mutex> g(g_noalloc_mu);
    if (g_noalloc_ptr[device]) return QOB_OK;
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return set_error(QOB_EHIP, "hipGetSymbolAddress(scratch) on device %d: %s", device,
                         hipGetErrorString(e));
    g_noalloc_ptr[device] = p;
    return QOB_OK;
}

static void *noalloc_ptr(int device) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> g(g_noalloc_mu);
    return g_noalloc_ptr[device];
}

}  // namespace qob

extern "C" qob_status qob_normalize_ket_noalloc(qob_ctx ctx, qob_dtype dt, int64_t n, void *x, qob_stream s) {
    using namespace qob;
    if (!ctx) return set_error(QOB_EARG, "null context");
    if (dt != QOB_C64 && dt != QOB_C32) return set_error(QOB_EARG, "bad dtype %d", (int)dt);
    if (n < 0) return set_error(QOB_EDIM, "negative length %lld", (long long)n);
    if (n == 0) return QOB_OK;
    if (!x) return set_error(QOB_EARG, "null state pointer");
    if (reduce_blocks(n) > NOALLOC_PARTS)
        return set_error(QOB_EUNSUPPORTED, "no-allocation normalize: %d partial sums exceed the %d-slot scratch",
                         reduce_blocks(n), NOALLOC_PARTS);
    double2 *base = (double2 *)noalloc_ptr(ctx_device(ctx));
    if (!base) return set_error(QOB_EHIP, "no-allocation scratch not set up on device %d", ctx_device(ctx));
    hipStream_t st = ctx_stream(ctx, s);
    double *nrm = (double *)(base + NOALLOC_PARTS);
    hipError_t e = launch_nrm2(dt, n, x, base, nrm, st);
    if (e == hipSuccess) e = launch_scale_by_inv_norm(dt, n, nrm, x, st);
    if (e != hipSuccess) return set_error(QOB_EHIP, "no-allocation normalize: %s", hipGetErrorString(e));
    return QOB_OK;
}
