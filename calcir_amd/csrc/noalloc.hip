// noalloc.hip -- cir_propagate_noalloc: the propagation step of cir_propagate
// (calcIR.cu:597-617) for callers that cannot allocate device memory (code
// under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc).  Its work buffer is `scratch`, a fixed device array of
// kScratchBytes that every call on a GPU shares, so calls that overlap on one
// GPU race.  The array keeps the last call's contents; the propagation
// kernels clear the parts they accumulate into on every call.

#include <hip/hip_runtime.h>

#include <mutex>

#include "common.hpp"
#include "spectra.hpp"

namespace {

constexpr size_t kScratchBytes = size_t(32) << 20;  // 32 MiB per GPU, held until the process ends
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// The array's device address on each GPU (each holds its own copy).  The
// host-side name `scratch` is not a device address: only these pointers are
// passed to kernels.
constexpr int kMaxDevices = 64;
void* g_scratch[kMaxDevices];
hipError_t g_status[kMaxDevices];
std::once_flag g_once[kMaxDevices];

int scratch_on_current_device(void** out) {
    int dev = -1;
    CIR_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return cir::fail(CIR_ERR_ARG, "cir_propagate_noalloc: device index out of range");
    std::call_once(g_once[dev], [dev] {
        void* p = nullptr;
        g_status[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_scratch[dev] = p;
    });
    if (g_status[dev] != hipSuccess || !g_scratch[dev])
        return cir::hip_fail(g_status[dev] != hipSuccess ? g_status[dev] : hipErrorInvalidSymbol,
                             "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
    *out = g_scratch[dev];
    return CIR_OK;
}

}  // namespace

extern "C" size_t cir_propagate_noalloc_max_bytes(void) { return kScratchBytes; }

extern "C" int cir_propagate_noalloc(int prec, int n, int ld, int ncols, int ldy, float dt, const void* V_d,
                                     const double* w_d, const void* Yin_d, void* Yout_d, void* stream) {
    if (n <= 0 || ncols <= 0 || ld < n || ldy < 2 * ncols || !V_d || !w_d || !Yin_d || !Yout_d)
        return cir::fail(CIR_ERR_ARG, "cir_propagate_noalloc: bad sizes or null pointer");
    const size_t esz = (prec == CIR_PREC_FP64) ? sizeof(double) : sizeof(float);
    const size_t need = cir::propagate_work_elems(n, ncols) * esz;
    if (need > kScratchBytes)
        return cir::fail(CIR_ERR_ARG, "cir_propagate_noalloc: this shape needs more than the 32 MiB scratch "
                                      "(use cir_propagate with a workspace)");
    void* G = nullptr;
    CIR_TRY(scratch_on_current_device(&G));
    const int rp = prec == CIR_PREC_FP64 ? CIR_PREC_FP64 : CIR_PREC_FP32;
    // no autotuning here: timing the K splits would create HIP events
    return cir::propagate(rp, prec, n, ld, ncols, ldy, double(dt), V_d, w_d, Yin_d, Yout_d, G, cir::as_stream(stream),
                          false);
}
