// common.hpp -- shared definitions of libcalcir's HIP sources (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "calcir.h"

namespace cir {

// calcIR.h:36-37
constexpr double kHbar = 5.308837367;  // cm^-1 ps
constexpr double kPi = 3.14159265359;
constexpr double kTwoPiExact = 6.283185307179586476925286766559;
// calcIR.cu:1036, 1197, 1198
constexpr double kBohrPerNm = 18.8973;
constexpr double kCmPerHartree = 2.1947463e5;
// calcIR.cu:1136 (TIP4P O-M), 1289 (TDM offset), 1299 (Raman anisotropy),
// 1152 and 1157 (site charges)
constexpr double kOMTip4p = 0.0150;
constexpr double kTdmOffset = 0.067;
constexpr double kRamanAniso = 4.6;
constexpr double kQH = 0.52;
constexpr double kQM = 1.04;

// Map coefficients (calcIR.cu:1203-1242), resolved once on the host.
struct MapCoef {
    double w[3], x[2], p[2], mup[3], wi[3];
};
bool map_coef(int imap, int ispecies, MapCoef* c);

// Error plumbing: the detail text of the last failure (thread-local).
void set_error(const std::string& msg);
int fail(int status, const std::string& msg);
int hip_fail(hipError_t e, const char* what, const char* file, int line);
// Status for a C++ exception caught at a C entry point (called from the
// entry point's catch handler): host allocation failure -> CIR_ERR_NOMEM,
// anything else -> CIR_ERR_ARG, with `where` in the detail text.  No C++
// exception crosses the C ABI.
int host_exception(const char* where) noexcept;

#define CIR_HIP(call)                                                        \
    do {                                                                     \
        hipError_t cir_e_ = (call);                                          \
        if (cir_e_ != hipSuccess) return ::cir::hip_fail(cir_e_, #call, __FILE__, __LINE__); \
    } while (0)

#define CIR_CHECK_LAUNCH() CIR_HIP(hipGetLastError())

#define CIR_TRY(call)                       \
    do {                                    \
        int cir_s_ = (call);                \
        if (cir_s_ != CIR_OK) return cir_s_; \
    } while (0)

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Per-phase device timing with HIP events recorded on the stream the phase's
// kernels are launched on.  Inactive (on == false) it records nothing.
struct Prof {
    bool on = false;
    struct Rec {
        int phase;
        hipEvent_t a, b;
    };
    std::vector<Rec> recs;
    Prof() = default;
    Prof(const Prof&) = delete;
    Prof& operator=(const Prof&) = delete;
    ~Prof() {  // records a failed run left uncollected
        for (Rec& r : recs) {
            (void)hipEventDestroy(r.a);
            (void)hipEventDestroy(r.b);
        }
    }
    int begin(int phase, hipStream_t s) {
        if (!on) return -1;
        Rec r{phase, nullptr, nullptr};
        if (hipEventCreate(&r.a) != hipSuccess) return -1;
        if (hipEventCreate(&r.b) != hipSuccess) {
            (void)hipEventDestroy(r.a);
            return -1;
        }
        (void)hipEventRecord(r.a, s);
        recs.push_back(r);
        return int(recs.size()) - 1;
    }
    void end(int idx, hipStream_t s) {
        if (idx >= 0) (void)hipEventRecord(recs[idx].b, s);
    }
    // Adds the elapsed time of every record into secs[phase] / counts[phase].
    void collect(double* secs, long long* counts, int nphase) {
        for (Rec& r : recs) {
            float ms = 0.f;
            if (hipEventSynchronize(r.b) == hipSuccess && hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess &&
                r.phase >= 0 && r.phase < nphase) {
                secs[r.phase] += ms * 1e-3;
                counts[r.phase] += 1;
            }
            (void)hipEventDestroy(r.a);
            (void)hipEventDestroy(r.b);
        }
        recs.clear();
    }
};

// Two HIP events that bracket a timed region, destroyed with the scope on
// every return path.
struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
    EventPair() = default;
    EventPair(const EventPair&) = delete;
    EventPair& operator=(const EventPair&) = delete;
    hipError_t create() {
        hipError_t e = hipEventCreate(&a);
        return e == hipSuccess ? hipEventCreate(&b) : e;
    }
    ~EventPair() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

struct ProfScope {
    Prof* p;
    int idx;
    hipStream_t s;
    ProfScope(Prof* prof, int phase, hipStream_t st) : p(prof), idx(prof ? prof->begin(phase, st) : -1), s(st) {}
    ~ProfScope() {
        if (p) p->end(idx, s);
    }
};

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// XCDs of an MI355X: a launch's workgroups are dispatched to them in turn
// (workgroup i to XCD i mod 8), each XCD with its own L2.
constexpr int kXcds = 8;

}  // namespace cir

// Device-side helpers (HIP translation units only) ----------------------------
#if defined(__HIPCC__)
namespace cir {

// calcIR.cu:1457-1460.  Contraction is disabled so that dx - L*round(dx/L)
// is rounded exactly as written (the FP32 cutoff decision depends on it).
template <typename T>
__device__ __forceinline__ T minimage(T dx, T boxl) {
#pragma clang fp contract(off)
    return dx - boxl * round(dx / boxl);
}

// Periodic box of one frame.  Rectangular boxes (the reference's case,
// calcIR.cu:404-405) hold the diagonal; triclinic boxes follow the GROMACS
// convention a = (a0,0,0), b = (b0,b1,0), c = (c0,c1,c2), stored per frame as
// the row-major 3x3 matrix of a, b, c.
template <typename T>
struct PBox {
    T a0, b0, b1, c0, c1, c2;
    bool tri;  // some off-diagonal element is nonzero
};

// boxk 0: 3 floats (diagonal) per frame; 1: 9 floats (a, b, c rows).
template <typename T>
__device__ __forceinline__ PBox<T> load_box(const float* box, int boxk, int f) {
    PBox<T> B;
    if (boxk == 0) {
        const float* q = box + 3 * size_t(f);
        B.a0 = T(q[0]);
        B.b1 = T(q[1]);
        B.c2 = T(q[2]);
        B.b0 = B.c0 = B.c1 = T(0);
        B.tri = false;
    } else {
        const float* q = box + 9 * size_t(f);
        B.a0 = T(q[0]);
        B.b0 = T(q[3]);
        B.b1 = T(q[4]);
        B.c0 = T(q[6]);
        B.c1 = T(q[7]);
        B.c2 = T(q[8]);
        B.tri = q[3] != 0.0f || q[6] != 0.0f || q[7] != 0.0f;
    }
    return B;
}

// Minimum image of a separation.  Rectangular: calcIR.cu:1457-1460 per axis
// (bit for bit the reference form).  Triclinic: the lattice vectors are
// removed in the order c, b, a, each by the rounded fractional coordinate
// along its own axis (the GROMACS form for reduced boxes).
template <typename T>
__device__ __forceinline__ void minimage3(T& dx, T& dy, T& dz, const PBox<T>& B) {
#pragma clang fp contract(off)
    if (!B.tri) {
        dx = minimage<T>(dx, B.a0);
        dy = minimage<T>(dy, B.b1);
        dz = minimage<T>(dz, B.c2);
        return;
    }
    T s = round(dz / B.c2);
    dx = dx - s * B.c0;
    dy = dy - s * B.c1;
    dz = dz - s * B.c2;
    s = round(dy / B.b1);
    dx = dx - s * B.b0;
    dy = dy - s * B.b1;
    s = round(dx / B.a0);
    dx = dx - s * B.a0;
}

// Neumaier-compensated accumulation.
template <typename T>
struct Compensated {
    T sum = T(0), c = T(0);
    __device__ __forceinline__ void add(T v) {
        T t = sum + v;
        if (fabs(sum) >= fabs(v))
            c += (sum - t) + v;
        else
            c += (v - t) + sum;
        sum = t;
    }
    __device__ __forceinline__ T value() const { return sum + c; }
};

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

}  // namespace cir
#endif  // __HIPCC__
