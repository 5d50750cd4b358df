// ctx.hpp -- the run context shared by the pipeline (pipeline.hip) and the
// lineshape adjoint (adjoint.hip): device buffers, the frame schedule and the
// context itself.
#pragma once

#include <memory>
#include <vector>

#include "common.hpp"

namespace cir {

// Stream-ordered allocation for one host thread.  hipFree waits for every
// stream of the device, so concurrent jobs that free their buffers with it
// stall one another until all queued work has drained.  While a thread has
// `tl_pool` set, DevBuf takes its buffers from that memory pool on that
// stream and frees them in the stream's order (cir_run_many's jobs).
struct PoolStream {
    hipMemPool_t pool;
    hipStream_t stream;
};
inline thread_local const PoolStream* tl_pool = nullptr;

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    hipStream_t fs = nullptr;  // stream a pooled buffer is freed on
    bool pooled = false;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) {
            if (pooled)
                (void)hipFreeAsync(p, fs);
            else
                (void)hipFree(p);
        }
        p = nullptr;
        bytes = 0;
        fs = nullptr;
        pooled = false;
    }
    int alloc(size_t n) {
        release();
        if (n == 0) n = 256;
        const PoolStream* ps = tl_pool;
        const hipError_t e = ps ? hipMallocFromPoolAsync(&p, n, ps->pool, ps->stream) : hipMalloc(&p, n);
        if (e != hipSuccess) {
            p = nullptr;
            return hip_fail(e, ps ? "hipMallocFromPoolAsync" : "hipMalloc", __FILE__, __LINE__);
        }
        bytes = n;
        if (ps) {
            pooled = true;
            fs = ps->stream;
        }
        return CIR_OK;
    }
    template <typename T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

// Which trajectory frame each (sample, tcf point) uses and, for the batched
// flow, the distinct frames in ascending order with their per-slot actions
// (calcIR.cu:360-395 frame choice).
struct Schedule {
    std::vector<int> fidx;        // ns*nt trajectory frame per (sample, point)
    std::vector<int> uniq;        // distinct frames, ascending
    std::vector<int> uof;         // trajectory frame -> distinct index (-1 if unused)
    std::vector<double> mult;     // (sample, point) pairs per distinct frame
    std::vector<double> orig;     // samples starting at the distinct frame
    int nslot = 0;
    std::vector<int> action;      // nuniq*nslot
    std::vector<int> target;      // nuniq*nslot
};

int build_schedule(const cir_params& p, int nframes, const float* times, Schedule& S);
int gather_frames(cir_ctx* c, const float* xyz_d, const float* box_d, const int* frames, int nb, float* xyzc,
                  float* boxc);
int auto_chunk(const cir_ctx* c, int nuniq);
int check_boxes(const cir_ctx* c, int nframes, const float* box);

// State the lineshape adjoint keeps between its forward and backward calls.
struct AdjState;
void adj_release(AdjState* a);
struct AdjDeleter {
    void operator()(AdjState* a) const { adj_release(a); }
};

// State of a streamed batched run between cir_stream_begin and cir_stream_end.
struct StreamRun;
void stream_release(StreamRun* s);
struct StreamDeleter {
    void operator()(StreamRun* s) const { stream_release(s); }
};

}  // namespace cir

struct cir_ctx {
    cir_params p;
    int imodel = 0, ispecies = 0;
    int natoms = 0, nmol = 0, N = 0, ld = 0, nomega = 0, nfft = 0;
    int real_prec = CIR_PREC_FP32;  // storage precision of real arrays
    int prop_prec = CIR_PREC_FP32;  // product precision of the propagator and the eigenvector back-transform
    size_t esz = 4;
    hipStream_t stream = nullptr;
    cir::MapCoef mc{};              // map coefficients in use (the table row unless overridden)
    int boxk = 0;                   // 0: 3 box floats per frame (rectangular), 1: 9 (triclinic)
    int mem_share = 1;              // jobs sharing the device memory (cir_run_many)
    std::vector<double> tcf[3], ftcf[3], sw, pw, rw, rmw;
    int64_t nsteps = 0, nframes_used = 0, hist_skipped = 0;
    double seconds = 0.0;
    bool have = false;
    cir::Prof prof;
    double phase_sec[CIR_NPHASE] = {};
    long long phase_cnt[CIR_NPHASE] = {};
    std::unique_ptr<cir::AdjState, cir::AdjDeleter> adj;
    std::unique_ptr<cir::StreamRun, cir::StreamDeleter> streamed;
};

namespace cir {
// Floats per frame in the context's box arrays.
inline int box_floats(const cir_ctx* c) { return c->boxk ? 9 : 3; }
}  // namespace cir
