// k_spectra.hip -- frequency distribution / IPR, spectral density,
// propagation helpers, correlation functions and the C2R lineshape
// transform (gfx950).

#include <algorithm>
#include <cfloat>
#include <map>
#include <mutex>
#include <tuple>

#include "gemm.hpp"
#include "spectra.hpp"
#include "tuner.hpp"

namespace cir {

constexpr int kSpNT = 256;
constexpr int kPropMaxSplit = 8;

// ---------------------------------------------------------------------------
// IPR and frequency distribution, calcIR.cu:494-538.  One lane per mode; V is
// row-major (component i, mode k) so lanes read consecutive modes of a row.
// Per-block LDS histograms, flushed with FP64 atomics.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(kSpNT)
ipr_hist_kernel(int n, int ld, int nchrom_mol, float avef, float omegaStart, int omegaStep, int nomega,
                const T* __restrict__ Vall, const double* __restrict__ wall, const double* __restrict__ weight,
                double* __restrict__ pw, double* __restrict__ rw, double* __restrict__ rmw,
                long long* __restrict__ skipped, int use_lds) {
    extern __shared__ double sh[];
    const int b = blockIdx.y, tid = threadIdx.x;
    const double wgt = weight ? weight[b] : 1.0;
    if (wgt == 0.0) return;
    if (use_lds) {
        for (int q = tid; q < 3 * nomega; q += kSpNT) sh[q] = 0.0;
        __syncthreads();
    }
    const int k = blockIdx.x * kSpNT + tid;
    if (k < n) {
        const T* V = Vall + size_t(b) * n * ld;
        T s4 = T(0), outer = T(0), inner = T(0);
        int cnt = 0;
        for (int i = 0; i < n; ++i) {
            const T c = V[size_t(i) * ld + k];
            const T c2 = c * c;
            s4 += c2 * c2;
            inner += c2;
            if (++cnt == nchrom_mol) {
                outer += inner * inner;
                inner = T(0);
                cnt = 0;
            }
        }
        const double ipr = double(T(1) / s4);
        const double mipr = double(T(1) / outer);
        const float wi = float(wall[size_t(b) * n + k]) + avef;
        const int bin = int(roundf((wi - omegaStart) / float(omegaStep)));
        if (bin < 0 || bin >= nomega) {
            // calcIR.cu:530-537 warns and then writes out of bounds; the
            // mode is skipped and counted here.
            atomicAdd(reinterpret_cast<unsigned long long*>(skipped), (unsigned long long)(llrint(wgt)));
        } else {
            const double inv = wgt / double(omegaStep);
            if (use_lds) {
                atomicAdd(&sh[bin], inv);
                atomicAdd(&sh[nomega + bin], inv * ipr);
                atomicAdd(&sh[2 * nomega + bin], inv * mipr);
            } else {
                atomicAdd(&pw[bin], inv);
                atomicAdd(&rw[bin], inv * ipr);
                atomicAdd(&rmw[bin], inv * mipr);
            }
        }
    }
    if (use_lds) {
        __syncthreads();
        for (int q = tid; q < nomega; q += kSpNT) {
            if (sh[q] != 0.0) {
                atomicAdd(&pw[q], sh[q]);
                atomicAdd(&rw[q], sh[nomega + q]);
                atomicAdd(&rmw[q], sh[2 * nomega + q]);
            }
        }
    }
}

// Deterministic form: per mode its bin (unclamped) and IPR / molecular IPR,
// then one lane per bin sums over the frames in order and over the frame's
// modes of that bin in order (the bins ascend with the eigenvalues, so those
// modes are one run, found by binary search).
template <typename T>
__global__ void __launch_bounds__(kSpNT)
ipr_modes_kernel(int n, int ld, int nchrom_mol, float avef, float omegaStart, int omegaStep, int nomega,
                 const T* __restrict__ Vall, const double* __restrict__ wall, const double* __restrict__ weight,
                 int* __restrict__ mbin, double* __restrict__ mval, long long* __restrict__ skipped) {
    const int b = blockIdx.y, k = blockIdx.x * kSpNT + threadIdx.x;
    if (k >= n) return;
    const double wgt = weight ? weight[b] : 1.0;
    const T* V = Vall + size_t(b) * n * ld;
    T s4 = T(0), outer = T(0), inner = T(0);
    int cnt = 0;
    for (int i = 0; i < n; ++i) {
        const T c = V[size_t(i) * ld + k];
        const T c2 = c * c;
        s4 += c2 * c2;
        inner += c2;
        if (++cnt == nchrom_mol) {
            outer += inner * inner;
            inner = T(0);
            cnt = 0;
        }
    }
    const float wi = float(wall[size_t(b) * n + k]) + avef;
    const int bin = int(roundf((wi - omegaStart) / float(omegaStep)));
    const size_t e = size_t(b) * n + k;
    mbin[e] = bin;
    mval[2 * e] = double(T(1) / s4);
    mval[2 * e + 1] = double(T(1) / outer);
    if ((bin < 0 || bin >= nomega) && wgt != 0.0)
        atomicAdd(reinterpret_cast<unsigned long long*>(skipped), (unsigned long long)(llrint(wgt)));  // integer: any order
}

__global__ void __launch_bounds__(kSpNT)
ipr_bins_kernel(int n, int batch, int omegaStep, int nomega, const double* __restrict__ weight,
                const int* __restrict__ mbin, const double* __restrict__ mval, double* __restrict__ pw,
                double* __restrict__ rw, double* __restrict__ rmw) {
    const int q = blockIdx.x * kSpNT + threadIdx.x;
    if (q >= nomega) return;
    double sp = 0.0, sr = 0.0, sm = 0.0;
    for (int b = 0; b < batch; ++b) {
        const double wgt = weight ? weight[b] : 1.0;
        if (wgt == 0.0) continue;
        const int* bins = mbin + size_t(b) * n;
        int lo = 0, hi = n;  // first mode with bin >= q
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (bins[mid] < q)
                lo = mid + 1;
            else
                hi = mid;
        }
        const double inv = wgt / double(omegaStep);
        for (int k = lo; k < n && bins[k] == q; ++k) {
            const double* v = mval + 2 * (size_t(b) * n + k);
            sp += inv;
            sr += inv * v[0];
            sm += inv * v[1];
        }
    }
    pw[q] += sp;
    rw[q] += sr;
    rmw[q] += sm;
}

size_t ipr_hist_det_bytes(int batch, int n) { return size_t(batch) * n * (sizeof(int) + 2 * sizeof(double)) + 256; }
size_t spectral_density_det_bytes(int batch, int nomega) { return size_t(batch) * nomega * sizeof(double); }

// |V^T mu|^2 per mode (calcIR.cu:462-467 fused).
template <typename T>
__global__ void __launch_bounds__(kSpNT)
mu2_kernel(int n, int ld, const T* __restrict__ Vall, const T* __restrict__ sites,
           const double* __restrict__ weight, double* __restrict__ m2all) {
    const int b = blockIdx.y;
    if (weight && weight[b] == 0.0) return;
    const int k = blockIdx.x * kSpNT + threadIdx.x;
    if (k >= n) return;
    const T* V = Vall + size_t(b) * n * ld;
    const T* mu = sites + size_t(b) * CIR_NSITE * n + size_t(CIR_SITE_MU) * n;
    T ax = T(0), ay = T(0), az = T(0);
    for (int i = 0; i < n; ++i) {
        const T v = V[size_t(i) * ld + k];
        ax += v * mu[i];
        ay += v * mu[n + i];
        az += v * mu[2 * n + i];
    }
    m2all[size_t(b) * n + k] = double(ax * ax + ay * ay + az * az);
}

// Lorentzian sum over modes per omega bin (calcIR.cu:1414-1444).
// (per_frame: frame b's term written to per_frame[b * nomega + q] instead of
// added to sw, for the fixed-order sum of sw_frames_kernel)
__global__ void __launch_bounds__(kSpNT)
sw_kernel(int n, int nomega, float omegaStart, int omegaStep, double gamma, float avef,
          const double* __restrict__ m2all, const double* __restrict__ wall,
          const double* __restrict__ weight, double* __restrict__ sw, double* __restrict__ per_frame) {
    const int b = blockIdx.y;
    const double wgt = weight ? weight[b] : 1.0;
    const int q = blockIdx.x * kSpNT + threadIdx.x;
    if (q >= nomega) return;
    if (wgt == 0.0) {
        if (per_frame) per_frame[size_t(b) * nomega + q] = 0.0;
        return;
    }
    const double om = double(float(omegaStart + float(omegaStep * q)));  // calcIR.cu:271
    const double* m2 = m2all + size_t(b) * n;
    const double* w = wall + size_t(b) * n;
    double acc = 0.0;
    for (int k = 0; k < n; ++k) {
        const double dw = om - (w[k] + double(avef));
        acc += m2[k] * gamma / (dw * dw + gamma * gamma) / kPi;
    }
    if (per_frame)
        per_frame[size_t(b) * nomega + q] = wgt * acc;
    else
        atomicAdd(&sw[q], wgt * acc);
}

__global__ void __launch_bounds__(kSpNT)
sw_frames_kernel(int batch, int nomega, const double* __restrict__ per_frame, double* __restrict__ sw) {
    const int q = blockIdx.x * kSpNT + threadIdx.x;
    if (q >= nomega) return;
    double acc = 0.0;
    for (int b = 0; b < batch; ++b) acc += per_frame[size_t(b) * nomega + q];
    sw[q] += acc;
}

// ---------------------------------------------------------------------------
// Propagation helpers
// ---------------------------------------------------------------------------
// Row phase G'_k = exp(+i w_k dt / hbar) G_k (calcIR.cu:1512-1514) on a
// row-major n x (2 ncols) block [Re | Im].
template <typename T>
__global__ void __launch_bounds__(kSpNT)
phase_kernel(int n, int ncols, int ldg, double dt, const double* __restrict__ w, T* __restrict__ G) {
    const int64_t idx = int64_t(blockIdx.x) * kSpNT + threadIdx.x;
    if (idx >= int64_t(n) * ncols) return;
    const int k = int(idx / ncols), c = int(idx % ncols);
    double s, co;
    sincos(w[k] * dt / kHbar, &s, &co);
    T* row = G + size_t(k) * ldg;
    const T gr = row[c], gi = row[ncols + c];
    row[c] = T(co) * gr - T(s) * gi;
    row[ncols + c] = T(s) * gr + T(co) * gi;
}

// Split-K reduction fused with the row phase: G'_k = exp(+i w_k dt/hbar)
// sum_s P[s]_k, P holding nsplit partial n x (2 ncols) products.
template <typename T>
__global__ void __launch_bounds__(kSpNT)
phase_reduce_kernel(int n, int ncols, int nsplit, double dt, const double* __restrict__ w,
                    const T* __restrict__ P, T* __restrict__ G) {
    const int64_t idx = int64_t(blockIdx.x) * kSpNT + threadIdx.x;
    if (idx >= int64_t(n) * ncols) return;
    const int k = int(idx / ncols), c = int(idx % ncols);
    const size_t stride = size_t(n) * 2 * ncols;
    const size_t o = size_t(k) * 2 * ncols + c;
    T gr = T(0), gi = T(0);
    for (int sp = 0; sp < nsplit; ++sp) {
        gr += P[sp * stride + o];
        gi += P[sp * stride + o + ncols];
    }
    double s, co;
    sincos(w[k] * dt / kHbar, &s, &co);
    G[o] = T(co) * gr - T(s) * gi;
    G[o + ncols] = T(s) * gr + T(co) * gi;
}

// Y[i][c] = sum_s P[s][i][c] for the 2 ncols columns of a split-K product.
template <typename T>
__global__ void __launch_bounds__(kSpNT)
sum_partials_kernel(int n, int width, int nsplit, int ldy, const T* __restrict__ P, T* __restrict__ Y) {
    const int64_t idx = int64_t(blockIdx.x) * kSpNT + threadIdx.x;
    if (idx >= int64_t(n) * width) return;
    const int i = int(idx / width), c = int(idx % width);
    const size_t stride = size_t(n) * width;
    T acc = T(0);
    for (int sp = 0; sp < nsplit; ++sp) acc += P[sp * stride + size_t(i) * width + c];
    Y[size_t(i) * ldy + c] = acc;
}

// F = I as [Re | Im] (calcIR.cu:1521-1535).
template <typename T>
__global__ void __launch_bounds__(kSpNT) identity_kernel(int n, int ldf, T* __restrict__ F) {
    const int64_t idx = int64_t(blockIdx.x) * kSpNT + threadIdx.x;
    if (idx >= int64_t(n) * 2 * n) return;
    const int r = int(idx / (2 * n)), c = int(idx % (2 * n));
    F[size_t(r) * ldf + c] = (c == r) ? T(1) : T(0);
}

// Per-slot state update after a propagation step.  Actions:
// 0 idle -> zero, 1 origin -> M(t) (real), 2 propagated -> keep, 3 hold -> old.
template <typename T>
__global__ void __launch_bounds__(kSpNT)
slot_fixup_kernel(int n, int nslot, int ldy, const int* __restrict__ action, const T* __restrict__ sites,
                  const T* __restrict__ Yold, T* __restrict__ Ynew) {
    const int ncols = 9 * nslot;
    const int64_t idx = int64_t(blockIdx.x) * kSpNT + threadIdx.x;
    if (idx >= int64_t(n) * ncols) return;
    const int i = int(idx / ncols), c = int(idx % ncols);
    const int s = c / 9, a = c % 9;
    const int act = action[s];
    T* yr = Ynew + size_t(i) * ldy + c;
    T* yi = yr + ncols;
    if (act == 0) {
        *yr = T(0);
        *yi = T(0);
    } else if (act == 1) {
        *yr = sites[size_t(CIR_SITE_MU + a) * n + i];
        *yi = T(0);
    } else if (act == 3) {
        *yr = Yold[size_t(i) * ldy + c];
        *yi = Yold[size_t(i) * ldy + ncols + c];
    }
}

// (b, a) pairs of calcIR.cu:628-733, value = M_b(t)^T F M_a(0), with the nine
// vectors ordered mu_x, mu_y, mu_z, a_xx, a_yy, a_zz, a_xy, a_yz, a_zx:
//   p = 0..2 IR (a,a); 3..5 iiFii (a,a); 6..8 ijFij (a,a);
//   p = 9..14 iiFjj: (4,3) (5,3) (3,4) (5,4) (3,5) (4,5).
// One row's terms of the 15 forms (Re, Im) into acc.
template <typename T>
__device__ __forceinline__ void corr_row(const T (&m)[9], const T (&yr)[9], const T (&yi)[9], double (&acc)[30]) {
#pragma unroll
    for (int p = 0; p < 15; ++p) {
        const int pb = (p < 9) ? p : ((p == 9 || p == 14) ? 4 : (p == 10 || p == 12) ? 5 : 3);
        const int pa = (p < 9) ? p : ((p == 9 || p == 10) ? 3 : (p == 11 || p == 12) ? 4 : 5);
        acc[2 * p] += double(m[pb] * yr[pa]);
        acc[2 * p + 1] += double(m[pb] * yi[pa]);
    }
}

// The block's sums of acc, combined into IR, VV and VH (calcIR.cu:646-648,
// 736-745) and stored at o[0..5].
template <int NT>
__device__ __forceinline__ void corr_finish(double (&acc)[30], double (*red)[30], double* o) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
#pragma unroll
    for (int p = 0; p < 30; ++p) {
        const double v = wave_sum(acc[p]);
        if (lane == 0) red[wave][p] = v;
    }
    __syncthreads();
    if (tid == 0) {
        double t[30];
        for (int p = 0; p < 30; ++p) {
            t[p] = 0.0;
            for (int w = 0; w < NT / 64; ++w) t[p] += red[w][p];
        }
        double ir[2] = {0, 0}, iiii[2] = {0, 0}, ijij[2] = {0, 0}, iijj[2] = {0, 0};
        for (int c = 0; c < 2; ++c) {
            for (int p = 0; p < 3; ++p) ir[c] += t[2 * p + c];
            for (int p = 3; p < 6; ++p) iiii[c] += t[2 * p + c];
            for (int p = 6; p < 9; ++p) ijij[c] += t[2 * p + c];
            for (int p = 9; p < 15; ++p) iijj[c] += t[2 * p + c];
        }
        for (int c = 0; c < 2; ++c) {
            o[c] = ir[c];
            o[2 + c] = (3.0 * iiii[c] + iijj[c] + 4.0 * ijij[c]) / 15.0;
            o[4 + c] = (2.0 * iiii[c] - iijj[c] + 6.0 * ijij[c]) / 30.0;
        }
    }
}

// One block per slot: the 15 bilinear forms, stored at out[6*target[s] ...].
template <typename T>
__global__ void __launch_bounds__(kSpNT)
correlate_kernel(int n, int nslot, int ldy, const int* __restrict__ action, const int* __restrict__ target,
                 const T* __restrict__ sites, const T* __restrict__ Y, double* __restrict__ out) {
    __shared__ double red[kSpNT / 64][30];
    const int s = blockIdx.x, tid = threadIdx.x;
    const int act = action ? action[s] : 2;
    if (act != 1 && act != 2) return;
    const int ncols = 9 * nslot;
    double acc[30];
#pragma unroll
    for (int p = 0; p < 30; ++p) acc[p] = 0.0;
    for (int i = tid; i < n; i += kSpNT) {
        const T* yr = Y + size_t(i) * ldy + 9 * s;
        const T* yi = yr + ncols;
        T m[9], vr[9], vi[9];
#pragma unroll
        for (int b = 0; b < 9; ++b) {
            m[b] = sites[size_t(CIR_SITE_MU + b) * n + i];
            vr[b] = yr[b];
            vi[b] = yi[b];
        }
        corr_row(m, vr, vi, acc);
    }
    corr_finish<kSpNT>(acc, red, out + 6 * size_t(target ? target[s] : 0));
}

// slot_fixup_kernel and correlate_kernel in one launch, one block per slot:
// each row's nine columns of the slot are set as the action says (idle: zero,
// origin: M(t), hold: the old state; propagated: left as the propagation wrote
// them) and, for an origin or a propagated slot, the forms summed from the
// same values in the same order as correlate_kernel sums them.
template <typename T>
__global__ void __launch_bounds__(kSpNT)
slot_correlate_kernel(int n, int nslot, int ldy, const int* __restrict__ action, const int* __restrict__ target,
                      const T* __restrict__ sites, const T* __restrict__ Yold, T* __restrict__ Ynew,
                      double* __restrict__ out) {
    __shared__ double red[kSpNT / 64][30];
    const int s = blockIdx.x, tid = threadIdx.x;
    const int act = action[s];
    const int ncols = 9 * nslot;
    double acc[30];
#pragma unroll
    for (int p = 0; p < 30; ++p) acc[p] = 0.0;
    for (int i = tid; i < n; i += kSpNT) {
        T* yr = Ynew + size_t(i) * ldy + 9 * s;
        T* yi = yr + ncols;
        if (act == 0 || act == 3) {
            const T* orr = Yold + size_t(i) * ldy + 9 * s;
#pragma unroll
            for (int b = 0; b < 9; ++b) {
                yr[b] = (act == 3) ? orr[b] : T(0);
                yi[b] = (act == 3) ? orr[ncols + b] : T(0);
            }
            continue;
        }
        T m[9], vr[9], vi[9];
#pragma unroll
        for (int b = 0; b < 9; ++b) m[b] = sites[size_t(CIR_SITE_MU + b) * n + i];
        if (act == 1) {
#pragma unroll
            for (int b = 0; b < 9; ++b) {
                vr[b] = m[b];
                vi[b] = T(0);
                yr[b] = vr[b];
                yi[b] = T(0);
            }
        } else {
#pragma unroll
            for (int b = 0; b < 9; ++b) {
                vr[b] = yr[b];
                vi[b] = yi[b];
            }
        }
        corr_row(m, vr, vi, acc);
    }
    if (act != 1 && act != 2) return;
    corr_finish<kSpNT>(acc, red, out + 6 * size_t(target[s]));
}

// cuFFT C2R contract (calcIR.cu:790-791): y[k] = Re X0 + 2 sum_j Re(X_j
// e^{+2 pi i jk/n}) over j = 1..min(nx-1, n/2), X[n/2] taken as real.
__global__ void __launch_bounds__(kSpNT)
c2r_kernel(int nfft, int nx, const double* __restrict__ X, double* __restrict__ y) {
    const int sp = blockIdx.y;
    const int k = blockIdx.x * kSpNT + threadIdx.x;
    if (k >= nfft) return;
    const double* x = X + 2 * size_t(sp) * nx;
    const int half = nfft / 2;
    const int jmax = min(nx - 1, half);
    double acc = nx > 0 ? x[0] : 0.0;
    for (int j = 1; j <= jmax; ++j) {
        const long long jk = (static_cast<long long>(j) * k) % nfft;
        double s, c;
        sincospi(2.0 * double(jk) / double(nfft), &s, &c);
        if ((nfft % 2 == 0) && j == half)
            acc += x[2 * j] * c;
        else
            acc += 2.0 * (x[2 * j] * c - x[2 * j + 1] * s);
    }
    y[size_t(sp) * nfft + k] = acc;
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
int ipr_hist(int prec, int batch, int n, int ld, int nchrom_mol, float avef, float omegaStart,
             int omegaStep, int nomega, const void* V, const double* w, const double* weight, double* pw,
             double* rw, double* rmw, long long* skipped, hipStream_t s, void* det) {
    if (batch <= 0 || n <= 0 || ld < n || nomega <= 0 || omegaStep == 0 || nchrom_mol <= 0)
        return fail(CIR_ERR_ARG, "ipr_hist: bad sizes");
    if (det) {
        double* mval = static_cast<double*>(det);
        int* mbin = reinterpret_cast<int*>(mval + 2 * size_t(batch) * n);
        dim3 grid(unsigned(ceil_div(n, kSpNT)), unsigned(batch));
        if (prec == CIR_PREC_FP64)
            hipLaunchKernelGGL(ipr_modes_kernel<double>, grid, dim3(kSpNT), 0, s, n, ld, nchrom_mol, avef, omegaStart,
                               omegaStep, nomega, static_cast<const double*>(V), w, weight, mbin, mval, skipped);
        else
            hipLaunchKernelGGL(ipr_modes_kernel<float>, grid, dim3(kSpNT), 0, s, n, ld, nchrom_mol, avef, omegaStart,
                               omegaStep, nomega, static_cast<const float*>(V), w, weight, mbin, mval, skipped);
        CIR_CHECK_LAUNCH();
        hipLaunchKernelGGL(ipr_bins_kernel, dim3(unsigned(ceil_div(nomega, kSpNT))), dim3(kSpNT), 0, s, n, batch,
                           omegaStep, nomega, weight, mbin, mval, pw, rw, rmw);
        CIR_CHECK_LAUNCH();
        return CIR_OK;
    }
    const size_t sm = 3 * size_t(nomega) * sizeof(double);
    const int use_lds = sm <= 96 * 1024 ? 1 : 0;
    dim3 grid(unsigned(ceil_div(n, kSpNT)), unsigned(batch));
    if (prec == CIR_PREC_FP64) {
        CIR_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&ipr_hist_kernel<double>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, int(use_lds ? sm : 0)));
        hipLaunchKernelGGL(ipr_hist_kernel<double>, grid, dim3(kSpNT), use_lds ? sm : 0, s, n, ld, nchrom_mol,
                           avef, omegaStart, omegaStep, nomega, static_cast<const double*>(V), w, weight, pw,
                           rw, rmw, skipped, use_lds);
    } else {
        CIR_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&ipr_hist_kernel<float>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, int(use_lds ? sm : 0)));
        hipLaunchKernelGGL(ipr_hist_kernel<float>, grid, dim3(kSpNT), use_lds ? sm : 0, s, n, ld, nchrom_mol,
                           avef, omegaStart, omegaStep, nomega, static_cast<const float*>(V), w, weight, pw, rw,
                           rmw, skipped, use_lds);
    }
    CIR_CHECK_LAUNCH();
    return CIR_OK;
}

int spectral_density(int prec, int batch, int n, int ld, float t1, float avef, float omegaStart,
                     int omegaStep, int nomega, const void* V, const double* w, const void* sites,
                     const double* weight, double* sw, double* m2work, hipStream_t s, void* det) {
    if (batch <= 0 || n <= 0 || ld < n || nomega <= 0) return fail(CIR_ERR_ARG, "spectral_density: bad sizes");
    dim3 g1(unsigned(ceil_div(n, kSpNT)), unsigned(batch));
    if (prec == CIR_PREC_FP64)
        hipLaunchKernelGGL(mu2_kernel<double>, g1, dim3(kSpNT), 0, s, n, ld, static_cast<const double*>(V),
                           static_cast<const double*>(sites), weight, m2work);
    else
        hipLaunchKernelGGL(mu2_kernel<float>, g1, dim3(kSpNT), 0, s, n, ld, static_cast<const float*>(V),
                           static_cast<const float*>(sites), weight, m2work);
    CIR_CHECK_LAUNCH();
    const double gamma = kHbar / (double(t1) * 2.0);  // calcIR.cu:1426
    dim3 g2(unsigned(ceil_div(nomega, kSpNT)), unsigned(batch));
    double* per_frame = static_cast<double*>(det);
    hipLaunchKernelGGL(sw_kernel, g2, dim3(kSpNT), 0, s, n, nomega, omegaStart, omegaStep, gamma, avef, m2work,
                       w, weight, sw, per_frame);
    CIR_CHECK_LAUNCH();
    if (per_frame) {
        hipLaunchKernelGGL(sw_frames_kernel, dim3(unsigned(ceil_div(nomega, kSpNT))), dim3(kSpNT), 0, s, batch, nomega,
                           per_frame, sw);
        CIR_CHECK_LAUNCH();
    }
    return CIR_OK;
}

// Number of K slices of the propagation products: enough workgroups to
// cover the chip (>= ~512) for the skinny origin products, 1 for the
// n x 2n F-matrix step.
// ---------------------------------------------------------------------------
// Propagation products of the batched path, Y(t+dt) = V exp(+i w dt/hbar)
// V^T Y(t): the F update of calcIR.cu:597-617 applied to the dipole and
// polarizability columns that calcIR.cu:628-733 multiplies by F.  n x n times
// n x 2 ncols with a few hundred columns at most, which the tiled GEMM splits
// poorly.  One
// 32 x 32 output block per workgroup over all of K: rows m0..m0+31, the Re
// columns 16 t..16 t+15 and the matching Im columns ncols + 16 t.., so that
// the row phase can be applied to (Re, Im) pairs held by lanes l and l ^ 16.
// The four waves take a quarter of K each and are summed through LDS.
// Blocks are numbered row tile fastest over a row-tile count padded to a
// multiple of 8, so every column tile of a row stripe lands on the same XCD
// and reads the stripe of V from one L2.
//   TRANS: op(V)(m, k) = V[k*ld + m]  (the projection G = V^T Y)
//   else:  op(V)(m, k) = V[m*ld + k]  (the return Y = V G')
//   PHASE: rotate row m by exp(+i w_m dt / hbar) before the store.
//   SPLIT3: split-BF16x3 MFMA 32x32x16 (a = hi + lo, a b ~ lo hi + hi lo +
//   hi hi); else FP32 MFMA 32x32x2.
// ---------------------------------------------------------------------------
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
constexpr int kPrNT = 256;
constexpr int kPrMaxCols = 512;  // wider products (the F-matrix form) use the tiled GEMM

template <bool TRANS, bool PHASE, bool SPLIT3>
__global__ void __launch_bounds__(kPrNT)
prop_mfma_kernel(int n, int ncols, int nmt_pad, int ld, const float* __restrict__ V, const float* __restrict__ B,
                 int ldb, float* __restrict__ C, int ldc, double dt, const double* __restrict__ w, int kseg,
                 int vecA) {
    __shared__ float red[3][16][64];
    __shared__ float ph[2][32];
    const int mt = blockIdx.x % nmt_pad, nt = blockIdx.x / nmt_pad;
    const int m0 = mt * 32;
    if (m0 >= n) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int h = lane >> 5, l32 = lane & 31;
    const int cc = 16 * nt + (l32 & 15);  // complex column of this lane's operand / result
    const bool cvalid = cc < ncols;
    const int bcol = (l32 < 16) ? cc : ncols + cc;
    const int m = m0 + l32;
    const bool mvalid = m < n;
    // K segment of this workgroup (blockIdx.y; kseg is a multiple of 64),
    // split in quarters over the 4 waves.  With more than one segment the
    // partial results are added into C (zeroed beforehand).
    const bool split_k = gridDim.y > 1;
    const int kq = kseg / 4;
    const int kbeg = blockIdx.y * kseg + wave * kq, kend = min(n, kbeg + kq);
    floatx16 acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.f;
    // four K steps of 16 per pass: the 64 loads of a pass are in flight together
    constexpr int kU = 4;
    for (int k0 = kbeg; k0 < kend; k0 += 16 * kU) {
        float av[kU][8], bv[kU][8];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int kb = k0 + 16 * u + 8 * h;
            if (!TRANS && vecA && mvalid && kb + 8 <= kend) {
                // V(m, kb..kb+7) is 32 contiguous, 16-byte aligned bytes of
                // row m (V aligned, ld and kb multiples of 4): two vector loads
                const float4* src = reinterpret_cast<const float4*>(V + size_t(m) * ld + kb);
                const float4 x0 = src[0], x1 = src[1];
                av[u][0] = x0.x;
                av[u][1] = x0.y;
                av[u][2] = x0.z;
                av[u][3] = x0.w;
                av[u][4] = x1.x;
                av[u][5] = x1.y;
                av[u][6] = x1.z;
                av[u][7] = x1.w;
            } else {
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    const int k = kb + q;
                    av[u][q] = (mvalid && k < kend) ? (TRANS ? V[size_t(k) * ld + m] : V[size_t(m) * ld + k]) : 0.f;
                }
            }
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int k = kb + q;
                bv[u][q] = (cvalid && k < kend) ? B[size_t(k) * ldb + bcol] : 0.f;
            }
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            if constexpr (SPLIT3) {
                bf16x8 ahi, alo, bhi, blo;
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    const __bf16 a = __bf16(av[u][q]), b = __bf16(bv[u][q]);
                    ahi[q] = a;
                    alo[q] = __bf16(av[u][q] - float(a));
                    bhi[q] = b;
                    blo[q] = __bf16(bv[u][q] - float(b));
                }
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(alo, bhi, acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi, blo, acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi, bhi, acc, 0, 0, 0);
            } else {
                // instruction q sums k = 16 u + q (lanes 0..31) and 16 u + 8 + q (lanes 32..63)
#pragma unroll
                for (int q = 0; q < 8; ++q)
                    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av[u][q], bv[u][q], acc, 0, 0, 0);
            }
        }
    }
    if (wave > 0) {
#pragma unroll
        for (int r = 0; r < 16; ++r) red[wave - 1][r][lane] = acc[r];
    }
    if (PHASE && tid < 32) {
        const int mm = m0 + tid;
        double sn = 0.0, co = 1.0;
        if (mm < n) sincos(w[mm] * dt / kHbar, &sn, &co);
        ph[0][tid] = float(co);
        ph[1][tid] = float(sn);
    }
    __syncthreads();
    if (wave == 0) {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int i = 8 * (r >> 2) + 4 * h + (r & 3);
            float v = acc[r] + red[0][r][lane] + red[1][r][lane] + red[2][r][lane];
            if (PHASE) {
                const float o = __shfl_xor(v, 16, 64);  // the other part of the same complex entry
                const float co = ph[0][i], sn = ph[1][i];
                v = (l32 < 16) ? co * v - sn * o : sn * o + co * v;
            }
            const int mm = m0 + i;
            if (mm < n && cvalid) {
                if (split_k)
                    atomicAdd(&C[size_t(mm) * ldc + bcol], v);
                else
                    C[size_t(mm) * ldc + bcol] = v;
            }
        }
    }
}

// C(0:n, 0:w2) = 0 for the split-K accumulation (row pitch ldc).
__global__ void __launch_bounds__(256) zero_rows_kernel(int n, int w2, int ldc, float* __restrict__ C) {
    const int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (e >= int64_t(n) * w2) return;
    C[size_t(e / w2) * ldc + size_t(e % w2)] = 0.f;
}

// Both split-K outputs zeroed in one launch: rows 0..n-1 of C0 (pitch ld0)
// and of C1 (pitch ld1), w2 columns each.
__global__ void __launch_bounds__(256) zero_rows2_kernel(int n, int w2, int ld0, float* __restrict__ C0, int ld1,
                                                         float* __restrict__ C1) {
    const int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x, ne = int64_t(n) * w2;
    if (e >= 2 * ne) return;
    const int64_t q = (e < ne) ? e : e - ne;
    if (e < ne)
        C0[size_t(q / w2) * ld0 + size_t(q % w2)] = 0.f;
    else
        C1[size_t(q / w2) * ld1 + size_t(q % w2)] = 0.f;
}

// K split of the propagation products, picked per (n, ncols, product form)
// by timing the candidates once per process (the first call of a shape whose
// output does not alias its input); the default guess aims at about 4
// workgroups per CU with K segments of at least 256.
static int prop_pair(bool split3, int n, int ld, int ncols, int ldy, double dt, const float* V, const double* w,
                     const float* Yin, float* Yout, float* G, int ks, hipStream_t s);

// The K split of the two products: 1, 2, 4 or 8 segments (at most n / 256),
// timed on the first call of each shape (tuner.hpp; the products overwrite
// Yout from Yin, so they can be run again when Yout does not alias Yin).
static TuneKey prop_key(int n, int ncols, bool split3) { return TuneKey{"prop_ksplit", {n, ncols, split3 ? 1 : 0, 0, 0}}; }
static int prop_ncand(int n) {
    int c = 0;
    for (int k = 1; k <= std::min(kPropMaxSplit, std::max(1, n / 256)); k *= 2) ++c;
    return c;
}

static int prop_products(bool split3, int n, int ld, int ncols, int ldy, double dt, const float* V, const double* w,
                         const float* Yin, float* Yout, float* G, hipStream_t s, bool may_tune) {
    const int nmt_pad = int(ceil_div(ceil_div(n, 32), 8) * 8);
    const int tiles = nmt_pad * int(ceil_div(ncols, 16));
    const int guess = std::max(1, std::min({kPropMaxSplit, int(ceil_div(1024, tiles)), std::max(1, n / 256)}));
    const TuneKey key = prop_key(n, ncols, split3);
    int ks = deterministic_mode() ? 1 : guess;  // one K segment: no atomic accumulation
    const int kept = tune_lookup(key);
    if (kept >= 0) {
        ks = 1 << kept;
    } else if (may_tune && Yin != Yout && prop_ncand(n) > 1 && tune_allowed(s)) {
        int err = CIR_OK;
        const int c = tune_time(key, prop_ncand(n), s, [&](int i) {
            return prop_pair(split3, n, ld, ncols, ldy, dt, V, w, Yin, Yout, G, 1 << i, s);
        }, &err);
        if (err != CIR_OK) return err;
        ks = 1 << c;
    }
    return prop_pair(split3, n, ld, ncols, ldy, dt, V, w, Yin, Yout, G, ks, s);
}

// The two products with ks K segments (ks > 1: atomic accumulation into
// zeroed outputs).
static int prop_pair(bool split3, int n, int ld, int ncols, int ldy, double dt, const float* V, const double* w,
                     const float* Yin, float* Yout, float* G, int ks, hipStream_t s) {
    const int nmt_pad = int(ceil_div(ceil_div(n, 32), 8) * 8);
    const int tiles = nmt_pad * int(ceil_div(ncols, 16));
    const int kseg = int(ceil_div(ceil_div(n, ks), 64) * 64);
    ks = int(ceil_div(n, kseg));
    const dim3 grid(static_cast<unsigned>(tiles), static_cast<unsigned>(ks));
    const int w2 = 2 * ncols;
    const int vecA = (ld % 4 == 0 && (reinterpret_cast<uintptr_t>(V) & 15) == 0) ? 1 : 0;
    auto zero = [&](float* C, int ldc) -> int {
        if (ks == 1) return CIR_OK;
        hipLaunchKernelGGL(zero_rows_kernel, dim3(unsigned(ceil_div(int64_t(n) * w2, 256))), dim3(256), 0, s, n, w2,
                           ldc, C);
        CIR_CHECK_LAUNCH();
        return CIR_OK;
    };
    // distinct input and output: both accumulators are cleared up front in
    // one launch; in place, the output is cleared after the first product
    // has read it
    const bool inplace = Yin == Yout;
    if (ks > 1 && !inplace) {
        hipLaunchKernelGGL(zero_rows2_kernel, dim3(unsigned(ceil_div(2 * int64_t(n) * w2, 256))), dim3(256), 0, s, n,
                           w2, w2, G, ldy, Yout);
        CIR_CHECK_LAUNCH();
    } else {
        CIR_TRY(zero(G, w2));
    }
    if (split3)
        hipLaunchKernelGGL((prop_mfma_kernel<true, true, true>), grid, dim3(kPrNT), 0, s, n, ncols, nmt_pad, ld, V,
                           Yin, ldy, G, w2, dt, w, kseg, vecA);
    else
        hipLaunchKernelGGL((prop_mfma_kernel<true, true, false>), grid, dim3(kPrNT), 0, s, n, ncols, nmt_pad, ld, V,
                           Yin, ldy, G, w2, dt, w, kseg, vecA);
    CIR_CHECK_LAUNCH();
    if (inplace) CIR_TRY(zero(Yout, ldy));
    if (split3)
        hipLaunchKernelGGL((prop_mfma_kernel<false, false, true>), grid, dim3(kPrNT), 0, s, n, ncols, nmt_pad, ld, V,
                           G, w2, Yout, ldy, dt, w, kseg, vecA);
    else
        hipLaunchKernelGGL((prop_mfma_kernel<false, false, false>), grid, dim3(kPrNT), 0, s, n, ncols, nmt_pad, ld, V,
                           G, w2, Yout, ldy, dt, w, kseg, vecA);
    CIR_CHECK_LAUNCH();
    return CIR_OK;
}

static int prop_splits(int n, int ncols, int* kslice) {
    const int64_t tiles = ceil_div(n, 64) * ceil_div(2 * ncols, 64);
    int ns = int(std::min<int64_t>(kPropMaxSplit, std::max<int64_t>(1, ceil_div(512, tiles))));
    ns = std::max(1, std::min(ns, n / 128));
    int ks = int(ceil_div(ceil_div(n, ns), 32) * 32);
    ns = int(ceil_div(n, ks));
    *kslice = ks;
    return ns;
}

size_t propagate_work_elems(int n, int ncols) {
    int ks;
    const int ns = prop_splits(n, ncols, &ks);
    return size_t(n) * 2 * ncols * (ns > 1 ? ns + 1 : 1);
}

// One K-sliced product P[s] = A[:, slice s] B[slice s, :] (s < ns).
static int split_gemm(int gemm_prec, int M, int N, int K, int ks, int ns, size_t esz, const void* A, int64_t a_m,
                      int64_t a_k, const void* B, int64_t b_k, int64_t b_n, void* P, hipStream_t s) {
    const int nfull = K / ks, rem = K - nfull * ks;
    GemmArgs g;
    g.M = M;
    g.N = N;
    g.K = ks;
    g.batch = nfull;
    g.A = A;
    g.a_m = a_m;
    g.a_k = a_k;
    g.a_b = int64_t(ks) * a_k;
    g.B = B;
    g.b_k = b_k;
    g.b_n = b_n;
    g.b_b = int64_t(ks) * b_k;
    g.C = P;
    g.c_m = N;
    g.c_n = 1;
    g.c_b = int64_t(M) * N;
    if (nfull > 0) CIR_TRY(gemm(gemm_prec, g, s));
    if (rem > 0) {
        g.K = rem;
        g.batch = 1;
        g.A = static_cast<const char*>(A) + size_t(nfull) * ks * a_k * esz;
        g.B = static_cast<const char*>(B) + size_t(nfull) * ks * b_k * esz;
        g.C = static_cast<char*>(P) + size_t(nfull) * M * N * esz;
        CIR_TRY(gemm(gemm_prec, g, s));
    }
    (void)ns;
    return CIR_OK;
}

bool propagate_needs_tuning(int prec, int gemm_prec, int n, int ncols) {
    if (prec == CIR_PREC_FP64 || ncols > kPrMaxCols || prop_ncand(n) < 2) return false;
    return tune_lookup(prop_key(n, ncols, gemm_prec == CIR_PREC_BF16X3)) < 0 && tune_allowed(nullptr);
}

int propagate(int prec, int gemm_prec, int n, int ld, int ncols, int ldy, double dt, const void* V,
              const double* w, const void* Yin, void* Yout, void* G, hipStream_t s, bool may_tune) {
    if (n <= 0 || ncols <= 0 || ldy < 2 * ncols || ld < n) return fail(CIR_ERR_ARG, "propagate: bad sizes");
    const size_t esz = (prec == CIR_PREC_FP64) ? sizeof(double) : sizeof(float);
    const int w2 = 2 * ncols;
    int ks;
    const int ns = prop_splits(n, ncols, &ks);
    const int64_t tot = int64_t(n) * ncols;
    if (prec != CIR_PREC_FP64 && ncols <= kPrMaxCols)
        return prop_products(gemm_prec == CIR_PREC_BF16X3, n, ld, ncols, ldy, dt, static_cast<const float*>(V), w,
                             static_cast<const float*>(Yin), static_cast<float*>(Yout), static_cast<float*>(G), s,
                             may_tune);
    if (ns == 1) {
        // G = V^T Yin (V row-major: V(i,k) at V[i*ld+k]; A(m=k, kk=i)), phase, Yout = V G'
        GemmArgs g;
        g.M = n;
        g.N = w2;
        g.K = n;
        g.A = V;
        g.a_m = 1;
        g.a_k = ld;
        g.B = Yin;
        g.b_k = ldy;
        g.b_n = 1;
        g.C = G;
        g.c_m = w2;
        g.c_n = 1;
        CIR_TRY(gemm(gemm_prec, g, s));
        if (prec == CIR_PREC_FP64)
            hipLaunchKernelGGL(phase_kernel<double>, dim3(unsigned(ceil_div(tot, kSpNT))), dim3(kSpNT), 0, s, n,
                               ncols, w2, dt, w, static_cast<double*>(G));
        else
            hipLaunchKernelGGL(phase_kernel<float>, dim3(unsigned(ceil_div(tot, kSpNT))), dim3(kSpNT), 0, s, n, ncols,
                               w2, dt, w, static_cast<float*>(G));
        CIR_CHECK_LAUNCH();
        g = GemmArgs{};
        g.M = n;
        g.N = w2;
        g.K = n;
        g.A = V;
        g.a_m = ld;
        g.a_k = 1;
        g.B = G;
        g.b_k = w2;
        g.b_n = 1;
        g.C = Yout;
        g.c_m = ldy;
        g.c_n = 1;
        return gemm(gemm_prec, g, s);
    }
    // split-K: G holds G' (n x w2), P the ns partial products behind it
    void* P = static_cast<char*>(G) + size_t(n) * w2 * esz;
    CIR_TRY(split_gemm(gemm_prec, n, w2, n, ks, ns, esz, V, 1, ld, Yin, ldy, 1, P, s));
    if (prec == CIR_PREC_FP64)
        hipLaunchKernelGGL(phase_reduce_kernel<double>, dim3(unsigned(ceil_div(tot, kSpNT))), dim3(kSpNT), 0, s, n,
                           ncols, ns, dt, w, static_cast<const double*>(P), static_cast<double*>(G));
    else
        hipLaunchKernelGGL(phase_reduce_kernel<float>, dim3(unsigned(ceil_div(tot, kSpNT))), dim3(kSpNT), 0, s, n,
                           ncols, ns, dt, w, static_cast<const float*>(P), static_cast<float*>(G));
    CIR_CHECK_LAUNCH();
    CIR_TRY(split_gemm(gemm_prec, n, w2, n, ks, ns, esz, V, ld, 1, G, w2, 1, P, s));
    const int64_t tot2 = int64_t(n) * w2;
    if (prec == CIR_PREC_FP64)
        hipLaunchKernelGGL(sum_partials_kernel<double>, dim3(unsigned(ceil_div(tot2, kSpNT))), dim3(kSpNT), 0, s, n, w2,
                           ns, ldy, static_cast<const double*>(P), static_cast<double*>(Yout));
    else
        hipLaunchKernelGGL(sum_partials_kernel<float>, dim3(unsigned(ceil_div(tot2, kSpNT))), dim3(kSpNT), 0, s, n, w2,
                           ns, ldy, static_cast<const float*>(P), static_cast<float*>(Yout));
    CIR_CHECK_LAUNCH();
    return CIR_OK;
}

int set_identity(int prec, int n, int ldf, void* F, hipStream_t s) {
    const int64_t tot = int64_t(n) * 2 * n;
    if (prec == CIR_PREC_FP64)
        hipLaunchKernelGGL(identity_kernel<double>, dim3(unsigned(ceil_div(tot, kSpNT))), dim3(kSpNT), 0, s, n,
                           ldf, static_cast<double*>(F));
    else
        hipLaunchKernelGGL(identity_kernel<float>, dim3(unsigned(ceil_div(tot, kSpNT))), dim3(kSpNT), 0, s, n, ldf,
                           static_cast<float*>(F));
    CIR_CHECK_LAUNCH();
    return CIR_OK;
}

int slot_fixup(int prec, int n, int nslot, int ldy, const int* action, const void* sites, const void* Yold,
               void* Ynew, hipStream_t s) {
    const int64_t tot = int64_t(n) * 9 * nslot;
    if (prec == CIR_PREC_FP64)
        hipLaunchKernelGGL(slot_fixup_kernel<double>, dim3(unsigned(ceil_div(tot, kSpNT))), dim3(kSpNT), 0, s, n,
                           nslot, ldy, action, static_cast<const double*>(sites), static_cast<const double*>(Yold),
                           static_cast<double*>(Ynew));
    else
        hipLaunchKernelGGL(slot_fixup_kernel<float>, dim3(unsigned(ceil_div(tot, kSpNT))), dim3(kSpNT), 0, s, n,
                           nslot, ldy, action, static_cast<const float*>(sites), static_cast<const float*>(Yold),
                           static_cast<float*>(Ynew));
    CIR_CHECK_LAUNCH();
    return CIR_OK;
}

int slot_correlate(int prec, int n, int nslot, int ldy, const int* action, const int* target, const void* sites,
                   const void* Yold, void* Ynew, double* out, hipStream_t s) {
    if (prec == CIR_PREC_FP64)
        hipLaunchKernelGGL(slot_correlate_kernel<double>, dim3(nslot), dim3(kSpNT), 0, s, n, nslot, ldy, action,
                           target, static_cast<const double*>(sites), static_cast<const double*>(Yold),
                           static_cast<double*>(Ynew), out);
    else
        hipLaunchKernelGGL(slot_correlate_kernel<float>, dim3(nslot), dim3(kSpNT), 0, s, n, nslot, ldy, action,
                           target, static_cast<const float*>(sites), static_cast<const float*>(Yold),
                           static_cast<float*>(Ynew), out);
    CIR_CHECK_LAUNCH();
    return CIR_OK;
}

int correlate(int prec, int n, int nslot, int ldy, const int* action, const int* target, const void* sites,
              const void* Y, double* out, hipStream_t s) {
    if (prec == CIR_PREC_FP64)
        hipLaunchKernelGGL(correlate_kernel<double>, dim3(nslot), dim3(kSpNT), 0, s, n, nslot, ldy, action, target,
                           static_cast<const double*>(sites), static_cast<const double*>(Y), out);
    else
        hipLaunchKernelGGL(correlate_kernel<float>, dim3(nslot), dim3(kSpNT), 0, s, n, nslot, ldy, action, target,
                           static_cast<const float*>(sites), static_cast<const float*>(Y), out);
    CIR_CHECK_LAUNCH();
    return CIR_OK;
}

// ---------------------------------------------------------------------------
// Re-unitarisation of the propagated origins.  An origin's nine columns obey
// Y_s = F M(0) with F unitary, so its Gram matrix Y_s^H Y_s stays M(0)^T M(0)
// = G0_s; rounding in the products lets it drift.  With G = Y_s^H Y_s =
// G0 + D (D small, Hermitian), Y_s <- Y_s (I - G0^{-1} D / 2) restores the
// Gram matrix to G0 up to O(D^2).  mode[s]: 0 nothing, 1 record G0_s from Y_s
// (an origin, Y_s = M(0) real), 2 correct Y_s.
// ---------------------------------------------------------------------------
constexpr int kGramRows = 512;  // rows per workgroup of the Gram sums

// G[s] += sum over this workgroup's rows of conj(y_a) y_b, thread (a, b)
template <typename T>
__global__ void __launch_bounds__(128)
gram_kernel(int n, int nslot, int ldy, int rpb, const int* __restrict__ mode, const T* __restrict__ Y,
            double* __restrict__ G) {
    const int s = blockIdx.y, t = threadIdx.x;
    if (mode[s] == 0 || t >= 81) return;
    const int a = t / 9, b = t % 9, ncols = 9 * nslot;
    const int ca = 9 * s + a, cb = 9 * s + b;
    const int i0 = blockIdx.x * rpb, i1 = min(n, i0 + rpb);
    double re = 0.0, im = 0.0;
    for (int i = i0; i < i1; ++i) {
        const T* y = Y + size_t(i) * ldy;
        const double ar = double(y[ca]), ai = double(y[ncols + ca]);
        const double br = double(y[cb]), bi = double(y[ncols + cb]);
        re += ar * br + ai * bi;  // conj(y_a) y_b
        im += ar * bi - ai * br;
    }
    atomicAdd(&G[(size_t(s) * 81 + t) * 2], re);
    atomicAdd(&G[(size_t(s) * 81 + t) * 2 + 1], im);
}

// one thread per slot: record G0 (mode 1) or form C = I - G0^{-1} (G - G0)/2
// (mode 2) through the Cholesky factor of the real G0
__global__ void __launch_bounds__(64)
reunit_coef_kernel(int nslot, const int* __restrict__ mode, const double* __restrict__ G, double* __restrict__ G0,
                   double* __restrict__ C) {
    const int s = blockIdx.x * 64 + threadIdx.x;
    if (s >= nslot) return;
    const double* g = G + size_t(s) * 162;
    double* g0 = G0 + size_t(s) * 81;
    if (mode[s] == 1) {
        for (int e = 0; e < 81; ++e) g0[e] = g[2 * e];
        return;
    }
    if (mode[s] != 2) return;
    double L[81];
    for (int e = 0; e < 81; ++e) L[e] = g0[e];
    for (int k = 0; k < 9; ++k) {  // lower Cholesky factor in place
        double d = L[k * 9 + k];
        for (int q = 0; q < k; ++q) d -= L[k * 9 + q] * L[k * 9 + q];
        d = sqrt(fmax(d, 1e-300));
        L[k * 9 + k] = d;
        for (int r = k + 1; r < 9; ++r) {
            double v = L[r * 9 + k];
            for (int q = 0; q < k; ++q) v -= L[r * 9 + q] * L[k * 9 + q];
            L[r * 9 + k] = v / d;
        }
    }
    double* c = C + size_t(s) * 162;
    for (int b = 0; b < 9; ++b)
        for (int part = 0; part < 2; ++part) {
            double x[9];
            for (int a = 0; a < 9; ++a) x[a] = g[(a * 9 + b) * 2 + part] - (part == 0 ? g0[a * 9 + b] : 0.0);
            for (int r = 0; r < 9; ++r) {  // L y = d
                double v = x[r];
                for (int q = 0; q < r; ++q) v -= L[r * 9 + q] * x[q];
                x[r] = v / L[r * 9 + r];
            }
            for (int r = 8; r >= 0; --r) {  // L^T x = y
                double v = x[r];
                for (int q = r + 1; q < 9; ++q) v -= L[q * 9 + r] * x[q];
                x[r] = v / L[r * 9 + r];
            }
            for (int a = 0; a < 9; ++a) c[(a * 9 + b) * 2 + part] = ((a == b && part == 0) ? 1.0 : 0.0) - 0.5 * x[a];
        }
}

// Y_s(i, :) <- Y_s(i, :) C_s for the slots with mode 2
template <typename T>
__global__ void __launch_bounds__(kSpNT)
reunit_apply_kernel(int n, int nslot, int ldy, const int* __restrict__ mode, const double* __restrict__ C,
                    T* __restrict__ Y) {
    const int64_t idx = int64_t(blockIdx.x) * kSpNT + threadIdx.x;
    if (idx >= int64_t(n) * nslot) return;
    const int i = int(idx / nslot), s = int(idx % nslot);
    if (mode[s] != 2) return;
    const int ncols = 9 * nslot;
    T* yr = Y + size_t(i) * ldy + 9 * s;
    T* yi = yr + ncols;
    double ar[9], ai[9];
    for (int a = 0; a < 9; ++a) {
        ar[a] = double(yr[a]);
        ai[a] = double(yi[a]);
    }
    const double* c = C + size_t(s) * 162;
    for (int b = 0; b < 9; ++b) {
        double re = 0.0, im = 0.0;
        for (int a = 0; a < 9; ++a) {
            const double cr = c[(a * 9 + b) * 2], ci = c[(a * 9 + b) * 2 + 1];
            re += ar[a] * cr - ai[a] * ci;
            im += ar[a] * ci + ai[a] * cr;
        }
        yr[b] = T(re);
        yi[b] = T(im);
    }
}

__global__ void __launch_bounds__(256) zero_doubles_kernel(int64_t cnt, double* __restrict__ p) {
    const int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (e < cnt) p[e] = 0.0;
}

size_t reunit_work_doubles(int nslot) { return size_t(nslot) * (162 + 162); }

int reunitarise(int prec, int n, int nslot, int ldy, const int* mode, double* G0, void* Y, double* work,
                hipStream_t s) {
    if (n <= 0 || nslot <= 0 || ldy < 18 * nslot || !mode || !G0 || !Y || !work)
        return fail(CIR_ERR_ARG, "reunitarise: bad arguments");
    double* G = work;
    double* C = work + size_t(nslot) * 162;
    const int64_t ng = int64_t(nslot) * 162;
    hipLaunchKernelGGL(zero_doubles_kernel, dim3(unsigned(ceil_div(ng, 256))), dim3(256), 0, s, ng, G);
    CIR_CHECK_LAUNCH();
    // rows per workgroup: all of them in deterministic mode (one adder per
    // entry, so the atomic sums have a fixed order)
    const int rpb = deterministic_mode() ? n : kGramRows;
    const dim3 gg(unsigned(ceil_div(n, rpb)), unsigned(nslot));
    if (prec == CIR_PREC_FP64)
        hipLaunchKernelGGL(gram_kernel<double>, gg, dim3(128), 0, s, n, nslot, ldy, rpb, mode,
                           static_cast<const double*>(Y), G);
    else
        hipLaunchKernelGGL(gram_kernel<float>, gg, dim3(128), 0, s, n, nslot, ldy, rpb, mode,
                           static_cast<const float*>(Y), G);
    CIR_CHECK_LAUNCH();
    hipLaunchKernelGGL(reunit_coef_kernel, dim3(unsigned(ceil_div(nslot, 64))), dim3(64), 0, s, nslot, mode, G, G0, C);
    CIR_CHECK_LAUNCH();
    const int64_t tot = int64_t(n) * nslot;
    if (prec == CIR_PREC_FP64)
        hipLaunchKernelGGL(reunit_apply_kernel<double>, dim3(unsigned(ceil_div(tot, kSpNT))), dim3(kSpNT), 0, s, n,
                           nslot, ldy, mode, C, static_cast<double*>(Y));
    else
        hipLaunchKernelGGL(reunit_apply_kernel<float>, dim3(unsigned(ceil_div(tot, kSpNT))), dim3(kSpNT), 0, s, n,
                           nslot, ldy, mode, C, static_cast<float*>(Y));
    CIR_CHECK_LAUNCH();
    return CIR_OK;
}

int spectrum_c2r(int nspec, int nfft, int nx, const double* X, double* y, hipStream_t s) {
    if (nspec <= 0 || nfft <= 0 || nx < 0 || nspec > 65535) return fail(CIR_ERR_ARG, "c2r: bad sizes");
    hipLaunchKernelGGL(c2r_kernel, dim3(unsigned(ceil_div(nfft, kSpNT)), unsigned(nspec)), dim3(kSpNT), 0, s, nfft,
                       nx, X, y);
    CIR_CHECK_LAUNCH();
    return CIR_OK;
}

}  // namespace cir
