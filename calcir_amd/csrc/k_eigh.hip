// k_eigh.hip -- batched symmetric eigensolver for gfx950, replacing
// magma_ssyevd_gpu (calcIR.cu:420-448).
//
//   1. tridiag_kernel: Householder reduction A = Q T Q^T, one workgroup per
//      matrix (a frame's Hamiltonian), panel-blocked: within a panel of
//      kTriNB columns the two-sided update is deferred and folded into each
//      column and each symmetric matrix-vector product through the panel's
//      V and W (A_now = A_panel - V W^T - W V^T); the trailing matrix is
//      updated once per panel.  Full symmetric storage, column-major.  The
//      reflector v_j (v_j[j+1] = 1) replaces column j of A below the
//      diagonal; rows 0..j of column j are zeroed.
//   2. bisect_kernel: every eigenvalue of T by FP64 Sturm-sequence
//      bisection, one lane per eigenvalue (eigenvalues come out ascending).
//   3. invit_kernel: eigenvectors of T by FP64 inverse iteration with
//      partial pivoting, one lane per eigenvalue; neighbours closer than
//      1e-7 * max(|lambda|, 1) form a cluster (at most kMaxCluster members)
//      that its first lane solves sequentially with modified Gram-Schmidt
//      against the cluster's earlier vectors.  Both this tolerance and the
//      bisection tolerance are relative to the eigenvalue itself, so a few
//      outlying diagonal entries (strong-field chromophores) neither merge the
//      bulk of the spectrum into one cluster nor coarsen its eigenvalues.
//   4. back-transform V = Q Z in compact-WY blocks of kBtNB reflectors on
//      the MFMA product (gemm.hpp): Z <- Z - Y (T (Y^T Z)), last block first.

#include <cfloat>

#include "gemm.hpp"

namespace cir {

constexpr int kTriNT = 512;  // threads per matrix in the tridiagonalisation
constexpr int kTriNW = kTriNT / 64;
constexpr int kTriNB = 32;   // panel width
constexpr int kTriRQ = 8;    // rows per thread per pass of the symmetric product
constexpr int kBtNB = 128;   // reflectors per compact-WY block in the back-transform
constexpr int kEigNT = 256;  // threads per block of bisection / inverse iteration
constexpr double kClusterRel = 1e-7;
constexpr int kMaxCluster = 128;

// Eigenvalue k joins the cluster of k-1 (ascending w).
__device__ __forceinline__ bool joins_cluster(const double* w, int k) {
    const double scale = fmax(fmax(fabs(w[k]), fabs(w[k - 1])), 1.0);
    return w[k] - w[k - 1] <= kClusterRel * scale;
}

// 16-byte vector of T for the streaming loads of the symmetric product.
template <typename T>
struct Vec16;
template <>
struct Vec16<float> {
    using type = float4;
    static constexpr int n = 4;
};
template <>
struct Vec16<double> {
    using type = double2;
    static constexpr int n = 2;
};
template <typename V, typename T>
__device__ __forceinline__ void vfma(T (&acc)[4], const V& a, T s);
template <>
__device__ __forceinline__ void vfma<float4, float>(float (&acc)[4], const float4& a, float s) {
    acc[0] += a.x * s;
    acc[1] += a.y * s;
    acc[2] += a.z * s;
    acc[3] += a.w * s;
}
template <>
__device__ __forceinline__ void vfma<double2, double>(double (&acc)[4], const double2& a, double s) {
    acc[0] += a.x * s;
    acc[1] += a.y * s;
}
constexpr int kTriCU = 8;  // columns of the symmetric product in flight per lane

template <typename T>
__device__ T block_sum1(T v, T* sred) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sred[wave] = v;
    __syncthreads();
    T t = T(0);
#pragma unroll
    for (int w = 0; w < kTriNW; ++w) t += sred[w];
    return t;
}

inline size_t tridiag_smem_elems(int n) {
    const size_t uni = std::max<size_t>(2 * size_t(n), size_t(2) * 2 * kTriNB * 64);
    return uni + size_t(kTriNW) + 2 * kTriNB + 2 * kTriNB + kTriNT;
}

template <typename T>
__global__ void __launch_bounds__(kTriNT)
tridiag_kernel(int n, int ld, T* __restrict__ Aall, T* __restrict__ dall, T* __restrict__ eall,
               T* __restrict__ tauall, T* __restrict__ VWall) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* const uni = reinterpret_cast<T*>(smem_raw);
    const size_t unisz = max(2 * size_t(n), size_t(2) * 2 * kTriNB * 64);
    T* const sv = uni;          // Householder vector (n)
    T* const sp = uni + n;      // column / product (n)
    T* const sL = uni;          // trailing update: row panel [2NB][64]
    T* const sR = uni + 2 * kTriNB * 64;  //                 col panel [2NB][64]
    T* const sred = uni + unisz;           // kTriNW
    T* const srow = sred + kTriNW;          // 2*kTriNB: row j of W then V
    T* const sdot = srow + 2 * kTriNB;      // 2*kTriNB: W^T v then V^T v
    T* const spart = sdot + 2 * kTriNB;     // kTriNT

    const int b = blockIdx.x;
    T* const A = Aall + size_t(b) * n * ld;
    T* const d = dall + size_t(b) * n;
    T* const e = eall + size_t(b) * n;
    T* const tau = tauall + size_t(b) * n;
    T* const Vp = VWall + size_t(b) * 2 * kTriNB * ld;
    T* const Wp = Vp + size_t(kTriNB) * ld;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

    for (int j0 = 0; j0 < n - 1; j0 += kTriNB) {
        const int nbp = min(kTriNB, n - 1 - j0);
        for (int i = 0; i < nbp; ++i) {
            const int j = j0 + i;
            __syncthreads();
            if (tid < i) {
                srow[tid] = Wp[size_t(tid) * ld + j];
                srow[kTriNB + tid] = Vp[size_t(tid) * ld + j];
            }
            __syncthreads();
            // column j with the panel's pending updates, rows j..n-1
            T sig = T(0);
            for (int r = j + tid; r < n; r += kTriNT) {
                T a = A[size_t(j) * ld + r];
                for (int l = 0; l < i; ++l)
                    a -= Vp[size_t(l) * ld + r] * srow[l] + Wp[size_t(l) * ld + r] * srow[kTriNB + l];
                sp[r] = a;
                if (r >= j + 2) sig += a * a;
            }
            sig = block_sum1(sig, sred);
            const T alpha = sp[j + 1];
            const T dj = sp[j];
            T tj = T(0), beta = alpha, scal = T(0);
            if (sig != T(0)) {
                const T nrm = sqrt(alpha * alpha + sig);
                beta = (alpha >= T(0)) ? -nrm : nrm;
                tj = (beta - alpha) / beta;
                scal = T(1) / (alpha - beta);
            }
            for (int r = j + 1 + tid; r < n; r += kTriNT) {
                const T vr = (r == j + 1) ? T(1) : sp[r] * scal;
                sv[r] = vr;
                A[size_t(j) * ld + r] = vr;
                Vp[size_t(i) * ld + r] = vr;
            }
            for (int r = tid; r <= j; r += kTriNT) A[size_t(j) * ld + r] = T(0);
            if (tid == 0) {
                d[j] = dj;
                e[j] = beta;
                tau[j] = tj;
            }
            __syncthreads();
            const int r0 = j + 1, m = n - r0;
            if (tj == T(0)) {
                for (int r = r0 + tid; r < n; r += kTriNT) Wp[size_t(i) * ld + r] = T(0);
                continue;
            }
            // p = A_panel(r0:, r0:) v: each lane streams 16-byte row vectors of
            // kTriCU columns at a time (rows below r0 are read and discarded
            // so that every load is aligned; ld is a multiple of 4)
            if (m >= kTriNT) {
                using VT = typename Vec16<T>::type;
                constexpr int VW = Vec16<T>::n;
                const int rstart = r0 & ~(VW - 1);
                for (int rb = rstart; rb < n; rb += kTriNT * VW) {
                    const int r = rb + tid * VW;
                    const bool act = r < n;
                    T acc[4] = {T(0), T(0), T(0), T(0)};
                    int c = r0;
                    for (; c + kTriCU <= n; c += kTriCU) {
                        VT a[kTriCU];
#pragma unroll
                        for (int u = 0; u < kTriCU; ++u)
                            a[u] = act ? *reinterpret_cast<const VT*>(A + size_t(c + u) * ld + r) : VT{};
#pragma unroll
                        for (int u = 0; u < kTriCU; ++u) vfma(acc, a[u], sv[c + u]);
                    }
                    for (; c < n; ++c) {
                        const VT a = act ? *reinterpret_cast<const VT*>(A + size_t(c) * ld + r) : VT{};
                        vfma(acc, a, sv[c]);
                    }
#pragma unroll
                    for (int q = 0; q < VW; ++q)
                        if (act && r + q >= r0 && r + q < n) sp[r + q] = acc[q];
                }
            } else {
                const int rnd = (m + 63) & ~63;
                const int tcn = kTriNT / rnd;
                const int tr = tid % rnd, tc = tid / rnd;
                T acc = T(0);
                if (tc < tcn && tr < m)
                    for (int c = r0 + tc; c < n; c += tcn) acc += A[size_t(c) * ld + r0 + tr] * sv[c];
                spart[tid] = acc;
                __syncthreads();
                if (tid < m) {
                    T s = T(0);
                    for (int g = 0; g < tcn; ++g) s += spart[g * rnd + tid];
                    sp[r0 + tid] = s;
                }
            }
            __syncthreads();
            // x_l = W_l . v and y_l = V_l . v (l < i), one column per wave at a time
            for (int q = wave; q < 2 * i; q += kTriNW) {
                const bool isw = q < i;
                const int l = isw ? q : q - i;
                const T* col = (isw ? Wp : Vp) + size_t(l) * ld;
                T a0 = T(0), a1 = T(0), a2 = T(0), a3 = T(0);
                int r = r0 + lane;
                for (; r + 192 < n; r += 256) {
                    a0 += col[r] * sv[r];
                    a1 += col[r + 64] * sv[r + 64];
                    a2 += col[r + 128] * sv[r + 128];
                    a3 += col[r + 192] * sv[r + 192];
                }
                for (; r < n; r += 64) a0 += col[r] * sv[r];
                const T tot = wave_sum((a0 + a1) + (a2 + a3));
                if (lane == 0) sdot[isw ? l : kTriNB + l] = tot;
            }
            __syncthreads();
            T wv = T(0);
            for (int r = r0 + tid; r < n; r += kTriNT) {
                T pr = sp[r];
                for (int l = 0; l < i; ++l)
                    pr -= Vp[size_t(l) * ld + r] * sdot[l] + Wp[size_t(l) * ld + r] * sdot[kTriNB + l];
                pr *= tj;
                sp[r] = pr;
                wv += pr * sv[r];
            }
            wv = block_sum1(wv, sred);
            const T K = T(-0.5) * tj * wv;
            for (int r = r0 + tid; r < n; r += kTriNT) Wp[size_t(i) * ld + r] = sp[r] + K * sv[r];
        }
        // trailing update: A(jn:, jn:) -= V W^T + W V^T
        const int jn = j0 + nbp;
        __syncthreads();
        if (jn < n) {
            const int mt = n - jn;
            const int ntile = (mt + 63) / 64;
            const int x = tid & 63, yb = tid >> 6;
            for (int tr = 0; tr < ntile; ++tr) {
                const int rbase = jn + tr * 64;
                __syncthreads();
                for (int q = tid; q < 2 * kTriNB * 64; q += kTriNT) {
                    const int l = q >> 6, xx = q & 63, r = rbase + xx;
                    T val = T(0);
                    if (r < n) {
                        if (l < kTriNB) {
                            if (l < nbp) val = Vp[size_t(l) * ld + r];
                        } else if (l - kTriNB < nbp) {
                            val = Wp[size_t(l - kTriNB) * ld + r];
                        }
                    }
                    sL[q] = val;
                }
                for (int tc = 0; tc < ntile; ++tc) {
                    const int cbase = jn + tc * 64;
                    __syncthreads();
                    for (int q = tid; q < 2 * kTriNB * 64; q += kTriNT) {
                        const int l = q >> 6, yy = q & 63, c = cbase + yy;
                        T val = T(0);
                        if (c < n) {
                            if (l < kTriNB) {
                                if (l < nbp) val = Wp[size_t(l) * ld + c];
                            } else if (l - kTriNB < nbp) {
                                val = Vp[size_t(l - kTriNB) * ld + c];
                            }
                        }
                        sR[q] = val;
                    }
                    __syncthreads();
                    const int r = rbase + x;
                    if (r < n) {
                        T a[8];
#pragma unroll
                        for (int q = 0; q < 8; ++q) {
                            const int c = cbase + yb + 8 * q;
                            a[q] = (c < n) ? A[size_t(c) * ld + r] : T(0);
                        }
                        for (int l = 0; l < 2 * kTriNB; ++l) {
                            const T lv = sL[l * 64 + x];
#pragma unroll
                            for (int q = 0; q < 8; ++q) a[q] -= lv * sR[l * 64 + yb + 8 * q];
                        }
#pragma unroll
                        for (int q = 0; q < 8; ++q) {
                            const int c = cbase + yb + 8 * q;
                            if (c < n) A[size_t(c) * ld + r] = a[q];
                        }
                    }
                }
            }
        }
        __syncthreads();
    }
    if (tid == 0) {
        d[n - 1] = A[size_t(n - 1) * ld + (n - 1)];
        e[n - 1] = T(0);
        tau[n - 1] = T(0);
    }
    if (n >= 1) {
        __syncthreads();
        if (tid == 0) A[size_t(n - 1) * ld + (n - 1)] = T(0);
    }
}

// Number of eigenvalues of T below x (Sturm sequence count).
__device__ __forceinline__ int sturm_count(int n, const double* sd, const double* se2, double x,
                                           double pivmin) {
    double q = sd[0] - x;
    if (fabs(q) < pivmin) q = -pivmin;
    int c = q < 0.0;
    for (int k = 1; k < n; ++k) {
        q = (sd[k] - x) - se2[k - 1] / q;
        if (fabs(q) < pivmin) q = -pivmin;
        c += q < 0.0;
    }
    return c;
}

template <typename T>
__global__ void __launch_bounds__(kEigNT)
bisect_kernel(int n, const T* __restrict__ dall, const T* __restrict__ eall, double* __restrict__ wall) {
    extern __shared__ double shd[];
    double* sd = shd;
    double* se2 = shd + n;
    double* sr = shd + 2 * n;  // 3 * (kEigNT/64)
    const int b = blockIdx.y, tid = threadIdx.x;
    const T* d = dall + size_t(b) * n;
    const T* e = eall + size_t(b) * n;
    double lo = DBL_MAX, hi = -DBL_MAX, emax = 0.0;
    for (int k = tid; k < n; k += kEigNT) {
        const double dk = double(d[k]);
        const double ek = (k < n - 1) ? double(e[k]) : 0.0;
        const double ekm = (k > 0) ? double(e[k - 1]) : 0.0;
        sd[k] = dk;
        se2[k] = ek * ek;
        const double rad = fabs(ek) + fabs(ekm);
        lo = fmin(lo, dk - rad);
        hi = fmax(hi, dk + rad);
        emax = fmax(emax, ek * ek);
    }
    for (int off = 32; off > 0; off >>= 1) {
        lo = fmin(lo, __shfl_xor(lo, off, 64));
        hi = fmax(hi, __shfl_xor(hi, off, 64));
        emax = fmax(emax, __shfl_xor(emax, off, 64));
    }
    const int lane = tid & 63, wave = tid >> 6;
    if (lane == 0) {
        sr[wave] = lo;
        sr[4 + wave] = hi;
        sr[8 + wave] = emax;
    }
    __syncthreads();
    double gl = sr[0], gu = sr[4], e2 = sr[8];
    for (int w = 1; w < kEigNT / 64; ++w) {
        gl = fmin(gl, sr[w]);
        gu = fmax(gu, sr[4 + w]);
        e2 = fmax(e2, sr[8 + w]);
    }
    const double tnorm = fmax(fabs(gl), fabs(gu));
    const double pivmin = DBL_MIN * fmax(1.0, e2);
    const double widen = 4.0 * DBL_EPSILON * tnorm + 4.0 * pivmin;
    gl -= widen;
    gu += widen;
    const int j = blockIdx.x * kEigNT + tid;
    if (j >= n) return;
    double a = gl, c = gu;
    // relative tolerance (absolute floor 4 eps) on the bracket of eigenvalue j
    for (int it = 0; it < 128 && (c - a) > 2.0 * DBL_EPSILON * (fabs(a) + fabs(c)) + 4.0 * DBL_EPSILON; ++it) {
        const double mid = 0.5 * (a + c);
        if (sturm_count(n, sd, se2, mid, pivmin) > j)
            c = mid;
        else
            a = mid;
    }
    wall[size_t(b) * n + j] = 0.5 * (a + c);
}

__device__ __forceinline__ double hash_uniform(uint32_t a, uint32_t b) {
    uint64_t z = (uint64_t(a) << 32) ^ uint64_t(b) ^ 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return double(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

// Scratch of one matrix: S[(a*n + k)*n + c], a = 0..2 the three diagonals of
// U (LU of T - lambda_c I with partial pivoting), a = 3 the vector.
template <typename T>
__global__ void __launch_bounds__(kEigNT)
invit_kernel(int n, int ld, const T* __restrict__ dall, const T* __restrict__ eall,
             const double* __restrict__ wall, T* __restrict__ Zall, double* __restrict__ scr, int b0) {
    extern __shared__ double shd[];
    double* sd = shd;
    double* se = shd + n;
    const int bl = blockIdx.y, b = b0 + bl, tid = threadIdx.x;
    const T* d = dall + size_t(b) * n;
    const T* e = eall + size_t(b) * n;
    const double* w = wall + size_t(b) * n;
    T* Z = Zall + size_t(b) * n * ld;
    double* S = scr + size_t(bl) * 4 * n * n;
    for (int k = tid; k < n; k += kEigNT) {
        sd[k] = double(d[k]);
        se[k] = (k < n - 1) ? double(e[k]) : 0.0;
    }
    __syncthreads();
    const int c = blockIdx.x * kEigNT + tid;
    if (c >= n) return;
    const double tnorm = fmax(fmax(fabs(w[0]), fabs(w[n - 1])), DBL_MIN);
    const double pivmin = DBL_EPSILON * tnorm;
    // cluster start and this lane's position in it; every kMaxCluster-th
    // member leads a group of the cluster
    int cs = c;
    while (cs > 0 && joins_cluster(w, cs)) --cs;
    if ((c - cs) % kMaxCluster != 0) return;  // solved by its group's first lane
    int cend = c;
    while (cend + 1 < n && cend + 1 - c < kMaxCluster && joins_cluster(w, cend + 1)) ++cend;

    const size_t nn = size_t(n) * n;
    double lam_prev = 0.0;
    for (int mb = c; mb <= cend; ++mb) {
        double lam = w[mb];
        if (mb > c) lam = fmax(lam, lam_prev + 10.0 * DBL_EPSILON * fmax(fabs(lam_prev), 1.0));
        lam_prev = lam;
        double* U0 = S + mb;
        double* U1 = S + nn + mb;
        double* U2 = S + 2 * nn + mb;
        double* X = S + 3 * nn + mb;
        for (int k = 0; k < n; ++k) X[size_t(k) * n] = hash_uniform(uint32_t(mb), uint32_t(k));
        double scale = 1.0;
        for (int it = 0; it < 2; ++it) {
            // forward: factor T - lam I = P L U and apply P L^{-1} to scale*X
            double a = sd[0] - lam;
            double bs = (n > 1) ? se[0] : 0.0;
            double yi = X[0] * scale;
            for (int i = 0; i < n - 1; ++i) {
                const double sub = se[i];
                const double dn = sd[i + 1] - lam;
                const double en = (i + 1 < n - 1) ? se[i + 1] : 0.0;
                const double yn = X[size_t(i + 1) * n] * scale;
                double u0, u1, u2, yo;
                if (fabs(a) >= fabs(sub)) {
                    if (a == 0.0) a = pivmin;
                    const double l = sub / a;
                    u0 = a;
                    u1 = bs;
                    u2 = 0.0;
                    yo = yi;
                    a = dn - l * bs;
                    bs = en;
                    yi = yn - l * yi;
                } else {
                    const double l = a / sub;
                    u0 = sub;
                    u1 = dn;
                    u2 = en;
                    yo = yn;
                    a = bs - l * dn;
                    bs = -l * en;
                    yi = yi - l * yn;
                }
                if (fabs(u0) < pivmin) u0 = (u0 < 0.0) ? -pivmin : pivmin;
                U0[size_t(i) * n] = u0;
                U1[size_t(i) * n] = u1;
                U2[size_t(i) * n] = u2;
                X[size_t(i) * n] = yo;
            }
            if (fabs(a) < pivmin) a = (a < 0.0) ? -pivmin : pivmin;
            // backward: U x = y
            double x1 = yi / a, x2 = 0.0;
            X[size_t(n - 1) * n] = x1;
            double nrm2 = x1 * x1;
            for (int i = n - 2; i >= 0; --i) {
                const double xi = (X[size_t(i) * n] - U1[size_t(i) * n] * x1 - U2[size_t(i) * n] * x2) /
                                  U0[size_t(i) * n];
                X[size_t(i) * n] = xi;
                nrm2 += xi * xi;
                x2 = x1;
                x1 = xi;
            }
            // orthogonalise against the cluster's earlier vectors
            if (mb > c) {
                for (int q = c; q < mb; ++q) {
                    const double* Xq = S + 3 * nn + q;
                    double dot = 0.0;
                    for (int k = 0; k < n; ++k) dot += X[size_t(k) * n] * Xq[size_t(k) * n];
                    nrm2 = 0.0;
                    for (int k = 0; k < n; ++k) {
                        const double v = X[size_t(k) * n] - dot * Xq[size_t(k) * n];
                        X[size_t(k) * n] = v;
                        nrm2 += v * v;
                    }
                }
            }
            scale = 1.0 / sqrt(nrm2);
        }
        // normalise in place, largest component positive
        double big = 0.0;
        for (int k = 0; k < n; ++k) {
            const double v = X[size_t(k) * n] * scale;
            X[size_t(k) * n] = v;
            if (fabs(v) > fabs(big)) big = v;
        }
        if (big < 0.0)
            for (int k = 0; k < n; ++k) X[size_t(k) * n] = -X[size_t(k) * n];
    }
    for (int mb = c; mb <= cend; ++mb) {
        const double* X = S + 3 * nn + mb;
        for (int k = 0; k < n; ++k) Z[size_t(k) * ld + mb] = T(X[size_t(k) * n]);
    }
}

// T factor of a compact-WY block from its Gram matrix G = Y^T Y:
// T(l,l) = tau_l, T(0:l, l) = -tau_l T(0:l, 0:l) G(0:l, l).
template <typename T>
__global__ void __launch_bounds__(128)
larft_kernel(int n, int kb, int j0, const T* __restrict__ tauall, const T* __restrict__ Gall,
             T* __restrict__ Tall) {
    __shared__ T col[kBtNB];
    const int b = blockIdx.x, tid = threadIdx.x;
    const T* tau = tauall + size_t(b) * n + j0;
    const T* G = Gall + size_t(b) * kBtNB * kBtNB;
    T* Tm = Tall + size_t(b) * kBtNB * kBtNB;
    for (int q = tid; q < kBtNB * kBtNB; q += 128) Tm[q] = T(0);
    __syncthreads();
    for (int l = 0; l < kb; ++l) {
        const T tl = tau[l];
        if (tid < l) col[tid] = -tl * G[tid * kBtNB + l];
        __syncthreads();
        if (tid < l) {
            T acc = T(0);
            for (int q = tid; q < l; ++q) acc += Tm[tid * kBtNB + q] * col[q];
            Tm[tid * kBtNB + l] = acc;
        }
        if (tid == 0) Tm[l * kBtNB + l] = tl;
        __syncthreads();
    }
}

struct EighWork {
    size_t off_d, off_e, off_tau, off_vw, off_g, off_t, off_w1, off_w2, off_scr, total;
    int sub;  // matrices per inverse-iteration pass
};

inline EighWork eigh_layout(size_t esz, int batch, int n, int ld) {
    EighWork L{};
    auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
    size_t o = 0;
    L.off_d = o;
    o = al(o + esz * batch * n);
    L.off_e = o;
    o = al(o + esz * batch * n);
    L.off_tau = o;
    o = al(o + esz * batch * n);
    L.off_vw = o;
    o = al(o + esz * batch * 2 * kTriNB * size_t(ld));
    L.off_g = o;
    o = al(o + esz * batch * kBtNB * kBtNB);
    L.off_t = o;
    o = al(o + esz * batch * kBtNB * kBtNB);
    L.off_w1 = o;
    o = al(o + esz * batch * kBtNB * size_t(n));
    L.off_w2 = o;
    o = al(o + esz * batch * kBtNB * size_t(n));
    const size_t per = 32 * size_t(n) * n;  // 4 n^2 doubles
    const size_t budget = size_t(8) << 30;
    int sub = int(std::max<size_t>(1, budget / per));
    L.sub = std::min(sub, batch);
    L.off_scr = o;
    o = al(o + per * L.sub);
    L.total = o;
    return L;
}

template <typename T>
int eigh_impl(int gemm_prec, int batch, int n, int ld, T* A, T* V, double* w, unsigned char* work,
              hipStream_t s, Prof* prof) {
    const EighWork L = eigh_layout(sizeof(T), batch, n, ld);
    T* dd = reinterpret_cast<T*>(work + L.off_d);
    T* ee = reinterpret_cast<T*>(work + L.off_e);
    T* tau = reinterpret_cast<T*>(work + L.off_tau);
    T* vw = reinterpret_cast<T*>(work + L.off_vw);
    T* G = reinterpret_cast<T*>(work + L.off_g);
    T* Tm = reinterpret_cast<T*>(work + L.off_t);
    T* W1 = reinterpret_cast<T*>(work + L.off_w1);
    T* W2 = reinterpret_cast<T*>(work + L.off_w2);
    double* scr = reinterpret_cast<double*>(work + L.off_scr);

    // 1. tridiagonalisation
    if (ld % 4 != 0) return fail(CIR_ERR_ARG, "eigh: the leading dimension must be a multiple of 4");
    const size_t sm = tridiag_smem_elems(n) * sizeof(T);
    if (sm > 160 * 1024) return fail(CIR_ERR_ARG, "eigh: n too large for the LDS of one workgroup");
    CIR_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&tridiag_kernel<T>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, int(sm)));
    {
        ProfScope ps(prof, CIR_PH_TRIDIAG, s);
        hipLaunchKernelGGL(tridiag_kernel<T>, dim3(batch), dim3(kTriNT), sm, s, n, ld, A, dd, ee, tau, vw);
        CIR_CHECK_LAUNCH();
    }
    // 2. eigenvalues
    const size_t smb = (2 * size_t(n) + 16) * sizeof(double);
    if (smb > 160 * 1024) return fail(CIR_ERR_ARG, "eigh: n too large for bisection LDS");
    CIR_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&bisect_kernel<T>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, int(smb)));
    {
        ProfScope ps(prof, CIR_PH_BISECT, s);
        hipLaunchKernelGGL(bisect_kernel<T>, dim3(unsigned(ceil_div(n, kEigNT)), batch), dim3(kEigNT), smb, s,
                           n, dd, ee, w);
        CIR_CHECK_LAUNCH();
    }
    // 3. eigenvectors of T into V (row-major)
    CIR_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&invit_kernel<T>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, int(smb)));
    ProfScope* pinv = new ProfScope(prof, CIR_PH_INVIT, s);
    for (int b0 = 0; b0 < batch; b0 += L.sub) {
        const int nb = std::min(L.sub, batch - b0);
        hipLaunchKernelGGL(invit_kernel<T>, dim3(unsigned(ceil_div(n, kEigNT)), nb), dim3(kEigNT), smb, s, n,
                           ld, dd, ee, w, V, scr, b0);
        CIR_CHECK_LAUNCH();
    }
    delete pinv;
    // 4. back-transform V <- Q V, last block of reflectors first
    ProfScope pbt(prof, CIR_PH_BACKTRANSFORM, s);
    const int nref = n - 1;
    const int nblk = int(ceil_div(nref, kBtNB));
    for (int blk = nblk - 1; blk >= 0; --blk) {
        const int j0 = blk * kBtNB;
        const int kb = std::min(kBtNB, nref - j0);
        const int mrows = n - j0 - 1;
        if (mrows <= 0 || kb <= 0) continue;
        const T* Y = A + size_t(j0) * ld + (j0 + 1);  // Y(r, l) = Y[l*ld + r]
        T* Zs = V + size_t(j0 + 1) * ld;                // rows j0+1.. of V
        GemmArgs g;
        // G = Y^T Y
        g.M = kb;
        g.N = kb;
        g.K = mrows;
        g.batch = batch;
        g.A = Y;
        g.a_m = ld;
        g.a_k = 1;
        g.a_b = int64_t(n) * ld;
        g.B = Y;
        g.b_k = 1;
        g.b_n = ld;
        g.b_b = int64_t(n) * ld;
        g.C = G;
        g.c_m = kBtNB;
        g.c_n = 1;
        g.c_b = kBtNB * kBtNB;
        CIR_TRY(gemm(gemm_prec, g, s));
        hipLaunchKernelGGL(larft_kernel<T>, dim3(batch), dim3(128), 0, s, n, kb, j0, tau, G, Tm);
        CIR_CHECK_LAUNCH();
        // W1 = Y^T Z
        g = GemmArgs{};
        g.M = kb;
        g.N = n;
        g.K = mrows;
        g.batch = batch;
        g.A = Y;
        g.a_m = ld;
        g.a_k = 1;
        g.a_b = int64_t(n) * ld;
        g.B = Zs;
        g.b_k = ld;
        g.b_n = 1;
        g.b_b = int64_t(n) * ld;
        g.C = W1;
        g.c_m = n;
        g.c_n = 1;
        g.c_b = int64_t(kBtNB) * n;
        CIR_TRY(gemm(gemm_prec, g, s));
        // W2 = T W1
        g = GemmArgs{};
        g.M = kb;
        g.N = n;
        g.K = kb;
        g.batch = batch;
        g.A = Tm;
        g.a_m = kBtNB;
        g.a_k = 1;
        g.a_b = kBtNB * kBtNB;
        g.B = W1;
        g.b_k = n;
        g.b_n = 1;
        g.b_b = int64_t(kBtNB) * n;
        g.C = W2;
        g.c_m = n;
        g.c_n = 1;
        g.c_b = int64_t(kBtNB) * n;
        CIR_TRY(gemm(gemm_prec, g, s));
        // Z -= Y W2
        g = GemmArgs{};
        g.M = mrows;
        g.N = n;
        g.K = kb;
        g.batch = batch;
        g.A = Y;
        g.a_m = 1;
        g.a_k = ld;
        g.a_b = int64_t(n) * ld;
        g.B = W2;
        g.b_k = n;
        g.b_n = 1;
        g.b_b = int64_t(kBtNB) * n;
        g.C = Zs;
        g.c_m = ld;
        g.c_n = 1;
        g.c_b = int64_t(n) * ld;
        g.alpha = -1.0;
        g.beta = 1.0;
        CIR_TRY(gemm(gemm_prec, g, s));
    }
    return CIR_OK;
}

size_t eigh_workspace(int prec, int batch, int n, int ld) {
    const size_t esz = (prec == CIR_PREC_FP64) ? sizeof(double) : sizeof(float);
    return eigh_layout(esz, batch, n, ld).total;
}

int eigh_batched(int prec, int batch, int n, int ld, void* A, void* V, double* w, void* work,
                 hipStream_t s, Prof* prof) {
    if (batch <= 0 || n <= 0 || ld < n || !A || !V || !w || !work)
        return fail(CIR_ERR_ARG, "cir_eigh_batched: bad sizes or null pointer");
    if (batch > 65535) return fail(CIR_ERR_ARG, "cir_eigh_batched: batch > 65535");
    unsigned char* wk = static_cast<unsigned char*>(work);
    if (prec == CIR_PREC_FP64)
        return eigh_impl<double>(CIR_PREC_FP64, batch, n, ld, static_cast<double*>(A), static_cast<double*>(V), w,
                                 wk, s, prof);
    return eigh_impl<float>(CIR_PREC_FP32, batch, n, ld, static_cast<float*>(A), static_cast<float*>(V), w, wk,
                            s, prof);
}

}  // namespace cir

extern "C" size_t cir_eigh_workspace(int prec, int batch, int n, int ld) {
    return cir::eigh_workspace(prec, batch, n, ld);
}
