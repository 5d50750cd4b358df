// k_band.hip -- two-stage reduction of the batched FP32 eigensolver (gfx950),
// replacing the one-stage Householder tridiagonalisation of magma_ssyevd_gpu
// (calcIR.cu:420-448) for batches of mid-size Hamiltonians.
//
//   Stage 1, dense -> band of half-width kBandNB (LAPACK sytrd_sy2sb form):
//     for each panel of kBandNB columns, sb_panel_kernel factors the panel
//     below the band, A(r0:n, j0:j0+NB) = Q R, by Householder reflectors in
//     registers (one workgroup per matrix); the trailing matrix then gets the
//     two-sided update A22 <- Q^T A22 Q as three batched MFMA products over
//     every matrix at once (X = A22 Yh with Yh = Y T, W = X - Y (Yh^T X)/2,
//     A22 -= [W Y] [Y W]^T).  A is kept symmetric in both triangles.
//   Stage 2, band -> tridiagonal (sb2st_kernel): bulge chasing, one workgroup
//     per matrix, sweeps in groups of kSbG run by the group's waves in
//     lock-step with a lag of two tasks (the tasks of one step touch disjoint
//     elements); the band rows the group touches live in a circular LDS
//     window that slides down the matrix.
//   Q2 back-transform (q2_*): the bulge-chasing reflectors of kQ2G
//     consecutive sweeps at one block position form a compact-WY block;
//     blocks are applied to column slabs of the eigenvector matrix held in
//     LDS, groups last to first, block positions in ascending order.
//
// Storage (per matrix):
//   BND   band rows: element (r, c), 0 <= r - c < 2 NB, at BND[r * 2NB + r - c]
//         (the bulge region r - c > NB is zero outside the chase)
//   A     column j0 + l of panel j0 holds reflector l below row r0 = j0 + NB:
//         rows r0..r0+l-1 zero, row r0+l one, rows below the reflector, the
//         layout the compact-WY back-transform reads (offset NB)

#include <cfloat>

#include "band.hpp"
#include "gemm.hpp"

namespace cir {

constexpr int kPanNT = 512;  // threads per matrix in the panel factorisation
constexpr int kPanNW = kPanNT / 64;

// DPP row rotation by CTRL (row_ror 8, 4, 2, 1) within 16-lane rows.
template <int CTRL>
__device__ __forceinline__ float bdpp(float x) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, 0xf, 0xf, false));
}
__device__ __forceinline__ float brow_sum16(float x) {
    x += bdpp<0x128>(x);
    x += bdpp<0x124>(x);
    x += bdpp<0x122>(x);
    x += bdpp<0x121>(x);
    return x;
}
__device__ __forceinline__ float bwave_sum(float x) {
    x = brow_sum16(x);
    // rows of 16 lanes: add the other row of the half, then the other half
    auto s16 = __builtin_amdgcn_permlane16_swap(__float_as_uint(x), __float_as_uint(x), false, false);
    x = __uint_as_float(s16[0]) + __uint_as_float(s16[1]);
    auto s32 = __builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(x), false, false);
    return __uint_as_float(s32[0]) + __uint_as_float(s32[1]);
}

// Wave sums of 32 per-lane values x[k], written to out[k] (LDS).  Two lane
// swaps halve the values per lane (after them lane l holds the partials of
// columns k + 8 (l >> 4), k < 8), 16-lane row sums finish them, and lane
// 16 R + i (i < 8) stores column 8 R + i.
__device__ __forceinline__ void wave_colsums32(const float (&x)[32], int lane, float* out) {
    float y[16], z[8];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const auto sw = __builtin_amdgcn_permlane32_swap(__float_as_uint(x[k]), __float_as_uint(x[k + 16]), false,
                                                         false);
        y[k] = __uint_as_float(sw[0]) + __uint_as_float(sw[1]);
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const auto sw = __builtin_amdgcn_permlane16_swap(__float_as_uint(y[k]), __float_as_uint(y[k + 8]), false,
                                                         false);
        z[k] = brow_sum16(__uint_as_float(sw[0]) + __uint_as_float(sw[1]));
    }
    const int i = lane & 15;
    float v = z[0];
#pragma unroll
    for (int k = 1; k < 8; ++k) v = (i == k) ? z[k] : v;
    if (i < 8) out[8 * (lane >> 4) + i] = v;
}

// Householder QR of one panel below the band and everything the trailing
// update needs.  Panel rows r0..n-1 (m = n - r0 of them) are spread over the
// threads, rows tid + 512 i (i < RPT) per thread, all NB columns in
// registers.  Per column c: one block sum for the norm, one for the 32
// products v^T P(:, c') (the trailing columns' w and, for c' < c, the Gram
// column Y^T v of the T factor).  Outputs: the diagonal block and R into BND,
// tau, Y into A (back-transform layout) and into P1 (columns NB..2NB) and P2
// (columns 0..NB), Yh = Y T into Yh; pitch pm.
template <int RPT>
__global__ void __launch_bounds__(kPanNT) sb_panel_kernel(int n, int ld, int j0, float* __restrict__ Aall,
                                                          float* __restrict__ BNDall, float* __restrict__ tauall,
                                                          float* __restrict__ P1all, float* __restrict__ P2all,
                                                          float* __restrict__ Yhall, int pm) {
    constexpr int NB = kBandNB;
    __shared__ float redA[kPanNW], redC[kPanNW][NB], dotv[NB], taus[NB], salpha;
    __shared__ float gram[NB][NB + 1], Tm[NB][NB + 1];
    const int b = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    float* A = Aall + size_t(b) * n * ld;
    float* BND = BNDall + size_t(b) * n * kBandW;
    float* P1 = P1all + size_t(b) * pm * 2 * NB;
    float* P2 = P2all + size_t(b) * pm * 2 * NB;
    float* Yh = Yhall + size_t(b) * pm * NB;
    const int r0 = j0 + NB, m = n - r0;

    // the panel's diagonal block is final: lower triangle into the band rows
    for (int q = tid; q < NB * NB; q += kPanNT) {
        const int i = q / NB, c = q % NB;
        if (c <= i && j0 + i < n) BND[size_t(j0 + i) * kBandW + (i - c)] = A[size_t(j0 + c) * ld + j0 + i];
    }
    float p[RPT][NB];
#pragma unroll
    for (int i = 0; i < RPT; ++i) {
        const int row = tid + kPanNT * i;
#pragma unroll
        for (int c = 0; c < NB; ++c) p[i][c] = (row < m) ? A[size_t(j0 + c) * ld + r0 + row] : 0.f;
    }

#pragma unroll 1
    for (int c = 0; c < NB; ++c) {
        // norm of the column below row c, and its head
        float s = 0.f;
#pragma unroll
        for (int i = 0; i < RPT; ++i) {
            const int row = tid + kPanNT * i;
            float x = 0.f;
#pragma unroll
            for (int q = 0; q < NB; ++q) x = (q == c) ? p[i][q] : x;
            if (row > c) s += x * x;
        }
        s = bwave_sum(s);
        if (lane == 0) redA[wave] = s;
        if (tid == c) {
            float x = 0.f;
#pragma unroll
            for (int q = 0; q < NB; ++q) x = (q == c) ? p[0][q] : x;
            salpha = x;
        }
        __syncthreads();
        float sigma = 0.f;
#pragma unroll
        for (int w = 0; w < kPanNW; ++w) sigma += redA[w];
        const float alpha = salpha;
        float tau = 0.f, beta = alpha, scale = 0.f;
        if (sigma > 0.f) {
            const float nrm = sqrtf(alpha * alpha + sigma);
            beta = alpha >= 0.f ? -nrm : nrm;
            tau = (beta - alpha) / beta;
            scale = 1.f / (alpha - beta);
        }
        // v below the unit; row c keeps beta (the R diagonal) and counts as 1
        float v[RPT];
#pragma unroll
        for (int i = 0; i < RPT; ++i) {
            const int row = tid + kPanNT * i;
#pragma unroll
            for (int q = 0; q < NB; ++q)
                if (q == c) {
                    if (row > c) p[i][q] *= scale;
                    if (row == c) p[i][q] = beta;
                    v[i] = (row > c) ? p[i][q] : (row == c ? 1.f : 0.f);
                }
        }
        // v^T P(:, c') over rows >= c, every c'
        float dots[NB];
#pragma unroll
        for (int q = 0; q < NB; ++q) dots[q] = 0.f;
#pragma unroll
        for (int i = 0; i < RPT; ++i)
#pragma unroll
            for (int q = 0; q < NB; ++q) dots[q] += v[i] * p[i][q];
        wave_colsums32(dots, lane, redC[wave]);
        __syncthreads();
        if (tid < NB) {
            float t = 0.f;
#pragma unroll
            for (int w = 0; w < kPanNW; ++w) t += redC[w][tid];
            dotv[tid] = t;
            if (tid < c) gram[tid][c] = t;  // (Y^T v_c) of the earlier reflectors
            if (tid == c) taus[c] = tau;
        }
        __syncthreads();
        // trailing panel columns: P(:, c') -= tau v (v^T P(:, c'))
#pragma unroll
        for (int q = 0; q < NB; ++q) {
            if (q <= c) continue;
            const float f = tau * dotv[q];
#pragma unroll
            for (int i = 0; i < RPT; ++i) p[i][q] -= f * v[i];
        }
    }

    // T factor (forward, columnwise): T(c, c) = tau_c, T(0:c, c) = -tau_c T(0:c, 0:c) Y^T v_c
    for (int c = 0; c < NB; ++c) {
        if (tid < NB) {
            const int t = tid;
            float acc = 0.f;
            if (t < c)
                for (int q = t; q < c; ++q) acc += Tm[t][q] * gram[q][c];
            if (t < c) Tm[t][c] = -taus[c] * acc;
            if (t == c) Tm[c][c] = taus[c];
            if (t > c) Tm[t][c] = 0.f;
        }
        __syncthreads();
    }
    if (tid < NB) tauall[size_t(b) * n + j0 + tid] = taus[tid];
    // R (upper triangle of rows r0..r0+NB-1) into the band rows
    if (tid < NB && tid < m) {
#pragma unroll
        for (int c = 0; c < NB; ++c)
            if (c >= tid) BND[size_t(r0 + tid) * kBandW + (NB + tid - c)] = p[0][c];
    }
    // column j0 + c above the reflector's unit is zero from row 0 (the
    // back-transform reads whole passes of columns from their first row, and
    // the band part of these columns is in BND now)
    for (int c = 0; c < NB; ++c)
        for (int r = tid; r < r0 + c && r < n; r += kPanNT) A[size_t(j0 + c) * ld + r] = 0.f;
    // Y into A, P1, P2 and Yh = Y T
#pragma unroll
    for (int i = 0; i < RPT; ++i) {
        const int row = tid + kPanNT * i;
        if (row >= m) continue;
        float y[NB];
#pragma unroll
        for (int c = 0; c < NB; ++c) y[c] = (row < c) ? 0.f : (row == c ? 1.f : p[i][c]);
#pragma unroll
        for (int c = 0; c < NB; ++c) {
            A[size_t(j0 + c) * ld + r0 + row] = y[c];
            P1[size_t(NB + c) * pm + row] = y[c];
            P2[size_t(c) * pm + row] = y[c];
        }
#pragma unroll
        for (int c = 0; c < NB; ++c) {
            float acc = 0.f;
#pragma unroll
            for (int q = 0; q <= c; ++q) acc += y[q] * Tm[q][c];
            Yh[size_t(c) * pm + row] = acc;
        }
    }
}

// W (P1 columns 0..NB) copied into P2 columns NB..2NB, so that the rank-2NB
// update reads [W Y] and [Y W] as plain column-major operands.
__global__ void __launch_bounds__(256) sb_copyw_kernel(int m, int pm, float* __restrict__ P1all,
                                                       float* __restrict__ P2all) {
    const int b = blockIdx.y;
    const int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (e >= int64_t(m) * kBandNB) return;
    const int c = int(e / m), r = int(e % m);
    const float* P1 = P1all + size_t(b) * pm * 2 * kBandNB;
    float* P2 = P2all + size_t(b) * pm * 2 * kBandNB;
    P2[size_t(kBandNB + c) * pm + r] = P1[size_t(c) * pm + r];
}

// Zero the band rows, then the last (unfactored) diagonal block's lower
// triangle into them: columns jl..n-1, rows jl..n-1 (at most NB + 1 of them).
__global__ void __launch_bounds__(256) sb_band_zero_kernel(int n, float* __restrict__ BNDall) {
    const int b = blockIdx.y;
    const int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (e >= int64_t(n) * kBandW) return;
    BNDall[size_t(b) * n * kBandW + e] = 0.f;
}
__global__ void __launch_bounds__(256) sb_band_tail_kernel(int n, int ld, int jl, const float* __restrict__ Aall,
                                                           float* __restrict__ BNDall) {
    const int b = blockIdx.x;
    const float* A = Aall + size_t(b) * n * ld;
    float* BND = BNDall + size_t(b) * n * kBandW;
    const int k = n - jl;
    for (int q = threadIdx.x; q < k * k; q += 256) {
        const int i = q / k, c = q % k;
        if (c <= i) BND[size_t(jl + i) * kBandW + (i - c)] = A[size_t(jl + c) * ld + jl + i];
    }
}

int sy2sb(int batch, int n, int ld, float* A, float* BND, float* tau, float* P1, float* P2, float* Yh, float* Mm,
          int pm, hipStream_t s) {
    constexpr int NB = kBandNB;
    if (n > sb_max_n()) return fail(CIR_ERR_ARG, "sy2sb: n too large for the register panel");
    if (pm < n) return fail(CIR_ERR_ARG, "sy2sb: panel pitch below n");
    {
        dim3 grid(unsigned(ceil_div(int64_t(n) * kBandW, 256)), unsigned(batch));
        hipLaunchKernelGGL(sb_band_zero_kernel, grid, dim3(256), 0, s, n, BND);
        CIR_CHECK_LAUNCH();
    }
    int j0 = 0;
    for (; j0 + NB < n - 1; j0 += NB) {
        const int r0 = j0 + NB, m = n - r0;
        const int rpt = int(ceil_div(m, kPanNT));
        if (rpt <= 1)
            hipLaunchKernelGGL(sb_panel_kernel<1>, dim3(unsigned(batch)), dim3(kPanNT), 0, s, n, ld, j0, A, BND, tau,
                               P1, P2, Yh, pm);
        else if (rpt <= 2)
            hipLaunchKernelGGL(sb_panel_kernel<2>, dim3(unsigned(batch)), dim3(kPanNT), 0, s, n, ld, j0, A, BND, tau,
                               P1, P2, Yh, pm);
        else if (rpt <= 3)
            hipLaunchKernelGGL(sb_panel_kernel<3>, dim3(unsigned(batch)), dim3(kPanNT), 0, s, n, ld, j0, A, BND, tau,
                               P1, P2, Yh, pm);
        else
            hipLaunchKernelGGL(sb_panel_kernel<4>, dim3(unsigned(batch)), dim3(kPanNT), 0, s, n, ld, j0, A, BND, tau,
                               P1, P2, Yh, pm);
        CIR_CHECK_LAUNCH();
        const int64_t mat = int64_t(n) * ld, pstride = int64_t(pm) * 2 * NB;
        float* A22 = A + size_t(r0) * ld + r0;
        GemmArgs g;
        g.batch = batch;
        // X = A22 Yh into P1 columns 0..NB
        g.M = m;
        g.N = NB;
        g.K = m;
        g.A = A22;
        g.a_m = 1;
        g.a_k = ld;
        g.a_b = mat;
        g.B = Yh;
        g.b_k = 1;
        g.b_n = pm;
        g.b_b = int64_t(pm) * NB;
        g.C = P1;
        g.c_m = 1;
        g.c_n = pm;
        g.c_b = pstride;
        CIR_TRY(gemm(CIR_PREC_FP32, g, s));
        // Mm = Yh^T X (NB x NB)
        g = GemmArgs{};
        g.batch = batch;
        g.M = NB;
        g.N = NB;
        g.K = m;
        g.A = Yh;
        g.a_m = pm;
        g.a_k = 1;
        g.a_b = int64_t(pm) * NB;
        g.B = P1;
        g.b_k = 1;
        g.b_n = pm;
        g.b_b = pstride;
        g.C = Mm;
        g.c_m = NB;
        g.c_n = 1;
        g.c_b = int64_t(NB) * NB;
        CIR_TRY(gemm(CIR_PREC_FP32, g, s));
        // W = X - Y Mm / 2 in place of X (Y: P2 columns 0..NB)
        g = GemmArgs{};
        g.batch = batch;
        g.M = m;
        g.N = NB;
        g.K = NB;
        g.A = P2;
        g.a_m = 1;
        g.a_k = pm;
        g.a_b = pstride;
        g.B = Mm;
        g.b_k = NB;
        g.b_n = 1;
        g.b_b = int64_t(NB) * NB;
        g.C = P1;
        g.c_m = 1;
        g.c_n = pm;
        g.c_b = pstride;
        g.alpha = -0.5;
        g.beta = 1.0;
        CIR_TRY(gemm(CIR_PREC_FP32, g, s));
        {
            dim3 grid(unsigned(ceil_div(int64_t(m) * NB, 256)), unsigned(batch));
            hipLaunchKernelGGL(sb_copyw_kernel, grid, dim3(256), 0, s, m, pm, P1, P2);
            CIR_CHECK_LAUNCH();
        }
        // A22 -= [W Y] [Y W]^T (both triangles)
        g = GemmArgs{};
        g.batch = batch;
        g.M = m;
        g.N = m;
        g.K = 2 * NB;
        g.A = P1;
        g.a_m = 1;
        g.a_k = pm;
        g.a_b = pstride;
        g.B = P2;
        g.b_k = pm;
        g.b_n = 1;
        g.b_b = pstride;
        g.C = A22;
        g.c_m = 1;
        g.c_n = ld;
        g.c_b = mat;
        g.alpha = -1.0;
        g.beta = 1.0;
        CIR_TRY(gemm(CIR_PREC_FP32, g, s));
    }
    // the last diagonal block (n - j0 <= NB + 1 rows) is already banded
    hipLaunchKernelGGL(sb_band_tail_kernel, dim3(unsigned(batch)), dim3(256), 0, s, n, ld, j0, A, BND);
    CIR_CHECK_LAUNCH();
    return CIR_OK;
}

}  // namespace cir

namespace cir {

// ---------------------------------------------------------------------------
// Stage 2: band -> tridiagonal by bulge chasing.
// ---------------------------------------------------------------------------
constexpr int kSbG = 8;                  // sweeps per group = waves per workgroup
constexpr int kSbNT = 64 * kSbG;
constexpr int kSbWR = 512;               // window rows (power of two)
constexpr int kSbWP = kBandW + 1;        // window row pitch (floats)
constexpr size_t kSbSmemFloats = size_t(kSbWR) * kSbWP + size_t(kSbG) * 2 * kBandNB;
static_assert(kBandNB * (kBandW / 4) <= kSbNT, "one step's new band rows must fit one 16-byte piece per thread");
static_assert((2 * kSbG - 1) * kBandNB - kSbG + kBandNB < kSbWR, "window must hold a step's rows plus the read-ahead");

// Make this wave's LDS stores visible to its own later LDS loads (a wave's LDS
// operations complete in order; this keeps the compiler from moving them).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

struct SbWin {
    float* w;
    __device__ float& at(int r, int c) const { return w[(r & (kSbWR - 1)) * kSbWP + (r - c)]; }
};

// First row a sweep still needs at its task k (k < 0: not started, so its
// first task's row), or INT_MAX once it is done.
__device__ __forceinline__ int sb_task_row(int n, int s, int k) {
    return (k < sb_tasks(n, s)) ? s + 1 + max(k, 0) * kBandNB : 0x7fffffff;
}

// One wave's share of a bulge-chasing step: task (s, k) of its sweep.
// Lane l: row i = l & 31 of the task's block, column half h = l >> 5
// (columns 16 h .. 16 h + 15).  pv / ptau: the sweep's previous reflector
// (columns of the previous block, this lane's half), replaced by this task's.
__device__ void sb_task(int n, int s, int k, const SbWin& W, float* vs, float* ws, float (&pv)[16], float& ptau,
                        float* q2v, float* q2t, int lane) {
    constexpr int NB = kBandNB;
    const int i = lane & 31, h = lane >> 5;
    const int r0 = s + 1 + k * NB;
    const int len = min(NB, n - r0);
    float bk[16];
    float x;
    if (k == 0) {
        x = (i < len) ? W.at(r0 + i, s) : 0.f;
    } else {
        // the block below the previous one, columns p0 .. p0 + NB - 1: apply
        // the previous reflector from the right (this fills the bulge)
        const int p0 = r0 - NB;
#pragma unroll
        for (int c = 0; c < 16; ++c) bk[c] = (i < len) ? W.at(r0 + i, p0 + 16 * h + c) : 0.f;
        float u = 0.f;
#pragma unroll
        for (int c = 0; c < 16; ++c) u += bk[c] * pv[c];
        u += __shfl_xor(u, 32, 64);
        const float f = ptau * u;
#pragma unroll
        for (int c = 0; c < 16; ++c) bk[c] -= f * pv[c];
        x = __shfl(bk[0], i, 64);  // the block's first column, from the h = 0 lanes
    }
    // reflector annihilating x[1..len)
    const float sq = (i >= 1 && i < len) ? x * x : 0.f;
    const float sigma = bwave_sum(h == 0 ? sq : 0.f);
    const float alpha = __shfl(x, 0, 64);
    float tau = 0.f, beta = alpha, scale = 0.f;
    if (sigma > 0.f) {
        const float nrm = sqrtf(alpha * alpha + sigma);
        beta = alpha >= 0.f ? -nrm : nrm;
        tau = (beta - alpha) / beta;
        scale = 1.f / (alpha - beta);
    }
    const float vi = (i == 0) ? 1.f : (i < len ? x * scale : 0.f);
    if (h == 0) vs[i] = vi;
    if (k == 0) {
        if (h == 0 && i < len) W.at(r0 + i, s) = (i == 0) ? beta : 0.f;
    } else {
        if (h == 0) bk[0] = (i == 0) ? beta : 0.f;
        // left-apply to the block's other columns: w(c) = v^T B(:, c)
        float t[16];
#pragma unroll
        for (int c = 0; c < 16; ++c) t[c] = vi * bk[c];
        float z[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const auto sw = __builtin_amdgcn_permlane16_swap(__float_as_uint(t[c]), __float_as_uint(t[c + 8]), false,
                                                             false);
            z[c] = brow_sum16(__uint_as_float(sw[0]) + __uint_as_float(sw[1]));
        }
        // 16-lane row R = l >> 4 holds columns 16 h + 8 (R & 1) + c
        const int ri = lane & 15;
        float zz = z[0];
#pragma unroll
        for (int c = 1; c < 8; ++c) zz = (ri == c) ? z[c] : zz;
        if (ri < 8) ws[16 * h + 8 * ((lane >> 4) & 1) + ri] = zz;
        wave_lds_sync();
        const float tv = tau * vi;
#pragma unroll
        for (int c = 0; c < 16; ++c) {
            if (h == 0 && c == 0) continue;  // the annihilated column
            bk[c] -= tv * ws[16 * h + c];
        }
        const int p0 = r0 - NB;
        if (i < len) {
#pragma unroll
            for (int c = 0; c < 16; ++c) W.at(r0 + i, p0 + 16 * h + c) = bk[c];
        }
        wave_lds_sync();  // ws is reused below
    }
    wave_lds_sync();
    // two-sided update of the diagonal block D (rows/cols r0 .. r0+len-1)
    float dd[16], vj[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) {
        const int j = 16 * h + c;
        vj[c] = vs[j];
        float v = 0.f;
        if (i < len && j < len) v = (i >= j) ? W.at(r0 + i, r0 + j) : W.at(r0 + j, r0 + i);
        dd[c] = v;
    }
    float pp = 0.f;
#pragma unroll
    for (int c = 0; c < 16; ++c) pp += dd[c] * vj[c];
    pp += __shfl_xor(pp, 32, 64);
    pp *= tau;
    const float kk = bwave_sum(h == 0 ? pp * vi : 0.f);
    const float K = -0.5f * tau * kk;
    const float wi = pp + K * vi;
    if (h == 0) ws[i] = wi;
    wave_lds_sync();
#pragma unroll
    for (int c = 0; c < 16; ++c) {
        const int j = 16 * h + c;
        const float wj = ws[j];
        dd[c] -= vi * wj + wi * vj[c];
        if (i < len && j < len && j <= i) W.at(r0 + i, r0 + j) = dd[c];
    }
    // the reflector: Q2 block column (48 rows, zero outside [j0, j0 + len))
    const int jq = s % kQ2G;
    const int rr = lane - jq;
    const float vv = __shfl(vi, rr & 31, 64);  // every lane takes part in the shuffle
    if (lane < kQ2R) q2v[jq * kQ2R + lane] = (rr >= 0 && rr < NB) ? vv : 0.f;
    if (lane == 0) q2t[jq] = tau;
#pragma unroll
    for (int c = 0; c < 16; ++c) pv[c] = vj[c];
    ptau = tau;
    wave_lds_sync();
}

// Band -> tridiagonal for one matrix per workgroup.  Sweep s annihilates
// column s below the subdiagonal; its task k (rows s+1+k NB ..) applies the
// previous reflector from the right to the block below it, annihilates the
// block's first column with a new reflector, applies that from the left to
// the rest of the block and to the diagonal block from both sides.  The
// sweeps of a group run one per wave, task k of wave j at step k + 2 j: the
// tasks of one step touch disjoint elements, and a task runs after every
// task of the sweep before it that touches its rows.  The rows the group
// needs at a step (at most (2 kSbG - 1) NB of them) live in a circular LDS
// window; rows below are read ahead of each step, rows above written back
// after it.
__global__ void __launch_bounds__(kSbNT) sb2st_kernel(int n, float* __restrict__ BNDall, float* __restrict__ dall,
                                                       float* __restrict__ eall, float* __restrict__ Q2all) {
    constexpr int NB = kBandNB;
    extern __shared__ __align__(16) float sb_smem[];
    const int b = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    float* BND = BNDall + size_t(b) * n * kBandW;
    float* Q2 = Q2all + size_t(b) * q2_block_offset(n, q2_groups(n)) * kQ2BlockFloats;
    SbWin W{sb_smem};
    float* vs = sb_smem + size_t(kSbWR) * kSbWP + wave * 2 * NB;
    float* ws = vs + NB;
    const int nsweep = n - 2;

    // rows [r_lo, r_hi] of the band to or from the window, all threads,
    // 16-byte pieces
    auto load_rows = [&](int r_lo, int r_hi) {
        const int cnt = (r_hi - r_lo + 1) * (kBandW / 4);
        for (int q = tid; q < cnt; q += kSbNT) {
            const int r = r_lo + q / (kBandW / 4), c4 = (q % (kBandW / 4)) * 4;
            const float4 v = *reinterpret_cast<const float4*>(BND + size_t(r) * kBandW + c4);
            float* d = W.w + (r & (kSbWR - 1)) * kSbWP + c4;
            d[0] = v.x;
            d[1] = v.y;
            d[2] = v.z;
            d[3] = v.w;
        }
    };
    auto store_rows = [&](int r_lo, int r_hi) {
        const int cnt = (r_hi - r_lo + 1) * (kBandW / 4);
        for (int q = tid; q < cnt; q += kSbNT) {
            const int r = r_lo + q / (kBandW / 4), c4 = (q % (kBandW / 4)) * 4;
            const float* d = W.w + (r & (kSbWR - 1)) * kSbWP + c4;
            *reinterpret_cast<float4*>(BND + size_t(r) * kBandW + c4) = float4{d[0], d[1], d[2], d[3]};
        }
    };

    for (int s0 = 0; s0 < nsweep; s0 += kSbG) {
        const int ng = min(kSbG, nsweep - s0);
        int nstep = 0;
        for (int j = 0; j < ng; ++j) nstep = max(nstep, 2 * j + sb_tasks(n, s0 + j));
        const int g = s0 / kQ2G;
        float* q2g = Q2 + q2_block_offset(n, g) * kQ2BlockFloats;
        // rows step t needs: top(t) .. hi(t) = s0 + (t + 1) NB
        auto top_of = [&](int t) {
            int top = 0x7fffffff;
            for (int j = 0; j < ng; ++j) top = min(top, sb_task_row(n, s0 + j, t - 2 * j));
            return top;
        };
        int lo = s0, hi = min(n - 1, s0 + NB);
        __syncthreads();
        load_rows(lo, hi);
        float pv[16], ptau = 0.f;
#pragma unroll
        for (int c = 0; c < 16; ++c) pv[c] = 0.f;
        for (int t = 0; t < nstep; ++t) {
            // the next step's rows (at most NB: one 16-byte piece per thread)
            // are read while this step's tasks run; nobody touches them yet
            const int nhi = (t + 1 < nstep) ? min(n - 1, s0 + (t + 2) * NB) : hi;
            float4 pf = {0.f, 0.f, 0.f, 0.f};
            const int pr = hi + 1 + tid / (kBandW / 4), pc = (tid % (kBandW / 4)) * 4;
            const bool pre = pr <= nhi;
            if (pre) pf = *reinterpret_cast<const float4*>(BND + size_t(pr) * kBandW + pc);
            __syncthreads();
            const int s = s0 + wave, k = t - 2 * wave;
            if (wave < ng && k >= 0 && k < sb_tasks(n, s)) {
                float* blk = q2g + size_t(k) * kQ2BlockFloats;
                sb_task(n, s, k, W, vs, ws, pv, ptau, blk, blk + kQ2R * kQ2G, lane);
            }
            __syncthreads();
            // slide: write back the rows no longer needed; the read-ahead rows
            // go to slots whose rows left the window at least a step ago
            // (span of the window + NB < kSbWR)
            const int top = (t + 1 < nstep) ? top_of(t + 1) : 0x7fffffff;
            const int wlo = lo, whi = min(hi, top - 1);
            if (whi >= wlo) {
                store_rows(wlo, whi);
                lo = whi + 1;
            }
            if (pre) {
                float* d = W.w + (pr & (kSbWR - 1)) * kSbWP + pc;
                d[0] = pf.x;
                d[1] = pf.y;
                d[2] = pf.z;
                d[3] = pf.w;
            }
            hi = nhi;
        }
        __syncthreads();
        if (hi >= lo) store_rows(lo, hi);
        __syncthreads();
    }
    __syncthreads();
    // d and e of T from the band
    float* d = dall + size_t(b) * n;
    float* e = eall + size_t(b) * n;
    for (int i = tid; i < n; i += kSbNT) {
        d[i] = BND[size_t(i) * kBandW];
        e[i] = (i + 1 < n) ? BND[size_t(i + 1) * kBandW + 1] : 0.f;
    }
}

int sb2st(int batch, int n, float* BND, float* d, float* e, float* Q2, hipStream_t s) {
    if (n < 1) return fail(CIR_ERR_ARG, "sb2st: bad size");
    const size_t sm = kSbSmemFloats * sizeof(float);
    CIR_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&sb2st_kernel),
                                hipFuncAttributeMaxDynamicSharedMemorySize, int(sm)));
    hipLaunchKernelGGL(sb2st_kernel, dim3(unsigned(batch)), dim3(kSbNT), sm, s, n, BND, d, e, Q2);
    CIR_CHECK_LAUNCH();
    return CIR_OK;
}

}  // namespace cir

namespace cir {

typedef float floatx4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// Q2: the bulge-chasing reflectors as compact-WY blocks.
// ---------------------------------------------------------------------------

// Group and position of block q of a matrix.
__device__ __forceinline__ void q2_block_gk(int n, int64_t q, int* g, int* k) {
    int gg = 0;
    int64_t o = 0;
    const int ng = q2_groups(n);
    for (; gg < ng; ++gg) {
        const int t = sb_tasks(n, gg * kQ2G);
        if (q < o + t) break;
        o += t;
    }
    *g = gg;
    *k = int(q - o);
}

// Per block (one wave): reflectors that do not exist (sweeps past n - 3, or
// past their last task) are zeroed; T is formed (forward, columnwise) and
// -V T stored beside V, so that applying the block is Z <- Z + (-V T) (V^T Z).
__global__ void __launch_bounds__(64) q2_prep_kernel(int n, int64_t nblk, float* __restrict__ Q2all) {
    __shared__ float sv[kQ2R][kQ2G + 1], st[kQ2G][kQ2G + 1], sg[kQ2G], stau[kQ2G];
    const int64_t q = blockIdx.x;
    const int b = blockIdx.y, lane = threadIdx.x;
    if (q >= nblk) return;
    float* blk = Q2all + (size_t(b) * nblk + q) * kQ2BlockFloats;
    float* V = blk;
    float* tau = blk + kQ2R * kQ2G;
    float* NVT = tau + kQ2G;
    int g, k;
    q2_block_gk(n, q, &g, &k);
    for (int e = lane; e < kQ2R * kQ2G; e += 64) {
        const int j = e / kQ2R, r = e % kQ2R;
        const int s = g * kQ2G + j;
        const bool ex = k < sb_tasks(n, s);
        const float v = ex ? V[e] : 0.f;
        sv[r][j] = v;
        if (!ex) V[e] = 0.f;
    }
    if (lane < kQ2G) {
        const int s = g * kQ2G + lane;
        const bool ex = k < sb_tasks(n, s);
        stau[lane] = ex ? tau[lane] : 0.f;
        if (!ex) tau[lane] = 0.f;
    }
    __syncthreads();
    for (int j = 0; j < kQ2G; ++j) {
        if (lane < j) {
            float acc = 0.f;
            for (int r = 0; r < kQ2R; ++r) acc += sv[r][lane] * sv[r][j];
            sg[lane] = acc;
        }
        __syncthreads();
        if (lane < j) {
            float acc = 0.f;
            for (int q2 = lane; q2 < j; ++q2) acc += st[lane][q2] * sg[q2];
            st[lane][j] = -stau[j] * acc;
        } else if (lane == j) {
            st[j][j] = stau[j];
        } else if (lane < kQ2G) {
            st[lane][j] = 0.f;
        }
        __syncthreads();
    }
    for (int e = lane; e < kQ2R * kQ2G; e += 64) {
        const int j = e / kQ2R, r = e % kQ2R;
        float acc = 0.f;
        for (int q2 = 0; q2 <= j; ++q2) acc += sv[r][q2] * st[q2][j];
        NVT[e] = -acc;
    }
}

// Z <- Q2 Z on one 16-column slab of one matrix, the slab in LDS (rows past n
// zero).  Blocks run group by group from the last group, positions in
// ascending order; group g takes position k at step k + 2 (G - 1 - g), so the
// blocks of one step touch disjoint rows and every block runs after those it
// must follow.  Group g belongs to wave (G - 1 - g) mod 8.  A block on a
// wave: W = V^T Zb (16 x 16, K = 48) and Zb += (-V T) W, as 16x16x4 FP32 MFMA.
constexpr int kQ2NT = 512;
__global__ void __launch_bounds__(kQ2NT) q2_apply_kernel(int n, int ld, const float* __restrict__ Q2all,
                                                          float* __restrict__ Zall, int64_t nblk) {
    extern __shared__ __align__(16) float q2_smem[];
    const int slab = blockIdx.x, b = blockIdx.y;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int c0 = slab * kQ2S;
    const int rows = n + kQ2R;  // slab rows incl. zero padding
    float* Zs = q2_smem;
    float* Ws = q2_smem + size_t(rows) * kQ2S + wave * kQ2G * kQ2S;
    float* Z = Zall + size_t(b) * n * ld;
    const float* Q2 = Q2all + size_t(b) * nblk * kQ2BlockFloats;
    for (int e = tid; e < rows * kQ2S; e += kQ2NT) {
        const int r = e / kQ2S, c = e % kQ2S;
        Zs[e] = (r < n && c0 + c < n) ? Z[size_t(r) * ld + c0 + c] : 0.f;
    }
    const int NG = q2_groups(n);
    int nstep = 0;
    for (int g = 0; g < NG; ++g) nstep = max(nstep, 2 * (NG - 1 - g) + sb_tasks(n, g * kQ2G));
    const int l16 = lane & 15, lq = lane >> 4;
    for (int t = 0; t < nstep; ++t) {
        __syncthreads();
        for (int gi = wave; gi < NG; gi += kQ2NT / 64) {
            const int g = NG - 1 - gi;
            const int k = t - 2 * gi;
            if (k < 0 || k >= sb_tasks(n, g * kQ2G)) continue;
            const float* blk = Q2 + (q2_block_offset(n, g) + k) * kQ2BlockFloats;
            const float* V = blk;
            const float* NVT = blk + kQ2R * kQ2G + kQ2G;
            const int rb = g * kQ2G + 1 + k * kBandNB;
            float va[kQ2R / 4], na[3][4];
#pragma unroll
            for (int s = 0; s < kQ2R / 4; ++s) va[s] = V[l16 * kQ2R + 4 * s + lq];
#pragma unroll
            for (int rt = 0; rt < 3; ++rt)
#pragma unroll
                for (int s = 0; s < 4; ++s) na[rt][s] = NVT[(4 * s + lq) * kQ2R + rt * 16 + l16];
            floatx4 w = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int s = 0; s < kQ2R / 4; ++s) {
                const float zb = Zs[(rb + 4 * s + lq) * kQ2S + l16];
                w = __builtin_amdgcn_mfma_f32_16x16x4f32(va[s], zb, w, 0, 0, 0);
            }
            // W (reflector q = 4 lq + r, column l16) to the wave's scratch
#pragma unroll
            for (int r = 0; r < 4; ++r) Ws[(4 * lq + r) * kQ2S + l16] = w[r];
            wave_lds_sync();
            float wb[4];
#pragma unroll
            for (int s = 0; s < 4; ++s) wb[s] = Ws[(4 * s + lq) * kQ2S + l16];
#pragma unroll
            for (int rt = 0; rt < 3; ++rt) {
                floatx4 acc;
#pragma unroll
                for (int r = 0; r < 4; ++r) acc[r] = Zs[(rb + rt * 16 + 4 * lq + r) * kQ2S + l16];
#pragma unroll
                for (int s = 0; s < 4; ++s) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(na[rt][s], wb[s], acc, 0, 0, 0);
#pragma unroll
                for (int r = 0; r < 4; ++r) Zs[(rb + rt * 16 + 4 * lq + r) * kQ2S + l16] = acc[r];
            }
            wave_lds_sync();
        }
    }
    __syncthreads();
    for (int e = tid; e < n * kQ2S; e += kQ2NT) {
        const int r = e / kQ2S, c = e % kQ2S;
        if (c0 + c < n) Z[size_t(r) * ld + c0 + c] = Zs[e];
    }
}

int q2_apply(int batch, int n, int ld, const float* Q2, float* Z, hipStream_t s) {
    const int64_t nblk = q2_blocks(n);
    if (nblk <= 0) return CIR_OK;
    {
        dim3 grid(static_cast<unsigned>(nblk), static_cast<unsigned>(batch));
        hipLaunchKernelGGL(q2_prep_kernel, grid, dim3(64), 0, s, n, nblk, const_cast<float*>(Q2));
        CIR_CHECK_LAUNCH();
    }
    const size_t sm = (size_t(n + kQ2R) * kQ2S + size_t(kQ2NT / 64) * kQ2G * kQ2S) * sizeof(float);
    if (sm > 160 * 1024) return fail(CIR_ERR_ARG, "q2_apply: n too large for the LDS slab");
    CIR_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&q2_apply_kernel),
                                hipFuncAttributeMaxDynamicSharedMemorySize, int(sm)));
    dim3 grid(unsigned(ceil_div(n, kQ2S)), unsigned(batch));
    hipLaunchKernelGGL(q2_apply_kernel, grid, dim3(kQ2NT), sm, s, n, ld, Q2, Z, nblk);
    CIR_CHECK_LAUNCH();
    return CIR_OK;
}

}  // namespace cir
