// k_band.hip -- two-stage reduction of the batched FP32 eigensolver (gfx950),
// replacing the one-stage Householder tridiagonalisation of magma_ssyevd_gpu
// (calcIR.cu:420-448) for batches of mid-size Hamiltonians.
//
//   Stage 1, dense -> band of half-width kBandNB (LAPACK sytrd_sy2sb form):
//     for each panel of kBandNB columns, sb_panel_kernel factors the panel
//     below the band, A(r0:n, j0:j0+NB) = Q R, by Householder reflectors in
//     registers (one workgroup per matrix); the trailing matrix then gets the
//     two-sided update A22 <- Q^T A22 Q on its lower triangle, batched over
//     every matrix: X = A22 Yh with Yh = Y T (sb_symm_kernel), W = X -
//     Y (Yh^T X)/2 (two small products), A22 -= W Y^T + Y W^T
//     (sb_syr2k_kernel), on 32x32x2 FP32 MFMA.
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

#include <algorithm>
#include <cfloat>

#include "band.hpp"
#include "gemm.hpp"
#include "split_bf16.hpp"
#include "tuner.hpp"

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

// Make this wave's LDS stores visible to its own later LDS loads (a wave's LDS
// operations complete in order; this keeps the compiler from moving them).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
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
// tau, Y into A (back-transform layout) and into the Y half of the row-major
// panel P (absolute rows, pitch 2 NB), Yh = Y T row-major (panel rows,
// pitch NB).
template <int RPT>
__global__ void __launch_bounds__(kPanNT) sb_panel_kernel(int n, int ld, int j0, float* __restrict__ Aall,
                                                          float* __restrict__ BNDall, float* __restrict__ tauall,
                                                          float* __restrict__ Pall, float* __restrict__ Yhall) {
    constexpr int NB = kBandNB;
    __shared__ float redA[kPanNW], redC[kPanNW][NB], dotv[NB], taus[NB], salpha;
    __shared__ float gram[NB][NB + 1], Tm[NB][NB + 1];
    const int b = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    float* A = Aall + size_t(b) * n * ld;
    float* BND = BNDall + size_t(b) * n * kBandW;
    float* P = Pall + size_t(b) * n * 2 * NB;
    float* Yh = Yhall + size_t(b) * n * NB;
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
        for (int c = 0; c < NB; ++c) A[size_t(j0 + c) * ld + r0 + row] = y[c];
        float4* pr = reinterpret_cast<float4*>(P + size_t(r0 + row) * 2 * NB);
        float4* yr = reinterpret_cast<float4*>(Yh + size_t(row) * NB);
#pragma unroll
        for (int q = 0; q < NB / 4; ++q) pr[q] = float4{y[4 * q], y[4 * q + 1], y[4 * q + 2], y[4 * q + 3]};
        float yh[NB];
#pragma unroll
        for (int c = 0; c < NB; ++c) {
            float acc = 0.f;
#pragma unroll
            for (int q = 0; q <= c; ++q) acc += y[q] * Tm[q][c];
            yh[c] = acc;
        }
#pragma unroll
        for (int q = 0; q < NB / 4; ++q) yr[q] = float4{yh[4 * q], yh[4 * q + 1], yh[4 * q + 2], yh[4 * q + 3]};
    }
}

typedef float floatx16 __attribute__((ext_vector_type(16)));

// Workgroup x of matrix b in a grid of gx workgroups per matrix over batch
// matrices, launched one-dimensional (xcd_grid): workgroup i runs on XCD i
// mod 8 (round-robin dispatch) and XCD x takes the x-th eighth of the
// (matrix, workgroup) list, so a matrix's workgroups run together on one XCD
// and share its panel through that XCD's L2.  False for the padding past the
// list.
__device__ __forceinline__ bool xcd_slot(int gx, int batch, int& x, int& b) {
    const int64_t total = int64_t(gx) * batch, per = (total + kXcds - 1) / kXcds;
    const int64_t pos = int64_t(blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
    if (pos >= total) return false;
    b = int(pos / gx);
    x = int(pos % gx);
    return true;
}
inline dim3 xcd_grid(int gx, int batch) { return dim3(unsigned(ceil_div(int64_t(gx) * batch, int64_t(kXcds)) * kXcds)); }

// X = A22 Yh on the lower triangle of A22 (m x m, column-major, pitch ld,
// first row and column r0), X row-major with pitch 2 NB in the W half of the
// row-major panel P (rows r0.., P[r * 2NB + NB + p]); Yh row-major, pitch NB
// (panel rows).  Wave w of workgroup x owns block row I = 4 x + w and reads
// every tile once from its own side: the row tiles (I, J < I) go straight
// into the MFMA A operand (lanes along rows, coalesced), the column tiles
// (J >= I, I) are read along their rows too and transposed through the
// wave's LDS tile.  The next tile's loads are in flight while a tile's MFMA
// chain runs.  32x32x2 FP32 MFMA, K = 32 per tile.
__global__ void __launch_bounds__(256) sb_symm_kernel(int n, int ld, int r0, const float* __restrict__ Aall,
                                                      const float* __restrict__ Yhall, float* __restrict__ Pall,
                                                      int batch) {
    constexpr int NB = kBandNB;
    __shared__ float tl[4][32][33];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int m = n - r0, nbt = (m + 31) / 32;
    int x, b;
    if (!xcd_slot((nbt + 3) / 4, batch, x, b)) return;
    const int I = x * 4 + wave;
    if (I >= nbt) return;
    const float* A = Aall + size_t(b) * n * ld + size_t(r0) * ld + r0;
    const float* Yh = Yhall + size_t(b) * n * NB;
    float* P = Pall + size_t(b) * n * 2 * NB + size_t(r0) * 2 * NB;
    const int l32 = lane & 31, h = lane >> 5;
    const int ri = I * 32 + l32;  // this lane's row of block row I (M index)
    float (*T)[33] = tl[wave];
    // raw loads of tile J: the B operand (Yh rows J*32 + 2q + h) and the A
    // tile, as the A operand itself (row tile) or along the tile's rows for
    // the LDS transpose (column tile, T[2q + h][l32] = A22(J*32 + l32, I*32 + 2q + h))
    // Every load is issued with its index clamped below m and nothing is
    // masked until the values are used (the loads of a tile go out back to
    // back, a tile ahead): rows and columns of the output past m are never
    // stored, and K indices past m are cancelled by zeroing the B operand.
    const int ric = min(ri, m - 1);
    auto load = [&](int J, float (&bv)[16], float (&raw)[16]) {
#pragma unroll
        for (int q = 0; q < 16; ++q) bv[q] = Yh[size_t(min(J * 32 + 2 * q + h, m - 1)) * NB + l32];
        // row tile (J < I): A(J*32 + 2q + h, ri); column tile: A(I*32 + 2q + h,
        // J*32 + l32) (columns of the column-major lower triangle)
        const bool rowt = J < I;
        const int c0 = (rowt ? J : I) * 32 + h;
        const float* base = A + (rowt ? ric : min(J * 32 + l32, m - 1));
#pragma unroll
        for (int q = 0; q < 16; ++q) raw[q] = base[size_t(min(c0 + 2 * q, m - 1)) * ld];
    };
    floatx16 acc;
#pragma unroll
    for (int q = 0; q < 16; ++q) acc[q] = 0.f;
    // tile J from (bv, raw) while tile J + 1 is read into (bn, rn) (past the
    // last tile: clamped reads, never used); the two buffer pairs alternate,
    // so no copy waits for a read in flight
    auto tile = [&](int J, float (&bv)[16], float (&raw)[16], float (&bn)[16], float (&rn)[16]) {
        load(J + 1, bn, rn);
        __builtin_amdgcn_sched_barrier(0);  // keep the reads ahead of the products
#pragma unroll
        for (int q = 0; q < 16; ++q) bv[q] = (J * 32 + 2 * q + h < m) ? bv[q] : 0.f;
        float av[16];
        if (J < I) {
#pragma unroll
            for (int q = 0; q < 16; ++q) av[q] = raw[q];
        } else {
#pragma unroll
            for (int q = 0; q < 16; ++q) T[2 * q + h][l32] = raw[q];
            wave_lds_sync();
            // A operand A22(ri, J*32 + k) = A(J*32 + k, ri): T[l32][k]; the
            // diagonal tile's upper part mirrored from its lower part
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int k = 2 * q + h;
                av[q] = (J > I || k >= l32) ? T[l32][k] : T[k][l32];
            }
            wave_lds_sync();
        }
#pragma unroll
        for (int q = 0; q < 16; ++q) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av[q], bv[q], acc, 0, 0, 0);
    };
    float b0[16], r0v[16], b1[16], r1v[16];
    load(0, b0, r0v);
    for (int J = 0; J < nbt; J += 2) {
        tile(J, b0, r0v, b1, r1v);
        if (J + 1 >= nbt) break;
        tile(J + 1, b1, r1v, b0, r0v);
    }
    // acc(i, p): lane l holds column p = l32, rows 8 a + 4 h + c
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int r = I * 32 + 8 * (q >> 2) + 4 * h + (q & 3);
        if (r < m) P[size_t(r) * 2 * NB + NB + l32] = acc[q];
    }
}

// Word 3 of a raw buffer resource on gfx9-family GPUs (32-bit data format,
// no swizzle, range checking on offsets).
constexpr int kBufferWord3 = 0x00020000;

// A22 -= W Y^T + Y W^T on the lower triangle, from the row-major panel P
// (P[r * 2NB + k]: Y(r, k) for k < NB, W(r, k - NB) after).  Wave w of
// workgroup x owns block row br = 4 x + w and every block column bc <= br.
// The tile is formed transposed (MFMA row = column of A, MFMA column = row
// of A) so that lanes run along rows of the column-major A; step s of the
// K = 2NB sum takes panel position 32 h + s from lane half h, the A operand
// from the panel row of the tile's column at that position, the B operand
// from the panel row of the tile's row at the swapped half (k ^ NB).  The
// next tile's A values and operand are in flight while a tile's MFMA chain
// runs.
__global__ void __launch_bounds__(256) sb_syr2k_kernel(int n, int ld, int r0, float* __restrict__ Aall,
                                                       const float* __restrict__ Pall, int batch) {
    constexpr int NB = kBandNB, PW = 2 * kBandNB;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int m = n - r0, nbt = (m + 31) / 32;
    int x, b;
    if (!xcd_slot((nbt + 3) / 4, batch, x, b)) return;
    const int br = x * 4 + wave;
    if (br >= nbt) return;
    float* A = Aall + size_t(b) * n * ld + size_t(r0) * ld + r0;
    const float* P = Pall + size_t(b) * n * PW + size_t(r0) * PW;
    const int h = lane >> 5, l32 = lane & 31;
    const int r = br * 32 + l32;  // row of A22 this lane's column of the output holds
    // Loads are issued with row and column indices clamped below m and never
    // masked: an output element past m, or above the diagonal, is not stored,
    // and no K index runs past the panel.
    auto panel_row = [&](int row, int half, float (&v)[32]) {
        const float4* src = reinterpret_cast<const float4*>(P + size_t(min(row, m - 1)) * PW + half * NB);
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const float4 t = src[q];
            v[4 * q] = t.x;
            v[4 * q + 1] = t.y;
            v[4 * q + 2] = t.z;
            v[4 * q + 3] = t.w;
        }
    };
    // A22 of this matrix as a raw buffer (offsets in bytes, at most n * ld * 4)
    const __amdgpu_buffer_rsrc_t rsA = __builtin_amdgcn_make_buffer_rsrc(A, 0, 0x7fffffff, kBufferWord3);
    const int rc = min(r, m - 1);
    auto load_old = [&](int bc, float (&o)[16]) {
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int c = bc * 32 + 8 * (q >> 2) + 4 * h + (q & 3);
            o[q] = A[size_t(min(c, m - 1)) * ld + rc];
        }
    };
    float bv[32];
    panel_row(r, 1 - h, bv);
    // tile bc from (av, old) while tile bc + 1's operand and old values are
    // read into (an, on) (past the last tile: clamped reads, never used); the
    // two buffer pairs alternate, so no copy waits for a read in flight
    auto tile = [&](int bc, float (&av)[32], float (&old)[16], float (&an)[32], float (&on)[16]) {
        panel_row((bc + 1) * 32 + l32, h, an);
        load_old(bc + 1, on);
        __builtin_amdgcn_sched_barrier(0);  // keep the reads ahead of the products
        floatx16 acc;
#pragma unroll
        for (int q = 0; q < 16; ++q) acc[q] = 0.f;
#pragma unroll
        for (int q = 0; q < 32; ++q) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av[q], bv[q], acc, 0, 0, 0);
        // buffer stores, an element outside the lower triangle of A22 given
        // an offset past the buffer's range (the store is dropped): no branch,
        // so the count of memory operations in flight stays known and no wait
        // lands on the reads of the next tile
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int c = bc * 32 + 8 * (q >> 2) + 4 * h + (q & 3);
            const bool ok = c < m && r < m && r >= c;
            __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(old[q] - acc[q]), rsA,
                                                  ok ? (c * ld + r) * 4 : int(0x80000000), 0, 0);
        }
    };
    float a0[32], o0[16], a1[32], o1[16];
    panel_row(l32, h, a0);
    load_old(0, o0);
    for (int bc = 0; bc <= br; bc += 2) {
        tile(bc, a0, o0, a1, o1);
        if (bc + 1 > br) break;
        tile(bc + 1, a1, o1, a0, o0);
    }
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

int sy2sb(int batch, int n, int ld, float* A, float* BND, float* tau, float* P, float* Yh, float* Mm, hipStream_t s) {
    constexpr int NB = kBandNB;
    if (n > sb_max_n()) return fail(CIR_ERR_ARG, "sy2sb: n too large for the register panel");
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
                               P, Yh);
        else if (rpt <= 2)
            hipLaunchKernelGGL(sb_panel_kernel<2>, dim3(unsigned(batch)), dim3(kPanNT), 0, s, n, ld, j0, A, BND, tau,
                               P, Yh);
        else if (rpt <= 3)
            hipLaunchKernelGGL(sb_panel_kernel<3>, dim3(unsigned(batch)), dim3(kPanNT), 0, s, n, ld, j0, A, BND, tau,
                               P, Yh);
        else
            hipLaunchKernelGGL(sb_panel_kernel<4>, dim3(unsigned(batch)), dim3(kPanNT), 0, s, n, ld, j0, A, BND, tau,
                               P, Yh);
        CIR_CHECK_LAUNCH();
        const int nbt = int(ceil_div(m, 32));
        const dim3 grid = xcd_grid(int(ceil_div(nbt, 4)), batch);
        // X = A22 Yh into the W half of P
        hipLaunchKernelGGL(sb_symm_kernel, grid, dim3(256), 0, s, n, ld, r0, A, Yh, P, batch);
        CIR_CHECK_LAUNCH();
        const int64_t ps = int64_t(n) * 2 * NB;
        GemmArgs g;
        // Mm = Yh^T X (NB x NB)
        g.batch = batch;
        g.M = NB;
        g.N = NB;
        g.K = m;
        g.A = Yh;
        g.a_m = 1;
        g.a_k = NB;
        g.a_b = int64_t(n) * NB;
        g.B = P + size_t(r0) * 2 * NB + NB;
        g.b_k = 2 * NB;
        g.b_n = 1;
        g.b_b = ps;
        g.C = Mm;
        g.c_m = NB;
        g.c_n = 1;
        g.c_b = int64_t(NB) * NB;
        CIR_TRY(gemm(CIR_PREC_FP32, g, s));
        // W = X - Y Mm / 2 in place of X
        g = GemmArgs{};
        g.batch = batch;
        g.M = m;
        g.N = NB;
        g.K = NB;
        g.A = P + size_t(r0) * 2 * NB;
        g.a_m = 2 * NB;
        g.a_k = 1;
        g.a_b = ps;
        g.B = Mm;
        g.b_k = NB;
        g.b_n = 1;
        g.b_b = int64_t(NB) * NB;
        g.C = P + size_t(r0) * 2 * NB + NB;
        g.c_m = 2 * NB;
        g.c_n = 1;
        g.c_b = ps;
        g.alpha = -0.5;
        g.beta = 1.0;
        CIR_TRY(gemm(CIR_PREC_FP32, g, s));
        // A22 -= W Y^T + Y W^T (lower triangle)
        hipLaunchKernelGGL(sb_syr2k_kernel, grid, dim3(256), 0, s, n, ld, r0, A, P, batch);
        CIR_CHECK_LAUNCH();
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
constexpr int kSbWP = kBandW;            // window row pitch (floats)
// window, per-wave task scratch (v, two reduced vectors, a per-lane sink),
// and one 16-byte sink for the read-ahead pieces of threads with no row to add
constexpr int kSbScratch = 3 * kBandNB + 64;
constexpr size_t kSbSmemFloats = size_t(kSbWR) * kSbWP + size_t(kSbG) * kSbScratch + 4;
static_assert(kBandNB * (kBandW / 4) <= kSbNT, "one step's new band rows must fit one 16-byte piece per thread");
static_assert((2 * kSbG - 1) * kBandNB - kSbG + kBandNB < kSbWR, "window must hold a step's rows plus the read-ahead");


// Band element (r, c), r - c < kBandW, of a row in the window is at
// sb_rowbase(r) - c.
__device__ __forceinline__ int sb_rowbase(int r) { return (r & (kSbWR - 1)) * kSbWP + r; }

// First row a sweep still needs at its task k (k < 0: not started, so its
// first task's row), or INT_MAX once it is done.
__device__ __forceinline__ int sb_task_row(int n, int s, int k) {
    return (k < sb_tasks(n, s)) ? s + 1 + max(k, 0) * kBandNB : 0x7fffffff;
}

// Sum of a value over the two lanes i and i + 32 of a wave, in both.
__device__ __forceinline__ float half_pair_sum(float x) {
    const auto sw = __builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(x), false, false);
    return __uint_as_float(sw[0]) + __uint_as_float(sw[1]);
}
// The value of lane i (i < 32) in lanes i and i + 32.
__device__ __forceinline__ float from_low_half(float x) {
    return __uint_as_float(__builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(x), false, false)[0]);
}

// One wave's share of a bulge-chasing step: task (s, k) of its sweep.
// Lane l: row i = l & 31 of the task's block, column half h = l >> 5
// (columns 16 h .. 16 h + 15).  pv / ptau: the sweep's previous reflector
// (columns of the previous block, this lane's half), replaced by this task's;
// q2val: this task's reflector as a Q2 block column.  The wave's scratch sc
// holds v, the two reduced vectors and a per-lane sink for masked-off stores.
// Every window access is unconditional (lanes outside the block read a row
// inside it and discard the value, or store to the sink), so the accesses of
// a task issue back to back.  The diagonal block is read at the start (no
// earlier part of the task writes it), and the left application to the block
// below and the two-sided update of the diagonal block share one pass of
// reductions.
__device__ void sb_task(int n, int s, int k, float* w, float* sc, float (&pv)[16], float& ptau, float& q2val,
                        int lane) {
    constexpr int NB = kBandNB;
    float* vs = sc;
    float* wsB = sc + NB;
    float* wsD = sc + 2 * NB;
    float* sink = sc + 3 * NB + lane;
    const int i = lane & 31, h = lane >> 5;
    const int r0 = s + 1 + k * NB;
    const int len = min(NB, n - r0);
    const int p0 = r0 - NB;
    const bool rowok = i < len;
    const int rb = sb_rowbase(r0 + min(i, len - 1));  // this lane's row (clamped into the block)
    float bk[16], dd[16];
    float x = 0.f;
    if (k == 0) {
        const float v = w[rb - s];
        x = rowok ? v : 0.f;
    } else {
        // the block below the previous one, columns p0 .. p0 + NB - 1
        const float* bp = w + (rb - p0 - 16 * h - 15);
#pragma unroll
        for (int c = 0; c < 16; ++c) {
            const float v = bp[15 - c];
            bk[c] = rowok ? v : 0.f;
        }
    }
    // the diagonal block D (rows/cols r0 .. r0+len-1), both triangles, from
    // its lower triangle
#pragma unroll
    for (int c = 0; c < 16; ++c) {
        const int j = 16 * h + c;
        const int ra = min(max(i, j), len - 1), cb = min(min(i, j), ra);
        const float v = w[sb_rowbase(r0 + ra) - (r0 + cb)];
        dd[c] = (rowok && j < len) ? v : 0.f;
    }
    if (k > 0) {
        // the previous reflector from the right (this fills the bulge)
        float u0 = 0.f, u1 = 0.f;
#pragma unroll
        for (int c = 0; c < 16; c += 2) {
            u0 += bk[c] * pv[c];
            u1 += bk[c + 1] * pv[c + 1];
        }
        const float f = ptau * half_pair_sum(u0 + u1);
#pragma unroll
        for (int c = 0; c < 16; ++c) bk[c] -= f * pv[c];
        x = from_low_half(bk[0]);  // the block's first column, from the h = 0 lanes
    }
    // reflector annihilating x[1..len)
    const float sq = (h == 0 && i >= 1 && i < len) ? x * x : 0.f;
    const float sigma = bwave_sum(sq);
    const float alpha = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), 0));
    float tau = 0.f, beta = alpha, scale = 0.f;
    if (sigma > 0.f) {
        const float nrm = sqrtf(alpha * alpha + sigma);
        beta = alpha >= 0.f ? -nrm : nrm;
        tau = (beta - alpha) / beta;
        scale = 1.f / (alpha - beta);
    }
    const float vi = (i == 0) ? 1.f : (i < len ? x * scale : 0.f);
    if (h == 0) vs[i] = vi;
    wave_lds_sync();
    float vj[16];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const float4 t = reinterpret_cast<const float4*>(vs + 16 * h)[q];
        vj[4 * q] = t.x;
        vj[4 * q + 1] = t.y;
        vj[4 * q + 2] = t.z;
        vj[4 * q + 3] = t.w;
    }
    // D: p = tau D v, K = -tau/2 p.v, w = p + K v
    float p0a = 0.f, p1a = 0.f;
#pragma unroll
    for (int c = 0; c < 16; c += 2) {
        p0a += dd[c] * vj[c];
        p1a += dd[c + 1] * vj[c + 1];
    }
    const float pp = tau * half_pair_sum(p0a + p1a);
    if (k > 0) {
        if (h == 0) bk[0] = (i == 0) ? beta : 0.f;
        // the new reflector from the left on the block's other columns:
        // w(c) = v^T B(:, c), reduce-scattered over the half's 32 lanes
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
        if (ri < 8) wsB[16 * h + 8 * ((lane >> 4) & 1) + ri] = zz;
    }
    const float kk = bwave_sum(h == 0 ? pp * vi : 0.f);
    const float K = -0.5f * tau * kk;
    const float wi = pp + K * vi;
    if (h == 0) wsD[i] = wi;
    wave_lds_sync();
    if (k > 0) {
        const float tv = tau * vi;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const float4 t = reinterpret_cast<const float4*>(wsB + 16 * h)[q];
            const float tw[4] = {t.x, t.y, t.z, t.w};
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int c = 4 * q + e;
                if (!(h == 0 && c == 0)) bk[c] -= tv * tw[e];  // column 0 is the annihilated one
            }
        }
        if (rowok) {
            float* bq = w + (rb - p0 - 16 * h - 15);
#pragma unroll
            for (int c = 0; c < 16; ++c) bq[15 - c] = bk[c];
        }
    } else if (h == 0 && rowok) {
        w[rb - s] = (i == 0) ? beta : 0.f;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const float4 t = reinterpret_cast<const float4*>(wsD + 16 * h)[q];
        const float tw[4] = {t.x, t.y, t.z, t.w};
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int c = 4 * q + e, j = 16 * h + c;
            dd[c] -= vi * tw[e] + wi * vj[c];
            float* dst = (rowok && j < len && j <= i) ? w + (rb - r0 - j) : sink;
            *dst = dd[c];
        }
    }
    // the reflector as its Q2 block column (48 rows, zero outside [jq, jq +
    // len)): lane r < 48 holds row r
    const int jq = s % kQ2G;
    const int rr = lane - jq;
    const float vv = __shfl(vi, rr & 31, 64);  // every lane takes part in the shuffle
    q2val = (rr >= 0 && rr < NB) ? vv : 0.f;
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
    float* const win = sb_smem;
    float* sc = sb_smem + size_t(kSbWR) * kSbWP + wave * kSbScratch;
    const int nsweep = n - 2;

    // rows [r_lo, r_hi] of the band to or from the window, all threads,
    // 16-byte pieces
    auto load_rows = [&](int r_lo, int r_hi) {
        const int cnt = (r_hi - r_lo + 1) * (kBandW / 4);
        for (int q = tid; q < cnt; q += kSbNT) {
            const int r = r_lo + q / (kBandW / 4), c4 = (q % (kBandW / 4)) * 4;
            *reinterpret_cast<float4*>(win + (r & (kSbWR - 1)) * kSbWP + c4) =
                *reinterpret_cast<const float4*>(BND + size_t(r) * kBandW + c4);
        }
    };
    auto store_rows = [&](int r_lo, int r_hi) {
        const int cnt = (r_hi - r_lo + 1) * (kBandW / 4);
        for (int q = tid; q < cnt; q += kSbNT) {
            const int r = r_lo + q / (kBandW / 4), c4 = (q % (kBandW / 4)) * 4;
            *reinterpret_cast<float4*>(BND + size_t(r) * kBandW + c4) =
                *reinterpret_cast<const float4*>(win + (r & (kSbWR - 1)) * kSbWP + c4);
        }
    };
    // this thread's 16-byte piece of the rows read ahead (row pr, columns pc ..)
    const int pr_off = tid / (kBandW / 4), pc = (tid % (kBandW / 4)) * 4;
    float* sink = sb_smem + kSbSmemFloats - 4;

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
        auto hi_of = [&](int t) { return min(n - 1, s0 + (t + 1) * NB); };
        int lo = s0, hi = hi_of(0);
        __syncthreads();
        load_rows(lo, hi);
        // read-ahead of the rows step t + 1 adds, (hi(t), hi(t + 1)]: at most NB
        // rows, one 16-byte piece per thread, in flight during step t's tasks
        // (read and written by every thread, so that the wait for it is on
        // every path: a thread with no row reads the last row and writes the
        // sink)
        int pr = hi + 1 + pr_off;
        bool pre = nstep > 1 && pr <= hi_of(1);
        float4 pf = *reinterpret_cast<const float4*>(BND + size_t(min(pr, n - 1)) * kBandW + pc);
        float pv[16], ptau = 0.f;
#pragma unroll
        for (int c = 0; c < 16; ++c) pv[c] = 0.f;
        for (int t = 0; t < nstep; ++t) {
            __syncthreads();
            const int s = s0 + wave, k = t - 2 * wave;
            const bool act = wave < ng && k >= 0 && k < sb_tasks(n, s);
            float q2val = 0.f;
            if (act) sb_task(n, s, k, win, sc, pv, ptau, q2val, lane);
            __syncthreads();
            // the read-ahead rows go to slots whose rows left the window at
            // least a step ago (span of the window + NB < kSbWR); they are
            // written before this step's stores are issued, so waiting for
            // the read waits only for memory operations issued before the
            // tasks
            *reinterpret_cast<float4*>(pre ? win + (pr & (kSbWR - 1)) * kSbWP + pc : sink) = pf;
            // write back the rows no longer needed
            const int top = (t + 1 < nstep) ? top_of(t + 1) : 0x7fffffff;
            const int whi = min(hi, top - 1);
            if (whi >= lo) {
                store_rows(lo, whi);
                lo = whi + 1;
            }
            if (t + 1 < nstep) hi = hi_of(t + 1);
            if (act) {
                float* blk = q2g + size_t(k) * kQ2BlockFloats;
                const int jq = s % kQ2G;
                if (lane < kQ2R) blk[jq * kQ2R + lane] = q2val;
                if (lane == 0) blk[kQ2R * kQ2G + jq] = ptau;
            }
            pr = hi + 1 + pr_off;
            pre = t + 2 < nstep && pr <= hi_of(t + 2);
            pf = *reinterpret_cast<const float4*>(BND + size_t(min(pr, n - 1)) * kBandW + pc);
        }
        __syncthreads();
        if (hi >= lo) store_rows(lo, hi);
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

bool deterministic_eigh_size(int n) { return n >= 2 * kBandNB + 2 && n <= sb_max_n(); }
int deterministic_eigh_max() { return sb_max_n(); }

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
// past their last task) are zeroed; T is formed (forward, columnwise), and V
// and -V T are stored in the operand order of the 16x16x4 MFMA products that
// apply the block (Z <- Z + (-V T) (V^T Z)): lane l's 12 values of each
// contiguous, V_op[l][s] = V(4 s + l/16, l%16) and NVT_op[l][4 rt + s] =
// (-V T)(16 rt + l%16, 4 s + l/16).  V_op replaces V in place.
__global__ void __launch_bounds__(64) q2_prep_kernel(int n, int64_t nblk, float* __restrict__ Q2all, int bf16) {
    __shared__ float sv[kQ2R][kQ2G + 1], st[kQ2G][kQ2G + 1], snv[kQ2R][kQ2G + 1], sgm[kQ2G][kQ2G + 1], stau[kQ2G];
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
        const bool ex = k < sb_tasks(n, g * kQ2G + j);
        sv[r][j] = ex ? V[e] : 0.f;
    }
    if (lane < kQ2G) stau[lane] = (k < sb_tasks(n, g * kQ2G + lane)) ? tau[lane] : 0.f;
    __syncthreads();
    // the Gram matrix V^T V above its diagonal, all entries at once (4 per
    // lane, each a 48-term dot product), then T column by column from it
    for (int e = lane; e < kQ2G * kQ2G; e += 64) {
        const int i = e / kQ2G, j = e % kQ2G;
        if (i < j) {
            float acc = 0.f;
            for (int r = 0; r < kQ2R; ++r) acc += sv[r][i] * sv[r][j];
            sgm[i][j] = acc;
        }
    }
    __syncthreads();
    for (int j = 0; j < kQ2G; ++j) {
        if (lane < j) {
            float acc = 0.f;
            for (int q2 = lane; q2 < j; ++q2) acc += st[lane][q2] * sgm[q2][j];
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
        snv[r][j] = -acc;
    }
    __syncthreads();
    const int l16 = lane & 15, lq = lane >> 4;
    // V^T as the A operand of the W product: register 4 q + e of lane (lq,
    // l16) holds V(16 q + 4 lq + e, l16)
#pragma unroll
    for (int s = 0; s < kQ2R / 4; ++s) V[lane * (kQ2R / 4) + s] = sv[16 * (s >> 2) + 4 * lq + (s & 3)][l16];
    // -V T as the A operand of the update, register 4 rt + s of lane (lq,
    // l16): FP32 16x16x4 steps take NV(16 rt + l16, 4 s + lq), split-BF16
    // 16x16x16 products NV(16 rt + l16, 4 lq + s)
#pragma unroll
    for (int rt = 0; rt < 3; ++rt)
#pragma unroll
        for (int s = 0; s < 4; ++s)
            NVT[lane * 12 + 4 * rt + s] = bf16 ? snv[16 * rt + l16][4 * lq + s] : snv[16 * rt + l16][4 * s + lq];
    if (lane < kQ2G) tau[lane] = stau[lane];
}

// Z <- Q2 Z on one 16-column slab of one matrix, the slab in LDS (rows past n
// zero).  Blocks apply group by group from the last group, positions in
// ascending order.  Block (g, k) spans rows 16 g + 1 + 32 k .. 16 g + 47 +
// 32 k, so it must follow (g, k - 1) and, of the group after it, (g + 1, k - 1)
// and (g + 1, k); group g takes position k at step k + (G - 1 - g), which
// orders all of these, and the blocks of one step, (g, k), (g + 1, k + 1), ...,
// start 48 rows apart and touch disjoint rows.  Group g belongs to wave
// (G - 1 - g) mod (waves per workgroup), which reads its next block's
// operands while it applies the current one.  A block: W = V^T Zb
// (16 x 16, K = 48) and Zb += (-V T) W, as 16x16x4 FP32 MFMA, W passed from
// one product to the next in registers.
constexpr int kQ2MaxGroups = 160;
static_assert(kQ2MaxGroups * kQ2G >= 4 * 512 + kBandNB, "group table too small for sb_max_n()");
// The slab is stored column by column, row r of Z at index r + 3 of its
// column (block rows then start at multiples of 4, so a lane's 4 rows are one
// 16-byte access), with a column pitch of 4 mod 64 floats (the 16 columns of
// one access fall in distinct banks) covering rows n .. n + 47 as zeros.
__host__ __device__ inline int q2_pitch(int n) {
    const int p = n + kQ2R + 3;
    return p + ((4 - p) & 63);
}
template <int kQ2NT, bool BF>
__global__ void __launch_bounds__(kQ2NT) q2_apply_kernel(int n, int ld, const float* __restrict__ Q2all,
                                                          float* __restrict__ Zall, int64_t nblk, int nmat) {
    extern __shared__ __align__(16) float q2_smem[];
    __shared__ int goff[kQ2MaxGroups], done[kQ2MaxGroups], snstep, stuck;
    // Workgroup i runs on XCD i mod 8 (round-robin dispatch); XCD x takes the
    // x-th eighth of the (matrix, slab) list in order, so the slabs that run
    // together on one XCD belong to one matrix, start together, and read the
    // same block operands through that XCD's L2.
    const int nslab = (n + kQ2S - 1) / kQ2S;
    const int64_t total = int64_t(nslab) * nmat, per = (total + kXcds - 1) / kXcds;
    const int64_t pos = int64_t(blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
    if (pos >= total) return;
    const int b = int(pos / nslab), slab = int(pos % nslab);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int c0 = slab * kQ2S;
    const int P = q2_pitch(n);
    float* Zs = q2_smem;
    float* Z = Zall + size_t(b) * n * ld;
    const float* Q2 = Q2all + size_t(b) * nblk * kQ2BlockFloats;
    const int NG = q2_groups(n);
    if (tid == 0) {
        int o = 0, ns = 0;
        for (int g = 0; g < NG; ++g) {
            const int t = sb_tasks(n, g * kQ2G);
            goff[g] = o;
            done[g] = 0;
            o += t;
            ns = max(ns, (NG - 1 - g) + t);
        }
        snstep = ns;
        stuck = 0;
    }
    for (int e = tid; e < P * kQ2S; e += kQ2NT) {
        const int r = e / kQ2S - 3, c = e % kQ2S;
        Zs[c * P + r + 3] = (r >= 0 && r < n && c0 + c < n) ? Z[size_t(r) * ld + c0 + c] : 0.f;
    }
    __syncthreads();
    const int nstep = snstep;
    const int l16 = lane & 15, lq = lane >> 4;
    constexpr int NW = kQ2NT / 64;
    const int wv = __builtin_amdgcn_readfirstlane(wave);
    // Group index gi = NG - 1 - g is active at step t when gi <= t and
    // t - gi < its task count; the task count grows with gi, so the active
    // groups of a step are an interval [lo, hi] whose ends never decrease.
    auto tasks_of = [&](int gi) { return sb_tasks(n, (NG - 1 - gi) * kQ2G); };
    int lo = 0;
    // this wave's first active group at step t (t never decreasing), or -1
    auto first_of = [&](int t) {
        const int hi = min(t, NG - 1);
        while (lo <= hi && t - lo >= tasks_of(lo)) ++lo;
        const int g = lo + ((wv - lo) & (NW - 1));
        return g <= hi ? g : -1;
    };
    auto load_ops = [&](int t, int gi, float4 (&va)[3], float4 (&na)[3]) {
        const int g = NG - 1 - gi;
        const float* blk = Q2 + size_t(goff[g] + t - gi) * kQ2BlockFloats;
        const float4* V4 = reinterpret_cast<const float4*>(blk + lane * (kQ2R / 4));
        const float4* N4 = reinterpret_cast<const float4*>(blk + kQ2R * kQ2G + kQ2G + lane * 12);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            va[i] = V4[i];
            na[i] = N4[i];
        }
    };
    // the wave's blocks in order (step by step, groups in order within a
    // step); the next block's operands are in flight while a block runs, also
    // across the step barrier
    auto advance = [&](int& tt, int& g) {  // the block after (tt, g), or tt = nstep
        if (g >= 0 && g + NW <= min(tt, NG - 1)) {
            g += NW;
            return;
        }
        g = -1;
        while (g < 0 && ++tt < nstep) g = first_of(tt);
    };
    // Blocks run in the step order without a barrier per step: block (gi,
    // k) waits until group gi - 1 has finished position k, or all of its
    // positions where it has no position k (done[gi - 1] >= min(k + 1, its
    // task count)); with each wave's own earlier blocks this orders it after
    // every block whose rows it shares and that must come first.  A wait
    // points to a block of an earlier step and every wave takes its blocks in
    // step order, so every wait ends.
    int t = 0, gi = first_of(0);
    if (gi < 0) advance(t, gi);
    // one block from operands (va, na) while the next block's are read into
    // (vn, nn); the two buffer pairs alternate, so no copy waits for a read
    // in flight
    auto block = [&](const float4 (&va)[3], const float4 (&na)[3], float4 (&vn)[3], float4 (&nn)[3]) {
        {
            int t2 = t, g2 = gi;
            advance(t2, g2);
            // the next block's operands (this block's again after the last)
            load_ops(t2 < nstep ? t2 : t, t2 < nstep ? g2 : gi, vn, nn);
            __builtin_amdgcn_sched_barrier(0);  // keep the reads ahead of the block
            const int k = t - gi;
            if (gi > 0) {
                // (bounded: a wait that ran far past any block's time marks
                // the slab, which is then written back as NaN, instead of
                // holding the GPU)
                const int need = min(k + 1, tasks_of(gi - 1));
                for (int spin = 0;
                     __hip_atomic_load(&done[gi - 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) < need; ++spin) {
                    if (spin > (1 << 22) || stuck) {
                        stuck = 1;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
            const int g = NG - 1 - gi;
            const int rb = g * kQ2G + 1 + k * kBandNB;
            const float vs[12] = {va[0].x, va[0].y, va[0].z, va[0].w, va[1].x, va[1].y,
                                  va[1].z, va[1].w, va[2].x, va[2].y, va[2].z, va[2].w};
            const float ns[12] = {na[0].x, na[0].y, na[0].z, na[0].w, na[1].x, na[1].y,
                                  na[1].z, na[1].w, na[2].x, na[2].y, na[2].z, na[2].w};
            // this lane's column, from the block's first row
            float* zc = Zs + l16 * P + rb + 3;
            floatx4 zq[3];
#pragma unroll
            for (int q = 0; q < 3; ++q) zq[q] = *reinterpret_cast<const floatx4*>(zc + 16 * q + 4 * lq);
            floatx4 acc[3];
            if constexpr (BF) {
                // split-BF16x3 (x = hi + lo, products lo*hi + hi*lo + hi*hi
                // accumulated in FP32) on 16x16x16 BF16 MFMA: lane (lq, l16)
                // holds k = 4 lq .. 4 lq + 3 of its row (A) or column (B), so
                // W = V^T Zb comes out as the update's B operand, W(4 lq + r,
                // l16) in register r, with no transpose
                floatx4 w = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    shortx4 vh, vl, zh, zl;
                    split_bf16(vs[4 * q], vs[4 * q + 1], vs[4 * q + 2], vs[4 * q + 3], vh, vl);
                    split_bf16(zq[q][0], zq[q][1], zq[q][2], zq[q][3], zh, zl);
                    w = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(vl, zh, w, 0, 0, 0);
                    w = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(vh, zl, w, 0, 0, 0);
                    w = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(vh, zh, w, 0, 0, 0);
                }
                shortx4 wh, wl;
                split_bf16(w[0], w[1], w[2], w[3], wh, wl);
#pragma unroll
                for (int rt = 0; rt < 3; ++rt) {
                    acc[rt] = *reinterpret_cast<const floatx4*>(zc + 16 * rt + 4 * lq);
                    shortx4 nh, nl;
                    split_bf16(ns[4 * rt], ns[4 * rt + 1], ns[4 * rt + 2], ns[4 * rt + 3], nh, nl);
                    acc[rt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(nl, wh, acc[rt], 0, 0, 0);
                    acc[rt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(nh, wl, acc[rt], 0, 0, 0);
                    acc[rt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(nh, wh, acc[rt], 0, 0, 0);
                }
            } else {
            floatx4 w0 = {0.f, 0.f, 0.f, 0.f}, w1 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                w0 = __builtin_amdgcn_mfma_f32_16x16x4f32(vs[4 * q], zq[q][0], w0, 0, 0, 0);
                w1 = __builtin_amdgcn_mfma_f32_16x16x4f32(vs[4 * q + 1], zq[q][1], w1, 0, 0, 0);
                w0 = __builtin_amdgcn_mfma_f32_16x16x4f32(vs[4 * q + 2], zq[q][2], w0, 0, 0, 0);
                w1 = __builtin_amdgcn_mfma_f32_16x16x4f32(vs[4 * q + 3], zq[q][3], w1, 0, 0, 0);
            }
            // W as the next product's B operand: register s of lane (lq, l16)
            // must hold W(4 s + lq, l16), the accumulator holds W(4 lq + r,
            // l16) in register r, so the 4 x 4 (lq, r) blocks of each column
            // are transposed across lanes: the off-diagonal 2 x 2 blocks by
            // v_permlane32_swap, then each 2 x 2 block by v_permlane16_swap
            float wb[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) wb[r] = w0[r] + w1[r];
            {
                auto p0 = __builtin_amdgcn_permlane32_swap(__float_as_uint(wb[0]), __float_as_uint(wb[2]), false, false);
                auto p1 = __builtin_amdgcn_permlane32_swap(__float_as_uint(wb[1]), __float_as_uint(wb[3]), false, false);
                auto q0 = __builtin_amdgcn_permlane16_swap(p0[0], p1[0], false, false);
                auto q1 = __builtin_amdgcn_permlane16_swap(p0[1], p1[1], false, false);
                wb[0] = __uint_as_float(q0[0]);
                wb[1] = __uint_as_float(q0[1]);
                wb[2] = __uint_as_float(q1[0]);
                wb[3] = __uint_as_float(q1[1]);
            }
#pragma unroll
            for (int rt = 0; rt < 3; ++rt) acc[rt] = *reinterpret_cast<const floatx4*>(zc + 16 * rt + 4 * lq);
#pragma unroll
            for (int s = 0; s < 4; ++s)
#pragma unroll
                for (int rt = 0; rt < 3; ++rt)
                    acc[rt] = __builtin_amdgcn_mfma_f32_16x16x4f32(ns[4 * rt + s], wb[s], acc[rt], 0, 0, 0);
            }
#pragma unroll
            for (int rt = 0; rt < 3; ++rt) *reinterpret_cast<floatx4*>(zc + 16 * rt + 4 * lq) = acc[rt];
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
            if (lane == 0) __hip_atomic_store(&done[gi], k + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            t = t2;
            gi = g2;
        }
    };
    float4 va[3], na[3], vb[3], nb[3];
    load_ops(gi < 0 ? 0 : t, gi < 0 ? 0 : gi, va, na);  // (0, 0) is a block of every matrix
    while (t < nstep) {
        block(va, na, vb, nb);
        if (t >= nstep) break;
        block(vb, nb, va, na);
    }
    __syncthreads();
    const bool bad = stuck != 0;
    for (int e = tid; e < n * kQ2S; e += kQ2NT) {
        const int r = e / kQ2S, c = e % kQ2S;
        if (c0 + c < n) Z[size_t(r) * ld + c0 + c] = bad ? __builtin_nanf("") : Zs[c * P + r + 3];
    }
}

int q2_prep(int batch, int n, float* Q2, hipStream_t s, bool bf16) {
    const int64_t nblk = q2_blocks(n);
    if (nblk <= 0) return CIR_OK;
    dim3 grid(static_cast<unsigned>(nblk), static_cast<unsigned>(batch));
    hipLaunchKernelGGL(q2_prep_kernel, grid, dim3(64), 0, s, n, nblk, Q2, bf16 ? 1 : 0);
    CIR_CHECK_LAUNCH();
    return CIR_OK;
}

// Workgroup width of q2_apply_kernel: 1024 threads (16 waves, candidate 0)
// or 512, timed per size bucket on the first call (tuner.hpp) over a copy of
// the first few matrices' eigenvectors (the kernel works in place).
template <int NT, bool BF>
static int q2_launch(int batch, int n, int ld, const float* Q2, float* Z, int64_t nblk, hipStream_t s) {
    const size_t sm = size_t(q2_pitch(n)) * kQ2S * sizeof(float);
    CIR_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&q2_apply_kernel<NT, BF>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, int(sm)));
    // the one-dimensional grid holds at most 2^32 - 1 threads: very large
    // batches go in several launches
    const int64_t nslab = ceil_div(n, kQ2S);
    const int64_t most = std::max<int64_t>(1, ((int64_t(1) << 32) - 1) / (int64_t(NT) * nslab) - kXcds);
    for (int64_t b0 = 0; b0 < batch; b0 += most) {
        const int nb = int(std::min<int64_t>(most, batch - b0));
        dim3 grid(unsigned(ceil_div(nslab * nb, int64_t(kXcds)) * kXcds));
        hipLaunchKernelGGL((q2_apply_kernel<NT, BF>), grid, dim3(NT), sm, s, n, ld, Q2 + size_t(b0) * nblk * kQ2BlockFloats,
                           Z + size_t(b0) * n * ld, nblk, nb);
        CIR_CHECK_LAUNCH();
    }
    return CIR_OK;
}

int q2_apply(int batch, int n, int ld, const float* Q2, float* Z, hipStream_t s, bool bf16) {
    const int64_t nblk = q2_blocks(n);
    if (nblk <= 0) return CIR_OK;
    const size_t sm = size_t(q2_pitch(n)) * kQ2S * sizeof(float);
    if (sm > 160 * 1024) return fail(CIR_ERR_ARG, "q2_apply: n too large for the LDS slab");
    auto launch = [&](int cand, int nb, float* Zp) {
        if (bf16)
            return cand == 0 ? q2_launch<1024, true>(nb, n, ld, Q2, Zp, nblk, s)
                             : q2_launch<512, true>(nb, n, ld, Q2, Zp, nblk, s);
        return cand == 0 ? q2_launch<1024, false>(nb, n, ld, Q2, Zp, nblk, s)
                         : q2_launch<512, false>(nb, n, ld, Q2, Zp, nblk, s);
    };
    const TuneKey key{"q2_apply_threads", {size_bucket(n), bf16 ? 1 : 0, 0, 0, 0}};
    int cand = tune_lookup(key);
    if (cand < 0 && tune_allowed(s)) {
        const int nb = std::min(batch, 8);
        const size_t bytes = size_t(nb) * n * ld * sizeof(float);
        float* scratch = nullptr;
        if (hipMalloc(&scratch, bytes) != hipSuccess) {
            (void)hipGetLastError();  // no room to time: the default width
            scratch = nullptr;
        } else {
            int err = CIR_OK;
            const hipError_t he = hipMemcpyAsync(scratch, Z, bytes, hipMemcpyDeviceToDevice, s);
            if (he != hipSuccess) err = hip_fail(he, "hipMemcpyAsync", __FILE__, __LINE__);
            if (err == CIR_OK) cand = tune_time(key, 2, s, [&](int c) { return launch(c, nb, scratch); }, &err);
            (void)hipStreamSynchronize(s);
            (void)hipFree(scratch);
            if (err != CIR_OK) return err;
        }
    }
    return launch(cand < 0 ? 0 : cand, batch, Z);
}

}  // namespace cir
