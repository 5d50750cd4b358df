// k_gemm.hip -- batched dense products on gfx950 (see gemm.hpp).
//
// MFMA path: 256 threads = 4 waves in a 2x2 arrangement; each wave owns a
// (BM/2) x (BN/2) sub-tile built from 32x32 MFMA fragments.  Operand tiles
// are staged k-major in LDS (As[k][m], Bs[k][n]) so that the MFMA operand
// fetch of a wave (32 consecutive m or n per k) is bank-conflict free for
// either global layout.  BF16x3 splits each FP32 operand held in LDS into a
// bf16 high part and a bf16 residual and issues three MFMAs per fragment
// pair, accumulating in FP32.

#include "gemm.hpp"

namespace cir {

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

template <int BM, int BN, int BK, bool SPLIT>
__global__ void __launch_bounds__(256) gemm_mfma_kernel(GemmArgs g) {
    constexpr int WM = BM / 2, WN = BN / 2;
    constexpr int FM = WM / 32, FN = WN / 32;
    static_assert(FM >= 1 && FN >= 1, "wave tile must hold a 32x32 fragment");
    static_assert(!SPLIT || BK % 16 == 0, "bf16 MFMA consumes k in steps of 16");
    __shared__ float As[BK][BM + 4];
    __shared__ float Bs[BK][BN + 4];

    const int64_t b = blockIdx.z;
    const float* A = static_cast<const float*>(g.A) + b * g.a_b;
    const float* B = static_cast<const float*>(g.B) + b * g.b_b;
    float* C = static_cast<float*>(g.C) + b * g.c_b;
    const int M = g.M, N = g.N, K = g.K;
    const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave & 1) * WM, wn = (wave >> 1) * WN;
    const bool a_mc = (g.a_m == 1), b_nc = (g.b_n == 1);

    floatx16 acc[FM][FN];
    for (int i = 0; i < FM; ++i)
        for (int j = 0; j < FN; ++j)
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;

    // The next k tile is loaded into registers while the current one is
    // multiplied from LDS.
    constexpr int AE = BM * BK / 256, BE = BN * BK / 256;
    static_assert(AE * 256 == BM * BK && BE * 256 == BN * BK, "tile must split evenly over 256 threads");
    float ra[AE], rb[BE];
    auto a_coord = [&](int e, int& mm, int& kk) {
        if (a_mc) {
            mm = e % BM;
            kk = e / BM;
        } else {
            kk = e % BK;
            mm = e / BK;
        }
    };
    auto b_coord = [&](int e, int& nn, int& kk) {
        if (b_nc) {
            nn = e % BN;
            kk = e / BN;
        } else {
            kk = e % BK;
            nn = e / BK;
        }
    };
    auto gload = [&](int k0) {
#pragma unroll
        for (int t = 0; t < AE; ++t) {
            int mm, kk;
            a_coord(tid + 256 * t, mm, kk);
            const int gm = m0 + mm, gk = k0 + kk;
            ra[t] = (gm < M && gk < K) ? A[int64_t(gm) * g.a_m + int64_t(gk) * g.a_k] : 0.f;
        }
#pragma unroll
        for (int t = 0; t < BE; ++t) {
            int nn, kk;
            b_coord(tid + 256 * t, nn, kk);
            const int gn = n0 + nn, gk = k0 + kk;
            rb[t] = (gn < N && gk < K) ? B[int64_t(gk) * g.b_k + int64_t(gn) * g.b_n] : 0.f;
        }
    };
    gload(0);
    for (int k0 = 0; k0 < K; k0 += BK) {
#pragma unroll
        for (int t = 0; t < AE; ++t) {
            int mm, kk;
            a_coord(tid + 256 * t, mm, kk);
            As[kk][mm] = ra[t];
        }
#pragma unroll
        for (int t = 0; t < BE; ++t) {
            int nn, kk;
            b_coord(tid + 256 * t, nn, kk);
            Bs[kk][nn] = rb[t];
        }
        __syncthreads();
        if (k0 + BK < K) gload(k0 + BK);
        if constexpr (!SPLIT) {
#pragma unroll
            for (int kk = 0; kk < BK; kk += 2) {
                float af[FM], bf[FN];
                for (int i = 0; i < FM; ++i) af[i] = As[kk + (lane >> 5)][wm + i * 32 + (lane & 31)];
                for (int j = 0; j < FN; ++j) bf[j] = Bs[kk + (lane >> 5)][wn + j * 32 + (lane & 31)];
                for (int i = 0; i < FM; ++i)
                    for (int j = 0; j < FN; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(af[i], bf[j], acc[i][j], 0, 0, 0);
            }
        } else {
#pragma unroll
            for (int kk = 0; kk < BK; kk += 16) {
                bf16x8 ahi[FM], alo[FM], bhi[FN], blo[FN];
                const int kb = kk + 8 * (lane >> 5);
                for (int i = 0; i < FM; ++i)
                    for (int q = 0; q < 8; ++q) {
                        const float v = As[kb + q][wm + i * 32 + (lane & 31)];
                        const __bf16 h = __bf16(v);
                        ahi[i][q] = h;
                        alo[i][q] = __bf16(v - float(h));
                    }
                for (int j = 0; j < FN; ++j)
                    for (int q = 0; q < 8; ++q) {
                        const float v = Bs[kb + q][wn + j * 32 + (lane & 31)];
                        const __bf16 h = __bf16(v);
                        bhi[j][q] = h;
                        blo[j][q] = __bf16(v - float(h));
                    }
                for (int i = 0; i < FM; ++i)
                    for (int j = 0; j < FN; ++j) {
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(alo[i], bhi[j], acc[i][j], 0, 0, 0);
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi[i], blo[j], acc[i][j], 0, 0, 0);
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi[i], bhi[j], acc[i][j], 0, 0, 0);
                    }
            }
        }
        __syncthreads();
    }

    const float alpha = float(g.alpha), beta = float(g.beta);
    for (int i = 0; i < FM; ++i)
        for (int j = 0; j < FN; ++j)
            for (int r = 0; r < 16; ++r) {
                const int m = m0 + wm + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                const int n = n0 + wn + j * 32 + (lane & 31);
                if (m < M && n < N) {
                    float* c = C + int64_t(m) * g.c_m + int64_t(n) * g.c_n;
                    float v = alpha * acc[i][j][r];
                    if (beta != 0.f) v += beta * *c;
                    *c = v;
                }
            }
}

// FP64: 64x64 tile, 256 threads, 4x4 outputs per thread.
__global__ void __launch_bounds__(256) gemm_f64_kernel(GemmArgs g) {
    constexpr int BM = 64, BN = 64, BK = 16;
    __shared__ double As[BK][BM + 1];
    __shared__ double Bs[BK][BN + 1];
    const int64_t b = blockIdx.z;
    const double* A = static_cast<const double*>(g.A) + b * g.a_b;
    const double* B = static_cast<const double*>(g.B) + b * g.b_b;
    double* C = static_cast<double*>(g.C) + b * g.c_b;
    const int M = g.M, N = g.N, K = g.K;
    const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    const bool a_mc = (g.a_m == 1), b_nc = (g.b_n == 1);
    double acc[4][4] = {};
    for (int k0 = 0; k0 < K; k0 += BK) {
        for (int e = tid; e < BM * BK; e += 256) {
            int mm, kk;
            if (a_mc) {
                mm = e % BM;
                kk = e / BM;
            } else {
                kk = e % BK;
                mm = e / BK;
            }
            const int gm = m0 + mm, gk = k0 + kk;
            As[kk][mm] = (gm < M && gk < K) ? A[int64_t(gm) * g.a_m + int64_t(gk) * g.a_k] : 0.0;
        }
        for (int e = tid; e < BN * BK; e += 256) {
            int nn, kk;
            if (b_nc) {
                nn = e % BN;
                kk = e / BN;
            } else {
                kk = e % BK;
                nn = e / BK;
            }
            const int gn = n0 + nn, gk = k0 + kk;
            Bs[kk][nn] = (gn < N && gk < K) ? B[int64_t(gk) * g.b_k + int64_t(gn) * g.b_n] : 0.0;
        }
        __syncthreads();
#pragma unroll 4
        for (int kk = 0; kk < BK; ++kk) {
            double a[4], bb[4];
            for (int i = 0; i < 4; ++i) a[i] = As[kk][ty + 16 * i];
            for (int j = 0; j < 4; ++j) bb[j] = Bs[kk][tx + 16 * j];
            for (int i = 0; i < 4; ++i)
                for (int j = 0; j < 4; ++j) acc[i][j] = fma(a[i], bb[j], acc[i][j]);
        }
        __syncthreads();
    }
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            const int m = m0 + ty + 16 * i, n = n0 + tx + 16 * j;
            if (m < M && n < N) {
                double* c = C + int64_t(m) * g.c_m + int64_t(n) * g.c_n;
                double v = g.alpha * acc[i][j];
                if (g.beta != 0.0) v += g.beta * *c;
                *c = v;
            }
        }
}

template <int BM, int BN, int BK, bool SPLIT>
static int launch_mfma(const GemmArgs& g, hipStream_t s) {
    dim3 grid(unsigned(ceil_div(g.M, BM)), unsigned(ceil_div(g.N, BN)), unsigned(g.batch));
    hipLaunchKernelGGL((gemm_mfma_kernel<BM, BN, BK, SPLIT>), grid, dim3(256), 0, s, g);
    CIR_CHECK_LAUNCH();
    return CIR_OK;
}

int gemm(int prec, const GemmArgs& g, hipStream_t s) {
    if (g.M <= 0 || g.N <= 0 || g.batch <= 0) return CIR_OK;
    if (g.K < 0 || !g.A || !g.B || !g.C) return fail(CIR_ERR_ARG, "gemm: bad arguments");
    if ((g.a_m != 1 && g.a_k != 1) || (g.b_k != 1 && g.b_n != 1))
        return fail(CIR_ERR_ARG, "gemm: one stride of A and of B must be 1");
    if (g.batch > 65535 || ceil_div(g.N, 64) > 65535) return fail(CIR_ERR_ARG, "gemm: grid too large");
    if (prec == CIR_PREC_FP64) {
        dim3 grid(unsigned(ceil_div(g.M, 64)), unsigned(ceil_div(g.N, 64)), unsigned(g.batch));
        hipLaunchKernelGGL(gemm_f64_kernel, grid, dim3(256), 0, s, g);
        CIR_CHECK_LAUNCH();
        return CIR_OK;
    }
    const bool big = g.M >= 128 && g.N >= 128;
    if (prec == CIR_PREC_BF16X3)
        return big ? launch_mfma<128, 128, 32, true>(g, s) : launch_mfma<64, 64, 32, true>(g, s);
    return big ? launch_mfma<128, 128, 16, false>(g, s) : launch_mfma<64, 64, 16, false>(g, s);
}

}  // namespace cir
