// k_gemm.hip -- batched dense products on gfx950 (see gemm.hpp).
//
// MFMA paths: 256 threads = 4 waves in a 2x2 arrangement; each wave owns a
// (BM/2) x (BN/2) sub-tile built from 32x32 MFMA fragments.  Kernels are
// instantiated per operand layout (which stride is 1), so unit strides are
// compile-time constants; the next k tile is prefetched into registers while
// the current one is multiplied from LDS.  BF16x3 splits each FP32 element
// once, when it is staged, into a bf16 high part and a bf16 residual and
// issues three MFMAs per fragment pair, accumulating in FP32.  Where strides
// and sizes are multiples of 4, BF16x3 products take gemm_x3v_kernel:
// 16-byte operand reads, two K tiles of reads in flight, double-buffered
// LDS, and a matrix's tiles on one XCD.

#include <algorithm>

#include "gemm.hpp"
#include "split_bf16.hpp"
#include "tuner.hpp"

namespace cir {

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// One thread's share of a BR x BK operand tile (BR rows of m or n, BK of k)
// whose element (row, k) sits at P[row * rs + k * ks]; RC says rows are the
// unit-stride direction (rs == 1).  Offsets inside one matrix are 32-bit
// (gemm() checks that they fit), so the base stays a scalar and each element
// costs one VGPR offset.
template <int BR, int BK, bool RC>
struct TileLoad {
    static constexpr int E = BR * BK / 256;
    static_assert(E * 256 == BR * BK, "tile must split evenly over 256 threads");
    float v[E];
    __device__ static void coord(int tid, int t, int& row, int& kk) {
        const int e = tid + 256 * t;
        if (RC) {
            row = e % BR;
            kk = e / BR;
        } else {
            kk = e % BK;
            row = e / BK;
        }
    }
    __device__ void load(const float* __restrict__ P, int stride, int row0, int nrow, int k0, int K, int tid) {
#pragma unroll
        for (int t = 0; t < E; ++t) {
            int row, kk;
            coord(tid, t, row, kk);
            const int gr = row0 + row, gk = k0 + kk;
            const int off = RC ? gr + gk * stride : gr * stride + gk;
            v[t] = (gr < nrow && gk < K) ? P[off] : 0.f;
        }
    }
};

// FP32 MFMA: 256 threads = 4 waves in a 2x2 arrangement; each wave owns a
// (BM/2) x (BN/2) sub-tile of 32x32 fragments.  Tiles are staged k-major in
// LDS (As[k][m], Bs[k][n]) so that a wave's operand fetch (32 consecutive m or
// n per k) is conflict free; the next k tile is loaded into registers while
// the current one is multiplied.
template <int BM, int BN, int BK, bool AMC, bool BNC>
__global__ void __launch_bounds__(256) gemm_mfma_kernel(GemmArgs g) {
    constexpr int WM = BM / 2, WN = BN / 2;
    constexpr int FM = WM / 32, FN = WN / 32;
    static_assert(FM >= 1 && FN >= 1, "wave tile must hold a 32x32 fragment");
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
    const int sa = int(AMC ? g.a_k : g.a_m), sb = int(BNC ? g.b_k : g.b_n);

    floatx16 acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;

    TileLoad<BM, BK, AMC> la;
    TileLoad<BN, BK, BNC> lb;
    la.load(A, sa, m0, M, 0, K, tid);
    lb.load(B, sb, n0, N, 0, K, tid);
    for (int k0 = 0; k0 < K; k0 += BK) {
#pragma unroll
        for (int t = 0; t < la.E; ++t) {
            int row, kk;
            la.coord(tid, t, row, kk);
            As[kk][row] = la.v[t];
        }
#pragma unroll
        for (int t = 0; t < lb.E; ++t) {
            int row, kk;
            lb.coord(tid, t, row, kk);
            Bs[kk][row] = lb.v[t];
        }
        __syncthreads();
        if (k0 + BK < K) {
            la.load(A, sa, m0, M, k0 + BK, K, tid);
            lb.load(B, sb, n0, N, k0 + BK, K, tid);
        }
#pragma unroll
        for (int kk = 0; kk < BK; kk += 2) {
            float af[FM], bf[FN];
#pragma unroll
            for (int i = 0; i < FM; ++i) af[i] = As[kk + (lane >> 5)][wm + i * 32 + (lane & 31)];
#pragma unroll
            for (int j = 0; j < FN; ++j) bf[j] = Bs[kk + (lane >> 5)][wn + j * 32 + (lane & 31)];
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int j = 0; j < FN; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(af[i], bf[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }

    // epilogue: a fragment's 16 values of C are read together before any store
    const float alpha = float(g.alpha), beta = float(g.beta);
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) {
            const int n = n0 + wn + j * 32 + (lane & 31);
            float cv[16];
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int m = m0 + wm + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                cv[r] = (beta != 0.f && m < M && n < N) ? C[int64_t(m) * g.c_m + int64_t(n) * g.c_n] : 0.f;
            }
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int m = m0 + wm + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                if (m < M && n < N) C[int64_t(m) * g.c_m + int64_t(n) * g.c_n] = alpha * acc[i][j][r] + beta * cv[r];
            }
        }
}

// Split-BF16x3: each FP32 element is split once, when its tile is staged,
// into hi = bf16(v) and lo = bf16(v - hi), stored k-contiguous per row of
// m (or n) so that an MFMA operand (8 consecutive k of one row) is a single
// 16-byte LDS read.  The product is lo*hi + hi*lo + hi*hi in FP32.
// EARLYC (short K with beta != 0): the C tile is read into registers before
// the K loop, so its latency hides behind the few K steps instead of
// trailing them.
template <int BM, int BN, int BK, bool AMC, bool BNC, bool EARLYC>
__global__ void __launch_bounds__(256, 2) gemm_bf16x3_kernel(GemmArgs g) {
    constexpr int WM = BM / 2, WN = BN / 2;
    constexpr int FM = WM / 32, FN = WN / 32;
    constexpr int KP = BK + 8;  // row pitch in bf16 (16-byte aligned rows, staggered banks)
    static_assert(FM >= 1 && FN >= 1 && BK % 16 == 0, "tile shape");
    __shared__ __align__(16) __bf16 Ah[BM * KP];
    __shared__ __align__(16) __bf16 Al[BM * KP];
    __shared__ __align__(16) __bf16 Bh[BN * KP];
    __shared__ __align__(16) __bf16 Bl[BN * KP];

    const int64_t b = blockIdx.z;
    const float* A = static_cast<const float*>(g.A) + b * g.a_b;
    const float* B = static_cast<const float*>(g.B) + b * g.b_b;
    float* C = static_cast<float*>(g.C) + b * g.c_b;
    const int M = g.M, N = g.N, K = g.K;
    const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int l32 = lane & 31;
    const int wm = (wave & 1) * WM, wn = (wave >> 1) * WN;
    const int sa = int(AMC ? g.a_k : g.a_m), sb = int(BNC ? g.b_k : g.b_n);

    floatx16 acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;

    const float alpha = float(g.alpha), beta = float(g.beta);
    float cpre[EARLYC ? FM : 1][EARLYC ? FN : 1][16];
    if constexpr (EARLYC) {
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j) {
                const int n = n0 + wn + j * 32 + l32;
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int m = m0 + wm + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                    cpre[i][j][r] = (m < M && n < N) ? C[int64_t(m) * g.c_m + int64_t(n) * g.c_n] : 0.f;
                }
            }
    }

    TileLoad<BM, BK, AMC> la;
    TileLoad<BN, BK, BNC> lb;
    la.load(A, sa, m0, M, 0, K, tid);
    lb.load(B, sb, n0, N, 0, K, tid);
    for (int k0 = 0; k0 < K; k0 += BK) {
#pragma unroll
        for (int t = 0; t < la.E; ++t) {
            int row, kk;
            la.coord(tid, t, row, kk);
            const __bf16 h = __bf16(la.v[t]);
            Ah[row * KP + kk] = h;
            Al[row * KP + kk] = __bf16(la.v[t] - float(h));
        }
#pragma unroll
        for (int t = 0; t < lb.E; ++t) {
            int row, kk;
            lb.coord(tid, t, row, kk);
            const __bf16 h = __bf16(lb.v[t]);
            Bh[row * KP + kk] = h;
            Bl[row * KP + kk] = __bf16(lb.v[t] - float(h));
        }
        __syncthreads();
        if (k0 + BK < K) {
            la.load(A, sa, m0, M, k0 + BK, K, tid);
            lb.load(B, sb, n0, N, k0 + BK, K, tid);
        }
#pragma unroll
        for (int kk = 0; kk < BK; kk += 16) {
            const int kb = kk + 8 * (lane >> 5);
            bf16x8 ahi[FM], alo[FM], bhi[FN], blo[FN];
#pragma unroll
            for (int i = 0; i < FM; ++i) {
                const int o = (wm + i * 32 + l32) * KP + kb;
                ahi[i] = *reinterpret_cast<const bf16x8*>(&Ah[o]);
                alo[i] = *reinterpret_cast<const bf16x8*>(&Al[o]);
            }
#pragma unroll
            for (int j = 0; j < FN; ++j) {
                const int o = (wn + j * 32 + l32) * KP + kb;
                bhi[j] = *reinterpret_cast<const bf16x8*>(&Bh[o]);
                blo[j] = *reinterpret_cast<const bf16x8*>(&Bl[o]);
            }
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int j = 0; j < FN; ++j) {
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(alo[i], bhi[j], acc[i][j], 0, 0, 0);
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi[i], blo[j], acc[i][j], 0, 0, 0);
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi[i], bhi[j], acc[i][j], 0, 0, 0);
                }
        }
        __syncthreads();
    }

    // epilogue: a fragment's 16 values of C are read together before any store
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) {
            const int n = n0 + wn + j * 32 + l32;
            float cv[16];
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int m = m0 + wm + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                if constexpr (EARLYC)
                    cv[r] = cpre[i][j][r];
                else
                    cv[r] = (beta != 0.f && m < M && n < N) ? C[int64_t(m) * g.c_m + int64_t(n) * g.c_n] : 0.f;
            }
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int m = m0 + wm + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                if (m < M && n < N) C[int64_t(m) * g.c_m + int64_t(n) * g.c_n] = alpha * acc[i][j][r] + beta * cv[r];
            }
        }
}

// Split-BF16x3 with 16-byte operand reads, for operands whose strides and
// bases are multiples of 4 floats (16 bytes) and whose contiguous extent (K
// for a k-contiguous operand, the rows for a row-contiguous one) is a
// multiple of 4.  Each thread reads float4s, splits them and writes 4 hi and
// 4 lo values of one row as two 8-byte LDS stores; a row-contiguous operand is
// read as 4 x 4 blocks (4 float4s at consecutive k) and transposed in
// registers on the way.  LDS rows hold BK = 32 k with a pitch of 40 bf16 (20
// dwords): the MFMA reads (16 bytes per lane, 8 lanes per row step) and the
// staging writes (lanes 8..15 four rows below lanes 0..7) both fall in
// distinct banks.  Rows past the operand's edge are read clamped to its last
// row (or row quad): they only reach outputs past the product's edge, which
// are not stored.  PART (K not a multiple of BK): reads past K are clamped
// into the operand and their values zeroed.
template <int R, bool RC, bool PART>
struct VecTile {
    static constexpr int BK = 32, KP = 40;
    static constexpr int NV = R * BK / 4 / 256;  // float4 reads per thread
    static_assert(NV >= 1 && NV * 1024 == R * BK, "tile must split evenly over 256 threads");
    static_assert(!RC || NV % 4 == 0, "a row-contiguous tile is read in 4 x 4 blocks");
    typedef float4 Regs[NV];
    int off[RC ? NV / 4 : NV];  // element offset of each read at k0 = 0 (row clamped)
    // k-contiguous: float4 q covers row rho(q / 8), k 4 (q % 8) .. +3, with
    // rho interleaving rows so lanes 8..15 sit 4 rows below lanes 0..7;
    // row-contiguous: 4 x 4 block q covers rows 4 (q / 8) .., k 4 (q % 8) ..,
    // its float4 t the 4 rows at k 4 (q % 8) + t
    __device__ static int krow(int q) {
        const int rr = q >> 3;
        return (rr & ~7) | ((rr & 1) << 2) | ((rr >> 1) & 3);
    }
    __device__ void init(int stride, int row0, int nrow, int tid) {
        if constexpr (!RC) {
#pragma unroll
            for (int t = 0; t < NV; ++t) {
                const int q = tid + 256 * t;
                off[t] = min(row0 + krow(q), nrow - 1) * stride + 4 * (q & 7);
            }
        } else {
#pragma unroll
            for (int t4 = 0; t4 < NV / 4; ++t4) {
                const int q = tid + 256 * t4;
                off[t4] = 4 * (q & 7) * stride + min(row0 + 4 * (q >> 3), nrow - 4);
            }
        }
    }
    __device__ void load(Regs& v, const float* __restrict__ P, int stride, int k0, int K, int tid) const {
        if constexpr (!RC) {
#pragma unroll
            for (int t = 0; t < NV; ++t) {
                int o = off[t] + k0;
                if (PART) o -= max(0, k0 + 4 * ((tid + 256 * t) & 7) - (K - 4));
                v[t] = *reinterpret_cast<const float4*>(P + o);
            }
        } else {
#pragma unroll
            for (int t = 0; t < NV; ++t) {
                const int q = tid + 256 * (t >> 2);
                int o = off[t >> 2] + (k0 + (t & 3)) * stride;
                if (PART) o -= max(0, k0 + 4 * (q & 7) + (t & 3) - (K - 1)) * stride;
                v[t] = *reinterpret_cast<const float4*>(P + o);
            }
        }
    }
    // the split values into LDS (PART: zero past K)
    __device__ void store(const Regs& v, __bf16* __restrict__ H, __bf16* __restrict__ L, int k0, int K,
                          int tid) const {
        auto put = [&](int row, int kq, float a, float b, float c, float d) {
            shortx4 hi, lo;
            split_bf16(a, b, c, d, hi, lo);
            *reinterpret_cast<shortx4*>(H + row * KP + 4 * kq) = hi;
            *reinterpret_cast<shortx4*>(L + row * KP + 4 * kq) = lo;
        };
        if constexpr (!RC) {
#pragma unroll
            for (int t = 0; t < NV; ++t) {
                const int q = tid + 256 * t, row = krow(q), kq = q & 7;
                float4 x = v[t];
                if (PART && k0 + 4 * kq >= K) x = make_float4(0.f, 0.f, 0.f, 0.f);  // K % 4 == 0
                put(row, kq, x.x, x.y, x.z, x.w);
            }
        } else {
#pragma unroll
            for (int t4 = 0; t4 < NV / 4; ++t4) {
                const int q = tid + 256 * t4, rq = q >> 3, kq = q & 7;
                float4 a = v[4 * t4], b = v[4 * t4 + 1], c = v[4 * t4 + 2], d = v[4 * t4 + 3];
                if (PART) {
                    const int k = k0 + 4 * kq;
                    const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
                    if (k >= K) a = z;
                    if (k + 1 >= K) b = z;
                    if (k + 2 >= K) c = z;
                    if (k + 3 >= K) d = z;
                }
                put(4 * rq, kq, a.x, b.x, c.x, d.x);
                put(4 * rq + 1, kq, a.y, b.y, c.y, d.y);
                put(4 * rq + 2, kq, a.z, b.z, c.z, d.z);
                put(4 * rq + 3, kq, a.w, b.w, c.w, d.w);
            }
        }
    }
};

// The LDS tiles are double-buffered (one barrier per K tile) and the reads
// run two K tiles ahead of the products, in two register sets that alternate
// (the loop is unrolled by two so that each set is named statically): a
// tile's reads have two tiles' products to arrive.  Reads of tiles past the
// last repeat the last tile, so every iteration issues the same reads and the
// waits count them exactly.
template <int BM, int BN, bool AMC, bool BNC, bool PART>
__global__ void __launch_bounds__(256, 2) gemm_x3v_kernel(GemmArgs g) {
    constexpr int WM = BM / 2, WN = BN / 2;
    constexpr int FM = WM / 32, FN = WN / 32;
    constexpr int BK = 32, KP = 40;
    static_assert(FM >= 1 && FN >= 1, "tile shape");
    __shared__ __align__(16) __bf16 sm[2][2 * (BM + BN) * KP];

    // Workgroup i runs on XCD i mod 8 (round-robin dispatch); XCD x takes the
    // x-th eighth of the (matrix, column tile, row tile) list in order, so a
    // matrix's tiles run together on one XCD and read its operands from HBM
    // once, through that XCD's L2, instead of once per XCD.
    const int M = g.M, N = g.N, K = g.K;
    const int mt = (M + BM - 1) / BM, nt = (N + BN - 1) / BN;
    const int64_t total = int64_t(mt) * nt * g.batch, per = (total + kXcds - 1) / kXcds;
    const int64_t pos = int64_t(blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
    if (pos >= total) return;
    const int64_t b = pos / (int64_t(mt) * nt);
    const int tile = int(pos % (int64_t(mt) * nt));
    const float* A = static_cast<const float*>(g.A) + b * g.a_b;
    const float* B = static_cast<const float*>(g.B) + b * g.b_b;
    float* C = static_cast<float*>(g.C) + b * g.c_b;
    const int m0 = (tile % mt) * BM, n0 = (tile / mt) * BN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int l32 = lane & 31;
    const int wm = (wave & 1) * WM, wn = (wave >> 1) * WN;
    const int sa = int(AMC ? g.a_k : g.a_m), sb = int(BNC ? g.b_k : g.b_n);

    floatx16 acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;

    typedef VecTile<BM, AMC, PART> TA;
    typedef VecTile<BN, BNC, PART> TB;
    TA la;
    TB lb;
    la.init(sa, m0, M, tid);
    lb.init(sb, n0, N, tid);
    typename TA::Regs ra0, ra1;
    typename TB::Regs rb0, rb1;
    const int ntile = (K + BK - 1) / BK;
    auto load = [&](typename TA::Regs& ra, typename TB::Regs& rb, int t) {
        const int k0 = min(t, ntile - 1) * BK;
        la.load(ra, A, sa, k0, K, tid);
        lb.load(rb, B, sb, k0, K, tid);
    };
    auto stage = [&](int s, const typename TA::Regs& ra, const typename TB::Regs& rb, int t) {
        __bf16* base = sm[s];
        la.store(ra, base, base + BM * KP, t * BK, K, tid);
        lb.store(rb, base + 2 * BM * KP, base + (2 * BM + BN) * KP, t * BK, K, tid);
    };
    auto compute = [&](int s) {
        const __bf16* Ah = sm[s];
        const __bf16* Al = Ah + BM * KP;
        const __bf16* Bh = Ah + 2 * BM * KP;
        const __bf16* Bl = Bh + BN * KP;
#pragma unroll
        for (int kk = 0; kk < BK; kk += 16) {
            const int kb = kk + 8 * (lane >> 5);
            bf16x8 ahi[FM], alo[FM], bhi[FN], blo[FN];
#pragma unroll
            for (int i = 0; i < FM; ++i) {
                const int o = (wm + i * 32 + l32) * KP + kb;
                ahi[i] = *reinterpret_cast<const bf16x8*>(&Ah[o]);
                alo[i] = *reinterpret_cast<const bf16x8*>(&Al[o]);
            }
#pragma unroll
            for (int j = 0; j < FN; ++j) {
                const int o = (wn + j * 32 + l32) * KP + kb;
                bhi[j] = *reinterpret_cast<const bf16x8*>(&Bh[o]);
                blo[j] = *reinterpret_cast<const bf16x8*>(&Bl[o]);
            }
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int j = 0; j < FN; ++j) {
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(alo[i], bhi[j], acc[i][j], 0, 0, 0);
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi[i], blo[j], acc[i][j], 0, 0, 0);
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi[i], bhi[j], acc[i][j], 0, 0, 0);
                }
        }
    };
    // tile t sits in register set t % 2 until it is staged into LDS buffer
    // t % 2; the buffer is written after the previous barrier, when its last
    // readers (tile t - 2) are done
    // (the scheduling barriers keep each phase in place: the compiler would
    // otherwise move a tile's staging, and the wait for its reads, ahead of
    // the previous tile's products)
    load(ra0, rb0, 0);
    load(ra1, rb1, 1);
    stage(0, ra0, rb0, 0);
    __syncthreads();
    load(ra0, rb0, 2);
    for (int t = 0; t < ntile; t += 2) {
        __builtin_amdgcn_sched_barrier(0);
        compute(0);
        __builtin_amdgcn_sched_barrier(0);
        if (t + 1 < ntile) stage(1, ra1, rb1, t + 1);
        __syncthreads();
        load(ra1, rb1, t + 3);
        if (t + 1 >= ntile) break;
        __builtin_amdgcn_sched_barrier(0);
        compute(1);
        __builtin_amdgcn_sched_barrier(0);
        if (t + 2 < ntile) stage(0, ra0, rb0, t + 2);
        __syncthreads();
        load(ra0, rb0, t + 4);
    }

    // epilogue: C read unconditionally (indices clamped into C), stored
    // without checks in tiles inside C
    const float alpha = float(g.alpha), beta = float(g.beta);
    const bool inside = m0 + BM <= M && n0 + BN <= N;
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) {
            const int n = n0 + wn + j * 32 + l32;
            const int64_t cn = int64_t(min(n, N - 1)) * g.c_n;
            float cv[16];
            if (beta != 0.f) {
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int m = m0 + wm + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                    cv[r] = C[int64_t(min(m, M - 1)) * g.c_m + cn];
                }
            } else {
#pragma unroll
                for (int r = 0; r < 16; ++r) cv[r] = 0.f;
            }
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int m = m0 + wm + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                if (inside || (m < M && n < N)) C[int64_t(m) * g.c_m + int64_t(n) * g.c_n] = alpha * acc[i][j][r] + beta * cv[r];
            }
        }
}

// FP64 on the FP64 MFMA (16x16x4): 64x64 tile, 256 threads = 4 waves in a
// 2x2 arrangement, each wave 2x2 fragments of 16x16.  Tiles are staged k-major
// in LDS (As[k][m], Bs[k][n]); fragment operand: lane l supplies row/column
// l % 16 at k = l / 16 of a 4-deep k step; result element r of lane l is
// row l / 16 + 4 r, column l % 16 of the fragment (measured on gfx950).
typedef double doublex4 __attribute__((ext_vector_type(4)));

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
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
    const int l16 = lane & 15, kq = lane >> 4;
    const bool a_mc = (g.a_m == 1), b_nc = (g.b_n == 1);
    doublex4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[i][j][r] = 0.0;
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
#pragma unroll
        for (int k4 = 0; k4 < BK; k4 += 4) {
            double a[2], bb[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) a[i] = As[k4 + kq][wm + 16 * i + l16];
#pragma unroll
            for (int j = 0; j < 2; ++j) bb[j] = Bs[k4 + kq][wn + 16 * j + l16];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], bb[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int m = m0 + wm + 16 * i + kq + 4 * r, n = n0 + wn + 16 * j + l16;
                if (m < M && n < N) {
                    double* c = C + int64_t(m) * g.c_m + int64_t(n) * g.c_n;
                    double v = g.alpha * acc[i][j][r];
                    if (g.beta != 0.0) v += g.beta * *c;
                    *c = v;
                }
            }
}

template <int BM, int BN, int BK, bool SPLIT, bool AMC, bool BNC, bool EARLYC>
static int launch_tile(const GemmArgs& g, hipStream_t s) {
    dim3 grid(unsigned(ceil_div(g.M, BM)), unsigned(ceil_div(g.N, BN)), unsigned(g.batch));
    if constexpr (SPLIT)
        hipLaunchKernelGGL((gemm_bf16x3_kernel<BM, BN, BK, AMC, BNC, EARLYC>), grid, dim3(256), 0, s, g);
    else
        hipLaunchKernelGGL((gemm_mfma_kernel<BM, BN, BK, AMC, BNC>), grid, dim3(256), 0, s, g);
    CIR_CHECK_LAUNCH();
    return CIR_OK;
}

template <int BM, int BN, bool AMC, bool BNC>
static int launch_x3v_tile(const GemmArgs& g, hipStream_t s) {
    const int64_t total = ceil_div(g.M, BM) * ceil_div(g.N, BN) * g.batch;
    dim3 grid(unsigned(ceil_div(total, int64_t(kXcds)) * kXcds));
    if (g.K % 32 == 0)
        hipLaunchKernelGGL((gemm_x3v_kernel<BM, BN, AMC, BNC, false>), grid, dim3(256), 0, s, g);
    else
        hipLaunchKernelGGL((gemm_x3v_kernel<BM, BN, AMC, BNC, true>), grid, dim3(256), 0, s, g);
    CIR_CHECK_LAUNCH();
    return CIR_OK;
}

template <int BM, int BN>
static int launch_x3v(const GemmArgs& g, hipStream_t s) {
    const bool amc = g.a_m == 1, bnc = g.b_n == 1;
    if (amc && bnc) return launch_x3v_tile<BM, BN, true, true>(g, s);
    if (amc) return launch_x3v_tile<BM, BN, true, false>(g, s);
    if (bnc) return launch_x3v_tile<BM, BN, false, true>(g, s);
    return launch_x3v_tile<BM, BN, false, false>(g, s);
}

// Whether gemm_x3v_kernel takes the product: 16-byte aligned bases, strides
// and batch strides multiples of 4 floats, the contiguous extent of each
// operand (K, or its rows) a multiple of 4 and at least 4, and a grid that
// fits one dimension.
static bool x3v_ok(const GemmArgs& g) {
    auto al = [](const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; };
    if (!al(g.A) || !al(g.B) || g.K < 4) return false;
    // one-dimensional grid of 256-thread workgroups: at most 2^32 threads
    if (ceil_div(g.M, 128) * ceil_div(g.N, 128) * g.batch >= (int64_t(1) << 24) - kXcds) return false;
    if (g.batch > 1 && (g.a_b % 4 != 0 || g.b_b % 4 != 0)) return false;
    const bool amc = g.a_m == 1, bnc = g.b_n == 1;
    if (amc ? (g.a_k % 4 != 0 || g.M % 4 != 0) : (g.a_m % 4 != 0 || g.K % 4 != 0)) return false;
    if (bnc ? (g.b_k % 4 != 0 || g.N % 4 != 0) : (g.b_n % 4 != 0 || g.K % 4 != 0)) return false;
    return true;
}

template <int BM, int BN, int BK, bool SPLIT, bool EARLYC = false>
static int launch_layout(const GemmArgs& g, hipStream_t s) {
    const bool amc = g.a_m == 1, bnc = g.b_n == 1;
    if (amc && bnc) return launch_tile<BM, BN, BK, SPLIT, true, true, EARLYC>(g, s);
    if (amc) return launch_tile<BM, BN, BK, SPLIT, true, false, EARLYC>(g, s);
    if (bnc) return launch_tile<BM, BN, BK, SPLIT, false, true, EARLYC>(g, s);
    return launch_tile<BM, BN, BK, SPLIT, false, false, EARLYC>(g, s);
}

// Largest in-matrix element offset of a strided operand.
static int64_t max_offset(int rows, int64_t rs, int cols, int64_t cs) {
    return int64_t(std::max(rows - 1, 0)) * rs + int64_t(std::max(cols - 1, 0)) * cs;
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
    if (max_offset(g.M, g.a_m, g.K, g.a_k) >= (int64_t(1) << 31) ||
        max_offset(g.K, g.b_k, g.N, g.b_n) >= (int64_t(1) << 31))
        return fail(CIR_ERR_ARG, "gemm: a matrix spans more than 2^31 elements");
    const bool big = g.M >= 128 && g.N >= 128;
    const bool split = prec == CIR_PREC_BF16X3;
    // short-K updates C += A B in BF16x3: 64x64 tiles with the C tile read
    // ahead of the K loop (the 128x128 tile has no registers left for it)
    if (split && g.beta != 0.0 && g.K <= 128) return launch_layout<64, 64, 32, true, true>(g, s);
    // tile 0: 128x128 (the default where both sides reach 128), 1: 64x64
    // (the default for smaller products); BF16x3 products whose operands
    // allow 16-byte reads have a third candidate, first and the default at
    // every size: the double-buffered 128x128 kernel
    const bool vec = split && x3v_ok(g);
    auto launch = [&](int tile) {
        if (vec) {
            if (tile == 0) return launch_x3v<128, 128>(g, s);
            --tile;
        }
        if (split)
            return tile == 0 ? launch_layout<128, 128, 32, true>(g, s) : launch_layout<64, 64, 32, true>(g, s);
        return tile == 0 ? launch_layout<128, 128, 16, false>(g, s) : launch_layout<64, 64, 16, false>(g, s);
    };
    if (!big) return launch(vec ? 0 : 1);
    // Products that overwrite C (beta = 0) can be run again, so the first of
    // each size bucket times every tile; updates of C take the choice of
    // their bucket's product, or the first candidate.
    TuneKey key{"gemm", {int64_t(prec), size_bucket(g.M), size_bucket(g.N), size_bucket(g.K),
                         size_bucket(g.batch) * 8 + (vec ? 4 : 0) + (g.a_m == 1 ? 2 : 0) + (g.b_n == 1 ? 1 : 0)}};
    int tile = tune_lookup(key);
    if (tile < 0 && g.beta == 0.0 && tune_allowed(s)) {
        int err = CIR_OK;
        tile = tune_time(key, vec ? 3 : 2, s, launch, &err);
        if (err != CIR_OK) return err;
        return launch(tile);
    }
    return launch(tile < 0 ? 0 : tile);
}

}  // namespace cir
