// k_eigh.hip -- batched symmetric eigensolver for gfx950, replacing
// magma_ssyevd_gpu (calcIR.cu:420-448).
//
//   1. tridiag_kernel: Householder reduction A = Q T Q^T, one workgroup per
//      matrix (a frame's Hamiltonian), panel-blocked: within a panel of
//      kTriNB columns the two-sided update is deferred and folded into each
//      column and each symmetric matrix-vector product through the panel's
//      V and W (A_now = A_panel - V W^T - W V^T); the trailing matrix is
//      updated once per panel.  Only the lower triangle (column-major, row
//      >= column) is read or written: the symmetric product streams each
//      lower element once and uses it for both A v and A^T v, with 16-byte
//      raw buffer loads whose column offsets are scalar and whose next tile
//      is in flight while a tile is consumed; v^T A v is summed alongside.
//      The panel's V and W live row-major (one 2*kTriNB row per matrix row):
//      one pass per reflector reads each panel row once to finish the
//      previous W column (its scale from v^T A v, without a further pass),
//      form the next column and sum the next panel dot products, and the
//      FP32 trailing update runs on MFMA 32x32x2 straight from it.
//      The reflector v_j (v_j[j+1] = 1) replaces column j of A below the
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
//   4. back-transform V = Q Z in compact-WY passes of kBtPass reflectors on
//      the MFMA product (gemm.hpp): Z <- Z - Y (T (Y^T Z)), last pass first.
//      A pass joins two blocks of kBtNB reflectors, T = [T1 T12; 0 T2] with
//      T12 = -T1 (Y1^T Y2) T2, so Z is read and written once per 256
//      reflectors.

#include <cfloat>
#include <cstdlib>

#include "band.hpp"
#include "gemm.hpp"
#include "tuner.hpp"

namespace cir {

constexpr int kTriNT = 512;  // threads per matrix in the tridiagonalisation
constexpr int kTriNW = kTriNT / 64;
constexpr int kTriNB = 32;   // panel width
constexpr int kBtNB = 128;   // reflectors per compact-WY block (T factor) in the back-transform
constexpr int kBtPass = 2 * kBtNB;  // reflectors per back-transform pass over Z
constexpr int kEigNT = 256;  // threads per block of bisection / inverse iteration
constexpr double kClusterRel = 1e-7;
constexpr int kMaxCluster = 128;
constexpr int kSingleMinN = 512;     // matrices at least this large run the whole-GPU tridiagonalisation
constexpr int kSingleMaxBatch = 16;  //   in batches of at most this many (one workgroup per
                                     //   matrix needs more matrices to fill the GPU)

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
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// Buffer resource over one matrix of `bytes` bytes.  A raw buffer load whose
// per-lane offset lies at or past `bytes` returns zero, which masks rows past
// the end of a column without a branch.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t matrix_rsrc(const void* base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, int(bytes), 0x00020000);
}
constexpr int kOffNone = int(0x80000000u);  // a lane offset past every matrix

// 16 bytes of a matrix column.  Cache policy nt: the matrix streams through
// once per reflector and is far larger than L2, so it is marked non-temporal
// and leaves the caches to the panel rows (679 -> 660 ms per 2048-OH launch).
template <typename VT>
__device__ __forceinline__ VT buffer_load16(__amdgpu_buffer_rsrc_t r, int lane_off, int col_off) {
    return __builtin_bit_cast(VT, __builtin_amdgcn_raw_buffer_load_b128(r, lane_off, col_off, 2));
}

// columns per wave strip of the lower-triangle symmetric product (16 loads of
// 16 bytes in flight per lane)
template <typename T>
constexpr int strip_cols() {
    return sizeof(T) == 4 ? 16 : 8;
}

// Sum over the 16 lanes of a row (lanes 16 R .. 16 R + 15) in every lane of
// the row, by row rotations (DPP row_ror 8, 4, 2, 1).  After the rotation by
// 8 the partial sums repeat with period 8, and so on, so every lane holds the
// same bits.
template <int CTRL>
__device__ __forceinline__ float dpp_f32(float x) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, 0xf, 0xf, false));
}
__device__ __forceinline__ float row_sum16(float x) {
    x += dpp_f32<0x128>(x);
    x += dpp_f32<0x124>(x);
    x += dpp_f32<0x122>(x);
    x += dpp_f32<0x121>(x);
    return x;
}

// Adds the wave sums of the 16 per-lane column partials of a symmetric-product
// strip to sp[c + k], k < nc.  Two lane swaps halve the values each lane
// carries (v_permlane32_swap moves lanes 32-63 of the first operand and lanes
// 0-31 of the second into each other, v_permlane16_swap the odd 16-lane rows of
// the first and the even rows of the second), so that 16-lane row R ends with
// the 4-lane partials of columns 4 R .. 4 R + 3; row rotations finish them and
// lane 16 R + i adds column 4 R + i.
__device__ __forceinline__ void strip_sums16(const float (&dotc)[16], int lane, int c, int nc, float* sp) {
    float s1[8], s2[4];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const auto sw =
            __builtin_amdgcn_permlane32_swap(__float_as_uint(dotc[k]), __float_as_uint(dotc[k + 8]), false, false);
        s1[k] = __uint_as_float(sw[0]) + __uint_as_float(sw[1]);  // column k + 8 (lane >> 5)
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const auto sw =
            __builtin_amdgcn_permlane16_swap(__float_as_uint(s1[k]), __float_as_uint(s1[k + 4]), false, false);
        s2[k] = row_sum16(__uint_as_float(sw[0]) + __uint_as_float(sw[1]));  // column k + 4 (lane >> 4)
    }
    const int i = lane & 15, col = i + 4 * (lane >> 4);
    const float v = i == 0 ? s2[0] : i == 1 ? s2[1] : i == 2 ? s2[2] : s2[3];
    if (i < 4 && col < nc) atomicAdd(&sp[c + col], v);
}

// FP64 forms of the above: DPP and the lane swaps move 32-bit halves, so a
// double moves as its two halves.
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double x) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), CTRL, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double row_sum16(double x) {
    x += dpp_f64<0x128>(x);
    x += dpp_f64<0x124>(x);
    x += dpp_f64<0x122>(x);
    x += dpp_f64<0x121>(x);
    return x;
}
// (X, Y) -> (X', Y') of v_permlane32_swap / v_permlane16_swap on doubles
template <bool P32>
__device__ __forceinline__ void swap_f64(double& x, double& y) {
    const auto lo = P32 ? __builtin_amdgcn_permlane32_swap(unsigned(__double2loint(x)), unsigned(__double2loint(y)),
                                                           false, false)
                        : __builtin_amdgcn_permlane16_swap(unsigned(__double2loint(x)), unsigned(__double2loint(y)),
                                                           false, false);
    const auto hi = P32 ? __builtin_amdgcn_permlane32_swap(unsigned(__double2hiint(x)), unsigned(__double2hiint(y)),
                                                           false, false)
                        : __builtin_amdgcn_permlane16_swap(unsigned(__double2hiint(x)), unsigned(__double2hiint(y)),
                                                           false, false);
    x = __hiloint2double(int(hi[0]), int(lo[0]));
    y = __hiloint2double(int(hi[1]), int(lo[1]));
}
// Sum over the 32 lanes of a half wave (lanes 32 H .. 32 H + 31), the same
// bits in every lane: the 16-lane row sums, then the two rows of the half
// added through v_permlane16_swap of the value with itself.
__device__ __forceinline__ double half_sum32(double x) {
    x = row_sum16(x);
    double y = x;
    swap_f64<false>(x, y);
    return x + y;
}
// The 8 column sums of an FP64 strip: as strip_sums16 with one swap stage
// fewer; lane 16 R + i (i < 2) adds column 2 R + i.
__device__ __forceinline__ void strip_sums8(const double (&dotc)[8], int lane, int c, int nc, double* sp) {
    double s1[4], s2[2];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        double x = dotc[k], y = dotc[k + 4];
        swap_f64<true>(x, y);
        s1[k] = x + y;  // column k + 4 (lane >> 5)
    }
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        double x = s1[k], y = s1[k + 2];
        swap_f64<false>(x, y);
        s2[k] = row_sum16(x + y);  // column k + 2 (lane >> 4)
    }
    const int i = lane & 15, col = i + 2 * (lane >> 4);
    const double v = i == 0 ? s2[0] : s2[1];
    if (i < 2 && col < nc) atomicAdd(&sp[c + col], v);
}

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

// LDS stride of a transposed 64-row panel tile (padded against bank conflicts)
constexpr int kTileLd = 65;

// Positions of V(:, l) and W(:, l) in a row of the workgroup-per-matrix
// kernel's panel.  FP32 interleaves 16-column halves, [V0..15 | W0..15 |
// V16..31 | W16..31], so that the first 16 reflectors of a panel use only
// the first 128 bytes of each row (the fused pass then reads half the row);
// FP64 keeps [V | W].  pswap maps a V position to the W position of the
// same column and back; pcol gives the column of a position.
template <typename T>
__device__ __forceinline__ constexpr int posV(int l) {
    return sizeof(T) == 4 ? (l & 15) + 32 * (l >> 4) : l;
}
template <typename T>
__device__ __forceinline__ constexpr int posW(int l) {
    return sizeof(T) == 4 ? (l & 15) + 32 * (l >> 4) + 16 : l + kTriNB;
}
template <typename T>
__device__ __forceinline__ constexpr int pswap(int p) {
    return sizeof(T) == 4 ? p ^ 16 : (p + kTriNB) & (2 * kTriNB - 1);
}
template <typename T>
__device__ __forceinline__ constexpr int pcol(int p) {
    return sizeof(T) == 4 ? (p & 15) + 16 * (p >> 5) : p & (kTriNB - 1);
}

// LDS of the workgroup-per-matrix kernel, in elements of T: the union of
// (v, p) and, for FP64, the trailing update's two staging tiles (FP32 updates
// on MFMA straight from the panel), rounded to 16 bytes, then the small
// reduction arrays.  The FP32 kernel also keeps the first `rl` rows of each
// panel in LDS (tridiag_panel_rows), ahead of all of this.
template <typename T>
constexpr size_t tridiag_uni_elems(int n) {
    return sizeof(T) == 4 ? (2 * size_t(n) + 3) & ~size_t(3)
                          : std::max<size_t>(2 * size_t(n), size_t(2) * 2 * kTriNB * kTileLd);
}
template <typename T>
inline size_t tridiag_smem_elems(int n) {
    return tridiag_uni_elems<T>(n) + size_t(kTriNW) + 2 * kTriNB + 2 * kTriNB + size_t(kTriNW) * 2 * kTriNB +
           size_t(kTriNW);
}
constexpr size_t kLdsBytes = 160 * 1024;  // LDS of a gfx950 CU, all of it one workgroup's
// Panel rows the FP32 kernel keeps in LDS: what the rest of the 160 KiB holds
// (rows j0 .. j0 + rl - 1 of the panel at j0; the fused pass reads them every
// reflector).  Fewer than 64 (a panel's rows j and j + 1 alone need 33) and
// the head is not worth its copy-out: then none.
template <typename T>
inline int tridiag_panel_rows(int n) {
    if (sizeof(T) != 4) return 0;
    const size_t base = tridiag_smem_elems<T>(n) * sizeof(T), row = 2 * kTriNB * sizeof(T);
    if (base >= kLdsBytes) return 0;
    const int rl = int(std::min<size_t>(size_t(n), (kLdsBytes - base) / row));
    return rl >= 64 ? rl : 0;
}

typedef float floatx16 __attribute__((ext_vector_type(16)));

// Trailing update of the FP32 tridiagonalisation, A(jn:, jn:) -= V W^T + W V^T
// on the lower triangle, with MFMA 32x32x2 and no LDS.  A wave takes the
// 32-row blocks br = br0, br0 + brstep, ... and every block column bc <= br.  The
// block product is formed transposed (MFMA row index = column of A, MFMA column
// index = row of A) so that lanes run along rows of the column-major A.  Step s
// of the K = 64 sum takes panel position k = 32 h + s from lane half h: the b
// operand is the panel row of row r at k and the a operand the panel row of
// column c at the W position of the same column (or the V position of a W
// position): k ^ 16 in the interleaved panel (ILV), k ^ 32 in [V | W].
template <bool ILV>
__device__ void trailing_update_f32(int n, int ld, int jn, float* __restrict__ A, const float* __restrict__ P,
                                    int lane, int br0, int brstep) {
    constexpr int PW = 2 * kTriNB;
    const int nblk = (n - jn + 31) / 32;
    const int h = lane >> 5, l32 = lane & 31;
    for (int br = br0; br < nblk; br += brstep) {
        const int r = jn + br * 32 + l32;
        float bv[32];
        {
            const float4* src = reinterpret_cast<const float4*>(P + size_t(r < n ? r : 0) * PW + h * 32);
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const float4 t = (r < n) ? src[q] : float4{};
                bv[4 * q] = t.x;
                bv[4 * q + 1] = t.y;
                bv[4 * q + 2] = t.z;
                bv[4 * q + 3] = t.w;
            }
        }
        // a operand of block column bc: panel row of column cb + l32, halves swapped
        auto load_a = [&](float(&av)[32], int bc) {
            const int cl = jn + bc * 32 + l32;
            const float4* src =
                reinterpret_cast<const float4*>(P + size_t(cl < n ? cl : 0) * PW + (ILV ? h : 1 - h) * 32);
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const float4 t = (cl < n) ? src[ILV ? (q ^ 4) : q] : float4{};  // ILV: 16-float halves swapped
                av[4 * q] = t.x;
                av[4 * q + 1] = t.y;
                av[4 * q + 2] = t.z;
                av[4 * q + 3] = t.w;
            }
        };
        float av[32], an[32];
        load_a(av, 0);
        for (int bc = 0; bc <= br; ++bc) {
            const int cb = jn + bc * 32;
            // this block of A and the next block's operand are in flight
            // while the MFMA chain runs
            float old[16];
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                const int c = cb + 8 * (i >> 2) + 4 * h + (i & 3);
                old[i] = (c < n && r < n && r >= c) ? A[size_t(c) * ld + r] : 0.f;
            }
            if (bc < br) load_a(an, bc + 1);
            floatx16 acc;
#pragma unroll
            for (int i = 0; i < 16; ++i) acc[i] = 0.f;
#pragma unroll
            for (int q = 0; q < 32; ++q) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av[q], bv[q], acc, 0, 0, 0);
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                const int c = cb + 8 * (i >> 2) + 4 * h + (i & 3);
                if (c < n && r < n && r >= c) A[size_t(c) * ld + r] = old[i] - acc[i];
            }
#pragma unroll
            for (int q = 0; q < 32; ++q) av[q] = an[q];
        }
    }
}


// Householder tridiagonalisation of a batch of symmetric matrices, one
// workgroup per matrix, reading and updating only the lower triangle
// (column-major, A(r, c) = A[c*ld + r], r >= c).  Panels of kTriNB
// reflectors (LAPACK latrd form): within a panel the trailing block is not
// updated; column j and the product A v are corrected with the panel's V and W
// (A_eff = A - V W^T - W V^T), and the trailing block gets the rank-2kTriNB
// update once per panel.  The panel lives row-major in P (row r: V(r, 0..NB)
// then W(r, 0..NB)), read as 16-byte vectors by the lanes of one row group.  Output: d, e (T = tridiag(d, e)), tau, and the reflectors
// v_j (v_j(j+1) = 1) in column j of A below the diagonal; the rest of A is
// zeroed.  Same factorisation as LAPACK ssytrd/dsytrd (lower) on the
// column-major view, which is what MAGMA's ssyevd "Upper" on the row-major
// Hamiltonian computes (calcIR.cu:616-646).
template <typename T>
__global__ void __launch_bounds__(kTriNT, 2)
tridiag_kernel(int n, int ld, T* __restrict__ Aall, T* __restrict__ dall, T* __restrict__ eall,
               T* __restrict__ tauall, T* __restrict__ VWall, int rl) {
    using VT = typename Vec16<T>::type;
    constexpr int VW = Vec16<T>::n;
    constexpr int PW = 2 * kTriNB;  // panel row width
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* const sP = reinterpret_cast<T*>(smem_raw);  // rl x PW: panel rows j0 .. j0 + rl - 1
    T* const uni = sP + size_t(rl) * PW;
    const size_t unisz = tridiag_uni_elems<T>(n);
    T* const sv = uni;                   // Householder vector (n)
    T* const sp = uni + n;               // column / product (n)
    T* const sL = uni;                   // trailing update: row tile [PW][kTileLd]
    T* const sR = uni + PW * kTileLd;    //                  column tile [PW][kTileLd]
    T* const sred = uni + unisz;         // kTriNW
    T* const srow = sred + kTriNW;       // PW: panel row j
    T* const sdot = srow + PW;           // PW: y = V^T v and x = W^T v at the V and W positions
    T* const sdp = sdot + PW;            // kTriNW * PW: per-wave partial dots
    T* const sred2 = sdp + kTriNW * PW;  // kTriNW: v^T A v partials

    const int b = blockIdx.x;
    T* const A = Aall + size_t(b) * n * ld;
    T* const d = dall + size_t(b) * n;
    T* const e = eall + size_t(b) * n;
    T* const tau = tauall + size_t(b) * n;
    T* const P = VWall + size_t(b) * PW * ld;
    // wave index as a wave-uniform (scalar) value, so per-strip column
    // addresses are scalar and the row offset is the only per-lane part
    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);

    for (int j0 = 0; j0 < n - 1; j0 += kTriNB) {
        const int nbp = min(kTriNB, n - 1 - j0);
        // rows below lim of this panel live in LDS (sP), the rest in P
        const int lim = min(n, j0 + rl);
        auto prow = [&](int r) -> T* { return (r < lim) ? sP + size_t(r - j0) * PW : P + size_t(r) * PW; };
        // clear the panel rows this panel uses (columns not formed yet are
        // read as zeros)
        for (size_t q = size_t(lim) * PW + size_t(tid) * VW; q < size_t(n) * PW; q += size_t(kTriNT) * VW)
            *reinterpret_cast<VT*>(P + q) = VT{};
        for (int q = tid * VW; q < (lim - j0) * PW; q += kTriNT * VW) *reinterpret_cast<VT*>(sP + q) = VT{};
        __syncthreads();
        // Per reflector (column j = j0 + i), with tau_prev, K_prev, x/y (sdot),
        // v (sv) and p = A_panel v (sp) of reflector i - 1 still in LDS:
        //   A0  row j of the panel into srow, finishing W(j, i-1)
        //   A   one pass over rows r >= j, LPR lanes per row (16-byte vector
        //       of the panel row each): finishes W(r, i-1) = tau_prev (p_r -
        //       V x - W y)_r + K_prev v_r, forms u_r = A(r, j) - (V W(j,:)^T +
        //       W V(j,:)^T)_r, and accumulates sum_{r>=j+2} u_r^2 and the raw
        //       panel dots sum_{r>=j+2} P(r, l) u_r
        //   B   reflector from u; v, x = W^T v and y = V^T v from the raw dots
        //   C   p = A_panel v with v^T A_panel v alongside, which gives
        //       p'.v = tau (v^T A v - 2 x.y) and so K = -tau p'.v / 2 without a
        //       further pass
        constexpr int LPR = PW / VW;      // lanes per panel row
        constexpr int RPI = 64 / LPR;     // rows per wave instruction
        const int g = lane % LPR, rsub = lane / LPR;
        T tau_prev = T(0), K_prev = T(0);
        // group sum over the LPR lanes of one row
        auto row_sum = [&](T x) {
            if constexpr (sizeof(T) == 4 && LPR == 16) {
                return row_sum16(x);
            } else if constexpr (sizeof(T) == 8 && LPR == 32) {
                return half_sum32(x);
            } else {
#pragma unroll
                for (int off = 1; off < LPR; off <<= 1) x += __shfl_xor(x, off, 64);
                return x;
            }
        };
        // W(r, i-1) from the lane's panel-row vector xr (entries updated in
        // place); returns nothing, writes the finished entry to the panel row
        // prw (in LDS or in P)
        auto finish_w = [&](VT& xr, T* prw, int r, int icol) {
            const T* xe = reinterpret_cast<const T*>(&xr);
            T part = T(0);
#pragma unroll
            for (int q = 0; q < VW; ++q) part += xe[q] * sdot[pswap<T>(g * VW + q)];
            const T corr = row_sum(part);
            const T wnew = tau_prev * (sp[r] - corr) + K_prev * sv[r];
            const int cw = posW<T>(icol) - g * VW;  // lane-local index of W(:, icol), if this lane holds it
            if (cw >= 0 && cw < VW) {
                T* xw = reinterpret_cast<T*>(&xr);
#pragma unroll
                for (int q = 0; q < VW; ++q)
                    if (q == cw) xw[q] = wnew;
                prw[posW<T>(icol)] = wnew;
            }
        };
        for (int i = 0; i < nbp; ++i) {
            const int j = j0 + i;
            // ---- A0: srow = panel row j (W(j, i-1) finished)
            if (wave == 0 && rsub == 0) {
                T* const prw = prow(j);
                VT xr = *reinterpret_cast<const VT*>(prw + g * VW);
                if (i > 0) finish_w(xr, prw, j, i - 1);
                *reinterpret_cast<VT*>(srow + g * VW) = xr;
            }
            __syncthreads();
            // ---- A: fused pass over rows j..n-1
            {
                T coefc[VW];
#pragma unroll
                for (int q = 0; q < VW; ++q) coefc[q] = srow[pswap<T>(g * VW + q)];
                T rd[VW];
#pragma unroll
                for (int q = 0; q < VW; ++q) rd[q] = T(0);
                T sig = T(0);
                // up to 16 formed columns (FP32) sit in the first half of the
                // row; the other half is still zero and is not read
                const bool need = sizeof(T) != 4 || i > 16 || g * VW < 32;
                constexpr int kU = 4;
                // rows [r0, r1) of the panel at base (sP with row offset j0,
                // or P): two loops, each with its address space known
                auto rows = [&](T* base, int roff, int r0, int r1) {
                for (int rb = r0 + wave * RPI; rb < r1; rb += kU * kTriNW * RPI) {
                    VT xr[kU];
                    T av[kU];
#pragma unroll
                    for (int u = 0; u < kU; ++u) {
                        const int r = rb + u * kTriNW * RPI + rsub;
                        xr[u] = (r < r1 && need) ? *reinterpret_cast<const VT*>(base + size_t(r - roff) * PW + g * VW)
                                                 : VT{};
                        av[u] = (r < r1) ? A[size_t(j) * ld + r] : T(0);
                    }
#pragma unroll
                    for (int u = 0; u < kU; ++u) {
                        const int r = rb + u * kTriNW * RPI + rsub;
                        const bool live = r < r1;
                        if (i > 0 && live) finish_w(xr[u], base + size_t(r - roff) * PW, r, i - 1);
                        const T* xe = reinterpret_cast<const T*>(&xr[u]);
                        T part = T(0);
#pragma unroll
                        for (int q = 0; q < VW; ++q) part += xe[q] * coefc[q];
                        const T uu = av[u] - row_sum(part);
                        if (live && r >= j + 2) {
#pragma unroll
                            for (int q = 0; q < VW; ++q) rd[q] += xe[q] * uu;
                            if (g == 0) sig += uu * uu;
                        }
                        if (live && g == 0) sp[r] = uu;
                    }
                }
                };
                rows(sP, j0, j, lim);
                rows(P, 0, max(j, lim), n);
                // partials: rows of a wave instruction, then waves
#pragma unroll
                for (int q = 0; q < VW; ++q)
                    for (int off = LPR; off < 64; off <<= 1) rd[q] += __shfl_xor(rd[q], off, 64);
                if (rsub == 0) {
#pragma unroll
                    for (int q = 0; q < VW; ++q) sdp[wave * PW + g * VW + q] = rd[q];
                }
                sig = wave_sum(sig);
                if (lane == 0) sred[wave] = sig;
            }
            __syncthreads();
            // ---- B: the reflector
            T sig = T(0);
#pragma unroll
            for (int w = 0; w < kTriNW; ++w) sig += sred[w];
            const T alpha = sp[j + 1];
            const T dj = sp[j];
            T tj = T(0), beta = alpha, scal = T(0);
            if (sig != T(0)) {
                const T nrm = sqrt(alpha * alpha + sig);
                beta = (alpha >= T(0)) ? -nrm : nrm;
                tj = (beta - alpha) / beta;
                scal = T(1) / (alpha - beta);
            }
            if (tid < PW) {
                // x_l = W(:, l)^T v, y_l = V(:, l)^T v over the formed columns l < i
                T raw = T(0);
#pragma unroll
                for (int w = 0; w < kTriNW; ++w) raw += sdp[w * PW + tid];
                const int l = pcol<T>(tid);
                sdot[tid] = (l < i) ? scal * raw + prow(j + 1)[tid] : T(0);
            }
            for (int r = j + 1 + tid; r < n; r += kTriNT) {
                const T vr = (r == j + 1) ? T(1) : sp[r] * scal;
                sv[r] = vr;
                A[size_t(j) * ld + r] = vr;
                prow(r)[posV<T>(i)] = vr;
            }
            for (int r = tid; r <= j; r += kTriNT) A[size_t(j) * ld + r] = T(0);
            if (tid == 0) {
                d[j] = dj;
                e[j] = beta;
                tau[j] = tj;
            }
            __syncthreads();
            // ---- C: p = A_panel v from the lower triangle only: each wave
            // takes strips of kStrip columns; a lane holds a 16-byte vector of
            // rows of a (64*VW) x kStrip tile, adds the row part
            // sum_c A(r,c) v_c to p_r with LDS atomics and keeps the
            // transposed part sum_r A(r,c) v_r (r > c) of every column in a
            // register, reduced over the wave when the strip ends.  Rows above
            // the diagonal are read (the tile is aligned) and masked.
            T Kj = T(0);
            if (tj != T(0)) {
                T vav = T(0);
            {
                constexpr int kStrip = strip_cols<T>();
                const __amdgpu_buffer_rsrc_t arsrc = matrix_rsrc(A, uint32_t(size_t(n) * ld * sizeof(T)));
                for (int r = j + 1 + tid; r < n; r += kTriNT) sp[r] = T(0);
                __syncthreads();
                // Strips are dealt to the waves in snake order (wave w takes
                // strips w, 2 kTriNW - 1 - w, 2 kTriNW + w, ...), which evens
                // out the rows each wave reads.  A wave runs through its
                // strips' tiles as one stream: the loads of the next tile, of
                // this strip or of the next one, are in flight while a tile is
                // consumed.
                constexpr int kTile = 64 * VW;
                const int nstrip = (n - j - 1 + kStrip - 1) / kStrip;
                auto strip_col = [&](int m) {
                    const int s = kTriNW * m + ((m & 1) ? kTriNW - 1 - wave : wave);
                    return s < nstrip ? j + 1 + s * kStrip : n;
                };
                T dotc[kStrip];
#pragma unroll
                for (int k = 0; k < kStrip; ++k) dotc[k] = T(0);
                // column offsets are wave-uniform (scalar); columns past n
                // are clamped to n - 1 and enter with zero weight
                auto load = [&](VT(&a)[kStrip], int c, int rt) {
                    const int r = rt + lane * VW;
                    const int loff = (r < n) ? r * int(sizeof(T)) : kOffNone;
#pragma unroll
                    for (int k = 0; k < kStrip; ++k)
                        a[k] = buffer_load16<VT>(arsrc, loff, min(c + k, n - 1) * ld * int(sizeof(T)));
                };
                auto tile = [&](const VT(&a)[kStrip], int c, int rt) {
                    const int nc = min(kStrip, n - c);
                    const bool diag = rt == (c & ~(VW - 1));  // the strip's first tile holds the diagonal
                    const int r = rt + lane * VW;
                    T vr[VW], rowp[VW], vcs[kStrip];
#pragma unroll
                    for (int q = 0; q < VW; ++q) {
                        vr[q] = (r + q < n && r + q > j) ? sv[r + q] : T(0);  // rows <= j hold stale v
                        rowp[q] = T(0);
                    }
                    // v over the strip's columns (zero past n), re-read from
                    // LDS for each tile to save registers
#pragma unroll
                    for (int k = 0; k < kStrip; ++k) vcs[k] = (k < nc) ? sv[c + k] : T(0);
                    if (diag) {
                        // tile on the diagonal: row >= column for the row
                        // part, row > column for the transposed part
#pragma unroll
                        for (int k = 0; k < kStrip; ++k) {
                            const T* ak = reinterpret_cast<const T*>(&a[k]);
#pragma unroll
                            for (int q = 0; q < VW; ++q) {
                                const int rr = r + q;
                                const T val = (rr >= c + k) ? ak[q] : T(0);
                                rowp[q] += val * vcs[k];
                                dotc[k] += (rr > c + k) ? val * vr[q] : T(0);
                            }
                        }
                    } else {
#pragma unroll
                        for (int k = 0; k < kStrip; ++k) {
                            const T* ak = reinterpret_cast<const T*>(&a[k]);
#pragma unroll
                            for (int q = 0; q < VW; ++q) {
                                rowp[q] += ak[q] * vcs[k];
                                dotc[k] += ak[q] * vr[q];
                            }
                        }
                    }
#pragma unroll
                    for (int q = 0; q < VW; ++q) {
                        vav += vr[q] * rowp[q];
                        if (r + q < n && r + q >= c) atomicAdd(&sp[r + q], rowp[q]);
                    }
                };
                // the column sums of a finished strip go to sp[c ..]
                auto strip_end = [&](int c) {
                    const int nc = min(kStrip, n - c);
#pragma unroll
                    for (int k = 0; k < kStrip; ++k)
                        if (k < nc) vav += sv[c + k] * dotc[k];
                    if constexpr (sizeof(T) == 4 && kStrip == 16) {
                        strip_sums16(dotc, lane, c, nc, sp);
                    } else if constexpr (sizeof(T) == 8 && kStrip == 8) {
                        strip_sums8(dotc, lane, c, nc, sp);
                    } else {
#pragma unroll
                        for (int k = 0; k < kStrip; ++k) {
                            const T dk = wave_sum(dotc[k]);
                            if (lane == k && k < nc) atomicAdd(&sp[c + k], dk);
                        }
                    }
#pragma unroll
                    for (int k = 0; k < kStrip; ++k) dotc[k] = T(0);
                };
                // cursor: strip m at column c, tile rows from rt
                int m = 0, c = strip_col(0);
                int rt = c & ~(VW - 1);
                // consume `cur` at the cursor while `nxt` loads the next tile
                auto step = [&](const VT(&cur)[kStrip], VT(&nxt)[kStrip]) {
                    int m2 = m, c2 = c, r2 = rt + kTile;
                    if (r2 >= n) {
                        c2 = strip_col(++m2);
                        r2 = c2 & ~(VW - 1);
                    }
                    if (c2 < n) load(nxt, c2, r2);
                    tile(cur, c, rt);
                    if (c2 != c) strip_end(c);
                    m = m2;
                    c = c2;
                    rt = r2;
                    return c < n;
                };
                VT a0[kStrip], a1[kStrip];
                if (c < n) {
                    load(a0, c, rt);
                    while (step(a0, a1) && step(a1, a0)) {
                    }
                }
            }
                vav = wave_sum(vav);
                if (lane == 0) sred2[wave] = vav;
                __syncthreads();  // the product's LDS atomics and the v^T A v partials
                T vAv = T(0);
#pragma unroll
                for (int w = 0; w < kTriNW; ++w) vAv += sred2[w];
                T xy = T(0);
                for (int l = 0; l < kTriNB; ++l) xy += sdot[posW<T>(l)] * sdot[posV<T>(l)];
                Kj = T(-0.5) * tj * tj * (vAv - T(2) * xy);
            }
            tau_prev = tj;
            K_prev = Kj;
        }
        // finish W(:, nbp - 1) before the trailing update
        {
            const int jl = j0 + nbp - 1;
            for (int rb = jl + 1 + wave * RPI; rb < n; rb += kTriNW * RPI) {
                const int r = rb + rsub;
                // a row group (LPR lanes) is live or dead as a whole, so the
                // row sums never read an inactive lane
                if (r < n) {
                    T* const prw = prow(r);
                    VT xr = *reinterpret_cast<const VT*>(prw + g * VW);
                    finish_w(xr, prw, r, nbp - 1);
                }
            }
        }
        // trailing update: A(jn:, jn:) -= V W^T + W V^T on lower tiles
        const int jn = j0 + nbp;
        __syncthreads();
        // the update reads the panel from P: the LDS rows from jn on go there
        for (int q = (jn - j0) * PW + tid * VW; q < (lim - j0) * PW; q += kTriNT * VW)
            *reinterpret_cast<VT*>(P + size_t(j0) * PW + q) = *reinterpret_cast<const VT*>(sP + q);
        __syncthreads();
        if constexpr (sizeof(T) == 4) {
            if (jn < n) trailing_update_f32<true>(n, ld, jn, A, P, lane, wave, kTriNW);
        } else if (jn < n) {
            const int mt = n - jn;
            const int ntile = (mt + 63) / 64;
            const int x = tid & 63, yb = tid >> 6;
            for (int tr = 0; tr < ntile; ++tr) {
                const int rbase = jn + tr * 64;
                __syncthreads();
                // sL[l][xx] = (V(r, l), W(r, l - NB)) of rows rbase + xx
                for (int q = tid; q < PW * 64; q += kTriNT) {
                    const int xx = q / PW, l = q % PW, r = rbase + xx;
                    sL[l * kTileLd + xx] = (r < n) ? P[size_t(r) * PW + l] : T(0);
                }
                for (int tc = 0; tc <= tr; ++tc) {
                    const int cbase = jn + tc * 64;
                    __syncthreads();
                    // sR[l][yy] = (W(c, l), V(c, l - NB)) of columns cbase + yy
                    for (int q = tid; q < PW * 64; q += kTriNT) {
                        const int yy = q / PW, l = q % PW, c = cbase + yy;
                        const int ls = l < kTriNB ? l + kTriNB : l - kTriNB;
                        sR[ls * kTileLd + yy] = (c < n) ? P[size_t(c) * PW + l] : T(0);
                    }
                    __syncthreads();
                    // thread (x, yb): row rbase + x, columns cbase + yb + kTriNW q
                    constexpr int kCols = 64 / kTriNW;
                    const int r = rbase + x;
                    if (r < n) {
                        T a[kCols];
#pragma unroll
                        for (int q = 0; q < kCols; ++q) {
                            const int c = cbase + yb + kTriNW * q;
                            a[q] = (c < n && r >= c) ? A[size_t(c) * ld + r] : T(0);
                        }
                        for (int l = 0; l < 2 * nbp; ++l) {
                            const int lr = l < nbp ? l : l - nbp + kTriNB;
                            const T lv = sL[lr * kTileLd + x];
#pragma unroll
                            for (int q = 0; q < kCols; ++q) a[q] -= lv * sR[lr * kTileLd + yb + kTriNW * q];
                        }
#pragma unroll
                        for (int q = 0; q < kCols; ++q) {
                            const int c = cbase + yb + kTriNW * q;
                            if (c < n && r >= c) A[size_t(c) * ld + r] = a[q];
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

// ---------------------------------------------------------------------------
// Single-matrix tridiagonalisation across the whole GPU (batch == 1: the
// reference control flow diagonalises one Hamiltonian at a time).  Same
// factorisation, panel layout and per-reflector steps as tridiag_kernel, with
// each step a grid of workgroups instead of one workgroup's phases:
//   tri1_rowpass    rows split over workgroups: finish W(:, i-1), next column
//                   u, partial panel dots and partial sum of u^2 (with FIN:
//                   only finish W(:, i-1), for the last column of a panel)
//   tri1_reflector  one workgroup: the reflector, v, x = W^T v, y = V^T v
//   tri1_symv       (row tile, strip group) items of the lower triangle:
//                   p = A_panel v and v^T A v with global atomics
//   tri1_update_*   the panel's trailing update over row blocks
// Global state between the launches: u (gu), v (gv), p (gp), x/y (gsdot),
// tau and v^T A v (gstate), per-workgroup partial sums (gpart).
// ---------------------------------------------------------------------------
constexpr int kS1NT = 256;    // threads per workgroup of the row pass, product and update
constexpr int kS1Rows = 64;   // matrix rows per row-pass workgroup
constexpr int kS1Group = 2;   // column strips per product item

inline size_t tri1_elems(int n) {
    return 3 * size_t(n) + 2 * kTriNB + 4 + size_t((n + kS1Rows - 1) / kS1Rows) * (2 * kTriNB + 1);
}

template <typename T, bool FIN>
__global__ void __launch_bounds__(kS1NT)
tri1_rowpass(int n, int ld, int j, int i, const T* __restrict__ A, T* __restrict__ P, const T* __restrict__ gp,
             const T* __restrict__ gv, const T* __restrict__ gsdot, const T* __restrict__ gstate,
             T* __restrict__ gu, T* __restrict__ gpart) {
    using VT = typename Vec16<T>::type;
    constexpr int VW = Vec16<T>::n;
    constexpr int PW = 2 * kTriNB;
    constexpr int LPR = PW / VW, RPI = 64 / LPR, NWV = kS1NT / 64;
    constexpr int NI = kS1Rows / (NWV * RPI);  // row instructions per lane
    __shared__ T sd[PW], srow[PW], spart[NWV][PW], ssig[NWV];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int g = lane % LPR, rsub = lane / LPR;
    if (tid < PW) sd[tid] = gsdot[tid];
    __syncthreads();
    T tau_prev = T(0), K_prev = T(0);
    if (i > 0) {
        tau_prev = gstate[0];
        T xy = T(0);
        for (int l = 0; l < kTriNB; ++l) xy += sd[kTriNB + l] * sd[l];
        K_prev = T(-0.5) * tau_prev * tau_prev * (gstate[1] - T(2) * xy);
    }
    auto row_sum = [&](T x) {
#pragma unroll
        for (int off = 1; off < LPR; off <<= 1) x += __shfl_xor(x, off, 64);
        return x;
    };
    auto finish_w = [&](VT& xr, int r, bool write) {
        const T* xe = reinterpret_cast<const T*>(&xr);
        T part = T(0);
#pragma unroll
        for (int q = 0; q < VW; ++q) part += xe[q] * sd[((g * VW + q) + kTriNB) & (PW - 1)];
        const T corr = row_sum(part);
        const T wnew = tau_prev * (gp[r] - corr) + K_prev * gv[r];
        const int cw = kTriNB + i - 1 - g * VW;
        if (cw >= 0 && cw < VW) {
            T* xw = reinterpret_cast<T*>(&xr);
#pragma unroll
            for (int q = 0; q < VW; ++q)
                if (q == cw) xw[q] = wnew;
            if (write) P[size_t(r) * PW + kTriNB + i - 1] = wnew;
        }
    };
    if (!FIN) {
        if (wave == 0 && rsub == 0) {
            VT xr = *reinterpret_cast<const VT*>(P + size_t(j) * PW + g * VW);
            if (i > 0) finish_w(xr, j, false);
            *reinterpret_cast<VT*>(srow + g * VW) = xr;
        }
        __syncthreads();
    }
    T coefc[VW], rd[VW];
#pragma unroll
    for (int q = 0; q < VW; ++q) {
        coefc[q] = FIN ? T(0) : srow[((g * VW + q) + kTriNB) & (PW - 1)];
        rd[q] = T(0);
    }
    T sig = T(0);
    const int r0 = j + blockIdx.x * kS1Rows;
    VT xr[NI];
    T av[NI];
#pragma unroll
    for (int t = 0; t < NI; ++t) {
        const int r = r0 + (t * NWV + wave) * RPI + rsub;
        xr[t] = (r < n) ? *reinterpret_cast<const VT*>(P + size_t(r) * PW + g * VW) : VT{};
        av[t] = (!FIN && r < n) ? A[size_t(j) * ld + r] : T(0);
    }
#pragma unroll
    for (int t = 0; t < NI; ++t) {
        const int r = r0 + (t * NWV + wave) * RPI + rsub;
        // a row group (LPR lanes) is live or dead as a whole, so the row sums
        // never read an inactive lane
        if (r < n) {
            if (i > 0) finish_w(xr[t], r, true);
            if (!FIN) {
                const T* xe = reinterpret_cast<const T*>(&xr[t]);
                T part = T(0);
#pragma unroll
                for (int q = 0; q < VW; ++q) part += xe[q] * coefc[q];
                const T u = av[t] - row_sum(part);
                if (r >= j + 2) {
#pragma unroll
                    for (int q = 0; q < VW; ++q) rd[q] += xe[q] * u;
                    if (g == 0) sig += u * u;
                }
                if (g == 0) gu[r] = u;
            }
        }
    }
    if (FIN) return;
#pragma unroll
    for (int q = 0; q < VW; ++q)
        for (int off = LPR; off < 64; off <<= 1) rd[q] += __shfl_xor(rd[q], off, 64);
    if (rsub == 0) {
#pragma unroll
        for (int q = 0; q < VW; ++q) spart[wave][g * VW + q] = rd[q];
    }
    sig = wave_sum(sig);
    if (lane == 0) ssig[wave] = sig;
    __syncthreads();
    T* out = gpart + size_t(blockIdx.x) * (PW + 1);
    if (tid < PW) {
        T s = T(0);
#pragma unroll
        for (int w = 0; w < NWV; ++w) s += spart[w][tid];
        out[tid] = s;
    }
    if (tid == PW) {
        T s = T(0);
#pragma unroll
        for (int w = 0; w < NWV; ++w) s += ssig[w];
        out[PW] = s;
    }
}

template <typename T>
__global__ void __launch_bounds__(1024)
tri1_reflector(int n, int ld, int j, int i, int nblk, T* __restrict__ A, T* __restrict__ P, const T* __restrict__ gu,
               T* __restrict__ gv, T* __restrict__ gp, T* __restrict__ gsdot, T* __restrict__ gstate,
               const T* __restrict__ gpart, T* __restrict__ d, T* __restrict__ e, T* __restrict__ tau) {
    constexpr int PW = 2 * kTriNB;
    __shared__ T red[PW + 1];
    const int tid = threadIdx.x;
    if (tid < PW + 1) {
        T s = T(0);
        for (int b = 0; b < nblk; ++b) s += gpart[size_t(b) * (PW + 1) + tid];
        red[tid] = s;
    }
    __syncthreads();
    const T sig = red[PW];
    const T alpha = gu[j + 1];
    const T dj = gu[j];
    T tj = T(0), beta = alpha, scal = T(0);
    if (sig != T(0)) {
        const T nrm = sqrt(alpha * alpha + sig);
        beta = (alpha >= T(0)) ? -nrm : nrm;
        tj = (beta - alpha) / beta;
        scal = T(1) / (alpha - beta);
    }
    if (tid < PW) {
        // x_l = W(:, l)^T v, y_l = V(:, l)^T v over the formed columns l < i
        const int l = tid & (kTriNB - 1);
        gsdot[tid] = (l < i) ? scal * red[tid] + P[size_t(j + 1) * PW + tid] : T(0);
    }
    for (int r = j + 1 + tid; r < n; r += 1024) {
        const T vr = (r == j + 1) ? T(1) : gu[r] * scal;
        gv[r] = vr;
        A[size_t(j) * ld + r] = vr;
        P[size_t(r) * PW + i] = vr;
        gp[r] = T(0);
    }
    for (int r = tid; r <= j; r += 1024) A[size_t(j) * ld + r] = T(0);
    if (tid == 0) {
        d[j] = dj;
        e[j] = beta;
        tau[j] = tj;
        gstate[0] = tj;
        gstate[1] = T(0);
    }
}

// Item w of the product: row tile t = w % ntile (64*VW rows from the aligned
// start of row j+1), strip group sg = w / ntile (kS1Group strips of kStrip
// columns from column j+1).  The row part of a tile is summed over the
// group's strips in registers before its atomics.
template <typename T>
__global__ void __launch_bounds__(kS1NT)
tri1_symv(int n, int ld, int j, int ntile, int nitems, const T* __restrict__ A, const T* __restrict__ gv,
          T* __restrict__ gp, T* __restrict__ gstate) {
    using VT = typename Vec16<T>::type;
    constexpr int VW = Vec16<T>::n;
    constexpr int kStrip = strip_cols<T>();
    constexpr int kTile = 64 * VW;
    if (gstate[0] == T(0)) return;  // tau = 0: no product needed
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(int(blockIdx.x) * (kS1NT / 64) + (int(threadIdx.x) >> 6));
    if (w >= nitems) return;
    const int t = w % ntile, sg = w / ntile;
    const int rt = ((j + 1) & ~(VW - 1)) + t * kTile;
    const int r = rt + lane * VW;
    const __amdgpu_buffer_rsrc_t arsrc = matrix_rsrc(A, uint32_t(size_t(n) * ld * sizeof(T)));
    const int loff = (r < n) ? r * int(sizeof(T)) : kOffNone;
    T vr[VW], rowp[VW];
#pragma unroll
    for (int q = 0; q < VW; ++q) {
        vr[q] = (r + q < n && r + q > j) ? gv[r + q] : T(0);
        rowp[q] = T(0);
    }
    T vav = T(0);
    for (int s = 0; s < kS1Group; ++s) {
        const int c = j + 1 + (sg * kS1Group + s) * kStrip;
        if (c >= n || rt + kTile <= c) break;  // later strips lie further right
        const int nc = min(kStrip, n - c);
        VT a[kStrip];
#pragma unroll
        for (int k = 0; k < kStrip; ++k)
            a[k] = buffer_load16<VT>(arsrc, loff, min(c + k, n - 1) * ld * int(sizeof(T)));
        T vcs[kStrip], dotc[kStrip];
#pragma unroll
        for (int k = 0; k < kStrip; ++k) {
            vcs[k] = (k < nc) ? gv[c + k] : T(0);
            dotc[k] = T(0);
        }
        if (rt < c + kStrip) {
#pragma unroll
            for (int k = 0; k < kStrip; ++k) {
                const T* ak = reinterpret_cast<const T*>(&a[k]);
#pragma unroll
                for (int q = 0; q < VW; ++q) {
                    const int rr = r + q;
                    const T val = (rr >= c + k) ? ak[q] : T(0);
                    rowp[q] += val * vcs[k];
                    dotc[k] += (rr > c + k) ? val * vr[q] : T(0);
                }
            }
        } else {
#pragma unroll
            for (int k = 0; k < kStrip; ++k) {
                const T* ak = reinterpret_cast<const T*>(&a[k]);
#pragma unroll
                for (int q = 0; q < VW; ++q) {
                    rowp[q] += ak[q] * vcs[k];
                    dotc[k] += ak[q] * vr[q];
                }
            }
        }
#pragma unroll
        for (int k = 0; k < kStrip; ++k) {
            vav += vcs[k] * dotc[k];
            const T dk = wave_sum(dotc[k]);
            if (lane == k && k < nc) atomicAdd(&gp[c + k], dk);
        }
    }
#pragma unroll
    for (int q = 0; q < VW; ++q) {
        vav += vr[q] * rowp[q];
        if (r + q < n && r + q > j && rowp[q] != T(0)) atomicAdd(&gp[r + q], rowp[q]);
    }
    vav = wave_sum(vav);
    if (lane == 0) atomicAdd(&gstate[1], vav);
}

// FP32 trailing update: wave w of the grid takes row blocks w, w + nwaves, ...
__global__ void __launch_bounds__(kS1NT)
tri1_update_f32(int n, int ld, int jn, float* __restrict__ A, const float* __restrict__ P) {
    const int lane = threadIdx.x & 63;
    const int w = int(blockIdx.x) * (kS1NT / 64) + (int(threadIdx.x) >> 6);
    trailing_update_f32<false>(n, ld, jn, A, P, lane, w, int(gridDim.x) * (kS1NT / 64));
}

// FP64 trailing update: one 64 x 64 lower tile per workgroup (blockIdx.x =
// row tile, blockIdx.y = column tile), panel tiles staged in LDS.
inline size_t tri1_update_f64_smem() { return size_t(2) * 2 * kTriNB * kTileLd * sizeof(double); }

__global__ void __launch_bounds__(kS1NT)
tri1_update_f64(int n, int ld, int jn, int nbp, double* __restrict__ A, const double* __restrict__ P) {
    constexpr int PW = 2 * kTriNB;
    const int tr = blockIdx.x, tc = blockIdx.y;
    if (tc > tr) return;
    extern __shared__ __align__(16) unsigned char smem_u[];
    double* const sL = reinterpret_cast<double*>(smem_u);
    double* const sR = sL + PW * kTileLd;
    const int tid = threadIdx.x;
    const int rbase = jn + tr * 64, cbase = jn + tc * 64;
    for (int q = tid; q < PW * 64; q += kS1NT) {
        const int xx = q / PW, l = q % PW;
        const int r = rbase + xx, c = cbase + xx;
        sL[l * kTileLd + xx] = (r < n) ? P[size_t(r) * PW + l] : 0.0;
        const int ls = l < kTriNB ? l + kTriNB : l - kTriNB;
        sR[ls * kTileLd + xx] = (c < n) ? P[size_t(c) * PW + l] : 0.0;
    }
    __syncthreads();
    constexpr int NWV = kS1NT / 64, kCols = 64 / NWV;
    const int x = tid & 63, yb = tid >> 6;
    const int r = rbase + x;
    if (r >= n) return;
    double a[kCols];
#pragma unroll
    for (int q = 0; q < kCols; ++q) {
        const int c = cbase + yb + NWV * q;
        a[q] = (c < n && r >= c) ? A[size_t(c) * ld + r] : 0.0;
    }
    for (int l = 0; l < 2 * nbp; ++l) {
        const int lr = l < nbp ? l : l - nbp + kTriNB;
        const double lv = sL[lr * kTileLd + x];
#pragma unroll
        for (int q = 0; q < kCols; ++q) a[q] -= lv * sR[lr * kTileLd + yb + NWV * q];
    }
#pragma unroll
    for (int q = 0; q < kCols; ++q) {
        const int c = cbase + yb + NWV * q;
        if (c < n && r >= c) A[size_t(c) * ld + r] = a[q];
    }
}

template <typename T>
__global__ void __launch_bounds__(kS1NT) tri1_clear(int n, int j0, T* __restrict__ P) {
    constexpr int PW = 2 * kTriNB;
    for (size_t q = size_t(j0) * PW + size_t(blockIdx.x) * kS1NT + threadIdx.x; q < size_t(n) * PW;
         q += size_t(gridDim.x) * kS1NT)
        P[q] = T(0);
}

template <typename T>
__global__ void tri1_last(int n, int ld, T* __restrict__ A, T* __restrict__ d, T* __restrict__ e,
                          T* __restrict__ tau) {
    if (threadIdx.x == 0) {
        d[n - 1] = A[size_t(n - 1) * ld + (n - 1)];
        e[n - 1] = T(0);
        tau[n - 1] = T(0);
        A[size_t(n - 1) * ld + (n - 1)] = T(0);
    }
}

template <typename T>
int tridiag_single(int n, int ld, T* A, T* d, T* e, T* tau, T* P, T* buf, hipStream_t s) {
    constexpr int VW = Vec16<T>::n;
    constexpr int kStrip = strip_cols<T>();
    constexpr int kTile = 64 * VW;
    if constexpr (sizeof(T) == 8)
        CIR_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&tri1_update_f64),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, int(tri1_update_f64_smem())));
    T* gu = buf;
    T* gv = gu + n;
    T* gp = gv + n;
    T* gsdot = gp + n;
    T* gstate = gsdot + 2 * kTriNB;
    T* gpart = gstate + 4;
    for (int j0 = 0; j0 < n - 1; j0 += kTriNB) {
        const int nbp = std::min(kTriNB, n - 1 - j0);
        hipLaunchKernelGGL(tri1_clear<T>, dim3(256), dim3(kS1NT), 0, s, n, j0, P);
        CIR_CHECK_LAUNCH();
        for (int i = 0; i < nbp; ++i) {
            const int j = j0 + i;
            const int g1 = int(ceil_div(n - j, kS1Rows));
            hipLaunchKernelGGL((tri1_rowpass<T, false>), dim3(g1), dim3(kS1NT), 0, s, n, ld, j, i, A, P, gp, gv,
                               gsdot, gstate, gu, gpart);
            CIR_CHECK_LAUNCH();
            hipLaunchKernelGGL(tri1_reflector<T>, dim3(1), dim3(1024), 0, s, n, ld, j, i, g1, A, P, gu, gv, gp,
                               gsdot, gstate, gpart, d, e, tau);
            CIR_CHECK_LAUNCH();
            if (j + 1 < n) {
                const int rb0 = (j + 1) & ~(VW - 1);
                const int ntile = int(ceil_div(n - rb0, kTile));
                const int nstrip = int(ceil_div(n - (j + 1), kStrip));
                const int nsg = int(ceil_div(nstrip, kS1Group));
                const int nitems = ntile * nsg;
                hipLaunchKernelGGL(tri1_symv<T>, dim3(unsigned(ceil_div(nitems, kS1NT / 64))), dim3(kS1NT), 0, s,
                                   n, ld, j, ntile, nitems, A, gv, gp, gstate);
                CIR_CHECK_LAUNCH();
            }
        }
        // the panel's last W column, then the trailing update
        const int jn = j0 + nbp;
        if (jn < n) {
            hipLaunchKernelGGL((tri1_rowpass<T, true>), dim3(unsigned(ceil_div(n - jn, kS1Rows))), dim3(kS1NT), 0, s,
                               n, ld, jn, nbp, A, P, gp, gv, gsdot, gstate, gu, gpart);
            CIR_CHECK_LAUNCH();
            const int nblk = int(ceil_div(n - jn, 32));
            if constexpr (sizeof(T) == 4) {
                hipLaunchKernelGGL(tri1_update_f32, dim3(unsigned(ceil_div(nblk, kS1NT / 64))), dim3(kS1NT), 0, s, n,
                                   ld, jn, reinterpret_cast<float*>(A), reinterpret_cast<const float*>(P));
            } else {
                const int nt = int(ceil_div(n - jn, 64));
                hipLaunchKernelGGL(tri1_update_f64, dim3(nt, nt), dim3(kS1NT), tri1_update_f64_smem(), s, n, ld, jn,
                                   nbp, reinterpret_cast<double*>(A), reinterpret_cast<const double*>(P));
            }
            CIR_CHECK_LAUNCH();
        }
    }
    hipLaunchKernelGGL(tri1_last<T>, dim3(1), dim3(64), 0, s, n, ld, A, d, e, tau);
    CIR_CHECK_LAUNCH();
    return CIR_OK;
}

// 1/q from the hardware reciprocal (about 24 bits) refined by Newton-Raphson
// steps r <- r + r (1 - q r): one step gives about 48 bits (measured max
// relative error 2.2e-15), two give full double precision.  The Sturm count
// below needs only the signs of its recurrence: a relative error of a few
// ulps in e^2/q is a perturbation of the off-diagonal of that size, far below
// the bisection tolerance of an FP32-derived T (1024 eps ||T||); an FP64 T
// (tolerance 4 eps ||T||) takes the second step.
template <bool FULL>
__device__ __forceinline__ double nr_rcp(double q) {
    double r = __builtin_amdgcn_rcp(q);
    r = fma(r, fma(-q, r, 1.0), r);
    if constexpr (FULL) r = fma(r, fma(-q, r, 1.0), r);
    return r;
}

// Number of eigenvalues of T below x (Sturm sequence count).  |q| >= pivmin
// keeps every reciprocal finite.
template <bool FULL>
__device__ __forceinline__ int sturm_count(int n, const double* sd, const double* se2, double x,
                                           double pivmin) {
    double q = sd[0] - x;
    if (fabs(q) < pivmin) q = -pivmin;
    int c = q < 0.0;
    for (int k = 1; k < n; ++k) {
        q = (sd[k] - x) - se2[k - 1] * nr_rcp<FULL>(q);
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
    double* sr = shd + 2 * n;  // 3 * (kEigNT/64) reductions, then kEigNT multisection counts
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
    // multisection: Sturm counts at kEigNT interior points of [gl, gu] give
    // every eigenvalue a starting bracket kEigNT + 1 times narrower
    int* scnt = reinterpret_cast<int*>(sr + 12);
    const double step = (gu - gl) / double(kEigNT + 1);
    scnt[tid] = sturm_count<sizeof(T) == 8>(n, sd, se2, gl + step * double(tid + 1), pivmin);
    __syncthreads();
    const int j = blockIdx.x * kEigNT + tid;
    if (j >= n) return;
    // eigenvalue j lies in (x_t, x_{t+1}] with t = #points whose count <= j
    int t = 0;
    for (int h = kEigNT; h >= 1; h >>= 1)
        if (t + h <= kEigNT && scnt[t + h - 1] <= j) t += h;
    double a = (t == 0) ? gl : gl + step * double(t);
    double c = (t == kEigNT) ? gu : gl + step * double(t + 1);
    // tolerance: relative to the eigenvalue, with an absolute floor scaled by
    // ||T|| (1024 eps ||T|| for an FP32-derived T, whose entries carry FP32
    // rounding, 4 eps ||T|| for FP64)
    const double abstol = (sizeof(T) == 4 ? 1024.0 : 4.0) * DBL_EPSILON * tnorm;
    for (int it = 0; it < 128 && (c - a) > fmax(2.0 * DBL_EPSILON * (fabs(a) + fabs(c)), abstol); ++it) {
        const double mid = 0.5 * (a + c);
        if (sturm_count<sizeof(T) == 8>(n, sd, se2, mid, pivmin) > j)
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

// 3 FP64 arrays and a byte array of n x n, rounded up to 256 bytes so that
// the next matrix's arrays stay aligned
__host__ __device__ inline size_t invit_scratch_bytes(int n) {
    return (25 * ((size_t(n) + 63) & ~size_t(63)) * n + 255) & ~size_t(255);
}

// Scratch of one matrix (invit_scratch_bytes): FP64 arrays X (the vectors),
// R (reciprocal pivots) and L (multipliers), then the pivot flags F (bytes),
// each of n rows by n eigenvalues padded to a multiple of 64: element (k, c),
// c the eigenvalue index, at [(c / 64) 64 n + 64 k + c % 64].  Lanes run along
// c, so every access is coalesced, and a wave's successive rows are adjacent
// 512-byte blocks (a row-major n x n layout put them n elements apart).  The LU of T - lambda_c I with partial pivoting
// has U = (u0, u1, u2) per row; only 1/u0 is kept: u1 and u2 are rebuilt from
// T, the flag and the previous row's multiplier (a pivoted row has u1 =
// d(i+1) - lambda, u2 = e(i+1); an unpivoted row u2 = 0 and u1 the carried
// e(i), times -l(i-1) after a pivoted row), which cuts the scratch traffic by
// a third.  The first iteration's right-hand side is regenerated from its
// seed, so X is first written by the back substitution.  The second iteration
// reuses the factorisation.  Factors and
// vectors are FP64 even for an FP32-derived T: eigenvectors of eigenvalues
// closer than ~1e-7 ||T|| (but not clustered) are only told apart by factors
// accurate beyond FP32, and near-degenerate pivots grow the vector past the
// FP32 range before it is normalised.
template <typename T>
__global__ void __launch_bounds__(kEigNT)
invit_kernel(int n, int ld, const T* __restrict__ dall, const T* __restrict__ eall,
             const double* __restrict__ wall, T* __restrict__ Zall, unsigned char* __restrict__ scr, int b0) {
    extern __shared__ double shd[];
    double* sd = shd;
    double* se = shd + n;
    const int bl = blockIdx.y, b = b0 + bl, tid = threadIdx.x;
    const T* d = dall + size_t(b) * n;
    const T* e = eall + size_t(b) * n;
    const double* w = wall + size_t(b) * n;
    T* Z = Zall + size_t(b) * n * ld;
    // element (k, c) at col(c) + 64 k: tiles of 64 eigenvalues (one wave),
    // rows contiguous within a tile, so a lane's walk down the rows streams
    // 512-byte blocks per wave
    const size_t nn = ((size_t(n) + 63) & ~size_t(63)) * n;
    auto col = [n](int c) { return size_t(c >> 6) * 64 * n + (c & 63); };
    double* const SX = reinterpret_cast<double*>(scr + size_t(bl) * invit_scratch_bytes(n));
    double* const SR = SX + nn;
    double* const SL = SR + nn;
    unsigned char* const SF = reinterpret_cast<unsigned char*>(SL + nn);
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
    const bool single = cend == c;

    double lam_prev = 0.0;
    for (int mb = c; mb <= cend; ++mb) {
        double lam = w[mb];
        if (mb > c) lam = fmax(lam, lam_prev + 10.0 * DBL_EPSILON * fmax(fabs(lam_prev), 1.0));
        lam_prev = lam;
        double* X = SX + col(mb);
        double* R = SR + col(mb);
        double* Lm = SL + col(mb);
        unsigned char* F = SF + col(mb);
        // iteration 1, forward: factor T - lam I = P L U.  Its right-hand
        // side is a seeded vector b taken as already transformed by P L^{-1}
        // (the start vector is P L b, as in EISPACK's tinvit), so the first
        // back substitution regenerates b and no vector is stored here.
        double a = sd[0] - lam;
        double bs = (n > 1) ? se[0] : 0.0;
        for (int i = 0; i < n - 1; ++i) {
            const double sub = se[i];
            const double dn = sd[i + 1] - lam;
            const double en = (i + 1 < n - 1) ? se[i + 1] : 0.0;
            double u0, l;
            unsigned char fl;
            if (fabs(a) >= fabs(sub)) {
                // a pivot below pivmin (0 included) is taken as +-pivmin, as
                // the back substitution's reciprocal takes it; the reciprocal
                // then stays finite and |l| <= 1
                if (fabs(a) < pivmin) a = (a < 0.0) ? -pivmin : pivmin;
                l = (sizeof(T) == 8) ? sub / a : sub * nr_rcp<false>(a);
                u0 = a;
                a = dn - l * bs;
                bs = en;
                fl = 0;
            } else {
                l = (sizeof(T) == 8) ? a / sub : a * nr_rcp<false>(sub);
                u0 = sub;
                a = bs - l * dn;
                bs = -l * en;
                fl = 1;
            }
            if (fabs(u0) < pivmin) u0 = (u0 < 0.0) ? -pivmin : pivmin;
            // reciprocal pivot: the back substitution multiplies; an FP32-derived
            // T takes the refined hardware reciprocal, an FP64 T the IEEE one
            R[size_t(i) * 64] = (sizeof(T) == 8) ? 1.0 / u0 : nr_rcp<false>(u0);
            Lm[size_t(i) * 64] = l;
            F[size_t(i) * 64] = fl;
        }
        if (fabs(a) < pivmin) a = (a < 0.0) ? -pivmin : pivmin;
        const double alast = a;
        double scale = 1.0, big = 0.0, yi = hash_uniform(uint32_t(mb), uint32_t(n - 1));
        // The solves' loops read each row's factors (and vector) from the
        // scratch kPf rows ahead, into a ring of kPf register slots: a slot is
        // read again for the row kPf further on as soon as its row is solved,
        // so a row's reads are in flight while the kPf - 1 rows before it are
        // solved.  Each pass is a chain of dependent rows, and at a few waves
        // per SIMD the read latency of one row at a time set the kernel's
        // time.  Rows of the last partial ring are solved one at a time.
        constexpr int kPf = 4;
        double rr[kPf], rl[kPf], rx[kPf];
        int rf[kPf];
        for (int it = 0; it < 2; ++it) {
            if (it == 1) {
                // forward with the kept factorisation: P L^{-1} (scale x)
                yi = X[0] * scale;
                // row i: X(i + 1), L(i), F(i)
                auto ld = [&](int j, int i) {
                    i = min(i, n - 2);
                    rx[j] = X[size_t(i + 1) * 64];
                    rl[j] = Lm[size_t(i) * 64];
                    rf[j] = F[size_t(i) * 64];
                };
                auto row = [&](int i, double xn, double l, int f) {
                    const double yn = xn * scale;
                    double yo;
                    if (f == 0) {
                        yo = yi;
                        yi = yn - l * yi;
                    } else {
                        yo = yn;
                        yi = yi - l * yn;
                    }
                    X[size_t(i) * 64] = yo;
                };
                const int nfull = ((n - 1) / kPf) * kPf;  // rows 0 .. nfull-1 through the ring
                if (nfull > 0) {
#pragma unroll
                    for (int j = 0; j < kPf; ++j) ld(j, j);
                }
                for (int i0 = 0; i0 < nfull; i0 += kPf) {
#pragma unroll
                    for (int j = 0; j < kPf; ++j) {
                        row(i0 + j, rx[j], rl[j], rf[j]);
                        ld(j, i0 + j + kPf);
                    }
                }
                for (int i = nfull; i < n - 1; ++i)
                    row(i, X[size_t(i + 1) * 64], Lm[size_t(i) * 64], F[size_t(i) * 64]);
            }
            // backward: U x = y, u1/u2 rebuilt from T and the flags
            double x1 = yi / alast, x2 = 0.0;
            X[size_t(n - 1) * 64] = x1;
            double nrm2 = x1 * x1;
            big = x1;
            int fi = (n > 1) ? F[size_t(n - 2) * 64] : 0;
            // row i: R(i), L(i - 1), F(i - 1) (rows i > 0), X(i) (second iteration)
            auto ld = [&](int j, int i) {
                i = max(i, 0);
                const int ip = max(i - 1, 0);
                rr[j] = R[size_t(i) * 64];
                rl[j] = Lm[size_t(ip) * 64];
                rf[j] = F[size_t(ip) * 64];
                rx[j] = (it == 1) ? X[size_t(i) * 64] : 0.0;
            };
            auto row = [&](int i, double r, double lpv, int fpv, double xr) {
                const int fp = (i > 0) ? fpv : 0;
                const double lp = (i > 0) ? lpv : 0.0;
                double u1, u2;
                if (fi) {
                    u1 = sd[i + 1] - lam;
                    u2 = (i + 1 < n - 1) ? se[i + 1] : 0.0;
                } else {
                    u1 = (i > 0 && fp) ? -lp * se[i] : se[i];
                    u2 = 0.0;
                }
                const double rhs = (it == 0) ? hash_uniform(uint32_t(mb), uint32_t(i)) : xr;
                const double xi = (rhs - u1 * x1 - u2 * x2) * r;
                X[size_t(i) * 64] = xi;
                nrm2 += xi * xi;
                if (fabs(xi) >= fabs(big)) big = xi;  // first largest component in ascending order
                x2 = x1;
                x1 = xi;
                fi = fp;
            };
            // rows n-2 .. n-1-nfull through the ring, then the rest one at a time
            const int nfull = ((n - 1) / kPf) * kPf;
            if (nfull > 0) {
#pragma unroll
                for (int j = 0; j < kPf; ++j) ld(j, n - 2 - j);
            }
            for (int i0 = n - 2; i0 > n - 2 - nfull; i0 -= kPf) {
#pragma unroll
                for (int j = 0; j < kPf; ++j) {
                    row(i0 - j, rr[j], rl[j], rf[j], rx[j]);
                    ld(j, i0 - j - kPf);
                }
            }
            for (int i = n - 2 - nfull; i >= 0; --i) {
                const int ip = max(i - 1, 0);
                row(i, R[size_t(i) * 64], Lm[size_t(ip) * 64], F[size_t(ip) * 64],
                    (it == 1) ? X[size_t(i) * 64] : 0.0);
            }
            // orthogonalise against the cluster's earlier vectors
            if (mb > c) {
                for (int q = c; q < mb; ++q) {
                    const double* Xq = SX + col(q);
                    double dot = 0.0;
                    for (int k = 0; k < n; ++k) dot += X[size_t(k) * 64] * Xq[size_t(k) * 64];
                    nrm2 = 0.0;
                    for (int k = 0; k < n; ++k) {
                        const double v = X[size_t(k) * 64] - dot * Xq[size_t(k) * 64];
                        X[size_t(k) * 64] = v;
                        nrm2 += v * v;
                    }
                }
            }
            scale = 1.0 / sqrt(nrm2);
        }
        if (single) {
            // normalise, largest component positive, straight into Z
            const double f = (big < 0.0) ? -scale : scale;
            for (int k = 0; k < n; ++k) Z[size_t(k) * ld + mb] = T(X[size_t(k) * 64] * f);
            continue;
        }
        // cluster members stay in scratch, normalised, for the later members
        big = 0.0;
        for (int k = 0; k < n; ++k) {
            const double v = X[size_t(k) * 64] * scale;
            X[size_t(k) * 64] = v;
            if (fabs(v) > fabs(big)) big = v;
        }
        if (big < 0.0)
            for (int k = 0; k < n; ++k) X[size_t(k) * 64] = -X[size_t(k) * 64];
    }
    if (single) return;
    for (int mb = c; mb <= cend; ++mb) {
        const double* X = SX + col(mb);
        for (int k = 0; k < n; ++k) Z[size_t(k) * ld + mb] = T(X[size_t(k) * 64]);
    }
}

// T factor of a compact-WY block of kb <= kBtNB reflectors from its Gram
// matrix G = Y^T Y: T(l,l) = tau_l, T(0:l, l) = -tau_l T(0:l, 0:l) G(0:l, l).
// G and T are kBtNB x kBtNB blocks with row pitch ldt inside per-matrix
// arrays of bstride elements.  Column l's triangular product takes four
// lanes per row (each a quarter of the terms, summed by two shuffles), so
// the longest chain per column is l / 4 products instead of l.
template <typename T>
__global__ void __launch_bounds__(4 * kBtNB)
larft_kernel(int n, int kb, int j0, const T* __restrict__ tauall, const T* __restrict__ Gall,
             T* __restrict__ Tall, int ldt, int bstride) {
    // T is built in LDS, transposed with a padded pitch (Tt[c*kLtP + r] =
    // T(r, c)): the column step reads a row of T with lanes along r and the
    // final copy reads it with lanes along c, both without bank conflicts
    constexpr int kLtP = kBtNB + 1;
    extern __shared__ __align__(16) unsigned char lsm[];
    T* const Tt = reinterpret_cast<T*>(lsm);
    T* const col = Tt + kBtNB * kLtP;
    const int b = blockIdx.x, tid = threadIdx.x;
    const int row = tid >> 2, part = tid & 3;
    const T* tau = tauall + size_t(b) * n + j0;
    const T* G = Gall + size_t(b) * bstride;
    T* Tm = Tall + size_t(b) * bstride;
    for (int q = tid; q < kBtNB * kLtP; q += 4 * kBtNB) Tt[q] = T(0);
    __syncthreads();
    for (int l = 0; l < kb; ++l) {
        const T tl = tau[l];
        if (tid < l) col[tid] = -tl * G[tid * ldt + l];
        __syncthreads();
        T acc = T(0);
        if (row < l)
            for (int q = row + part; q < l; q += 4) acc += Tt[q * kLtP + row] * col[q];
        acc += __shfl_xor(acc, 1, 64);
        acc += __shfl_xor(acc, 2, 64);
        if (row < l && part == 0) Tt[l * kLtP + row] = acc;
        if (tid == 0) Tt[l * kLtP + l] = tl;
        __syncthreads();
    }
    for (int q = tid; q < kBtNB * kBtNB; q += 4 * kBtNB) {
        const int r = q / kBtNB, c = q % kBtNB;
        Tm[r * ldt + c] = Tt[c * kLtP + r];
    }
}

// Y^T of a back-transform pass into a k-contiguous copy: Yt(m, k) =
// Y(m, k) = A[(j0 + k) * ld + j0 + off + m] at Yt[m * kBtPass + k], m < mrows,
// k < kb, through a padded 64 x 64 LDS tile (reads and writes both along
// memory).  The update Z -= Y W then stages both operands k-contiguous.
template <typename T>
__global__ void __launch_bounds__(256)
yt_kernel(int n, int ld, int j0, int off, int kb, int mrows, const T* __restrict__ Aall, T* __restrict__ Ytall,
          int64_t ystride) {
    __shared__ T tile[64][65];
    const T* A = Aall + size_t(blockIdx.z) * n * ld;
    T* Yt = Ytall + size_t(blockIdx.z) * ystride;
    const int m0 = blockIdx.x * 64, k0 = blockIdx.y * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int kk = ty; kk < 64; kk += 4) {
        const int m = m0 + tx, k = k0 + kk;
        tile[kk][tx] = (m < mrows && k < kb) ? A[size_t(j0 + k) * ld + (j0 + off + m)] : T(0);
    }
    __syncthreads();
    for (int mm = ty; mm < 64; mm += 4) {
        const int m = m0 + mm, k = k0 + tx;
        if (m < mrows && k < kb) Yt[size_t(m) * kBtPass + k] = tile[tx][mm];
    }
}

inline size_t larft_smem_bytes(size_t esz) { return (size_t(kBtNB) * (kBtNB + 1) + kBtNB) * esz; }

// The FP32 batched path can reduce to tridiagonal form in two stages
// (k_band.hip); the few-matrix whole-GPU path and FP64 reduce in one stage.
// CIR_EIGH_STAGES=1 or 2 in the environment forces one form where both apply
// (tests compare them); otherwise the two-stage form is taken from
// kTwoStageMinN on (at 256 matrices per batch the two forms take the same
// time near n = 1024; at n = 2048 the two-stage form takes 12% less).
constexpr int kTwoStageMinN = 1024;
inline bool use_two_stage(size_t esz, int batch, int n) {
    if (esz != sizeof(float) || n > sb_max_n() || n < 2 * kBandNB + 2) return false;
    if (deterministic_mode()) return true;  // the form without float atomics
    if (n >= kSingleMinN && batch <= kSingleMaxBatch) return false;
    const char* fv = std::getenv("CIR_EIGH_STAGES");
    const int forced = fv ? std::atoi(fv) : 0;
    if (forced == 1) return false;
    if (forced == 2) return true;
    return n >= kTwoStageMinN;
}

struct EighWork {
    size_t off_d, off_e, off_tau, off_vw, off_g, off_t, off_w1, off_w2, off_s1, off_scr, total;
    size_t off_bnd, off_p1, off_p2, off_yh, off_mm, off_q2;  // two-stage buffers
    bool two;
    int sub;  // matrices per inverse-iteration pass
};

inline EighWork eigh_layout(size_t esz, int batch, int n, int ld) {
    EighWork L{};
    auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
    size_t o = 0;
    L.two = use_two_stage(esz, batch, n);
    if (L.two) {
        L.off_bnd = o;
        o = al(o + esz * batch * size_t(n) * kBandW);
        L.off_p1 = o;
        o = al(o + esz * batch * size_t(n) * 2 * kBandNB);
        L.off_p2 = o;
        o = al(o + esz * batch * size_t(n) * 2 * kBandNB);
        L.off_yh = o;
        o = al(o + esz * batch * size_t(n) * kBandNB);
        L.off_mm = o;
        o = al(o + esz * batch * size_t(kBandNB) * kBandNB);
        L.off_q2 = o;
        o = al(o + esz * batch * size_t(q2_blocks(n)) * kQ2BlockFloats);
    }
    L.off_d = o;
    o = al(o + esz * batch * n);
    L.off_e = o;
    o = al(o + esz * batch * n);
    L.off_tau = o;
    o = al(o + esz * batch * n);
    L.off_vw = o;
    o = al(o + esz * batch * 2 * kTriNB * size_t(ld));
    L.off_g = o;
    o = al(o + esz * batch * kBtPass * kBtPass);
    L.off_t = o;
    o = al(o + esz * batch * kBtPass * kBtPass);
    L.off_w1 = o;
    o = al(o + esz * batch * kBtPass * size_t(n));
    L.off_w2 = o;
    o = al(o + esz * batch * kBtPass * size_t(n));
    L.off_s1 = o;  // single-matrix tridiagonalisation state
    o = al(o + esz * tri1_elems(n));
    const size_t per = invit_scratch_bytes(n);
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
    unsigned char* scr = work + L.off_scr;

    // 1. tridiagonalisation: a batch runs one workgroup per matrix; a few
    // large matrices (too few workgroups to fill the GPU) run one at a time
    // across the whole GPU
    if (ld % 4 != 0) return fail(CIR_ERR_ARG, "eigh: the leading dimension must be a multiple of 4");
    if constexpr (sizeof(T) == sizeof(float)) {
        if (L.two) {
            float* bnd = reinterpret_cast<float*>(work + L.off_bnd);
            float* q2 = reinterpret_cast<float*>(work + L.off_q2);
            {
                ProfScope ps(prof, CIR_PH_TRIDIAG, s);
                CIR_TRY(sy2sb(batch, n, ld, A, bnd, tau, reinterpret_cast<float*>(work + L.off_p1),
                              reinterpret_cast<float*>(work + L.off_yh), reinterpret_cast<float*>(work + L.off_mm),
                              s));
            }
            ProfScope ps(prof, CIR_PH_BULGE, s);
            CIR_TRY(sb2st(batch, n, bnd, dd, ee, q2, s));
            CIR_TRY(q2_prep(batch, n, q2, s, gemm_prec == CIR_PREC_BF16X3));
        }
    }
    if (L.two) {
    } else if (n >= kSingleMinN && batch <= kSingleMaxBatch) {
        ProfScope ps(prof, CIR_PH_TRIDIAG, s);
        T* s1 = reinterpret_cast<T*>(work + L.off_s1);
        for (int b = 0; b < batch; ++b)
            CIR_TRY(tridiag_single<T>(n, ld, A + size_t(b) * n * ld, dd + size_t(b) * n, ee + size_t(b) * n,
                                      tau + size_t(b) * n, vw + size_t(b) * 2 * kTriNB * ld, s1, s));
    } else {
        const int rl = tridiag_panel_rows<T>(n);
        const size_t sm = (tridiag_smem_elems<T>(n) + size_t(rl) * 2 * kTriNB) * sizeof(T);
        if (sm > kLdsBytes) return fail(CIR_ERR_ARG, "eigh: n too large for the LDS of one workgroup");
        CIR_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&tridiag_kernel<T>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, int(sm)));
        ProfScope ps(prof, CIR_PH_TRIDIAG, s);
        hipLaunchKernelGGL(tridiag_kernel<T>, dim3(batch), dim3(kTriNT), sm, s, n, ld, A, dd, ee, tau, vw, rl);
        CIR_CHECK_LAUNCH();
    }
    // 2. eigenvalues
    const size_t smb = (2 * size_t(n) + 12 + kEigNT / 2 + 1) * sizeof(double);
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
    // 4. back-transform V <- Q V: the bulge-chasing reflectors first (two
    // stages), then the Householder panels, last block of reflectors first
    CIR_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&larft_kernel<T>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, int(larft_smem_bytes(sizeof(T)))));
    if constexpr (sizeof(T) == sizeof(float)) {
        if (L.two) {
            ProfScope ps(prof, CIR_PH_BULGE_BT, s);
            CIR_TRY(q2_apply(batch, n, ld, reinterpret_cast<const float*>(work + L.off_q2), V, s,
                             gemm_prec == CIR_PREC_BF16X3));
        }
    }
    ProfScope pbt(prof, CIR_PH_BACKTRANSFORM, s);
    // reflector of column j: unit at row j + off (one stage: the
    // subdiagonal; two stages: the band edge), columns with reflectors
    // 0 .. nref-1
    const int off = L.two ? kBandNB : 1;
    int nref = n - 1;
    if (L.two) {
        int np = 0;
        for (int j0 = 0; j0 + kBandNB < n - 1; j0 += kBandNB) ++np;
        nref = np * kBandNB;
    }
    const int npass = int(ceil_div(nref, kBtPass));
    constexpr int P = kBtPass;
    const int64_t pp = int64_t(P) * P, pn = int64_t(P) * n;
    for (int pass = npass - 1; pass >= 0; --pass) {
        const int j0 = pass * P;
        const int kb = std::min(P, nref - j0);
        const int kb1 = std::min(kBtNB, kb), kb2 = kb - kb1;
        const int mrows = n - j0 - off;
        if (mrows <= 0 || kb <= 0) continue;
        const T* Y = A + size_t(j0) * ld + (j0 + off);  // Y(r, l) = Y[l*ld + r]
        T* Zs = V + size_t(j0 + off) * ld;                // rows j0+off.. of V
        // C(m, n) = alpha op(A) B + beta C on the per-matrix arrays; the
        // strides are (m, k) of A, (k, n) of B, (m, n) of C
        auto mm = [&](int M, int N, int K, const T* a, int64_t a_m, int64_t a_k, int64_t a_b, const T* b,
                      int64_t b_k, int64_t b_n, int64_t b_b, T* c, int64_t c_m, int64_t c_n, int64_t c_b,
                      double alpha, double beta) {
            GemmArgs g;
            g.M = M;
            g.N = N;
            g.K = K;
            g.batch = batch;
            g.A = a;
            g.a_m = a_m;
            g.a_k = a_k;
            g.a_b = a_b;
            g.B = b;
            g.b_k = b_k;
            g.b_n = b_n;
            g.b_b = b_b;
            g.C = c;
            g.c_m = c_m;
            g.c_n = c_n;
            g.c_b = c_b;
            g.alpha = alpha;
            g.beta = beta;
            return gemm(gemm_prec, g, s);
        };
        const int64_t mat = int64_t(n) * ld;
        // G = Y^T Y (kb x kb, pitch P)
        CIR_TRY(mm(kb, kb, mrows, Y, ld, 1, mat, Y, 1, ld, mat, G, P, 1, pp, 1.0, 0.0));
        // T1, T2 from the diagonal blocks of G; T12 = -T1 (G12 T2), with
        // G12 T2 staged in W2
        hipLaunchKernelGGL(larft_kernel<T>, dim3(batch), dim3(4 * kBtNB), larft_smem_bytes(sizeof(T)), s, n, kb1, j0, tau,
                           G, Tm, P, int(pp));
        CIR_CHECK_LAUNCH();
        if (kb2 > 0) {
            hipLaunchKernelGGL(larft_kernel<T>, dim3(batch), dim3(4 * kBtNB), larft_smem_bytes(sizeof(T)), s, n, kb2,
                               j0 + kb1, tau, G + int64_t(kb1) * P + kb1, Tm + int64_t(kb1) * P + kb1, P, int(pp));
            CIR_CHECK_LAUNCH();
            CIR_TRY(mm(kb1, kb2, kb2, G + kb1, P, 1, pp, Tm + int64_t(kb1) * P + kb1, P, 1, pp, W2, P, 1, pn, 1.0,
                       0.0));
            CIR_TRY(mm(kb1, kb2, kb1, Tm, P, 1, pp, W2, P, 1, pn, Tm + kb1, P, 1, pp, -1.0, 0.0));
        }
        // W1 = Y^T Z (kb x n)
        CIR_TRY(mm(kb, n, mrows, Y, ld, 1, mat, Zs, ld, 1, mat, W1, n, 1, pn, 1.0, 0.0));
        // W2 = T W1, stored transposed (W2(k, c) at W2[c * P + k]) so that the
        // update reads it k-contiguous: formed as W2^T = W1^T T^T by block
        // columns of T^T (T21 = 0 is never formed), which keeps the stores
        // unit-stride
        CIR_TRY(mm(n, kb1, kb, W1, 1, n, pn, Tm, 1, P, pp, W2, P, 1, pn, 1.0, 0.0));
        if (kb2 > 0)
            CIR_TRY(mm(n, kb2, kb2, W1 + int64_t(kb1) * n, 1, n, pn, Tm + int64_t(kb1) * P + kb1, 1, P, pp,
                       W2 + kb1, P, 1, pn, 1.0, 0.0));
        // Z -= Y W2, with Y copied k-contiguous into W1 (free once W2 is formed)
        {
            dim3 grid(unsigned(ceil_div(mrows, 64)), unsigned(ceil_div(kb, 64)), unsigned(batch));
            hipLaunchKernelGGL(yt_kernel<T>, grid, dim3(256), 0, s, n, ld, j0, off, kb, mrows, A, W1, pn);
            CIR_CHECK_LAUNCH();
        }
        CIR_TRY(mm(mrows, n, kb, W1, P, 1, pn, W2, 1, P, pn, Zs, ld, 1, mat, -1.0, 1.0));
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
    // float arrays; the back-transform products in FP32 or split-BF16x3
    const int gemm_prec = (prec == CIR_PREC_BF16X3) ? CIR_PREC_BF16X3 : CIR_PREC_FP32;
    return eigh_impl<float>(gemm_prec, batch, n, ld, static_cast<float*>(A), static_cast<float*>(V), w, wk, s,
                            prof);
}

}  // namespace cir

extern "C" size_t cir_eigh_workspace(int prec, int batch, int n, int ld) {
    return cir::eigh_workspace(prec, batch, n, ld);
}
